// pk_kernels.hip -- gfx950 kernels of the split Riemann sweep.
//
// Device layout of one grid block: SoA planes of FP64, one plane per
// variable in xyz order [rho, mx, my, mz, E (, Bx, By, Bz, psi)], each plane
// the ghost-padded box (x fastest):  a[v*sv + x + sy*y + sz*z].
//
// The sweep kernel maps one lane to one cell of a tile of TP pencils x TS
// cells along the sweep direction and exchanges neighbour values (primitive
// states, face states, fluxes) through LDS.  Lanes along x are consecutive,
// so every global load and store is coalesced in all three directions.  In
// FUSED mode both RK2 stages (rk_coef = 1/2, 1; constants.F90:237-238) run in
// one pass: the tile recomputes stage 1 on a 4-cell halo, so U is read once
// and U1 written once; intermediates never leave the CU.  In STAGE mode one
// stage runs over whole pencils exactly as one solve_cg_riemann call does
// (solve_cg_riemann.F90:170-228), pencil ends included.
#include <hip/hip_runtime.h>
#include "pk_math.hpp"
#include "pk_mass.hpp"
#include "pk_kernels.hpp"

namespace PK_KNS {

using pk::DtArgs;
using pk::GhostArgs;
using pk::ListArgs;
using pk::StepScalars;
using pk::SweepArgs;

template <int DIR>
__device__ __forceinline__ int plane_of(int k)
{
    // sweep-order variable k -> xyz plane (fluidtypes.F90:318-320, fluidindex.F90:190-192)
    if (DIR == 0) return k;
    if (DIR == 1) {
        constexpr int p[9] = {0, 2, 1, 3, 4, 6, 5, 7, 8};
        return p[k];
    }
    constexpr int p[9] = {0, 3, 2, 1, 4, 7, 6, 5, 8};
    return p[k];
}

// Fine/coarse flux points of pencil pf at the right face of cell cs, in
// solve's order (solve_cg_riemann.F90:353-363): a flux sent by a finer
// neighbour replaces the computed one (li, ri), then the flux at a face that
// borders a coarser neighbour is saved (lo, ro), by the one lane that owns
// the cell.  Sweep-order u-fluxes of this fluid, then [B, psi].
template <class P>
__device__ __forceinline__ void fc_points(const SweepArgs &a, int pf, int cs, bool own, double (&F)[P::NV])
{
    const long long np = a.fc_npen;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        if (a.fc_at[k * np + pf] != cs) continue;
        const double *r = a.fc_in + (long long)a.fc_slot[k * np + pf] * a.fc_rec;
#pragma unroll
        for (int v = 0; v < P::NV; v++) F[v] = r[v < 5 ? a.fc_beg + v : a.fc_boff + (v - 5)];
    }
    if (!own) return;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        if (a.fc_at[(2 + k) * np + pf] != cs) continue;
        double *r = a.fc_out + (long long)a.fc_slot[(2 + k) * np + pf] * a.fc_rec;
#pragma unroll
        for (int v = 0; v < P::NV; v++) r[v < 5 ? a.fc_beg + v : a.fc_boff + (v - 5)] = F[v];
    }
}

// One RK stage of the tile.  A holds primitives then fluxes, B face states.
// LIM >= 0 fixes both limiters at compile time; LIM < 0 reads them from a.
// FC: apply the fine/coarse flux points of pencil fc_pf (-1: none) at the
// lane's cell fc_cs; fc_own marks the lane that writes saved fluxes.
template <class P, int LIM, bool FC = false>
__device__ __forceinline__ void tile_stage(const SweepArgs &a, const double (&W)[P::NV],
                                           const double (&W0)[P::NV], double coef, bool upd,
                                           bool lo_end, bool hi_end, double *__restrict__ A,
                                           double *__restrict__ B, int t, int tm, int tp, int nt,
                                           double (&Wn)[P::NV], int fc_pf = -1, int fc_cs = 0,
                                           bool fc_own = false)
{
    constexpr int NV = P::NV;
    double q[NV];
    prim<P>(W, a.gam_1, q);
#pragma unroll
    for (int v = 0; v < NV; v++) A[v * nt + t] = q[v];
    __syncthreads();

    double fp[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        const int lim = (LIM >= 0) ? LIM : ((v < 5) ? a.lim : a.limb);
        const double dl = q[v] - A[v * nt + tm];
        const double dr = A[v * nt + tp] - q[v];
        double dq;
        if (lo_end)
            dq = (lim == LIM_VANLEER) ? 0.0 : limit(lim, dr, dr);
        else if (hi_end)
            dq = (lim == LIM_VANLEER) ? 0.0 : limit(lim, dl, dl);
        else
            dq = limit(lim, dl, dr);
        const double h = 0.5 * dq;
        fp[v] = q[v] + h;
        B[v * nt + t] = q[v] - h;
    }
    __syncthreads();

    double fr[NV], F[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) fr[v] = B[v * nt + tp];
    if constexpr (P::MAG)
        flux_hlld(fp, fr, a.gam, a.ch, F);
    else
        flux_hllc(fp, fr, a.gam, F);
    if constexpr (FC) {
        if (a.fc_at && fc_pf >= 0) fc_points<P>(a, fc_pf, fc_cs, fc_own, F);
    }
#pragma unroll
    for (int v = 0; v < NV; v++) A[v * nt + t] = F[v];
    __syncthreads();

#pragma unroll
    for (int v = 0; v < NV; v++) {
        const double fm = A[v * nt + tm];
        const double un = upd ? W0[v] + coef * (fm - F[v]) : W[v];
        Wn[v] = (v < 5) ? un + 0.0 : un;  // internal_sources adds +0 to the fluid part
    }
}

// timestep_fluid (timestep.F90:460-490, ion_cs / neu_cs) of one written
// cell, folded into a lane's running min of dl/c per direction and max of c.
template <int DIR, class P>
__device__ __forceinline__ void dt_cell(const SweepArgs &a, const double (&Wn)[P::NV], double (&mn)[3], double &cmax)
{
    const double small = 1.17549435082228750797e-38;
    const double d = Wn[0];
    const double mx = Wn[plane_of<DIR>(1)], my = Wn[plane_of<DIR>(2)], mz = Wn[plane_of<DIR>(3)];
    const double e = Wn[4];
    const double m[3] = {mx, my, mz};
    double v[3] = {0.0, 0.0, 0.0};
    if (d > 0.0) {
#pragma unroll
        for (int q = 0; q < 3; q++) v[q] = fabs(qdiv(m[q], d));
    }
    const double ek = qdiv(0.5 * ((mx * mx + my * my) + mz * mz), d);
    double cs;
    if constexpr (P::MAG) {
        const double bx = Wn[plane_of<DIR>(5)], by = Wn[plane_of<DIR>(6)], bz = Wn[plane_of<DIR>(7)];
        const double pmag = 0.5 * ((bx * bx + by * by) + bz * bz);
        const double ps = (e - ek) * a.gam_1 + (2.0 - a.gam) * pmag;
        const double pr = ps - pmag;
        cs = qsqrt(fabs(qdiv(2.0 * pmag + a.gam * pr, d)));
    } else {
        const double pr = (e - ek) * a.gam_1;
        cs = qsqrt(fabs(qdiv(a.gam * pr, d)));
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const double c = fmx(v[q] + cs, small);
        cmax = fmx(cmax, c);
        if (a.has_dir[q]) mn[q] = fmn(mn[q], qdiv(a.dl[q], c));
    }
}

// The lanes' running values reduced over the wave and folded into one slot
// with 64-bit atomics (the values are positive, so their bit patterns order
// as the numbers do; min and max do not depend on the order of reduction).
__device__ __forceinline__ void dt_reduce(const SweepArgs &a, double (&mn)[3], double &cmax, unsigned slot_id)
{
    // xor butterfly over the lanes that exist (the last wave of a workgroup
    // may be partial); partners beyond it are skipped, lane 0 ends with all.
    const int lane = threadIdx.x & 63;
    const int nact = min(64, (int)blockDim.x - (int)(threadIdx.x & ~63u));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const bool take = (lane ^ o) < nact;
        double t[4];
#pragma unroll
        for (int q = 0; q < 3; q++) t[q] = __shfl_xor(mn[q], o, 64);
        t[3] = __shfl_xor(cmax, o, 64);
        if (take) {
#pragma unroll
            for (int q = 0; q < 3; q++) mn[q] = fmn(mn[q], t[q]);
            cmax = fmx(cmax, t[3]);
        }
    }
    if ((threadIdx.x & 63) == 0) {
        const int slot = (int)((slot_id * (blockDim.x / 64u) + threadIdx.x / 64u) % (unsigned)a.nslots);
        unsigned long long *s = a.dtslots + 4 * slot;
#pragma unroll
        for (int q = 0; q < 3; q++) atomicMin(s + q, (unsigned long long)__double_as_longlong(mn[q]));
        atomicMax(s + 3, (unsigned long long)__double_as_longlong(cmax));
    }
}

// timestep_fluid of the cell a lane has just written (valid lanes), reduced
// over the wave into one slot.
template <int DIR, class P>
__device__ __forceinline__ void dt_contribution(const SweepArgs &a, const double (&Wn)[P::NV], bool valid,
                                                unsigned slot_id)
{
    const double huge = 1.7976931348623157e308;
    double mn[3] = {huge, huge, huge}, cmax = 0.0;
    if (valid) dt_cell<DIR, P>(a, Wn, mn, cmax);
    dt_reduce(a, mn, cmax, slot_id);
}

// Periodic images of a written cell: each ghost cell has exactly one interior
// pre-image, so writing every output cell's images refreshes all ghosts.
template <int DIR, class P>
__device__ __forceinline__ void write_images(const SweepArgs &a, const int (&xyz)[3], long long idx,
                                             const double (&Wn)[P::NV])
{
    int off[3][3], cnt[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int ic = xyz[d] - a.g[d];
        cnt[d] = 1;
        off[d][0] = 0;
        if (a.g[d] > 0) {
            if (ic < a.g[d]) off[d][cnt[d]++] = a.n[d];
            if (ic >= a.n[d] - a.g[d]) off[d][cnt[d]++] = -a.n[d];
        }
    }
    for (int i0 = 0; i0 < cnt[0]; i0++)
        for (int i1 = 0; i1 < cnt[1]; i1++)
            for (int i2 = 0; i2 < cnt[2]; i2++) {
                if ((i0 | i1 | i2) == 0) continue;
                const long long j = idx + off[0][i0] + (long long)a.sy * off[1][i1] + (long long)a.sz * off[2][i2];
#pragma unroll
                for (int k = 0; k < P::NV; k++) a.dst[(long long)plane_of<DIR>(k) * a.sv + j] = Wn[k];
            }
}

// One tile of a sweep: chunk `bid % chunks` along the sweep direction times
// TP consecutive pencils of the transverse interior, flattened with d1
// fastest (tile `bid / chunks`).  Lanes t >= TP*TS of a larger workgroup
// take part in the barriers only.  slot_id picks the timestep slot.
template <int DIR, class P, bool FUSED, int LIM>
__device__ __forceinline__ void sweep_tile(const SweepArgs &a, int bid, unsigned slot_id, double *__restrict__ LA,
                                           double *__restrict__ LB)
{
    constexpr int NV = P::NV;
    constexpr int HALO = FUSED ? 4 : 2;
    const int nt = blockDim.x;
    const int t = threadIdx.x;

    constexpr int ds = DIR;
    constexpr int d1 = (DIR == 0) ? 1 : 0;
    constexpr int d2 = (DIR == 2) ? 1 : 2;

    int p, s, tm, tp;
    if (DIR == 0) {
        s = t % a.TS;
        p = t / a.TS;
        tm = t - 1;
        tp = t + 1;
    } else {
        p = t % a.TP;
        s = t / a.TP;
        tm = t - a.TP;
        tp = t + a.TP;
    }
    tm = tm < 0 ? 0 : tm;
    tp = tp >= nt ? nt - 1 : tp;

    const int chunk = bid % a.chunks;
    const int ptile = bid / a.chunks;
    const int n1 = a.n[d1];
    const int npen = n1 * a.n[d2];
    const int pf = ptile * a.TP + p;
    const bool pvalid = t < a.TP * a.TS && pf < npen;
    const int pfc = pf < npen ? pf : npen - 1;
    const int c1 = a.g[d1] + pfc % n1;
    const int c2 = a.g[d2] + pfc / n1;
    const int ng = a.ng[ds];
    const int os = (FUSED ? a.g[ds] : 0) + chunk * a.out;
    const int oe = FUSED ? a.g[ds] + a.n[ds] : ng;
    const int cs = os - HALO + s;
    const bool svalid = cs >= 0 && cs < ng;
    const int csc = cs < 0 ? 0 : (cs >= ng ? ng - 1 : cs);
    const bool outl = pvalid && s >= HALO && s < HALO + a.out && cs < oe;

    int xyz[3];
    xyz[ds] = csc;
    xyz[d1] = c1;
    xyz[d2] = c2;
    const long long idx = (long long)xyz[0] + (long long)a.sy * xyz[1] + (long long)a.sz * xyz[2];

    double W[NV], W0[NV], Wn[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) W[k] = a.src[(long long)plane_of<DIR>(k) * a.sv + idx];

    // Floor bookkeeping.  Flags: stage 1 over every cell whose stage-1 value
    // the tile completes within two cells of the interior (lanes 2..TS-3;
    // the reference's stage computes these cells with the same stencil), the
    // last stage over the cells the tile writes.  Mass: the interior cells,
    // each stage's events added exactly (pk_mass.hpp).
    unsigned flags = 0u;
    if constexpr (FUSED) {
        double Ws[NV];
        tile_stage<P, LIM>(a, W, W, a.c1, true, false, false, LA, LB, t, tm, tp, nt, Ws);
        double dm = 0.0;
        unsigned fl = 0u;
        floors<P>(Ws, a.smalld, a.smallei, a.use_smalld, a.use_smallei, a.ei_flag, fl, dm);
        const int g = a.g[ds];
        if (pvalid && s >= 2 && s <= a.TS - 3 && cs >= g - 2 && cs < g + a.n[ds] + 2) flags |= fl;
        pk::mass_add(a.dmass, outl ? dm : 0.0);
        tile_stage<P, LIM>(a, Ws, W, a.c2, true, false, false, LB, LA, t, tm, tp, nt, Wn);
    } else {
        const bool lo = cs == 0, hi = cs == ng - 1;
        if (outl && a.base != a.src) {
#pragma unroll
            for (int k = 0; k < NV; k++) W0[k] = a.base[(long long)plane_of<DIR>(k) * a.sv + idx];
        } else {
#pragma unroll
            for (int k = 0; k < NV; k++) W0[k] = W[k];
        }
        tile_stage<P, LIM, true>(a, W, W0, a.c1, !(lo || hi), lo, hi, LA, LB, t, tm, tp, nt, Wn, pvalid ? pf : -1,
                                 cs, outl && svalid);
    }
    {
        double dm = 0.0;
        unsigned fl = 0u;
        floors<P>(Wn, a.smalld, a.smallei, a.use_smalld, a.use_smallei, a.ei_flag, fl, dm);
        const bool inner = cs >= a.g[ds] && cs < a.g[ds] + a.n[ds];
        if (outl) flags |= fl;
        pk::mass_add(a.dmass, (outl && inner) ? dm : 0.0);
    }
    if constexpr (FUSED) {
        if (a.dtslots) dt_contribution<DIR, P>(a, Wn, outl && svalid, slot_id);
    }
    if constexpr (P::MAG) Wn[8] = Wn[8] * a.psi_fac;

    if (outl && svalid) {
#pragma unroll
        for (int k = 0; k < NV; k++) a.dst[(long long)plane_of<DIR>(k) * a.sv + idx] = Wn[k];
        if (FUSED && a.images) write_images<DIR, P>(a, xyz, idx, Wn);
    }
    // negative values are rare: one atomic per lane that saw one
    if (flags) atomicOr(a.flags, flags);
}

// ---------------------------------------------------------------- march ----
// The fused RK2 sweep with one lane per pencil.  A lane walks its pencil
// cell by cell and carries a register window instead of sharing neighbours
// through LDS: the primitives of the last two cells and the last face state
// and flux of each stage.  Reading cell t it completes the stage-1 flux at
// (t-2)+1/2, the stage-1 state of cell t-2, the stage-2 flux at (t-4)+1/2 and
// the new state of cell t-4.  No cell is computed twice, no barrier is needed,
// and for y and z sweeps consecutive lanes are consecutive x, so every load
// and store is coalesced.  Each cell sees the operands and operations of the
// tile kernel in the same order: the results are bit-identical.

template <class P, int LIM>
__device__ __forceinline__ void cell_faces(const SweepArgs &a, const double (&qm)[P::NV], const double (&q)[P::NV],
                                           const double (&qp)[P::NV], double (&fp)[P::NV], double (&fm)[P::NV])
{
#pragma unroll
    for (int v = 0; v < P::NV; v++) {
        const int lim = (LIM >= 0) ? LIM : ((v < 5) ? a.lim : a.limb);
        const double dl = q[v] - qm[v];
        const double dr = qp[v] - q[v];
        const double h = 0.5 * limit(lim, dl, dr);
        fp[v] = q[v] + h;
        fm[v] = q[v] - h;
    }
}

template <class P>
__device__ __forceinline__ void riemann(const SweepArgs &a, const double (&l)[P::NV], const double (&r)[P::NV],
                                        double (&F)[P::NV])
{
    if constexpr (P::MAG)
        flux_hlld(l, r, a.gam, a.ch, F);
    else
        flux_hllc(l, r, a.gam, F);
}

template <class P>
__device__ __forceinline__ void cell_update(const double (&W0)[P::NV], double coef, const double (&Fm)[P::NV],
                                            const double (&Fp)[P::NV], double (&Wn)[P::NV])
{
#pragma unroll
    for (int v = 0; v < P::NV; v++) {
        const double un = W0[v] + coef * (Fm[v] - Fp[v]);
        Wn[v] = (v < 5) ? un + 0.0 : un;  // internal_sources adds +0 to the fluid part
    }
}

template <int NV>
__device__ __forceinline__ void copy_nv(double (&d)[NV], const double (&s)[NV])
{
#pragma unroll
    for (int v = 0; v < NV; v++) d[v] = s[v];
}

template <int DIR, class P, int LIM>
__global__ __launch_bounds__(64) void k_march(const SweepArgs a)
{
    constexpr int NV = P::NV;
    constexpr int ds = DIR;
    constexpr int d1 = (DIR == 0) ? 1 : 0;
    constexpr int d2 = (DIR == 2) ? 1 : 2;
    const int n1 = a.n[d1];
    const int npen = n1 * a.n[d2];
    const int pf = blockIdx.x * blockDim.x + threadIdx.x;
    const bool pvalid = pf < npen;
    const int pfc = pvalid ? pf : npen - 1;
    int xyz[3];
    xyz[d1] = a.g[d1] + pfc % n1;
    xyz[d2] = a.g[d2] + pfc / n1;
    xyz[ds] = 0;
    const long long step = (DIR == 0) ? 1ll : ((DIR == 1) ? (long long)a.sy : (long long)a.sz);
    const long long base = (long long)xyz[0] + (long long)a.sy * xyz[1] + (long long)a.sz * xyz[2];
    const int g = a.g[ds], n = a.n[ds];
    const int t0 = g - 4, t1 = g + n + 4;  // cells read: [t0, t1)

    auto load = [&](int c, double (&W)[NV]) {
        c = c < t0 ? t0 : (c >= t1 ? t1 - 1 : c);
        const long long o = base + step * c;
#pragma unroll
        for (int k = 0; k < NV; k++) W[k] = a.src[(long long)plane_of<DIR>(k) * a.sv + o];
    };

    // the window (values before the pipeline fills are never stored)
    double qa[NV], qb[NV];    // primitives of cells t-2, t-1
    double fpp[NV], F1m[NV];  // stage 1: face state on the right of cell t-2; flux at (t-2)-1/2
    double sa[NV], sb[NV];    // stage-1 primitives of cells t-4, t-3
    double fps[NV], F2m[NV];  // stage 2: face state on the right of cell t-4; flux at (t-4)-1/2
#pragma unroll
    for (int v = 0; v < NV; v++) qa[v] = qb[v] = fpp[v] = F1m[v] = sa[v] = sb[v] = fps[v] = F2m[v] = 1.0;

    unsigned flags = 0u;
    const double huge = 1.7976931348623157e308;
    double mn[3] = {huge, huge, huge}, cmax = 0.0;
    double Wnext[NV];
    load(t0, Wnext);
    for (int t = t0; t < t1; t++) {
        double W[NV], W2[NV], W4[NV];
        copy_nv(W, Wnext);
        load(t + 1, Wnext);
        load(t - 2, W2);
        load(t - 4, W4);
        double q[NV];
        prim<P>(W, a.gam_1, q);
        // stage 1: face states of cell t-1, flux at (t-2)+1/2, state of cell t-2
        double fp[NV], fm[NV], F1[NV];
        cell_faces<P, LIM>(a, qa, qb, q, fp, fm);
        riemann<P>(a, fpp, fm, F1);
        double Ws[NV];
        cell_update<P>(W2, a.c1, F1m, F1, Ws);
        {
            // the stage-1 value of cell t-2 is complete from t = g on (cells
            // from g-2: the two cells next to the interior, as the tiles)
            unsigned fl = 0u;
            double dm = 0.0;
            floors<P>(Ws, a.smalld, a.smallei, a.use_smalld, a.use_smallei, a.ei_flag, fl, dm);
            const int c = t - 2;
            if (c >= g - 2) flags |= fl;
            pk::mass_add(a.dmass, (pvalid && c >= g && c < g + n) ? dm : 0.0);
        }
        double qs[NV];
        prim<P>(Ws, a.gam_1, qs);
        // stage 2: face states of cell t-3, flux at (t-4)+1/2, new state of cell t-4
        double gp[NV], gm[NV], F2[NV];
        cell_faces<P, LIM>(a, sa, sb, qs, gp, gm);
        riemann<P>(a, fps, gm, F2);
        double Wn[NV];
        cell_update<P>(W4, a.c2, F2m, F2, Wn);
        const int c = t - 4;
        {
            unsigned fl = 0u;
            double dm = 0.0;
            floors<P>(Wn, a.smalld, a.smallei, a.use_smalld, a.use_smallei, a.ei_flag, fl, dm);
            if (c >= g) flags |= fl;
            pk::mass_add(a.dmass, (pvalid && c >= g) ? dm : 0.0);
        }
        if (c >= g) {
            if (a.dtslots && pvalid) dt_cell<DIR, P>(a, Wn, mn, cmax);
            if constexpr (P::MAG) Wn[8] = Wn[8] * a.psi_fac;
            if (pvalid) {
                xyz[ds] = c;
                const long long idx = base + step * c;
#pragma unroll
                for (int k = 0; k < NV; k++) a.dst[(long long)plane_of<DIR>(k) * a.sv + idx] = Wn[k];
                if (a.images) write_images<DIR, P>(a, xyz, idx, Wn);
            }
        }
        // advance the window
        copy_nv(qa, qb);
        copy_nv(qb, q);
        copy_nv(fpp, fp);
        copy_nv(F1m, F1);
        copy_nv(sa, sb);
        copy_nv(sb, qs);
        copy_nv(fps, gp);
        copy_nv(F2m, F2);
    }
    if (a.dtslots) dt_reduce(a, mn, cmax, blockIdx.x);
    if (pvalid && flags) atomicOr(a.flags, flags);
}

// The kernel arguments of list item `item` (its block's table entry with
// the step's scalars and buffer parity applied); returns the block's first
// item.  The tables are read-only while a kernel runs: they are read through
// the constant address space, so the uniform loads go to scalar registers.
template <int DIR>
__device__ __forceinline__ unsigned load_list_args(const ListArgs &L, unsigned item, SweepArgs &a)
{
    typedef const __attribute__((address_space(4))) unsigned cu32;
    const int blk = __builtin_amdgcn_readfirstlane((int)((cu32 *)L.items)[item]);
    {
        static_assert(sizeof(SweepArgs) % 4 == 0, "SweepArgs is copied in dwords");
        cu32 *w = (cu32 *)(L.args + blk);
        unsigned *aw = reinterpret_cast<unsigned *>(&a);
#pragma unroll
        for (int i = 0; i < (int)(sizeof(SweepArgs) / 4); i++) aw[i] = w[i];
    }
    const unsigned first = ((cu32 *)L.first)[blk];
    if (L.no_wrap && a.images == 2) a.images = 1;
    if (L.parity) {
        const double *s0 = a.src;
        a.src = a.dst;
        a.dst = const_cast<double *>(s0);
    }
    a.base = a.src;
    typedef const __attribute__((address_space(4))) double cf64;
    const int inst = L.per_block ? (L.bpos ? (int)((cu32 *)L.bpos)[blk] : blk) : 0;
    const cf64 *stp = (const cf64 *)(L.step + inst);
    const double sdt = stp[0], sch = stp[1], sfac = stp[2];
    a.c1 = (0.5 * sdt) / a.dl[DIR];  // rk_coef * dt / dl, constants.F90:237-238
    a.c2 = (1.0 * sdt) / a.dl[DIR];
    a.ch = sch;
    a.psi_fac = L.apply_fac ? sfac : 1.0;
    a.dtslots = L.dtslots ? L.dtslots + 4ll * L.nslots * inst : nullptr;
    a.nslots = L.nslots;
    return first;
}

#include "pk_stream.hpp"

// The hydro sweep fits four waves per SIMD in 128 VGPRs (2 spilled; 5 % faster
// at 128^3 than three waves); the MHD sweep keeps its two (227-256 VGPRs).
template <int DIR, class P, bool FUSED, int LIM>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(P::MAG ? 1 : 4))) void k_sweep(const SweepArgs a)
{
    extern __shared__ double lds[];
    // XCD-aware order: workgroup i runs on XCD i % 8, so give each XCD a
    // contiguous range of tiles; tiles that share halo cells then share an L2.
    int bid = blockIdx.x;
    const int nfull = (int)(gridDim.x / 8u) * 8;
    if (bid < nfull) bid = (bid % 8) * (nfull / 8) + bid / 8;
    sweep_tile<DIR, P, FUSED, LIM>(a, bid, blockIdx.x, lds, lds + P::NV * blockDim.x);
}

// Persistent fused sweep over a block list.  Each workgroup takes the next
// `grab` tiles from the queue (one atomic per fetch), and for each loads
// that block's arguments and runs the tile; it leaves when the queue is
// empty, so every workgroup ends after at most nitems / grab + gridDim.x
// fetches in total.
template <int DIR, class P, int LIM>
__global__ __launch_bounds__(512) void k_sweep_list(const ListArgs L)
{
    extern __shared__ double lds[];
    __shared__ unsigned s_item;
    const unsigned grab = L.grab > 1u ? L.grab : 1u;
    unsigned item = 0u, end = 0u;
    for (;;) {
        __syncthreads();  // the previous tile is done with LDS and s_item
        if (item >= end) {
            if (threadIdx.x == 0) s_item = atomicAdd(L.queue, grab);
            __syncthreads();
            item = (unsigned)__builtin_amdgcn_readfirstlane((int)s_item);
            if (item >= L.nitems) break;
            end = min(item + grab, L.nitems);
        }
        SweepArgs a;
        const unsigned first = load_list_args<DIR>(L, item, a);
        sweep_tile<DIR, P, true, LIM>(a, (int)(item - first), item, lds, lds + P::NV * blockDim.x);
        item++;
    }
}

__global__ void k_fill_items(const unsigned *__restrict__ first, int nblocks, unsigned *__restrict__ items)
{
    // one workgroup per block writes its range of the item table
    const int b = blockIdx.x;
    if (b >= nblocks) return;
    for (unsigned i = first[b] + threadIdx.x; i < first[b + 1]; i += blockDim.x) items[i] = (unsigned)b;
}

// one workgroup per block: that block's dt slots -> its four results
__global__ __launch_bounds__(256) void k_dtslots_final_blocks(const unsigned long long *__restrict__ slots, int nslots,
                                                              double *__restrict__ out)
{
    __shared__ double red[4][256];
    const int t = threadIdx.x;
    const unsigned long long *s = slots + 4ll * nslots * blockIdx.x;
    double m0 = 1.7976931348623157e308, m1 = m0, m2 = m0, cmax = 0.0;
    for (int i = t; i < nslots; i += 256) {
        m0 = fmn(m0, __longlong_as_double((long long)s[4 * i + 0]));
        m1 = fmn(m1, __longlong_as_double((long long)s[4 * i + 1]));
        m2 = fmn(m2, __longlong_as_double((long long)s[4 * i + 2]));
        cmax = fmx(cmax, __longlong_as_double((long long)s[4 * i + 3]));
    }
    red[0][t] = m0;
    red[1][t] = m1;
    red[2][t] = m2;
    red[3][t] = cmax;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            red[0][t] = fmn(red[0][t], red[0][t + w]);
            red[1][t] = fmn(red[1][t], red[1][t + w]);
            red[2][t] = fmn(red[2][t], red[2][t + w]);
            red[3][t] = fmx(red[3][t], red[3][t + w]);
        }
        __syncthreads();
    }
    if (t == 0) {
        double *o = out + 4 * blockIdx.x;
        o[0] = red[0][0];
        o[1] = red[1][0];
        o[2] = red[2][0];
        o[3] = red[3][0];
    }
}

// a batch of device-to-device copies: workgroups stride over the jobs
__global__ __launch_bounds__(256) void k_copy_blocks(const pk::CopyJob *__restrict__ jobs, int njobs)
{
    for (int j = blockIdx.y; j < njobs; j += gridDim.y) {
        const pk::CopyJob c = jobs[j];
        for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < c.n;
             i += (long long)gridDim.x * blockDim.x)
            c.dst[i] = c.src[i];
    }
}

// Ghost exchange between blocks (internal_boundaries_local,
// cg_list_bnd.F90:269-322): one job per face, cut into chunks of kFaceChunk
// values, one workgroup per chunk (wgjob names its job); values run x
// fastest, so rows of y and z faces are coalesced.
__global__ __launch_bounds__(256) void k_exchange_faces(const pk::FaceJob *__restrict__ jobs,
                                                      const int *__restrict__ wgjob)
{
    const pk::FaceJob &J = jobs[wgjob[blockIdx.x]];
    const int e0 = J.ext[0], e01 = e0 * J.ext[1], ncell = e01 * J.ext[2];
    const int total = ncell * J.nplanes;
    const int base = ((int)blockIdx.x - J.wg0) * pk::kFaceChunk;
    const int end = min(base + pk::kFaceChunk, total);
    for (int e = base + (int)threadIdx.x; e < end; e += 256) {
        const int v = e / ncell, c = e - v * ncell;
        const int k = c / e01, r = c - k * e01;
        const int j = r / e0, i = r - j * e0;
        const long long s = v * J.ssv + (J.sorg[0] + i) + (long long)J.ssy * (J.sorg[1] + j) +
                            (long long)J.ssz * (J.sorg[2] + k);
        const long long d = v * J.dsv + (J.dorg[0] + i) + (long long)J.dsy * (J.dorg[1] + j) +
                            (long long)J.dsz * (J.dorg[2] + k);
        J.dst[d] = J.src[s];
    }
}

__global__ void k_zero_u32(unsigned *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Periodic refresh of the ghost layers along direction DIR over the full
// extent of the other two directions (run x, then y, then z for corners).
template <int DIR>
__global__ void k_fill_ghosts(const GhostArgs a)
{
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    constexpr int d1 = (DIR == 0) ? 1 : 0;
    constexpr int d2 = (DIR == 2) ? 1 : 2;
    const int g = a.g[DIR];
    const long long per_plane = (long long)(2 * g) * a.ng[d1] * a.ng[d2];
    if (tid >= per_plane * a.nplanes) return;
    const int v = (int)(tid / per_plane);
    long long r = tid % per_plane;
    // fastest index: x when DIR != 0, else the ghost layer index
    int xyz[3];
    int gi;
    if (DIR == 0) {
        gi = (int)(r % (2 * g));
        r /= (2 * g);
        xyz[d1] = (int)(r % a.ng[d1]);
        xyz[d2] = (int)(r / a.ng[d1]);
    } else {
        xyz[d1] = (int)(r % a.ng[d1]);
        r /= a.ng[d1];
        gi = (int)(r % (2 * g));
        xyz[d2] = (int)(r / (2 * g));
    }
    const int n = a.n[DIR];
    const int dstc = (gi < g) ? gi : n + gi;          // low ghosts 0..g-1, high ghosts g+n..
    const int srcc = (gi < g) ? gi + n : gi;          // periodic images inside the interior
    int s3[3] = {xyz[0], xyz[1], xyz[2]};
    xyz[DIR] = dstc;
    s3[DIR] = srcc;
    const long long di = (long long)xyz[0] + (long long)a.sy * xyz[1] + (long long)a.sz * xyz[2];
    const long long si = (long long)s3[0] + (long long)a.sy * s3[1] + (long long)a.sz * s3[2];
    double *base = a.u + (long long)(a.plane0 + v) * a.sv;
    base[di] = base[si];
}

__global__ void k_scale(double *__restrict__ x, long long n, double fac)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = x[i] * fac;
}

// timestep_fluid (timestep.F90:460-490) with ion_cs / neu_cs
// (initionized.F90:137-141, initneutral.F90:102-105): per-workgroup partial
// min of dl_d/c_d for d = x, y, z and max of c.
template <bool MAG>
__global__ __launch_bounds__(256) void k_dt_partial(const DtArgs a, double *__restrict__ part)
{
    __shared__ double red[4][256];
    const int t = threadIdx.x;
    const long long ncell = (long long)a.n[0] * a.n[1] * a.n[2];
    double m0 = 1.7976931348623157e308, m1 = m0, m2 = m0, cmax = 1.17549435082228750797e-38;
    for (long long c = (long long)blockIdx.x * blockDim.x + t; c < ncell; c += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(c % a.n[0]);
        const long long r = c / a.n[0];
        const int j = (int)(r % a.n[1]);
        const int k = (int)(r / a.n[1]);
        const long long idx = (long long)(i + a.g[0]) + (long long)a.sy * (j + a.g[1]) + (long long)a.sz * (k + a.g[2]);
        const double *u = a.u + (long long)a.plane0 * a.sv + idx;
        const double d = u[0], mx = u[a.sv], my = u[2 * a.sv], mz = u[3 * a.sv], e = u[4 * a.sv];
        double v0 = 0.0, v1 = 0.0, v2 = 0.0;
        if (d > 0.0) {
            v0 = fabs(qdiv(mx, d));
            v1 = fabs(qdiv(my, d));
            v2 = fabs(qdiv(mz, d));
        }
        const double ek = qdiv(0.5 * ((mx * mx + my * my) + mz * mz), d);
        double cs;
        if (MAG) {
            const double *b = a.u + (long long)a.bplane * a.sv + idx;
            const double bx = b[0], by = b[a.sv], bz = b[2 * a.sv];
            const double pmag = 0.5 * ((bx * bx + by * by) + bz * bz);
            const double ps = (e - ek) * a.gam_1 + (2.0 - a.gam) * pmag;
            const double pr = ps - pmag;
            cs = qsqrt(fabs(qdiv(2.0 * pmag + a.gam * pr, d)));
        } else {
            const double pr = (e - ek) * a.gam_1;
            cs = qsqrt(fabs(qdiv(a.gam * pr, d)));
        }
        const double small = 1.17549435082228750797e-38;
        const double c0 = fmx(v0 + cs, small), c1 = fmx(v1 + cs, small), c2 = fmx(v2 + cs, small);
        cmax = fmx(cmax, fmx(fmx(c0, c1), c2));
        if (a.has_dir[0]) m0 = fmn(m0, qdiv(a.dl[0], c0));
        if (a.has_dir[1]) m1 = fmn(m1, qdiv(a.dl[1], c1));
        if (a.has_dir[2]) m2 = fmn(m2, qdiv(a.dl[2], c2));
    }
    red[0][t] = m0;
    red[1][t] = m1;
    red[2][t] = m2;
    red[3][t] = cmax;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            red[0][t] = fmn(red[0][t], red[0][t + w]);
            red[1][t] = fmn(red[1][t], red[1][t + w]);
            red[2][t] = fmn(red[2][t], red[2][t + w]);
            red[3][t] = fmx(red[3][t], red[3][t + w]);
        }
        __syncthreads();
    }
    if (t == 0) {
        part[4 * blockIdx.x + 0] = red[0][0];
        part[4 * blockIdx.x + 1] = red[1][0];
        part[4 * blockIdx.x + 2] = red[2][0];
        part[4 * blockIdx.x + 3] = red[3][0];
    }
}

__global__ __launch_bounds__(256) void k_dt_final(const double *__restrict__ part, int nparts,
                                                  double *__restrict__ out)
{
    __shared__ double red[4][256];
    const int t = threadIdx.x;
    double m0 = 1.7976931348623157e308, m1 = m0, m2 = m0, cmax = 0.0;
    for (int i = t; i < nparts; i += 256) {
        m0 = fmn(m0, part[4 * i + 0]);
        m1 = fmn(m1, part[4 * i + 1]);
        m2 = fmn(m2, part[4 * i + 2]);
        cmax = fmx(cmax, part[4 * i + 3]);
    }
    red[0][t] = m0;
    red[1][t] = m1;
    red[2][t] = m2;
    red[3][t] = cmax;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            red[0][t] = fmn(red[0][t], red[0][t + w]);
            red[1][t] = fmn(red[1][t], red[1][t + w]);
            red[2][t] = fmn(red[2][t], red[2][t + w]);
            red[3][t] = fmx(red[3][t], red[3][t + w]);
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = red[0][0];
        out[1] = red[1][0];
        out[2] = red[2][0];
        out[3] = red[3][0];
    }
}

// Fortran AoS block (nc, nxg, nyg, nzg) <-> device planes m.planes[c].
__global__ void k_aos_to_soa(const double *__restrict__ aos, const pk::PlaneMap m, double *__restrict__ soa,
                             long long sv, int nx, int ny, int nz, int sy, int sz)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long ncell = (long long)nx * ny * nz;
    if (i >= ncell) return;
    const int x = (int)(i % nx);
    const long long r = i / nx;
    const int y = (int)(r % ny);
    const int z = (int)(r / ny);
    const long long di = (long long)x + (long long)sy * y + (long long)sz * z;
    for (int c = 0; c < m.nc; c++) soa[(long long)m.planes[c] * sv + di] = aos[i * m.nc + c];
}

__global__ void k_soa_to_aos(const double *__restrict__ soa, long long sv, const pk::PlaneMap m,
                             double *__restrict__ aos, int nx, int ny, int nz, int sy, int sz)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long ncell = (long long)nx * ny * nz;
    if (i >= ncell) return;
    const int x = (int)(i % nx);
    const long long r = i / nx;
    const int y = (int)(r % ny);
    const int z = (int)(r / ny);
    const long long si = (long long)x + (long long)sy * y + (long long)sz * z;
    for (int c = 0; c < m.nc; c++) aos[i * m.nc + c] = soa[(long long)m.planes[c] * sv + si];
}

__global__ void k_dtslots_init(unsigned long long *slots, int nslots)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nslots) return;
    const unsigned long long huge = (unsigned long long)__double_as_longlong(1.7976931348623157e308);
    slots[4 * i + 0] = huge;
    slots[4 * i + 1] = huge;
    slots[4 * i + 2] = huge;
    slots[4 * i + 3] = 0ull;
}

__global__ __launch_bounds__(256) void k_dtslots_final(const unsigned long long *__restrict__ slots, int nslots,
                                                       double *__restrict__ out)
{
    __shared__ double red[4][256];
    const int t = threadIdx.x;
    double m0 = 1.7976931348623157e308, m1 = m0, m2 = m0, cmax = 0.0;
    for (int i = t; i < nslots; i += 256) {
        m0 = fmn(m0, __longlong_as_double((long long)slots[4 * i + 0]));
        m1 = fmn(m1, __longlong_as_double((long long)slots[4 * i + 1]));
        m2 = fmn(m2, __longlong_as_double((long long)slots[4 * i + 2]));
        cmax = fmx(cmax, __longlong_as_double((long long)slots[4 * i + 3]));
    }
    red[0][t] = m0;
    red[1][t] = m1;
    red[2][t] = m2;
    red[3][t] = cmax;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            red[0][t] = fmn(red[0][t], red[0][t + w]);
            red[1][t] = fmn(red[1][t], red[1][t + w]);
            red[2][t] = fmn(red[2][t], red[2][t + w]);
            red[3][t] = fmx(red[3][t], red[3][t + w]);
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = red[0][0];
        out[1] = red[1][0];
        out[2] = red[2][0];
        out[3] = red[3][0];
    }
}

hipError_t launch_dtslots_init(unsigned long long *slots, int nslots, hipStream_t st)
{
    hipLaunchKernelGGL(k_dtslots_init, dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0, st, slots, nslots);
    return hipGetLastError();
}

hipError_t launch_dtslots_final(const unsigned long long *slots, int nslots, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_dtslots_final, dim3(1), dim3(256), 0, st, slots, nslots, out);
    return hipGetLastError();
}

// ------------------------------------------------------------ launchers ----

template <int DIR, class P, bool FUSED, int LIM>
static hipError_t launch_sweep_t(const SweepArgs &a, hipStream_t st)
{
    const int nt = a.TP * a.TS;
    const size_t lds = sizeof(double) * 2 * P::NV * nt;
    static bool lds_raised = false;  // allow dynamic LDS beyond 64 KiB (gfx950 has 160 KiB per CU)
    if (!lds_raised) {
        hipError_t e = hipFuncSetAttribute((const void *)k_sweep<DIR, P, FUSED, LIM>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        lds_raised = true;
    }
    if (lds > 160 * 1024 || nt > 512) return hipErrorInvalidValue;
    const long long nblk = (long long)a.chunks * a.ptiles;
    hipLaunchKernelGGL((k_sweep<DIR, P, FUSED, LIM>), dim3((unsigned)nblk), dim3(nt), lds, st, a);
    return hipGetLastError();
}

template <int DIR, class P, int LIM>
static hipError_t launch_march_t(const SweepArgs &a, hipStream_t st)
{
    const int d1 = (DIR == 0) ? 1 : 0, d2 = (DIR == 2) ? 1 : 2;
    const long long npen = (long long)a.n[d1] * a.n[d2];
    const long long nblk = (npen + 63) / 64;
    if (a.g[DIR] < 4 || nblk < 1 || nblk > 0x7fffffffll) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_march<DIR, P, LIM>), dim3((unsigned)nblk), dim3(64), 0, st, a);
    return hipGetLastError();
}

constexpr int kMaxDevices = 64;

// Workgroups of `fn` (one wave each) resident on the current device at
// once, 0 if unknown; cached per device.
static int stream_slots(const void *fn, int *cache)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return 0;
    if (cache[dev] == 0) {
        int per = 0, ncu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, 64, 0) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
            (void)hipGetLastError();  // launch without rounding; leave no error for the launch check
            return 0;
        }
        cache[dev] = per * ncu;
    }
    return cache[dev];
}

// The stream kernel: a.TP pencils per lane group, a.ptiles ranges of whole
// pencil groups (a.out positions each, the last one shorter) asked for.  A
// launch of more ranges than the device holds at once runs in rounds; it is
// rounded to a whole number of rounds, the ranges made equal to within one
// group, so that the last round does not leave the device partly idle.
// IMG: the kernel that writes periodic ghost images (a.images == 1).
template <int DIR, class P, int LIM, int TP, bool IMG>
static hipError_t launch_stream_tp(const SweepArgs &a, hipStream_t st)
{
    constexpr int d1 = (DIR == 0) ? 1 : 0, d2 = (DIR == 2) ? 1 : 2;
    const long long npen = (long long)a.n[d1] * a.n[d2];
    const long long L = a.n[DIR] + 8;
    const long long ngroups = (npen + TP - 1) / TP;
    const long long npos = ngroups * L;
    const int span = a.out;
    if (a.g[DIR] < 4 || span <= 0 || span % L != 0 || npos + span + 64 > 0x7fffffffll || !pk::stream_stores_fit(a.sv))
        return hipErrorInvalidValue;
    long long nblk = (npos + span - 1) / span;
    if (nblk != a.ptiles) return hipErrorInvalidValue;
    const void *fn = (const void *)k_stream<DIR, P, LIM, TP, IMG>;
    static int cache[kMaxDevices] = {};
    const long long slots = stream_slots(fn, cache);
    if (slots > 0 && nblk >= slots) {
        // nblk <= ngroups, so one round of slots ranges has a group each
        long long r = (nblk + slots / 2) / slots;
        while (r > 1 && r * slots > ngroups) r--;
        nblk = r * slots;
    }
    hipLaunchKernelGGL((k_stream<DIR, P, LIM, TP, IMG>), dim3((unsigned)nblk), dim3(64), 0, st, a, (int)ngroups);
    return hipGetLastError();
}

template <int DIR, class P, int LIM>
static hipError_t launch_stream_t(const SweepArgs &a, hipStream_t st)
{
    const bool img = a.images == 1;
    switch (a.TP) {
    case 1: return img ? launch_stream_tp<DIR, P, LIM, 1, true>(a, st) : launch_stream_tp<DIR, P, LIM, 1, false>(a, st);
    case 2: return img ? launch_stream_tp<DIR, P, LIM, 2, true>(a, st) : launch_stream_tp<DIR, P, LIM, 2, false>(a, st);
    case 4: return img ? launch_stream_tp<DIR, P, LIM, 4, true>(a, st) : launch_stream_tp<DIR, P, LIM, 4, false>(a, st);
    case 8: return img ? launch_stream_tp<DIR, P, LIM, 8, true>(a, st) : launch_stream_tp<DIR, P, LIM, 8, false>(a, st);
    default: return hipErrorInvalidValue;
    }
}

template <int DIR, class P, bool FUSED>
static hipError_t launch_lim(const SweepArgs &a, hipStream_t st)
{
    if constexpr (FUSED) {
        if (a.TS <= -2) {  // the stream kernel (k_stream)
            if (a.lim == a.limb) {
                switch (a.lim) {
                case LIM_VANLEER: return launch_stream_t<DIR, P, LIM_VANLEER>(a, st);
                case LIM_MINMOD: return launch_stream_t<DIR, P, LIM_MINMOD>(a, st);
                case LIM_MONCEN: return launch_stream_t<DIR, P, LIM_MONCEN>(a, st);
                case LIM_SUPERBEE: return launch_stream_t<DIR, P, LIM_SUPERBEE>(a, st);
                default: break;
                }
            }
            return launch_stream_t<DIR, P, -1>(a, st);
        }
        if (a.TS < 0) {  // whole pencils, one lane each (k_march)
            if (a.lim == a.limb) {
                switch (a.lim) {
                case LIM_VANLEER: return launch_march_t<DIR, P, LIM_VANLEER>(a, st);
                case LIM_MINMOD: return launch_march_t<DIR, P, LIM_MINMOD>(a, st);
                case LIM_MONCEN: return launch_march_t<DIR, P, LIM_MONCEN>(a, st);
                case LIM_SUPERBEE: return launch_march_t<DIR, P, LIM_SUPERBEE>(a, st);
                default: break;
                }
            }
            return launch_march_t<DIR, P, -1>(a, st);
        }
    }
    if (a.lim == a.limb) {
        switch (a.lim) {
        case LIM_VANLEER: return launch_sweep_t<DIR, P, FUSED, LIM_VANLEER>(a, st);
        case LIM_MINMOD: return launch_sweep_t<DIR, P, FUSED, LIM_MINMOD>(a, st);
        case LIM_MONCEN: return launch_sweep_t<DIR, P, FUSED, LIM_MONCEN>(a, st);
        case LIM_SUPERBEE: return launch_sweep_t<DIR, P, FUSED, LIM_SUPERBEE>(a, st);
        default: break;
        }
    }
    return launch_sweep_t<DIR, P, FUSED, -1>(a, st);
}

hipError_t launch_sweep(int dir, bool mhd, bool fused, const SweepArgs &a, hipStream_t st)
{
    // reading ghost positions from their interior pre-images is the stream kernel's only
    if (a.images == 2 && !(fused && a.TS <= -2)) return hipErrorInvalidValue;
#define PK_CASE(D, PP, F) \
    if (dir == D && (mhd == PP::MAG) && fused == F) return launch_lim<D, PP, F>(a, st);
    PK_CASE(0, HD, true) PK_CASE(1, HD, true) PK_CASE(2, HD, true)
    PK_CASE(0, HD, false) PK_CASE(1, HD, false) PK_CASE(2, HD, false)
    PK_CASE(0, MHD, true) PK_CASE(1, MHD, true) PK_CASE(2, MHD, true)
    PK_CASE(0, MHD, false) PK_CASE(1, MHD, false) PK_CASE(2, MHD, false)
#undef PK_CASE
    return hipErrorInvalidValue;
}

// Persistent list sweep: launch (resident == nullptr) or report how many
// workgroups of nt lanes the device holds at once.
// Persistent stream over a list (nt == 64: k_stream_list, one pencil per
// lane group along x, eight along y and z).
template <int DIR, class P, int LIM>
static hipError_t stream_list_t(const ListArgs *L, int grid, hipStream_t st, int *resident)
{
    constexpr int TP = pk::stream_list_tp(DIR);
    const void *fn = (const void *)k_stream_list<DIR, P, LIM, TP>;
    if (resident) {
        int per = 0, dev = 0, ncu = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, 64, 0);
        if (e == hipSuccess) e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        *resident = per * ncu;
        return e;
    }
    if (grid <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_stream_list<DIR, P, LIM, TP>), dim3((unsigned)grid), dim3(64), 0, st, *L);
    return hipGetLastError();
}

template <int DIR, class P, int LIM>
static hipError_t list_t(const ListArgs *L, int nt, int grid, hipStream_t st, int *resident)
{
    if (nt == pk::kStreamListNt) return stream_list_t<DIR, P, LIM>(L, grid, st, resident);
    const void *fn = (const void *)k_sweep_list<DIR, P, LIM>;
    const size_t lds = sizeof(double) * 2 * P::NV * nt;
    const size_t lds_max = 160 * 1024 - 256;  // static LDS (the queue slot) comes out of the same 160 KiB
    static bool lds_raised = false;
    if (!lds_raised) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_max);
        if (e != hipSuccess) return e;
        lds_raised = true;
    }
    if (nt < 64 || nt > 512 || nt % 64 != 0 || lds > lds_max) return hipErrorInvalidValue;
    if (resident) {
        int per = 0, dev = 0, ncu = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, nt, lds);
        if (e == hipSuccess) e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        *resident = per * ncu;
        return e;
    }
    if (grid <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_sweep_list<DIR, P, LIM>), dim3((unsigned)grid), dim3(nt), lds, st, *L);
    return hipGetLastError();
}

template <int DIR, class P>
static hipError_t list_lim(int lim, int limb, const ListArgs *L, int nt, int grid, hipStream_t st, int *resident)
{
    if (lim == limb) {
        switch (lim) {
        case LIM_VANLEER: return list_t<DIR, P, LIM_VANLEER>(L, nt, grid, st, resident);
        case LIM_MINMOD: return list_t<DIR, P, LIM_MINMOD>(L, nt, grid, st, resident);
        case LIM_MONCEN: return list_t<DIR, P, LIM_MONCEN>(L, nt, grid, st, resident);
        case LIM_SUPERBEE: return list_t<DIR, P, LIM_SUPERBEE>(L, nt, grid, st, resident);
        default: break;
        }
    }
    return list_t<DIR, P, -1>(L, nt, grid, st, resident);
}

static hipError_t list_dispatch(int dir, bool mhd, int lim, int limb, const ListArgs *L, int nt, int grid,
                                hipStream_t st, int *resident)
{
    if (mhd) {
        if (dir == 0) return list_lim<0, MHD>(lim, limb, L, nt, grid, st, resident);
        if (dir == 1) return list_lim<1, MHD>(lim, limb, L, nt, grid, st, resident);
        if (dir == 2) return list_lim<2, MHD>(lim, limb, L, nt, grid, st, resident);
    } else {
        if (dir == 0) return list_lim<0, HD>(lim, limb, L, nt, grid, st, resident);
        if (dir == 1) return list_lim<1, HD>(lim, limb, L, nt, grid, st, resident);
        if (dir == 2) return list_lim<2, HD>(lim, limb, L, nt, grid, st, resident);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_sweep_list(int dir, bool mhd, int lim, int limb, const ListArgs &L, int nt, int grid,
                             hipStream_t st)
{
    if (grid <= 0) {
        int r = 0;
        hipError_t e = list_dispatch(dir, mhd, lim, limb, &L, nt, 0, st, &r);
        if (e != hipSuccess) return e;
        grid = r > 0 ? r : 1;
    }
    return list_dispatch(dir, mhd, lim, limb, &L, nt, grid, st, nullptr);
}

int sweep_list_resident(int dir, bool mhd, int lim, int limb, int nt)
{
    int r = 0;
    if (list_dispatch(dir, mhd, lim, limb, nullptr, nt, 0, nullptr, &r) != hipSuccess) return 0;
    return r;
}

size_t sweep_lds_bytes(bool mhd, int TP, int TS) { return sizeof(double) * 2 * (mhd ? 9 : 5) * TP * TS; }

hipError_t launch_fill_ghosts(int dir, const GhostArgs &a, hipStream_t st)
{
    const int d1 = (dir == 0) ? 1 : 0, d2 = (dir == 2) ? 1 : 2;
    const long long total = (long long)(2 * a.g[dir]) * a.ng[d1] * a.ng[d2] * a.nplanes;
    if (total == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((total + 255) / 256);
    if (dir == 0) hipLaunchKernelGGL(k_fill_ghosts<0>, dim3(blocks), dim3(256), 0, st, a);
    if (dir == 1) hipLaunchKernelGGL(k_fill_ghosts<1>, dim3(blocks), dim3(256), 0, st, a);
    if (dir == 2) hipLaunchKernelGGL(k_fill_ghosts<2>, dim3(blocks), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_scale(double *x, long long n, double fac, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scale, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, n, fac);
    return hipGetLastError();
}

int dt_partial_blocks(const DtArgs &a)
{
    const long long ncell = (long long)a.n[0] * a.n[1] * a.n[2];
    long long b = (ncell + 255) / 256;
    if (b > 2048) b = 2048;
    if (b < 1) b = 1;
    return (int)b;
}

hipError_t launch_dt(const DtArgs &a, double *part, double *out, hipStream_t st)
{
    const int nb = PK_KNS::dt_partial_blocks(a);
    if (a.magnetized)
        hipLaunchKernelGGL(k_dt_partial<true>, dim3(nb), dim3(256), 0, st, a, part);
    else
        hipLaunchKernelGGL(k_dt_partial<false>, dim3(nb), dim3(256), 0, st, a, part);
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(256), 0, st, part, nb, out);
    return hipGetLastError();
}

hipError_t launch_aos_to_soa(const double *aos, const pk::PlaneMap &m, double *soa, long long sv, const int ng[3],
                             int sy, int sz, hipStream_t st)
{
    const long long n = (long long)ng[0] * ng[1] * ng[2];
    hipLaunchKernelGGL(k_aos_to_soa, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, aos, m, soa, sv, ng[0],
                       ng[1], ng[2], sy, sz);
    return hipGetLastError();
}

hipError_t launch_soa_to_aos(const double *soa, long long sv, const pk::PlaneMap &m, double *aos, const int ng[3],
                             int sy, int sz, hipStream_t st)
{
    const long long n = (long long)ng[0] * ng[1] * ng[2];
    hipLaunchKernelGGL(k_soa_to_aos, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, soa, sv, m, aos, ng[0],
                       ng[1], ng[2], sy, sz);
    return hipGetLastError();
}

hipError_t launch_fill_items(const unsigned *first, int nblocks, unsigned *items, hipStream_t st)
{
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_items, dim3((unsigned)nblocks), dim3(256), 0, st, first, nblocks, items);
    return hipGetLastError();
}

hipError_t launch_dtslots_final_blocks(const unsigned long long *slots, int nslots, int nblocks, double *out,
                                       hipStream_t st)
{
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dtslots_final_blocks, dim3((unsigned)nblocks), dim3(256), 0, st, slots, nslots, out);
    return hipGetLastError();
}

hipError_t launch_exchange_faces(const pk::FaceJob *jobs, const int *wgjob, int nwg, hipStream_t st)
{
    if (nwg <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_exchange_faces, dim3((unsigned)nwg), dim3(256), 0, st, jobs, wgjob);
    return hipGetLastError();
}

hipError_t launch_copy_blocks(const pk::CopyJob *jobs, int njobs, hipStream_t st)
{
    if (njobs <= 0) return hipSuccess;
    const unsigned gy = (unsigned)(njobs < 65535 ? njobs : 65535);
    hipLaunchKernelGGL(k_copy_blocks, dim3(256, gy), dim3(256), 0, st, jobs, njobs);
    return hipGetLastError();
}

hipError_t launch_zero_u32(unsigned *p, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

}  // namespace PK_KNS
