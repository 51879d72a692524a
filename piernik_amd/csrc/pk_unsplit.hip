// pk_unsplit.hip -- gfx950 kernels of the unsplit Riemann stage
// (solve_cg_ub / solve_cg_u of src/scheme/unsplit_schemes/, one RK stage).
//
// k_unsplit_flux_tile: the interface fluxes of the transverse-interior
// pencils along one direction, from tiles of TP pencils x TS cells.  A lane
// computes the primitives (sweep order) and limited slopes of its cell once,
// the face states go through LDS, and the lane of cell c evaluates HLLD + GLM
// or HLLC at c+1/2, exactly as interpol and riemann_wrap do for that
// interface (unsplit_mag_modules.F90:166-209, solve_cg_unsplit.F90:157-187).
// Lanes run x fastest, so y and z loads are coalesced.
//
// k_unsplit_update: one lane per interior cell.  The flux differences of
// x, then y, then z are added to the base state (apply_flux and update_psi,
// unsplit_mag_modules.F90:211-328), then apply_source's passes run over the
// directions in turn (unsplit_source.F90:42-132): +0 from the vanishing
// internal sources and, at the last stage, care_for_positives in that
// direction's sweep order.  Ghost cells are not written: update_boundaries
// replaces them after every stage.
//
// k_unsplit_fused_tile: both in one pass per direction, for stages whose
// output is not their input (the two RK2 stages): the flux tile adds each
// cell's flux difference to a running sum kept in the output buffer, and the
// last direction's pass runs the sources.  Same operands in the same order,
// so the same bits, without writing and re-reading the flux arrays.
//
// Built without FMA contraction (namespace pk): bit-identical to the
// oracle's unsplit driver.  pk_unsplit_fast.hip builds the same source with
// contraction and fast division into namespace pk_fast.
#include <hip/hip_runtime.h>
#include "pk_math.hpp"
#include "pk_mass.hpp"
#include "pk_unsplit.hpp"

namespace PK_KNS {

using pk::UnsplitArgs;

namespace {

// sweep-order variable k -> xyz plane (fluidtypes.F90:318-320, fluidindex.F90:190-192)
__device__ __forceinline__ int uplane(int dir, int k)
{
    if (dir == 0 || k == 0 || k == 4 || k == 8) return k;
    if (dir == 1) return (k == 1 || k == 2) ? 3 - k : ((k == 5 || k == 6) ? 11 - k : k);
    return (k == 1 || k == 3) ? 4 - k : ((k == 5 || k == 7) ? 12 - k : k);
}

// Tile of TP pencils x TS cells along DIR, lanes x fastest.  The lane of
// cell c evaluates the Riemann flux at c+1/2 from its right face state and
// the left face state of cell c+1: l = q(c) + 0.5*limit(q(c)-q(c-1),
// q(c+1)-q(c)) and r = q(c+1) - 0.5*limit(...), the operands and operations
// of the reference's per-interface formula.  The tile starts one cell before
// the left cell of its first interface; TS-3 fluxes per tile, row c+1 of F
// holding the flux at c+1/2.
template <int DIR, class P, int LIM>
__global__ __launch_bounds__(256) void k_unsplit_flux_tile(const UnsplitArgs a, double *__restrict__ F, int TP,
                                                           int TS, int chunks)
{
    constexpr int NV = P::NV;
    constexpr int d1 = (DIR == 0) ? 1 : 0;
    constexpr int d2 = (DIR == 2) ? 1 : 2;
    extern __shared__ double lds[];
    double *A = lds;                   // primitives
    double *B = lds + NV * blockDim.x; // left face states
    const int nt = blockDim.x;
    const int t = threadIdx.x;
    // XCD-aware order: each XCD gets a contiguous range of tiles
    int bid = blockIdx.x;
    const int nfull = (int)(gridDim.x / 8u) * 8;
    if (bid < nfull) bid = (bid % 8) * (nfull / 8) + bid / 8;
    int p, sl, tm, tp;
    if (DIR == 0) {
        sl = t % TS;
        p = t / TS;
        tm = t - 1;
        tp = t + 1;
    } else {
        p = t % TP;
        sl = t / TP;
        tm = t - TP;
        tp = t + TP;
    }
    tm = tm < 0 ? 0 : tm;
    tp = tp >= nt ? nt - 1 : tp;
    const int chunk = bid % chunks;
    const int ptile = bid / chunks;
    const int n1 = a.n[d1];
    const int npen = n1 * a.n[d2];
    const int pf = ptile * TP + p;
    const bool pvalid = t < TP * TS && pf < npen;
    const int pfc = pf < npen ? pf : npen - 1;
    const int out = TS - 3;
    const int c0 = a.g[DIR] - 1 + chunk * out;  // left cell of the tile's first interface
    const int cs = c0 - 1 + sl;                 // this lane's cell (the tile starts at c0 - 1)
    const int ng = a.ng[DIR];
    const int csc = cs < 0 ? 0 : (cs >= ng ? ng - 1 : cs);
    int xyz[3];
    xyz[DIR] = csc;
    xyz[d1] = a.g[d1] + pfc % n1;
    xyz[d2] = a.g[d2] + pfc / n1;
    const long long idx = (long long)xyz[0] + (long long)a.sy * xyz[1] + (long long)a.sz * xyz[2];

    double W[NV], q[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) W[k] = a.src[(long long)uplane(DIR, k) * a.sv + idx];
    prim<P>(W, a.gam_1, q);
#pragma unroll
    for (int v = 0; v < NV; v++) A[v * nt + t] = q[v];
    __syncthreads();
    double fp[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        const int lim = (LIM >= 0) ? LIM : ((v < 5) ? a.lim : a.limb);
        const double h = 0.5 * limit(lim, q[v] - A[v * nt + tm], A[v * nt + tp] - q[v]);
        fp[v] = q[v] + h;          // right face of this cell
        B[v * nt + t] = q[v] - h;  // left face of this cell
    }
    __syncthreads();
    // the flux at cs+1/2, row cs+1, for interfaces g-1/2 .. g+n-1/2
    const bool own = pvalid && sl >= 1 && sl < 1 + out && cs >= a.g[DIR] - 1 && cs <= a.g[DIR] + a.n[DIR] - 1;
    if (!own) return;
    double fr[NV], f[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) fr[v] = B[v * nt + tp];
    if constexpr (P::MAG)
        flux_hlld(fp, fr, a.gam, a.ch, f);
    else
        flux_hllc(fp, fr, a.gam, f);
    const long long st = (DIR == 0) ? 1ll : ((DIR == 1) ? (long long)a.sy : (long long)a.sz);
    const long long row = idx + st;
#pragma unroll
    for (int k = 0; k < NV; k++) F[(long long)uplane(DIR, k) * a.sv + row] = f[k];
}

// Flux along DIR and its difference in one tile (dst != src).  The tile
// starts two cells before its first updated cell (TS-4 per tile); lanes 1 ..
// TS-3 evaluate the flux at their cell's right face as k_unsplit_flux_tile
// does, and the lane of cell c adds coef*(F(c-1/2) - F(c+1/2)), the left
// flux from its neighbour through LDS, to the cell's running sum.  The sum
// starts from base at the first active direction and waits in dst between
// directions, each cell read and written by its own lane only; after the
// last active direction the source passes run, dst holds the stage's
// result and the pass also writes each cell's periodic ghost images (the
// stage's ghost refresh).  The operands and their order are those of
// k_unsplit_update, so the result is the same bit for bit, without the flux
// arrays.
// Three waves per SIMD: the relaxed build needs 172 VGPRs unconstrained (two
// waves); held to 168 it spills 7 and runs 10 % faster.
template <int DIR, class P, int LIM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_unsplit_fused_tile(const UnsplitArgs a, int TP, int TS, int chunks,
                                                            int first, int last)
{
    constexpr int NV = P::NV;
    constexpr int d1 = (DIR == 0) ? 1 : 0;
    constexpr int d2 = (DIR == 2) ? 1 : 2;
    extern __shared__ double lds[];
    double *A = lds;                    // primitives, then fluxes
    double *B = lds + NV * blockDim.x;  // left face states
    const int nt = blockDim.x;
    const int t = threadIdx.x;
    int bid = blockIdx.x;
    const int nfull = (int)(gridDim.x / 8u) * 8;
    if (bid < nfull) bid = (bid % 8) * (nfull / 8) + bid / 8;
    int p, sl, tm, tp;
    if (DIR == 0) {
        sl = t % TS;
        p = t / TS;
        tm = t - 1;
        tp = t + 1;
    } else {
        p = t % TP;
        sl = t / TP;
        tm = t - TP;
        tp = t + TP;
    }
    tm = tm < 0 ? 0 : tm;
    tp = tp >= nt ? nt - 1 : tp;
    const int chunk = bid % chunks;
    const int ptile = bid / chunks;
    const int n1 = a.n[d1];
    const int npen = n1 * a.n[d2];
    const int pf = ptile * TP + p;
    const bool pvalid = t < TP * TS && pf < npen;
    const int pfc = pf < npen ? pf : npen - 1;
    const int out = TS - 4;
    const int c0 = a.g[DIR] + chunk * out;  // the tile's first updated cell
    const int cs = c0 - 2 + sl;             // this lane's cell
    const int ng = a.ng[DIR];
    const int csc = cs < 0 ? 0 : (cs >= ng ? ng - 1 : cs);
    int xyz[3];
    xyz[DIR] = csc;
    xyz[d1] = a.g[d1] + pfc % n1;
    xyz[d2] = a.g[d2] + pfc / n1;
    const long long idx = (long long)xyz[0] + (long long)a.sy * xyz[1] + (long long)a.sz * xyz[2];

    double W[NV], q[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) W[k] = a.src[(long long)uplane(DIR, k) * a.sv + idx];
    prim<P>(W, a.gam_1, q);
#pragma unroll
    for (int v = 0; v < NV; v++) A[v * nt + t] = q[v];
    __syncthreads();
    double fp[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        const int lim = (LIM >= 0) ? LIM : ((v < 5) ? a.lim : a.limb);
        const double h = 0.5 * limit(lim, q[v] - A[v * nt + tm], A[v * nt + tp] - q[v]);
        fp[v] = q[v] + h;
        B[v * nt + t] = q[v] - h;
    }
    __syncthreads();
    double f[NV];
    if (sl >= 1 && sl < TS - 2) {
        double fr[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) fr[v] = B[v * nt + tp];
        if constexpr (P::MAG)
            flux_hlld(fp, fr, a.gam, a.ch, f);
        else
            flux_hllc(fp, fr, a.gam, f);
    } else {
#pragma unroll
        for (int v = 0; v < NV; v++) f[v] = 0.0;
    }
    // every read of A (the primitives) came before the last barrier
#pragma unroll
    for (int v = 0; v < NV; v++) A[v * nt + t] = f[v];
    __syncthreads();
    const bool upd = pvalid && sl >= 2 && sl < 2 + out && cs < a.g[DIR] + a.n[DIR];
    const long long ci = idx;  // an interior cell when upd: csc == cs
    double T[NV];
    unsigned flags = 0u;
    double dmass = 0.0;
    if (upd) {
        const double *S = first ? a.base : a.dst;
#pragma unroll
        for (int v = 0; v < NV; v++) T[v] = S[(long long)v * a.sv + ci];
        const double coef = (a.dt / a.dl[DIR]) * a.rk;
#pragma unroll
        for (int k = 0; k < NV; k++) {
            const int v = uplane(DIR, k);
            T[v] = T[v] + coef * (A[k * nt + tm] - f[k]);
        }
        if (last) {
            for (int d = 0; d < 3; d++) {
                if (!a.has_dir[d]) continue;
                double Ws[NV];
#pragma unroll
                for (int kk = 0; kk < NV; kk++) Ws[kk] = T[uplane(d, kk)];
#pragma unroll
                for (int v = 0; v < 5; v++) Ws[v] = Ws[v] + 0.0;
                if (a.last) {
                    unsigned fl = 0u;
                    double dm = 0.0;
                    floors<P>(Ws, a.smalld, a.smallei, a.use_smalld, a.use_smallei, a.ei_flag, fl, dm);
                    flags |= fl;
                    dmass += dm;  // one cell's floor events, in the direction order
                }
#pragma unroll
                for (int kk = 0; kk < NV; kk++) T[uplane(d, kk)] = Ws[kk];
            }
            if constexpr (P::MAG) T[8] = T[8] * a.psi_fac;  // glmdamping after the stages (1 at the first stage)
        }
#pragma unroll
        for (int v = 0; v < NV; v++) a.dst[(long long)v * a.sv + ci] = T[v];
    }
    // last and a.last are the same for every lane: the whole wave takes part
    if (last && a.last) pk::mass_add(a.dmass, dmass);
    if (!upd) return;
    if (flags) atomicOr(a.flags, flags);
    if (!last) return;
    // periodic images of the finished cell (the stage's ghost refresh): every
    // ghost has exactly one interior pre-image, since n >= nb
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
                const long long j = ci + off[0][i0] + (long long)a.sy * off[1][i1] + (long long)a.sz * off[2][i2];
#pragma unroll
                for (int v = 0; v < NV; v++) a.dst[(long long)v * a.sv + j] = T[v];
            }
}

template <class P>
__global__ __launch_bounds__(256) void k_unsplit_update(const UnsplitArgs a, const double *__restrict__ F0,
                                                        const double *__restrict__ F1,
                                                        const double *__restrict__ F2)
{
    constexpr int NV = P::NV;
    const long long nx = a.n[0], nxy = nx * a.n[1];
    const long long total = nxy * a.n[2];
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned flags = 0u;
    double dmass = 0.0;
    if (e < total) {
        const int i = (int)(e % nx), j = (int)((e / nx) % a.n[1]), k = (int)(e / nxy);
        const long long idx =
            (long long)(i + a.g[0]) + (long long)a.sy * (j + a.g[1]) + (long long)a.sz * (k + a.g[2]);
        double T[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) T[v] = a.base[(long long)v * a.sv + idx];
        const double *Fd[3] = {F0, F1, F2};
        const long long st[3] = {1, a.sy, a.sz};
        for (int d = 0; d < 3; d++) {
            if (!a.has_dir[d]) continue;
            const double coef = (a.dt / a.dl[d]) * a.rk;
            const double *F = Fd[d];
#pragma unroll
            for (int v = 0; v < NV; v++)
                T[v] = T[v] + coef * (F[(long long)v * a.sv + idx] - F[(long long)v * a.sv + idx + st[d]]);
        }
        for (int d = 0; d < 3; d++) {
            if (!a.has_dir[d]) continue;
            double W[NV];
#pragma unroll
            for (int kk = 0; kk < NV; kk++) W[kk] = T[uplane(d, kk)];
#pragma unroll
            for (int v = 0; v < 5; v++) W[v] = W[v] + 0.0;
            if (a.last) {
                unsigned fl = 0u;
                double dm = 0.0;
                floors<P>(W, a.smalld, a.smallei, a.use_smalld, a.use_smallei, a.ei_flag, fl, dm);
                flags |= fl;
                dmass += dm;
            }
#pragma unroll
            for (int kk = 0; kk < NV; kk++) T[uplane(d, kk)] = W[kk];
        }
#pragma unroll
        for (int v = 0; v < NV; v++) a.dst[(long long)v * a.sv + idx] = T[v];
    }
    if (a.last) pk::mass_add(a.dmass, dmass);
    if (flags) atomicOr(a.flags, flags);
}

// Tile shapes of the flux kernel, pencils x cells: along x one wave per
// pencil; along y and z 8 x-consecutive pencils (64 B rows) x 32 cells
// (measured at 256^3 MHD against 4 x 64, 16 x 16 and 32 x 8).
constexpr int kTile[3][2] = {{4, 64}, {8, 32}, {8, 32}};

template <int DIR, class P, int LIM>
hipError_t fused_tile(const UnsplitArgs &a, int first, int last, hipStream_t st)
{
    const int d1 = (DIR == 0) ? 1 : 0, d2 = (DIR == 2) ? 1 : 2;
    const long long npen = (long long)a.n[d1] * a.n[d2];
    const int TP = kTile[DIR][0], TS = kTile[DIR][1];
    const int chunks = (a.n[DIR] + (TS - 4) - 1) / (TS - 4);
    const long long tiles = (npen + TP - 1) / TP * chunks;
    if (tiles < 1 || tiles > 0x7fffffffll) return hipErrorInvalidValue;
    const size_t shm = sizeof(double) * 2 * P::NV * TP * TS;
    hipLaunchKernelGGL((k_unsplit_fused_tile<DIR, P, LIM>), dim3((unsigned)tiles), dim3(TP * TS), shm, st, a, TP,
                       TS, chunks, first, last);
    return hipGetLastError();
}

template <int DIR, class P>
hipError_t fused_lim(const UnsplitArgs &a, int first, int last, hipStream_t st)
{
    if (a.lim == a.limb) {
        switch (a.lim) {
        case LIM_VANLEER: return fused_tile<DIR, P, LIM_VANLEER>(a, first, last, st);
        case LIM_MINMOD: return fused_tile<DIR, P, LIM_MINMOD>(a, first, last, st);
        case LIM_MONCEN: return fused_tile<DIR, P, LIM_MONCEN>(a, first, last, st);
        case LIM_SUPERBEE: return fused_tile<DIR, P, LIM_SUPERBEE>(a, first, last, st);
        default: break;
        }
    }
    return fused_tile<DIR, P, -1>(a, first, last, st);
}

template <int DIR, class P, int LIM>
hipError_t flux_tile(const UnsplitArgs &a, double *F, hipStream_t st)
{
    const int d1 = (DIR == 0) ? 1 : 0, d2 = (DIR == 2) ? 1 : 2;
    const long long npen = (long long)a.n[d1] * a.n[d2];
    const int TP = kTile[DIR][0], TS = kTile[DIR][1];
    const int chunks = (a.n[DIR] + 1 + (TS - 3) - 1) / (TS - 3);
    const long long tiles = (npen + TP - 1) / TP * chunks;
    if (tiles < 1 || tiles > 0x7fffffffll) return hipErrorInvalidValue;
    const size_t shm = sizeof(double) * 2 * P::NV * TP * TS;
    hipLaunchKernelGGL((k_unsplit_flux_tile<DIR, P, LIM>), dim3((unsigned)tiles), dim3(TP * TS), shm, st, a, F, TP,
                       TS, chunks);
    return hipGetLastError();
}

template <int DIR, class P>
hipError_t flux_lim(const UnsplitArgs &a, double *F, unsigned blocks, hipStream_t st)
{
    if (a.lim == a.limb) {
        switch (a.lim) {
        case LIM_VANLEER: return flux_tile<DIR, P, LIM_VANLEER>(a, F, st);
        case LIM_MINMOD: return flux_tile<DIR, P, LIM_MINMOD>(a, F, st);
        case LIM_MONCEN: return flux_tile<DIR, P, LIM_MONCEN>(a, F, st);
        case LIM_SUPERBEE: return flux_tile<DIR, P, LIM_SUPERBEE>(a, F, st);
        default: break;
        }
    }
    return flux_tile<DIR, P, -1>(a, F, st);
}

}  // namespace

hipError_t launch_unsplit_flux(int dir, bool mhd, const UnsplitArgs &a, double *F, hipStream_t st)
{
    if (dir < 0 || dir > 2 || !a.has_dir[dir] || a.g[dir] < 2) return hipErrorInvalidValue;
    const int d1 = (dir == 0) ? 1 : 0, d2 = (dir == 2) ? 1 : 2;
    const long long total = ((long long)a.n[dir] + 1) * a.n[d1] * a.n[d2];
    const long long blocks = (total + 255) / 256;
    if (blocks < 1 || blocks > 0x7fffffffll) return hipErrorInvalidValue;
    if (mhd) {
        if (dir == 0) return flux_lim<0, MHD>(a, F, (unsigned)blocks, st);
        if (dir == 1) return flux_lim<1, MHD>(a, F, (unsigned)blocks, st);
        return flux_lim<2, MHD>(a, F, (unsigned)blocks, st);
    }
    if (dir == 0) return flux_lim<0, HD>(a, F, (unsigned)blocks, st);
    if (dir == 1) return flux_lim<1, HD>(a, F, (unsigned)blocks, st);
    return flux_lim<2, HD>(a, F, (unsigned)blocks, st);
}

hipError_t launch_unsplit_fused(int dir, bool mhd, const UnsplitArgs &a, int first, int last, hipStream_t st)
{
    if (dir < 0 || dir > 2 || !a.has_dir[dir] || a.g[dir] < 2 || a.src == a.dst) return hipErrorInvalidValue;
    if (mhd) {
        if (dir == 0) return fused_lim<0, MHD>(a, first, last, st);
        if (dir == 1) return fused_lim<1, MHD>(a, first, last, st);
        return fused_lim<2, MHD>(a, first, last, st);
    }
    if (dir == 0) return fused_lim<0, HD>(a, first, last, st);
    if (dir == 1) return fused_lim<1, HD>(a, first, last, st);
    return fused_lim<2, HD>(a, first, last, st);
}

hipError_t launch_unsplit_update(bool mhd, const UnsplitArgs &a, const double *const F[3], hipStream_t st)
{
    const long long total = (long long)a.n[0] * a.n[1] * a.n[2];
    const long long blocks = (total + 255) / 256;
    if (blocks < 1 || blocks > 0x7fffffffll) return hipErrorInvalidValue;
    if (mhd)
        hipLaunchKernelGGL(k_unsplit_update<MHD>, dim3((unsigned)blocks), dim3(256), 0, st, a, F[0], F[1], F[2]);
    else
        hipLaunchKernelGGL(k_unsplit_update<HD>, dim3((unsigned)blocks), dim3(256), 0, st, a, F[0], F[1], F[2]);
    return hipGetLastError();
}

}  // namespace PK_KNS
