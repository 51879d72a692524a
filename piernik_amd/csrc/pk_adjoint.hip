// pk_adjoint.hip -- reverse mode of the fused RK2 hydro sweep (gfx950).
//
// For one directional sweep of a periodic block with hydro fluids
//   U* = floors(U + c1 R(U)),  U' = floors(U + c2 R(U*)),
//   R(W)_i = F(W)_{i-1/2} - F(W)_{i+1/2},  c1 = (dt/2)/dl,  c2 = dt/dl,
// (solve_cg_riemann.F90:366 inside the RK2 loop of sweeps.F90:217-272), the
// kernel takes lam' = dL/dU' and returns lam = dL/dU, plus dL/d(dt), with
// F = HLLC of the limited piecewise-linear face states (hlld.F90:615-841,
// interpolations.F90:112-204).
//
// One lane is one cell of a tile of TP pencils x TS cells along the sweep.
// The tile first recomputes both stages from U (forward values are the
// forward sweep's bit for bit: same functions, same order, no contraction),
// then runs the chain rule backwards; neighbour terms go through LDS.  The
// output cells are lanes 8 .. TS-9: eight cells of halo on each side.
// Reads wrap periodically along the sweep, so ghost cells are not used.
#include <hip/hip_runtime.h>
#include "pk_adjoint.hpp"
#include "pk_adjoint_kernels.hpp"

namespace pk {

template <int DIR>
__device__ __forceinline__ int plane_of_adj(int k)
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

// limiter of variable v: fluid variables use 'limiter', B and psi 'limiter_b'
template <int LIM>
__device__ __forceinline__ int lim_of(const AdjArgs &a, int v)
{
    return (LIM >= 0) ? LIM : ((v < 5) ? a.lim : a.limb);
}

// limited face states of cell k from q_{k-1}, q_k, q_{k+1}
template <class P, int LIM>
__device__ __forceinline__ void faces(const AdjArgs &a, const double (&qm)[P::NV], const double (&q)[P::NV],
                                      const double (&qp)[P::NV], double (&fp)[P::NV], double (&fm)[P::NV])
{
#pragma unroll
    for (int v = 0; v < P::NV; v++) {
        const double dq = limit(lim_of<LIM>(a, v), q[v] - qm[v], qp[v] - q[v]);
        const double h = 0.5 * dq;
        fp[v] = q[v] + h;
        fm[v] = q[v] - h;
    }
}

template <int NV>
__device__ __forceinline__ void lds_get(const double *A, int nt, int t, double (&x)[NV])
{
#pragma unroll
    for (int v = 0; v < NV; v++) x[v] = A[v * nt + t];
}
template <int NV>
__device__ __forceinline__ void lds_put(double *A, int nt, int t, const double (&x)[NV])
{
#pragma unroll
    for (int v = 0; v < NV; v++) A[v * nt + t] = x[v];
}

// LDS arrays of NV doubles per lane: five shared neighbour buffers, and for
// MHD three lane-private ones that hold values live across both flux VJPs
// (registers for hydro, whose occupancy the larger LDS would halve).
template <class P>
constexpr int adj_lds_arrays()
{
    return P::MAG ? 8 : 5;
}

// q_bar of cell k from its face-state adjoints (fpb: left state of k+1/2,
// fmb: right state of k-1/2) through the limiter: the own-cell part and the
// dl/dr adjoints that the neighbours take.
template <class P, int LIM>
__device__ __forceinline__ void faces_vjp(const AdjArgs &a, const double (&qm)[P::NV], const double (&q)[P::NV],
                                          const double (&qp)[P::NV], const double (&fpb)[P::NV],
                                          const double (&fmb)[P::NV], double (&qb)[P::NV], double (&dlb)[P::NV],
                                          double (&drb)[P::NV])
{
#pragma unroll
    for (int v = 0; v < P::NV; v++) {
        const double dqb = 0.5 * fpb[v] - 0.5 * fmb[v];
        limit_vjp(lim_of<LIM>(a, v), q[v] - qm[v], qp[v] - q[v], dqb, dlb[v], drb[v]);
        qb[v] = (fpb[v] + fmb[v]) + (dlb[v] - drb[v]);
    }
}

template <class P>
__device__ __forceinline__ void flux(const AdjArgs &a, const double *l, const double *r, double *F)
{
    if constexpr (P::MAG)
        flux_hlld(l, r, a.gam, a.ch, F);
    else
        flux_hllc(l, r, a.gam, F);
}

// returns dL/dch of this interface
template <class P>
__device__ __forceinline__ double flux_vjp(const AdjArgs &a, const double *l, const double *r, const double *fb,
                                           double *lb, double *rb)
{
    if constexpr (P::MAG) {
        return hlld_vjp(l, r, a.gam, a.ch, fb, lb, rb);
    } else {
        hllc_vjp(l, r, a.gam, fb, lb, rb);
        return 0.0;
    }
}

template <class P>
__device__ __forceinline__ void floors_vjp(const AdjArgs &a, const double (&u)[P::NV], double (&ub)[P::NV])
{
    if constexpr (P::MAG)
        floors_vjp_mhd(u, a.smalld, a.smallei, a.use_smalld, a.use_smallei, ub);
    else
        floors_vjp_hd(u, a.smalld, a.smallei, a.use_smalld, a.use_smallei, ub);
}

template <class P>
__device__ __forceinline__ void prim_vjp(const AdjArgs &a, const double (&u)[P::NV], const double (&qb)[P::NV],
                                         double (&ub)[P::NV])
{
    if constexpr (P::MAG)
        prim_vjp_mhd(u, a.gam_1, qb, ub);
    else
        prim_vjp_hd(u, a.gam_1, qb, ub);
}

// One tile (bid) of the adjoint sweep; lanes t >= TP*TS of a larger
// workgroup take part in the barriers only.
template <int DIR, class P, int LIM>
__device__ __forceinline__ void adjoint_tile(const AdjArgs &a, int bid, double *lds)
{
    constexpr int NV = P::NV;
    const int nt = blockDim.x;
    const int t = threadIdx.x;
    double *Q1 = lds, *Q2 = lds + NV * nt, *S1 = lds + 2 * NV * nt, *S2 = lds + 3 * NV * nt, *S3 = lds + 4 * NV * nt;
    double *K1 = lds + 5 * NV * nt, *K2 = lds + 6 * NV * nt, *K3 = lds + 7 * NV * nt;  // MHD: lane-private
    // park x in a lane-private LDS array (MHD) until fetch() reloads it
    auto park = [&](double *K, const double (&x)[NV]) {
        if constexpr (P::MAG) lds_put<NV>(K, nt, t, x);
    };
    auto fetch = [&](double *K, double (&x)[NV]) {
        if constexpr (P::MAG) lds_get<NV>(K, nt, t, x);
    };

    constexpr int d1 = (DIR == 0) ? 1 : 0;
    constexpr int d2 = (DIR == 2) ? 1 : 2;
    const int step = (DIR == 0) ? 1 : a.TP;
    int p, s;
    if (DIR == 0) {
        s = t % a.TS;
        p = t / a.TS;
    } else {
        p = t % a.TP;
        s = t / a.TP;
    }
    const int tm = max(t - step, 0), tp = min(t + step, nt - 1);

    const int chunk = bid % a.chunks;
    const int ptile = bid / a.chunks;
    const int n1 = a.n[d1], nd = a.n[DIR];
    const int npen = n1 * a.n[d2];
    const int pf = ptile * a.TP + p;
    const bool pvalid = t < a.TP * a.TS && pf < npen;
    const int pfc = pf < npen ? pf : npen - 1;
    const int c1i = a.g[d1] + pfc % n1;
    const int c2i = a.g[d2] + pfc / n1;
    const int cs = chunk * a.out - 8 + s;  // interior coordinate along the sweep, before wrapping
    const int ci = ((cs % nd) + nd) % nd;
    const bool outl = pvalid && s >= 8 && s < 8 + a.out && cs < nd;

    int xyz[3];
    xyz[DIR] = a.g[DIR] + ci;
    xyz[d1] = c1i;
    xyz[d2] = c2i;
    const long long idx = (long long)xyz[0] + (long long)a.sy * xyz[1] + (long long)a.sz * xyz[2];

    double U[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) U[k] = a.u[(long long)plane_of_adj<DIR>(k) * a.sv + idx];

    // ---- forward: stage 1 ----
    double q[NV];
    prim<P>(U, a.gam_1, q);
    lds_put<NV>(Q1, nt, t, q);
    __syncthreads();
    double qm[NV], qp[NV], fp[NV], fm[NV], F[NV], R1[NV], Us_pre[NV], Us[NV];
    lds_get<NV>(Q1, nt, tm, qm);
    lds_get<NV>(Q1, nt, tp, qp);
    faces<P, LIM>(a, qm, q, qp, fp, fm);
    lds_put<NV>(S1, nt, t, fm);
    __syncthreads();
    {
        double r[NV];
        lds_get<NV>(S1, nt, tp, r);
        flux<P>(a, fp, r, F);
        lds_put<NV>(S2, nt, t, F);
    }
    __syncthreads();
    {
        double Fm[NV];
        lds_get<NV>(S2, nt, tm, Fm);
#pragma unroll
        for (int v = 0; v < NV; v++) {
            R1[v] = Fm[v] - F[v];
            const double un = U[v] + a.c1 * R1[v];
            Us_pre[v] = (v < 5) ? un + 0.0 : un;
            Us[v] = Us_pre[v];
        }
        unsigned fl = 0u;
        double dm = 0.0;
        floors<P>(Us, a.smalld, a.smallei, a.use_smalld, a.use_smallei, false, fl, dm);
    }
    park(K1, R1);
    park(K2, Us_pre);
    // ---- forward: stage 2 ----
    double qs[NV];
    prim<P>(Us, a.gam_1, qs);
    lds_put<NV>(Q2, nt, t, qs);
    __syncthreads();
    lds_get<NV>(Q2, nt, tm, qm);
    lds_get<NV>(Q2, nt, tp, qp);
    faces<P, LIM>(a, qm, qs, qp, fp, fm);
    lds_put<NV>(S1, nt, t, fm);
    __syncthreads();
    {
        double r[NV];
        lds_get<NV>(S1, nt, tp, r);
        flux<P>(a, fp, r, F);
        lds_put<NV>(S2, nt, t, F);
    }
    __syncthreads();
    double lp[NV];  // lam' through the psi damping and the floors of stage 2
    double dtb = 0.0, chb = 0.0;
    {
        double Fm[NV], R2[NV], Up[NV];
        lds_get<NV>(S2, nt, tm, Fm);
#pragma unroll
        for (int v = 0; v < NV; v++) {
            R2[v] = Fm[v] - F[v];
            const double un = U[v] + a.c2 * R2[v];
            Up[v] = (v < 5) ? un + 0.0 : un;
        }
        // ---- reverse: stage 2 ----
#pragma unroll
        for (int k = 0; k < NV; k++) lp[k] = a.lam_out[(long long)plane_of_adj<DIR>(k) * a.sv + idx];
        if constexpr (P::MAG) lp[8] = lp[8] * a.psi_fac;
        floors_vjp<P>(a, Up, lp);
        if (outl) {
            double acc = 0.0;
#pragma unroll
            for (int v = 0; v < NV; v++) acc += lp[v] * R2[v];
            dtb += acc / a.dl;  // c2 = dt/dl
        }
    }
    lds_put<NV>(S3, nt, t, lp);  // S1 still holds every cell's stage-2 right-face state fm
    park(K3, lp);
    __syncthreads();
    double fpb[NV];
    {
        // interface k+1/2 of stage 2: l = left state of cell k (recomputed),
        // r = right state of cell k+1 (the value the forward flux took)
        double lpn[NV], fb[NV], q2[NV], l[NV], r[NV], junk[NV], rb[NV];
        lds_get<NV>(S3, nt, tp, lpn);
#pragma unroll
        for (int v = 0; v < NV; v++) fb[v] = a.c2 * (lpn[v] - lp[v]);
        double qc[NV];
        lds_get<NV>(Q2, nt, t, qc);
        lds_get<NV>(Q2, nt, tm, qm);
        lds_get<NV>(Q2, nt, tp, q2);
        faces<P, LIM>(a, qm, qc, q2, l, junk);
        lds_get<NV>(S1, nt, tp, r);
#pragma unroll
        for (int v = 0; v < NV; v++) fpb[v] = rb[v] = 0.0;
        const double cb = flux_vjp<P>(a, l, r, fb, fpb, rb);
        if (outl) chb += cb;
        lds_put<NV>(S2, nt, t, rb);  // belongs to cell k+1
    }
    __syncthreads();
    double lsp[NV];  // lam* through the floors of stage 1
    double l1[NV];   // stage-1 left state of interface k+1/2
    {
        double fmb[NV], qb[NV], dlb[NV], drb[NV];
        double qc[NV];
        lds_get<NV>(S2, nt, tm, fmb);
        lds_get<NV>(Q2, nt, t, qc);
        lds_get<NV>(Q2, nt, tm, qm);
        lds_get<NV>(Q2, nt, tp, qp);
        faces_vjp<P, LIM>(a, qm, qc, qp, fpb, fmb, qb, dlb, drb);
        lds_put<NV>(S1, nt, t, dlb);
        lds_put<NV>(S3, nt, t, drb);
        __syncthreads();
        double dln[NV], drp[NV];
        lds_get<NV>(S1, nt, tp, dln);
        lds_get<NV>(S3, nt, tm, drp);
#pragma unroll
        for (int v = 0; v < NV; v++) {
            qb[v] += drp[v] - dln[v];
            lsp[v] = 0.0;
        }
        fetch(K2, Us_pre);
        {
            double Usr[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) Usr[v] = Us_pre[v];
            unsigned fl = 0u;
            double dm = 0.0;
            floors<P>(Usr, a.smalld, a.smallei, a.use_smalld, a.use_smallei, false, fl, dm);
            prim_vjp<P>(a, Usr, qb, lsp);
        }
        floors_vjp<P>(a, Us_pre, lsp);
        fetch(K1, R1);
        if (outl) {
            double acc = 0.0;
#pragma unroll
            for (int v = 0; v < NV; v++) acc += lsp[v] * R1[v];
            dtb += (0.5 * acc) / a.dl;  // c1 = (dt/2)/dl
        }
        lds_put<NV>(S2, nt, t, lsp);
        // stage-1 face states of cell k: l1 stays here, fm goes to Q2 (whose
        // reads all came before the last barrier) for the left neighbour
        double qc1[NV], qp1[NV], fm1[NV];
        lds_get<NV>(Q1, nt, t, qc1);
        lds_get<NV>(Q1, nt, tm, qm);
        lds_get<NV>(Q1, nt, tp, qp1);
        faces<P, LIM>(a, qm, qc1, qp1, l1, fm1);
        lds_put<NV>(Q2, nt, t, fm1);
    }
    __syncthreads();
    // ---- reverse: stage 1 ----
    {
        double lsn[NV], fb[NV], r[NV], rb[NV];
        lds_get<NV>(S2, nt, tp, lsn);
        fetch(K3, lp);
#pragma unroll
        for (int v = 0; v < NV; v++) {
            fb[v] = a.c1 * (lsn[v] - lsp[v]);
            lsp[v] = lsp[v] + lp[v];  // from here on lam* + lam' (the two direct terms)
        }
        lds_get<NV>(Q2, nt, tp, r);  // stage-1 right state of cell k+1
#pragma unroll
        for (int v = 0; v < NV; v++) fpb[v] = rb[v] = 0.0;
        const double cb = flux_vjp<P>(a, l1, r, fb, fpb, rb);
        if (outl) chb += cb;
        lds_put<NV>(S1, nt, t, rb);
    }
    __syncthreads();
    double lam[NV];
    {
        double fmb[NV], qb[NV], dlb[NV], drb[NV];
        double qc[NV];
        lds_get<NV>(S1, nt, tm, fmb);
        lds_get<NV>(Q1, nt, t, qc);
        lds_get<NV>(Q1, nt, tm, qm);
        lds_get<NV>(Q1, nt, tp, qp);
        faces_vjp<P, LIM>(a, qm, qc, qp, fpb, fmb, qb, dlb, drb);
        lds_put<NV>(S2, nt, t, dlb);
        lds_put<NV>(S3, nt, t, drb);
        __syncthreads();
        double dln[NV], drp[NV];
        lds_get<NV>(S2, nt, tp, dln);
        lds_get<NV>(S3, nt, tm, drp);
#pragma unroll
        for (int v = 0; v < NV; v++) {
            qb[v] += drp[v] - dln[v];
            lam[v] = 0.0;
        }
        double Ur[NV];  // reloaded: cheaper than holding U across the kernel
#pragma unroll
        for (int k = 0; k < NV; k++) Ur[k] = a.u[(long long)plane_of_adj<DIR>(k) * a.sv + idx];
        prim_vjp<P>(a, Ur, qb, lam);
#pragma unroll
        for (int v = 0; v < NV; v++) lam[v] = lam[v] + lsp[v];
    }
    if (outl) {
#pragma unroll
        for (int k = 0; k < NV; k++) a.lam_in[(long long)plane_of_adj<DIR>(k) * a.sv + idx] = lam[k];
    }
    // dL/d(dt) and dL/d(ch): per-wave sums in a fixed order, one slot per wave of the grid
    const int lane = t & 63;
    const int nact = min(64, nt - (t & ~63));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double o1 = __shfl_xor(dtb, o, 64);
        const double o2 = __shfl_xor(chb, o, 64);
        // the lower lane of each pair accumulates; lane 0 ends with the sum
        if ((lane ^ o) < nact && lane < (lane ^ o)) {
            dtb += o1;
            chb += o2;
        }
    }
    if (lane == 0) {
        const long long w = (long long)bid * ((nt + 63) / 64) + t / 64;
        a.dtb_part[w] = dtb;
        a.chb_part[w] = chb;
    }
}

template <int DIR, class P, int LIM>
__global__ __launch_bounds__(256) void k_sweep_adjoint(const AdjArgs a)
{
    extern __shared__ double lds[];
    adjoint_tile<DIR, P, LIM>(a, blockIdx.x, lds);
}

// Persistent adjoint sweep over a block list (independent instances): as
// k_sweep_list, tiles from a device queue, per-block arguments from a table.
template <int DIR, class P, int LIM>
__global__ __launch_bounds__(256) void k_sweep_adjoint_list(const AdjListArgs L)
{
    extern __shared__ double lds[];
    __shared__ unsigned s_item;
    for (;;) {
        __syncthreads();  // the previous tile is done with LDS and s_item
        if (threadIdx.x == 0) s_item = atomicAdd(L.queue, 1u);
        __syncthreads();
        const unsigned item = (unsigned)__builtin_amdgcn_readfirstlane((int)s_item);
        if (item >= L.nitems) break;
        typedef const __attribute__((address_space(4))) unsigned cu32;
        const int blk = __builtin_amdgcn_readfirstlane((int)((cu32 *)L.items)[item]);
        AdjArgs a;
        {
            static_assert(sizeof(AdjArgs) % 4 == 0, "AdjArgs is copied in dwords");
            cu32 *w = (cu32 *)(L.args + blk);
            unsigned *aw = reinterpret_cast<unsigned *>(&a);
#pragma unroll
            for (int i = 0; i < (int)(sizeof(AdjArgs) / 4); i++) aw[i] = w[i];
        }
        const unsigned first = ((cu32 *)L.first)[blk];
        adjoint_tile<DIR, P, LIM>(a, (int)(item - first), lds);
    }
}

// Deterministic sum of n per-wave partials: kRedParts fixed ranges
// [n*g/kRedParts, n*(g+1)/kRedParts), each summed by one 256-lane workgroup
// (per-lane strided sums, then a tree), then the range sums added in order g.
// The one-workgroup-per-job kernel repeats the same ranges and order, so both
// give the same bits.
__device__ double wg_range_sum(const double *__restrict__ part, long long lo, long long hi, double *red)
{
    const int t = threadIdx.x;
    double acc = 0.0;
    for (long long i = lo + t; i < hi; i += 256) acc += part[i];
    red[t] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) red[t] = red[t] + red[t + w];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

__device__ __forceinline__ long long range_lo(long long n, int g) { return n * g / kRedParts; }

// one workgroup per job
__global__ __launch_bounds__(256) void k_sum_parts_jobs(const SumJob *__restrict__ jobs)
{
    __shared__ double red[256];
    const SumJob j = jobs[blockIdx.x];
    double tot = 0.0;
    for (int g = 0; g < kRedParts; g++) tot += wg_range_sum(j.part, range_lo(j.n, g), range_lo(j.n, g + 1), red);
    if (threadIdx.x == 0) j.out[0] = tot;
}

// grid (kRedParts, 2): range g of partial array y, into part_y[n + g]
__global__ __launch_bounds__(256) void k_sum_ranges(double *__restrict__ p0, double *__restrict__ p1, long long n)
{
    __shared__ double red[256];
    double *p = blockIdx.y ? p1 : p0;
    const int g = blockIdx.x;
    const double r = wg_range_sum(p, range_lo(n, g), range_lo(n, g + 1), red);
    if (threadIdx.x == 0) p[n + g] = r;
}

// grid 2: the range sums in order
__global__ void k_sum_final(const double *__restrict__ p0, const double *__restrict__ p1, long long n,
                            double *__restrict__ out)
{
    if (threadIdx.x != 0) return;
    const double *p = blockIdx.x ? p1 : p0;
    double tot = 0.0;
    for (int g = 0; g < kRedParts; g++) tot += p[n + g];
    out[blockIdx.x] = tot;
}

// Adjoint of the periodic ghost refresh along DIR (k_fill_ghosts): every
// ghost cell's adjoint is added to its interior image and cleared.  Run for
// z, then y, then x (the refresh runs x, y, z).
template <int DIR>
__global__ void k_fold_ghosts(double *__restrict__ lam, long long sv, int sy, int sz, int ng0, int ng1, int ng2,
                              int g0, int g1, int g2, int n0, int n1, int n2, int nplanes)
{
    const int ng[3] = {ng0, ng1, ng2}, g[3] = {g0, g1, g2}, n[3] = {n0, n1, n2};
    constexpr int e1 = (DIR == 0) ? 1 : 0;
    constexpr int e2 = (DIR == 2) ? 1 : 2;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    // one thread per (plane, transverse cell, interior image along DIR); it
    // owns the ghosts that copy that image, so no two threads touch one cell
    const long long per = (long long)n[DIR] * ng[e1] * ng[e2];
    if (tid >= per * nplanes) return;
    const int v = (int)(tid / per);
    long long r = tid % per;
    const int i = (int)(r % n[DIR]);
    r /= n[DIR];
    int xyz[3];
    xyz[e1] = (int)(r % ng[e1]);
    xyz[e2] = (int)(r / ng[e1]);
    double *base = lam + (long long)v * sv;
    xyz[DIR] = g[DIR] + i;
    const long long dst = (long long)xyz[0] + (long long)sy * xyz[1] + (long long)sz * xyz[2];
    double acc = base[dst];
    // ghost layers whose image is i: low ghost j (0 <= j < g) copies g + n - g + j ... i.e. image n + j - g + g
    for (int j = 0; j < g[DIR]; j++) {
        // low ghost at j copies interior index n - g + j; high ghost at g + n + j copies interior index j
        if (i == n[DIR] - g[DIR] + j) {
            xyz[DIR] = j;
            const long long gi = (long long)xyz[0] + (long long)sy * xyz[1] + (long long)sz * xyz[2];
            acc += base[gi];
            base[gi] = 0.0;
        }
        if (i == j) {
            xyz[DIR] = g[DIR] + n[DIR] + j;
            const long long gi = (long long)xyz[0] + (long long)sy * xyz[1] + (long long)sz * xyz[2];
            acc += base[gi];
            base[gi] = 0.0;
        }
    }
    base[dst] = acc;
}

__global__ void k_zero_best(unsigned long long *best)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *best = ~0ull;
}

// timestep adjoint: the cell and direction where dl_d/c_d is smallest (the
// smallest linear index among exact ties), then dL/dU there.
__device__ __forceinline__ double cell_cs(const DtAdjArgs &a, const double *u)
{
    const double d = u[0], e = u[4 * a.sv];
    const double m0 = u[a.sv], m1 = u[2 * a.sv], m2 = u[3 * a.sv];
    const double ek = (0.5 * ((m0 * m0 + m1 * m1) + m2 * m2)) / d;
    if (a.magnetized) {
        const double bx = u[5 * a.sv], by = u[6 * a.sv], bz = u[7 * a.sv];
        const double pmag = 0.5 * ((bx * bx + by * by) + bz * bz);
        const double ps = (e - ek) * a.gam_1 + (2.0 - a.gam) * pmag;
        const double pr = ps - pmag;
        return sqrt(fabs((2.0 * pmag + a.gam * pr) / d));
    }
    const double pr = (e - ek) * a.gam_1;
    return sqrt(fabs((a.gam * pr) / d));
}

__global__ __launch_bounds__(256) void k_dt_argmin(const DtAdjArgs a, unsigned long long *best)
{
    const long long ncell = (long long)a.n[0] * a.n[1] * a.n[2];
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
         c += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(c % a.n[0]);
        const long long r = c / a.n[0];
        const int j = (int)(r % a.n[1]);
        const int k = (int)(r / a.n[1]);
        const long long idx =
            (long long)(i + a.g[0]) + (long long)a.sy * (j + a.g[1]) + (long long)a.sz * (k + a.g[2]);
        const double *u = a.u + idx;
        const double d = u[0];
        const double cs = cell_cs(a, u);
        for (int q = 0; q < 3; q++) {
            if (!a.has_dir[q]) continue;
            const double v = (d > 0.0) ? fabs(u[(1 + q) * a.sv] / d) : 0.0;
            const double cc = fmx(v + cs, 1.17549435082228750797e-38);
            if (a.dl[q] / cc == a.mn) atomicMin(best, (unsigned long long)(c * 3 + q));
        }
    }
}

__device__ void k_dt_apply_body(const DtAdjArgs &a, const unsigned long long *best, double *lam)
{
    const unsigned long long b = *best;
    if (b == ~0ull) return;
    const long long c = (long long)(b / 3);
    const int q = (int)(b % 3);
    const int i = (int)(c % a.n[0]);
    const long long r = c / a.n[0];
    const int j = (int)(r % a.n[1]);
    const int k = (int)(r / a.n[1]);
    const long long idx = (long long)(i + a.g[0]) + (long long)a.sy * (j + a.g[1]) + (long long)a.sz * (k + a.g[2]);
    // dt = cfl * dl_q / c_q
    if (a.magnetized) {
        double u[9], ub[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int v = 0; v < 9; v++) u[v] = a.u[(long long)v * a.sv + idx];
        dt_vjp_mhd(u, q, a.dl[q], a.gam, a.gam_1, a.cfl * a.dtb, ub);
        for (int v = 0; v < 8; v++) lam[(long long)v * a.sv + idx] += ub[v];
    } else {
        double u[5], ub[5] = {0, 0, 0, 0, 0};
        for (int v = 0; v < 5; v++) u[v] = a.u[(long long)v * a.sv + idx];
        dt_vjp_hd(u, q, a.dl[q], a.gam, a.gam_1, a.cfl * a.dtb, ub);
        for (int v = 0; v < 5; v++) lam[(long long)v * a.sv + idx] += ub[v];
    }
}

__global__ void k_dt_apply(const DtAdjArgs a, const unsigned long long *best, double *lam)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    k_dt_apply_body(a, best, lam);
}

// one workgroup per job: min over the job's cells and directions of dl/c
__global__ __launch_bounds__(256) void k_dt_mn_jobs(const DtBatchJob *__restrict__ jobs)
{
    __shared__ double red[256];
    const DtBatchJob &J = jobs[blockIdx.x];
    const DtAdjArgs &a = J.a;
    const long long ncell = (long long)a.n[0] * a.n[1] * a.n[2];
    double m = 1.7976931348623157e308;
    for (long long c = threadIdx.x; c < ncell; c += 256) {
        const int i = (int)(c % a.n[0]);
        const long long r = c / a.n[0];
        const int j = (int)(r % a.n[1]);
        const int k = (int)(r / a.n[1]);
        const long long idx =
            (long long)(i + a.g[0]) + (long long)a.sy * (j + a.g[1]) + (long long)a.sz * (k + a.g[2]);
        const double *u = a.u + idx;
        const double d = u[0];
        const double cs = cell_cs(a, u);
        for (int q = 0; q < 3; q++) {
            if (!a.has_dir[q]) continue;
            const double v = (d > 0.0) ? fabs(u[(1 + q) * a.sv] / d) : 0.0;
            m = fmn(m, a.dl[q] / fmx(v + cs, 1.17549435082228750797e-38));
        }
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmn(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *J.mn = red[0];
        *J.best = ~0ull;
    }
}

// grid (chunks, jobs): the smallest cell/direction index attaining the minimum
__global__ __launch_bounds__(256) void k_dt_argmin_jobs(const DtBatchJob *__restrict__ jobs)
{
    const DtBatchJob &J = jobs[blockIdx.y];
    const DtAdjArgs &a = J.a;
    const double mn = *J.mn;
    const long long ncell = (long long)a.n[0] * a.n[1] * a.n[2];
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
         c += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(c % a.n[0]);
        const long long r = c / a.n[0];
        const int j = (int)(r % a.n[1]);
        const int k = (int)(r / a.n[1]);
        const long long idx =
            (long long)(i + a.g[0]) + (long long)a.sy * (j + a.g[1]) + (long long)a.sz * (k + a.g[2]);
        const double *u = a.u + idx;
        const double d = u[0];
        const double cs = cell_cs(a, u);
        for (int q = 0; q < 3; q++) {
            if (!a.has_dir[q]) continue;
            const double v = (d > 0.0) ? fabs(u[(1 + q) * a.sv] / d) : 0.0;
            if (a.dl[q] / fmx(v + cs, 1.17549435082228750797e-38) == mn)
                atomicMin(J.best, (unsigned long long)(c * 3 + q));
        }
    }
}

// one thread per job
__global__ void k_dt_apply_jobs(const DtBatchJob *__restrict__ jobs, int njobs)
{
    const int jb = blockIdx.x * blockDim.x + threadIdx.x;
    if (jb >= njobs) return;
    const DtBatchJob &J = jobs[jb];
    DtAdjArgs a = J.a;
    a.dtb = J.dtb;
    k_dt_apply_body(a, J.best, J.lam);
}

// ---- launchers ------------------------------------------------------------------

template <int DIR, class P, int LIM>
static hipError_t adj_t(const AdjArgs &a, int nt, long long nblk, hipStream_t st)
{
    const size_t lds = sizeof(double) * adj_lds_arrays<P>() * P::NV * nt;
    static bool raised = false;
    if (!raised) {
        hipError_t e = hipFuncSetAttribute((const void *)k_sweep_adjoint<DIR, P, LIM>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        raised = true;
    }
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_sweep_adjoint<DIR, P, LIM>), dim3((unsigned)nblk), dim3(nt), lds, st, a);
    return hipGetLastError();
}

template <int DIR, class P>
static hipError_t adj_lim(const AdjArgs &a, int nt, long long nblk, hipStream_t st)
{
    if (a.lim == a.limb || !P::MAG) {
        switch (a.lim) {
        case LIM_VANLEER: return adj_t<DIR, P, LIM_VANLEER>(a, nt, nblk, st);
        case LIM_MINMOD: return adj_t<DIR, P, LIM_MINMOD>(a, nt, nblk, st);
        case LIM_MONCEN: return adj_t<DIR, P, LIM_MONCEN>(a, nt, nblk, st);
        case LIM_SUPERBEE: return adj_t<DIR, P, LIM_SUPERBEE>(a, nt, nblk, st);
        default: return hipErrorInvalidValue;
        }
    }
    return adj_t<DIR, P, -1>(a, nt, nblk, st);
}

template <int DIR, class P, int LIM>
static hipError_t adjl_t(const AdjListArgs &L, int grid, hipStream_t st, int *resident)
{
    const int nt = 256;
    const void *fn = (const void *)k_sweep_adjoint_list<DIR, P, LIM>;
    const size_t lds = sizeof(double) * adj_lds_arrays<P>() * P::NV * nt;
    const size_t lds_max = 160 * 1024 - 256;
    static bool raised = false;
    if (!raised) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_max);
        if (e != hipSuccess) return e;
        raised = true;
    }
    if (lds > lds_max) return hipErrorInvalidValue;
    if (resident) {
        int per = 0, dev = 0, ncu = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, nt, lds);
        if (e == hipSuccess) e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        *resident = per * ncu;
        return e;
    }
    hipLaunchKernelGGL((k_sweep_adjoint_list<DIR, P, LIM>), dim3((unsigned)grid), dim3(nt), lds, st, L);
    return hipGetLastError();
}

template <int DIR, class P>
static hipError_t adjl_lim(int lim, int limb, const AdjListArgs &L, int grid, hipStream_t st, int *resident)
{
    if (lim == limb || !P::MAG) {
        switch (lim) {
        case LIM_VANLEER: return adjl_t<DIR, P, LIM_VANLEER>(L, grid, st, resident);
        case LIM_MINMOD: return adjl_t<DIR, P, LIM_MINMOD>(L, grid, st, resident);
        case LIM_MONCEN: return adjl_t<DIR, P, LIM_MONCEN>(L, grid, st, resident);
        case LIM_SUPERBEE: return adjl_t<DIR, P, LIM_SUPERBEE>(L, grid, st, resident);
        default: return hipErrorInvalidValue;
        }
    }
    return adjl_t<DIR, P, -1>(L, grid, st, resident);
}

static hipError_t adjl_dispatch(int dir, bool mhd, int lim, int limb, const AdjListArgs &L, int grid, hipStream_t st,
                                int *resident)
{
    if (mhd) {
        if (dir == 0) return adjl_lim<0, MHD>(lim, limb, L, grid, st, resident);
        if (dir == 1) return adjl_lim<1, MHD>(lim, limb, L, grid, st, resident);
        if (dir == 2) return adjl_lim<2, MHD>(lim, limb, L, grid, st, resident);
    } else {
        if (dir == 0) return adjl_lim<0, HD>(lim, limb, L, grid, st, resident);
        if (dir == 1) return adjl_lim<1, HD>(lim, limb, L, grid, st, resident);
        if (dir == 2) return adjl_lim<2, HD>(lim, limb, L, grid, st, resident);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_sweep_adjoint_list(int dir, bool mhd, int lim, int limb, const AdjListArgs &L, hipStream_t st)
{
    int r = 0;
    hipError_t e = adjl_dispatch(dir, mhd, lim, limb, L, 0, st, &r);
    if (e != hipSuccess) return e;
    return adjl_dispatch(dir, mhd, lim, limb, L, r > 0 ? r : 1, st, nullptr);
}

hipError_t launch_sum_parts_jobs(const SumJob *jobs, int njobs, hipStream_t st)
{
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sum_parts_jobs, dim3((unsigned)njobs), dim3(256), 0, st, jobs);
    return hipGetLastError();
}

template <class P>
static hipError_t adj_dir(int dir, const AdjArgs &a, int nt, long long nblk, hipStream_t st)
{
    if (dir == 0) return adj_lim<0, P>(a, nt, nblk, st);
    if (dir == 1) return adj_lim<1, P>(a, nt, nblk, st);
    if (dir == 2) return adj_lim<2, P>(a, nt, nblk, st);
    return hipErrorInvalidValue;
}

// Every adjoint tile runs on a 256-lane workgroup (lanes past TP*TS idle),
// as in the persistent list kernel: four partial slots per tile in both, so
// their sums of the same partials agree bit for bit.
long long sweep_adjoint_parts(const AdjArgs &a) { return (long long)a.chunks * a.ptiles * 4; }

hipError_t launch_sweep_adjoint(int dir, bool mhd, const AdjArgs &a, double *dtb_out, hipStream_t st)
{
    const int nt = 256;
    if (a.TP * a.TS < 1 || a.TP * a.TS > nt || a.out < 1 || a.out != a.TS - 16) return hipErrorInvalidValue;
    const long long nblk = (long long)a.chunks * a.ptiles;
    hipError_t e = mhd ? adj_dir<MHD>(dir, a, nt, nblk, st) : adj_dir<HD>(dir, a, nt, nblk, st);
    if (e != hipSuccess) return e;
    const long long n = sweep_adjoint_parts(a);
    hipLaunchKernelGGL(k_sum_ranges, dim3(kRedParts, 2), dim3(256), 0, st, a.dtb_part, a.chb_part, n);
    hipLaunchKernelGGL(k_sum_final, dim3(2), dim3(64), 0, st, a.dtb_part, a.chb_part, n, dtb_out);
    return hipGetLastError();
}

hipError_t launch_fold_ghosts(double *lam, long long sv, int sy, int sz, const int ng[3], const int g[3],
                              const int n[3], int nplanes, hipStream_t st)
{
    for (int d = 2; d >= 0; d--) {
        if (g[d] == 0) continue;
        const int e1 = (d == 0) ? 1 : 0, e2 = (d == 2) ? 1 : 2;
        const long long total = (long long)n[d] * ng[e1] * ng[e2] * nplanes;
        const unsigned blocks = (unsigned)((total + 255) / 256);
        if (d == 0)
            hipLaunchKernelGGL(k_fold_ghosts<0>, dim3(blocks), dim3(256), 0, st, lam, sv, sy, sz, ng[0], ng[1], ng[2],
                               g[0], g[1], g[2], n[0], n[1], n[2], nplanes);
        if (d == 1)
            hipLaunchKernelGGL(k_fold_ghosts<1>, dim3(blocks), dim3(256), 0, st, lam, sv, sy, sz, ng[0], ng[1], ng[2],
                               g[0], g[1], g[2], n[0], n[1], n[2], nplanes);
        if (d == 2)
            hipLaunchKernelGGL(k_fold_ghosts<2>, dim3(blocks), dim3(256), 0, st, lam, sv, sy, sz, ng[0], ng[1], ng[2],
                               g[0], g[1], g[2], n[0], n[1], n[2], nplanes);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_dt_adjoint_batch(const DtBatchJob *jobs, int njobs, hipStream_t st)
{
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dt_mn_jobs, dim3((unsigned)njobs), dim3(256), 0, st, jobs);
    hipLaunchKernelGGL(k_dt_argmin_jobs, dim3(64, (unsigned)njobs), dim3(256), 0, st, jobs);
    hipLaunchKernelGGL(k_dt_apply_jobs, dim3((unsigned)((njobs + 63) / 64)), dim3(64), 0, st, jobs, njobs);
    return hipGetLastError();
}

hipError_t launch_dt_adjoint(const DtAdjArgs &a, unsigned long long *best, double *lam, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_best, dim3(1), dim3(1), 0, st, best);
    const long long ncell = (long long)a.n[0] * a.n[1] * a.n[2];
    long long blocks = (ncell + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_dt_argmin, dim3((unsigned)blocks), dim3(256), 0, st, a, best);
    hipLaunchKernelGGL(k_dt_apply, dim3(1), dim3(64), 0, st, a, best, lam);
    return hipGetLastError();
}

}  // namespace pk
