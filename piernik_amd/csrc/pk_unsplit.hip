// pk_unsplit.hip -- gfx950 kernels of the unsplit Riemann stage
// (solve_cg_ub / solve_cg_u of src/scheme/unsplit_schemes/, one RK stage).
//
// k_unsplit_flux: one lane per interface of the transverse-interior
// pencils along one direction.  A lane reads the four cells of its stencil
// (sweep order), forms the two face states with the configured limiter and
// evaluates HLLD + GLM or HLLC, exactly as interpol and riemann_wrap do for
// that interface (unsplit_mag_modules.F90:166-209, solve_cg_unsplit.F90:157-187).
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
// Built without FMA contraction: bit-identical to the oracle's unsplit driver.
#include <hip/hip_runtime.h>
#include "pk_math.hpp"
#include "pk_unsplit.hpp"

namespace pk {

namespace {

// sweep-order variable k -> xyz plane (fluidtypes.F90:318-320, fluidindex.F90:190-192)
__device__ __forceinline__ int uplane(int dir, int k)
{
    if (dir == 0 || k == 0 || k == 4 || k == 8) return k;
    if (dir == 1) return (k == 1 || k == 2) ? 3 - k : ((k == 5 || k == 6) ? 11 - k : k);
    return (k == 1 || k == 3) ? 4 - k : ((k == 5 || k == 7) ? 12 - k : k);
}

template <int DIR, class P, int LIM>
__global__ __launch_bounds__(256) void k_unsplit_flux(const UnsplitArgs a, double *__restrict__ F)
{
    constexpr int NV = P::NV;
    constexpr int d1 = (DIR == 0) ? 1 : 0;
    constexpr int d2 = (DIR == 2) ? 1 : 2;
    const long long nrow = (long long)a.n[DIR] + 1;  // interfaces g-1/2 .. g+n-1/2
    const long long n1 = a.n[d1];
    const long long total = nrow * n1 * a.n[d2];
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    long long row, i1, i2;
    if (DIR == 0) {
        row = e % nrow;
        const long long r = e / nrow;
        i1 = r % n1;
        i2 = r / n1;
    } else {
        i1 = e % n1;
        const long long r = e / n1;
        row = r % nrow;
        i2 = r / nrow;
    }
    int xyz[3];
    xyz[DIR] = a.g[DIR] + (int)row;
    xyz[d1] = a.g[d1] + (int)i1;
    xyz[d2] = a.g[d2] + (int)i2;
    const long long idx = (long long)xyz[0] + (long long)a.sy * xyz[1] + (long long)a.sz * xyz[2];
    const long long st = (DIR == 0) ? 1ll : ((DIR == 1) ? (long long)a.sy : (long long)a.sz);

    // primitives of cells c-2 .. c+1 around the interface c-1/2
    double q[4][NV];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        double W[NV];
        const long long o = idx + (j - 2) * st;
#pragma unroll
        for (int k = 0; k < NV; k++) W[k] = a.src[(long long)uplane(DIR, k) * a.sv + o];
        prim<P>(W, a.gam_1, q[j]);
    }
    double l[NV], r[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        const int lim = (LIM >= 0) ? LIM : ((v < 5) ? a.lim : a.limb);
        // right face of cell c-1, left face of cell c
        l[v] = q[1][v] + 0.5 * limit(lim, q[1][v] - q[0][v], q[2][v] - q[1][v]);
        r[v] = q[2][v] - 0.5 * limit(lim, q[2][v] - q[1][v], q[3][v] - q[2][v]);
    }
    double f[NV];
    if constexpr (P::MAG)
        flux_hlld(l, r, a.gam, a.ch, f);
    else
        flux_hllc(l, r, a.gam, f);
#pragma unroll
    for (int k = 0; k < NV; k++) F[(long long)uplane(DIR, k) * a.sv + idx] = f[k];
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
    if (e >= total) return;
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
    unsigned flags = 0u;
    double dmass = 0.0;
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
            floors<P>(W, a.smalld, a.smallei, a.use_smalld, a.use_smallei, fl, dm);
            flags |= fl;
            dmass += dm;
        }
#pragma unroll
        for (int kk = 0; kk < NV; kk++) T[uplane(d, kk)] = W[kk];
    }
#pragma unroll
    for (int v = 0; v < NV; v++) a.dst[(long long)v * a.sv + idx] = T[v];
    if (flags) atomicOr(a.flags, flags);
    if (dmass != 0.0) atomicAdd(a.dmass, dmass);
}

template <int DIR, class P>
hipError_t flux_lim(const UnsplitArgs &a, double *F, unsigned blocks, hipStream_t st)
{
    if (a.lim == a.limb) {
        switch (a.lim) {
        case LIM_VANLEER:
            hipLaunchKernelGGL((k_unsplit_flux<DIR, P, LIM_VANLEER>), dim3(blocks), dim3(256), 0, st, a, F);
            return hipGetLastError();
        case LIM_MINMOD:
            hipLaunchKernelGGL((k_unsplit_flux<DIR, P, LIM_MINMOD>), dim3(blocks), dim3(256), 0, st, a, F);
            return hipGetLastError();
        case LIM_MONCEN:
            hipLaunchKernelGGL((k_unsplit_flux<DIR, P, LIM_MONCEN>), dim3(blocks), dim3(256), 0, st, a, F);
            return hipGetLastError();
        case LIM_SUPERBEE:
            hipLaunchKernelGGL((k_unsplit_flux<DIR, P, LIM_SUPERBEE>), dim3(blocks), dim3(256), 0, st, a, F);
            return hipGetLastError();
        default: break;
        }
    }
    hipLaunchKernelGGL((k_unsplit_flux<DIR, P, -1>), dim3(blocks), dim3(256), 0, st, a, F);
    return hipGetLastError();
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

}  // namespace pk
