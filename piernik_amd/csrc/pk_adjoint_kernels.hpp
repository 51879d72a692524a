// pk_adjoint_kernels.hpp -- argument blocks and launchers of the adjoint
// (reverse-mode) kernels in pk_adjoint.hip.
#pragma once
#include <hip/hip_runtime.h>

namespace pk {

// Reverse mode of one fused RK2 hydro sweep of a periodic block.
struct AdjArgs {
    const double *u;        // state at the start of the sweep (block layout)
    const double *lam_out;  // dL/d(sweep output), interior cells
    double *lam_in;         // dL/d(sweep input), interior cells (written)
    double *dtb_part;       // per-wave partial sums of dL/d(dt)
    long long sv;
    int sy, sz;
    int ng[3], g[3], n[3];
    int TP, TS;             // tile: TP pencils x TS cells; TS - 16 output cells
    int chunks, ptiles, out;
    int lim;
    double gam, gam_1;
    double c1, c2;          // (dt/2)/dl and dt/dl
    double dl;              // cell size along the sweep
    double smalld, smallei;
    int use_smalld, use_smallei;
};

// Reverse mode of timestep_fluid (one hydro fluid) at its minimising cell.
struct DtAdjArgs {
    const double *u;
    long long sv;
    int sy, sz;
    int n[3], g[3], has_dir[3];
    double dl[3];
    double gam, gam_1;
    double mn;   // min over cells and directions of dl/c (the forward reduction)
    double cfl;  // dt = cfl * mn
    double dtb;  // dL/d(dt)
};

long long sweep_adjoint_parts(const AdjArgs &a);
hipError_t launch_sweep_adjoint(int dir, const AdjArgs &a, double *dtb_out, hipStream_t st);
hipError_t launch_fold_ghosts(double *lam, long long sv, int sy, int sz, const int ng[3], const int g[3],
                              const int n[3], int nplanes, hipStream_t st);
hipError_t launch_dt_adjoint(const DtAdjArgs &a, unsigned long long *best, double *lam, hipStream_t st);

}  // namespace pk
