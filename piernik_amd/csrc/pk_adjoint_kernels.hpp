// pk_adjoint_kernels.hpp -- argument blocks and launchers of the adjoint
// (reverse-mode) kernels in pk_adjoint.hip.
#pragma once
#include <hip/hip_runtime.h>

namespace pk {

// Reverse mode of one fused RK2 sweep of a periodic block (hydro or MHD+GLM).
struct AdjArgs {
    const double *u;        // state at the start of the sweep (block layout)
    const double *lam_out;  // dL/d(sweep output), interior cells
    double *lam_in;         // dL/d(sweep input), interior cells (written)
    double *dtb_part;       // per-wave partial sums of dL/d(dt) at fixed ch
    double *chb_part;       // per-wave partial sums of dL/d(ch) (MHD)
    long long sv;
    int sy, sz;
    int ng[3], g[3], n[3];
    int TP, TS;             // tile: TP pencils x TS cells; TS - 16 output cells
    int chunks, ptiles, out;
    int lim, limb;          // limiter of the fluid variables, of B and psi
    double gam, gam_1;
    double c1, c2;          // (dt/2)/dl and dt/dl
    double dl;              // cell size along the sweep
    double ch;              // GLM speed of the sweep (MHD)
    double psi_fac;         // the damping factor the forward sweep applied to psi (MHD; else 1)
    double smalld, smallei;
    int use_smalld, use_smallei;
};

// Reverse mode of timestep_fluid (one fluid; neu_cs or ion_cs) at its
// minimising cell.
struct DtAdjArgs {
    const double *u;        // rho plane of the fluid; B planes follow at +5 when magnetized
    int magnetized;
    long long sv;
    int sy, sz;
    int n[3], g[3], has_dir[3];
    double dl[3];
    double gam, gam_1;
    double mn;   // min over cells and directions of dl/c (the forward reduction)
    double cfl;  // dt = cfl * mn
    double dtb;  // dL/d(dt)
};

// Persistent adjoint sweep over a block list (workgroups of 256 lanes take
// tiles from a queue; tile item belongs to block items[item] and is its
// tile item - first[block]).  Each block's AdjArgs has its tile shape with
// TP*TS <= 256 and its dtb/chb partial arrays sized for 4 waves per tile.
struct AdjListArgs {
    const AdjArgs *args;
    const unsigned *first;
    const unsigned *items;
    unsigned nitems;
    unsigned *queue;
};

// out[0] = sum of part[0 .. n-1] in a fixed order
struct SumJob {
    const double *part;
    long long n;
    double *out;
};

// Batched timestep adjoint over independent instances: per job the minimum
// of dl/c (written to *mn), its cell and direction (*best), then
// lam += dtb * d(cfl*dl/c)/du there.  a.mn and a.dtb are ignored (dtb here).
struct DtBatchJob {
    DtAdjArgs a;
    double dtb;
    double *lam;
    double *mn;
    unsigned long long *best;
};

// Partial-sum buffers (dtb_part, chb_part) hold sweep_adjoint_parts(a)
// partials followed by kRedParts doubles of reduction scratch.
constexpr int kRedParts = 64;
long long sweep_adjoint_parts(const AdjArgs &a);
hipError_t launch_dt_adjoint_batch(const DtBatchJob *jobs, int njobs, hipStream_t st);
hipError_t launch_sweep_adjoint_list(int dir, bool mhd, int lim, int limb, const AdjListArgs &L, hipStream_t st);
hipError_t launch_sum_parts_jobs(const SumJob *jobs, int njobs, hipStream_t st);
// dtb_out[0] = dL/d(dt) at fixed ch, dtb_out[1] = dL/d(ch)
hipError_t launch_sweep_adjoint(int dir, bool mhd, const AdjArgs &a, double *dtb_out, hipStream_t st);
hipError_t launch_fold_ghosts(double *lam, long long sv, int sy, int sz, const int ng[3], const int g[3],
                              const int n[3], int nplanes, hipStream_t st);
hipError_t launch_dt_adjoint(const DtAdjArgs &a, unsigned long long *best, double *lam, hipStream_t st);

}  // namespace pk
