// pk_kernels.hpp -- kernel argument blocks and launchers (host side).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace pk {

struct SweepArgs {
    const double *src;   // state read with halo (current stage input)
    const double *base;  // U0 of the update (start-of-step state); == src at the first stage
    double *dst;         // output state (never aliases src)
    long long sv;        // plane stride (elements)
    int sy, sz;          // y and z strides (elements), x stride is 1
    int ng[3];           // ghost-padded extents
    int g[3];            // ghost layers per direction (0 where the direction is inactive)
    int n[3];            // interior extents
    int TP, TS;          // tile: TP pencils x TS cells along the sweep
    int chunks;          // tiles along the sweep direction
    int tiles1;          // tiles along the fastest transverse direction
    int out;             // output cells per tile along the sweep (TS-8 fused, TS-4 per stage)
    int lim, limb;       // limiter for fluid primitives / for B and psi
    double gam, gam_1;
    double c1;           // dt/dx coefficient of the (first) stage
    double c2;           // dt/dx coefficient of the second stage (fused mode)
    double ch;           // GLM cleaning speed
    double psi_fac;      // multiplier applied to psi on output (fused GLM damping), else 1
    double smalld, smallei;
    int use_smalld, use_smallei;
    unsigned *flags;     // device: bit 0 negative density, bit 1 negative internal energy
    double *dmass;       // device: floor-added mass / cell volume
    // fused mode only:
    int images;          // also write every output cell's periodic ghost images (one periodic block)
    unsigned long long *dtslots;  // if set: per slot, min of dl_d/c_d (d = x, y, z) and max c of the
                                  // cells written, as FP64 bit patterns (timestep_fluid of the next step)
    int nslots;
    int has_dir[3];
    double dl[3];
};

struct GhostArgs {
    double *u;
    long long sv;
    int sy, sz;
    int ng[3], g[3], n[3];
    int plane0, nplanes;
};

struct DtArgs {
    const double *u;
    long long sv;
    int sy, sz;
    int ng[3], g[3], n[3];
    int plane0;    // rho plane of the fluid
    int bplane;    // Bx plane (magnetized fluid)
    int magnetized;
    int has_dir[3];
    double dl[3];
    double gam, gam_1;
};

hipError_t launch_sweep(int dir, bool mhd, bool fused, const SweepArgs &a, hipStream_t st);
size_t sweep_lds_bytes(bool mhd, int TP, int TS);
hipError_t launch_fill_ghosts(int dir, const GhostArgs &a, hipStream_t st);
hipError_t launch_scale(double *x, long long n, double fac, hipStream_t st);
int dt_partial_blocks(const DtArgs &a);
hipError_t launch_dtslots_init(unsigned long long *slots, int nslots, hipStream_t st);
hipError_t launch_dtslots_final(const unsigned long long *slots, int nslots, double *out, hipStream_t st);
hipError_t launch_dt(const DtArgs &a, double *part, double *out, hipStream_t st);
hipError_t launch_aos_to_soa(const double *aos, int nc, double *soa, long long sv, int plane0, const int ng[3],
                             int sy, int sz, hipStream_t st);
hipError_t launch_soa_to_aos(const double *soa, long long sv, int plane0, int nc, double *aos, const int ng[3],
                             int sy, int sz, hipStream_t st);

}  // namespace pk

namespace pk_fast {
// the sweep of pk_kernels_fast.hip (FMA contraction, fast division)
hipError_t launch_sweep(int dir, bool mhd, bool fused, const pk::SweepArgs &a, hipStream_t st);
}  // namespace pk_fast
