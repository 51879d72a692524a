// pk_unsplit.hpp -- host/device interface of the unsplit Riemann stage
// (RIEMANN_UNSPLIT, src/scheme/unsplit_schemes/).
#pragma once
#include <hip/hip_runtime.h>

namespace pk {

struct UnsplitArgs {
    const double *src;   // the stage's input state: the fluxes are computed from it
    const double *base;  // the state the flux differences are added to (u)
    double *dst;         // output state (uh at the first stage, u at the last; may equal base)
    long long sv;        // plane stride (elements)
    int sy, sz;
    int ng[3], g[3], n[3];
    int has_dir[3];
    double dl[3];
    int lim, limb;
    double gam, gam_1;
    double ch;           // GLM cleaning speed
    double dt, rk;       // dt and rk_coef of the stage
    double psi_fac;      // fused passes: psi of the finished state times this (glmdamping after the stages)
    int last;            // last stage: care_for_positives runs
    double smalld, smallei;
    int use_smalld, use_smallei;
    int ei_flag;         // record negative internal energies (disallow_negatives)
    unsigned *flags;     // device: bit 0 negative density, bit 1 negative internal energy
    unsigned long long *dmass;  // device: floor-added mass / cell volume, exact counters (pk_mass.hpp)
};

// Interface fluxes along dir of the transverse-interior pencils into F (tiled)
// (the state's planes and strides; row c holds the flux at c-1/2, rows
// g..g+n written).
hipError_t launch_unsplit_flux(int dir, bool mhd, const UnsplitArgs &a, double *F, hipStream_t st);
// The flux differences of every active direction and the source passes,
// over the interior cells; F[d] from launch_unsplit_flux (unused if d inactive).
hipError_t launch_unsplit_update(bool mhd, const UnsplitArgs &a, const double *const F[3], hipStream_t st);
// Flux along dir and its difference added in one pass (src != dst): the
// running sum starts from base when `first`, is kept in dst between
// directions, and gets the source passes and the periodic ghost images when
// `last` (first and last active direction of the stage).  Same result as the
// two launchers above followed by the ghost refresh.
hipError_t launch_unsplit_fused(int dir, bool mhd, const UnsplitArgs &a, int first, int last, hipStream_t st);

}  // namespace pk

// The same kernels built with FMA contraction and fast division and square
// root (pk_unsplit_fast.hip).
namespace pk_fast {
hipError_t launch_unsplit_flux(int dir, bool mhd, const pk::UnsplitArgs &a, double *F, hipStream_t st);
hipError_t launch_unsplit_update(bool mhd, const pk::UnsplitArgs &a, const double *const F[3], hipStream_t st);
hipError_t launch_unsplit_fused(int dir, bool mhd, const pk::UnsplitArgs &a, int first, int last, hipStream_t st);
}  // namespace pk_fast
