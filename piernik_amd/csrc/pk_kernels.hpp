// pk_kernels.hpp -- kernel argument blocks and launchers (host side).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace pk {

// fc_at value of a pencil without that flux point: below every cell a lane
// of a tile can hold (tiles start at most a halo before the pencil).
constexpr int kFcNone = -(1 << 20);

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
    int ptiles;          // tiles over the transverse interior, pencils flattened (d1 fastest)
    int out;             // output cells per tile along the sweep (TS-8 fused, TS-4 per stage)
    int lim, limb;       // limiter for fluid primitives / for B and psi
    double gam, gam_1;
    double c1;           // dt/dx coefficient of the (first) stage
    double c2;           // dt/dx coefficient of the second stage (fused mode)
    double ch;           // GLM cleaning speed
    double psi_fac;      // multiplier applied to psi on output (fused GLM damping), else 1
    double smalld, smallei;
    int use_smalld, use_smallei;
    int ei_flag;         // record negative internal energies (disallow_negatives, sources.F90:425)
    unsigned *flags;     // device: bit 0 negative density, bit 1 negative internal energy
    unsigned long long *dmass;  // device: floor-added mass / cell volume, exact counters (pk_mass.hpp)
    // fused mode only:
    int images;          // 1: also write every output cell's periodic ghost images (one periodic block);
                         // 2 (stream kernels): read each ghost position's interior pre-image instead
                         // and write no images (the ghosts are left stale)
    unsigned long long *dtslots;  // if set: per slot, min of dl_d/c_d (d = x, y, z) and max c of the
                                  // cells written, as FP64 bit patterns (timestep_fluid of the next step)
    int nslots;
    int has_dir[3];
    double dl[3];
    // stage mode only: fine/coarse flux points of the pencils (null when none)
    const int *fc_at;    // [4][fc_npen]: the cell whose right-face flux is replaced (li, ri) or saved (lo, ro);
                         // kFcNone (below every lane's cell) where absent
    const int *fc_slot;  // [4][fc_npen]: the record of each active point in fc_in (li, ri) or fc_out (lo, ro)
    const double *fc_in; // records of fc_rec values: replacement fluxes of the active li, ri points
    double *fc_out;      // records: the saved fluxes of the active lo, ro points
    int fc_npen;         // pencils of the transverse interior (d1 fastest)
    int fc_rec;          // record length: the u-fluxes of every fluid, then [B, psi] with a magnetic field
    int fc_beg;          // this fluid's first u-flux in a record
    int fc_boff;         // first [B, psi] flux in a record
};

struct GhostArgs {
    double *u;
    long long sv;
    int sy, sz;
    int ng[3], g[3], n[3];
    int plane0, nplanes;
};

// Fortran AoS component c of nc <-> device plane planes[c].
struct PlaneMap {
    int nc;
    int planes[16];
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

// Per-step scalars of the batched sweep, in device memory so that a captured
// graph replays every step unchanged.
struct StepScalars {
    double dt;       // the step's dt; the sweep forms (rk_coef*dt)/dl per block
    double ch;       // GLM cleaning speed
    double psi_fac;  // GLM damping factor applied by the last sweep of each substep
    double pad;
};

// One persistent launch over a block list: workgroups take tiles from a
// device queue until it is empty.  Tile `item` belongs to block items[item]
// and is that block's tile item - first[block].
struct ListArgs {
    const SweepArgs *args;   // per block: geometry, pointers (src = buffer 0), tile shape
    const unsigned *first;   // per block: first item (nblocks + 1 entries)
    const unsigned *items;   // per item: block index
    unsigned nitems;
    unsigned *queue;         // item counter, zero before the launch
    int parity;              // 1: read buffer 1, write buffer 0
    int apply_fac;           // multiply psi by step->psi_fac
    const StepScalars *step;      // one for the whole list, or one per block (per_block)
    unsigned long long *dtslots;  // if set: fold timestep_fluid of the cells written into these slots
    int nslots;                   // slots (per block when per_block)
    int per_block;                // independent instances: each block has its own dt, ch, psi factor
                                  // and its own nslots dt slots
    const unsigned *bpos;         // per_block: each block's position among the instances (null: its index)
    int no_wrap;                  // blocks with images == 2 write their images instead (ghosts kept current)
    unsigned grab;                // consecutive items taken per queue fetch (0 or 1: one)
};

// One device-to-device copy of a batch (k_copy_blocks).
struct CopyJob {
    const double *src;
    double *dst;
    long long n;  // doubles
};

// One face of a ghost exchange (k_exchange_faces): the box ext[0] x ext[1]
// x ext[2] of every plane, from the neighbour's state at sorg to this
// block's at dorg (ghosted coordinates, x fastest; src and dst point at
// cell (0, 0, 0) of plane 0).
struct FaceJob {
    const double *src;
    double *dst;
    long long ssv, dsv;  // plane strides
    int ssy, ssz, dsy, dsz;
    int sorg[3], dorg[3], ext[3];
    int nplanes;
    int wg0;             // the job's first workgroup in its launch (kFaceChunk values each)
};
// Values one workgroup of k_exchange_faces copies.
constexpr int kFaceChunk = 4096;

hipError_t launch_sweep(int dir, bool mhd, bool fused, const SweepArgs &a, hipStream_t st);
// Workgroup size that selects the stream form of the persistent list sweep
// (k_stream_list: one-wave workgroups, each list item a range of a block's
// stream of a.out positions, a.TP pencils per lane group: 1 along x, 8 along
// y and z).
constexpr int kStreamListNt = 64;
// Pencils per lane group of the list stream along direction d.
constexpr int stream_list_tp(int d) { return d == 0 ? 1 : 8; }
// The stream kernels store a fluid's planes of sv cells through raw buffer
// resources over three planes each, whose byte ranges (24 sv) must stay
// below 2^31.
constexpr bool stream_stores_fit(long long sv) { return sv >= 0 && 24 * sv < (1ll << 31); }
// Persistent batched fused sweep with workgroups of nt lanes; grid = 0 picks
// the number of workgroups the device holds at once.
hipError_t launch_sweep_list(int dir, bool mhd, int lim, int limb, const ListArgs &L, int nt, int grid,
                             hipStream_t st);
int sweep_list_resident(int dir, bool mhd, int lim, int limb, int nt);
hipError_t launch_fill_items(const unsigned *first, int nblocks, unsigned *items, hipStream_t st);
// per-block dt slots -> per-block (min dl_x/c, min dl_y/c, min dl_z/c, max c)
hipError_t launch_dtslots_final_blocks(const unsigned long long *slots, int nslots, int nblocks, double *out,
                                       hipStream_t st);
hipError_t launch_copy_blocks(const CopyJob *jobs, int njobs, hipStream_t st);
// the faces of one exchange; maxvals = the largest ext product times nplanes
// jobs[wgjob[w]] is the job of workgroup w (nwg workgroups)
hipError_t launch_exchange_faces(const FaceJob *jobs, const int *wgjob, int nwg, hipStream_t st);
hipError_t launch_zero_u32(unsigned *p, int n, hipStream_t st);
size_t sweep_lds_bytes(bool mhd, int TP, int TS);
hipError_t launch_fill_ghosts(int dir, const GhostArgs &a, hipStream_t st);
hipError_t launch_scale(double *x, long long n, double fac, hipStream_t st);
int dt_partial_blocks(const DtArgs &a);
hipError_t launch_dtslots_init(unsigned long long *slots, int nslots, hipStream_t st);
hipError_t launch_dtslots_final(const unsigned long long *slots, int nslots, double *out, hipStream_t st);
hipError_t launch_dt(const DtArgs &a, double *part, double *out, hipStream_t st);
hipError_t launch_aos_to_soa(const double *aos, const PlaneMap &m, double *soa, long long sv, const int ng[3],
                             int sy, int sz, hipStream_t st);
hipError_t launch_soa_to_aos(const double *soa, long long sv, const PlaneMap &m, double *aos, const int ng[3],
                             int sy, int sz, hipStream_t st);

}  // namespace pk

namespace pk_fast {
// the sweeps of pk_kernels_fast.hip (FMA contraction, fast division)
hipError_t launch_sweep(int dir, bool mhd, bool fused, const pk::SweepArgs &a, hipStream_t st);
hipError_t launch_sweep_list(int dir, bool mhd, int lim, int limb, const pk::ListArgs &L, int nt, int grid,
                             hipStream_t st);
int sweep_list_resident(int dir, bool mhd, int lim, int limb, int nt);
}  // namespace pk_fast
