// pk_capi.hip -- host side of the C-ABI declared in include/piernik_hip.h.
//
// Keeps the per-run configuration (the reference's flind / &NUMERICAL_SETUP
// globals), a registry of grid blocks with device mirrors, and drives the
// kernels of pk_kernels.hip on one HIP stream.  The stepping driver
// (pk_fluid_update, pk_run) restates fluid_update_full / make_3sweeps /
// sweep / time_step (src/scheme/fluidupdate.F90:99-232,
// src/scheme/sweeps.F90:114-279, src/base/timestep.F90:112-240) for one
// periodic block with cflcontrol='warn'; the list driver (pk_*_all) does the
// same for the whole registered block list, every block periodic on itself,
// with one persistent launch per sweep and fluid, captured in a HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <set>
#include <unordered_map>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "../../include/piernik_hip.h"
#include "pk_adjoint_kernels.hpp"
#include "pk_kernels.hpp"
#include "pk_unsplit.hpp"
#include "pk_datafields.hpp"
#include "pk_mass.hpp"
#include "pk_xfer.hpp"

namespace {

struct Block {
    int64_t id = 0;
    int n[3] = {1, 1, 1}, nb = 0, g[3] = {0, 0, 0}, ng[3] = {1, 1, 1};
    double dl[3] = {1, 1, 1};
    long long sv = 0;
    int sy = 0, sz = 0;
    double *buf[2] = {nullptr, nullptr};
    int xo = 0;  // leading offset that aligns every interior row start
    int cur = 0;
    int uh = -1;  // buffer holding uh after a first stage, -1 when none
    double *h_u = nullptr, *h_uh = nullptr, *h_b = nullptr, *h_bh = nullptr, *h_psi = nullptr,
           *h_psih = nullptr;
    unsigned *d_flags = nullptr;
    unsigned long long *d_mass = nullptr;  // per fluid: the floor's mass counters (pk_mass.hpp)
    double *d_dtout = nullptr;
    unsigned long long *d_dtslots = nullptr;  // timestep reduction written by the last sweep of a step
    bool dt_pending = false;                  // d_dtout holds timestep_fluid of the current state
    bool ghosts_stale = false;                // stream sweeps read the periodic interior and left the ghosts old
    // reverse mode (pk_adjoint_setup): saved states and the adjoint state
    std::vector<double *> slots;
    double *lam[2] = {nullptr, nullptr};
    int lcur = 0;
    double *d_dtb_part = nullptr;
    double *d_chb_part = nullptr;
    long long dtb_parts = 0;
    double *d_dtb = nullptr;
    unsigned long long *d_best = nullptr;
    // the block whose interior adjoins each ghost face (pk_set_neighbour):
    // nbr[d][0] low side, nbr[d][1] high side; the block's own id when it is
    // periodic on itself
    int64_t nbr[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    // unsplit stages (pk_unsplit_stage): interface fluxes of x, y, z, each
    // with the state's planes and strides
    double *d_flux = nullptr;
    // the state at the start of a step that may be redone (cflcontrol 'redo')
    double *backup = nullptr;
    unsigned saved_flags = 0u;
    unsigned long long saved_mass[2 * pk::kMassDigits] = {};
    // fine/coarse flux points (pk_register_fcflx): host arrays per direction,
    // kind (0 finebnd, 1 coarsebnd) and side, and their device staging
    struct FcReg {
        const int32_t *index = nullptr;
        double *uflx = nullptr, *bflx = nullptr;
        int64_t first = 0;
    } fc[3][2][2];
    struct FcStage {
        int *d_tab = nullptr;                      // [2][4][npen]: cell (fc_at), then record slot
        double *d_in = nullptr, *d_out = nullptr;  // records of the active points
        size_t in_cap = 0, out_cap = 0;            // records
        double *h_in = nullptr, *h_out = nullptr;  // pinned staging of the records
        size_t h_in_cap = 0, h_out_cap = 0;
        std::vector<int> tab;                      // the table last uploaded
        std::vector<long long> out_at;             // per saved record: side * 2^40 + reference position
        hipEvent_t in_done = nullptr;              // the last upload from h_in has been read
    } fcs[3];
};

const int kDtSlots = 1024;

pk_config g_cfg;
bool g_init = false;
int g_ei_flag = 1;        // disallow_negatives (global.F90:237): the floors record negative internal energies
constexpr size_t kMassBytes = 2 * pk::kMassDigits * sizeof(unsigned long long);  // both fluids' counters
bool g_mhd = false;       // some fluid is magnetized: B and psi planes exist
int g_nfl = 1;
int g_fplane[2] = {0, 5};  // first plane of each fluid's group
bool g_fmag[2] = {false, false};
int g_magf = -1;           // the magnetized fluid, -1 if none
int g_nplanes = 5;
std::map<int64_t, Block *> g_blocks;
hipStream_t g_stream = nullptr;
std::string g_err;
double *g_pinned = nullptr;  // small pinned buffer for scalar read-backs
double *g_stage = nullptr;   // AoS staging for host transfers, grown on demand
size_t g_stage_bytes = 0;
double *g_part = nullptr;    // partial results of the standalone timestep reduction
pk_xfer::Ring g_ring;        // pinned chunks for host <-> device copies of the caller's arrays

// Tile shapes chosen by pk_autotune, per block shape, direction and physics.
struct TuneKey {
    int n0, n1, n2, dir, kind, lim, limb;
    bool operator<(const TuneKey &o) const
    {
        const int a[7] = {n0, n1, n2, dir, kind, lim, limb}, b[7] = {o.n0, o.n1, o.n2, o.dir, o.kind, o.lim, o.limb};
        for (int i = 0; i < 7; i++)
            if (a[i] != b[i]) return a[i] < b[i];
        return false;
    }
};
std::map<TuneKey, std::pair<int, int>> g_tuned;

TuneKey tune_key(const Block *b, int d, bool mag, int nt = 0)
{
    return TuneKey{b->n[0], b->n[1], b->n[2], d, (mag ? 1 : 0) + (g_cfg.fast_math ? 2 : 0) + 4 * nt, g_cfg.limiter,
                   g_cfg.limiter_b};
}

// Event pairs recorded around sweep launches while profiling is on.
struct ProfRec {
    hipEvent_t start, stop;
    int dir;
    double bytes;
};
bool g_prof = false;
std::vector<ProfRec> g_prof_recs;
std::vector<hipEvent_t> g_event_pool;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                                       \
    do {                                                                                                \
        hipError_t e_ = (x);                                                                            \
        if (e_ != hipSuccess) return fail(PK_EHIP, "%s failed: %s", #x, hipGetErrorString(e_));       \
    } while (0)

#define RCCHK(x)                \
    do {                        \
        int rc_ = (x);          \
        if (rc_) return rc_;    \
    } while (0)

// Address of cell (0,0,0) of plane 0 of state buffer i, and of fluid f's group.
double *state(Block *b, int i) { return b->buf[i] + b->xo; }
double *fluid_state(Block *b, int i, int f) { return state(b, i) + (long long)g_fplane[f] * b->sv; }
int nv_of(int f) { return g_fmag[f] ? 9 : 5; }

void list_invalidate();
void list_touch();
int settle_ghosts();

// Every entry point that takes a block reads or replaces its state, so the
// ghosts a list's stream sweeps left old are refilled first.
int get_block(int64_t id, Block **b)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    auto it = g_blocks.find(id);
    if (it == g_blocks.end()) return fail(PK_ENOBLOCK, "unknown cg_id %lld", (long long)id);
    RCCHK(settle_ghosts());
    *b = it->second;
    return PK_OK;
}

double min_dl(const Block *b)
{
    double m = DBL_MAX;
    for (int d = 0; d < 3; d++)
        if (g_cfg.has_dir[d] && b->dl[d] < m) m = b->dl[d];
    return m;
}

int first_stage() { return g_cfg.integration_order == 1 ? PK_EULER : PK_RK2_1; }
double rk_coef(int istep) { return istep == PK_RK2_1 ? 0.5 : 1.0; }  // constants.F90:237-238

void fill_common(const Block *b, int f, pk::SweepArgs &a)
{
    a.sv = b->sv;
    a.sy = b->sy;
    a.sz = b->sz;
    for (int d = 0; d < 3; d++) {
        a.ng[d] = b->ng[d];
        a.g[d] = b->g[d];
        a.n[d] = b->n[d];
    }
    a.lim = g_cfg.limiter;
    a.limb = g_cfg.limiter_b;
    a.gam = g_cfg.fluid[f].gamma;
    a.gam_1 = g_cfg.fluid[f].gamma - 1.0;
    a.smalld = g_cfg.smalld;
    a.smallei = g_cfg.smallei;
    a.use_smalld = g_cfg.use_smalld;
    a.use_smallei = g_cfg.use_smallei;
    a.ei_flag = g_ei_flag;
    a.flags = b->d_flags;
    a.dmass = b->d_mass + pk::kMassDigits * f;
    a.psi_fac = 1.0;
    a.ch = 0.0;
    a.images = 0;
    a.dtslots = nullptr;
    a.nslots = kDtSlots;
    for (int d = 0; d < 3; d++) {
        a.has_dir[d] = g_cfg.has_dir[d];
        a.dl[d] = b->dl[d];
    }
}

// Tile shape along direction d: TP pencils x TS cells (TP*TS <= 512 lanes),
// halo 4 (fused) or 2 (one stage); pencils of the transverse interior are
// flattened, so the last tile of a block is the only partial one.
void tile_geometry(const Block *b, int d, bool fused, int tp, int ts, int &TP, int &TS, int &out, int &chunks,
                   int &ptiles)
{
    const int maxt = 512;
    const int halo = fused ? 4 : 2;
    const int span = fused ? b->n[d] : b->ng[d];  // cells that receive output
    const int d1 = (d == 0) ? 1 : 0, d2 = (d == 2) ? 1 : 2;
    const int npen = b->n[d1] * b->n[d2];
    if (ts > span + 2 * halo) ts = span + 2 * halo;
    if (ts > maxt) ts = maxt;
    if (tp <= 0) tp = maxt / ts;
    if (tp < 1) tp = 1;
    if (tp > npen) tp = npen;
    while (tp * ts > maxt && tp > 1) tp--;
    if (tp * ts > maxt) ts = maxt / tp;
    if (ts < 2 * halo + 1) ts = 2 * halo + 1;
    TP = tp;
    TS = ts;
    out = ts - 2 * halo;
    chunks = (span + out - 1) / out;
    ptiles = (npen + tp - 1) / tp;
}

// Tile length that selects the march kernel: one lane per whole pencil.
const int kMarch = -1;

// Tile lengths of -2 and below select the stream kernel: a one-wave
// workgroup walks m = -TS iterations of the stream of TP-pencil groups
// (64/TP positions each).
bool is_stream(int ts) { return ts <= -2; }

void stream_geometry(const Block *b, int d, int tp, int m, int &TP, int &out, int &chunks, int &ptiles)
{
    const int d1 = (d == 0) ? 1 : 0, d2 = (d == 2) ? 1 : 2;
    if (tp != 2 && tp != 4 && tp != 8) tp = 1;
    if (m < 2) m = 2;
    TP = tp;
    // ranges of whole pencil groups (L positions each), about m iterations of 64 / tp positions
    const int L = b->n[d] + 8;
    const long long k = std::max(1ll, ((long long)m * (64 / tp) + L / 2) / L);
    out = (int)(k * L);  // positions per range
    const long long npen = (long long)b->n[d1] * b->n[d2];
    const long long ngroups = (npen + tp - 1) / tp;
    chunks = 1;
    ptiles = (int)((ngroups + k - 1) / k);
}

// The march kernel: whole pencils, one lane each, 64-lane workgroups (k_march).
void march_geometry(const Block *b, int d, int &TP, int &out, int &chunks, int &ptiles)
{
    const int d1 = (d == 0) ? 1 : 0, d2 = (d == 2) ? 1 : 2;
    TP = 64;
    ptiles = (b->n[d1] * b->n[d2] + 63) / 64;
    chunks = 1;
    out = b->n[d];
}

// Precedence: the shape given in pk_config, then the one pk_autotune
// measured for this block shape, then 3 x 72 (4-wave workgroups, two per CU).
// A tile length of kMarch selects the march kernel for the fused sweep.
void choose_tile(const Block *b, int d, bool fused, bool mag, int &TP, int &TS, int &out, int &chunks, int &ptiles)
{
    const int *ov = (d == 0) ? g_cfg.tile_x : g_cfg.tile_yz;
    int tp = ov[0], ts = ov[1];
    if ((ts == kMarch || is_stream(ts)) && !fused) ts = 0;
    if (ts == 0 && fused) {
        auto it = g_tuned.find(tune_key(b, d, mag));
        if (it != g_tuned.end()) {
            tp = it->second.first;
            ts = it->second.second;
        }
    }
    if (is_stream(ts) && fused && pk::stream_stores_fit(b->sv)) {
        TS = ts;
        stream_geometry(b, d, tp, -ts, TP, out, chunks, ptiles);
        return;
    }
    if (is_stream(ts)) ts = 0;  // planes too large for the stream's store ranges: the tile kernel
    if (ts == kMarch && fused) {
        TS = kMarch;
        march_geometry(b, d, TP, out, chunks, ptiles);
        return;
    }
    if (ts <= 0) {
        ts = 72;
        if (tp <= 0) tp = 3;
    }
    tile_geometry(b, d, fused, tp, ts, TP, TS, out, chunks, ptiles);
}

hipError_t sweep_kernel(int d, bool mag, bool fused, const pk::SweepArgs &a, hipStream_t st = nullptr)
{
    if (!st) st = g_stream;
    return g_cfg.fast_math ? pk_fast::launch_sweep(d, mag, fused, a, st) : pk::launch_sweep(d, mag, fused, a, st);
}

// ---- scratch of pk_sweep_noalloc -------------------------------------------
// The one entry point that allocates nothing keeps its scratch in a static
// device array: the timestep slots, both fluids' floor-mass counters and the
// floor flags word (the low 32 bits of its last element).  Each GPU holds its
// own copy, at the address hipGetSymbolAddress gives with that GPU current;
// host code uses only that address (the array's name is not a device
// address in host code).
constexpr int kNaSlots = 1024;                            // timestep slots, 4 words each
constexpr int kNaMass = 4 * kNaSlots;                     // first mass counter (2 fluids x kMassDigits)
constexpr int kNaFlags = kNaMass + 2 * pk::kMassDigits;   // the flags word
constexpr int kNaWords = kNaFlags + 1;
static __device__ __align__(256) unsigned long long scratch[kNaWords];
constexpr int kNaMaxDevices = 64;
unsigned long long *g_scratch_dev[kNaMaxDevices] = {};

// Device address of `scratch` on the current GPU, looked up once per GPU.
int scratch_ptr(unsigned long long **out)
{
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNaMaxDevices) return fail(PK_EINVAL, "device %d beyond the %d supported", dev, kNaMaxDevices);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(PK_EHIP, "hipGetSymbolAddress gave no address for the scratch array");
        g_scratch_dev[dev] = (unsigned long long *)p;
    }
    *out = g_scratch_dev[dev];
    return PK_OK;
}

// report[4] = the floor flags, report[5 + f] = fluid f's floor-added mass
// divided by the cell volume (the exact counters rounded once).
__global__ void k_noalloc_report(const unsigned long long *s, int nfl, double *report)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    report[4] = (double)(unsigned)s[kNaFlags];
    for (int f = 0; f < nfl; f++) report[5 + f] = pk::mass_value(s + kNaMass + pk::kMassDigits * f);
}

hipError_t list_kernel(int d, bool mag, const pk::ListArgs &L, int nt, int grid)
{
    return g_cfg.fast_math ? pk_fast::launch_sweep_list(d, mag, g_cfg.limiter, g_cfg.limiter_b, L, nt, grid, g_stream)
                           : pk::launch_sweep_list(d, mag, g_cfg.limiter, g_cfg.limiter_b, L, nt, grid, g_stream);
}

int list_resident(int d, bool mag, int nt)
{
    return g_cfg.fast_math ? pk_fast::sweep_list_resident(d, mag, g_cfg.limiter, g_cfg.limiter_b, nt)
                           : pk::sweep_list_resident(d, mag, g_cfg.limiter, g_cfg.limiter_b, nt);
}

int take_event(hipEvent_t *e)
{
    if (!g_event_pool.empty()) {
        *e = g_event_pool.back();
        g_event_pool.pop_back();
        return PK_OK;
    }
    HIPCHK(hipEventCreate(e));
    return PK_OK;
}

// Launch one sweep kernel (tile shape chosen here), timed while profiling.
int launch_timed(int d, double bytes, const std::function<hipError_t()> &go);

int launch(int d, bool fused, const pk::SweepArgs &a0, Block *b, int f)
{
    pk::SweepArgs a = a0;
    choose_tile(b, d, fused, g_fmag[f], a.TP, a.TS, a.out, a.chunks, a.ptiles);
    const double cells = (double)b->n[0] * b->n[1] * b->n[2];
    return launch_timed(d, cells * nv_of(f) * 8.0 * 2.0, [&]() { return sweep_kernel(d, g_fmag[f], fused, a); });
}

int launch_timed(int d, double bytes, const std::function<hipError_t()> &go)
{
    ProfRec rec{};
    if (g_prof) {
        RCCHK(take_event(&rec.start));
        RCCHK(take_event(&rec.stop));
        rec.dir = d;
        rec.bytes = bytes;
        HIPCHK(hipEventRecord(rec.start, g_stream));
    }
    HIPCHK(go());
    if (g_prof) {
        HIPCHK(hipEventRecord(rec.stop, g_stream));
        g_prof_recs.push_back(rec);
    }
    return PK_OK;
}

// Copy the cells outside the transverse interior (for sweep direction d)
// from src to dst, so that a one-stage call leaves them as the reference
// does (it only writes transverse-interior pencils).
__global__ void k_copy_transverse_ghosts(const double *__restrict__ src, double *__restrict__ dst, long long sv,
                                         int nplanes, int sy, int sz, int d, int nx, int ny, int nz, int gx, int gy,
                                         int gz, int ix, int iy, int iz)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long ncell = (long long)nx * ny * nz;
    if (i >= ncell) return;
    const int x = (int)(i % nx);
    const long long r = i / nx;
    const int y = (int)(r % ny);
    const int z = (int)(r / ny);
    const bool inx = x >= gx && x < gx + ix, iny = y >= gy && y < gy + iy, inz = z >= gz && z < gz + iz;
    bool interior;
    if (d == 0)
        interior = iny && inz;
    else if (d == 1)
        interior = inx && inz;
    else
        interior = inx && iny;
    if (interior) return;
    const long long idx = (long long)x + (long long)sy * y + (long long)sz * z;
    for (int v = 0; v < nplanes; v++) dst[(long long)v * sv + idx] = src[(long long)v * sv + idx];
}

int copy_transverse_ghosts(Block *b, int d, const double *src, double *dst)
{
    const long long ncell = (long long)b->ng[0] * b->ng[1] * b->ng[2];
    hipLaunchKernelGGL(k_copy_transverse_ghosts, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, g_stream, src,
                       dst, b->sv, g_nplanes, b->sy, b->sz, d, b->ng[0], b->ng[1], b->ng[2], b->g[0], b->g[1],
                       b->g[2], b->n[0], b->n[1], b->n[2]);
    HIPCHK(hipGetLastError());
    return PK_OK;
}

// ---- fine/coarse flux points (set_fluxpointers / save_outfluxes,
// grid_container_fcflx.F90:148-235) ------------------------------------

bool has_fc(const Block *b)
{
    for (int d = 0; d < 3; d++)
        for (int k = 0; k < 2; k++)
            for (int s = 0; s < 2; s++)
                if (b->fc[d][k][s].index) return true;
    return false;
}

int no_fc(const Block *b, const char *what)
{
    if (has_fc(b))
        return fail(PK_EINVAL, "block %lld has fine/coarse flux points, which only pk_solve_cg applies (%s)",
                    (long long)b->id, what);
    return PK_OK;
}

// u-fluxes of every fluid (flind%all), then [B, psi] with a magnetic field
int fc_rec() { return 5 * g_nfl + (g_mhd ? 4 : 0); }

// Kernel pencil pf = i1 + n1*i2 over (d1, d2) -> position in the reference's
// face arrays over (pdims(d, ORTHO1), pdims(d, ORTHO2)) (constants.F90:127-128):
// the same pair for x and z sweeps, (z, x) for y sweeps.
long long fc_ref_at(int d, int i1, int i2, int n1, int n2)
{
    return d == 1 ? (long long)i2 + (long long)n2 * i1 : (long long)i1 + (long long)n1 * i2;
}

// Grow a device or pinned record buffer to hold n records of rec values.
int fc_grow(double **p, size_t *cap, size_t n, int rec, bool pinned)
{
    if (n <= *cap && *p) return PK_OK;
    const size_t want = std::max<size_t>(n, 64);
    if (*p) {
        HIPCHK(hipStreamSynchronize(g_stream));  // no kernel or copy still uses the old buffer
        if (pinned)
            (void)hipHostFree(*p);
        else
            (void)hipFree(*p);
        *p = nullptr;
    }
    if (pinned)
        HIPCHK(hipHostMalloc((void **)p, sizeof(double) * want * rec, hipHostMallocDefault));
    else
        HIPCHK(hipMalloc((void **)p, sizeof(double) * want * rec));
    *cap = want;
    return PK_OK;
}

// Stage the points of direction d for the kernels.  Per pencil and point the
// table holds the cell whose right-face flux the point concerns (li: c, ri:
// c-1, lo: c-1, ro: c for the local cell c of its index, active where
// set_fluxpointers finds the index inside the interior) and the point's
// record; only active points have records.  The table goes to the device
// only when it changed (the indices change at regrids); the replacement
// fluxes go every call.  *active: some point is.
int fc_begin(Block *b, int d, bool *active)
{
    *active = false;
    bool reg = false;
    for (int k = 0; k < 2; k++)
        for (int s = 0; s < 2; s++) reg = reg || b->fc[d][k][s].index != nullptr;
    if (!reg) return PK_OK;
    Block::FcStage &F = b->fcs[d];
    const int d1 = (d == 0) ? 1 : 0, d2 = (d == 2) ? 1 : 2;
    const int n1 = b->n[d1], n2 = b->n[d2];
    const long long np = (long long)n1 * n2;
    const int rec = fc_rec(), nvar = 5 * g_nfl;
    const int g = b->g[d], n = b->n[d], ng = b->ng[d];
    std::vector<int> tab((size_t)(8 * np), pk::kFcNone);
    int *at = tab.data(), *slot = tab.data() + 4 * np;
    // count the active points, then fill the records
    long long nin = 0, nout = 0;
    F.out_at.clear();
    std::vector<std::pair<long long, const double *>> ins;  // replacement record sources
    std::vector<const double *> insb;
    for (int kind = 0; kind < 2; kind++)
        for (int side = 0; side < 2; side++) {
            const Block::FcReg &r = b->fc[d][kind][side];
            if (!r.index) continue;
            const int k = 2 * kind + side;  // 0 li, 1 ri, 2 lo, 3 ro
            for (int i2 = 0; i2 < n2; i2++)
                for (int i1 = 0; i1 < n1; i1++) {
                    const long long pf = i1 + (long long)n1 * i2, ra = fc_ref_at(d, i1, i2, n1, n2);
                    const long long idx = r.index[ra];
                    const bool act = (side == 0) ? idx >= r.first : idx <= r.first + n - 1;
                    if (!act) continue;
                    const long long c = idx - r.first + g;
                    const long long cs = (k == 0 || k == 3) ? c : c - 1;
                    if (cs < 0 || cs > ng - 2)
                        return fail(PK_EINVAL, "block %lld: fine/coarse face at index %lld along %d is outside the pencil",
                                    (long long)b->id, idx, d);
                    at[k * np + pf] = (int)cs;
                    if (kind == 0) {
                        slot[k * np + pf] = (int)nin++;
                        ins.push_back({ra, r.uflx});
                        insb.push_back(r.bflx);
                    } else {
                        slot[k * np + pf] = (int)nout++;
                        F.out_at.push_back(((long long)side << 40) + ra);
                    }
                }
        }
    if (nin + nout == 0) return PK_OK;
    if (nin > 0x7fffffffll || nout > 0x7fffffffll) return fail(PK_EINVAL, "too many fine/coarse points");
    if (!F.d_tab) HIPCHK(hipMalloc((void **)&F.d_tab, sizeof(int) * 8 * np));
    if (!F.in_done) HIPCHK(hipEventCreateWithFlags(&F.in_done, hipEventDisableTiming));
    RCCHK(fc_grow(&F.d_in, &F.in_cap, (size_t)nin, rec, false));
    RCCHK(fc_grow(&F.d_out, &F.out_cap, (size_t)nout, rec, false));
    HIPCHK(hipEventSynchronize(F.in_done));  // the previous call's upload from h_in has been read
    RCCHK(fc_grow(&F.h_in, &F.h_in_cap, (size_t)nin, rec, true));
    RCCHK(fc_grow(&F.h_out, &F.h_out_cap, (size_t)nout, rec, true));
    for (long long s = 0; s < nin; s++) {
        double *dst = F.h_in + s * rec;
        const long long ra = ins[(size_t)s].first;
        memcpy(dst, ins[(size_t)s].second + ra * nvar, sizeof(double) * nvar);
        if (g_mhd) memcpy(dst + nvar, insb[(size_t)s] + ra * 4, sizeof(double) * 4);
    }
    if (F.tab != tab) {
        F.tab.swap(tab);
        HIPCHK(hipMemcpyAsync(F.d_tab, F.tab.data(), sizeof(int) * 8 * np, hipMemcpyHostToDevice, g_stream));
    }
    if (nin) HIPCHK(hipMemcpyAsync(F.d_in, F.h_in, sizeof(double) * nin * rec, hipMemcpyHostToDevice, g_stream));
    HIPCHK(hipEventRecord(F.in_done, g_stream));
    *active = true;
    return PK_OK;
}

void fc_args(const Block *b, int d, int f, pk::SweepArgs &a)
{
    const int d1 = (d == 0) ? 1 : 0, d2 = (d == 2) ? 1 : 2;
    const long long np = (long long)b->n[d1] * b->n[d2];
    a.fc_at = b->fcs[d].d_tab;
    a.fc_slot = b->fcs[d].d_tab + 4 * np;
    a.fc_in = b->fcs[d].d_in;
    a.fc_out = b->fcs[d].d_out;
    a.fc_npen = (int)np;
    a.fc_rec = fc_rec();
    a.fc_beg = 5 * f;  // flind order: fluid f's u-fluxes at 5f..5f+4
    a.fc_boff = 5 * g_nfl;
}

// save_outfluxes: the saved fluxes of the active points into the
// registered coarsebnd arrays.
int fc_end(Block *b, int d)
{
    Block::FcStage &F = b->fcs[d];
    const size_t nout = F.out_at.size();
    if (!nout) return PK_OK;
    const int rec = fc_rec(), nvar = 5 * g_nfl;
    HIPCHK(hipMemcpyAsync(F.h_out, F.d_out, sizeof(double) * nout * rec, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    for (size_t s = 0; s < nout; s++) {
        const int side = (int)(F.out_at[s] >> 40);
        const long long ra = F.out_at[s] & ((1ll << 40) - 1);
        const Block::FcReg &r = b->fc[d][1][side];
        const double *src = F.h_out + s * rec;
        memcpy(r.uflx + ra * nvar, src, sizeof(double) * nvar);
        if (g_mhd) memcpy(r.bflx + ra * 4, src + nvar, sizeof(double) * 4);
    }
    return PK_OK;
}

int do_stage(Block *b, int ddim, int istep, double dt, double ch)
{
    b->dt_pending = false;
    const bool first = istep == first_stage();
    if (!first && (b->uh < 0 || b->uh == b->cur))
        return fail(PK_ESTATE, "stage %d called without a preceding first stage", istep);
    const int src_i = b->cur, dst_i = first ? 1 - b->cur : b->uh;
    bool fc = false;
    RCCHK(fc_begin(b, ddim, &fc));
    for (int f = 0; f < g_nfl; f++) {
        pk::SweepArgs a{};
        fill_common(b, f, a);
        a.ch = ch;
        a.c1 = (rk_coef(istep) * dt) / b->dl[ddim];
        a.c2 = a.c1;
        a.src = fluid_state(b, src_i, f);
        a.base = first ? a.src : fluid_state(b, dst_i, f);
        a.dst = fluid_state(b, dst_i, f);
        if (fc) fc_args(b, ddim, f, a);
        RCCHK(launch(ddim, false, a, b, f));
    }
    RCCHK(copy_transverse_ghosts(b, ddim, state(b, src_i), state(b, dst_i)));
    if (fc) RCCHK(fc_end(b, ddim));
    if (first) {
        b->uh = b->cur;
        b->cur = 1 - b->cur;
    } else {
        b->cur = b->uh;
        b->uh = -1;
    }
    return PK_OK;
}

// One directional sweep, both RK2 stages, every fluid.  images: write the
// periodic ghost images too (the block is periodic on itself); want_dt:
// reduce the next step's timestep_fluid from the cells written.  With
// src/dst (whole state buffers, one fluid) the sweep reads src and writes
// dst instead of the block's ping-pong pair, whose state is left alone.
int do_sweep(Block *b, int ddim, double dt, double ch, double psi_fac, int images = 0, bool want_dt = false,
             double *src = nullptr, double *dst = nullptr)
{
    RCCHK(no_fc(b, "a whole sweep runs both stages in one pass"));
    if (g_cfg.integration_order == 1) {
        RCCHK(do_stage(b, ddim, PK_EULER, dt, ch));
        if (psi_fac != 1.0 && g_mhd) {
            HIPCHK(pk::launch_scale(fluid_state(b, b->cur, g_magf) + 8 * b->sv, (long long)b->sz * b->ng[2], psi_fac,
                                    g_stream));
        }
        return PK_OK;
    }
    if (b->g[ddim] < 4) return fail(PK_EINVAL, "fused RK2 sweep needs nb >= 4 (nb = %d)", b->g[ddim]);
    if (want_dt) HIPCHK(pk::launch_dtslots_init(b->d_dtslots, kDtSlots, g_stream));
    for (int f = 0; f < g_nfl; f++) {
        pk::SweepArgs a{};
        fill_common(b, f, a);
        a.ch = ch;
        a.c1 = (0.5 * dt) / b->dl[ddim];
        a.c2 = (1.0 * dt) / b->dl[ddim];
        a.psi_fac = psi_fac;
        a.src = src ? src + b->xo + (long long)g_fplane[f] * b->sv : fluid_state(b, b->cur, f);
        a.base = a.src;
        a.dst = dst ? dst + b->xo + (long long)g_fplane[f] * b->sv : fluid_state(b, 1 - b->cur, f);
        a.images = images;
        a.dtslots = want_dt ? b->d_dtslots : nullptr;
        RCCHK(launch(ddim, true, a, b, f));
    }
    if (images == 2) b->ghosts_stale = true;
    if (src || dst) return PK_OK;
    b->cur = 1 - b->cur;
    b->uh = -1;
    b->dt_pending = false;
    if (want_dt) {
        HIPCHK(pk::launch_dtslots_final(b->d_dtslots, kDtSlots, b->d_dtout, g_stream));
        b->dt_pending = true;
    }
    return PK_OK;
}

int do_fill(Block *b)
{
    pk::GhostArgs a{};
    a.u = state(b, b->cur);
    a.sv = b->sv;
    a.sy = b->sy;
    a.sz = b->sz;
    for (int d = 0; d < 3; d++) {
        a.ng[d] = b->ng[d];
        a.g[d] = b->g[d];
        a.n[d] = b->n[d];
    }
    a.plane0 = 0;
    a.nplanes = g_nplanes;
    for (int d = 0; d < 3; d++)
        if (b->g[d] > 0) HIPCHK(pk::launch_fill_ghosts(d, a, g_stream));
    b->ghosts_stale = false;
    return PK_OK;
}

// timestep_fluid of every fluid of one block, min over fluids; with
// cfl > 0 the min commutes with the multiplication by cfl, so one combined
// reduction over all fluids' cells gives the same dt.
int enqueue_dt(Block *b, double *out)
{
    for (int f = 0; f < g_nfl; f++) {
        pk::DtArgs a{};
        a.u = state(b, b->cur);
        a.sv = b->sv;
        a.sy = b->sy;
        a.sz = b->sz;
        for (int d = 0; d < 3; d++) {
            a.ng[d] = b->ng[d];
            a.g[d] = b->g[d];
            a.n[d] = b->n[d];
            a.has_dir[d] = g_cfg.has_dir[d];
            a.dl[d] = b->dl[d];
        }
        a.plane0 = g_fplane[f];
        a.bplane = g_magf >= 0 ? g_fplane[g_magf] + 5 : 0;
        a.magnetized = g_fmag[f];
        a.gam = g_cfg.fluid[f].gamma;
        a.gam_1 = g_cfg.fluid[f].gamma - 1.0;
        HIPCHK(pk::launch_dt(a, g_part, out + 4 * f, g_stream));
    }
    return PK_OK;
}

void combine_dt(const double *v, int nsets, double *dt_out, double *c_out)
{
    double dt = DBL_MAX, c = 0.0;
    for (int f = 0; f < nsets; f++) {
        double m = v[4 * f + 0];
        if (v[4 * f + 1] < m) m = v[4 * f + 1];
        if (v[4 * f + 2] < m) m = v[4 * f + 2];
        const double dtf = g_cfg.cfl * m;  // timestep.F90:492
        if (dtf < dt) dt = dtf;
        if (v[4 * f + 3] > c) c = v[4 * f + 3];
    }
    *dt_out = dt;
    *c_out = c;
}

int do_timestep(Block *b, double *dt_out, double *c_out)
{
    if (b->dt_pending) {
        // reduced by the last sweep from the cells it wrote (same formula, same state)
        HIPCHK(hipMemcpyAsync(g_pinned, b->d_dtout, 4 * sizeof(double), hipMemcpyDeviceToHost, g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
        combine_dt(g_pinned, 1, dt_out, c_out);
        return PK_OK;
    }
    RCCHK(enqueue_dt(b, b->d_dtout));  // d_dtout holds 4 doubles per fluid
    HIPCHK(hipMemcpyAsync(g_pinned, b->d_dtout, 4 * g_nfl * sizeof(double), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    combine_dt(g_pinned, g_nfl, dt_out, c_out);
    return PK_OK;
}

// update_chspeed with ch_grid (hdc.F90:112-122) and the glmdamping factor
// (hdc.F90:313-331) over blocks with the given smallest cell sizes.
void glm_scalars(const std::vector<double> &mdls, double dt, double *ch_out, double *fac_out)
{
    double ch = DBL_MAX;
    for (double mdl : mdls) {
        const double c = (g_cfg.cfl_glm * mdl) / dt;
        if (c < ch) ch = c;
    }
    double f = 0.0;
    for (double mdl : mdls) {
        const double ff = (g_cfg.glm_alpha * ch) / (mdl / dt);
        if (ff > f) f = ff;
    }
    *ch_out = ch;
    *fac_out = std::exp(-f);
}

// The fused sweep along every active direction of every fluid of b runs the
// stream kernel (its shape in pk_config, tuned or given): then its sweeps may
// read ghost positions from their periodic pre-images and leave the ghosts
// old (images == 2), to be refilled once by the caller.
bool stream_all_dirs(const Block *b)
{
    for (int f = 0; f < g_nfl; f++)
        for (int d = 0; d < 3; d++) {
            if (!g_cfg.has_dir[d]) continue;
            int TP, TS, out, chunks, ptiles;
            choose_tile(b, d, true, g_fmag[f], TP, TS, out, chunks, ptiles);
            if (!is_stream(TS)) return false;
        }
    return true;
}

// One step of fluid_update.  allow_wrap: the sweeps may leave the periodic
// ghosts old when every one of them is a stream sweep (b->ghosts_stale; the
// caller refills them with do_fill before they are read).
int do_fluid_update(Block *b, double dt, bool allow_wrap = false)
{
    double ch = 0.0, fac = 1.0;
    if (g_mhd) glm_scalars({min_dl(b)}, dt, &ch, &fac);
    const int images = (allow_wrap && g_cfg.integration_order == 2 && stream_all_dirs(b)) ? 2 : 1;
    for (int half = 0; half < 2; half++) {
        int last = -1;
        for (int s = 0; s < 3; s++) {
            const int d = half == 0 ? s : 2 - s;
            if (g_cfg.has_dir[d]) last = d;
        }
        for (int s = 0; s < 3; s++) {
            const int d = half == 0 ? s : 2 - s;
            if (!g_cfg.has_dir[d]) continue;
            if (g_cfg.integration_order == 2) {
                // the sweep refreshes the periodic ghosts itself; the last one of
                // the step also reduces the next step's timestep
                RCCHK(do_sweep(b, d, dt, ch, (d == last) ? fac : 1.0, images, half == 1 && d == last));
            } else {
                RCCHK(do_sweep(b, d, dt, ch, (d == last) ? fac : 1.0));
                RCCHK(do_fill(b));
            }
        }
    }
    return PK_OK;
}

int alloc_block(Block *b)
{
    // interior rows start on a 128 B line for wide blocks, 64 B for narrow ones
    const int al = b->ng[0] >= 128 ? 16 : 8;
    b->xo = (al - b->g[0] % al) % al;
    const int pitch = (b->xo + b->ng[0] + al - 1) / al * al;
    b->sy = pitch;
    b->sz = pitch * b->ng[1];
    long long sv = (long long)b->xo + (long long)b->sz * b->ng[2];
    sv = (sv + 31) / 32 * 32;
    b->sv = sv;
    const size_t bytes = sizeof(double) * (size_t)sv * g_nplanes;
    for (int i = 0; i < 2; i++) {
        HIPCHK(hipMalloc((void **)&b->buf[i], bytes));
        HIPCHK(hipMemsetAsync(b->buf[i], 0, bytes, g_stream));
    }
    HIPCHK(hipMalloc((void **)&b->d_flags, sizeof(unsigned)));
    HIPCHK(hipMalloc((void **)&b->d_mass, kMassBytes));
    HIPCHK(hipMemsetAsync(b->d_flags, 0, sizeof(unsigned), g_stream));
    HIPCHK(hipMemsetAsync(b->d_mass, 0, kMassBytes, g_stream));
    HIPCHK(hipMalloc((void **)&b->d_dtout, sizeof(double) * 8));
    HIPCHK(hipMalloc((void **)&b->d_dtslots, sizeof(unsigned long long) * 4 * kDtSlots));
    return PK_OK;
}

void free_adjoint(Block *b)
{
    for (double *p : b->slots)
        if (p) (void)hipFree(p);
    b->slots.clear();
    for (int i = 0; i < 2; i++) {
        if (b->lam[i]) (void)hipFree(b->lam[i]);
        b->lam[i] = nullptr;
    }
    if (b->d_dtb_part) (void)hipFree(b->d_dtb_part);
    if (b->d_chb_part) (void)hipFree(b->d_chb_part);
    if (b->d_dtb) (void)hipFree(b->d_dtb);
    if (b->d_best) (void)hipFree(b->d_best);
    b->d_dtb_part = nullptr;
    b->d_chb_part = nullptr;
    b->d_dtb = nullptr;
    b->d_best = nullptr;
    b->dtb_parts = 0;
}

void free_block(Block *b)
{
    free_adjoint(b);
    for (int i = 0; i < 2; i++)
        if (b->buf[i]) (void)hipFree(b->buf[i]);
    if (b->d_flags) (void)hipFree(b->d_flags);
    if (b->d_mass) (void)hipFree(b->d_mass);
    if (b->d_dtout) (void)hipFree(b->d_dtout);
    if (b->d_dtslots) (void)hipFree(b->d_dtslots);
    if (b->d_flux) (void)hipFree(b->d_flux);
    if (b->backup) (void)hipFree(b->backup);
    for (int d = 0; d < 3; d++) {
        Block::FcStage &f = b->fcs[d];
        if (f.d_tab) (void)hipFree(f.d_tab);
        if (f.d_in) (void)hipFree(f.d_in);
        if (f.d_out) (void)hipFree(f.d_out);
        if (f.h_in) (void)hipHostFree(f.h_in);
        if (f.h_out) (void)hipHostFree(f.h_out);
        if (f.in_done) (void)hipEventDestroy(f.in_done);
    }
    delete b;
}

// The pinned ring, set up on first use; if pinned memory cannot be had the
// transfers fall back to plain copies from the pageable arrays.
bool ring_ready()
{
    if (g_ring.ensure() == hipSuccess) return true;
    (void)hipGetLastError();  // clear the failed allocation's error
    return false;
}

int ensure_stage(size_t bytes)
{
    if (bytes <= g_stage_bytes) return PK_OK;
    HIPCHK(hipStreamSynchronize(g_stream));
    if (g_stage) (void)hipFree(g_stage);
    g_stage = nullptr;
    g_stage_bytes = 0;
    HIPCHK(hipMalloc((void **)&g_stage, bytes));
    g_stage_bytes = bytes;
    return PK_OK;
}

// Plane maps of the reference arrays: u(nvar) with fluid f at components
// 5f..5f+4 (flind%all_fluids(f)%fl%beg), b(3) and psi of the magnetized fluid.
pk::PlaneMap map_u()
{
    pk::PlaneMap m{};
    m.nc = 5 * g_nfl;
    for (int f = 0; f < g_nfl; f++)
        for (int k = 0; k < 5; k++) m.planes[5 * f + k] = g_fplane[f] + k;
    return m;
}
pk::PlaneMap map_b()
{
    pk::PlaneMap m{};
    m.nc = 3;
    for (int k = 0; k < 3; k++) m.planes[k] = g_fplane[g_magf] + 5 + k;
    return m;
}
pk::PlaneMap map_psi()
{
    pk::PlaneMap m{};
    m.nc = 1;
    m.planes[0] = g_fplane[g_magf] + 8;
    return m;
}

// from_device: the source array is in device memory (no host staging)
int upload_array(Block *b, const double *h, const pk::PlaneMap &m, double *dst, bool from_device)
{
    const size_t ncell = (size_t)b->ng[0] * b->ng[1] * b->ng[2];
    const double *src = h;
    if (!from_device) {
        RCCHK(ensure_stage(sizeof(double) * ncell * m.nc));
        if (ring_ready())
            HIPCHK(g_ring.h2d(g_stage, h, sizeof(double) * ncell * m.nc, g_stream));
        else
            HIPCHK(hipMemcpyAsync(g_stage, h, sizeof(double) * ncell * m.nc, hipMemcpyHostToDevice, g_stream));
        src = g_stage;
    }
    HIPCHK(pk::launch_aos_to_soa(src, m, dst, b->sv, b->ng, b->sy, b->sz, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return PK_OK;
}

int download_array(Block *b, double *h, const pk::PlaneMap &m, const double *src, bool to_device)
{
    const size_t ncell = (size_t)b->ng[0] * b->ng[1] * b->ng[2];
    if (to_device) {
        HIPCHK(pk::launch_soa_to_aos(src, b->sv, m, h, b->ng, b->sy, b->sz, g_stream));
    } else {
        RCCHK(ensure_stage(sizeof(double) * ncell * m.nc));
        HIPCHK(pk::launch_soa_to_aos(src, b->sv, m, g_stage, b->ng, b->sy, b->sz, g_stream));
        if (ring_ready())
            HIPCHK(g_ring.d2h(h, g_stage, sizeof(double) * ncell * m.nc, g_stream));
        else
            HIPCHK(hipMemcpyAsync(h, g_stage, sizeof(double) * ncell * m.nc, hipMemcpyDeviceToHost, g_stream));
    }
    HIPCHK(hipStreamSynchronize(g_stream));
    return PK_OK;
}

int upload_state(Block *b, const double *u, const double *bb, const double *psi, double *dst, bool dev = false)
{
    b->dt_pending = false;
    if (u) RCCHK(upload_array(b, u, map_u(), dst, dev));
    if (g_mhd) {
        if (bb) RCCHK(upload_array(b, bb, map_b(), dst, dev));
        if (psi) RCCHK(upload_array(b, psi, map_psi(), dst, dev));
    }
    return PK_OK;
}

int download_state(Block *b, double *u, double *bb, double *psi, const double *src, bool dev = false)
{
    if (u) RCCHK(download_array(b, u, map_u(), src, dev));
    if (g_mhd) {
        if (bb) RCCHK(download_array(b, bb, map_b(), src, dev));
        if (psi) RCCHK(download_array(b, psi, map_psi(), src, dev));
    }
    return PK_OK;
}

// ------------------------------------------------------------- reverse mode
//
// Adjoint of the fused RK2 sweep for hydro fluids on periodic blocks.  The
// state before every sweep of a recorded step is kept in a slot; the
// adjoint state lam (same layout as the state) runs backwards through the
// sweeps, each linearised at its saved input.

int adj_check(Block *b)
{
    if (b->lam[0] == nullptr) return fail(PK_ESTATE, "pk_adjoint_setup has not been called for this block");
    return PK_OK;
}

int slot_ptr(Block *b, int slot, double **p)
{
    if (slot < 0 || slot >= (int)b->slots.size())
        return fail(PK_EINVAL, "slot %d outside 0..%d", slot, (int)b->slots.size() - 1);
    *p = b->slots[slot];
    return PK_OK;
}

size_t state_bytes(const Block *b) { return sizeof(double) * (size_t)b->sv * g_nplanes; }

// tile of the adjoint sweep along d: TS - 16 output cells, TP pencils, <= 256 lanes
void adj_tile(const Block *b, int d, int &TP, int &TS, int &out, int &chunks, int &ptiles)
{
    const int nt = 256;
    const int span = b->n[d];
    const int d1 = (d == 0) ? 1 : 0, d2 = (d == 2) ? 1 : 2;
    const long long npen = (long long)b->n[d1] * b->n[d2];
    long long best = -1;
    for (int k = 1; k <= span; k++) {
        const int o = (span + k - 1) / k;
        const int ts = o + 16;
        if (ts > nt) continue;
        const int tp = nt / ts;
        const long long tiles = (long long)((span + o - 1) / o) * ((npen + tp - 1) / tp);
        if (best < 0 || tiles < best) {
            best = tiles;
            TS = ts;
            TP = tp;
            out = o;
        }
        if (o <= 1) break;
    }
    if (TP > npen) TP = (int)npen;
    chunks = (span + out - 1) / out;
    ptiles = (int)((npen + TP - 1) / TP);
}

// lam <- adjoint of the fused sweep along d linearised at state u (a slot).
// dL/d(dt) at fixed ch and dL/d(ch) land in d_dtb[2*out_slot ..]; with dtb
// set they are also read back and added to *dtb, *chb.
int do_sweep_adjoint(Block *b, const double *u, int d, double dt, double ch, double psi_fac, double *dtb,
                     double *chb, int out_slot = 0)
{
    pk::AdjArgs a{};
    a.u = u + b->xo;
    a.lam_out = b->lam[b->lcur] + b->xo;
    a.lam_in = b->lam[1 - b->lcur] + b->xo;
    a.dtb_part = b->d_dtb_part;
    a.chb_part = b->d_chb_part;
    a.sv = b->sv;
    a.sy = b->sy;
    a.sz = b->sz;
    for (int k = 0; k < 3; k++) {
        a.ng[k] = b->ng[k];
        a.g[k] = b->g[k];
        a.n[k] = b->n[k];
    }
    adj_tile(b, d, a.TP, a.TS, a.out, a.chunks, a.ptiles);
    if (pk::sweep_adjoint_parts(a) > b->dtb_parts) return fail(PK_EINVAL, "adjoint partial-sum buffer too small");
    a.lim = g_cfg.limiter;
    a.limb = g_cfg.limiter_b;
    a.gam = g_cfg.fluid[0].gamma;
    a.gam_1 = g_cfg.fluid[0].gamma - 1.0;
    a.c1 = (0.5 * dt) / b->dl[d];
    a.c2 = (1.0 * dt) / b->dl[d];
    a.dl = b->dl[d];
    a.ch = ch;
    a.psi_fac = psi_fac;
    a.smalld = g_cfg.smalld;
    a.smallei = g_cfg.smallei;
    a.use_smalld = g_cfg.use_smalld;
    a.use_smallei = g_cfg.use_smallei;
    HIPCHK(pk::launch_sweep_adjoint(d, g_mhd, a, b->d_dtb + 2 * out_slot, g_stream));
    b->lcur = 1 - b->lcur;
    if (dtb) {
        HIPCHK(hipMemcpyAsync(g_pinned, b->d_dtb + 2 * out_slot, 2 * sizeof(double), hipMemcpyDeviceToHost,
                              g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
        *dtb += g_pinned[0];
        *chb += g_pinned[1];
    }
    return PK_OK;
}

// ch and the damping factor fluid_update_full uses on one block for dt
void block_glm(Block *b, double dt, double *ch, double *fac)
{
    *ch = 0.0;
    *fac = 1.0;
    if (g_mhd) glm_scalars({min_dl(b)}, dt, ch, fac);
}

// The directions of the sweeps of one step in order (XYZ then ZYX); damp[i]:
// sweep i is the last of its substep (it applies the GLM damping).
std::vector<int> step_dirs(std::vector<bool> *damp = nullptr)
{
    std::vector<int> v;
    if (damp) damp->clear();
    for (int half = 0; half < 2; half++) {
        int last = -1;
        for (int s = 0; s < 3; s++) {
            const int d = half == 0 ? s : 2 - s;
            if (g_cfg.has_dir[d]) last = d;
        }
        for (int s = 0; s < 3; s++) {
            const int d = half == 0 ? s : 2 - s;
            if (!g_cfg.has_dir[d]) continue;
            v.push_back(d);
            if (damp) damp->push_back(d == last);
        }
    }
    return v;
}

// ---------------------------------------------------------------- block list
//
// The whole registered list (ascending cg_id), swept by one persistent
// launch per direction and fluid.  Per block and (direction, fluid) the
// host builds the kernel arguments once; the per-step scalars live in
// device memory (StepScalars), so the step's launches are captured once in a
// HIP graph and replayed.

struct LaunchTab {
    int nt = 256;                       // workgroup lanes of the persistent launch
    std::vector<Block *> blocks;        // the blocks of the launch, in item order
    std::vector<pk::SweepArgs> h_args;  // per block, src = buffer of the block's cur at build time
    std::vector<unsigned> h_first;      // per block first item, nblocks + 1 entries, then each
                                        // block's position in the list (nblocks entries)
    pk::SweepArgs *d_args = nullptr;
    unsigned *d_first = nullptr;        // h_first on the device
    unsigned *d_items = nullptr;
    size_t cap_blocks = 0, cap_items = 0;
    unsigned nitems = 0;
    int grid = 0;
};

const int kQueues = 32;

// The face jobs of one exchange direction over the list.
struct FaceTab {
    std::vector<pk::FaceJob> h;
    pk::FaceJob *d_jobs = nullptr;
    size_t cap = 0;
    std::vector<int> wgjob;  // per workgroup of the launch: its job
    int *d_wgjob = nullptr;
    size_t wcap = 0;
};

// Cuts the jobs of t into workgroups of pk::kFaceChunk values and copies the
// tables to the device.
int finish_faces(FaceTab &t)
{
    t.wgjob.clear();
    for (size_t i = 0; i < t.h.size(); i++) {
        pk::FaceJob &j = t.h[i];
        const long long vals = (long long)j.ext[0] * j.ext[1] * j.ext[2] * j.nplanes;
        if (vals >= (1ll << 31)) return fail(PK_EINVAL, "ghost face of %lld values exceeds one exchange job", vals);
        j.wg0 = (int)t.wgjob.size();
        const long long nwg = (vals + pk::kFaceChunk - 1) / pk::kFaceChunk;
        if ((long long)t.wgjob.size() + nwg >= (1ll << 31)) return fail(PK_EINVAL, "too many ghost faces for one launch");
        t.wgjob.insert(t.wgjob.end(), (size_t)nwg, (int)i);
    }
    if (t.cap < t.h.size()) {
        if (t.d_jobs) (void)hipFree(t.d_jobs);
        t.d_jobs = nullptr;
        HIPCHK(hipMalloc((void **)&t.d_jobs, sizeof(pk::FaceJob) * t.h.size()));
        t.cap = t.h.size();
    }
    if (t.wcap < t.wgjob.size()) {
        if (t.d_wgjob) (void)hipFree(t.d_wgjob);
        t.d_wgjob = nullptr;
        HIPCHK(hipMalloc((void **)&t.d_wgjob, sizeof(int) * t.wgjob.size()));
        t.wcap = t.wgjob.size();
    }
    if (!t.h.empty()) {
        HIPCHK(hipMemcpyAsync(t.d_jobs, t.h.data(), sizeof(pk::FaceJob) * t.h.size(), hipMemcpyHostToDevice, g_stream));
        HIPCHK(hipMemcpyAsync(t.d_wgjob, t.wgjob.data(), sizeof(int) * t.wgjob.size(), hipMemcpyHostToDevice,
                              g_stream));
    }
    return PK_OK;
}

int launch_faces(const FaceTab &t)
{
    HIPCHK(pk::launch_exchange_faces(t.d_jobs, t.d_wgjob, (int)t.wgjob.size(), g_stream));
    return PK_OK;
}

struct ListState {
    bool built = false;
    std::vector<Block *> blocks;
    std::vector<int> cur0;     // each block's cur when the tables were built
    std::vector<double> mdl;   // min(dl) per block
    LaunchTab tab[3][2];
    // per direction and fluid kind: block shapes swept by a second launch of
    // the stream form (pk_autotune_all), the rest by tab's workgroup size
    std::set<std::vector<int>> split_cls[3][2];
    int split_ts[3][2] = {{0, 0}, {0, 0}, {0, 0}};  // their iterations per item set by pk_set_list_split
    LaunchTab stab[3][2];
    int nt_pref[3][2] = {{0, 0}, {0, 0}, {0, 0}};  // workgroup lanes chosen by pk_autotune_all (0: default)
    int ts_force[3][2] = {{0, 0}, {0, 0}, {0, 0}};  // tile length (stream: -iterations) set by pk_set_list_tiles
    unsigned *d_queue = nullptr;
    pk::StepScalars *d_step = nullptr;
    pk::StepScalars *h_step = nullptr;  // pinned
    unsigned long long *d_dtslots = nullptr;
    double *d_dtout = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool fresh = false;  // ghosts current and d_dtout holds the timestep of the current state
    // blocks connected to other blocks (pk_set_neighbour): ghost faces come from
    // the neighbours through an exchange before every sweep instead of the
    // sweep's periodic images
    bool linked = false;
    bool ghosts_stale = false;  // linked list after a step: the ghosts other than the swept faces are old
    FaceTab ex[3][2];           // per direction and buffer parity
    // blocks periodic on themselves: each ghost face from the block's own
    // opposite interior (one launch per direction fills the whole list), and
    // whether some block's stream sweeps read ghost positions from their
    // pre-images and leave the ghosts old (images == 2 in its tables)
    FaceTab self[3];
    bool self_built = false;
    bool wrap_any = false;
    bool wrap_stale = false;
};
ListState g_list;

// Independent instances (an ensemble): the registered blocks advanced
// together, each with its own dt, ch and GLM damping factor.
const int kEnsSlots = 64;  // dt slots per block
struct EnsembleState {
    size_t nblk = 0;
    pk::StepScalars *d_step = nullptr, *h_step = nullptr;  // per block
    unsigned long long *d_dtslots = nullptr;               // per block kEnsSlots slots
    double *d_dtout = nullptr, *h_dtout = nullptr;         // per block 4 values
    std::vector<double> dt_old;
    bool fresh = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    pk::CopyJob *d_jobs = nullptr;
    std::vector<pk::CopyJob> h_jobs;
    // reverse mode
    pk::AdjArgs *d_adj = nullptr, *h_adj = nullptr;  // per block (pinned host copy)
    unsigned *d_afirst[3] = {nullptr, nullptr, nullptr}, *d_aitems[3] = {nullptr, nullptr, nullptr};
    unsigned anitems[3] = {0, 0, 0};
    bool adj_built = false;
    pk::SumJob *d_sum = nullptr;
    std::vector<pk::SumJob> h_sum;
    pk::DtBatchJob *d_dtj = nullptr;
    std::vector<pk::DtBatchJob> h_dtj;
    double *d_mn = nullptr;
    unsigned long long *d_best = nullptr;
    double *d_sums = nullptr, *h_sums = nullptr;  // per block, per sweep: dtb, chb
    unsigned *d_queue = nullptr;
};
EnsembleState g_ens;

void destroy_graph()
{
    if (g_list.exec) (void)hipGraphExecDestroy(g_list.exec);
    if (g_list.graph) (void)hipGraphDestroy(g_list.graph);
    g_list.exec = nullptr;
    g_list.graph = nullptr;
}

void ens_free();

void list_invalidate()
{
    (void)settle_ghosts();  // the tables that know which ghosts are old go away
    g_list.built = false;
    g_list.fresh = false;
    destroy_graph();
    ens_free();
}

void list_touch()
{
    g_list.fresh = false;
    g_ens.fresh = false;
}

void free_tab(LaunchTab &t)
{
    if (t.d_args) (void)hipFree(t.d_args);
    if (t.d_first) (void)hipFree(t.d_first);
    if (t.d_items) (void)hipFree(t.d_items);
    t = LaunchTab();
}

void ens_free()
{
    if (g_stream) (void)hipStreamSynchronize(g_stream);
    if (g_ens.exec) (void)hipGraphExecDestroy(g_ens.exec);
    if (g_ens.graph) (void)hipGraphDestroy(g_ens.graph);
    if (g_ens.d_step) (void)hipFree(g_ens.d_step);
    if (g_ens.h_step) (void)hipHostFree(g_ens.h_step);
    if (g_ens.d_dtslots) (void)hipFree(g_ens.d_dtslots);
    if (g_ens.d_dtout) (void)hipFree(g_ens.d_dtout);
    if (g_ens.h_dtout) (void)hipHostFree(g_ens.h_dtout);
    if (g_ens.d_jobs) (void)hipFree(g_ens.d_jobs);
    if (g_ens.d_adj) (void)hipFree(g_ens.d_adj);
    if (g_ens.h_adj) (void)hipHostFree(g_ens.h_adj);
    for (int d = 0; d < 3; d++) {
        if (g_ens.d_afirst[d]) (void)hipFree(g_ens.d_afirst[d]);
        if (g_ens.d_aitems[d]) (void)hipFree(g_ens.d_aitems[d]);
    }
    if (g_ens.d_sum) (void)hipFree(g_ens.d_sum);
    if (g_ens.d_dtj) (void)hipFree(g_ens.d_dtj);
    if (g_ens.d_mn) (void)hipFree(g_ens.d_mn);
    if (g_ens.d_best) (void)hipFree(g_ens.d_best);
    if (g_ens.d_sums) (void)hipFree(g_ens.d_sums);
    if (g_ens.h_sums) (void)hipHostFree(g_ens.h_sums);
    if (g_ens.d_queue) (void)hipFree(g_ens.d_queue);
    g_ens = EnsembleState();
}

void list_free()
{
    (void)hipStreamSynchronize(g_stream);
    destroy_graph();
    ens_free();
    for (auto &row : g_list.tab)
        for (auto &t : row) free_tab(t);
    for (auto &row : g_list.stab)
        for (auto &t : row) free_tab(t);
    for (auto &t : g_list.self) {
        if (t.d_jobs) (void)hipFree(t.d_jobs);
        if (t.d_wgjob) (void)hipFree(t.d_wgjob);
    }
    if (g_list.d_queue) (void)hipFree(g_list.d_queue);
    if (g_list.d_step) (void)hipFree(g_list.d_step);
    if (g_list.h_step) (void)hipHostFree(g_list.h_step);
    if (g_list.d_dtslots) (void)hipFree(g_list.d_dtslots);
    if (g_list.d_dtout) (void)hipFree(g_list.d_dtout);
    for (auto &row : g_list.ex)
        for (auto &t : row) {
            if (t.d_jobs) (void)hipFree(t.d_jobs);
            if (t.d_wgjob) (void)hipFree(t.d_wgjob);
        }
    g_list = ListState();
}

// Default list tile along d for nt lanes: the split of the pencil into
// chunks that needs the fewest nt-lane tiles; ties go to the longer tile.
// Tiles a workgroup of the persistent list sweep takes per queue fetch: one
// atomic per tile on the shared counter cost 9 % of a 4096-block step; four
// per fetch (or sixteen) remove it.
constexpr unsigned kListGrab = 4;

int list_default_ts(const Block *b, int d, int nt)
{
    if (nt == pk::kStreamListNt) return -16;  // the stream: 16 iterations per item
    const int span = b->n[d];
    const int d1 = (d == 0) ? 1 : 0, d2 = (d == 2) ? 1 : 2;
    const long long npen = (long long)b->n[d1] * b->n[d2];
    long long best = -1;
    int best_ts = 0;
    for (int k = 1; k <= span; k++) {
        const int out = (span + k - 1) / k;
        const int ts = out + 8;
        if (ts > nt) continue;
        const int tp = nt / ts;
        const long long tiles = (long long)((span + out - 1) / out) * ((npen + tp - 1) / tp);
        if (best < 0 || tiles < best) {
            best = tiles;
            best_ts = ts;
        }
        if (out <= 1) break;
    }
    return best_ts > 0 ? best_ts : 9;
}

int list_ts(const Block *b, int d, int f, int nt)
{
    if (nt == pk::kStreamListNt && nt != g_list.nt_pref[d][g_fmag[f]] && g_list.split_ts[d][g_fmag[f]] != 0)
        return g_list.split_ts[d][g_fmag[f]];
    const int forced = g_list.ts_force[d][g_fmag[f]];
    if (forced != 0 && nt == g_list.nt_pref[d][g_fmag[f]]) return forced;
    auto it = g_tuned.find(tune_key(b, d, g_fmag[f], nt));
    if (it != g_tuned.end()) return it->second.second;
    return list_default_ts(b, d, nt);
}

// Kernel arguments and item table of one persistent launch over `blks`.
// Block b's fluid f is swept in the stream form along every active direction
// (the list's workgroup size, or its split shapes).
bool list_wrap_ok(const Block *b, int f)
{
    const int mag = g_fmag[f];
    for (int d = 0; d < 3; d++) {
        if (!g_cfg.has_dir[d]) continue;
        const int nt = g_list.nt_pref[d][mag] > 0 ? g_list.nt_pref[d][mag] : 256;
        if (nt != pk::kStreamListNt && !g_list.split_cls[d][mag].count({b->n[0], b->n[1], b->n[2]})) return false;
    }
    return true;
}

// wrap: 0 every block writes its images; 1 the blocks whose every sweep is a
// stream sweep read their ghost positions from the pre-images (list_wrap_ok);
// 2 every block of a stream table does (timing only: the autotuner's
// candidates, whose results are discarded)
int build_tab(LaunchTab &t, const std::vector<Block *> &blks, int d, int f, int nt,
              const std::vector<int> *ts_override = nullptr, int wrap = 0)
{
    t.nt = nt;
    t.blocks = blks;
    if (blks.empty()) {  // nothing to launch
        t.h_args.clear();
        t.h_first.assign(1, 0u);
        t.nitems = 0;
        return PK_OK;
    }
    t.h_args.resize(blks.size());
    t.h_first.resize(2 * blks.size() + 1);
    {
        std::unordered_map<const Block *, unsigned> at;
        for (size_t i = 0; i < g_list.blocks.size(); i++) at[g_list.blocks[i]] = (unsigned)i;
        for (size_t i = 0; i < blks.size(); i++) {
            auto it = at.find(blks[i]);
            t.h_first[blks.size() + 1 + i] = it != at.end() ? it->second : (unsigned)i;
        }
    }
    unsigned items = 0;
    for (size_t i = 0; i < blks.size(); i++) {
        Block *b = blks[i];
        pk::SweepArgs &a = t.h_args[i];
        a = pk::SweepArgs{};
        fill_common(b, f, a);
        a.src = fluid_state(b, b->cur, f);
        a.base = a.src;
        a.dst = fluid_state(b, 1 - b->cur, f);
        const bool wr = wrap == 2 ? nt == pk::kStreamListNt : (wrap == 1 && list_wrap_ok(b, f));
        a.images = g_list.linked ? 0 : (wr ? 2 : 1);
        const int ts = ts_override ? (*ts_override)[i] : list_ts(b, d, f, nt);
        if (nt == pk::kStreamListNt) {
            // the stream form: items are ranges of -ts iterations of the block's stream
            if (!is_stream(ts)) return fail(PK_EINVAL, "stream list item of %d iterations", -ts);
            if (!pk::stream_stores_fit(b->sv)) return fail(PK_EINVAL, "block planes of %lld cells exceed the stream's buffer range", b->sv);
            a.TS = ts;
            stream_geometry(b, d, pk::stream_list_tp(d), -ts, a.TP, a.out, a.chunks, a.ptiles);
        } else {
            tile_geometry(b, d, true, nt / ts, ts, a.TP, a.TS, a.out, a.chunks, a.ptiles);
            if (a.TP * a.TS > nt) return fail(PK_EINVAL, "tile %d x %d exceeds %d lanes", a.TP, a.TS, nt);
        }
        t.h_first[i] = items;
        const unsigned long long n = (unsigned long long)a.chunks * (unsigned long long)a.ptiles;
        // headroom for the fetches past the end (at most grab per resident workgroup)
        if (items + n > 0xffffffffull - (1ull << 24)) return fail(PK_EINVAL, "block list has too many tiles");
        items += (unsigned)n;
    }
    t.h_first[blks.size()] = items;
    t.nitems = items;
    if (t.cap_blocks < blks.size() + 1) {
        if (t.d_args) (void)hipFree(t.d_args);
        if (t.d_first) (void)hipFree(t.d_first);
        t.d_args = nullptr;
        t.d_first = nullptr;
        HIPCHK(hipMalloc((void **)&t.d_args, sizeof(pk::SweepArgs) * (blks.size() + 1)));
        HIPCHK(hipMalloc((void **)&t.d_first, sizeof(unsigned) * (2 * blks.size() + 1)));
        t.cap_blocks = blks.size() + 1;
    }
    if (t.cap_items < items) {
        if (t.d_items) (void)hipFree(t.d_items);
        t.d_items = nullptr;
        HIPCHK(hipMalloc((void **)&t.d_items, sizeof(unsigned) * items));
        t.cap_items = items;
    }
    HIPCHK(hipMemcpyAsync(t.d_args, t.h_args.data(), sizeof(pk::SweepArgs) * blks.size(), hipMemcpyHostToDevice,
                          g_stream));
    HIPCHK(hipMemcpyAsync(t.d_first, t.h_first.data(), sizeof(unsigned) * (2 * blks.size() + 1), hipMemcpyHostToDevice,
                          g_stream));
    HIPCHK(pk::launch_fill_items(t.d_first, (int)blks.size(), t.d_items, g_stream));
    t.grid = list_resident(d, g_fmag[f], nt);
    if (t.grid <= 0) return fail(PK_EHIP, "no resident workgroups for the list sweep (nt = %d)", nt);
    HIPCHK(hipStreamSynchronize(g_stream));
    return PK_OK;
}

bool block_linked(const Block *b)
{
    for (int d = 0; d < 3; d++)
        if (b->g[d] > 0 && (b->nbr[d][0] != b->id || b->nbr[d][1] != b->id)) return true;
    return false;
}

// o may fill b's ghost faces along d: same faces, ghost width and cell size
int check_face(const Block *b, const Block *o, int d)
{
    const int d1 = (d == 0) ? 1 : 0, d2 = (d == 2) ? 1 : 2;
    if (o->ng[d1] != b->ng[d1] || o->ng[d2] != b->ng[d2])
        return fail(PK_EINVAL, "blocks %lld and %lld do not share a face along %c", (long long)b->id,
                    (long long)o->id, "xyz"[d]);
    if (o->nb != b->nb) return fail(PK_EINVAL, "blocks %lld and %lld differ in nb", (long long)b->id, (long long)o->id);
    for (int q = 0; q < 3; q++)
        if (o->dl[q] != b->dl[q])
            return fail(PK_EINVAL, "blocks %lld and %lld differ in cell size (no fine/coarse boundaries)",
                        (long long)b->id, (long long)o->id);
    return PK_OK;
}

// Face jobs of the exchange along each direction and buffer parity: block b's
// ghost layers on side s come from the g adjoining interior layers of its
// neighbour, over the whole ghosted extent of the other two directions.
int build_faces()
{
    for (int d = 0; d < 3; d++) {
        if (!g_cfg.has_dir[d]) continue;
        for (int p = 0; p < 2; p++) {
            FaceTab &t = g_list.ex[d][p];
            t.h.clear();
            for (Block *b : g_list.blocks) {
                for (int side = 0; side < 2; side++) {
                    auto it = g_blocks.find(b->nbr[d][side]);
                    if (it == g_blocks.end())
                        return fail(PK_ESTATE, "neighbour %lld of block %lld is not registered",
                                    (long long)b->nbr[d][side], (long long)b->id);
                    const Block *o = it->second;
                    RCCHK(check_face(b, o, d));
                    pk::FaceJob j{};
                    j.src = state(const_cast<Block *>(o), o->cur ^ p);
                    j.dst = state(b, b->cur ^ p);
                    j.ssv = o->sv;
                    j.dsv = b->sv;
                    j.ssy = o->sy;
                    j.ssz = o->sz;
                    j.dsy = b->sy;
                    j.dsz = b->sz;
                    for (int q = 0; q < 3; q++) {
                        j.ext[q] = b->ng[q];
                        j.sorg[q] = 0;
                        j.dorg[q] = 0;
                    }
                    j.ext[d] = b->g[d];
                    j.dorg[d] = side == 0 ? 0 : b->g[d] + b->n[d];
                    j.sorg[d] = side == 0 ? o->n[d] : o->g[d];
                    j.nplanes = g_nplanes;
                    t.h.push_back(j);
                }
            }
            RCCHK(finish_faces(t));
        }
    }
    HIPCHK(hipStreamSynchronize(g_stream));
    return PK_OK;
}

// Face jobs that fill every ghost layer of each block of an unlinked list
// from the block's own opposite interior (periodic on itself), x then y then
// z over the full ghosted extents, as k_fill_ghosts does block by block.
int build_self_faces()
{
    g_list.self_built = false;
    for (int d = 0; d < 3; d++) {
        FaceTab &t = g_list.self[d];
        t.h.clear();
        t.wgjob.clear();
        if (!g_cfg.has_dir[d]) continue;
        for (Block *b : g_list.blocks) {
            for (int side = 0; side < 2; side++) {
                pk::FaceJob j{};
                j.src = state(b, b->cur);
                j.dst = state(b, b->cur);
                j.ssv = j.dsv = b->sv;
                j.ssy = j.dsy = b->sy;
                j.ssz = j.dsz = b->sz;
                for (int q = 0; q < 3; q++) {
                    j.ext[q] = b->ng[q];
                    j.sorg[q] = 0;
                    j.dorg[q] = 0;
                }
                j.ext[d] = b->g[d];
                j.dorg[d] = side == 0 ? 0 : b->g[d] + b->n[d];
                j.sorg[d] = side == 0 ? b->n[d] : b->g[d];
                j.nplanes = g_nplanes;
                t.h.push_back(j);
            }
        }
        RCCHK(finish_faces(t));
    }
    HIPCHK(hipStreamSynchronize(g_stream));
    g_list.self_built = true;
    return PK_OK;
}

int exchange(int d, int parity)
{
    return launch_faces(g_list.ex[d][parity]);
}

// every ghost cell of every block of a linked list, x then y then z
int exchange_all()
{
    for (int d = 0; d < 3; d++)
        if (g_cfg.has_dir[d]) RCCHK(exchange(d, 0));
    g_list.ghosts_stale = false;
    return PK_OK;
}

// ghost refresh of the whole list: periodic images, or the exchange
int list_fill()
{
    if (g_list.linked) return exchange_all();
    if (g_list.self_built) {
        for (int d = 0; d < 3; d++) {
            const FaceTab &t = g_list.self[d];
            if (!t.h.empty()) RCCHK(launch_faces(t));
        }
        for (Block *b : g_list.blocks) b->ghosts_stale = false;
    } else {
        for (Block *b : g_list.blocks) RCCHK(do_fill(b));
    }
    g_list.wrap_stale = false;
    return PK_OK;
}

// The ghosts the list's stream sweeps left old (every block of the list, as
// their images would have left them), before anything reads them.
int settle_ghosts()
{
    if (!g_list.wrap_stale || !g_list.built) return PK_OK;
    return list_fill();
}

int list_build()
{
    if (g_list.built) {
        bool same = g_list.blocks.size() == g_blocks.size();
        for (size_t i = 0; same && i < g_list.blocks.size(); i++) same = g_list.blocks[i]->cur == g_list.cur0[i];
        if (same) return PK_OK;
        destroy_graph();
        if (g_ens.exec) (void)hipGraphExecDestroy(g_ens.exec);
        if (g_ens.graph) (void)hipGraphDestroy(g_ens.graph);
        g_ens.exec = nullptr;
        g_ens.graph = nullptr;
    }
    if (g_blocks.empty()) return fail(PK_ESTATE, "no registered blocks");
    if (g_cfg.integration_order != 2) return fail(PK_EINVAL, "the batched list sweep is the fused RK2 sweep");
    g_list.blocks.clear();
    g_list.wrap_any = false;
    g_list.cur0.clear();
    g_list.mdl.clear();
    for (auto &kv : g_blocks) {
        Block *b = kv.second;
        for (int d = 0; d < 3; d++)
            if (g_cfg.has_dir[d] && b->g[d] < 4)
                return fail(PK_EINVAL, "block %lld: the fused RK2 sweep needs nb >= 4", (long long)b->id);
        RCCHK(no_fc(b, "the batched list sweep"));
        g_list.blocks.push_back(b);
        g_list.cur0.push_back(b->cur);
        g_list.mdl.push_back(min_dl(b));
    }
    g_list.linked = false;
    for (Block *b : g_list.blocks) g_list.linked = g_list.linked || block_linked(b);
    g_list.self_built = false;
    if (g_list.linked)
        RCCHK(build_faces());
    else
        RCCHK(build_self_faces());
    for (int d = 0; d < 3; d++) {
        if (!g_cfg.has_dir[d]) continue;
        for (int f = 0; f < g_nfl; f++) {
            const int nt = g_list.nt_pref[d][g_fmag[f]] > 0 ? g_list.nt_pref[d][g_fmag[f]] : 256;
            const std::set<std::vector<int>> &sc = g_list.split_cls[d][g_fmag[f]];
            std::vector<Block *> main, split;
            for (Block *b : g_list.blocks)
                (nt != pk::kStreamListNt && sc.count({b->n[0], b->n[1], b->n[2]}) ? split : main).push_back(b);
            const int wrap = (!g_list.linked && g_list.self_built) ? 1 : 0;
            RCCHK(build_tab(g_list.tab[d][f], main, d, f, nt, nullptr, wrap));
            RCCHK(build_tab(g_list.stab[d][f], split, d, f, pk::kStreamListNt, nullptr, wrap));
            for (const LaunchTab *t : {&g_list.tab[d][f], &g_list.stab[d][f]})
                for (const pk::SweepArgs &a : t->h_args) g_list.wrap_any = g_list.wrap_any || a.images == 2;
        }
    }
    if (!g_list.d_queue) HIPCHK(hipMalloc((void **)&g_list.d_queue, sizeof(unsigned) * kQueues));
    if (!g_list.d_step) HIPCHK(hipMalloc((void **)&g_list.d_step, sizeof(pk::StepScalars)));
    if (!g_list.h_step) HIPCHK(hipHostMalloc((void **)&g_list.h_step, sizeof(pk::StepScalars), hipHostMallocDefault));
    if (!g_list.d_dtslots) HIPCHK(hipMalloc((void **)&g_list.d_dtslots, sizeof(unsigned long long) * 4 * kDtSlots));
    if (!g_list.d_dtout) HIPCHK(hipMalloc((void **)&g_list.d_dtout, sizeof(double) * 4));
    g_list.built = true;
    return PK_OK;
}

// The sweeps of one step in order: (direction, apply psi damping, reduce dt).
struct SweepStep {
    int d;
    bool fac, dt;
};
std::vector<SweepStep> step_sweeps()
{
    std::vector<SweepStep> v;
    for (int half = 0; half < 2; half++) {
        int last = -1;
        for (int s = 0; s < 3; s++) {
            const int d = half == 0 ? s : 2 - s;
            if (g_cfg.has_dir[d]) last = d;
        }
        for (int s = 0; s < 3; s++) {
            const int d = half == 0 ? s : 2 - s;
            if (g_cfg.has_dir[d]) v.push_back({d, d == last, half == 1 && d == last});
        }
    }
    return v;
}

// Enqueue the launches of one step on g_stream.  mode PK_BATCH_PER_BLOCK
// launches the single-block sweep once per block with the same tile shapes
// (for items of the stream form, the block's single-block shape); the
// persistent modes launch the list sweep once per sweep and fluid.
int enqueue_step(int mode, const pk::StepScalars &sc, bool ens = false)
{
    const std::vector<SweepStep> sw = step_sweeps();
    const bool persistent = mode != PK_BATCH_PER_BLOCK;
    if (sw.size() * (size_t)g_nfl * 2 > (size_t)kQueues) return fail(PK_EINVAL, "too many launches per step");
    if (persistent) HIPCHK(pk::launch_zero_u32(g_list.d_queue, kQueues, g_stream));
    unsigned long long *slots = ens ? g_ens.d_dtslots : g_list.d_dtslots;
    const int nslots = ens ? kEnsSlots : kDtSlots;
    HIPCHK(pk::launch_dtslots_init(slots, nslots * (ens ? (int)g_ens.nblk : 1), g_stream));
    int q = 0;
    for (size_t k = 0; k < sw.size(); k++) {
        const int d = sw[k].d;
        const int parity = (int)(k % 2);
        if (g_list.linked && !ens) RCCHK(exchange(d, parity));
        for (int f = 0; f < g_nfl; f++) {
            for (LaunchTab *tp : {&g_list.tab[d][f], &g_list.stab[d][f]}) {
                LaunchTab &t = *tp;
                if (t.nitems == 0) continue;
                double bytes = 0.0;
                for (Block *b : t.blocks) bytes += (double)b->n[0] * b->n[1] * b->n[2] * nv_of(f) * 16.0;
                if (persistent) {
                    pk::ListArgs L{};
                    L.args = t.d_args;
                    L.first = t.d_first;
                    L.items = t.d_items;
                    L.nitems = t.nitems;
                    L.queue = g_list.d_queue + q++;
                    L.grab = kListGrab;
                    L.parity = parity;
                    L.apply_fac = sw[k].fac && g_fmag[f];
                    L.step = ens ? g_ens.d_step : g_list.d_step;
                    L.dtslots = sw[k].dt ? slots : nullptr;
                    L.nslots = nslots;
                    L.per_block = ens ? 1 : 0;
                    L.bpos = t.d_first + t.blocks.size() + 1;
                    L.no_wrap = ens ? 1 : 0;  // an ensemble keeps its ghosts current (the adjoint reads them)
                    if (mode == PK_BATCH_GRAPH) {
                        HIPCHK(list_kernel(d, g_fmag[f], L, t.nt, t.grid));
                    } else {
                        RCCHK(launch_timed(d, bytes, [&]() { return list_kernel(d, g_fmag[f], L, t.nt, t.grid); }));
                    }
                } else {
                    for (size_t i = 0; i < t.blocks.size(); i++) {
                        pk::SweepArgs a = t.h_args[i];
                        if (a.images == 2) a.images = 1;  // one block at a time writes its images
                        if (is_stream(a.TS)) {
                            // list items of the stream form are sized for the whole list
                            // sharing the device; one block alone gets the single-block shape
                            choose_tile(t.blocks[i], d, true, g_fmag[f], a.TP, a.TS, a.out, a.chunks, a.ptiles);
                        }
                        if (parity) {
                            const double *s0 = a.src;
                            a.src = a.dst;
                            a.dst = const_cast<double *>(s0);
                        }
                        const size_t bi = ens ? t.h_first[t.blocks.size() + 1 + i] : 0;
                        const pk::StepScalars &bs = ens ? g_ens.h_step[bi] : sc;
                        a.base = a.src;
                        a.c1 = (0.5 * bs.dt) / a.dl[d];
                        a.c2 = (1.0 * bs.dt) / a.dl[d];
                        a.ch = bs.ch;
                        a.psi_fac = (sw[k].fac && g_fmag[f]) ? bs.psi_fac : 1.0;
                        a.dtslots = sw[k].dt ? slots + (ens ? 4ll * kEnsSlots * bi : 0ll) : nullptr;
                        a.nslots = nslots;
                        HIPCHK(sweep_kernel(d, g_fmag[f], true, a));
                    }
                }
            }
        }
    }
    if (ens)
        HIPCHK(pk::launch_dtslots_final_blocks(g_ens.d_dtslots, kEnsSlots, (int)g_ens.nblk, g_ens.d_dtout, g_stream));
    else
        HIPCHK(pk::launch_dtslots_final(g_list.d_dtslots, kDtSlots, g_list.d_dtout, g_stream));
    return PK_OK;
}

int ensure_graph()
{
    if (g_list.exec) return PK_OK;
    // kernel attributes were set when the tables were built (outside capture)
    pk::StepScalars sc{};
    HIPCHK(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step(PK_BATCH_GRAPH, sc);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(g_stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    HIPCHK(e);
    g_list.graph = g;
    HIPCHK(hipGraphInstantiate(&g_list.exec, g_list.graph, nullptr, nullptr, 0));
    return PK_OK;
}

// Ghost refresh of every block and the list timestep of the current state.
int list_prepare(double *dt_out, double *c_out)
{
    double dt = DBL_MAX, c = 0.0;
    RCCHK(list_fill());
    for (Block *b : g_list.blocks) {
        double dtb, cb;
        RCCHK(do_timestep(b, &dtb, &cb));
        if (dtb < dt) dt = dtb;
        if (cb > c) c = cb;
    }
    *dt_out = dt;
    *c_out = c;
    return PK_OK;
}

int list_dt(double *dt_out, double *c_out)
{
    if (g_list.fresh) {
        HIPCHK(hipMemcpyAsync(g_pinned, g_list.d_dtout, 4 * sizeof(double), hipMemcpyDeviceToHost, g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
        combine_dt(g_pinned, 1, dt_out, c_out);
        return PK_OK;
    }
    return list_prepare(dt_out, c_out);
}

int list_step(int mode, double dt)
{
    if (mode == PK_BATCH_PER_BLOCK) RCCHK(settle_ghosts());  // one block at a time reads its ghosts
    pk::StepScalars sc{};
    sc.dt = dt;
    sc.ch = 0.0;
    sc.psi_fac = 1.0;
    if (g_mhd) glm_scalars(g_list.mdl, dt, &sc.ch, &sc.psi_fac);
    // the pinned slot may still feed the previous step's copy
    HIPCHK(hipStreamSynchronize(g_stream));
    *g_list.h_step = sc;
    HIPCHK(hipMemcpyAsync(g_list.d_step, g_list.h_step, sizeof sc, hipMemcpyHostToDevice, g_stream));
    if (mode == PK_BATCH_GRAPH) {
        RCCHK(ensure_graph());
        HIPCHK(hipGraphLaunch(g_list.exec, g_stream));
    } else {
        RCCHK(enqueue_step(mode, sc));
    }
    for (Block *b : g_list.blocks) {
        b->uh = -1;
        b->dt_pending = false;
    }
    g_list.fresh = true;
    g_list.ghosts_stale = g_list.linked;
    if (g_list.wrap_any && mode != PK_BATCH_PER_BLOCK) g_list.wrap_stale = true;
    g_ens.fresh = false;
    return PK_OK;
}

int check_mode(int mode)
{
    if (mode != PK_BATCH_GRAPH && mode != PK_BATCH_PERSISTENT && mode != PK_BATCH_PER_BLOCK)
        return fail(PK_EINVAL, "unknown batch mode %d", mode);
    return PK_OK;
}

// ------------------------------------------------------------------ ensemble

int ens_build()
{
    RCCHK(settle_ghosts());  // the ensemble's sweeps read the ghosts
    RCCHK(list_build());
    if (g_list.linked) return fail(PK_EINVAL, "ensemble instances must be periodic on themselves (no neighbours)");
    const size_t n = g_list.blocks.size();
    if (g_ens.nblk == n && g_ens.d_step) return PK_OK;
    ens_free();
    g_ens.nblk = n;
    HIPCHK(hipMalloc((void **)&g_ens.d_step, sizeof(pk::StepScalars) * n));
    HIPCHK(hipHostMalloc((void **)&g_ens.h_step, sizeof(pk::StepScalars) * n, hipHostMallocDefault));
    HIPCHK(hipMalloc((void **)&g_ens.d_dtslots, sizeof(unsigned long long) * 4 * kEnsSlots * n));
    HIPCHK(hipMalloc((void **)&g_ens.d_dtout, sizeof(double) * 4 * n));
    HIPCHK(hipHostMalloc((void **)&g_ens.h_dtout, sizeof(double) * 4 * n, hipHostMallocDefault));
    HIPCHK(hipMalloc((void **)&g_ens.d_jobs, sizeof(pk::CopyJob) * n));
    HIPCHK(hipMalloc((void **)&g_ens.d_queue, sizeof(unsigned) * 2));
    g_ens.h_jobs.resize(n);
    g_ens.dt_old.assign(n, -1.0);
    g_ens.fresh = false;
    return PK_OK;
}

int ens_graph()
{
    if (g_ens.exec) return PK_OK;
    pk::StepScalars sc{};
    HIPCHK(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step(PK_BATCH_GRAPH, sc, true);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(g_stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    HIPCHK(e);
    g_ens.graph = g;
    HIPCHK(hipGraphInstantiate(&g_ens.exec, g_ens.graph, nullptr, nullptr, 0));
    return PK_OK;
}

// per-block dt of the current state (timestep_fluid), before any caps
int ens_dt_raw(std::vector<double> &dt)
{
    const size_t n = g_ens.nblk;
    dt.assign(n, 0.0);
    if (g_ens.fresh) {
        HIPCHK(hipMemcpyAsync(g_ens.h_dtout, g_ens.d_dtout, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
        for (size_t i = 0; i < n; i++) {
            double c;
            combine_dt(g_ens.h_dtout + 4 * i, 1, &dt[i], &c);
        }
        return PK_OK;
    }
    for (size_t i = 0; i < n; i++) {
        Block *b = g_list.blocks[i];
        RCCHK(do_fill(b));
        double c;
        RCCHK(do_timestep(b, &dt[i], &c));
    }
    return PK_OK;
}

// the per-block scalars of a step with per-block dts, uploaded
int ens_set_step(const double *dts)
{
    HIPCHK(hipStreamSynchronize(g_stream));  // the pinned scalars may still feed a copy
    for (size_t i = 0; i < g_ens.nblk; i++) {
        pk::StepScalars sc{};
        sc.dt = dts[i];
        sc.ch = 0.0;
        sc.psi_fac = 1.0;
        if (g_mhd) glm_scalars({g_list.mdl[i]}, dts[i], &sc.ch, &sc.psi_fac);
        g_ens.h_step[i] = sc;
    }
    HIPCHK(hipMemcpyAsync(g_ens.d_step, g_ens.h_step, sizeof(pk::StepScalars) * g_ens.nblk, hipMemcpyHostToDevice,
                          g_stream));
    return PK_OK;
}

// every block's current buffer <-> its slot; parity 1: the data sits in the
// block's other buffer (between the sweeps of a step)
int ens_copy_slots(int slot, bool save, int parity = 0)
{
    for (size_t i = 0; i < g_ens.nblk; i++) {
        Block *b = g_list.blocks[i];
        if (b->lam[0] == nullptr || slot < 0 || slot >= (int)b->slots.size())
            return fail(PK_EINVAL, "block %lld has no slot %d (pk_adjoint_setup)", (long long)b->id, slot);
        pk::CopyJob j{};
        j.src = save ? b->buf[b->cur ^ parity] : b->slots[slot];
        j.dst = save ? b->slots[slot] : b->buf[b->cur ^ parity];
        j.n = (long long)b->sv * g_nplanes;
        g_ens.h_jobs[i] = j;
    }
    HIPCHK(hipMemcpyAsync(g_ens.d_jobs, g_ens.h_jobs.data(), sizeof(pk::CopyJob) * g_ens.nblk, hipMemcpyHostToDevice,
                          g_stream));
    HIPCHK(pk::launch_copy_blocks(g_ens.d_jobs, (int)g_ens.nblk, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));  // h_jobs is reused by the next call
    return PK_OK;
}

// tile tables of the persistent adjoint sweep (per direction) and the sum jobs
int ens_adj_build()
{
    if (g_ens.adj_built) return PK_OK;
    const size_t n = g_ens.nblk;
    HIPCHK(hipMalloc((void **)&g_ens.d_adj, sizeof(pk::AdjArgs) * n));
    HIPCHK(hipHostMalloc((void **)&g_ens.h_adj, sizeof(pk::AdjArgs) * n, hipHostMallocDefault));
    HIPCHK(hipMalloc((void **)&g_ens.d_sum, sizeof(pk::SumJob) * 2 * n));
    HIPCHK(hipMalloc((void **)&g_ens.d_sums, sizeof(double) * 12 * n));
    HIPCHK(hipHostMalloc((void **)&g_ens.h_sums, sizeof(double) * 12 * n, hipHostMallocDefault));
    g_ens.h_sum.resize(2 * n);
    for (int d = 0; d < 3; d++) {
        if (!g_cfg.has_dir[d]) continue;
        std::vector<unsigned> first(n + 1);
        unsigned items = 0;
        for (size_t i = 0; i < n; i++) {
            int TP = 1, TS = 17, out = 1, chunks = 1, ptiles = 1;
            adj_tile(g_list.blocks[i], d, TP, TS, out, chunks, ptiles);
            first[i] = items;
            items += (unsigned)(chunks * ptiles);
        }
        first[n] = items;
        g_ens.anitems[d] = items;
        HIPCHK(hipMalloc((void **)&g_ens.d_afirst[d], sizeof(unsigned) * (n + 1)));
        HIPCHK(hipMalloc((void **)&g_ens.d_aitems[d], sizeof(unsigned) * (items ? items : 1)));
        HIPCHK(hipMemcpyAsync(g_ens.d_afirst[d], first.data(), sizeof(unsigned) * (n + 1), hipMemcpyHostToDevice,
                              g_stream));
        HIPCHK(pk::launch_fill_items(g_ens.d_afirst[d], (int)n, g_ens.d_aitems[d], g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
    }
    g_ens.adj_built = true;
    return PK_OK;
}

// ---- the unsplit Riemann driver (RIEMANN_UNSPLIT) ----

// periodic ghost refresh of one of the block's two buffers
int fill_buffer(Block *b, int i)
{
    pk::GhostArgs a{};
    a.u = state(b, i);
    a.sv = b->sv;
    a.sy = b->sy;
    a.sz = b->sz;
    for (int d = 0; d < 3; d++) {
        a.ng[d] = b->ng[d];
        a.g[d] = b->g[d];
        a.n[d] = b->n[d];
    }
    a.plane0 = 0;
    a.nplanes = g_nplanes;
    for (int d = 0; d < 3; d++)
        if (b->g[d] > 0) HIPCHK(pk::launch_fill_ghosts(d, a, g_stream));
    return PK_OK;
}

// One RK stage of unsplit_sweep on a periodic block (solve_cg_ub / solve_cg_u,
// then update_boundaries of the stage's arrays, unsplit_sweeps.F90:126-190):
// the first stage reads u and writes uh (the other buffer), the last reads
// uh and adds its fluxes to u in place.
// psi_fac: with the fused passes, psi of the stage's result is multiplied by
// it in the last pass (*fac_applied set); the in-place Euler stage leaves it.
int unsplit_stage(Block *b, int istep, double dt, double ch, double psi_fac = 1.0, bool *fac_applied = nullptr)
{
    if (fac_applied) *fac_applied = false;
    RCCHK(no_fc(b, "the unsplit solver"));
    if (g_nfl != 1) return fail(PK_EINVAL, "the unsplit solver covers one fluid");
    if (block_linked(b)) return fail(PK_EINVAL, "the unsplit solver covers blocks periodic on themselves");
    const bool euler = g_cfg.integration_order == 1;
    const bool first = !euler && istep == PK_RK2_1;
    const bool last = euler ? istep == PK_EULER : istep == PK_RK2_2;
    if (!first && !last) return fail(PK_EINVAL, "istep %d does not match integration_order %d", istep,
                                     g_cfg.integration_order);
    for (int d = 0; d < 3; d++)
        if (g_cfg.has_dir[d] && b->g[d] < 2) return fail(PK_EINVAL, "the unsplit stage needs nb >= 2");
    if (!(dt > 0.0)) return fail(PK_EINVAL, "dt must be positive");
    int src, dst;
    if (first) {
        src = b->cur;
        dst = 1 - b->cur;
    } else if (euler) {
        src = dst = b->cur;
    } else {
        if (b->uh < 0) return fail(PK_ESTATE, "the first stage has not run on block %lld", (long long)b->id);
        src = b->uh;
        dst = b->cur;
    }
    pk::UnsplitArgs a{};
    a.src = state(b, src);
    a.base = state(b, b->cur);
    a.dst = state(b, dst);
    a.sv = b->sv;
    a.sy = b->sy;
    a.sz = b->sz;
    for (int d = 0; d < 3; d++) {
        a.ng[d] = b->ng[d];
        a.g[d] = b->g[d];
        a.n[d] = b->n[d];
        a.has_dir[d] = g_cfg.has_dir[d];
        a.dl[d] = b->dl[d];
    }
    a.lim = g_cfg.limiter;
    a.limb = g_cfg.limiter_b;
    a.gam = g_cfg.fluid[0].gamma;
    a.gam_1 = g_cfg.fluid[0].gamma - 1.0;
    a.ch = ch;
    a.dt = dt;
    a.rk = rk_coef(istep);
    a.last = last ? 1 : 0;
    a.psi_fac = 1.0;
    a.smalld = g_cfg.smalld;
    a.smallei = g_cfg.smallei;
    a.use_smalld = g_cfg.use_smalld;
    a.use_smallei = g_cfg.use_smallei;
    a.ei_flag = g_ei_flag;
    a.flags = b->d_flags;
    a.dmass = b->d_mass;
    if (src != dst) {
        // one pass per direction: flux and flux difference, no flux arrays
        int d_first = -1, d_last = -1;
        for (int d = 0; d < 3; d++)
            if (g_cfg.has_dir[d]) {
                if (d_first < 0) d_first = d;
                d_last = d;
            }
        for (int d = 0; d < 3; d++) {
            if (!g_cfg.has_dir[d]) continue;
            const int fi = d == d_first, la = d == d_last;
            pk::UnsplitArgs ad = a;
            if (la) ad.psi_fac = psi_fac;
            HIPCHK(g_cfg.fast_math ? pk_fast::launch_unsplit_fused(d, g_mhd, ad, fi, la, g_stream)
                                   : pk::launch_unsplit_fused(d, g_mhd, ad, fi, la, g_stream));
        }
    } else {
        // in place (the Euler stage): every flux before any update
        if (!b->d_flux) HIPCHK(hipMalloc((void **)&b->d_flux, sizeof(double) * 3 * (size_t)g_nplanes * b->sv));
        const double *F[3] = {nullptr, nullptr, nullptr};
        for (int d = 0; d < 3; d++) {
            if (!g_cfg.has_dir[d]) continue;
            double *Fd = b->d_flux + (size_t)d * g_nplanes * b->sv;
            HIPCHK(g_cfg.fast_math ? pk_fast::launch_unsplit_flux(d, g_mhd, a, Fd, g_stream)
                                   : pk::launch_unsplit_flux(d, g_mhd, a, Fd, g_stream));
            F[d] = Fd;
        }
        HIPCHK(g_cfg.fast_math ? pk_fast::launch_unsplit_update(g_mhd, a, F, g_stream)
                               : pk::launch_unsplit_update(g_mhd, a, F, g_stream));
    }
    if (src == dst) RCCHK(fill_buffer(b, dst));  // the fused passes wrote the periodic images already
    else if (fac_applied) *fac_applied = true;
    b->uh = euler ? -1 : (first ? dst : b->uh);
    b->dt_pending = false;
    return PK_OK;
}

// psi *= fac over the block's current buffer, ghosts included (hdc.F90:338)
int damp_psi(Block *b, double fac)
{
    HIPCHK(pk::launch_scale(fluid_state(b, b->cur, g_magf) + 8 * b->sv, (long long)b->sz * b->ng[2], fac, g_stream));
    return PK_OK;
}

// fluid_update_unsplit (unsplit_fluidupdate.F90:45-126) on one periodic block:
// update_chspeed, glmdamping(half), the stages, glmdamping(half)
int fluid_update_unsplit(Block *b, double dt)
{
    double ch = 0.0, fac = 1.0;
    if (g_mhd) {
        const double mdl = min_dl(b);
        ch = (g_cfg.cfl_glm * mdl) / dt;  // hdc.F90:122
        // hdc.F90:315-331 with dt_eff = dt/2
        double f = 0.0;
        const double r = (g_cfg.glm_alpha * ch) / (mdl / (dt / 2.0));
        if (r > f) f = r;
        fac = exp(-f);
        RCCHK(damp_psi(b, fac));
    }
    if (g_cfg.integration_order == 1) {
        RCCHK(unsplit_stage(b, PK_EULER, dt, ch));
    } else {
        RCCHK(unsplit_stage(b, PK_RK2_1, dt, ch));
    }
    bool damped = false;  // the last fused pass applies glmdamping's second half itself
    if (g_cfg.integration_order != 1) RCCHK(unsplit_stage(b, PK_RK2_2, dt, ch, g_mhd ? fac : 1.0, &damped));
    if (g_mhd && !damped) RCCHK(damp_psi(b, fac));
    return PK_OK;
}

// ---- step retries (cflcontrol = 'redo' / 'flexible') ----

// timestep.F90:73-90 with repetitive_steps; the other knobs of &NUMERICAL_SETUP
// (global.F90:233,237,245,258)
struct CflControl {
    int mode = PK_CFL_WARN;
    double cfl_max = 0.9;
    double dt_shrink = 0.5;
    int max_attempts = 10;
    int disallow_negatives = 1;
};
CflControl g_cflc;

// dt_initial, dt_max_grow, dt_min, dt_max of &NUMERICAL_SETUP
// (global.F90:243-247) and tend, as time_step applies them (timestep.F90:198-224)
struct TimeControl {
    double dt_initial = -1.0;
    double dt_max_grow = 2.0;
    double dt_min = DBL_MIN;  // tiny(1.)
    double dt_max = DBL_MAX;  // huge(1.)
    double tend = DBL_MAX;    // simulated time left at the start of a pk_run / pk_run_all
};
TimeControl g_tc;

// cfl_max as init_global sanitizes it (global.F90:304)
double cfl_max_sane()
{
    const double c = g_cfg.cfl;
    return fmin(fmax(g_cflc.cfl_max, fmin(fmin(c * 1.1, c + 0.05), (1.0 + c) / 2.0)), 1.0);
}

// repeat_fluidstep's backup of every block (timestep_retry.F90:86-273), and
// the floor bookkeeping so that a redone step is accounted once
// (downgrade_magic_mass)
int step_save(const std::vector<Block *> &bl)
{
    for (Block *b : bl) {
        if (!b->backup) HIPCHK(hipMalloc((void **)&b->backup, state_bytes(b)));
        HIPCHK(hipMemcpyAsync(b->backup, b->buf[b->cur], state_bytes(b), hipMemcpyDeviceToDevice, g_stream));
        HIPCHK(hipMemcpyAsync(&b->saved_flags, b->d_flags, sizeof(unsigned), hipMemcpyDeviceToHost, g_stream));
        HIPCHK(hipMemcpyAsync(b->saved_mass, b->d_mass, kMassBytes, hipMemcpyDeviceToHost, g_stream));
    }
    HIPCHK(hipStreamSynchronize(g_stream));
    return PK_OK;
}

int step_restore(const std::vector<Block *> &bl)
{
    for (Block *b : bl) {
        HIPCHK(hipMemcpyAsync(b->buf[b->cur], b->backup, state_bytes(b), hipMemcpyDeviceToDevice, g_stream));
        HIPCHK(hipMemcpyAsync(b->d_mass, b->saved_mass, kMassBytes, hipMemcpyHostToDevice, g_stream));
        b->dt_pending = false;
        b->uh = -1;
    }
    list_touch();  // the list's pending dt and ghosts belong to the discarded step
    HIPCHK(hipStreamSynchronize(g_stream));
    return PK_OK;
}

// this step's negative-value flags only: cleared before the step, read after
// and merged into the accumulated report (check_cfl_violation resets
// dn_negative / ei_negative every step, timestep.F90:291-300)
int step_clear_flags(const std::vector<Block *> &bl)
{
    for (Block *b : bl) HIPCHK(hipMemsetAsync(b->d_flags, 0, sizeof(unsigned), g_stream));
    return PK_OK;
}

int step_flags(const std::vector<Block *> &bl, unsigned *neg)
{
    std::vector<unsigned> f(bl.size(), 0u);
    for (size_t i = 0; i < bl.size(); i++)
        HIPCHK(hipMemcpyAsync(&f[i], bl[i]->d_flags, sizeof(unsigned), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    *neg = 0u;
    for (size_t i = 0; i < bl.size(); i++) {
        *neg |= f[i];
        const unsigned merged = f[i] | bl[i]->saved_flags;
        HIPCHK(hipMemcpyAsync(bl[i]->d_flags, &merged, sizeof(unsigned), hipMemcpyHostToDevice, g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
    }
    return PK_OK;
}

// The main loop (piernik.F90:125-170) for nsteps steps: time_step with its
// caps, one step, and with repetitive steps check_cfl_violation after it --
// stepcfl = c_new/c_old*cfl*dt_cur_shrink > cfl_max, or negatives seen with
// disallow_negatives, redoes the step from its saved state with
// dt = dt_full*dt_cur_shrink (cfl_warn, timestep.F90:327-361; 'redo':
// dt_shrink**attempt, 'flexible': cfl_max/stepcfl, times dt_shrink from the
// second attempt).  A PIERNIK step is two substeps: t += 2 dt.
struct StepLoop {
    std::vector<Block *> blocks;
    std::function<int(double *, double *)> timestep;
    std::function<int(double)> update;
};

int run_loop(StepLoop &L, int nsteps, double *t_out, double *dt_out, int *redone)
{
    const bool redo = g_cflc.mode == PK_CFL_REDO || g_cflc.mode == PK_CFL_FLEXIBLE;
    const bool aut = g_cflc.mode == PK_CFL_AUTO;
    const double cfl_max = cfl_max_sane();
    double t = 0.0, dt = -1.0, dt_old = -1.0, dt_full = 0.0, c_old = 0.0, shrink = 1.0;
    double c_main_old = 0.0, stepcfl = 0.0, cfl_c = 1.0;  // timestep_pub.F90:42
    int s = 0, attempt = 0, nredo = 0;
    bool repeat = false;
    while (s < nsteps && t < g_tc.tend) {
        if (!repeat) {
            double c;
            RCCHK(L.timestep(&dt, &c));
            // time_step caps (timestep.F90:198-207); a call is a main loop from nstep = 0
            if (s < 1) {
                if (g_tc.dt_initial > 0.0) {
                    if (g_tc.dt_initial < dt) dt = g_tc.dt_initial;
                } else if (g_tc.dt_initial < 0.0) {
                    const double cap = -g_tc.dt_initial * dt;
                    if (cap < dt) dt = cap;
                }
            } else if (dt_old > 0.0) {
                const double cap = dt_old * g_tc.dt_max_grow;
                if (cap < dt) dt = cap;
            }
            if (aut) {
                // cfl_auto (timestep.F90:368-405), called from the main time_step
                double stepcfl_old = stepcfl;
                stepcfl = g_cfg.cfl;
                if (c_main_old > 0.0)
                    stepcfl = c / c_main_old * g_cfg.cfl;
                else
                    stepcfl_old = g_cfg.cfl;
                if (stepcfl > 0.0 && dt_old > 0.0) {
                    cfl_c = fmin(1.0, 0.5 * (cfl_c + fmin(1.0, stepcfl_old / stepcfl * dt / dt_old)));
                    dt = dt * cfl_c;
                } else {
                    cfl_c = 1.0;
                }
            }
            c_main_old = c;
            if (dt < g_tc.dt_min)  // time_step writes a crash dump here (timestep.F90:212-219)
                return fail(PK_EINVAL, "step %d: dt = %.6e is below dt_min = %.6e", s, dt, g_tc.dt_min);
            {
                // the split scheme's clamp to the time left (timestep.F90:224)
                const double left = g_tc.tend - t;
                const double lim = (0.5 * left) + ((2.0 * DBL_EPSILON) * left);
                dt = fmin(fmin(dt, g_tc.dt_max), lim);
            }
            dt_full = dt;
            c_old = c;
            if (redo) RCCHK(step_save(L.blocks));
        } else {
            RCCHK(step_restore(L.blocks));
            dt = dt_full * shrink;
        }
        if (redo) RCCHK(step_clear_flags(L.blocks));
        RCCHK(L.update(dt));
        dt_old = dt;
        if (redo) {
            double dtn, c_new;
            unsigned neg;
            RCCHK(L.timestep(&dtn, &c_new));
            RCCHK(step_flags(L.blocks, &neg));
            double stepcfl = g_cfg.cfl * shrink;
            if (c_old > 0.0) stepcfl = c_new / c_old * g_cfg.cfl * shrink;
            repeat = stepcfl > cfl_max || (g_cflc.disallow_negatives && neg != 0u);
            if (repeat) {
                if (g_cflc.mode == PK_CFL_FLEXIBLE) {
                    shrink = cfl_max / stepcfl;
                    if (attempt >= 1) shrink = shrink * g_cflc.dt_shrink;
                } else {
                    double f = 1.0;
                    for (int k = 0; k <= attempt; k++) f = f * g_cflc.dt_shrink;
                    shrink = f;
                }
                attempt++;
                nredo++;
                if (attempt > g_cflc.max_attempts)
                    return fail(PK_EINVAL, "step %d redone more than max_redostep_attempts = %d times", s,
                                g_cflc.max_attempts);
                continue;
            }
            shrink = 1.0;
            attempt = 0;
        }
        t = t + dt;
        t = t + dt;
        s++;
    }
    HIPCHK(hipStreamSynchronize(g_stream));
    if (t_out) *t_out = t;
    if (dt_out) *dt_out = dt;
    if (redone) *redone = nredo;
    return PK_OK;
}

int g_last_redone = 0;

// device-resident stepping of one block: only for blocks periodic on themselves
int solo(const Block *b)
{
    if (block_linked(b))
        return fail(PK_EINVAL, "block %lld has neighbours: advance it with the block-list functions",
                    (long long)b->id);
    return PK_OK;
}

// Before a download: the ghosts a connected list's steps left old (the
// steps exchange only the faces each sweep reads) are refreshed from the
// neighbours, so a download returns what update_boundaries would leave.
int ghosts_for_download()
{
    RCCHK(settle_ghosts());
    if (!g_list.ghosts_stale) return PK_OK;
    RCCHK(list_build());
    return g_list.linked ? exchange_all() : PK_OK;
}

}  // namespace

extern "C" {

const char *pk_last_error(void) { return g_err.c_str(); }

int pk_init(const pk_config *cfg)
{
    if (!cfg) return fail(PK_EINVAL, "null config");
    if (cfg->nfluids < 1 || cfg->nfluids > 2) return fail(PK_EINVAL, "nfluids = %d (1 or 2)", cfg->nfluids);
    if (cfg->limiter < 0 || cfg->limiter > 3 || cfg->limiter_b < 0 || cfg->limiter_b > 3)
        return fail(PK_EINVAL, "unknown limiter");
    if (cfg->integration_order != 1 && cfg->integration_order != 2)
        return fail(PK_EINVAL, "integration_order must be 1 or 2");
    int nmag = 0;
    for (int f = 0; f < cfg->nfluids; f++) {
        if (!(cfg->fluid[f].gamma > 1.0)) return fail(PK_EINVAL, "adiabatic fluid %d needs gamma > 1", f);
        if (cfg->fluid[f].magnetized) nmag++;
    }
    // solve_cg_riemann.F90:85-87: at most one magnetized fluid
    if (nmag > 1) return fail(PK_EINVAL, "at most one fluid may be magnetized");
    if (!(cfg->has_dir[0] || cfg->has_dir[1] || cfg->has_dir[2])) return fail(PK_EINVAL, "no active direction");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ndev < 1) return fail(PK_EHIP, "no HIP device");
    if (g_init) {
        for (auto &kv : g_blocks) free_block(kv.second);
        g_blocks.clear();
        list_free();
    }
    if (!g_stream) HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    if (!g_pinned) HIPCHK(hipHostMalloc((void **)&g_pinned, 16 * sizeof(double), hipHostMallocDefault));
    if (!g_part) HIPCHK(hipMalloc((void **)&g_part, sizeof(double) * 4 * 2048));
    unsigned long long *sp = nullptr;
    RCCHK(scratch_ptr(&sp));  // pk_sweep_noalloc's scratch on this GPU
    g_cfg = *cfg;
    g_nfl = cfg->nfluids;
    g_magf = -1;
    int p = 0;
    for (int f = 0; f < g_nfl; f++) {
        g_fmag[f] = cfg->fluid[f].magnetized != 0;
        g_fplane[f] = p;
        p += g_fmag[f] ? 9 : 5;
        if (g_fmag[f]) g_magf = f;
    }
    g_nplanes = p;
    g_mhd = g_magf >= 0;
    g_cflc = CflControl();
    g_ei_flag = g_cflc.disallow_negatives;
    g_tc = TimeControl();
    g_init = true;
    return PK_OK;
}

int pk_finalize(void)
{
    list_free();
    for (auto &kv : g_blocks) free_block(kv.second);
    g_blocks.clear();
    if (g_stream) {
        (void)hipStreamSynchronize(g_stream);
        (void)hipStreamDestroy(g_stream);
        g_stream = nullptr;
    }
    if (g_pinned) {
        (void)hipHostFree(g_pinned);
        g_pinned = nullptr;
    }
    if (g_stage) (void)hipFree(g_stage);
    g_stage = nullptr;
    g_stage_bytes = 0;
    if (g_part) (void)hipFree(g_part);
    g_part = nullptr;
    g_ring.release();
    g_init = false;
    return PK_OK;
}

int pk_register_block(int64_t cg_id, const int32_t n[3], int32_t nb, const double dl[3], double *u, double *uh,
                      double *b, double *bh, double *psi, double *psih)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    if (g_blocks.count(cg_id)) return fail(PK_EINVAL, "cg_id %lld already registered", (long long)cg_id);
    if (!n || !dl) return fail(PK_EINVAL, "null extents");
    if (nb < 1) return fail(PK_EINVAL, "nb = %d", nb);
    Block *blk = new Block();
    blk->id = cg_id;
    blk->nb = nb;
    for (int d = 0; d < 3; d++) blk->nbr[d][0] = blk->nbr[d][1] = cg_id;
    for (int d = 0; d < 3; d++) {
        if (n[d] < 1) {
            delete blk;
            return fail(PK_EINVAL, "n[%d] = %d", d, n[d]);
        }
        if (!g_cfg.has_dir[d] && n[d] != 1) {
            delete blk;
            return fail(PK_EINVAL, "inactive direction %d must have n = 1", d);
        }
        if (g_cfg.has_dir[d] && n[d] < nb) {
            delete blk;
            return fail(PK_EINVAL, "n[%d] = %d smaller than nb = %d", d, n[d], nb);
        }
        if (!(dl[d] > 0.0)) {
            delete blk;
            return fail(PK_EINVAL, "dl[%d] must be positive", d);
        }
        blk->n[d] = n[d];
        blk->g[d] = g_cfg.has_dir[d] ? nb : 0;
        blk->ng[d] = n[d] + 2 * blk->g[d];
        blk->dl[d] = dl[d];
    }
    blk->h_u = u;
    blk->h_uh = uh;
    blk->h_b = b;
    blk->h_bh = bh;
    blk->h_psi = psi;
    blk->h_psih = psih;
    int rc = alloc_block(blk);
    if (rc) {
        free_block(blk);
        return rc;
    }
    g_blocks[cg_id] = blk;
    list_invalidate();
    return PK_OK;
}

int pk_unregister_block(int64_t cg_id)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    (void)hipStreamSynchronize(g_stream);
    list_invalidate();
    g_blocks.erase(cg_id);
    free_block(b);
    return PK_OK;
}

int pk_upload(int64_t cg_id)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    if (!b->h_u) return fail(PK_EINVAL, "block %lld has no host arrays", (long long)cg_id);
    if (g_mhd && (!b->h_b || !b->h_psi)) return fail(PK_EINVAL, "magnetized block needs b and psi host arrays");
    list_touch();
    return upload_state(b, b->h_u, b->h_b, b->h_psi, state(b, b->cur));
}

int pk_download(int64_t cg_id)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    if (!b->h_u) return fail(PK_EINVAL, "block %lld has no host arrays", (long long)cg_id);
    RCCHK(ghosts_for_download());
    RCCHK(download_state(b, b->h_u, b->h_b, b->h_psi, state(b, b->cur)));
    if (b->uh >= 0 && b->h_uh) return download_state(b, b->h_uh, b->h_bh, b->h_psih, state(b, b->uh));
    return PK_OK;
}

int pk_upload_from(int64_t cg_id, const double *u, const double *bb, const double *psi)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    if (!u || (g_mhd && (!bb || !psi))) return fail(PK_EINVAL, "missing host array");
    list_touch();
    return upload_state(b, u, bb, psi, state(b, b->cur));
}

int pk_download_to(int64_t cg_id, double *u, double *bb, double *psi)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    RCCHK(ghosts_for_download());
    return download_state(b, u, bb, psi, state(b, b->cur));
}

int pk_upload_device(int64_t cg_id, const double *u, const double *bb, const double *psi)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    if (!u || (g_mhd && (!bb || !psi))) return fail(PK_EINVAL, "missing device array");
    list_touch();
    return upload_state(b, u, bb, psi, state(b, b->cur), true);
}

int pk_download_device(int64_t cg_id, double *u, double *bb, double *psi)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    RCCHK(ghosts_for_download());
    return download_state(b, u, bb, psi, state(b, b->cur), true);
}

int pk_solve_cg(int64_t cg_id, int32_t ddim, int32_t istep, double dt, double chspeed)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    if (ddim < 0 || ddim > 2 || !g_cfg.has_dir[ddim]) return fail(PK_EINVAL, "inactive or bad direction %d", ddim);
    const bool ok = (g_cfg.integration_order == 1) ? istep == PK_EULER : (istep == PK_RK2_1 || istep == PK_RK2_2);
    if (!ok) return fail(PK_EINVAL, "istep %d does not match integration_order %d", istep, g_cfg.integration_order);
    if (!(dt > 0.0)) return fail(PK_EINVAL, "dt must be positive");
    list_touch();
    return do_stage(b, ddim, istep, dt, chspeed);
}

int pk_sweep(int64_t cg_id, int32_t ddim, double dt, double chspeed, double psi_fac)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    RCCHK(solo(b));
    if (ddim < 0 || ddim > 2 || !g_cfg.has_dir[ddim]) return fail(PK_EINVAL, "inactive or bad direction %d", ddim);
    if (!(dt > 0.0)) return fail(PK_EINVAL, "dt must be positive");
    list_touch();
    return do_sweep(b, ddim, dt, chspeed, psi_fac);
}

int pk_sweep_noalloc(const double *src, double *dst, int64_t sv, int32_t sy, int32_t sz, const int32_t n[3],
                     int32_t nb, const double dl[3], int32_t ddim, double dt, double chspeed, double psi_fac,
                     double *report, void *stream)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    if (!src || !dst || !n || !dl) return fail(PK_EINVAL, "null argument");
    if (ddim < 0 || ddim > 2 || !g_cfg.has_dir[ddim]) return fail(PK_EINVAL, "inactive or bad direction %d", ddim);
    if (!(dt > 0.0)) return fail(PK_EINVAL, "dt must be positive");
    if (g_cfg.integration_order != 2) return fail(PK_EINVAL, "the fused sweep is RK2 only");
    if (nb < 4) return fail(PK_EINVAL, "fused RK2 sweep needs nb >= 4 (nb = %d)", nb);
    // the geometry as a block of the library's own (no buffers): the tile
    // choice and the kernel arguments come from it
    Block b;
    b.nb = nb;
    for (int d = 0; d < 3; d++) {
        if (!g_cfg.has_dir[d] && n[d] != 1) return fail(PK_EINVAL, "inactive direction %d must have n = 1", d);
        if (g_cfg.has_dir[d] && n[d] < nb) return fail(PK_EINVAL, "n[%d] = %d smaller than nb = %d", d, n[d], nb);
        if (!(dl[d] > 0.0)) return fail(PK_EINVAL, "dl[%d] must be positive", d);
        b.n[d] = n[d];
        b.g[d] = g_cfg.has_dir[d] ? nb : 0;
        b.ng[d] = n[d] + 2 * b.g[d];
        b.dl[d] = dl[d];
    }
    const long long plane = (long long)sz * b.ng[2];
    if (sy < b.ng[0] || (long long)sz < (long long)sy * b.ng[1] || sv < plane)
        return fail(PK_EINVAL, "strides (sy %d, sz %d, sv %lld) do not hold the ghosted box %d x %d x %d", sy, sz,
                    (long long)sv, b.ng[0], b.ng[1], b.ng[2]);
    const long long span = sv * g_nplanes;
    if ((dst < src ? src - dst : dst - src) < span)
        return fail(PK_EINVAL, "src and dst overlap (each holds %d planes of %lld values)", g_nplanes, (long long)sv);
    b.sv = sv;
    b.sy = sy;
    b.sz = sz;
    unsigned long long *s = nullptr;
    RCCHK(scratch_ptr(&s));
    hipStream_t st = stream ? (hipStream_t)stream : g_stream;
    // the counters and flags start at zero (a fill kernel: safe under graph capture)
    HIPCHK(pk::launch_zero_u32((unsigned *)(s + kNaMass), 2 * (kNaWords - kNaMass), st));
    if (report) HIPCHK(pk::launch_dtslots_init(s, kNaSlots, st));
    for (int f = 0; f < g_nfl; f++) {
        pk::SweepArgs a{};
        fill_common(&b, f, a);
        a.flags = (unsigned *)(s + kNaFlags);
        a.dmass = s + kNaMass + pk::kMassDigits * f;
        a.nslots = kNaSlots;
        a.dtslots = report ? s : nullptr;
        a.ch = chspeed;
        a.c1 = (0.5 * dt) / b.dl[ddim];
        a.c2 = (1.0 * dt) / b.dl[ddim];
        a.psi_fac = psi_fac;
        a.src = src + (long long)g_fplane[f] * sv;
        a.base = a.src;
        a.dst = dst + (long long)g_fplane[f] * sv;
        a.images = 0;
        choose_tile(&b, ddim, true, g_fmag[f], a.TP, a.TS, a.out, a.chunks, a.ptiles);
        HIPCHK(sweep_kernel(ddim, g_fmag[f], true, a, st));
    }
    if (report) {
        HIPCHK(pk::launch_dtslots_final(s, kNaSlots, report, st));
        hipLaunchKernelGGL(k_noalloc_report, dim3(1), dim3(64), 0, st, s, g_nfl, report);
        HIPCHK(hipGetLastError());
    }
    return PK_OK;
}

int pk_fill_ghosts(int64_t cg_id)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    if (!block_linked(b)) return do_fill(b);
    // a connected block takes its ghosts from the neighbours' current state
    RCCHK(list_build());
    return exchange_all();
}

}  // extern "C"

namespace {

// The fluid a field name's last letter selects (common_shortcuts,
// common_hdf5.F90:418-542): 'i' the ionized fluid (the magnetized one, else
// the first), 'n' the neutral one (the other of two fluids; the only fluid
// of a hydro run).  -1 if there is none.
int named_fluid(char c)
{
    if (c == 'i') return g_magf >= 0 ? g_magf : 0;
    if (c != 'n') return -1;
    if (g_nfl == 2) return g_magf >= 0 ? 1 - g_magf : 1;
    return g_magf >= 0 ? -1 : 0;
}

// datafields_hdf5 (data_hdf5.F90:496-580) for the fields computed from u, b
// and psi; psi is the qna array the fallback of get_data_from_cg finds.
int datafield_args(Block *b, const char *name, pk::DatafieldArgs &a)
{
    a = pk::DatafieldArgs{};
    a.state = state(b, b->cur);
    a.sv = b->sv;
    a.sy = b->sy;
    a.sz = b->sz;
    for (int d = 0; d < 3; d++) {
        a.g[d] = b->g[d];
        a.n[d] = b->n[d];
    }
    a.mag = -1;
    const std::string v = name ? name : "";
    const int mplane = g_mhd ? g_fplane[g_magf] + 5 : -1;
    auto fluid = [&](char c, int *f) {
        *f = named_fluid(c);
        if (*f < 0) return fail(PK_EINVAL, "field '%s': no such fluid in this configuration", v.c_str());
        a.fluid = g_fplane[*f];
        a.gam_1 = g_cfg.fluid[*f].gamma - 1.0;
        return PK_OK;
    };
    int f = -1;
    if (v.size() == 4 && (v.compare(0, 3, "den") == 0 || v.compare(0, 3, "ene") == 0)) {
        RCCHK(fluid(v[3], &f));
        a.kind = pk::DF_COPY;
        a.plane = a.fluid + (v[0] == 'd' ? 0 : 4);
    } else if (v.size() == 4 && v.compare(0, 2, "vl") == 0 && v[2] >= 'x' && v[2] <= 'z') {
        RCCHK(fluid(v[3], &f));
        a.kind = pk::DF_VEL;
        a.plane = a.fluid + 1 + (v[2] - 'x');
    } else if (v.size() == 5 && v.compare(0, 3, "mom") == 0 && v[3] >= 'x' && v[3] <= 'z') {
        RCCHK(fluid(v[4], &f));
        a.kind = pk::DF_COPY;
        a.plane = a.fluid + 1 + (v[3] - 'x');
    } else if (v.size() == 4 && (v.compare(0, 3, "pre") == 0 || v.compare(0, 3, "eth") == 0)) {
        RCCHK(fluid(v[3], &f));
        a.kind = v[0] == 'p' ? pk::DF_PRE : pk::DF_ETH;
        if (v[3] == 'i') a.mag = mplane;  // emag_c enters the ionized fluid's pressure
    } else if (v == "v") {
        RCCHK(fluid(g_nfl == 2 || g_magf >= 0 ? 'i' : 'n', &f));
        a.kind = pk::DF_VTOT;
    } else if (g_mhd && v.size() == 4 && v.compare(0, 3, "mag") == 0 && v[3] >= 'x' && v[3] <= 'z') {
        a.kind = pk::DF_COPY;
        a.plane = mplane + (v[3] - 'x');
    } else if (g_mhd && v == "magB") {
        a.kind = pk::DF_MAGB;
        a.mag = mplane;
    } else if (g_mhd && v == "psi") {
        a.kind = pk::DF_COPY;
        a.plane = mplane + 3;
    } else {
        return fail(PK_EINVAL, "'%s' is not an output field of this configuration", v.c_str());
    }
    return PK_OK;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

int datafield(int64_t cg_id, const char *name, void *out, bool fp32)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    if (!out) return fail(PK_EINVAL, "no output array");
    pk::DatafieldArgs a;
    RCCHK(datafield_args(b, name, a));
    const size_t cells = (size_t)b->n[0] * b->n[1] * b->n[2];
    const size_t bytes = (fp32 ? sizeof(float) : sizeof(double)) * cells;
    RCCHK(ensure_stage(bytes));
    if (fp32)
        a.out32 = (float *)g_stage;
    else
        a.out = g_stage;
    // planes read per cell: the field's inputs, then the field written
    int reads = 1;
    if (a.kind == pk::DF_VEL) reads = 2;
    if (a.kind == pk::DF_VTOT) reads = 4;
    if (a.kind == pk::DF_PRE || a.kind == pk::DF_ETH) reads = 5 + (a.mag >= 0 ? 3 : 0);
    if (a.kind == pk::DF_MAGB) reads = 3;
    const double algo = 8.0 * cells * reads + (double)bytes;
    RCCHK(launch_timed(3, algo, [&]() { return pk::launch_datafield(a, g_stream); }));
    if (ring_ready())
        HIPCHK(g_ring.d2h(out, g_stage, bytes, g_stream));
    else
        HIPCHK(hipMemcpyAsync(out, g_stage, bytes, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return PK_OK;
}

}  // namespace

extern "C" {

int pk_datafield(int64_t cg_id, const char *name, double *out) { return datafield(cg_id, name, out, false); }

int pk_datafield_f32(int64_t cg_id, const char *name, float *out) { return datafield(cg_id, name, out, true); }

int pk_register_fcflx(int64_t cg_id, int32_t ddim, int32_t side, int32_t kind, int64_t first, const int32_t *index,
                      double *uflx, double *bflx)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    if (ddim < 0 || ddim > 2 || !g_cfg.has_dir[ddim]) return fail(PK_EINVAL, "inactive or bad direction %d", ddim);
    if (side != 0 && side != 1) return fail(PK_EINVAL, "side must be 0 (low) or 1 (high), got %d", side);
    if (kind != PK_FC_FINEBND && kind != PK_FC_COARSEBND)
        return fail(PK_EINVAL, "kind must be PK_FC_FINEBND or PK_FC_COARSEBND, got %d", kind);
    if (index && !uflx) return fail(PK_EINVAL, "uflx is required with index");
    if (index && g_mhd && !bflx) return fail(PK_EINVAL, "bflx (B and psi fluxes) is required with a magnetic field");
    (void)hipStreamSynchronize(g_stream);
    Block::FcReg &r = b->fc[ddim][kind][side];
    r.index = index;
    r.uflx = index ? uflx : nullptr;
    r.bflx = (index && g_mhd) ? bflx : nullptr;
    r.first = first;
    list_invalidate();
    return PK_OK;
}

int pk_set_neighbour(int64_t cg_id, int32_t dir, int32_t side, int64_t nb_id)
{
    Block *b, *o;
    RCCHK(get_block(cg_id, &b));
    if (dir < 0 || dir > 2 || !g_cfg.has_dir[dir]) return fail(PK_EINVAL, "inactive or bad direction %d", dir);
    if (side != 0 && side != 1) return fail(PK_EINVAL, "side must be 0 (low) or 1 (high), got %d", side);
    RCCHK(get_block(nb_id, &o));
    RCCHK(check_face(b, o, dir));
    if (nb_id != cg_id && b->lam[0])
        return fail(PK_EINVAL, "block %lld is set up for the adjoint, which needs it periodic on itself",
                    (long long)cg_id);
    (void)hipStreamSynchronize(g_stream);
    b->nbr[dir][side] = nb_id;
    list_invalidate();
    return PK_OK;
}

int pk_glm_damp(int64_t cg_id, double fac)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    if (!g_mhd) return PK_OK;
    list_touch();
    b->dt_pending = false;
    HIPCHK(pk::launch_scale(fluid_state(b, b->cur, g_magf) + 8 * b->sv, (long long)b->sz * b->ng[2], fac, g_stream));
    return PK_OK;
}

int pk_timestep_fluid(int64_t cg_id, double *dt_out, double *c_out)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    if (!dt_out || !c_out) return fail(PK_EINVAL, "null output");
    return do_timestep(b, dt_out, c_out);
}

int pk_fluid_update(int64_t cg_id, double dt)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    RCCHK(solo(b));
    if (!(dt > 0.0)) return fail(PK_EINVAL, "dt must be positive");
    list_touch();
    return do_fluid_update(b, dt);
}

int pk_run(int64_t cg_id, int32_t nsteps, double *t_out, double *dt_out)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    RCCHK(solo(b));
    if (nsteps < 0) return fail(PK_EINVAL, "nsteps < 0");
    list_touch();
    RCCHK(do_fill(b));
    StepLoop L;
    L.blocks = {b};
    L.timestep = [b](double *dt, double *c) { return do_timestep(b, dt, c); };
    L.update = [b](double dt) { return do_fluid_update(b, dt, true); };
    RCCHK(run_loop(L, nsteps, t_out, dt_out, &g_last_redone));
    if (b->ghosts_stale) RCCHK(do_fill(b));  // the ghosts as the sweeps' images would have left them
    return PK_OK;
}

int pk_set_cfl_control(int32_t mode, double cfl_max, double dt_shrink, int32_t max_attempts,
                       int32_t disallow_negatives)
{
    if (mode != PK_CFL_WARN && mode != PK_CFL_REDO && mode != PK_CFL_FLEXIBLE && mode != PK_CFL_AUTO)
        return fail(PK_EINVAL, "unknown cfl control mode %d", mode);
    if (!(cfl_max > 0.0) || !(dt_shrink > 0.0 && dt_shrink < 1.0) || max_attempts < 0)
        return fail(PK_EINVAL, "cfl_max > 0, 0 < dt_shrink < 1 and max_attempts >= 0 required");
    g_cflc.mode = mode;
    g_cflc.cfl_max = cfl_max;
    g_cflc.dt_shrink = dt_shrink;
    g_cflc.max_attempts = max_attempts;
    g_cflc.disallow_negatives = disallow_negatives ? 1 : 0;
    if (g_ei_flag != g_cflc.disallow_negatives) {
        g_ei_flag = g_cflc.disallow_negatives;
        list_invalidate();  // the list's argument tables carry the flag
    }
    return PK_OK;
}

int pk_set_time_control(double dt_initial, double dt_max_grow, double dt_min, double dt_max, double tend)
{
    if (!(dt_max_grow > 0.0) || !(dt_min >= 0.0) || !(dt_max > 0.0) || !(tend > 0.0))
        return fail(PK_EINVAL, "dt_max_grow, dt_max and tend must be positive, dt_min not negative");
    g_tc.dt_initial = dt_initial;
    g_tc.dt_max_grow = dt_max_grow;
    g_tc.dt_min = dt_min;
    g_tc.dt_max = dt_max;
    g_tc.tend = tend;
    return PK_OK;
}

int pk_redone_steps(int32_t *count)
{
    if (!count) return fail(PK_EINVAL, "null output");
    *count = g_last_redone;
    return PK_OK;
}

int pk_unsplit_stage(int64_t cg_id, int32_t istep, double dt, double chspeed)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    list_touch();
    return unsplit_stage(b, istep, dt, chspeed);
}

int pk_fluid_update_unsplit(int64_t cg_id, double dt)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    if (!(dt > 0.0)) return fail(PK_EINVAL, "dt must be positive");
    RCCHK(solo(b));
    list_touch();
    return fluid_update_unsplit(b, dt);
}

int pk_run_unsplit(int64_t cg_id, double tend, int32_t nmax, double *t_out, int32_t *nsteps_out)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    RCCHK(solo(b));
    if (nmax < 0) return fail(PK_EINVAL, "nmax < 0");
    list_touch();
    double t = 0.0, dt = -1.0, dt_old = -1.0;
    int s = 0;
    RCCHK(do_fill(b));
    while (t < tend && s < nmax) {
        double c;
        RCCHK(do_timestep(b, &dt, &c));
        // time_step caps (timestep.F90:198-207), then the unsplit clamp to tend (:221-222)
        if (s >= 1 && dt_old > 0.0) {
            const double cap = dt_old * 2.0;
            if (cap < dt) dt = cap;
        }
        const double rem = tend - t;
        const double cl = rem + (2.0 * DBL_EPSILON) * rem;
        if (cl < dt) dt = cl;
        RCCHK(fluid_update_unsplit(b, dt));
        t = t + dt;
        dt_old = dt;
        s++;
    }
    HIPCHK(hipStreamSynchronize(g_stream));
    if (t_out) *t_out = t;
    if (nsteps_out) *nsteps_out = s;
    return PK_OK;
}

int pk_get_report(int64_t cg_id, pk_stage_report *rep, int32_t reset)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    unsigned fl = 0;
    unsigned long long acc[2][pk::kMassDigits] = {};
    HIPCHK(hipMemcpyAsync(&fl, b->d_flags, sizeof fl, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipMemcpyAsync(acc, b->d_mass, kMassBytes, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    if (rep) {
        double dvol = 1.0;
        for (int d = 0; d < 3; d++)
            if (g_cfg.has_dir[d]) dvol *= b->dl[d];
        rep->dn_negative = (fl & 1u) ? 1 : 0;
        rep->ei_negative = (fl & 2u) ? 1 : 0;
        // dvol times the exact sum of the floor's additions, rounded once
        for (int f = 0; f < 2; f++) rep->magic_mass[f] = f < g_nfl ? pk::mass_value(acc[f]) * dvol : 0.0;
    }
    if (reset) {
        HIPCHK(hipMemsetAsync(b->d_flags, 0, sizeof(unsigned), g_stream));
        HIPCHK(hipMemsetAsync(b->d_mass, 0, kMassBytes, g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
    }
    return PK_OK;
}

int pk_autotune(int64_t cg_id, double dt, double chspeed, int32_t tiles_out[6])
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    if (!(dt > 0.0)) return fail(PK_EINVAL, "dt must be positive");
    if (g_cfg.integration_order != 2) return fail(PK_EINVAL, "autotuning covers the fused RK2 sweep only");
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const int ts_list[] = {24, 40, 48, 56, 64, 72, 80, 96, 104, 136, 200, 264};
    const int waves_list[] = {1, 2, 3, 4, 6, 8};
    for (int d = 0; d < 3; d++) {
        if (tiles_out) tiles_out[2 * d] = tiles_out[2 * d + 1] = 0;
        if (!g_cfg.has_dir[d]) continue;
        if (b->g[d] < 4) return fail(PK_EINVAL, "fused RK2 sweep needs nb >= 4");
        for (int f = 0; f < g_nfl; f++) {
            pk::SweepArgs a{};
            fill_common(b, f, a);
            a.ch = chspeed;
            a.c1 = (0.5 * dt) / b->dl[d];
            a.c2 = (1.0 * dt) / b->dl[d];
            a.src = fluid_state(b, b->cur, f);
            a.base = a.src;
            a.dst = fluid_state(b, 1 - b->cur, f);
            std::vector<std::pair<int, int>> seen;
            float best = 1e30f;
            std::pair<int, int> best_t(0, 0);
            // the march kernel is a candidate for y and z sweeps with enough
            // pencils to give every SIMD a wave (x pencils are rows: uncoalesced)
            const int d1 = (d == 0) ? 1 : 0, d2 = (d == 2) ? 1 : 2;
            const bool try_march = d > 0 && (long long)b->n[d1] * b->n[d2] >= 65536;
            for (int ts : ts_list) {
                for (int w : waves_list) {
                    int TP, TS, out, chunks, ptiles;
                    int tp = (64 * w) / (ts < b->n[d] + 8 ? ts : b->n[d] + 8);
                    if (tp < 1) continue;
                    tile_geometry(b, d, true, tp, ts, TP, TS, out, chunks, ptiles);
                    if (TP * TS > 512) continue;
                    bool dup = false;
                    for (auto &p : seen) dup = dup || (p.first == TP && p.second == TS);
                    if (dup) continue;
                    seen.push_back({TP, TS});
                    a.TP = TP;
                    a.TS = TS;
                    a.out = out;
                    a.chunks = chunks;
                    a.ptiles = ptiles;
                    HIPCHK(sweep_kernel(d, g_fmag[f], true, a));  // warm-up
                    HIPCHK(hipEventRecord(e0, g_stream));
                    for (int rep = 0; rep < 3; rep++) HIPCHK(sweep_kernel(d, g_fmag[f], true, a));
                    HIPCHK(hipEventRecord(e1, g_stream));
                    HIPCHK(hipEventSynchronize(e1));
                    float ms = 0.0f;
                    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                    if (ms < best) {
                        best = ms;
                        best_t = {TP, TS};
                    }
                }
            }
            // the stream kernel: 1, 2, 4 or 8 pencils per lane group,
            // workgroups of about 8 to 64 iterations
            for (int tp : {1, 2, 4, 8}) {
                for (int m : {8, 16, 32, 64}) {
                    a.TS = -m;
                    stream_geometry(b, d, tp, m, a.TP, a.out, a.chunks, a.ptiles);
                    HIPCHK(sweep_kernel(d, g_fmag[f], true, a));  // warm-up
                    HIPCHK(hipEventRecord(e0, g_stream));
                    for (int rep = 0; rep < 3; rep++) HIPCHK(sweep_kernel(d, g_fmag[f], true, a));
                    HIPCHK(hipEventRecord(e1, g_stream));
                    HIPCHK(hipEventSynchronize(e1));
                    float ms = 0.0f;
                    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                    if (ms < best) {
                        best = ms;
                        best_t = {tp, -m};
                    }
                }
            }
            if (try_march) {
                a.TS = kMarch;
                march_geometry(b, d, a.TP, a.out, a.chunks, a.ptiles);
                HIPCHK(sweep_kernel(d, g_fmag[f], true, a));  // warm-up
                HIPCHK(hipEventRecord(e0, g_stream));
                for (int rep = 0; rep < 3; rep++) HIPCHK(sweep_kernel(d, g_fmag[f], true, a));
                HIPCHK(hipEventRecord(e1, g_stream));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0.0f;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) {
                    best = ms;
                    best_t = {64, kMarch};
                }
            }
            g_tuned[tune_key(b, d, g_fmag[f])] = best_t;
            if (tiles_out && f == 0) {
                tiles_out[2 * d] = best_t.first;
                tiles_out[2 * d + 1] = best_t.second;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // the candidates wrote the other buffer only; the state is unchanged.
    // The floor bookkeeping they accumulated is discarded.
    HIPCHK(hipMemsetAsync(b->d_flags, 0, sizeof(unsigned), g_stream));
    HIPCHK(hipMemsetAsync(b->d_mass, 0, kMassBytes, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return PK_OK;
}

// ---- the block list ----------------------------------------------------------

int pk_timestep_all(double *dt_out, double *c_out)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    if (!dt_out || !c_out) return fail(PK_EINVAL, "null output");
    RCCHK(list_build());
    return list_dt(dt_out, c_out);
}

int pk_fluid_update_all(int32_t mode, double dt)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    RCCHK(check_mode(mode));
    if (!(dt > 0.0)) return fail(PK_EINVAL, "dt must be positive");
    RCCHK(list_build());
    if (!g_list.fresh) RCCHK(list_fill());
    return list_step(mode, dt);  // ghosts left old are refilled when something reads them
}

int pk_run_all(int32_t mode, int32_t nsteps, double *t_out, double *dt_out)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    RCCHK(check_mode(mode));
    if (nsteps < 0) return fail(PK_EINVAL, "nsteps < 0");
    RCCHK(list_build());
    StepLoop L;
    L.blocks = g_list.blocks;
    L.timestep = [](double *dt, double *c) { return list_dt(dt, c); };
    L.update = [mode](double dt) {
        // a redone step starts from restored states: ghosts and dt are recomputed
        if (!g_list.fresh) RCCHK(list_fill());
        return list_step(mode, dt);
    };
    return run_loop(L, nsteps, t_out, dt_out, &g_last_redone);  // ghosts left old are refilled when read
}

int pk_autotune_all(double dt, double chspeed)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    if (!(dt > 0.0)) return fail(PK_EINVAL, "dt must be positive");
    RCCHK(settle_ghosts());
    RCCHK(list_build());
    pk::StepScalars sc{};
    sc.dt = dt;
    sc.ch = chspeed;
    sc.psi_fac = 1.0;
    if (!(chspeed > 0.0) && g_mhd) {
        double fac;
        glm_scalars(g_list.mdl, dt, &sc.ch, &fac);
    }
    HIPCHK(hipStreamSynchronize(g_stream));
    *g_list.h_step = sc;
    HIPCHK(hipMemcpyAsync(g_list.d_step, g_list.h_step, sizeof sc, hipMemcpyHostToDevice, g_stream));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    // block shapes present in the list, with up to ~8M cells of each for timing
    std::map<std::vector<int>, std::vector<Block *>> classes;
    for (Block *b : g_list.blocks) {
        auto &v = classes[{b->n[0], b->n[1], b->n[2]}];
        long long cells = 0;
        for (Block *o : v) cells += (long long)o->n[0] * o->n[1] * o->n[2];
        if (cells < 8ll * 1024 * 1024) v.push_back(b);
    }
    // the stream form stores through buffer ranges of 24 sv bytes: every block's planes must fit
    bool streams_fit = true;
    for (Block *b : g_list.blocks) streams_fit = streams_fit && pk::stream_stores_fit(b->sv);
    LaunchTab tmp;
    int rc = PK_OK;
    const int nt_list[] = {pk::kStreamListNt, 128, 256, 512};
    for (int d = 0; d < 3 && !rc; d++) {
        if (!g_cfg.has_dir[d]) continue;
        for (int f = 0; f < g_nfl && !rc; f++) {
            // per block shape and workgroup size: the best tile length and its time
            std::map<std::vector<int>, std::map<int, std::pair<int, float>>> timed;
            for (int nt : nt_list) {
                if (nt == pk::kStreamListNt && !streams_fit) continue;
                if (list_resident(d, g_fmag[f], nt) <= 0) continue;
                for (auto &cl : classes) {
                    Block *b0 = cl.second[0];
                    std::vector<int> cand;
                    if (nt == pk::kStreamListNt) {
                        cand = {-4, -8, -16, -32};  // iterations of the stream per item
                    } else {
                        for (int k = 1; k <= 8; k++) {
                            const int out = (b0->n[d] + k - 1) / k;
                            const int ts = out + 8;
                            if (ts <= nt && std::find(cand.begin(), cand.end(), ts) == cand.end()) cand.push_back(ts);
                        }
                    }
                    float best = 1e30f;
                    int best_ts = cand.empty() ? 9 : cand[0];
                    for (int ts : cand) {
                        std::vector<int> tsv(cl.second.size(), ts);
                        // stream candidates timed as they run when every sweep of the block is one:
                        // ghost positions from their pre-images, no images written
                        if ((rc = build_tab(tmp, cl.second, d, f, nt, &tsv, g_list.self_built ? 2 : 0))) break;
                        pk::ListArgs L{};
                        L.args = tmp.d_args;
                        L.first = tmp.d_first;
                        L.items = tmp.d_items;
                        L.nitems = tmp.nitems;
                        L.queue = g_list.d_queue;
                        L.grab = kListGrab;
                        L.parity = 0;
                        L.apply_fac = 0;
                        L.step = g_list.d_step;
                        L.dtslots = nullptr;
                        L.nslots = kDtSlots;
                        float ms = 0.0f;
                        for (int rep = 0; rep < 2; rep++) {  // warm-up, then timed
                            HIPCHK(pk::launch_zero_u32(g_list.d_queue, 1, g_stream));
                            HIPCHK(hipEventRecord(e0, g_stream));
                            HIPCHK(list_kernel(d, g_fmag[f], L, nt, tmp.grid));
                            HIPCHK(hipEventRecord(e1, g_stream));
                            HIPCHK(hipEventSynchronize(e1));
                            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                        }
                        if (ms < best) {
                            best = ms;
                            best_ts = ts;
                        }
                    }
                    if (rc) break;
                    timed[cl.first][nt] = {best_ts, best};
                }
                if (rc) break;
            }
            if (rc) break;
            // one workgroup size for the whole list, or the stream form for the
            // shapes where it is faster and one tile size for the rest (two
            // launches per sweep), whichever the per-shape times favour
            double best_total = 1e300;
            int best_nt = 256;
            bool best_split = false;
            for (int nt : nt_list) {
                for (int split = 0; split < 2; split++) {
                    if (split && nt == pk::kStreamListNt) continue;
                    double total = 0.0;
                    bool ok = true;
                    for (auto &cl : classes) {
                        auto &tm = timed[cl.first];
                        const auto it = tm.find(nt);
                        const auto is = tm.find(pk::kStreamListNt);
                        if (it == tm.end() && !(split && is != tm.end())) {
                            ok = false;
                            break;
                        }
                        double t = it != tm.end() ? it->second.second : 1e30;
                        if (split && is != tm.end()) t = std::min(t, (double)is->second.second);
                        total += t;
                    }
                    if (ok && total < best_total) {
                        best_total = total;
                        best_nt = nt;
                        best_split = split != 0;
                    }
                }
            }
            g_list.nt_pref[d][g_fmag[f]] = best_nt;
            g_list.ts_force[d][g_fmag[f]] = 0;
            g_list.split_cls[d][g_fmag[f]].clear();
            g_list.split_ts[d][g_fmag[f]] = 0;
            for (auto &cl : classes) {
                auto &tm = timed[cl.first];
                int nt = best_nt;
                if (best_split) {
                    const auto it = tm.find(best_nt), is = tm.find(pk::kStreamListNt);
                    if (is != tm.end() && (it == tm.end() || is->second.second < it->second.second)) {
                        nt = pk::kStreamListNt;
                        g_list.split_cls[d][g_fmag[f]].insert(cl.first);
                    }
                }
                const auto it = tm.find(nt);
                if (it == tm.end()) continue;
                const int ts = it->second.first;
                g_tuned[tune_key(cl.second[0], d, g_fmag[f], nt)] = {is_stream(ts) ? pk::stream_list_tp(d) : nt / ts, ts};
            }
        }
    }
    free_tab(tmp);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    // the candidates wrote the other buffers only; their floor bookkeeping is discarded
    for (Block *b : g_list.blocks) {
        HIPCHK(hipMemsetAsync(b->d_flags, 0, sizeof(unsigned), g_stream));
        HIPCHK(hipMemsetAsync(b->d_mass, 0, kMassBytes, g_stream));
    }
    HIPCHK(hipStreamSynchronize(g_stream));
    const bool fresh = g_list.fresh;
    list_invalidate();
    RCCHK(list_build());
    g_list.fresh = fresh;
    return PK_OK;
}

int pk_set_list_tiles(int32_t ddim, int32_t ifluid, int32_t lanes, int32_t ts)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    if (ddim < 0 || ddim > 2 || !g_cfg.has_dir[ddim]) return fail(PK_EINVAL, "inactive or bad direction %d", ddim);
    if (ifluid < 0 || ifluid >= g_nfl) return fail(PK_EINVAL, "bad fluid %d", ifluid);
    if (lanes == pk::kStreamListNt) {
        if (ts != 0 && !is_stream(ts)) return fail(PK_EINVAL, "stream list items of %d iterations", -ts);
    } else if (lanes == 128 || lanes == 256 || lanes == 512) {
        if (ts != 0 && (ts < 9 || ts > lanes)) return fail(PK_EINVAL, "tile length %d for %d lanes", ts, lanes);
    } else {
        return fail(PK_EINVAL, "list workgroups of %d lanes (64: the stream; 128, 256, 512: tiles)", lanes);
    }
    if (lanes == pk::kStreamListNt)
        for (Block *b : g_list.blocks)
            if (!pk::stream_stores_fit(b->sv))
                return fail(PK_EINVAL, "block planes of %lld cells exceed the stream's buffer range", b->sv);
    g_list.nt_pref[ddim][g_fmag[ifluid]] = lanes;
    g_list.ts_force[ddim][g_fmag[ifluid]] = ts;
    g_list.split_cls[ddim][g_fmag[ifluid]].clear();
    g_list.split_ts[ddim][g_fmag[ifluid]] = 0;
    list_invalidate();
    return PK_OK;
}

int pk_set_list_split(int32_t ddim, int32_t ifluid, const int32_t *shapes, int32_t nshapes, int32_t ts)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    if (ddim < 0 || ddim > 2 || !g_cfg.has_dir[ddim]) return fail(PK_EINVAL, "inactive or bad direction %d", ddim);
    if (ifluid < 0 || ifluid >= g_nfl) return fail(PK_EINVAL, "bad fluid %d", ifluid);
    if (nshapes < 0 || (nshapes > 0 && !shapes)) return fail(PK_EINVAL, "null shapes");
    if (ts != 0 && !is_stream(ts)) return fail(PK_EINVAL, "stream list items of %d iterations", -ts);
    if (g_list.nt_pref[ddim][g_fmag[ifluid]] == pk::kStreamListNt)
        return fail(PK_EINVAL, "the list already sweeps every block in the stream form");
    std::set<std::vector<int>> cls;
    for (int i = 0; i < nshapes; i++) cls.insert({shapes[3 * i], shapes[3 * i + 1], shapes[3 * i + 2]});
    for (Block *b : g_list.blocks)
        if (cls.count({b->n[0], b->n[1], b->n[2]}) && !pk::stream_stores_fit(b->sv))
            return fail(PK_EINVAL, "block planes of %lld cells exceed the stream's buffer range", b->sv);
    g_list.split_cls[ddim][g_fmag[ifluid]] = cls;
    g_list.split_ts[ddim][g_fmag[ifluid]] = ts;
    list_invalidate();
    return PK_OK;
}

int pk_list_tiles(int32_t ddim, int32_t ifluid, int64_t cg_id, int32_t out[3])
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    if (ddim < 0 || ddim > 2 || !g_cfg.has_dir[ddim]) return fail(PK_EINVAL, "inactive or bad direction %d", ddim);
    if (ifluid < 0 || ifluid >= g_nfl) return fail(PK_EINVAL, "bad fluid %d", ifluid);
    RCCHK(list_build());
    for (const LaunchTab *t : {&g_list.tab[ddim][ifluid], &g_list.stab[ddim][ifluid]}) {
        for (size_t i = 0; i < t->blocks.size(); i++) {
            if (t->blocks[i]->id != cg_id) continue;
            out[0] = t->nt;
            out[1] = t->h_args[i].TP;
            out[2] = t->h_args[i].TS;
            return PK_OK;
        }
    }
    return fail(PK_ENOBLOCK, "unknown cg_id %lld", (long long)cg_id);
}

// ---- reverse mode -------------------------------------------------------------

int pk_adjoint_setup(int64_t cg_id, int32_t nslots)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    if (g_nfl != 1) return fail(PK_EINVAL, "the adjoint covers one fluid per block");
    if (g_cfg.integration_order != 2) return fail(PK_EINVAL, "the adjoint covers the RK2 sweep");
    if (block_linked(b)) return fail(PK_EINVAL, "the adjoint covers blocks periodic on themselves");
    if (nslots < 1) return fail(PK_EINVAL, "nslots = %d", nslots);
    for (int d = 0; d < 3; d++)
        if (g_cfg.has_dir[d] && b->g[d] < 4) return fail(PK_EINVAL, "the fused RK2 sweep needs nb >= 4");
    HIPCHK(hipStreamSynchronize(g_stream));
    free_adjoint(b);
    const size_t bytes = state_bytes(b);
    for (int i = 0; i < 2; i++) {
        HIPCHK(hipMalloc((void **)&b->lam[i], bytes));
        HIPCHK(hipMemsetAsync(b->lam[i], 0, bytes, g_stream));
    }
    b->slots.assign(nslots, nullptr);
    for (int i = 0; i < nslots; i++) HIPCHK(hipMalloc((void **)&b->slots[i], bytes));
    long long parts = 1;
    for (int d = 0; d < 3; d++) {
        if (!g_cfg.has_dir[d]) continue;
        int TP = 1, TS = 17, out = 1, chunks = 1, ptiles = 1;
        adj_tile(b, d, TP, TS, out, chunks, ptiles);
        const long long p = (long long)chunks * ptiles * 4;  // 4 waves per tile: 256-lane list workgroups too
        if (p > parts) parts = p;
    }
    HIPCHK(hipMalloc((void **)&b->d_dtb_part, sizeof(double) * (parts + pk::kRedParts)));
    HIPCHK(hipMalloc((void **)&b->d_chb_part, sizeof(double) * (parts + pk::kRedParts)));
    b->dtb_parts = parts;
    HIPCHK(hipMalloc((void **)&b->d_dtb, 2 * 6 * sizeof(double)));
    HIPCHK(hipMalloc((void **)&b->d_best, sizeof(unsigned long long)));
    b->lcur = 0;
    HIPCHK(hipStreamSynchronize(g_stream));
    return PK_OK;
}

int pk_state_save(int64_t cg_id, int32_t slot)
{
    Block *b;
    double *p;
    RCCHK(get_block(cg_id, &b));
    RCCHK(adj_check(b));
    RCCHK(slot_ptr(b, slot, &p));
    HIPCHK(hipMemcpyAsync(p, b->buf[b->cur], state_bytes(b), hipMemcpyDeviceToDevice, g_stream));
    return PK_OK;
}

int pk_state_load(int64_t cg_id, int32_t slot)
{
    Block *b;
    double *p;
    RCCHK(get_block(cg_id, &b));
    RCCHK(adj_check(b));
    RCCHK(slot_ptr(b, slot, &p));
    list_touch();
    HIPCHK(hipMemcpyAsync(b->buf[b->cur], p, state_bytes(b), hipMemcpyDeviceToDevice, g_stream));
    b->dt_pending = false;
    b->uh = -1;
    return PK_OK;
}

int pk_adjoint_set(int64_t cg_id, const double *lam_u, const double *lam_b, const double *lam_psi)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    RCCHK(adj_check(b));
    if (!lam_u || (g_mhd && (!lam_b || !lam_psi))) return fail(PK_EINVAL, "missing adjoint array");
    double *dst = b->lam[b->lcur];
    HIPCHK(hipMemsetAsync(dst, 0, state_bytes(b), g_stream));
    RCCHK(upload_state(b, lam_u, lam_b, lam_psi, dst + b->xo, true));
    // the ghost cells are periodic copies of interior cells
    HIPCHK(pk::launch_fold_ghosts(dst + b->xo, b->sv, b->sy, b->sz, b->ng, b->g, b->n, g_nplanes, g_stream));
    return PK_OK;
}

int pk_adjoint_get(int64_t cg_id, double *lam_u, double *lam_b, double *lam_psi)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    RCCHK(adj_check(b));
    return download_state(b, lam_u, lam_b, lam_psi, b->lam[b->lcur] + b->xo, true);
}

int pk_sweep_adjoint(int64_t cg_id, int32_t slot, int32_t ddim, double dt, double chspeed, double psi_fac,
                     double *dt_bar, double *ch_bar)
{
    Block *b;
    double *p;
    RCCHK(get_block(cg_id, &b));
    RCCHK(adj_check(b));
    RCCHK(slot_ptr(b, slot, &p));
    if (ddim < 0 || ddim > 2 || !g_cfg.has_dir[ddim]) return fail(PK_EINVAL, "inactive or bad direction %d", ddim);
    if (!(dt > 0.0)) return fail(PK_EINVAL, "dt must be positive");
    double dtb = 0.0, chb = 0.0;
    RCCHK(do_sweep_adjoint(b, p, ddim, dt, chspeed, psi_fac, &dtb, &chb));
    if (dt_bar) *dt_bar += dtb;
    if (ch_bar) *ch_bar += chb;
    return PK_OK;
}

int pk_fluid_update_record(int64_t cg_id, int32_t slot0, double dt)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    RCCHK(adj_check(b));
    if (!(dt > 0.0)) return fail(PK_EINVAL, "dt must be positive");
    const std::vector<int> dirs = step_dirs();
    if (slot0 < 0 || slot0 + (int)dirs.size() > (int)b->slots.size())
        return fail(PK_EINVAL, "slots %d..%d do not exist", slot0, slot0 + (int)dirs.size() - 1);
    list_touch();
    std::vector<bool> damp;
    step_dirs(&damp);
    double ch, fac;
    block_glm(b, dt, &ch, &fac);
    // the state before sweep i lands in slot0 + i: one copy, then each sweep
    // but the last writes the next slot directly (periodic images included)
    HIPCHK(hipMemcpyAsync(b->slots[slot0], b->buf[b->cur], state_bytes(b), hipMemcpyDeviceToDevice, g_stream));
    for (size_t i = 0; i + 1 < dirs.size(); i++)
        RCCHK(do_sweep(b, dirs[i], dt, ch, damp[i] ? fac : 1.0, true, false, b->slots[slot0 + i],
                       b->slots[slot0 + i + 1]));
    const size_t l = dirs.size() - 1;
    RCCHK(do_sweep(b, dirs[l], dt, ch, damp[l] ? fac : 1.0, true, true, b->slots[slot0 + l], b->buf[1 - b->cur]));
    b->cur = 1 - b->cur;
    b->uh = -1;
    b->dt_pending = false;
    HIPCHK(pk::launch_dtslots_final(b->d_dtslots, kDtSlots, b->d_dtout, g_stream));
    b->dt_pending = true;
    return PK_OK;
}

int pk_fluid_update_adjoint(int64_t cg_id, int32_t slot0, double dt, double *dt_bar)
{
    Block *b;
    RCCHK(get_block(cg_id, &b));
    RCCHK(adj_check(b));
    if (!(dt > 0.0)) return fail(PK_EINVAL, "dt must be positive");
    const std::vector<int> dirs = step_dirs();
    if (slot0 < 0 || slot0 + (int)dirs.size() > (int)b->slots.size())
        return fail(PK_EINVAL, "slots %d..%d do not exist", slot0, slot0 + (int)dirs.size() - 1);
    std::vector<bool> damp;
    step_dirs(&damp);
    double ch, fac;
    block_glm(b, dt, &ch, &fac);
    double dtb = 0.0, chb = 0.0;
    for (int i = (int)dirs.size() - 1; i >= 0; i--)
        RCCHK(do_sweep_adjoint(b, b->slots[slot0 + i], dirs[i], dt, ch, damp[i] ? fac : 1.0, nullptr, nullptr, i));
    // the sweeps' sums, read back once
    HIPCHK(hipMemcpyAsync(g_pinned, b->d_dtb, 2 * dirs.size() * sizeof(double), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    for (size_t i = 0; i < dirs.size(); i++) {
        dtb += g_pinned[2 * i];
        chb += g_pinned[2 * i + 1];
    }
    // ch = cfl_glm*min(dl)/dt (hdc.F90:122); the damping factor exp(-glm_alpha*ch*dt/min(dl))
    // does not depend on dt
    if (g_mhd) dtb += chb * (-(ch / dt));
    if (dt_bar) *dt_bar += dtb;
    return PK_OK;
}

int pk_timestep_adjoint(int64_t cg_id, int32_t slot, double dt_bar)
{
    Block *b;
    double *p;
    RCCHK(get_block(cg_id, &b));
    RCCHK(adj_check(b));
    RCCHK(slot_ptr(b, slot, &p));
    pk::DtArgs a{};
    a.u = p + b->xo;
    a.sv = b->sv;
    a.sy = b->sy;
    a.sz = b->sz;
    for (int d = 0; d < 3; d++) {
        a.ng[d] = b->ng[d];
        a.g[d] = b->g[d];
        a.n[d] = b->n[d];
        a.has_dir[d] = g_cfg.has_dir[d];
        a.dl[d] = b->dl[d];
    }
    a.plane0 = 0;
    a.bplane = 5;
    a.magnetized = g_mhd ? 1 : 0;
    a.gam = g_cfg.fluid[0].gamma;
    a.gam_1 = g_cfg.fluid[0].gamma - 1.0;
    HIPCHK(pk::launch_dt(a, g_part, b->d_dtout, g_stream));
    HIPCHK(hipMemcpyAsync(g_pinned, b->d_dtout, 4 * sizeof(double), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    double mn = g_pinned[0];
    if (g_pinned[1] < mn) mn = g_pinned[1];
    if (g_pinned[2] < mn) mn = g_pinned[2];
    b->dt_pending = false;
    pk::DtAdjArgs q{};
    q.u = a.u;
    q.magnetized = g_mhd ? 1 : 0;
    q.sv = b->sv;
    q.sy = b->sy;
    q.sz = b->sz;
    for (int d = 0; d < 3; d++) {
        q.n[d] = b->n[d];
        q.g[d] = b->g[d];
        q.has_dir[d] = g_cfg.has_dir[d];
        q.dl[d] = b->dl[d];
    }
    q.gam = a.gam;
    q.gam_1 = a.gam_1;
    q.mn = mn;
    q.cfl = g_cfg.cfl;
    q.dtb = dt_bar;
    HIPCHK(pk::launch_dt_adjoint(q, b->d_best, b->lam[b->lcur] + b->xo, g_stream));
    return PK_OK;
}

// ---- ensembles of independent instances -------------------------------------------

int pk_ensemble_step(int32_t mode, int32_t first, double *dt_out, int32_t *capped_out)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    RCCHK(check_mode(mode));
    RCCHK(ens_build());
    std::vector<double> dt;
    RCCHK(ens_dt_raw(dt));
    for (size_t i = 0; i < g_ens.nblk; i++) {
        // time_step caps (timestep.F90:200-208) with dt_initial = -1, dt_max_grow = 2
        int capped = 0;
        if (!first && g_ens.dt_old[i] > 0.0) {
            const double cap = g_ens.dt_old[i] * 2.0;
            if (cap < dt[i]) {
                dt[i] = cap;
                capped = 1;
            }
        }
        if (dt_out) dt_out[i] = dt[i];
        if (capped_out) capped_out[i] = capped;
        g_ens.dt_old[i] = dt[i];
    }
    RCCHK(ens_set_step(dt.data()));
    if (mode == PK_BATCH_GRAPH) {
        RCCHK(ens_graph());
        HIPCHK(hipGraphLaunch(g_ens.exec, g_stream));
    } else {
        pk::StepScalars sc{};
        RCCHK(enqueue_step(mode, sc, true));
    }
    for (Block *b : g_list.blocks) {
        b->uh = -1;
        b->dt_pending = false;
    }
    g_list.fresh = false;
    g_ens.fresh = true;
    return PK_OK;
}

int pk_ensemble_run(int32_t mode, int32_t nsteps, double *t_out, double *dt_out)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    if (nsteps < 0) return fail(PK_EINVAL, "nsteps < 0");
    RCCHK(ens_build());
    const size_t n = g_ens.nblk;
    std::vector<double> t(n, 0.0), dt(n, 0.0);
    for (int s = 0; s < nsteps; s++) {
        RCCHK(pk_ensemble_step(mode, s == 0, dt.data(), nullptr));
        for (size_t i = 0; i < n; i++) {
            t[i] = t[i] + dt[i];
            t[i] = t[i] + dt[i];
        }
    }
    HIPCHK(hipStreamSynchronize(g_stream));
    for (size_t i = 0; i < n; i++) {
        if (t_out) t_out[i] = t[i];
        if (dt_out) dt_out[i] = dt[i];
    }
    return PK_OK;
}

int pk_ensemble_state_save(int32_t slot)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    RCCHK(ens_build());
    return ens_copy_slots(slot, true);
}

int pk_ensemble_state_load(int32_t slot)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    RCCHK(ens_build());
    RCCHK(ens_copy_slots(slot, false));
    list_touch();
    for (Block *b : g_list.blocks) {
        b->dt_pending = false;
        b->uh = -1;
    }
    return PK_OK;
}

int pk_ensemble_record_step(const double *dts, int32_t slot0)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    if (!dts) return fail(PK_EINVAL, "null dts");
    RCCHK(ens_build());
    std::vector<bool> damp;
    const std::vector<int> dirs = step_dirs(&damp);
    RCCHK(ens_set_step(dts));
    HIPCHK(pk::launch_zero_u32(g_list.d_queue, kQueues, g_stream));
    for (size_t k = 0; k < dirs.size(); k++) {
        RCCHK(ens_copy_slots(slot0 + (int)k, true, (int)(k % 2)));
        const int d = dirs[k];
        for (int f = 0; f < g_nfl; f++) {
            for (int part = 0; part < 2; part++) {
                LaunchTab &t = part ? g_list.stab[d][f] : g_list.tab[d][f];
                if (t.nitems == 0) continue;
                pk::ListArgs L{};
                L.args = t.d_args;
                L.first = t.d_first;
                L.items = t.d_items;
                L.nitems = t.nitems;
                L.queue = g_list.d_queue + (int)((k * g_nfl + f) * 2 + part);
                L.grab = kListGrab;
                L.parity = (int)(k % 2);
                L.apply_fac = damp[k] && g_fmag[f];
                L.step = g_ens.d_step;
                L.dtslots = nullptr;
                L.nslots = kEnsSlots;
                L.per_block = 1;
                L.bpos = t.d_first + t.blocks.size() + 1;
                L.no_wrap = 1;
                HIPCHK(list_kernel(d, g_fmag[f], L, t.nt, t.grid));
            }
        }
    }
    list_touch();
    for (Block *b : g_list.blocks) {
        b->dt_pending = false;
        b->uh = -1;
    }
    return PK_OK;
}

int pk_ensemble_step_adjoint(const double *dts, int32_t slot0, double *dt_bar)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    if (!dts) return fail(PK_EINVAL, "null dts");
    if (g_nfl != 1) return fail(PK_EINVAL, "the adjoint covers one fluid per block");
    RCCHK(ens_build());
    RCCHK(ens_adj_build());
    const size_t n = g_ens.nblk;
    const int lc = g_list.blocks[0]->lcur;
    for (Block *b : g_list.blocks) {
        RCCHK(adj_check(b));
        if (b->lcur != lc) return fail(PK_ESTATE, "the blocks' adjoint states are out of step");
    }
    std::vector<bool> damp;
    const std::vector<int> dirs = step_dirs(&damp);
    std::vector<double> ch(n, 0.0), fac(n, 1.0);
    for (size_t i = 0; i < n; i++)
        if (g_mhd) glm_scalars({g_list.mdl[i]}, dts[i], &ch[i], &fac[i]);
    for (int k = (int)dirs.size() - 1; k >= 0; k--) {
        const int d = dirs[k];
        HIPCHK(hipStreamSynchronize(g_stream));  // the pinned tables may still feed the previous copy
        for (size_t i = 0; i < n; i++) {
            Block *b = g_list.blocks[i];
            if (slot0 + k >= (int)b->slots.size()) return fail(PK_EINVAL, "slot %d does not exist", slot0 + k);
            pk::AdjArgs a{};
            a.u = b->slots[slot0 + k] + b->xo;
            a.lam_out = b->lam[b->lcur] + b->xo;
            a.lam_in = b->lam[1 - b->lcur] + b->xo;
            a.dtb_part = b->d_dtb_part;
            a.chb_part = b->d_chb_part;
            a.sv = b->sv;
            a.sy = b->sy;
            a.sz = b->sz;
            for (int q = 0; q < 3; q++) {
                a.ng[q] = b->ng[q];
                a.g[q] = b->g[q];
                a.n[q] = b->n[q];
            }
            adj_tile(b, d, a.TP, a.TS, a.out, a.chunks, a.ptiles);
            a.lim = g_cfg.limiter;
            a.limb = g_cfg.limiter_b;
            a.gam = g_cfg.fluid[0].gamma;
            a.gam_1 = g_cfg.fluid[0].gamma - 1.0;
            a.c1 = (0.5 * dts[i]) / b->dl[d];
            a.c2 = (1.0 * dts[i]) / b->dl[d];
            a.dl = b->dl[d];
            a.ch = ch[i];
            a.psi_fac = damp[k] ? fac[i] : 1.0;
            a.smalld = g_cfg.smalld;
            a.smallei = g_cfg.smallei;
            a.use_smalld = g_cfg.use_smalld;
            a.use_smallei = g_cfg.use_smallei;
            g_ens.h_adj[i] = a;
            const long long parts = (long long)a.chunks * a.ptiles * 4;
            g_ens.h_sum[2 * i] = pk::SumJob{b->d_dtb_part, parts, g_ens.d_sums + 12 * i + 2 * k};
            g_ens.h_sum[2 * i + 1] = pk::SumJob{b->d_chb_part, parts, g_ens.d_sums + 12 * i + 2 * k + 1};
        }
        HIPCHK(hipMemcpyAsync(g_ens.d_adj, g_ens.h_adj, sizeof(pk::AdjArgs) * n, hipMemcpyHostToDevice, g_stream));
        HIPCHK(hipMemcpyAsync(g_ens.d_sum, g_ens.h_sum.data(), sizeof(pk::SumJob) * 2 * n, hipMemcpyHostToDevice,
                              g_stream));
        HIPCHK(pk::launch_zero_u32(g_ens.d_queue, 1, g_stream));
        pk::AdjListArgs L{};
        L.args = g_ens.d_adj;
        L.first = g_ens.d_afirst[d];
        L.items = g_ens.d_aitems[d];
        L.nitems = g_ens.anitems[d];
        L.queue = g_ens.d_queue;
        HIPCHK(pk::launch_sweep_adjoint_list(d, g_mhd, g_cfg.limiter, g_cfg.limiter_b, L, g_stream));
        HIPCHK(pk::launch_sum_parts_jobs(g_ens.d_sum, (int)(2 * n), g_stream));
        for (Block *b : g_list.blocks) b->lcur = 1 - b->lcur;
    }
    HIPCHK(hipMemcpyAsync(g_ens.h_sums, g_ens.d_sums, sizeof(double) * 12 * n, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    for (size_t i = 0; i < n; i++) {
        double dtb = 0.0, chb = 0.0;
        for (size_t k = 0; k < dirs.size(); k++) {
            dtb += g_ens.h_sums[12 * i + 2 * k];
            chb += g_ens.h_sums[12 * i + 2 * k + 1];
        }
        // ch = cfl_glm*min(dl)/dt (hdc.F90:122)
        if (g_mhd) dtb += chb * (-(ch[i] / dts[i]));
        if (dt_bar) dt_bar[i] += dtb;
    }
    return PK_OK;
}

int pk_ensemble_timestep_adjoint(int32_t slot, const double *dt_bar)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    if (!dt_bar) return fail(PK_EINVAL, "null dt_bar");
    if (g_nfl != 1) return fail(PK_EINVAL, "the adjoint covers one fluid per block");
    RCCHK(ens_build());
    const size_t n = g_ens.nblk;
    if (!g_ens.d_dtj) {
        HIPCHK(hipMalloc((void **)&g_ens.d_dtj, sizeof(pk::DtBatchJob) * n));
        HIPCHK(hipMalloc((void **)&g_ens.d_mn, sizeof(double) * n));
        HIPCHK(hipMalloc((void **)&g_ens.d_best, sizeof(unsigned long long) * n));
        g_ens.h_dtj.resize(n);
    }
    // timestep_fluid of each block at the state in slot: all blocks in three launches
    HIPCHK(hipStreamSynchronize(g_stream));  // h_dtj may still feed the previous copy
    for (size_t i = 0; i < n; i++) {
        Block *b = g_list.blocks[i];
        RCCHK(adj_check(b));
        if (slot < 0 || slot >= (int)b->slots.size()) return fail(PK_EINVAL, "slot %d does not exist", slot);
        pk::DtBatchJob j{};
        j.a.u = b->slots[slot] + b->xo;
        j.a.magnetized = g_mhd ? 1 : 0;
        j.a.sv = b->sv;
        j.a.sy = b->sy;
        j.a.sz = b->sz;
        for (int d = 0; d < 3; d++) {
            j.a.n[d] = b->n[d];
            j.a.g[d] = b->g[d];
            j.a.has_dir[d] = g_cfg.has_dir[d];
            j.a.dl[d] = b->dl[d];
        }
        j.a.gam = g_cfg.fluid[0].gamma;
        j.a.gam_1 = g_cfg.fluid[0].gamma - 1.0;
        j.a.cfl = g_cfg.cfl;
        j.dtb = dt_bar[i];
        j.lam = b->lam[b->lcur] + b->xo;
        j.mn = g_ens.d_mn + i;
        j.best = g_ens.d_best + i;
        g_ens.h_dtj[i] = j;
    }
    HIPCHK(hipMemcpyAsync(g_ens.d_dtj, g_ens.h_dtj.data(), sizeof(pk::DtBatchJob) * n, hipMemcpyHostToDevice,
                          g_stream));
    HIPCHK(pk::launch_dt_adjoint_batch(g_ens.d_dtj, (int)n, g_stream));
    for (Block *b : g_list.blocks) b->dt_pending = false;
    return PK_OK;
}

int pk_ensemble_ids(int64_t *ids, int32_t n)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    RCCHK(list_build());
    if (n < (int32_t)g_list.blocks.size()) return fail(PK_EINVAL, "room for %d ids, %d blocks", n,
                                                      (int)g_list.blocks.size());
    for (size_t i = 0; i < g_list.blocks.size(); i++) ids[i] = g_list.blocks[i]->id;
    return PK_OK;
}

int pk_profile_begin(void)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    for (auto &r : g_prof_recs) {
        g_event_pool.push_back(r.start);
        g_event_pool.push_back(r.stop);
    }
    g_prof_recs.clear();
    g_prof = true;
    return PK_OK;
}

int pk_profile_end(pk_profile_stats *out)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    HIPCHK(hipStreamSynchronize(g_stream));
    pk_profile_stats st{};
    for (auto &r : g_prof_recs) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, r.start, r.stop));
        st.launches[r.dir] += 1;
        st.ms[r.dir] += ms;
        st.algo_bytes[r.dir] += r.bytes;
    }
    g_prof = false;
    if (out) *out = st;
    return PK_OK;
}

int pk_synchronize(void)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    HIPCHK(hipStreamSynchronize(g_stream));
    return PK_OK;
}

}  // extern "C"
