// pk_capi.cpp -- host side of the C-ABI declared in include/piernik_hip.h.
//
// Keeps the per-run configuration (the reference's flind / &NUMERICAL_SETUP
// globals), a registry of grid blocks with device mirrors, and drives the
// kernels of pk_kernels.hip on one HIP stream.  The stepping driver
// (pk_fluid_update, pk_run) restates fluid_update_full / make_3sweeps /
// sweep / time_step (src/scheme/fluidupdate.F90:99-232,
// src/scheme/sweeps.F90:114-279, src/base/timestep.F90:112-240) for one
// periodic block with cflcontrol='warn'.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>
#include <string>

#include "../../include/piernik_hip.h"
#include "pk_kernels.hpp"

namespace {

struct Block {
    int64_t id = 0;
    int n[3] = {1, 1, 1}, nb = 0, g[3] = {0, 0, 0}, ng[3] = {1, 1, 1};
    double dl[3] = {1, 1, 1};
    int nplanes = 5;
    long long sv = 0;
    int sy = 0, sz = 0;
    double *buf[2] = {nullptr, nullptr};
    int xo = 0;  // leading offset that puts every interior row start on a 128 B line
    int cur = 0;
    int uh = -1;  // buffer holding uh after a first stage, -1 when none
    double *h_u = nullptr, *h_uh = nullptr, *h_b = nullptr, *h_bh = nullptr, *h_psi = nullptr,
           *h_psih = nullptr;
    unsigned *d_flags = nullptr;
    double *d_dmass = nullptr;
    double *d_part = nullptr;
    double *d_dtout = nullptr;
    double *d_stage = nullptr;  // AoS staging for host transfers
    unsigned long long *d_dtslots = nullptr;  // timestep reduction written by the last sweep of a step
    bool dt_pending = false;                  // d_dtout holds timestep_fluid of the current state
};

const int kDtSlots = 1024;

pk_config g_cfg;
bool g_init = false;
bool g_mhd = false;
std::map<int64_t, Block *> g_blocks;
hipStream_t g_stream = nullptr;
std::string g_err;
double *g_pinned = nullptr;  // small pinned buffer for scalar read-backs

// Tile shapes chosen by pk_autotune, per block shape, direction and physics.
struct TuneKey {
    int n0, n1, n2, dir, mhd, lim, limb;
    bool operator<(const TuneKey &o) const
    {
        const int a[7] = {n0, n1, n2, dir, mhd, lim, limb}, b[7] = {o.n0, o.n1, o.n2, o.dir, o.mhd, o.lim, o.limb};
        for (int i = 0; i < 7; i++)
            if (a[i] != b[i]) return a[i] < b[i];
        return false;
    }
};
std::map<TuneKey, std::pair<int, int>> g_tuned;

TuneKey tune_key(const Block *b, int d)
{
    return TuneKey{b->n[0], b->n[1], b->n[2], d, (g_mhd ? 1 : 0) + (g_cfg.fast_math ? 2 : 0), g_cfg.limiter,
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

// Address of cell (0,0,0) of plane 0 of state buffer i.
double *state(Block *b, int i) { return b->buf[i] + b->xo; }

int get_block(int64_t id, Block **b)
{
    if (!g_init) return fail(PK_ESTATE, "pk_init has not been called");
    auto it = g_blocks.find(id);
    if (it == g_blocks.end()) return fail(PK_ENOBLOCK, "unknown cg_id %lld", (long long)id);
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

void fill_common(const Block *b, pk::SweepArgs &a)
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
    a.gam = g_cfg.fluid[0].gamma;
    a.gam_1 = g_cfg.fluid[0].gamma - 1.0;
    a.smalld = g_cfg.smalld;
    a.smallei = g_cfg.smallei;
    a.use_smalld = g_cfg.use_smalld;
    a.use_smallei = g_cfg.use_smallei;
    a.flags = b->d_flags;
    a.dmass = b->d_dmass;
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

// Tile shape along direction d: TP pencils x TS cells, within the launch
// bound (512 lanes MHD, 1024 HD); halo = 4 (fused) or 2 (one stage).
// Tile shape along direction d: TP pencils x TS cells (TP*TS <= 512 lanes),
// halo 4 (fused) or 2 (one stage).  Precedence: the shape given in pk_config,
// then the one pk_autotune measured for this block shape, then 3 x 72
// (4-wave workgroups, two per CU; measured best at 256^3 MHD).
void tile_geometry(const Block *b, int d, bool fused, int tp, int ts, int &TP, int &TS, int &out, int &chunks,
                   int &tiles1)
{
    const int maxt = 512;
    const int halo = fused ? 4 : 2;
    const int span = fused ? b->n[d] : b->ng[d];  // cells that receive output
    const int d1 = (d == 0) ? 1 : 0;
    if (ts > span + 2 * halo) ts = span + 2 * halo;
    if (ts > maxt) ts = maxt;
    if (tp <= 0) tp = maxt / ts;
    if (tp < 1) tp = 1;
    if (tp > b->n[d1]) tp = b->n[d1];
    while (tp * ts > maxt && tp > 1) tp--;
    if (tp * ts > maxt) ts = maxt / tp;
    if (ts < 2 * halo + 1) ts = 2 * halo + 1;
    TP = tp;
    TS = ts;
    out = ts - 2 * halo;
    chunks = (span + out - 1) / out;
    tiles1 = (b->n[d1] + tp - 1) / tp;
}

void choose_tile(const Block *b, int d, bool fused, int &TP, int &TS, int &out, int &chunks, int &tiles1)
{
    const int *ov = (d == 0) ? g_cfg.tile_x : g_cfg.tile_yz;
    int tp = ov[0], ts = ov[1];
    if (ts <= 0 && fused) {
        auto it = g_tuned.find(tune_key(b, d));
        if (it != g_tuned.end()) {
            tp = it->second.first;
            ts = it->second.second;
        }
    }
    if (ts <= 0) {
        ts = 72;
        if (tp <= 0) tp = 3;
    }
    tile_geometry(b, d, fused, tp, ts, TP, TS, out, chunks, tiles1);
}

hipError_t sweep_kernel(int d, bool fused, const pk::SweepArgs &a)
{
    return g_cfg.fast_math ? pk_fast::launch_sweep(d, g_mhd, fused, a, g_stream)
                           : pk::launch_sweep(d, g_mhd, fused, a, g_stream);
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

int launch(int d, bool fused, const pk::SweepArgs &a0, Block *b)
{
    pk::SweepArgs a = a0;
    choose_tile(b, d, fused, a.TP, a.TS, a.out, a.chunks, a.tiles1);
    ProfRec rec{};
    if (g_prof) {
        int rc;
        if ((rc = take_event(&rec.start)) || (rc = take_event(&rec.stop))) return rc;
        rec.dir = d;
        const double cells = (double)b->n[0] * b->n[1] * b->n[2];
        rec.bytes = cells * b->nplanes * 8.0 * 2.0;
        HIPCHK(hipEventRecord(rec.start, g_stream));
    }
    HIPCHK(sweep_kernel(d, fused, a));
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
                       dst, b->sv, b->nplanes, b->sy, b->sz, d, b->ng[0], b->ng[1], b->ng[2], b->g[0], b->g[1],
                       b->g[2], b->n[0], b->n[1], b->n[2]);
    HIPCHK(hipGetLastError());
    return PK_OK;
}

int do_stage(Block *b, int ddim, int istep, double dt, double ch)
{
    b->dt_pending = false;
    pk::SweepArgs a{};
    fill_common(b, a);
    a.ch = ch;
    a.c1 = (rk_coef(istep) * dt) / b->dl[ddim];
    a.c2 = a.c1;
    const bool first = istep == first_stage();
    double *src = state(b, b->cur);
    double *other = state(b, 1 - b->cur);
    if (first) {
        a.src = src;
        a.base = src;
        a.dst = other;
        int rc = launch(ddim, false, a, b);
        if (rc) return rc;
        if ((rc = copy_transverse_ghosts(b, ddim, src, other))) return rc;
        b->uh = b->cur;
        b->cur = 1 - b->cur;
    } else {
        if (b->uh < 0 || b->uh == b->cur)
            return fail(PK_ESTATE, "stage %d called without a preceding first stage", istep);
        double *uh = state(b, b->uh);
        a.src = src;
        a.base = uh;
        a.dst = uh;
        int rc = launch(ddim, false, a, b);
        if (rc) return rc;
        if ((rc = copy_transverse_ghosts(b, ddim, src, uh))) return rc;
        b->cur = b->uh;
        b->uh = -1;
    }
    return PK_OK;
}

// One directional sweep, both RK2 stages.  images: write the periodic ghost
// images too (the block is periodic on itself); want_dt: reduce the next
// step's timestep_fluid from the cells written.
int do_sweep(Block *b, int ddim, double dt, double ch, double psi_fac, bool images = false, bool want_dt = false)
{
    if (g_cfg.integration_order == 1) {
        int rc = do_stage(b, ddim, PK_EULER, dt, ch);
        if (rc) return rc;
        if (psi_fac != 1.0 && g_mhd) {
            HIPCHK(pk::launch_scale(state(b, b->cur) + 8 * b->sv, b->sz * b->ng[2], psi_fac, g_stream));
        }
        return PK_OK;
    }
    if (b->g[ddim] < 4) return fail(PK_EINVAL, "fused RK2 sweep needs nb >= 4 (nb = %d)", b->g[ddim]);
    pk::SweepArgs a{};
    fill_common(b, a);
    a.ch = ch;
    a.c1 = (0.5 * dt) / b->dl[ddim];
    a.c2 = (1.0 * dt) / b->dl[ddim];
    a.psi_fac = psi_fac;
    a.src = state(b, b->cur);
    a.base = a.src;
    a.dst = state(b, 1 - b->cur);
    a.images = images ? 1 : 0;
    if (want_dt) {
        HIPCHK(pk::launch_dtslots_init(b->d_dtslots, kDtSlots, g_stream));
        a.dtslots = b->d_dtslots;
    }
    int rc = launch(ddim, true, a, b);
    if (rc) return rc;
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
    a.nplanes = b->nplanes;
    for (int d = 0; d < 3; d++)
        if (b->g[d] > 0) HIPCHK(pk::launch_fill_ghosts(d, a, g_stream));
    return PK_OK;
}

int do_timestep(Block *b, double *dt_out, double *c_out)
{
    double dt = DBL_MAX, c = 0.0;
    if (b->dt_pending && g_cfg.nfluids == 1) {
        // reduced by the last sweep from the cells it wrote (same formula, same state)
        HIPCHK(hipMemcpyAsync(g_pinned, b->d_dtout, 4 * sizeof(double), hipMemcpyDeviceToHost, g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
        double m = g_pinned[0];
        if (g_pinned[1] < m) m = g_pinned[1];
        if (g_pinned[2] < m) m = g_pinned[2];
        *dt_out = g_cfg.cfl * m;
        *c_out = g_pinned[3];
        return PK_OK;
    }
    for (int f = 0; f < g_cfg.nfluids; f++) {
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
        a.plane0 = 5 * f;
        a.bplane = 5;
        a.magnetized = g_cfg.fluid[f].magnetized;
        a.gam = g_cfg.fluid[f].gamma;
        a.gam_1 = g_cfg.fluid[f].gamma - 1.0;
        HIPCHK(pk::launch_dt(a, b->d_part, b->d_dtout, g_stream));
        HIPCHK(hipMemcpyAsync(g_pinned, b->d_dtout, 4 * sizeof(double), hipMemcpyDeviceToHost, g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
        double m = g_pinned[0];
        if (g_pinned[1] < m) m = g_pinned[1];
        if (g_pinned[2] < m) m = g_pinned[2];
        const double dtf = g_cfg.cfl * m;  // timestep.F90:492
        if (dtf < dt) dt = dtf;
        if (g_pinned[3] > c) c = g_pinned[3];
    }
    *dt_out = dt;
    *c_out = c;
    return PK_OK;
}

int do_fluid_update(Block *b, double dt)
{
    double ch = 0.0, fac = 1.0;
    if (g_mhd) {
        // update_chspeed with ch_grid (hdc.F90:112-122) and the glmdamping factor (hdc.F90:313-331)
        const double mdl = min_dl(b);
        ch = DBL_MAX;
        const double c = (g_cfg.cfl_glm * mdl) / dt;
        if (c < ch) ch = c;
        double f = 0.0;
        const double ff = (g_cfg.glm_alpha * ch) / (mdl / dt);
        if (ff > f) f = ff;
        fac = std::exp(-f);
    }
    for (int half = 0; half < 2; half++) {
        int last = -1;
        for (int s = 0; s < 3; s++) {
            const int d = half == 0 ? s : 2 - s;
            if (g_cfg.has_dir[d]) last = d;
        }
        for (int s = 0; s < 3; s++) {
            const int d = half == 0 ? s : 2 - s;
            if (!g_cfg.has_dir[d]) continue;
            int rc;
            if (g_cfg.integration_order == 2) {
                // the sweep refreshes the periodic ghosts itself; the last one of
                // the step also reduces the next step's timestep
                rc = do_sweep(b, d, dt, ch, (d == last) ? fac : 1.0, true, half == 1 && d == last);
            } else {
                rc = do_sweep(b, d, dt, ch, (d == last) ? fac : 1.0);
                if (!rc) rc = do_fill(b);
            }
            if (rc) return rc;
        }
    }
    return PK_OK;
}

int alloc_block(Block *b)
{
    b->xo = (16 - b->g[0] % 16) % 16;
    const int pitch = (b->xo + b->ng[0] + 15) / 16 * 16;
    b->sy = pitch;
    b->sz = pitch * b->ng[1];
    long long sv = (long long)b->xo + (long long)b->sz * b->ng[2];
    sv = (sv + 31) / 32 * 32;
    b->sv = sv;
    const size_t bytes = sizeof(double) * (size_t)sv * b->nplanes;
    for (int i = 0; i < 2; i++) {
        HIPCHK(hipMalloc((void **)&b->buf[i], bytes));
        HIPCHK(hipMemsetAsync(b->buf[i], 0, bytes, g_stream));
    }
    HIPCHK(hipMalloc((void **)&b->d_flags, sizeof(unsigned)));
    HIPCHK(hipMalloc((void **)&b->d_dmass, sizeof(double)));
    HIPCHK(hipMemsetAsync(b->d_flags, 0, sizeof(unsigned), g_stream));
    HIPCHK(hipMemsetAsync(b->d_dmass, 0, sizeof(double), g_stream));
    HIPCHK(hipMalloc((void **)&b->d_part, sizeof(double) * 4 * 2048));
    HIPCHK(hipMalloc((void **)&b->d_dtout, sizeof(double) * 4));
    HIPCHK(hipMalloc((void **)&b->d_dtslots, sizeof(unsigned long long) * 4 * kDtSlots));
    const size_t ncell = (size_t)b->ng[0] * b->ng[1] * b->ng[2];
    HIPCHK(hipMalloc((void **)&b->d_stage, sizeof(double) * ncell * 5 * g_cfg.nfluids));
    return PK_OK;
}

void free_block(Block *b)
{
    for (int i = 0; i < 2; i++)
        if (b->buf[i]) (void)hipFree(b->buf[i]);
    if (b->d_flags) (void)hipFree(b->d_flags);
    if (b->d_dmass) (void)hipFree(b->d_dmass);
    if (b->d_part) (void)hipFree(b->d_part);
    if (b->d_dtout) (void)hipFree(b->d_dtout);
    if (b->d_stage) (void)hipFree(b->d_stage);
    if (b->d_dtslots) (void)hipFree(b->d_dtslots);
    delete b;
}

int upload_array(Block *b, const double *h, int nc, int plane0, double *dst)
{
    const size_t ncell = (size_t)b->ng[0] * b->ng[1] * b->ng[2];
    HIPCHK(hipMemcpyAsync(b->d_stage, h, sizeof(double) * ncell * nc, hipMemcpyHostToDevice, g_stream));
    HIPCHK(pk::launch_aos_to_soa(b->d_stage, nc, dst, b->sv, plane0, b->ng, b->sy, b->sz, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return PK_OK;
}

int download_array(Block *b, double *h, int nc, int plane0, const double *src)
{
    const size_t ncell = (size_t)b->ng[0] * b->ng[1] * b->ng[2];
    HIPCHK(pk::launch_soa_to_aos(src, b->sv, plane0, nc, b->d_stage, b->ng, b->sy, b->sz, g_stream));
    HIPCHK(hipMemcpyAsync(h, b->d_stage, sizeof(double) * ncell * nc, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return PK_OK;
}

int upload_state(Block *b, const double *u, const double *bb, const double *psi, double *dst)
{
    int rc;
    b->dt_pending = false;
    if (u && (rc = upload_array(b, u, 5 * g_cfg.nfluids, 0, dst))) return rc;
    if (g_mhd) {
        if (bb && (rc = upload_array(b, bb, 3, 5, dst))) return rc;
        if (psi && (rc = upload_array(b, psi, 1, 8, dst))) return rc;
    }
    return PK_OK;
}

int download_state(Block *b, double *u, double *bb, double *psi, const double *src)
{
    int rc;
    if (u && (rc = download_array(b, u, 5 * g_cfg.nfluids, 0, src))) return rc;
    if (g_mhd) {
        if (bb && (rc = download_array(b, bb, 3, 5, src))) return rc;
        if (psi && (rc = download_array(b, psi, 1, 8, src))) return rc;
    }
    return PK_OK;
}

}  // namespace

extern "C" {

const char *pk_last_error(void) { return g_err.c_str(); }

int pk_init(const pk_config *cfg)
{
    if (!cfg) return fail(PK_EINVAL, "null config");
    if (cfg->nfluids != 1) return fail(PK_EINVAL, "nfluids = %d: this build runs one fluid per block", cfg->nfluids);
    if (cfg->limiter < 0 || cfg->limiter > 3 || cfg->limiter_b < 0 || cfg->limiter_b > 3)
        return fail(PK_EINVAL, "unknown limiter");
    if (cfg->integration_order != 1 && cfg->integration_order != 2)
        return fail(PK_EINVAL, "integration_order must be 1 or 2");
    if (!(cfg->fluid[0].gamma > 1.0)) return fail(PK_EINVAL, "adiabatic fluid needs gamma > 1");
    if (!(cfg->has_dir[0] || cfg->has_dir[1] || cfg->has_dir[2])) return fail(PK_EINVAL, "no active direction");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ndev < 1) return fail(PK_EHIP, "no HIP device");
    if (!g_stream) HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    if (!g_pinned) HIPCHK(hipHostMalloc((void **)&g_pinned, 8 * sizeof(double), hipHostMallocDefault));
    g_cfg = *cfg;
    g_mhd = cfg->fluid[0].magnetized != 0;
    g_init = true;
    return PK_OK;
}

int pk_finalize(void)
{
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
        blk->n[d] = n[d];
        blk->g[d] = g_cfg.has_dir[d] ? nb : 0;
        blk->ng[d] = n[d] + 2 * blk->g[d];
        blk->dl[d] = dl[d];
    }
    blk->nplanes = g_mhd ? 9 : 5;
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
    return PK_OK;
}

int pk_unregister_block(int64_t cg_id)
{
    Block *b;
    int rc = get_block(cg_id, &b);
    if (rc) return rc;
    (void)hipStreamSynchronize(g_stream);
    g_blocks.erase(cg_id);
    free_block(b);
    return PK_OK;
}

int pk_upload(int64_t cg_id)
{
    Block *b;
    int rc = get_block(cg_id, &b);
    if (rc) return rc;
    if (!b->h_u) return fail(PK_EINVAL, "block %lld has no host arrays", (long long)cg_id);
    if (g_mhd && (!b->h_b || !b->h_psi)) return fail(PK_EINVAL, "magnetized block needs b and psi host arrays");
    return upload_state(b, b->h_u, b->h_b, b->h_psi, state(b, b->cur));
}

int pk_download(int64_t cg_id)
{
    Block *b;
    int rc = get_block(cg_id, &b);
    if (rc) return rc;
    if (!b->h_u) return fail(PK_EINVAL, "block %lld has no host arrays", (long long)cg_id);
    if ((rc = download_state(b, b->h_u, b->h_b, b->h_psi, state(b, b->cur)))) return rc;
    if (b->uh >= 0 && b->h_uh) return download_state(b, b->h_uh, b->h_bh, b->h_psih, state(b, b->uh));
    return PK_OK;
}

int pk_upload_from(int64_t cg_id, const double *u, const double *bb, const double *psi)
{
    Block *b;
    int rc = get_block(cg_id, &b);
    if (rc) return rc;
    if (!u || (g_mhd && (!bb || !psi))) return fail(PK_EINVAL, "missing host array");
    return upload_state(b, u, bb, psi, state(b, b->cur));
}

int pk_download_to(int64_t cg_id, double *u, double *bb, double *psi)
{
    Block *b;
    int rc = get_block(cg_id, &b);
    if (rc) return rc;
    return download_state(b, u, bb, psi, state(b, b->cur));
}

int pk_solve_cg(int64_t cg_id, int32_t ddim, int32_t istep, double dt, double chspeed)
{
    Block *b;
    int rc = get_block(cg_id, &b);
    if (rc) return rc;
    if (ddim < 0 || ddim > 2 || !g_cfg.has_dir[ddim]) return fail(PK_EINVAL, "inactive or bad direction %d", ddim);
    const bool ok = (g_cfg.integration_order == 1) ? istep == PK_EULER : (istep == PK_RK2_1 || istep == PK_RK2_2);
    if (!ok) return fail(PK_EINVAL, "istep %d does not match integration_order %d", istep, g_cfg.integration_order);
    if (!(dt > 0.0)) return fail(PK_EINVAL, "dt must be positive");
    return do_stage(b, ddim, istep, dt, chspeed);
}

int pk_sweep(int64_t cg_id, int32_t ddim, double dt, double chspeed, double psi_fac)
{
    Block *b;
    int rc = get_block(cg_id, &b);
    if (rc) return rc;
    if (ddim < 0 || ddim > 2 || !g_cfg.has_dir[ddim]) return fail(PK_EINVAL, "inactive or bad direction %d", ddim);
    if (!(dt > 0.0)) return fail(PK_EINVAL, "dt must be positive");
    return do_sweep(b, ddim, dt, chspeed, psi_fac);
}

int pk_fill_ghosts(int64_t cg_id)
{
    Block *b;
    int rc = get_block(cg_id, &b);
    if (rc) return rc;
    return do_fill(b);
}

int pk_glm_damp(int64_t cg_id, double fac)
{
    Block *b;
    int rc = get_block(cg_id, &b);
    if (rc) return rc;
    if (!g_mhd) return PK_OK;
    HIPCHK(pk::launch_scale(state(b, b->cur) + 8 * b->sv, (long long)b->sz * b->ng[2], fac, g_stream));
    return PK_OK;
}

int pk_timestep_fluid(int64_t cg_id, double *dt_out, double *c_out)
{
    Block *b;
    int rc = get_block(cg_id, &b);
    if (rc) return rc;
    if (!dt_out || !c_out) return fail(PK_EINVAL, "null output");
    return do_timestep(b, dt_out, c_out);
}

int pk_fluid_update(int64_t cg_id, double dt)
{
    Block *b;
    int rc = get_block(cg_id, &b);
    if (rc) return rc;
    if (!(dt > 0.0)) return fail(PK_EINVAL, "dt must be positive");
    return do_fluid_update(b, dt);
}

int pk_run(int64_t cg_id, int32_t nsteps, double *t_out, double *dt_out)
{
    Block *b;
    int rc = get_block(cg_id, &b);
    if (rc) return rc;
    if (nsteps < 0) return fail(PK_EINVAL, "nsteps < 0");
    double t = 0.0, dt = -1.0, dt_old = -1.0;
    if ((rc = do_fill(b))) return rc;
    for (int s = 0; s < nsteps; s++) {
        double c;
        if ((rc = do_timestep(b, &dt, &c))) return rc;
        // time_step caps (timestep.F90:200-208) with dt_initial = -1, dt_max_grow = 2
        if (s >= 1 && dt_old > 0.0) {
            const double cap = dt_old * 2.0;
            if (cap < dt) dt = cap;
        }
        if ((rc = do_fluid_update(b, dt))) return rc;
        t = t + dt;
        t = t + dt;
        dt_old = dt;
    }
    HIPCHK(hipStreamSynchronize(g_stream));
    if (t_out) *t_out = t;
    if (dt_out) *dt_out = dt;
    return PK_OK;
}

int pk_get_report(int64_t cg_id, pk_stage_report *rep, int32_t reset)
{
    Block *b;
    int rc = get_block(cg_id, &b);
    if (rc) return rc;
    unsigned fl = 0;
    double dm = 0.0;
    HIPCHK(hipMemcpyAsync(&fl, b->d_flags, sizeof fl, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipMemcpyAsync(&dm, b->d_dmass, sizeof dm, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    if (rep) {
        double dvol = 1.0;
        for (int d = 0; d < 3; d++)
            if (g_cfg.has_dir[d]) dvol *= b->dl[d];
        rep->dn_negative = (fl & 1u) ? 1 : 0;
        rep->ei_negative = (fl & 2u) ? 1 : 0;
        rep->magic_mass = dm * dvol;
    }
    if (reset) {
        HIPCHK(hipMemsetAsync(b->d_flags, 0, sizeof(unsigned), g_stream));
        HIPCHK(hipMemsetAsync(b->d_dmass, 0, sizeof(double), g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
    }
    return PK_OK;
}

int pk_autotune(int64_t cg_id, double dt, double chspeed, int32_t tiles_out[6])
{
    Block *b;
    int rc = get_block(cg_id, &b);
    if (rc) return rc;
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
        pk::SweepArgs a{};
        fill_common(b, a);
        a.ch = chspeed;
        a.c1 = (0.5 * dt) / b->dl[d];
        a.c2 = (1.0 * dt) / b->dl[d];
        a.src = state(b, b->cur);
        a.base = a.src;
        a.dst = state(b, 1 - b->cur);
        std::vector<std::pair<int, int>> seen;
        float best = 1e30f;
        std::pair<int, int> best_t(0, 0);
        for (int ts : ts_list) {
            for (int w : waves_list) {
                int TP, TS, out, chunks, tiles1;
                int tp = (64 * w) / (ts < b->n[d] + 8 ? ts : b->n[d] + 8);
                if (tp < 1) continue;
                tile_geometry(b, d, true, tp, ts, TP, TS, out, chunks, tiles1);
                if (TP * TS > 512) continue;
                bool dup = false;
                for (auto &p : seen) dup = dup || (p.first == TP && p.second == TS);
                if (dup) continue;
                seen.push_back({TP, TS});
                a.TP = TP;
                a.TS = TS;
                a.out = out;
                a.chunks = chunks;
                a.tiles1 = tiles1;
                HIPCHK(sweep_kernel(d, true, a));  // warm-up
                HIPCHK(hipEventRecord(e0, g_stream));
                for (int rep = 0; rep < 3; rep++) HIPCHK(sweep_kernel(d, true, a));
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
        g_tuned[tune_key(b, d)] = best_t;
        if (tiles_out) {
            tiles_out[2 * d] = best_t.first;
            tiles_out[2 * d + 1] = best_t.second;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // the candidates wrote the other buffer only; the state is unchanged.
    // The floor bookkeeping they accumulated is discarded.
    HIPCHK(hipMemsetAsync(b->d_flags, 0, sizeof(unsigned), g_stream));
    HIPCHK(hipMemsetAsync(b->d_dmass, 0, sizeof(double), g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
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
