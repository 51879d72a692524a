/*
 * piernik_hip.h -- C-ABI of the MI355X (gfx950) split Riemann sweep.
 *
 * Drop-in for PIERNIK's per-block solver operator.  The reference binds
 *   procedure(solve_cg_sub), pointer :: solve_cg          (src/scheme/sweeps.F90:167)
 * to solve_cg_riemann (src/scheme/Riemann_Solver/solve_cg_riemann.F90:68-107)
 * and calls it once per block, direction and RK stage (sweeps.F90:245).
 * pk_solve_cg() below is that call; the other entry points carry the state
 * the reference keeps in Fortran module globals (flind, global, hdc) and the
 * device-resident stepping that replaces the host loops around it.
 *
 * Conventions
 *   - Plain C types only; every entry point returns 0 on success and a
 *     PK_E* code otherwise; pk_last_error() gives the message (the Fortran
 *     adapter passes it to die(), as the reference does on any error).
 *   - Host arrays use the reference layout (column-major, ghosts included):
 *       u(nvar, nxg, nyg, nzg), b(3, nxg, nyg, nzg), psi(nxg, nyg, nzg)
 *     with nxg = nx + 2*nb in every active direction (src/base/named_array.F90).
 *   - The library owns device mirrors of registered blocks; they stay
 *     resident between calls.  pk_upload / pk_download move data explicitly.
 *   - Directions: 0 = x, 1 = y, 2 = z (the reference's xdim..zdim minus one).
 *   - RK stages use the reference's enumerators: EULER = 10000,
 *     RK2_1 = 10001, RK2_2 = 10002 (src/base/constants.F90:230-238).
 *   - One host thread drives the library; all work is ordered on one HIP
 *     stream.  Calls that return values to the host synchronise that stream.
 */
#ifndef PIERNIK_HIP_H
#define PIERNIK_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PK_OK 0
#define PK_EINVAL 1      /* bad argument or configuration */
#define PK_ENOBLOCK 2    /* unknown cg_id */
#define PK_EHIP 3        /* HIP runtime error */
#define PK_ESTATE 4      /* call out of order (e.g. before pk_init) */
#define PK_ENOMEM 5
#define PK_EIO 6         /* file I/O (HDF5, libpiernik_gdf) */

#define PK_EULER 10000
#define PK_RK2_1 10001
#define PK_RK2_2 10002

/* limiter names of src/scheme/Riemann_Solver/fluxlimiters.F90:90-103 */
#define PK_LIM_VANLEER 0
#define PK_LIM_MINMOD 1
#define PK_LIM_MONCEN 2
#define PK_LIM_SUPERBEE 3

/* One fluid of flind (src/fluids/fluidtypes.F90:279-343).  Adiabatic only. */
typedef struct pk_fluid {
    double gamma;          /* fl%gam */
    int32_t magnetized;    /* fl%is_magnetized: HLLD + GLM (riemann_hlld); else HLLC (riemann_hlld_u) */
    int32_t reserved;
} pk_fluid;

/* The &NUMERICAL_SETUP knobs the sweep reads (src/base/global.F90:118-120,205-262). */
typedef struct pk_config {
    int32_t nfluids;           /* 1 or 2 (e.g. ionized + neutral); at most one magnetized
                                  (solve_cg_riemann.F90:85-87).  Fluid f occupies components
                                  5f..5f+4 of u, as flind%all_fluids(f)%fl%beg */
    int32_t limiter;           /* 'limiter'   -> PK_LIM_* */
    int32_t limiter_b;         /* 'limiter_b' -> PK_LIM_* */
    int32_t integration_order; /* 1 (EULER) or 2 (RK2) */
    pk_fluid fluid[2];
    double smalld;             /* density floor (sources.F90:357) */
    double smallei;            /* internal-energy floor (sources.F90:426) */
    int32_t use_smalld;
    int32_t use_smallei;
    double cfl;                /* CFL number of time_step (timestep.F90:492) */
    double cfl_glm;            /* GLM speed factor, ch = cfl_glm*min(dl)/dt (hdc.F90:122) */
    double glm_alpha;          /* GLM damping (hdc.F90:325) */
    int32_t has_dir[3];        /* dom%has_dir: active directions */
    int32_t tile_x[2];         /* sweep tile shape (pencils, cells) along x; 0 = autotuned default;
                                  cells = -1: one lane per whole pencil (march kernel);
                                  cells = -m <= -2: the stream kernel, pencils per lane group
                                  (1, 2, 4 or 8), about m iterations per workgroup */
    int32_t tile_yz[2];        /* sweep tile shape (pencils, cells) along y and z; 0 = default */
    int32_t fast_math;         /* 0: every operation rounds as the reference's FP64 code (bit-exact);
                                  1: FMA contraction and fast division/sqrt in the split sweeps and
                                     the unsplit passes (~1e-15 per op) */
    int32_t reserved;
} pk_config;

/* Floor bookkeeping accumulated since the last reset
 * (dn_negative, ei_negative: sources.F90:350-351,425; magic mass :356-358). */
typedef struct pk_stage_report {
    int32_t dn_negative;
    int32_t ei_negative;
    double magic_mass[2];  /* per fluid: mass added by the density floor, interior cells
                              (local_magic_mass, mass_defect.F90) */
} pk_stage_report;

/* Execution of the block-list entry points (pk_*_all). */
#define PK_BATCH_GRAPH 0       /* one persistent launch per sweep and fluid, the step captured in a HIP graph */
#define PK_BATCH_PERSISTENT 1  /* the same launches issued directly on the stream */
#define PK_BATCH_PER_BLOCK 2   /* one launch per block, sweep and fluid, same tile shapes */

/* ---- life cycle ---------------------------------------------------------- */

/* init_fluids/init_global/set_interpolations in one call (initpiernik.F90:190-195). */
int pk_init(const pk_config *cfg);
int pk_finalize(void);
const char *pk_last_error(void);

/* ---- block registry (grid containers) ------------------------------------ */

/* Register one cg: n = interior cells per direction, nb = ghost layers,
 * dl = cell sizes.  Host pointers may be NULL (device-only block); psi and
 * psih are required when a fluid is magnetized.  Allocates device mirrors. */
int pk_register_block(int64_t cg_id, const int32_t n[3], int32_t nb, const double dl[3],
                      double *u, double *uh, double *b, double *bh, double *psi, double *psih);
int pk_unregister_block(int64_t cg_id);

/* Host <-> device copies of the registered host arrays (u, b, psi). */
int pk_upload(int64_t cg_id);
int pk_download(int64_t cg_id);
/* Same from / to caller-given host arrays in the reference layout. */
int pk_upload_from(int64_t cg_id, const double *u, const double *b, const double *psi);
int pk_download_to(int64_t cg_id, double *u, double *b, double *psi);
/* Same from / to caller-given DEVICE arrays in the reference layout (e.g.
 * tensors of shape (nzg, nyg, nxg, nvar) on this GPU); no host staging. */
int pk_upload_device(int64_t cg_id, const double *u, const double *b, const double *psi);
int pk_download_device(int64_t cg_id, double *u, double *b, double *psi);

/* One output field of the data files over the block's interior, computed on
 * the device from its current state as datafields_hdf5 computes it
 * (data_hdf5.F90:345-691; ekin, emag, sq_sum3 of func.F90:55-88), into
 * out[n_b(1)*n_b(2)*n_b(3)] in the reference's (nx, ny, nz) order.  Names are
 * the reference's short names (hdf_vars): den?, vlx? vly? vlz?, momx? momy?
 * momz?, ene?, pre?, eth? for ? = i (the ionized fluid: the magnetized one,
 * else the first) or n (the neutral one: the other of two fluids, the only
 * one of a hydro run), v (|m|/density of the ionized fluid, else the
 * neutral), and with a magnetic field magx magy magz, magB and psi.  pre?
 * and eth? of the ionized fluid subtract the magnetic energy (cell-centred
 * B).  The GDF writer (piernik_gdf.h) calls this per block and field. */
int pk_datafield(int64_t cg_id, const char *name, double *out);
/* The same rounded to nearest FP32 on the device (the conversion HDF5 applies
 * when FP64 values are written to the reference's default FP32 datasets). */
int pk_datafield_f32(int64_t cg_id, const char *name, float *out);

/* ---- the operator --------------------------------------------------------- */

/* solve_cg_sub(cg, ddim, istep): one RK stage of the split Riemann update of
 * one block along ddim, on the device mirror (sweeps.F90:140-151 interface;
 * solve_cg_riemann.F90:68-307 semantics: at the first stage the state is
 * kept as uh, every transverse-interior pencil is updated over its whole
 * length except the two end cells, floors apply to all cells).  chspeed is
 * the GLM speed of hdc.F90:90-140. */
int pk_solve_cg(int64_t cg_id, int32_t ddim, int32_t istep, double dt, double chspeed);

/* Fine/coarse flux points of pk_solve_cg on a refined level: one side of
 * cg%finebnd(ddim, side) (kind PK_FC_FINEBND: fluxes a finer neighbour sent,
 * which replace the block's flux at their face) or cg%coarsebnd(ddim, side)
 * (PK_FC_COARSEBND: the block's own flux at a face bordering a coarser
 * neighbour, saved for it), as set_fluxpointers and save_outfluxes use them
 * (grid_container_fcflx.F90:148-235; flx_arr.F90:38-42 for the arrays;
 * applied in solve, solve_cg_riemann.F90:353-363 and :405-408).  The host
 * arrays stay the caller's, in the reference layout over the block's
 * interior pencils, (i1, i2) along pdims(ddim, ORTHO1:ORTHO2)
 * (constants.F90:127-128: x -> (y, z), y -> (z, x), z -> (x, y)):
 * index(n1, n2) (int32, the reference's cell index), uflx(nvar, n1, n2) and,
 * with a magnetic field, bflx(4, n1, n2) = [B, psi] fluxes in sweep order.
 * first is cg%ijkse(ddim, LO), the index of the block's first interior cell:
 * a LO side is active on a pencil where index >= first, a HI side where
 * index <= first + n - 1.  Each pk_solve_cg along ddim reads the finebnd
 * arrays and, on return, has written the coarsebnd fluxes of the active
 * pencils.  index == NULL removes the side.  The fused and batched paths
 * (pk_sweep, pk_fluid_update, pk_run, the block-list and unsplit functions)
 * refuse a block with flux points. */
#define PK_FC_FINEBND 0
#define PK_FC_COARSEBND 1
int pk_register_fcflx(int64_t cg_id, int32_t ddim, int32_t side, int32_t kind, int64_t first,
                      const int32_t *index, double *uflx, double *bflx);

/* Both RK2 stages of one directional sweep in one pass (requires nb >= 4
 * and ghost cells consistent with the neighbours along ddim, as after
 * update_boundaries).  psi_fac multiplies psi on output (GLM damping fused
 * into the last sweep of a substep; 1 otherwise). */
int pk_sweep(int64_t cg_id, int32_t ddim, double dt, double chspeed, double psi_fac);

/* pk_sweep over caller-owned device memory, allocating nothing.  It has no
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  src and dst are distinct, non-overlapping device
 * arrays holding every plane of one block in the library's device layout:
 * plane p of ghosted cell (x, y, z) at [p*sv + x + sy*y + sz*z], the planes
 * of each fluid in flind order (rho, m_x, m_y, m_z, E, then B_x, B_y, B_z,
 * psi after the magnetized fluid's five), n interior cells and nb >= 4 ghost
 * layers in each active direction (n = 1, no ghosts, in an inactive one).
 * dst receives pk_sweep's interior result (both RK2 stages along ddim, psi
 * times psi_fac); its ghost cells are not written.  report (device, may be
 * NULL) receives 5 + nfluids doubles: min over the cells written of dl_d / c
 * for d = x, y, z and max c (timestep_fluid of the new state, cfl not
 * applied; as pk_timestep_fluid combines them), the floor flags (bit 0
 * negative density, bit 1 negative internal energy) and per fluid the mass
 * the density floor added divided by the cell volume.  The work is queued on
 * `stream` (a hipStream_t; NULL: the library's stream) and not waited for.
 * Scratch: one static device array of 33 KiB per GPU (timestep slots, floor
 * counters, flags), held until the process ends; its size does not depend on
 * the block, so no block size is refused for it.  Every call on a GPU uses
 * that GPU's one array: calls that overlap on one GPU race. */
int pk_sweep_noalloc(const double *src, double *dst, int64_t sv, int32_t sy, int32_t sz, const int32_t n[3],
                     int32_t nb, const double dl[3], int32_t ddim, double dt, double chspeed, double psi_fac,
                     double *report, void *stream);

/* Ghost refresh of all variables of one block, all directions, corners
 * included (update_boundaries, sweeps.F90:66-107): periodic on itself, or
 * for a block with neighbours (pk_set_neighbour) the exchange over the
 * whole registered list from the neighbours' current state. */
int pk_fill_ghosts(int64_t cg_id);

/* Connect blocks of the registered list (internal boundaries,
 * cg_list_bnd.F90:169-322, which the reference finds from the grid's
 * segment lists, cg%i_bnd / cg%o_bnd): the ghost layers of cg_id on side
 * `side` (0 low, 1 high) of direction `dir` (0 x, 1 y, 2 z) take the
 * adjoining nb interior layers of block nb_id (across the domain for a
 * periodic boundary).  A block registered is its own neighbour on every side
 * (periodic on itself).  Both blocks must share the face (the same ghosted
 * extents in the other two directions), nb and the cell size (no fine/coarse
 * boundaries).  A list with such links is advanced only by the block-list
 * functions (pk_fluid_update_all, pk_run_all), which exchange the faces each
 * sweep reads before that sweep; pk_sweep, pk_fluid_update, pk_run and the
 * reverse-mode and ensemble functions refuse a block with neighbours.
 * Downloads refresh every ghost cell first. */
int pk_set_neighbour(int64_t cg_id, int32_t dir, int32_t side, int64_t nb_id);

/* psi *= fac over the whole block (glmdamping, hdc.F90:338). */
int pk_glm_damp(int64_t cg_id, double fac);

/* timestep_fluid for every fluid of one block (timestep.F90:431-495):
 * dt_out = cfl * min over cells and directions of dl/c, c_out = max c. */
int pk_timestep_fluid(int64_t cg_id, double *dt_out, double *c_out);

/* fluid_update_full on one periodic block (fluidupdate.F90:99-136):
 * update_chspeed, X->Y->Z sweeps, GLM damping, Z->Y->X sweeps, GLM damping. */
int pk_fluid_update(int64_t cg_id, double dt);

/* nsteps iterations of the main loop (piernik.F90:124-170) with
 * cflcontrol='warn': time_step, then fluid_update.  *t_out = simulated time,
 * *dt_out = the last dt. */
int pk_run(int64_t cg_id, int32_t nsteps, double *t_out, double *dt_out);

/* Floor bookkeeping of one block since the last reset. */
int pk_get_report(int64_t cg_id, pk_stage_report *rep, int32_t reset);

/* Wait for all queued work. */
int pk_synchronize(void);

/* ---- the whole block list --------------------------------------------------- */
/* The registered blocks (ascending cg_id) advanced together, each periodic on
 * itself, with one dt for all: the block loops of sweep()
 * (sweeps.F90:230-260), time_step (timestep.F90:112-240: min over blocks),
 * update_chspeed (hdc.F90:112-140: min over blocks) and glmdamping
 * (hdc.F90:313-331: one factor from the max over blocks).  RK2 only.
 * mode is a PK_BATCH_* value; every mode gives bit-identical results. */

/* timestep_fluid over the list: dt = cfl * min over blocks, fluids, cells and
 * directions of dl/c; c_out = max c. */
int pk_timestep_all(double *dt_out, double *c_out);
/* fluid_update_full over the list with the given dt. */
int pk_fluid_update_all(int32_t mode, double dt);
/* nsteps main-loop iterations over the list (as pk_run for one block). */
int pk_run_all(int32_t mode, int32_t nsteps, double *t_out, double *dt_out);

/* ---- ensembles of independent instances (no counterpart in the reference) -- */
/* The registered blocks (ascending cg_id; pk_ensemble_ids lists them) as
 * independent problems: each takes its own time_step (its own dt, caps, ch
 * and GLM damping), all advanced by the same persistent launches as the list
 * entry points.  Arrays below have one entry per block in that order. */
int pk_ensemble_ids(int64_t *ids, int32_t n);
/* One main-loop iteration of every block; first != 0: no growth cap (the
 * first step of a run).  dt_out, capped_out (may be NULL): each block's dt
 * and whether dt_max_grow capped it. */
int pk_ensemble_step(int32_t mode, int32_t first, double *dt_out, int32_t *capped_out);
int pk_ensemble_run(int32_t mode, int32_t nsteps, double *t_out, double *dt_out);
/* every block's state <-> its slot (pk_adjoint_setup on every block) */
int pk_ensemble_state_save(int32_t slot);
int pk_ensemble_state_load(int32_t slot);
/* one step of every block with the given dts, keeping the state before each
 * sweep in slots slot0, slot0+1, ... */
int pk_ensemble_record_step(const double *dts, int32_t slot0);
/* the recorded step of every block in reverse (one persistent adjoint launch
 * per sweep); dt_bar[i] += dL/d(dt_i), the dependence of ch on dt included */
int pk_ensemble_step_adjoint(const double *dts, int32_t slot0, double *dt_bar);
/* lam_i += dt_bar[i] * d(dt_i)/du_i at the state in slot (timestep_fluid) */
int pk_ensemble_timestep_adjoint(int32_t slot, const double *dt_bar);

/* ---- step retries (cflcontrol) --------------------------------------------- */

#define PK_CFL_WARN 0      /* 'warn' / 'none': steps are never redone */
#define PK_CFL_REDO 1      /* 'redo': redo with dt_full * dt_shrink**attempt */
#define PK_CFL_FLEXIBLE 2  /* 'flexible': redo with dt_full * cfl_max/stepcfl (times dt_shrink from attempt 2) */
#define PK_CFL_AUTO 3      /* 'auto' / 'adaptive': dt scaled by cfl_c from the change of c_max (cfl_auto) */

/* cflcontrol, cfl_max, dt_shrink, max_redostep_attempts and
 * disallow_negatives of &NUMERICAL_SETUP (global.F90:118-120,216-258;
 * timestep.F90:73-90).  With PK_CFL_REDO or PK_CFL_FLEXIBLE, pk_run and
 * pk_run_all save the state before every step and redo the step as
 * repeat_fluidstep / check_cfl_violation / cfl_warn do
 * (timestep_retry.F90:86-273, timestep.F90:247-361): when c_max of the new
 * state over c_max of the old, times cfl and the current shrink factor,
 * exceeds cfl_max (sanitized as global.F90:304 does), or when the floors met
 * negative values and disallow_negatives is set.  With PK_CFL_AUTO no step
 * is redone: each step's dt is multiplied by cfl_c = min(1, (cfl_c +
 * min(1, stepcfl_old/stepcfl*dt/dt_old))/2), stepcfl = c_max/c_max_old*cfl
 * (cfl_auto, timestep.F90:368-405; cfl_c starts at 1 with each pk_run).
 * Default PK_CFL_WARN.  The backups take one extra state per block. */
int pk_set_cfl_control(int32_t mode, double cfl_max, double dt_shrink, int32_t max_attempts,
                       int32_t disallow_negatives);

/* dt_initial, dt_max_grow, dt_min, dt_max of &NUMERICAL_SETUP (global.F90:
 * 243-247, defaults -1, 2, tiny, huge) as time_step applies them to pk_run
 * and pk_run_all (timestep.F90:198-224).  A call is one main loop from
 * nstep = 0: dt_initial > 0 caps its first dt, dt_initial < 0 scales it by
 * -dt_initial, later steps grow by at most dt_max_grow.  A dt below dt_min
 * fails the call (the reference writes a crash dump).  tend is the simulated
 * time left at the start of the call: dt <= (tend - t)/2 (+2 epsilon) as
 * the split scheme clamps it, and the loop ends once t >= tend
 * (piernik.F90:125).  pk_init restores the defaults (tend = huge). */
int pk_set_time_control(double dt_initial, double dt_max_grow, double dt_min, double dt_max, double tend);

/* Steps redone by the last pk_run / pk_run_all (no counterpart in the reference;
 * its log prints "Redoing previous step"). */
int pk_redone_steps(int32_t *count);

/* ---- the unsplit Riemann solver (RIEMANN_UNSPLIT) --------------------------- */

/* One RK stage of the unsplit solver on a block periodic on itself:
 * solve_cg_unsplit(cg, istep) (src/scheme/unsplit_schemes/solve_cg_unsplit.F90:14-45,
 * called at unsplit_sweeps.F90:161) -- the interface fluxes of every active
 * direction from one state (u at RK2_1, uh at RK2_2), their differences added
 * x, y, z in turn to uh (a copy of u) or to u, B and psi included, the
 * vanishing sources and, at the last stage, care_for_positives once per
 * direction in its sweep order -- followed by update_boundaries of the
 * stage's arrays (unsplit_sweeps.F90:47-82, periodic).  One fluid; nb >= 2. */
int pk_unsplit_stage(int64_t cg_id, int32_t istep, double dt, double chspeed);

/* fluid_update_unsplit (unsplit_fluidupdate.F90:45-126) on a periodic block:
 * c_h = cfl_glm*min(dl)/dt, psi damped by exp(-glm_alpha*c_h*(dt/2)/min(dl))
 * before and after the stages. */
int pk_fluid_update_unsplit(int64_t cg_id, double dt);

/* The main loop with the unsplit solver until t >= tend or nmax steps:
 * time_step with its caps and the unsplit clamp dt <= (tend - t)*(1 + 2 eps)
 * (timestep.F90:198-222), then pk_fluid_update_unsplit; t += dt. */
int pk_run_unsplit(int64_t cg_id, double tend, int32_t nmax, double *t_out, int32_t *nsteps_out);

/* ---- tuning (no counterpart in the reference) ------------------------------ */

/* Time candidate tile shapes of the fused sweep on this block (the state is
 * read, never advanced) and keep the fastest per direction for every block
 * of the same shape and physics.  dt and chspeed should be representative.
 * tiles_out (may be NULL) receives (pencils, cells) for x, y, z; 0 for an
 * inactive direction; cells = -1 is the march kernel (one lane per whole
 * pencil, 64 pencils per workgroup), a candidate for y and z sweeps of
 * blocks with at least 65536 pencils; cells = -m <= -2 is the stream kernel
 * (a one-wave workgroup walks the pencils of m * 64 stream positions, each
 * cell's two stages computed once).  A shape given in pk_config (tile_x,
 * tile_yz; (64, -1) selects the march kernel, (tp, -m) the stream kernel)
 * overrides the tuned one. */
int pk_autotune(int64_t cg_id, double dt, double chspeed, int32_t tiles_out[6]);

/* Tile shapes of the persistent list sweep: per direction and fluid kind the
 * workgroup size (64 lanes: the stream form, one-wave workgroups walking
 * ranges of a block's pencils; 128, 256 or 512 lanes: tiles) and, per block
 * shape, the tile length (stream: iterations per item), timed on the
 * registered list (the state is read, never advanced; the floor bookkeeping
 * of every block is reset).  chspeed <= 0: the list's own update_chspeed
 * for this dt. */
int pk_autotune_all(double dt, double chspeed);
/* Sets the list sweep's shape for one direction and the kind (magnetized or
 * not) of fluid ifluid, for every block, until the next pk_autotune_all:
 * lanes 64 (the stream form; ts = -m, m >= 2 iterations per item) or 128,
 * 256, 512 (tiles of ts cells with the 8 halo cells, 9 <= ts <= lanes);
 * ts = 0: the default for each block.  No counterpart in the reference (a
 * launch-shape control, as tile_x and tile_yz of pk_config_t). */
int pk_set_list_tiles(int32_t ddim, int32_t ifluid, int32_t lanes, int32_t ts);
/* Sweeps the blocks whose interior shape is one of shapes[0..3*nshapes) by a
 * second launch of the stream form (ts = -m iterations per item, 0: the
 * default), the others as before, for one direction and fluid kind, until
 * the next pk_autotune_all or pk_set_list_tiles (nshapes = 0: one launch
 * again).  pk_autotune_all makes this split itself where the per-shape
 * times favour it.  No counterpart in the reference (a launch-shape
 * control). */
int pk_set_list_split(int32_t ddim, int32_t ifluid, const int32_t *shapes, int32_t nshapes, int32_t ts);
/* The list sweep's workgroup lanes and tile for one block, direction and
 * fluid: out = {lanes, pencils, cells} (stream form: {64, pencils per lane
 * group, -iterations per item}). */
int pk_list_tiles(int32_t ddim, int32_t ifluid, int64_t cg_id, int32_t out[3]);

/* ---- reverse mode (no counterpart in the reference) ------------------------ */
/* Adjoint of the fused RK2 sweep for one fluid (HLLC, or HLLD + GLM) on
 * blocks that are periodic on themselves: the limiter and wave-speed branches are
 * differentiated along the branch the forward took (exact ties of min/max
 * split evenly), floors included.  The adjoint state lam has the state's
 * layout; each call below moves it backwards through one operation,
 * linearised at a state saved in a slot. */

/* Allocate nslots state slots and the adjoint state of a block. */
int pk_adjoint_setup(int64_t cg_id, int32_t nslots);
/* slot <- current state; current state <- slot. */
int pk_state_save(int64_t cg_id, int32_t slot);
int pk_state_load(int64_t cg_id, int32_t slot);
/* lam <- dL/d(u, b, psi) from DEVICE arrays in the reference layout
 * (b and psi of the magnetized fluid; NULL for hydro); entries of ghost
 * cells are added to their periodic images.  pk_adjoint_get writes lam back
 * (ghost entries 0). */
int pk_adjoint_set(int64_t cg_id, const double *lam_u, const double *lam_b, const double *lam_psi);
int pk_adjoint_get(int64_t cg_id, double *lam_u, double *lam_b, double *lam_psi);
/* fluid_update_full with the given dt, saving the state before each of its
 * sweeps into slots slot0, slot0+1, ... (6 in 3D); results are those of
 * pk_fluid_update. */
int pk_fluid_update_record(int64_t cg_id, int32_t slot0, double dt);
/* lam <- (d sweep / d u)^T lam for the sweep pk_sweep(ddim, dt, chspeed,
 * psi_fac) linearised at the state in slot; *dt_bar += dL/d(dt) at fixed
 * chspeed, *ch_bar += dL/d(chspeed) (either may be NULL). */
int pk_sweep_adjoint(int64_t cg_id, int32_t slot, int32_t ddim, double dt, double chspeed, double psi_fac,
                     double *dt_bar, double *ch_bar);
/* The sweeps of a step recorded by pk_fluid_update_record, in reverse;
 * *dt_bar includes the dependence of chspeed on dt (hdc.F90:122). */
int pk_fluid_update_adjoint(int64_t cg_id, int32_t slot0, double dt, double *dt_bar);
/* lam += dt_bar * d(dt)/du for dt = cfl * min(dl/c) (timestep_fluid) at
 * the state in slot, through the minimising cell and direction. */
int pk_timestep_adjoint(int64_t cg_id, int32_t slot, double dt_bar);

/* ---- measurement (no counterpart in the reference) ------------------------ */

/* Totals of the kernels launched between pk_profile_begin() and
 * pk_profile_end(), timed with HIP events recorded on the library's stream
 * around each launch: entries 0-2 the sweep kernels of x, y, z (algo_bytes:
 * every interior cell read once and written once), entry 3 the output-field
 * kernel of pk_datafield (algo_bytes: the planes the field reads and the
 * field written, once per interior cell). */
typedef struct pk_profile_stats {
    int64_t launches[4];
    double ms[4];
    double algo_bytes[4];
} pk_profile_stats;

int pk_profile_begin(void);
int pk_profile_end(pk_profile_stats *out);

#ifdef __cplusplus
}
#endif
#endif
