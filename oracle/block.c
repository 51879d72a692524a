/*
 * block.c -- block-level part of the oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Restates, for one grid block ("cg") with periodic boundaries:
 *   solve_cg_riemann / solve_cg_ub / solve_cg_u
 *                       src/scheme/Riemann_Solver/solve_cg_riemann.F90:68-307
 *   set_fluxpointers / save_outfluxes (fine/coarse flux points)
 *                       src/grid/grid_container_fcflx.F90:148-235
 *   get_sweep + sweep permutations
 *                       src/base/named_array.F90:327-348,
 *                       src/fluids/fluidtypes.F90:318-320, fluidindex.F90:190-192
 *   periodic ghost refresh (update_boundaries on a single periodic block)
 *                       src/scheme/sweeps.F90:66-107, src/grid/cg_list_bnd.F90:251-328
 *   update_chspeed, glmdamping
 *                       src/scheme/Riemann_Solver/hdc.F90:90-140, 289-349
 *   time_step, timestep_fluid, ion_cs, neu_cs
 *                       src/base/timestep.F90:112-240, 431-497,
 *                       src/fluids/ionized/initionized.F90:89-145,
 *                       src/fluids/neutral/initneutral.F90:82-109
 *   fluid_update_full, make_3sweeps, sweep
 *                       src/scheme/fluidupdate.F90:99-232, src/scheme/sweeps.F90:114-279
 * with cflcontrol='warn' (no step retries), tend never binding, and no
 * source terms other than the vanishing internal ones.
 */
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define EULER 10000
#define RK2_1 10001
#define RK2_2 10002

static const double small_c = 1.17549435082228750797e-38; /* tiny(real(1.0,4)), constants.F90:104 */

static int gh(const or_params *p, const or_block *cg, int d) { return p->has_dir[d] ? cg->nb : 0; }

static void gdims(const or_params *p, const or_block *cg, int ng[3])
{
    for (int d = 0; d < 3; d++) ng[d] = cg->n[d] + 2 * gh(p, cg, d);
}

static int any_mag(const or_params *p)
{
    for (int f = 0; f < p->nfluids; f++)
        if (p->fl[f].magnetized) return 1;
    return 0;
}

/* index of sweep-order variable k of fluid f in xyz order (involutions) */
static int swp_var(int ddim, int k)
{
    static const int perm[3][5] = {{0, 1, 2, 3, 4}, {0, 2, 1, 3, 4}, {0, 3, 2, 1, 4}};
    return perm[ddim][k];
}
static int swp_mag(int ddim, int k)
{
    static const int perm[3][3] = {{0, 1, 2}, {1, 0, 2}, {2, 1, 0}};
    return perm[ddim][k];
}

static double rk_coef(int istep) { return istep == RK2_1 ? 0.5 : 1.0; } /* constants.F90:237-238 */
static int first_stage(const or_params *p) { return p->integration_order == 1 ? EULER : RK2_1; }

/* solve_cg_riemann for one block, one direction, one stage (in place). */
void or_solve_cg(const or_params *p, or_block *cg, int ddim, int istep,
                 double dt, double chspeed, or_report *rep)
{
    or_solve_cg_fc(p, cg, ddim, istep, dt, chspeed, rep, NULL);
}

/*
 * set_fluxpointers for one pencil (grid_container_fcflx.F90:148-209, the
 * cartesian case): a side is active where its index lies inside the
 * interior, ijkse(ddim, LO) for LO sides and ijkse(ddim, HI) for HI sides.
 * With c the pencil's local cell of that index, li replaces the flux at
 * c+1/2, ri at c-1/2, lo saves the flux at c-1/2 and ro at c+1/2.
 */
static void fc_pencil(const or_params *p, const or_block *cg, int ddim, const or_fcflx *fc, int o1, int o2,
                      or_peflx *e)
{
    static const int ortho[3][2] = {{1, 2}, {2, 0}, {0, 1}}; /* pdims(:, ORTHO1:ORTHO2) */
    const int n1 = cg->n[ortho[ddim][0]];
    const size_t at = (size_t)o1 + (size_t)n1 * o2;
    const int g = gh(p, cg, ddim), n = cg->n[ddim];
    const int mhd = any_mag(p);
    for (int k = 0; k < 4; k++) {
        const or_fcface *f = (k < 2) ? &fc->fine[k] : &fc->coarse[k - 2];
        const int side = k & 1; /* 0 LO, 1 HI */
        e->row[k] = -1;
        e->u[k] = e->b[k] = NULL;
        if (!f->index) continue;
        const int64_t idx = f->index[at];
        const int active = side == 0 ? idx >= fc->first : idx <= fc->first + n - 1;
        if (!active) continue;
        const int c = (int)(idx - fc->first) + g;
        const int right = (k == 0 || k == 3); /* li and ro act on the right face of c */
        e->row[k] = right ? c : c - 1;
        e->u[k] = f->uflx + at * (size_t)p->nvar;
        e->b[k] = mhd ? f->bflx + at * 4 : NULL;
    }
}

/* or_solve_cg with the block's fine/coarse flux arrays (fc may be NULL):
 * the pencils' external fluxes as set_fluxpointers / save_outfluxes give
 * them (solve_cg_riemann.F90:201,224,289,300). */
void or_solve_cg_fc(const or_params *p, or_block *cg, int ddim, int istep,
                    double dt, double chspeed, or_report *rep, const or_fcflx *fc)
{
    int ng[3];
    gdims(p, cg, ng);
    int nv = p->nvar, mhd = any_mag(p);
    int n = ng[ddim];
    int t1 = (ddim == 0) ? 1 : 0, t2 = (ddim == 2) ? 1 : 2;
    size_t sd[3] = {1, (size_t)ng[0], (size_t)ng[0] * ng[1]};
    double dtodx = (rk_coef(istep) * dt) / cg->dl[ddim];
    double dvol = 1.0;
    for (int d = 0; d < 3; d++)
        if (p->has_dir[d]) dvol *= cg->dl[d];
    int first = (istep == first_stage(p));
    int nbp = gh(p, cg, ddim);

    double *u0 = (double *)malloc(sizeof(double) * (size_t)n * nv);
    double *u = (double *)malloc(sizeof(double) * (size_t)n * nv);
    double *u1 = (double *)malloc(sizeof(double) * (size_t)n * nv);
    double *b0 = (double *)malloc(sizeof(double) * (size_t)n * 4);
    double *b = (double *)malloc(sizeof(double) * (size_t)n * 4);
    double *b1 = (double *)malloc(sizeof(double) * (size_t)n * 4);

    for (int c2 = gh(p, cg, t2); c2 < gh(p, cg, t2) + cg->n[t2]; c2++) {
        for (int c1 = gh(p, cg, t1); c1 < gh(p, cg, t1) + cg->n[t1]; c1++) {
            size_t base = (size_t)c1 * sd[t1] + (size_t)c2 * sd[t2];
            for (int i = 0; i < n; i++) {
                size_t cell = base + (size_t)i * sd[ddim];
                if (first) {
                    for (int v = 0; v < nv; v++) cg->uh[cell * nv + v] = cg->u[cell * nv + v];
                    if (mhd) {
                        for (int c = 0; c < 3; c++) cg->bh[cell * 3 + c] = cg->b[cell * 3 + c];
                        cg->psih[cell] = cg->psi[cell];
                    }
                }
                for (int f = 0; f < p->nfluids; f++) {
                    int o = p->fl[f].beg;
                    for (int k = 0; k < 5; k++) {
                        u0[(size_t)(o + k) * n + i] = cg->uh[cell * nv + o + swp_var(ddim, k)];
                        u[(size_t)(o + k) * n + i] = cg->u[cell * nv + o + swp_var(ddim, k)];
                    }
                }
                if (mhd) {
                    for (int k = 0; k < 3; k++) {
                        b0[(size_t)k * n + i] = cg->bh[cell * 3 + swp_mag(ddim, k)];
                        b[(size_t)k * n + i] = cg->b[cell * 3 + swp_mag(ddim, k)];
                    }
                    b0[(size_t)3 * n + i] = cg->psih[cell];
                    b[(size_t)3 * n + i] = cg->psi[cell];
                }
            }
            or_peflx e, *pe = NULL;
            if (fc) {
                const int i1 = c1 - gh(p, cg, t1), i2 = c2 - gh(p, cg, t2);
                /* (ORTHO1, ORTHO2) is (t1, t2) for x and z sweeps, (t2, t1) for y */
                if (ddim == 1) fc_pencil(p, cg, ddim, fc, i2, i1, &e);
                else fc_pencil(p, cg, ddim, fc, i1, i2, &e);
                pe = &e;
            }
            or_pencil_stage_ext(p, n, dtodx, chspeed, u0, u, b0, b, u1, b1, nbp, dvol, rep, pe);
            for (int i = 0; i < n; i++) {
                size_t cell = base + (size_t)i * sd[ddim];
                for (int f = 0; f < p->nfluids; f++) {
                    int o = p->fl[f].beg;
                    for (int k = 0; k < 5; k++)
                        cg->u[cell * nv + o + swp_var(ddim, k)] = u1[(size_t)(o + k) * n + i];
                }
                if (mhd) {
                    for (int k = 0; k < 3; k++) cg->b[cell * 3 + swp_mag(ddim, k)] = b1[(size_t)k * n + i];
                    cg->psi[cell] = b1[(size_t)3 * n + i];
                }
            }
        }
    }
    free(u0); free(u); free(u1); free(b0); free(b); free(b1);
}

/* Copy every ghost cell from its periodic image (all_fluid_boundaries +
 * all_mag_boundaries on one periodic block, corners included). */
void or_fill_ghosts_periodic(const or_params *p, or_block *cg)
{
    int ng[3], g[3];
    gdims(p, cg, ng);
    for (int d = 0; d < 3; d++) g[d] = gh(p, cg, d);
    int nv = p->nvar, mhd = any_mag(p);
    for (int k = 0; k < ng[2]; k++)
        for (int j = 0; j < ng[1]; j++)
            for (int i = 0; i < ng[0]; i++) {
                int idx[3] = {i, j, k}, src[3], ghost = 0;
                for (int d = 0; d < 3; d++) {
                    int r = idx[d] - g[d];
                    if (r < 0 || r >= cg->n[d]) ghost = 1;
                    r = ((r % cg->n[d]) + cg->n[d]) % cg->n[d];
                    src[d] = r + g[d];
                }
                if (!ghost) continue;
                size_t dc = (size_t)i + (size_t)ng[0] * ((size_t)j + (size_t)ng[1] * k);
                size_t sc = (size_t)src[0] + (size_t)ng[0] * ((size_t)src[1] + (size_t)ng[1] * src[2]);
                for (int v = 0; v < nv; v++) cg->u[dc * nv + v] = cg->u[sc * nv + v];
                if (mhd) {
                    for (int c = 0; c < 3; c++) cg->b[dc * 3 + c] = cg->b[sc * 3 + c];
                    cg->psi[dc] = cg->psi[sc];
                }
            }
}

static double min_dl(const or_params *p, const or_block *cg)
{
    double m = DBL_MAX;
    for (int d = 0; d < 3; d++)
        if (p->has_dir[d] && cg->dl[d] < m) m = cg->dl[d];
    return m;
}

/* hdc.F90:112-122 with ch_grid = .true. */
double or_update_chspeed(const or_params *p, const or_block *cg, double dt)
{
    double ch = DBL_MAX;
    double c = (p->cfl_glm * min_dl(p, cg)) / dt;
    if (c < ch) ch = c;
    return ch;
}

/* hdc.F90:313-331 */
double or_glm_factor(const or_params *p, const or_block *cg, double chspeed, double dt)
{
    double fac = 0.0;
    double f = (p->glm_alpha * chspeed) / (min_dl(p, cg) / dt);
    if (f > fac) fac = f;
    return exp(-fac);
}

/* hdc.F90:338: psi (ghosts included) *= fac */
void or_glmdamping(const or_params *p, or_block *cg, double fac)
{
    int ng[3];
    gdims(p, cg, ng);
    size_t nc = (size_t)ng[0] * ng[1] * ng[2];
    for (size_t c = 0; c < nc; c++) cg->psi[c] = cg->psi[c] * fac;
}

/* timestep_fluid (timestep.F90:431-495) for fluid ifl with ion_cs / neu_cs */
void or_timestep_fluid(const or_params *p, const or_block *cg, int ifl, double *dt_out, double *c_out)
{
    int ng[3], g[3];
    gdims(p, cg, ng);
    for (int d = 0; d < 3; d++) g[d] = gh(p, cg, d);
    const or_fluid *fl = &p->fl[ifl];
    int nv = p->nvar, o = fl->beg;
    double gam = fl->gam, gam_1 = fl->gam - 1.0;
    double c_fl = small_c, dt_proc[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    for (int k = g[2]; k < g[2] + cg->n[2]; k++)
        for (int j = g[1]; j < g[1] + cg->n[1]; j++)
            for (int i = g[0]; i < g[0] + cg->n[0]; i++) {
                size_t cell = (size_t)i + (size_t)ng[0] * ((size_t)j + (size_t)ng[1] * k);
                const double *uc = cg->u + cell * nv + o;
                double v[3];
                if (uc[0] > 0.0) {
                    for (int d = 0; d < 3; d++) v[d] = fabs(uc[1 + d] / uc[0]);
                } else {
                    v[0] = v[1] = v[2] = 0.0;
                }
                double ek = (0.5 * ((uc[1] * uc[1] + uc[2] * uc[2]) + uc[3] * uc[3])) / uc[0];
                double cs;
                if (fl->magnetized) {
                    const double *bc = cg->b + cell * 3;
                    double pmag = 0.5 * ((bc[0] * bc[0] + bc[1] * bc[1]) + bc[2] * bc[2]);
                    double ps = (uc[4] - ek) * gam_1 + (2.0 - gam) * pmag;
                    double pr = ps - pmag;
                    cs = sqrt(fabs((2.0 * pmag + gam * pr) / uc[0]));
                } else {
                    double pr = (uc[4] - ek) * gam_1;
                    cs = sqrt(fabs((gam * pr) / uc[0]));
                }
                double c[3];
                for (int d = 0; d < 3; d++) {
                    c[d] = v[d] + cs;
                    if (small_c > c[d]) c[d] = small_c; /* max(v + cs, small) */
                }
                double cm = c[0];
                if (c[1] > cm) cm = c[1];
                if (c[2] > cm) cm = c[2];
                if (cm > c_fl) c_fl = cm;
                for (int d = 0; d < 3; d++) {
                    if (p->has_dir[d] && c[d] > 0.0) {
                        double t = cg->dl[d] / c[d];
                        if (t < dt_proc[d]) dt_proc[d] = t;
                    } else {
                        dt_proc[d] = DBL_MAX;
                    }
                }
            }
    double mn = dt_proc[0];
    if (dt_proc[1] < mn) mn = dt_proc[1];
    if (dt_proc[2] < mn) mn = dt_proc[2];
    *dt_out = p->cfl * mn;
    *c_out = c_fl;
}

/* time_step (timestep.F90:112-225) with dt_initial=-1, dt_max_grow=2,
 * dt_max=huge and tend never binding.  nstep counts completed steps. */
double or_time_step(const or_params *p, const or_block *cg, int nstep, double dt_old, double *c_all)
{
    double dt = DBL_MAX, c = 0.0;
    for (int f = 0; f < p->nfluids; f++) {
        double dt_, c_;
        or_timestep_fluid(p, cg, f, &dt_, &c_);
        if (dt_ < dt) dt = dt_;
        if (c_ > c) c = c_;
    }
    if (nstep < 1) {
        double t = 1.0 * dt; /* dt_initial = -1: min(dt, -dt_initial*dt) */
        if (t < dt) dt = t;
    } else if (dt_old > 0.0) {
        double t = dt_old * 2.0;
        if (t < dt) dt = t;
    }
    if (c_all) *c_all = c;
    return dt;
}

static void sweep(const or_params *p, or_block *cg, int ddim, double dt, double chspeed, or_report *rep)
{
    int s0 = first_stage(p), s1 = (p->integration_order == 1) ? EULER : RK2_2;
    for (int istep = s0; istep <= s1; istep++) {
        or_solve_cg(p, cg, ddim, istep, dt, chspeed, rep);
        or_fill_ghosts_periodic(p, cg);
    }
}

/* fluid_update_full (fluidupdate.F90:99-136) for one periodic block. */
void or_fluid_update(const or_params *p, or_block *cg, double dt, or_report *rep)
{
    int mhd = any_mag(p);
    double ch = mhd ? or_update_chspeed(p, cg, dt) : 0.0;
    for (int half = 0; half < 2; half++) {
        for (int s = 0; s < 3; s++) {
            int d = half == 0 ? s : 2 - s;
            if (p->has_dir[d]) sweep(p, cg, d, dt, ch, rep);
        }
        if (mhd) or_glmdamping(p, cg, or_glm_factor(p, cg, ch, dt));
    }
}

/* nsteps iterations of the main loop (piernik.F90:124-170): time_step then
 * fluid_update.  Returns the simulated time 2*sum(dt). */
double or_run(const or_params *p, or_block *cg, int nsteps, or_report *rep)
{
    double t = 0.0, dt = -1.0;
    or_fill_ghosts_periodic(p, cg);
    for (int s = 0; s < nsteps; s++) {
        double c;
        dt = or_time_step(p, cg, s, dt, &c);
        or_fluid_update(p, cg, dt, rep);
        t = t + dt;
        t = t + dt;
    }
    return t;
}

/* ---- the block list ---------------------------------------------------- */

/* fluid_update_full over a list of blocks, each periodic on itself: one
 * update_chspeed over the list (hdc.F90:112-140, min over blocks), the sweep
 * loops of make_3sweeps/sweep over every block (sweeps.F90:217-272), and
 * glmdamping with one factor from the max over blocks (hdc.F90:313-331). */
void or_fluid_update_list(const or_params *p, or_block *const *cgs, int nblk, double dt, or_report *reps)
{
    int mhd = any_mag(p);
    double ch = DBL_MAX, fac = 1.0;
    if (mhd) {
        for (int i = 0; i < nblk; i++) {
            double c = or_update_chspeed(p, cgs[i], dt);
            if (c < ch) ch = c;
        }
        double f = 0.0;
        for (int i = 0; i < nblk; i++) {
            double r = (p->glm_alpha * ch) / (min_dl(p, cgs[i]) / dt);
            if (r > f) f = r;
        }
        fac = exp(-f);
    } else {
        ch = 0.0;
    }
    for (int half = 0; half < 2; half++) {
        for (int s = 0; s < 3; s++) {
            int d = half == 0 ? s : 2 - s;
            if (!p->has_dir[d]) continue;
            for (int i = 0; i < nblk; i++) sweep(p, cgs[i], d, dt, ch, &reps[i]);
        }
        if (mhd)
            for (int i = 0; i < nblk; i++) or_glmdamping(p, cgs[i], fac);
    }
}

/* nsteps main-loop iterations over the list: time_step takes the min over
 * blocks (timestep.F90:112-225), then fluid_update over the list. */
double or_run_list(const or_params *p, or_block *const *cgs, int nblk, int nsteps, or_report *reps)
{
    return or_run_list_nb(p, cgs, NULL, nblk, nsteps, reps);
}

/* Ghost exchange between the blocks of a list (internal_boundaries_local,
 * cg_list_bnd.F90:169-322, as all_fluid_boundaries and all_mag_boundaries
 * call it, all_boundaries.F90:82-175).  nbr[6*i + 2*d + s] is the index of
 * the block whose interior adjoins block i's ghost layers on side s (0 low,
 * 1 high) of direction d; a block periodic on itself is its own neighbour.
 * The ghost layers of direction d are copied from the neighbour's nb
 * adjoining interior layers over the whole extent of the other two
 * directions, ghost cells included, x then y then z, so edges and corners
 * come out as the corner-complete refresh gives them.  Blocks of one list
 * share nb and their extents across each face. */
void or_fill_ghosts_list(const or_params *p, or_block *const *cgs, const int32_t *nbr, int nblk)
{
    int nv = p->nvar, mhd = any_mag(p);
    for (int d = 0; d < 3; d++) {
        if (!p->has_dir[d]) continue;
        for (int i = 0; i < nblk; i++) {
            or_block *cg = cgs[i];
            int ng[3];
            gdims(p, cg, ng);
            int g = gh(p, cg, d);
            for (int side = 0; side < 2; side++) {
                const or_block *o = cgs[nbr[6 * i + 2 * d + side]];
                int ngo[3];
                gdims(p, o, ngo);
                /* own ghost layers and the neighbour's adjoining interior layers */
                int dst0 = side == 0 ? 0 : g + cg->n[d];
                int src0 = side == 0 ? o->n[d] : g;
                for (int k = 0; k < ng[2]; k++)
                    for (int j = 0; j < ng[1]; j++)
                        for (int ii = 0; ii < ng[0]; ii++) {
                            int c[3] = {ii, j, k};
                            if (c[d] >= g) continue;
                            int dc3[3] = {ii, j, k}, sc3[3] = {ii, j, k};
                            dc3[d] = dst0 + c[d];
                            sc3[d] = src0 + c[d];
                            size_t dc = (size_t)dc3[0] + (size_t)ng[0] * ((size_t)dc3[1] + (size_t)ng[1] * dc3[2]);
                            size_t sc = (size_t)sc3[0] + (size_t)ngo[0] * ((size_t)sc3[1] + (size_t)ngo[1] * sc3[2]);
                            for (int v = 0; v < nv; v++) cg->u[dc * nv + v] = o->u[sc * nv + v];
                            if (mhd) {
                                for (int q = 0; q < 3; q++) cg->b[dc * 3 + q] = o->b[sc * 3 + q];
                                cg->psi[dc] = o->psi[sc];
                            }
                        }
            }
        }
    }
}

static void fill_list(const or_params *p, or_block *const *cgs, const int32_t *nbr, int nblk)
{
    if (nbr)
        or_fill_ghosts_list(p, cgs, nbr, nblk);
    else
        for (int i = 0; i < nblk; i++) or_fill_ghosts_periodic(p, cgs[i]);
}

/* fluid_update_full over a list of blocks connected by nbr (as in
 * or_fill_ghosts_list): sweep() runs each stage on every block, then
 * refreshes the ghosts of the whole list (sweeps.F90:217-272, :66-107). */
void or_fluid_update_list_nb(const or_params *p, or_block *const *cgs, const int32_t *nbr, int nblk, double dt,
                             or_report *reps)
{
    int mhd = any_mag(p);
    double ch = DBL_MAX, fac = 1.0;
    if (mhd) {
        for (int i = 0; i < nblk; i++) {
            double c = or_update_chspeed(p, cgs[i], dt);
            if (c < ch) ch = c;
        }
        double f = 0.0;
        for (int i = 0; i < nblk; i++) {
            double r = (p->glm_alpha * ch) / (min_dl(p, cgs[i]) / dt);
            if (r > f) f = r;
        }
        fac = exp(-f);
    } else {
        ch = 0.0;
    }
    int s0 = first_stage(p), s1 = (p->integration_order == 1) ? EULER : RK2_2;
    for (int half = 0; half < 2; half++) {
        for (int s = 0; s < 3; s++) {
            int d = half == 0 ? s : 2 - s;
            if (!p->has_dir[d]) continue;
            for (int istep = s0; istep <= s1; istep++) {
                for (int i = 0; i < nblk; i++) or_solve_cg(p, cgs[i], d, istep, dt, ch, &reps[i]);
                fill_list(p, cgs, nbr, nblk);
            }
        }
        if (mhd)
            for (int i = 0; i < nblk; i++) or_glmdamping(p, cgs[i], fac);
    }
}

/* or_run_list over blocks connected by nbr (NULL: each periodic on itself). */
double or_run_list_nb(const or_params *p, or_block *const *cgs, const int32_t *nbr, int nblk, int nsteps,
                      or_report *reps)
{
    double t = 0.0, dt = -1.0;
    fill_list(p, cgs, nbr, nblk);
    for (int s = 0; s < nsteps; s++) {
        double dtn = DBL_MAX;
        for (int i = 0; i < nblk; i++) {
            double c;
            double d = or_time_step(p, cgs[i], s, dt, &c);
            if (d < dtn) dtn = d;
        }
        dt = dtn;
        if (nbr)
            or_fluid_update_list_nb(p, cgs, nbr, nblk, dt, reps);
        else
            or_fluid_update_list(p, cgs, nblk, dt, reps);
        t = t + dt;
        t = t + dt;
    }
    return t;
}
