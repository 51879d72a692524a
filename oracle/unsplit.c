/*
 * unsplit.c -- the unsplit Riemann driver of the oracle (TEST INFRASTRUCTURE
 * ONLY).
 *
 * Restates, for one periodic block:
 *   fluid_update_unsplit   src/scheme/unsplit_schemes/unsplit_fluidupdate.F90:45-126
 *   unsplit_sweep          src/scheme/unsplit_schemes/unsplit_sweeps.F90:84-195
 *                          (update_boundaries :47-82)
 *   solve_cg_ub / solve    src/scheme/unsplit_schemes/unsplit_mag_modules.F90:40-209
 *   solve_cg_u / solve_u   src/scheme/unsplit_schemes/solve_cg_unsplit.F90:79-187
 *   apply_flux, update_psi, bounds_for_flux
 *                          unsplit_mag_modules.F90:211-358, solve_cg_unsplit.F90:189-274
 *   apply_source           src/scheme/unsplit_schemes/unsplit_source.F90:42-132
 *   glmdamping(half)       src/scheme/Riemann_Solver/hdc.F90:289-349
 *   time_step, unsplit clamp to tend
 *                          src/base/timestep.F90:112-240 (:221-222)
 * with cflcontrol never redoing a step, no sources other than the vanishing
 * internal ones and use_eglm off (hdc.F90:380).
 *
 * One RK stage computes the interface fluxes of every direction from one
 * state (u at the first stage, uh at the second), then adds each direction's
 * flux difference in turn to uh (a copy of u) or u, and to B and psi; the
 * floors run at the last stage only, once per direction over that
 * direction's pencils, in its sweep order.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define EULER 10000
#define RK2_1 10001
#define RK2_2 10002

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
static int last_stage(const or_params *p) { return p->integration_order == 1 ? EULER : RK2_2; }

typedef struct {
    double *f[3];  /* (nvar, ng) per direction: row c holds the flux at c-1/2 */
    double *bf[3]; /* (3, ng) */
    double *pf[3]; /* (ng) */
} fluxes_t;

/* The interface fluxes of every active direction from the state (u, b, psi)
 * over the pencils of the transverse interior, full length (the flux
 * arrays' first row is 0, as tflux(:, 1) = 0). */
static void block_fluxes(const or_params *p, const or_block *cg, const double *u, const double *b, const double *psi,
                         double ch, fluxes_t *F)
{
    int ng[3];
    gdims(p, cg, ng);
    const size_t ncell = (size_t)ng[0] * ng[1] * ng[2];
    const int nv = p->nvar, mhd = any_mag(p);
    const size_t sd[3] = {1, (size_t)ng[0], (size_t)ng[0] * ng[1]};
    for (int d = 0; d < 3; d++) {
        if (!p->has_dir[d]) continue;
        memset(F->f[d], 0, sizeof(double) * ncell * nv);
        if (mhd) {
            memset(F->bf[d], 0, sizeof(double) * ncell * 3);
            memset(F->pf[d], 0, sizeof(double) * ncell);
        }
        const int n = ng[d], m = n - 1;
        const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
        double *pu = (double *)malloc(sizeof(double) * (size_t)n * nv);
        double *pb = (double *)malloc(sizeof(double) * (size_t)n * 4);
        double *flx = (double *)malloc(sizeof(double) * (size_t)m * nv);
        double *bflx = (double *)malloc(sizeof(double) * (size_t)m * 4);
        for (int c2 = gh(p, cg, t2); c2 < gh(p, cg, t2) + cg->n[t2]; c2++) {
            for (int c1 = gh(p, cg, t1); c1 < gh(p, cg, t1) + cg->n[t1]; c1++) {
                const size_t base = (size_t)c1 * sd[t1] + (size_t)c2 * sd[t2];
                for (int i = 0; i < n; i++) {
                    const size_t cell = base + (size_t)i * sd[d];
                    for (int f = 0; f < p->nfluids; f++) {
                        const int o = p->fl[f].beg;
                        for (int k = 0; k < 5; k++) pu[(size_t)(o + k) * n + i] = u[cell * nv + o + swp_var(d, k)];
                    }
                    if (mhd) {
                        for (int k = 0; k < 3; k++) pb[(size_t)k * n + i] = b[cell * 3 + swp_mag(d, k)];
                        pb[(size_t)3 * n + i] = psi[cell];
                    }
                }
                or_pencil_flux(p, n, ch, pu, mhd ? pb : NULL, flx, mhd ? bflx : NULL);
                for (int i = 1; i < n; i++) {
                    const size_t cell = base + (size_t)i * sd[d];
                    for (int f = 0; f < p->nfluids; f++) {
                        const int o = p->fl[f].beg;
                        for (int k = 0; k < 5; k++)
                            F->f[d][cell * nv + o + swp_var(d, k)] = flx[(size_t)(o + k) * m + i - 1];
                    }
                    if (mhd) {
                        for (int k = 0; k < 3; k++) F->bf[d][cell * 3 + swp_mag(d, k)] = bflx[(size_t)k * m + i - 1];
                        F->pf[d][cell] = bflx[(size_t)3 * m + i - 1];
                    }
                }
            }
        }
        free(pu); free(pb); free(flx); free(bflx);
    }
}

/* T += dt/dl(d) * rk_coef * (F(c) - F(c + e_d)) for every active direction
 * d in turn, over bounds_for_flux: along d all cells but the outermost
 * layer, across d the interior (unsplit_mag_modules.F90:266-276, 330-358). */
static void apply_flux(const or_params *p, const or_block *cg, double *T, int nc, double *const *F, double dt,
                       int istep)
{
    int ng[3];
    gdims(p, cg, ng);
    const size_t sd[3] = {1, (size_t)ng[0], (size_t)ng[0] * ng[1]};
    for (int a = 0; a < 3; a++) {
        if (!p->has_dir[a]) continue;
        int lo[3], hi[3];
        for (int d = 0; d < 3; d++) {
            lo[d] = 0;
            hi[d] = ng[d] - 1;
            if (p->has_dir[d]) {
                lo[d] += 1;
                hi[d] -= 1;
                if (d != a) {
                    lo[d] += cg->nb - 1;
                    hi[d] -= cg->nb - 1;
                }
            }
        }
        const double coef = (dt / cg->dl[a]) * rk_coef(istep);
        for (int k = lo[2]; k <= hi[2]; k++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int i = lo[0]; i <= hi[0]; i++) {
                    const size_t c = (size_t)i + sd[1] * j + sd[2] * k;
                    const size_t cn = c + sd[a];
                    for (int v = 0; v < nc; v++)
                        T[c * nc + v] = T[c * nc + v] + coef * (F[a][c * nc + v] - F[a][cn * nc + v]);
                }
    }
}

/* apply_source (unsplit_source.F90:42-132): for every active direction, over
 * its transverse-interior pencils, u1 = u + 0 (the vanishing internal
 * sources) and, at the last stage, care_for_positives in that direction's
 * sweep order; u and b are uh and bh at the first stage. */
static void apply_source(const or_params *p, or_block *cg, int istep, double *U, const double *B, or_report *rep)
{
    int ng[3];
    gdims(p, cg, ng);
    const int nv = p->nvar, mhd = any_mag(p);
    const size_t sd[3] = {1, (size_t)ng[0], (size_t)ng[0] * ng[1]};
    double dvol = 1.0;
    for (int d = 0; d < 3; d++)
        if (p->has_dir[d]) dvol *= cg->dl[d];
    for (int d = 0; d < 3; d++) {
        if (!p->has_dir[d]) continue;
        const int n = ng[d];
        const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
        double *u1 = (double *)malloc(sizeof(double) * (size_t)n * nv);
        double *b1 = (double *)malloc(sizeof(double) * (size_t)n * 4);
        for (int c2 = gh(p, cg, t2); c2 < gh(p, cg, t2) + cg->n[t2]; c2++) {
            for (int c1 = gh(p, cg, t1); c1 < gh(p, cg, t1) + cg->n[t1]; c1++) {
                const size_t base = (size_t)c1 * sd[t1] + (size_t)c2 * sd[t2];
                for (int i = 0; i < n; i++) {
                    const size_t cell = base + (size_t)i * sd[d];
                    for (int f = 0; f < p->nfluids; f++) {
                        const int o = p->fl[f].beg;
                        for (int k = 0; k < 5; k++) u1[(size_t)(o + k) * n + i] = U[cell * nv + o + swp_var(d, k)];
                    }
                    if (mhd)
                        for (int k = 0; k < 3; k++) b1[(size_t)k * n + i] = B[cell * 3 + swp_mag(d, k)];
                }
                for (size_t q = 0; q < (size_t)n * nv; q++) u1[q] = u1[q] + 0.0;
                if (istep == last_stage(p)) or_floors(p, n, u1, mhd ? b1 : NULL, gh(p, cg, d), dvol, rep);
                for (int i = 0; i < n; i++) {
                    const size_t cell = base + (size_t)i * sd[d];
                    for (int f = 0; f < p->nfluids; f++) {
                        const int o = p->fl[f].beg;
                        for (int k = 0; k < 5; k++) U[cell * nv + o + swp_var(d, k)] = u1[(size_t)(o + k) * n + i];
                    }
                }
            }
        }
        free(u1); free(b1);
    }
}

/* One RK stage of unsplit_sweep for one periodic block (solve_cg_ub or
 * solve_cg_u, then update_boundaries of the stage's arrays). */
void or_unsplit_stage(const or_params *p, or_block *cg, int istep, double dt, double chspeed, or_report *rep)
{
    int ng[3];
    gdims(p, cg, ng);
    const size_t ncell = (size_t)ng[0] * ng[1] * ng[2];
    const int nv = p->nvar, mhd = any_mag(p);
    const int first = istep == first_stage(p) && p->integration_order >= 2;
    const int last = istep == last_stage(p);
    fluxes_t F;
    for (int d = 0; d < 3; d++) {
        F.f[d] = (double *)malloc(sizeof(double) * ncell * nv);
        F.bf[d] = mhd ? (double *)malloc(sizeof(double) * ncell * 3) : NULL;
        F.pf[d] = mhd ? (double *)malloc(sizeof(double) * ncell) : NULL;
    }
    /* fluxes from u at the first stage, from uh at the second */
    if (first || p->integration_order < 2)
        block_fluxes(p, cg, cg->u, cg->b, cg->psi, chspeed, &F);
    else
        block_fluxes(p, cg, cg->uh, cg->bh, cg->psih, chspeed, &F);
    /* apply_flux(mag), apply_flux(fluid), update_psi: into copies at the first stage */
    double *Tu = cg->u, *Tb = cg->b, *Tp = cg->psi;
    if (!last) {
        memcpy(cg->uh, cg->u, sizeof(double) * ncell * nv);
        Tu = cg->uh;
        if (mhd) {
            memcpy(cg->bh, cg->b, sizeof(double) * ncell * 3);
            memcpy(cg->psih, cg->psi, sizeof(double) * ncell);
            Tb = cg->bh;
            Tp = cg->psih;
        }
    }
    if (mhd) apply_flux(p, cg, Tb, 3, F.bf, dt, istep);
    apply_flux(p, cg, Tu, nv, F.f, dt, istep);
    if (mhd) apply_flux(p, cg, Tp, 1, F.pf, dt, istep);
    apply_source(p, cg, istep, Tu, Tb, rep);
    /* update_boundaries(istep): the stage's arrays */
    or_block view = *cg;
    view.u = Tu;
    view.b = Tb;
    view.psi = Tp;
    or_fill_ghosts_periodic(p, &view);
    for (int d = 0; d < 3; d++) {
        free(F.f[d]);
        free(F.bf[d]);
        free(F.pf[d]);
    }
}

/* fluid_update_unsplit: update_chspeed, glmdamping(half), the stages,
 * glmdamping(half). */
void or_fluid_update_unsplit(const or_params *p, or_block *cg, double dt, or_report *rep)
{
    const int mhd = any_mag(p);
    double ch = 0.0, fac = 1.0;
    if (mhd) {
        ch = or_update_chspeed(p, cg, dt);
        fac = or_glm_factor(p, cg, ch, dt / 2.0); /* dt_eff = dt/2 (hdc.F90:315-318) */
        or_glmdamping(p, cg, fac);
    }
    for (int istep = first_stage(p); istep <= last_stage(p); istep++) or_unsplit_stage(p, cg, istep, dt, ch, rep);
    if (mhd) or_glmdamping(p, cg, fac);
}

/* The main loop (piernik.F90:125) with the unsplit solver until t >= tend
 * or nmax steps: time_step (caps of timestep.F90:198-207, then the unsplit
 * clamp dt <= (tend - t) + 2*epsilon*(tend - t), :221-222), one
 * fluid_update_unsplit, t += dt.  Returns t. */
double or_run_unsplit(const or_params *p, or_block *cg, double tend, int nmax, int32_t *nsteps, or_report *rep)
{
    double t = 0.0, dt = -1.0;
    int s = 0;
    or_fill_ghosts_periodic(p, cg);
    while (t < tend && s < nmax) {
        double c;
        dt = or_time_step(p, cg, s, dt, &c);
        const double rem = tend - t;
        const double cl = rem + (2.0 * DBL_EPSILON) * rem;
        if (cl < dt) dt = cl;
        or_fluid_update_unsplit(p, cg, dt, rep);
        t = t + dt;
        s++;
    }
    if (nsteps) *nsteps = s;
    return t;
}
