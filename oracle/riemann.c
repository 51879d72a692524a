/*
 * riemann.c -- pencil-level part of the oracle (TEST INFRASTRUCTURE ONLY).
 *
 * A scalar C restatement of PIERNIK's Riemann-solver pencil operator:
 *   utoq            src/scheme/Riemann_Solver/interpolations.F90:69-106
 *   limiters        src/scheme/Riemann_Solver/fluxlimiters.F90:109-216
 *   linear (PLM)    src/scheme/Riemann_Solver/interpolations.F90:170-204
 *   HLLC            src/scheme/Riemann_Solver/hlld.F90:615-841 (riemann_hlld_u)
 *   HLLD + GLM      src/scheme/Riemann_Solver/hlld.F90:161-609 (riemann_hlld)
 *   update          src/scheme/Riemann_Solver/solve_cg_riemann.F90:318-414
 *   sources, floors src/fluids/sources.F90:151-244, 299-433
 * Evaluation order follows the Fortran expressions term by term: Fortran
 * evaluates a+b+c as (a+b)+c, a*b/c as (a*b)/c, x**2 as x*x and SUM() as a
 * left-to-right accumulation.  Build with -ffp-contract=off.
 *
 * Pencil arrays are column-major (n, nv): element (i, v) is a[v*n + i],
 * i = 0..n-1 along the sweep.  Variables are in sweep order
 * [rho, m_n, m_t1, m_t2, E] per fluid and [B_n, B_t1, B_t2, psi].
 */
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define IDN 0
#define IMX 1
#define IMY 2
#define IMZ 3
#define IEN 4

/* Fortran SPACING(x) for real(8): 2**max(exponent(x)-digits(x), minexponent(x)-1),
 * and TINY(x) for x == 0 (F2008 13.7.157). */
double or_spacing(double x)
{
    if (x == 0.0) return DBL_MIN;
    if (isnan(x) || isinf(x)) return NAN;
    int e;
    (void)frexp(x, &e);
    int k = e - 53;
    if (k < -1022) k = -1022;
    return ldexp(1.0, k);
}

/* Fortran SIGN(1., x): copysign semantics (signed zero honoured). */
static inline double fsign1(double x) { return copysign(1.0, x); }
static inline double fmin2(double a, double b) { return (b < a) ? b : a; }
static inline double fmax2(double a, double b) { return (b > a) ? b : a; }

/* fluxlimiters.F90:109-216: one column q(0..n-1) -> dq(0..n-1). */
void or_limiter(int kind, int n, const double *u, double *dq)
{
    if (n < 2) { for (int i = 0; i < n; i++) dq[i] = 0.0; return; }
    if (kind == OR_VANLEER) {
        /* calculate_slope_vanleer, fluxlimiters.F90:109-136 */
        dq[0] = 0.0;
        dq[n - 1] = 0.0;
        double dlft = u[1] - u[0];
        for (int i = 1; i < n - 1; i++) {
            double drgt = u[i + 1] - u[i];
            double dcen = dlft * drgt;
            if (dcen > 0.0)
                dq[i] = (2.0 * dcen) / (dlft + drgt);
            else
                dq[i] = 0.0;
            dlft = drgt;
        }
        return;
    }
    /* minmod / moncen / superbee share the one-sided difference arrays,
     * fluxlimiters.F90:157-158 (and :184-185, :209-210) */
    double dl[n], dr[n];
    for (int i = 1; i < n; i++) dl[i] = u[i] - u[i - 1];
    dl[0] = dl[1];
    for (int i = 0; i < n - 1; i++) dr[i] = dl[i + 1];
    dr[n - 1] = dr[n - 2];
    for (int i = 0; i < n; i++) {
        double s = fsign1(dl[i]) + fsign1(dr[i]);
        double a = fabs(dl[i]), b = fabs(dr[i]);
        switch (kind) {
        case OR_MINMOD: /* :159 */
            dq[i] = (s * fmin2(a, b)) * 0.5;
            break;
        case OR_MONCEN: { /* :186 */
            double m = fmin2(fmin2(2.0 * a, 2.0 * b), 0.5 * fabs(dl[i] + dr[i]));
            dq[i] = (s * m) * 0.5;
            break;
        }
        case OR_SUPERBEE: { /* :212-213 */
            double m = (a > b) ? fmin2(a, fabs(2.0 * dr[i])) : fmin2(fabs(2.0 * dl[i]), b);
            dq[i] = (s * 0.5) * m;
            break;
        }
        default:
            dq[i] = 0.0;
        }
    }
}

/* interpolations.F90:170-204: ql(i) = q(i) + dq(i)/2, qr(i) = q(i+1) - dq(i+1)/2
 * for interfaces i = 0..n-2 (between cells i and i+1). */
void or_linear(int kind, int n, int nv, const double *q, double *ql, double *qr)
{
    int m = n - 1;
    double dq[n];
    for (int v = 0; v < nv; v++) {
        const double *qc = q + (size_t)v * n;
        or_limiter(kind, n, qc, dq);
        for (int i = 0; i < n; i++) dq[i] = 0.5 * dq[i];
        for (int i = 0; i < m; i++) {
            ql[(size_t)v * m + i] = qc[i] + dq[i];
            qr[(size_t)v * m + i] = qc[i + 1] - dq[i + 1];
        }
    }
}

/* interpolations.F90:69-106 with func.F90:55-88 (ekin = emag(m)/rho,
 * emag = half*sq_sum3).  b (n, >=3) is read only for a magnetized fluid. */
void or_utoq(const or_params *p, int n, const double *u, const double *b, double *q)
{
    for (int f = 0; f < p->nfluids; f++) {
        const or_fluid *fl = &p->fl[f];
        int o = fl->beg;
        double gam_1 = fl->gam - 1.0;
        const double *dn = u + (size_t)(o + IDN) * n, *mx = u + (size_t)(o + IMX) * n,
                     *my = u + (size_t)(o + IMY) * n, *mz = u + (size_t)(o + IMZ) * n,
                     *en = u + (size_t)(o + IEN) * n;
        for (int i = 0; i < n; i++) {
            q[(size_t)(o + IDN) * n + i] = dn[i];
            q[(size_t)(o + IMX) * n + i] = mx[i] / dn[i];
            q[(size_t)(o + IMY) * n + i] = my[i] / dn[i];
            q[(size_t)(o + IMZ) * n + i] = mz[i] / dn[i];
            double ek = (0.5 * ((mx[i] * mx[i] + my[i] * my[i]) + mz[i] * mz[i])) / dn[i];
            double pr = gam_1 * (en[i] - ek);
            if (fl->magnetized) {
                double b2 = (b[i] * b[i] + b[n + i] * b[n + i]) + b[2 * n + i] * b[2 * n + i];
                pr = pr - (0.5 * gam_1) * b2;
            }
            q[(size_t)(o + IEN) * n + i] = pr;
        }
    }
}

/* riemann_hlld_u, hlld.F90:615-841 (adiabatic, has_energy).  m interfaces. */
void or_hllc(int m, const double *ql, const double *qr, double gam, double *f)
{
#define UL(v) ql[(size_t)(v) * m + i]
#define UR(v) qr[(size_t)(v) * m + i]
#define F(v) f[(size_t)(v) * m + i]
    for (int i = 0; i < m; i++) {
        double prl = UL(IEN), prr = UR(IEN);
        double enl = (UL(IEN) / (gam - 1.0)) +
                     (0.5 * UL(IDN)) * ((UL(IMX) * UL(IMX) + UL(IMY) * UL(IMY)) + UL(IMZ) * UL(IMZ));
        double enr = (UR(IEN) / (gam - 1.0)) +
                     (0.5 * UR(IDN)) * ((UR(IMX) * UR(IMX) + UR(IMY) * UR(IMY)) + UR(IMZ) * UR(IMZ));
        double gampr_l = gam * UL(IEN), gampr_r = gam * UR(IEN);
        double c_fastm = sqrt(fmax2(gampr_l / UL(IDN), gampr_r / UR(IDN)));
        double sl = fmin2(UL(IMX), UR(IMX)) - c_fastm;
        double sr = fmax2(UL(IMX), UR(IMX)) + c_fastm;

        double fl[5], fr[5];
        fl[IDN] = UL(IDN) * UL(IMX);
        fl[IMX] = UL(IDN) * (UL(IMX) * UL(IMX)) + prl;
        fl[IMY] = (UL(IDN) * UL(IMY)) * UL(IMX);
        fl[IMZ] = (UL(IDN) * UL(IMZ)) * UL(IMX);
        fl[IEN] = (enl + prl) * UL(IMX);
        fr[IDN] = UR(IDN) * UR(IMX);
        fr[IMX] = UR(IDN) * (UR(IMX) * UR(IMX)) + prr;
        fr[IMY] = (UR(IDN) * UR(IMY)) * UR(IMX);
        fr[IMZ] = (UR(IDN) * UR(IMZ)) * UR(IMX);
        fr[IEN] = (enr + prr) * UR(IMX);

        if (sl >= 0.0) {
            for (int v = 0; v < 5; v++) F(v) = fl[v];
        } else if (sr <= 0.0) {
            for (int v = 0; v < 5; v++) F(v) = fr[v];
        } else {
            double sm;
            double auxr = (sr - UR(IMX)) * UR(IDN), auxl = (sl - UL(IMX)) * UL(IDN);
            if (fabs(auxr - auxl) < or_spacing(auxr + auxl))
                sm = (sl + sr) / 2.0;
            else
                sm = ((auxr * UR(IMX) - prr) - (auxl * UL(IMX) - prl)) / (auxr - auxl);
            double slvxl = sl - UL(IMX), srvxr = sr - UR(IMX);
            double dn_l = UL(IDN) * slvxl, dn_r = UR(IDN) * srvxr;
            double prt_star = 0.5 * ((prl + dn_l * (sm - UL(IMX))) + (prr + dn_r * (sm - UR(IMX))));
            double usl[5] = {0}, usr[5] = {0};
            double slsm = 0.0, srsm = 0.0;
            if (sm >= 0.0) {
                slsm = sl - sm;
                usl[IDN] = dn_l / slsm;
                usl[IMX] = usl[IDN] * sm;
                usl[IMY] = usl[IDN] * UL(IMY);
                usl[IMZ] = usl[IDN] * UL(IMZ);
                usl[IEN] = (((slvxl * enl) - (prl * UL(IMX))) + (prt_star * sm)) / slsm;
            }
            if (sm <= 0.0) {
                srsm = sr - sm;
                usr[IDN] = dn_r / srsm;
                usr[IMX] = usr[IDN] * sm;
                usr[IMY] = usr[IDN] * UR(IMY);
                usr[IMZ] = usr[IDN] * UR(IMZ);
                usr[IEN] = (((srvxr * enr) - (prr * UR(IMX))) + (prt_star * sm)) / srsm;
            }
            double Ul[5] = {UL(IDN), UL(IDN) * UL(IMX), UL(IDN) * UL(IMY), UL(IDN) * UL(IMZ), enl};
            double Ur[5] = {UR(IDN), UR(IDN) * UR(IMX), UR(IDN) * UR(IMY), UR(IDN) * UR(IMZ), enr};
            if (sm > 0.0) {
                for (int v = 0; v < 5; v++) F(v) = fl[v] + sl * (usl[v] - Ul[v]);
            } else if (sm < 0.0) {
                for (int v = 0; v < 5; v++) F(v) = fr[v] + sr * (usr[v] - Ur[v]);
            } else {
                for (int v = 0; v < 5; v++)
                    F(v) = 0.5 * (((fl[v] + sl * (usl[v] - Ul[v])) + fr[v]) + sr * (usr[v] - Ur[v]));
            }
            (void)slsm; (void)srsm;
        }
    }
#undef UL
#undef UR
#undef F
}

/* riemann_hlld with DIVB_HDC, hlld.F90:161-609.  m interfaces.
 * bl, br, bf are (m, 4): [B_n, B_t1, B_t2, psi]. */
void or_hlld(int m, const double *ql, const double *qr, const double *bl,
             const double *br, double gam, double chspeed, double *f, double *bf)
{
#define UL(v) ql[(size_t)(v) * m + i]
#define UR(v) qr[(size_t)(v) * m + i]
#define BL(v) bl[(size_t)(v) * m + i]
#define BR(v) br[(size_t)(v) * m + i]
#define F(v) f[(size_t)(v) * m + i]
#define BF(v) bf[(size_t)(v) * m + i]
    /* GLM fluxes, Dedner et al. eq. 42 (hlld.F90:228-230), whole vector first */
    for (int i = 0; i < m; i++) {
        BF(3) = ((chspeed * chspeed) * 0.5) * ((BR(0) + BL(0)) - (BR(3) - BL(3)) / chspeed);
        BF(0) = 0.5 * ((BR(3) + BL(3)) - chspeed * (BR(0) - BL(0)));
    }
    for (int i = 0; i < m; i++) {
        double magprl = 0.5 * ((BL(0) * BL(0) + BL(1) * BL(1)) + BL(2) * BL(2));
        double magprr = 0.5 * ((BR(0) * BR(0) + BR(1) * BR(1)) + BR(2) * BR(2));
        double prl = UL(IEN) + magprl;
        double prr = UR(IEN) + magprr;
        double b2l = (BL(0) * BL(0) + BL(1) * BL(1)) + BL(2) * BL(2);
        double b2r = (BR(0) * BR(0) + BR(1) * BR(1)) + BR(2) * BR(2);
        double enl = ((UL(IEN) / (gam - 1.0)) +
                      (0.5 * UL(IDN)) * ((UL(IMX) * UL(IMX) + UL(IMY) * UL(IMY)) + UL(IMZ) * UL(IMZ))) +
                     0.5 * b2l;
        double enr = ((UR(IEN) / (gam - 1.0)) +
                      (0.5 * UR(IDN)) * ((UR(IMX) * UR(IMX) + UR(IMY) * UR(IMY)) + UR(IMZ) * UR(IMZ))) +
                     0.5 * b2r;
        double gampr_l = gam * UL(IEN), gampr_r = gam * UR(IEN);
        double tl = gampr_l + b2l, tr = gampr_r + b2r;
        double cfl2 = (tl + sqrt(tl * tl - (4.0 * gampr_l) * (BL(0) * BL(0)))) / UL(IDN);
        double cfr2 = (tr + sqrt(tr * tr - (4.0 * gampr_r) * (BR(0) * BR(0)))) / UR(IDN);
        double c_fastm = sqrt(0.5 * fmax2(cfl2, cfr2));

        double sl = fmin2(UL(IMX), UR(IMX)) - c_fastm;
        double sr = fmax2(UL(IMX), UR(IMX)) + c_fastm;

        double fl[5], fr[5], bcl[3], bcr[3];
        fl[IDN] = UL(IDN) * UL(IMX);
        fl[IMX] = (UL(IDN) * (UL(IMX) * UL(IMX)) + prl) - BL(0) * BL(0);
        fl[IMY] = (UL(IDN) * UL(IMY)) * UL(IMX) - BL(0) * BL(1);
        fl[IMZ] = (UL(IDN) * UL(IMZ)) * UL(IMX) - BL(0) * BL(2);
        fl[IEN] = (enl + prl) * UL(IMX) - BL(0) * ((UL(IMX) * BL(0) + UL(IMY) * BL(1)) + UL(IMZ) * BL(2));
        bcl[1] = BL(1) * UL(IMX) - BL(0) * UL(IMY);
        bcl[2] = BL(2) * UL(IMX) - BL(0) * UL(IMZ);

        fr[IDN] = UR(IDN) * UR(IMX);
        fr[IMX] = (UR(IDN) * (UR(IMX) * UR(IMX)) + prr) - BR(0) * BR(0);
        fr[IMY] = (UR(IDN) * UR(IMY)) * UR(IMX) - BR(0) * BR(1);
        fr[IMZ] = (UR(IDN) * UR(IMZ)) * UR(IMX) - BR(0) * BR(2);
        fr[IEN] = (enr + prr) * UR(IMX) - BR(0) * ((UR(IMX) * BR(0) + UR(IMY) * BR(1)) + UR(IMZ) * BR(2));
        bcr[1] = BR(1) * UR(IMX) - BR(0) * UR(IMY);
        bcr[2] = BR(2) * UR(IMX) - BR(0) * UR(IMZ);

        if (sl >= 0.0) {
            for (int v = 0; v < 5; v++) F(v) = fl[v];
            BF(1) = bcl[1];
            BF(2) = bcl[2];
            continue;
        }
        if (sr <= 0.0) {
            for (int v = 0; v < 5; v++) F(v) = fr[v];
            BF(1) = bcr[1];
            BF(2) = bcr[2];
            continue;
        }
        double sm;
        double auxr = (sr - UR(IMX)) * UR(IDN), auxl = (sl - UL(IMX)) * UL(IDN);
        if (fabs(auxr - auxl) < or_spacing(auxr + auxl))
            sm = (sl + sr) / 2.0;
        else
            sm = ((auxr * UR(IMX) - prr) - (auxl * UL(IMX) - prl)) / (auxr - auxl);

        double slsm = sl - sm, srsm = sr - sm;
        double slvxl = sl - UL(IMX), srvxr = sr - UR(IMX);
        double srmsl = sr - sl;
        double dn_l = UL(IDN) * slvxl, dn_r = UR(IDN) * srvxr;
        double b_lr = BL(0) * BR(0);
        double b_lrgam = b_lr / gam;
        double prt_star = 0.5 * ((prl + dn_l * (sm - UL(IMX))) + (prr + dn_r * (sm - UR(IMX))));

        double vsl[3], vsr[3], bsl[3], bsr[3];
        vsl[0] = sm;
        vsr[0] = sm;
        bsl[0] = BL(0);
        bsr[0] = BR(0);

        double coeff_1 = dn_l * slsm - b_lr;
        double ue = UL(IEN);
        if ((fabs(coeff_1) > or_spacing(b_lr)) && b_lrgam <= ue) {
            bsl[1] = (BL(1) * (dn_l * slvxl - b_lr)) / coeff_1;
            bsl[2] = (BL(2) * (dn_l * slvxl - b_lr)) / coeff_1;
        } else {
            bsl[1] = ((sr * BR(1) - sl * BL(1)) - (bcr[1] - bcl[1])) / srmsl;
            bsl[2] = ((sr * BR(2) - sl * BL(2)) - (bcr[2] - bcl[2])) / srmsl;
        }
        coeff_1 = dn_r * srsm - b_lr;
        ue = UR(IEN);
        if ((fabs(coeff_1) > or_spacing(b_lr)) && b_lrgam <= ue) {
            bsr[1] = (BR(1) * (dn_r * srvxr - b_lr)) / coeff_1;
            bsr[2] = (BR(2) * (dn_r * srvxr - b_lr)) / coeff_1;
        } else {
            bsr[1] = ((sr * BR(1) - sl * BL(1)) - (bcr[1] - bcl[1])) / srmsl;
            bsr[2] = ((sr * BR(2) - sl * BL(2)) - (bcr[2] - bcl[2])) / srmsl;
        }

        vsl[1] = UL(IMY) + (BL(0) / dn_l) * (BL(1) - bsl[1]);
        vsl[2] = UL(IMZ) + (BL(0) / dn_l) * (BL(2) - bsl[2]);
        vsr[1] = UR(IMY) + (BR(0) / dn_r) * (BR(1) - bsr[1]);
        vsr[2] = UR(IMZ) + (BR(0) / dn_r) * (BR(2) - bsr[2]);

        double vb_l = (UL(IMX) * BL(0) + UL(IMY) * BL(1)) + UL(IMZ) * BL(2);
        double vb_r = (UR(IMX) * BR(0) + UR(IMY) * BR(1)) + UR(IMZ) * BR(2);
        double vb_starl = (vsl[0] * bsl[0] + vsl[1] * bsl[1]) + vsl[2] * bsl[2];
        double vb_starr = (vsr[0] * bsr[0] + vsr[1] * bsr[1]) + vsr[2] * bsr[2];

        double usl[5], usr[5];
        usl[IDN] = dn_l / slsm;
        usl[IMX] = usl[IDN] * vsl[0];
        usl[IMY] = usl[IDN] * vsl[1];
        usl[IMZ] = usl[IDN] * vsl[2];
        usr[IDN] = dn_r / srsm;
        usr[IMX] = usr[IDN] * vsr[0];
        usr[IMY] = usr[IDN] * vsr[1];
        usr[IMZ] = usr[IDN] * vsr[2];
        usl[IEN] = ((((slvxl * enl) - (prl * UL(IMX))) + (prt_star * sm)) + BL(0) * (vb_l - vb_starl)) / slsm;
        usr[IEN] = ((((srvxr * enr) - (prr * UR(IMX))) + (prt_star * sm)) + BR(0) * (vb_r - vb_starr)) / srsm;

        double Ul[5] = {UL(IDN), UL(IDN) * UL(IMX), UL(IDN) * UL(IMY), UL(IDN) * UL(IMZ), enl};
        double Ur[5] = {UR(IDN), UR(IDN) * UR(IMX), UR(IDN) * UR(IMY), UR(IDN) * UR(IMZ), enr};

        if (fabs(BL(0)) > 0.0) {
            double dn_lsqt = sqrt(usl[IDN]);
            double dn_rsqt = sqrt(usr[IDN]);
            double alfven_l = sm - fabs(BL(0)) / dn_lsqt;
            double alfven_r = sm + fabs(BR(0)) / dn_rsqt;
            if (alfven_l > 0.0) {
                for (int v = 0; v < 5; v++) F(v) = fl[v] + sl * (usl[v] - Ul[v]);
                BF(1) = bcl[1] + sl * (bsl[1] - BL(1));
                BF(2) = bcl[2] + sl * (bsl[2] - BL(2));
            } else if (alfven_r < 0.0) {
                for (int v = 0; v < 5; v++) F(v) = fr[v] + sr * (usr[v] - Ur[v]);
                BF(1) = bcr[1] + sr * (bsr[1] - BR(1));
                BF(2) = bcr[2] + sr * (bsr[2] - BR(2));
            } else {
                double b_sig = (BL(0) >= 0.0) ? 1.0 : -1.0;
                double add_dnsq = dn_lsqt + dn_rsqt;
                double mul_dnsq = dn_lsqt * dn_rsqt;
                double v2[3], b2[3];
                v2[0] = sm;
                for (int c = 1; c < 3; c++)
                    v2[c] = ((dn_lsqt * vsl[c] + dn_rsqt * vsr[c]) + b_sig * (bsr[c] - bsl[c])) / add_dnsq;
                b2[0] = 0.5 * (BL(0) + BR(0));
                for (int c = 1; c < 3; c++)
                    b2[c] = ((dn_lsqt * bsr[c] + dn_rsqt * bsl[c]) + (b_sig * mul_dnsq) * (vsr[c] - vsl[c])) / add_dnsq;
                double vb_2star = (v2[0] * b2[0] + v2[1] * b2[1]) + v2[2] * b2[2];
                double u2l[5] = {0}, u2r[5] = {0};
                if (sm >= 0.0) {
                    u2l[IDN] = usl[IDN];
                    u2l[IMX] = usl[IDN] * v2[0];
                    u2l[IMY] = usl[IDN] * v2[1];
                    u2l[IMZ] = usl[IDN] * v2[2];
                    u2l[IEN] = usl[IEN] - (b_sig * dn_lsqt) * (vb_starl - vb_2star);
                }
                if (sm <= 0.0) {
                    u2r[IDN] = usr[IDN];
                    u2r[IMX] = usr[IDN] * v2[0];
                    u2r[IMY] = usr[IDN] * v2[1];
                    u2r[IMZ] = usr[IDN] * v2[2];
                    u2r[IEN] = usr[IEN] + (b_sig * dn_rsqt) * (vb_starr - vb_2star);
                }
                if (sm > 0.0) {
                    for (int v = 0; v < 5; v++)
                        F(v) = ((fl[v] + alfven_l * u2l[v]) - (alfven_l - sl) * usl[v]) - sl * Ul[v];
                    for (int c = 1; c < 3; c++)
                        BF(c) = ((bcl[c] + alfven_l * b2[c]) - (alfven_l - sl) * bsl[c]) - sl * BL(c);
                } else if (sm < 0.0) {
                    for (int v = 0; v < 5; v++)
                        F(v) = ((fr[v] + alfven_r * u2r[v]) - (alfven_r - sr) * usr[v]) - sr * Ur[v];
                    for (int c = 1; c < 3; c++)
                        BF(c) = ((bcr[c] + alfven_r * b2[c]) - (alfven_r - sr) * bsr[c]) - sr * BR(c);
                } else {
                    for (int v = 0; v < 5; v++)
                        F(v) = 0.5 * ((((fl[v] + alfven_l * u2l[v]) - (alfven_l - sl) * usl[v]) - sl * Ul[v]) +
                                      (((fr[v] + alfven_r * u2r[v]) - (alfven_r - sr) * usr[v]) - sr * Ur[v]));
                    for (int c = 1; c < 3; c++)
                        BF(c) = 0.5 * ((((bcl[c] + alfven_l * b2[c]) - (alfven_l - sl) * bsl[c]) - sl * BL(c)) +
                                       (((bcr[c] + alfven_r * b2[c]) - (alfven_r - sr) * bsr[c]) - sr * BR(c)));
                }
            }
        } else {
            if (sm > 0.0) {
                for (int v = 0; v < 5; v++) F(v) = fl[v] + sl * (usl[v] - Ul[v]);
                BF(1) = bcl[1] + sl * (bsl[1] - BL(1));
                BF(2) = bcl[2] + sl * (bsl[2] - BL(2));
            } else if (sm < 0.0) {
                for (int v = 0; v < 5; v++) F(v) = fr[v] + sr * (usr[v] - Ur[v]);
                BF(1) = bcr[1] + sr * (bsr[1] - BR(1));
                BF(2) = bcr[2] + sr * (bsr[2] - BR(2));
            } else {
                for (int v = 0; v < 5; v++)
                    F(v) = 0.5 * (((fl[v] + sl * (usl[v] - Ul[v])) + fr[v]) + sr * (usr[v] - Ur[v]));
                for (int c = 1; c < 3; c++)
                    BF(c) = 0.5 * (((bcl[c] + sl * (bsl[c] - BL(c))) + bcr[c]) + sr * (bsr[c] - BR(c)));
            }
        }
    }
#undef UL
#undef UR
#undef BL
#undef BR
#undef F
#undef BF
}

static int any_magnetized(const or_params *p)
{
    for (int f = 0; f < p->nfluids; f++)
        if (p->fl[f].magnetized) return 1;
    return 0;
}

/* care_for_positives (sources.F90:299-324) on a pencil u1 (n, nvar) with
 * the updated B b1 (n, >=3, sweep order) for a magnetized fluid. */
void or_floors(const or_params *p, int n, double *u1, const double *b1, int nbp, double dvol,
               or_report *rep)
{
    /* care_for_positives -> limit_minimal_density, sources.F90:326-392 */
    for (int f = 0; f < p->nfluids; f++) {
        double *dn = u1 + (size_t)(p->fl[f].beg + IDN) * n;
        for (int i = 0; i < n; i++)
            if (dn[i] < 0.0) rep->dn_negative = 1;
    }
    if (p->use_smalld) {
        /* local_magic_mass (sources.F90:356-358): per fluid, the pencil-interior
         * mass nbp..n-nbp-1 (nb+1:n-nb) before the floor is taken off and the
         * mass after it added back, each sum left to right times dvol */
        for (int f = 0; f < p->nfluids; f++) {
            double *dn = u1 + (size_t)(p->fl[f].beg + IDN) * n;
            double s = 0.0;
            for (int i = nbp; i < n - nbp; i++) s += dn[i];
            rep->magic_mass[f] = rep->magic_mass[f] - s * dvol;
            for (int i = 0; i < n; i++) dn[i] = fmax2(dn[i], p->smalld);
            s = 0.0;
            for (int i = nbp; i < n - nbp; i++) s += dn[i];
            rep->magic_mass[f] = rep->magic_mass[f] + s * dvol;
        }
    }
    /* limit_minimal_intener, sources.F90:394-433 */
    for (int f = 0; f < p->nfluids; f++) {
        const or_fluid *fl = &p->fl[f];
        int o = fl->beg;
        double *dn = u1 + (size_t)(o + IDN) * n, *mx = u1 + (size_t)(o + IMX) * n,
               *my = u1 + (size_t)(o + IMY) * n, *mz = u1 + (size_t)(o + IMZ) * n,
               *en = u1 + (size_t)(o + IEN) * n;
        for (int i = 0; i < n; i++) {
            double kin = (0.5 * ((mx[i] * mx[i] + my[i] * my[i]) + mz[i] * mz[i])) / dn[i];
            double mag = 0.0, ei;
            if (fl->magnetized) {
                mag = 0.5 * ((b1[i] * b1[i] + b1[n + i] * b1[n + i]) + b1[2 * n + i] * b1[2 * n + i]);
                ei = (en[i] - kin) - mag;
            } else {
                ei = en[i] - kin;
            }
            if (p->disallow_negatives && ei < 0.0) rep->ei_negative = 1; /* sources.F90:425 */
            if (p->use_smallei) ei = fmax2(ei, p->smallei);
            en[i] = ei + kin;
            if (fl->magnetized) en[i] = en[i] + mag;
        }
    }
}

/*
 * The interface fluxes of one pencil (solve / solve_u of the unsplit driver,
 * unsplit_mag_modules.F90:166-209, solve_cg_unsplit.F90:157-187):
 * interpol -> riemann_wrap.  u (n, nvar), b (n, 4) = [B_n, B_t1, B_t2, psi]
 * or NULL; flx (n-1, nvar), bflx (n-1, 4): interface i+1/2 at row i.
 */
void or_pencil_flux(const or_params *p, int n, double chspeed, const double *u, const double *b, double *flx,
                    double *bflx)
{
    int nv = p->nvar, m = n - 1;
    int mhd = any_magnetized(p);
    /* pencil work arrays on the stack: no allocation per pencil */
    double q[n * nv], ql[m * nv], qr[m * nv], bl[m * 4], br[m * 4];
    or_utoq(p, n, u, mhd ? b : NULL, q);
    or_linear(p->limiter, n, nv, q, ql, qr);
    if (mhd) or_linear(p->limiter_b, n, 4, b, bl, br);
    for (int f = 0; f < p->nfluids; f++) {
        int o = p->fl[f].beg;
        if (p->fl[f].magnetized)
            or_hlld(m, ql + (size_t)o * m, qr + (size_t)o * m, bl, br, p->fl[f].gam, chspeed, flx + (size_t)o * m,
                    bflx);
        else
            or_hllc(m, ql + (size_t)o * m, qr + (size_t)o * m, p->fl[f].gam, flx + (size_t)o * m);
    }
}

/*
 * One stage of the pencil operator as run for one pencil by solve_cg_ub /
 * solve_cg_u (solve_cg_riemann.F90:202-222 and :290-298):
 *   u1 = u; b1 = b
 *   solve/solve_u:   interpol(u1) -> riemann -> u1(2:n-1) = u0 + dtodx*(f(i-1)-f(i))
 *   internal_sources: u1 = u1 + 0*coeffdt   (all sources vanish here)
 *   care_for_positives (density floor, internal-energy floor, energy re-sum)
 * u0, u, u1 are (n, nvar); b0, b, b1 are (n, 4) = [B_n, B_t1, B_t2, psi] or NULL.
 */
void or_pencil_stage(const or_params *p, int n, double dtodx, double chspeed,
                     const double *u0, const double *u, const double *b0,
                     const double *b, double *u1, double *b1, int nbp, double dvol,
                     or_report *rep)
{
    or_pencil_stage_ext(p, n, dtodx, chspeed, u0, u, b0, b, u1, b1, nbp, dvol, rep, NULL);
}

/*
 * or_pencil_stage with the pencil's external fluxes (ext_fluxes,
 * fluxtypes.F90:44-51; e may be NULL): after riemann_wrap the flux rows li
 * and ri are replaced by the fluxes a finer neighbour sent, then rows lo and
 * ro are saved for a coarser neighbour, u-fluxes first, then with HDC the
 * [B, psi] fluxes (solve_cg_riemann.F90:353-363, solve_u :405-408).
 */
void or_pencil_stage_ext(const or_params *p, int n, double dtodx, double chspeed,
                         const double *u0, const double *u, const double *b0,
                         const double *b, double *u1, double *b1, int nbp, double dvol,
                         or_report *rep, const or_peflx *e)
{
    int nv = p->nvar, m = n - 1;
    int mhd = any_magnetized(p);
    memcpy(u1, u, sizeof(double) * (size_t)n * nv);
    if (mhd) memcpy(b1, b, sizeof(double) * (size_t)n * 4);

    /* pencil work arrays on the stack: no allocation per pencil */
    double q[n * nv], ql[m * nv], qr[m * nv], flx[m * nv], bl[m * 4], br[m * 4], bflx[m * 4];

    or_utoq(p, n, u1, mhd ? b1 : NULL, q);
    or_linear(p->limiter, n, nv, q, ql, qr);
    if (mhd) or_linear(p->limiter_b, n, 4, b1, bl, br);
    /* riemann_wrap / riemann_wrap_u: per-fluid column slices (hlld.F90:84-111, 135-153) */
    for (int f = 0; f < p->nfluids; f++) {
        int o = p->fl[f].beg;
        if (p->fl[f].magnetized)
            or_hlld(m, ql + (size_t)o * m, qr + (size_t)o * m, bl, br, p->fl[f].gam, chspeed,
                    flx + (size_t)o * m, bflx);
        else
            or_hllc(m, ql + (size_t)o * m, qr + (size_t)o * m, p->fl[f].gam, flx + (size_t)o * m);
    }
    if (e) {
        for (int k = 0; k < 4; k++) {
            int r = e->row[k];
            if (r < 0) continue;
            for (int v = 0; v < nv; v++) {
                if (k < 2) flx[(size_t)v * m + r] = e->u[k][v];
                else e->u[k][v] = flx[(size_t)v * m + r];
            }
        }
        if (mhd)
            for (int k = 0; k < 4; k++) {
                int r = e->row[k];
                if (r < 0) continue;
                for (int v = 0; v < 4; v++) {
                    if (k < 2) bflx[(size_t)v * m + r] = e->b[k][v];
                    else e->b[k][v] = bflx[(size_t)v * m + r];
                }
            }
    }
    /* solve_cg_riemann.F90:366,368,411 */
    for (int v = 0; v < nv; v++)
        for (int i = 1; i < n - 1; i++)
            u1[(size_t)v * n + i] = u0[(size_t)v * n + i] +
                                    dtodx * (flx[(size_t)v * m + i - 1] - flx[(size_t)v * m + i]);
    if (mhd)
        for (int v = 0; v < 4; v++)
            for (int i = 1; i < n - 1; i++)
                b1[(size_t)v * n + i] = b0[(size_t)v * n + i] +
                                        dtodx * (bflx[(size_t)v * m + i - 1] - bflx[(size_t)v * m + i]);

    /* internal_sources, sources.F90:239: usrc == +0 for these setups */
    for (size_t k = 0; k < (size_t)n * nv; k++) u1[k] = u1[k] + 0.0;

    or_floors(p, n, u1, mhd ? b1 : NULL, nbp, dvol, rep);
}
