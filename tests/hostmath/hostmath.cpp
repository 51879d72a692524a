// Host build of the sweep's per-interface arithmetic (piernik_amd/csrc/pk_math.hpp)
// for bitwise comparison with the oracle on the CPU.  Test helper only.
#include "../../piernik_amd/csrc/pk_adjoint.hpp"

extern "C" {

double hm_spacing(double x) { return pk::spacing(x); }

double hm_limit(int lim, double dl, double dr) { return pk::limit(lim, dl, dr); }

// u: (m, 5) or (m, 9) column-major conserved states -> q (same shape)
void hm_prim(int m, int mhd, double gam_1, const double *u, double *q)
{
    for (int i = 0; i < m; i++) {
        if (mhd) {
            double a[9], b[9];
            for (int v = 0; v < 9; v++) a[v] = u[v * m + i];
            pk::prim<pk::MHD>(a, gam_1, b);
            for (int v = 0; v < 9; v++) q[v * m + i] = b[v];
        } else {
            double a[5], b[5];
            for (int v = 0; v < 5; v++) a[v] = u[v * m + i];
            pk::prim<pk::HD>(a, gam_1, b);
            for (int v = 0; v < 5; v++) q[v * m + i] = b[v];
        }
    }
}

// ql, qr: (m, 5) column-major primitive face states -> f (m, 5)
void hm_hllc(int m, const double *ql, const double *qr, double gam, double *f)
{
    for (int i = 0; i < m; i++) {
        double l[5], r[5], o[5];
        for (int v = 0; v < 5; v++) {
            l[v] = ql[v * m + i];
            r[v] = qr[v * m + i];
        }
        pk::flux_hllc(l, r, gam, o);
        for (int v = 0; v < 5; v++) f[v * m + i] = o[v];
    }
}

// ql, qr: (m, 9) column-major [rho, v_n, v_t1, v_t2, p, B_n, B_t1, B_t2, psi] -> f (m, 9)
void hm_hlld(int m, const double *ql, const double *qr, double gam, double ch, double *f)
{
    for (int i = 0; i < m; i++) {
        double l[9], r[9], o[9];
        for (int v = 0; v < 9; v++) {
            l[v] = ql[v * m + i];
            r[v] = qr[v * m + i];
        }
        pk::flux_hlld(l, r, gam, ch, o);
        for (int v = 0; v < 9; v++) f[v * m + i] = o[v];
    }
}

// u: (m, nv) column-major; applies the floors cell by cell
void hm_floors(int m, int mhd, double smalld, double smallei, int use_smalld, int use_smallei, double *u,
               unsigned *flags)
{
    for (int i = 0; i < m; i++) {
        double dm = 0.0;
        if (mhd) {
            double a[9];
            for (int v = 0; v < 9; v++) a[v] = u[v * m + i];
            pk::floors<pk::MHD>(a, smalld, smallei, use_smalld, use_smallei, *flags, dm);
            for (int v = 0; v < 9; v++) u[v * m + i] = a[v];
        } else {
            double a[5];
            for (int v = 0; v < 5; v++) a[v] = u[v * m + i];
            pk::floors<pk::HD>(a, smalld, smallei, use_smalld, use_smallei, *flags, dm);
            for (int v = 0; v < 5; v++) u[v * m + i] = a[v];
        }
    }
}

// ---- reverse mode (pk_adjoint.hpp) ----
void hm_prim_vjp_hd(double gam_1, const double *u, const double *qb, double *ub)
{
    double a[5], b[5], o[5] = {0, 0, 0, 0, 0};
    for (int v = 0; v < 5; v++) {
        a[v] = u[v];
        b[v] = qb[v];
    }
    pk::prim_vjp_hd(a, gam_1, b, o);
    for (int v = 0; v < 5; v++) ub[v] = o[v];
}

void hm_limit_vjp(int lim, double dl, double dr, double dqb, double *out)
{
    pk::limit_vjp(lim, dl, dr, dqb, out[0], out[1]);
}

void hm_hllc_vjp(const double *l, const double *r, double gam, const double *fb, double *lb, double *rb)
{
    for (int v = 0; v < 5; v++) lb[v] = rb[v] = 0.0;
    pk::hllc_vjp(l, r, gam, fb, lb, rb);
}

void hm_floors_hd(double smalld, double smallei, int use_smalld, int use_smallei, double *u)
{
    double a[5];
    for (int v = 0; v < 5; v++) a[v] = u[v];
    unsigned fl = 0;
    double dm = 0.0;
    pk::floors<pk::HD>(a, smalld, smallei, use_smalld, use_smallei, fl, dm);
    for (int v = 0; v < 5; v++) u[v] = a[v];
}

void hm_floors_vjp_hd(double smalld, double smallei, int use_smalld, int use_smallei, const double *u, double *ub)
{
    double a[5], b[5];
    for (int v = 0; v < 5; v++) {
        a[v] = u[v];
        b[v] = ub[v];
    }
    pk::floors_vjp_hd(a, smalld, smallei, use_smalld, use_smallei, b);
    for (int v = 0; v < 5; v++) ub[v] = b[v];
}

void hm_dt_vjp_hd(const double *u, int qdir, double dl, double gam, double tb, double *ub)
{
    double a[5], b[5] = {0, 0, 0, 0, 0};
    for (int v = 0; v < 5; v++) a[v] = u[v];
    pk::dt_vjp_hd(a, qdir, dl, gam, gam - 1.0, tb, b);
    for (int v = 0; v < 5; v++) ub[v] = b[v];
}
}
