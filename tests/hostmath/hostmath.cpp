// Host build of the sweep's per-interface arithmetic (piernik_amd/csrc/pk_math.hpp)
// for bitwise comparison with the oracle on the CPU.  Test helper only.
#include "../../piernik_amd/csrc/pk_adjoint.hpp"
#include "../../piernik_amd/csrc/pk_mass.hpp"

extern "C" {

double hm_spacing(double x) { return pk::spacing(x); }

// The floor's mass counters (pk_mass.hpp) over n values, each added digit by
// digit as a lane adds it on the device; returns the rounded total.
double hm_mass_sum(int n, const double *x)
{
    unsigned long long acc[pk::kMassDigits] = {};
    for (int k = 0; k < n; k++) {
        uint32_t d[3];
        const int i = pk::mass_place(x[k], d[0], d[1], d[2]);
        if (i < 0) continue;
        for (int j = 0; j < 3; j++)
            if (i + j < pk::kMassDigits) acc[i + j] += d[j];
    }
    return pk::mass_value(acc);
}

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
void hm_floors(int m, int mhd, double smalld, double smallei, int use_smalld, int use_smallei, int ei_flag,
               double *u, unsigned *flags)
{
    for (int i = 0; i < m; i++) {
        double dm = 0.0;
        if (mhd) {
            double a[9];
            for (int v = 0; v < 9; v++) a[v] = u[v * m + i];
            pk::floors<pk::MHD>(a, smalld, smallei, use_smalld, use_smallei, ei_flag, *flags, dm);
            for (int v = 0; v < 9; v++) u[v * m + i] = a[v];
        } else {
            double a[5];
            for (int v = 0; v < 5; v++) a[v] = u[v * m + i];
            pk::floors<pk::HD>(a, smalld, smallei, use_smalld, use_smallei, ei_flag, *flags, dm);
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
    pk::floors<pk::HD>(a, smalld, smallei, use_smalld, use_smallei, true, fl, dm);
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

void hm_hlld_vjp(const double *l, const double *r, double gam, double ch, const double *fb, double *lb, double *rb,
                 double *chb)
{
    for (int v = 0; v < 9; v++) lb[v] = rb[v] = 0.0;
    *chb = pk::hlld_vjp(l, r, gam, ch, fb, lb, rb);
}

void hm_hlld1(const double *l, const double *r, double gam, double ch, double *f) { pk::flux_hlld(l, r, gam, ch, f); }

void hm_prim_vjp_mhd(double gam_1, const double *u, const double *qb, double *ub)
{
    double a[9], b[9], o[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int v = 0; v < 9; v++) {
        a[v] = u[v];
        b[v] = qb[v];
    }
    pk::prim_vjp_mhd(a, gam_1, b, o);
    for (int v = 0; v < 9; v++) ub[v] = o[v];
}

void hm_floors_mhd(double smalld, double smallei, int use_smalld, int use_smallei, double *u)
{
    double a[9];
    for (int v = 0; v < 9; v++) a[v] = u[v];
    unsigned fl = 0;
    double dm = 0.0;
    pk::floors<pk::MHD>(a, smalld, smallei, use_smalld, use_smallei, true, fl, dm);
    for (int v = 0; v < 9; v++) u[v] = a[v];
}

void hm_floors_vjp_mhd(double smalld, double smallei, int use_smalld, int use_smallei, const double *u, double *ub)
{
    double a[9], b[9];
    for (int v = 0; v < 9; v++) {
        a[v] = u[v];
        b[v] = ub[v];
    }
    pk::floors_vjp_mhd(a, smalld, smallei, use_smalld, use_smallei, b);
    for (int v = 0; v < 9; v++) ub[v] = b[v];
}

void hm_dt_vjp_mhd(const double *u, int qdir, double dl, double gam, double tb, double *ub)
{
    double a[9], b[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int v = 0; v < 9; v++) a[v] = u[v];
    pk::dt_vjp_mhd(a, qdir, dl, gam, gam - 1.0, tb, b);
    for (int v = 0; v < 9; v++) ub[v] = b[v];
}

// branch of flux_hlld: out = {code, min margin}; code 0/1 supersonic L/R,
// 2/3 single star L/R, 4/5 double star L/R, 6 both (Alfven, sm = 0),
// 7 both (B_n = 0, sm = 0), 8/9 B_n = 0 one side L/R; +16 if the star B took
// the HLL fallback
void hm_hlld_branch(const double *l, const double *r, double gam, double *out)
{
    const double bxl = l[5], byl = l[6], bzl = l[7], bxr = r[5], byr = r[6], bzr = r[7];
    const double b2l = (bxl * bxl + byl * byl) + bzl * bzl, b2r = (bxr * bxr + byr * byr) + bzr * bzr;
    const double prl = l[4] + 0.5 * b2l, prr = r[4] + 0.5 * b2r;
    const double gpl = gam * l[4], gpr = gam * r[4];
    const double tl = gpl + b2l, tr = gpr + b2r;
    const double cl = (tl + sqrt(tl * tl - (4.0 * gpl) * (bxl * bxl))) / l[0];
    const double cr = (tr + sqrt(tr * tr - (4.0 * gpr) * (bxr * bxr))) / r[0];
    const double c = sqrt(0.5 * fmax(cl, cr));
    const double sl = fmin(l[1], r[1]) - c, sr = fmax(l[1], r[1]) + c;
    double m = fmin(fabs(sl), fabs(sr));
    m = fmin(m, fabs(l[1] - r[1]));
    m = fmin(m, fabs(cl - cr));
    if (sl >= 0.0 || sr <= 0.0) {
        out[0] = sl >= 0.0 ? 0 : 1;
        out[1] = m;
        return;
    }
    const double auxr = (sr - r[1]) * r[0], auxl = (sl - l[1]) * l[0];
    const double sm = ((auxr * r[1] - prr) - (auxl * l[1] - prl)) / (auxr - auxl);
    const double slsm = sl - sm, srsm = sr - sm, slvxl = sl - l[1], srvxr = sr - r[1];
    const double dn_l = l[0] * slvxl, dn_r = r[0] * srvxr, b_lr = bxl * bxr;
    const double c1 = dn_l * slsm - b_lr, c2 = dn_r * srsm - b_lr;
    const bool okl = (fabs(c1) > pk::spacing(b_lr)) && b_lr / gam <= l[4];
    const bool okr = (fabs(c2) > pk::spacing(b_lr)) && b_lr / gam <= r[4];
    m = fmin(m, fabs(b_lr / gam - l[4]));
    m = fmin(m, fabs(b_lr / gam - r[4]));
    m = fmin(m, fabs(c1));
    m = fmin(m, fabs(c2));
    const int hll = (okl && okr) ? 0 : 16;
    const bool bxnz = fabs(bxl) > 0.0;
    if (sm != 0.0) m = fmin(m, fabs(sm));
    if (!bxnz) {
        out[0] = (sm == 0.0 ? 7 : (sm > 0.0 ? 8 : 9)) + hll;
        out[1] = m;
        return;
    }
    const double dsl = dn_l / slsm, dsr = dn_r / srsm;
    const double al = sm - fabs(bxl) / sqrt(dsl), ar = sm + fabs(bxr) / sqrt(dsr);
    m = fmin(m, fabs(al));
    m = fmin(m, fabs(ar));
    m = fmin(m, fabs(bxl));
    const bool ls = al > 0.0, rs = !ls && ar < 0.0, dbl = !ls && !rs;
    int code;
    if (dbl)
        code = sm == 0.0 ? 6 : (sm > 0.0 ? 4 : 5);
    else
        code = ls ? 2 : 3;
    out[0] = code + hll;
    out[1] = m;
}
}
