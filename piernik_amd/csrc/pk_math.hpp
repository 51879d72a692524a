// pk_math.hpp -- per-cell and per-interface FP64 arithmetic of the split
// Riemann sweep, for gfx950.
//
// Every function keeps the reference's floating-point evaluation order
// (left-to-right sums, explicit parentheses, x**2 as x*x) and is compiled with
// -ffp-contract=off, so each operation rounds exactly as PIERNIK's FP64 path:
//   prim()        src/scheme/Riemann_Solver/interpolations.F90:69-106 (utoq)
//   limit()       src/scheme/Riemann_Solver/fluxlimiters.F90:109-216
//   flux_hllc()   src/scheme/Riemann_Solver/hlld.F90:615-841 (riemann_hlld_u)
//   flux_hlld()   src/scheme/Riemann_Solver/hlld.F90:161-609 (riemann_hlld, DIVB_HDC)
//   floors()      src/fluids/sources.F90:299-433 (care_for_positives)
//
// Variables are in sweep order: the normal component comes first.
//   HD  (5):  rho, m_n, m_t1, m_t2, E            prims: rho, v_n, v_t1, v_t2, p
//   MHD (9):  HD + B_n, B_t1, B_t2, psi          prims: HD prims + B, psi
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef PK_RELAXED
#define PK_RELAXED 0
#endif

// Callable from device code and, for host-side checks of the same
// arithmetic, from host code.
#define PK_HD __host__ __device__ __forceinline__

// The same source is compiled twice: exactly (namespace pk) and with FMA
// contraction and fast division (namespace pk_fast, pk_kernels_fast.hip).
#ifndef PK_KNS
#define PK_KNS pk
#endif

namespace PK_KNS {

enum Limiter : int { LIM_VANLEER = 0, LIM_MINMOD = 1, LIM_MONCEN = 2, LIM_SUPERBEE = 3 };

// Fortran min/max.  The operands here are never NaN and never a +0/-0 pair
// whose sign could reach a result, so the hardware min/max is exact.
PK_HD double fmn(double a, double b) { return fmin(a, b); }
PK_HD double fmx(double a, double b) { return fmax(a, b); }

// a / b and sqrt(x).  The relaxed build (PK_RELAXED: the FMA-contracted
// kernels) forms them from the hardware reciprocal and reciprocal square root
// with one Newton-Raphson step each and a final residual correction, which
// leaves them within an ulp or two of the correctly rounded values for the
// finite, nonzero (division) and positive normal (sqrt) operands of the
// sweep, in fewer instructions than the compiler's fast expansions (no
// range scaling, one refinement less); sqrt of 0 is 0.  The exact build
// uses the correctly rounded operations, as the reference does.
PK_HD double qdiv(double a, double b)
{
#if PK_RELAXED && defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(b);
    r = fma(fma(-b, r, 1.0), r, r);
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);
#else
    return a / b;
#endif
}

PK_HD double qsqrt(double x)
{
#if PK_RELAXED && defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    g = fma(fma(-g, g, x), h, g);
    return x > 0.0 ? g : 0.0;
#else
    return sqrt(x);
#endif
}

// Fortran SPACING(x) for real(8): 2**max(exponent(x)-53, -1022); TINY for 0.
// For a biased exponent field E >= 53 that is the power of two with exponent
// field E-52, otherwise 2**-1022.
PK_HD double spacing(double x)
{
    const uint64_t bits = __builtin_bit_cast(uint64_t, x);
    const uint64_t e = (bits >> 52) & 0x7ffull;
    return (e >= 53) ? __builtin_bit_cast(double, (e - 52) << 52) : 2.2250738585072014e-308;
}

// One slope of the configured limiter from the backward (dl) and forward (dr)
// differences.  fluxlimiters.F90: vanleer :125-131, minmod :159,
// moncen :186, superbee :212-213.
PK_HD double limit(int lim, double dl, double dr)
{
    if (lim == LIM_VANLEER) {
        const double dcen = dl * dr;
        return (dcen > 0.0) ? qdiv(2.0 * dcen, dl + dr) : 0.0;
    }
    const double s = copysign(1.0, dl) + copysign(1.0, dr);
    const double a = fabs(dl), b = fabs(dr);
    if (lim == LIM_MINMOD) return (s * fmn(a, b)) * 0.5;
    if (lim == LIM_MONCEN) return (s * fmn(fmn(2.0 * a, 2.0 * b), 0.5 * fabs(dl + dr))) * 0.5;
    const double m = (a > b) ? fmn(a, fabs(2.0 * dr)) : fmn(fabs(2.0 * dl), b);
    return (s * 0.5) * m;
}

struct HD {
    static constexpr int NV = 5;
    static constexpr bool MAG = false;
};
struct MHD {
    static constexpr int NV = 9;
    static constexpr bool MAG = true;
};

// Conserved -> primitive (utoq).  gam_1 = gamma - 1 as stored by the
// reference (fluidtypes.F90:336).
template <class P>
PK_HD void prim(const double (&u)[P::NV], double gam_1, double (&q)[P::NV])
{
    const double d = u[0];
    q[0] = d;
    q[1] = qdiv(u[1], d);
    q[2] = qdiv(u[2], d);
    q[3] = qdiv(u[3], d);
    const double ek = qdiv(0.5 * ((u[1] * u[1] + u[2] * u[2]) + u[3] * u[3]), d);
    double p = gam_1 * (u[4] - ek);
    if constexpr (P::MAG) {
        const double b2 = (u[5] * u[5] + u[6] * u[6]) + u[7] * u[7];
        p = p - (0.5 * gam_1) * b2;
        q[5] = u[5];
        q[6] = u[6];
        q[7] = u[7];
        q[8] = u[8];
    }
    q[4] = p;
}

// HLLC (riemann_hlld_u, adiabatic) for one interface.  l, r: primitive face
// states; f: fluxes of rho, m_n, m_t1, m_t2, E.
PK_HD void flux_hllc(const double *l, const double *r, double gam, double *f)
{
    const double prl = l[4], prr = r[4];
    const double enl = qdiv(l[4], gam - 1.0) + (0.5 * l[0]) * ((l[1] * l[1] + l[2] * l[2]) + l[3] * l[3]);
    const double enr = qdiv(r[4], gam - 1.0) + (0.5 * r[0]) * ((r[1] * r[1] + r[2] * r[2]) + r[3] * r[3]);
    const double c = qsqrt(fmx(qdiv(gam * l[4], l[0]), qdiv(gam * r[4], r[0])));
    const double sl = fmn(l[1], r[1]) - c;
    const double sr = fmx(l[1], r[1]) + c;

    double fl[5], fr[5];
    fl[0] = l[0] * l[1];
    fl[1] = l[0] * (l[1] * l[1]) + prl;
    fl[2] = (l[0] * l[2]) * l[1];
    fl[3] = (l[0] * l[3]) * l[1];
    fl[4] = (enl + prl) * l[1];
    fr[0] = r[0] * r[1];
    fr[1] = r[0] * (r[1] * r[1]) + prr;
    fr[2] = (r[0] * r[2]) * r[1];
    fr[3] = (r[0] * r[3]) * r[1];
    fr[4] = (enr + prr) * r[1];

    if (sl >= 0.0) {
#pragma unroll
        for (int v = 0; v < 5; v++) f[v] = fl[v];
        return;
    }
    if (sr <= 0.0) {
#pragma unroll
        for (int v = 0; v < 5; v++) f[v] = fr[v];
        return;
    }
    const double auxr = (sr - r[1]) * r[0], auxl = (sl - l[1]) * l[0];
    const double sm = (fabs(auxr - auxl) < spacing(auxr + auxl))
                          ? (sl + sr) / 2.0
                          : qdiv((auxr * r[1] - prr) - (auxl * l[1] - prl), auxr - auxl);
    const double slvxl = sl - l[1], srvxr = sr - r[1];
    const double dn_l = l[0] * slvxl, dn_r = r[0] * srvxr;
    const double prt_star = 0.5 * ((prl + dn_l * (sm - l[1])) + (prr + dn_r * (sm - r[1])));
    if (sm != 0.0) {
        // one side only: pick its inputs, then evaluate the reference's
        // formula once (hlld.F90:771-815)
        const bool L = sm > 0.0;
        const double s_ = L ? sl : sr;
        const double svx = L ? slvxl : srvxr;
        const double en = L ? enl : enr;
        const double pr = L ? prl : prr;
        const double dn = L ? dn_l : dn_r;
        const double* q = L ? l : r;
        const double* f0 = L ? fl : fr;
        const double ssm = s_ - sm;
        const double d = qdiv(dn, ssm);
        const double us[5] = {d, d * sm, d * q[2], d * q[3], qdiv(((svx * en) - (pr * q[1])) + (prt_star * sm), ssm)};
        const double U[5] = {q[0], q[0] * q[1], q[0] * q[2], q[0] * q[3], en};
#pragma unroll
        for (int v = 0; v < 5; v++) f[v] = f0[v] + s_ * (us[v] - U[v]);
    } else {
        // sm == 0: half*(fl + sl*(us_l - U_l) + fr + sr*(us_r - U_r)), hlld.F90:822
        const double slsm = sl - sm, srsm = sr - sm;
        const double dl_ = dn_l / slsm, dr_ = dn_r / srsm;
        const double usl[5] = {dl_, dl_ * sm, dl_ * l[2], dl_ * l[3],
                               (((slvxl * enl) - (prl * l[1])) + (prt_star * sm)) / slsm};
        const double usr[5] = {dr_, dr_ * sm, dr_ * r[2], dr_ * r[3],
                               (((srvxr * enr) - (prr * r[1])) + (prt_star * sm)) / srsm};
        const double Ul[5] = {l[0], l[0] * l[1], l[0] * l[2], l[0] * l[3], enl};
        const double Ur[5] = {r[0], r[0] * r[1], r[0] * r[2], r[0] * r[3], enr};
#pragma unroll
        for (int v = 0; v < 5; v++)
            f[v] = 0.5 * (((fl[v] + sl * (usl[v] - Ul[v])) + fr[v]) + sr * (usr[v] - Ur[v]));
    }
}

// One side of an HLLD interface: face state q = [rho, v_n, v_t1, v_t2, p,
// B_n, B_t1, B_t2], |B|^2 and total pressure.  Gives the total energy
// (hlld.F90:263-264), the physical flux (:296-308) and the conserved state.
struct HlldSide {
    double en;
    double f[5], fby, fbz;
    double U[5];
};

PK_HD void hlld_side(const double *q, double b2, double pr, double gam, HlldSide &s)
{
    s.en = (qdiv(q[4], gam - 1.0) + (0.5 * q[0]) * ((q[1] * q[1] + q[2] * q[2]) + q[3] * q[3])) + 0.5 * b2;
    s.f[0] = q[0] * q[1];
    s.f[1] = (q[0] * (q[1] * q[1]) + pr) - q[5] * q[5];
    s.f[2] = (q[0] * q[2]) * q[1] - q[5] * q[6];
    s.f[3] = (q[0] * q[3]) * q[1] - q[5] * q[7];
    s.f[4] = (s.en + pr) * q[1] - q[5] * ((q[1] * q[5] + q[2] * q[6]) + q[3] * q[7]);
    s.fby = q[6] * q[1] - q[5] * q[2];
    s.fbz = q[7] * q[1] - q[5] * q[3];
    s.U[0] = q[0];
    s.U[1] = q[0] * q[1];
    s.U[2] = q[0] * q[2];
    s.U[3] = q[0] * q[3];
    s.U[4] = s.en;
}

// HLLD with the GLM psi flux (riemann_hlld, DIVB_HDC) for one interface.
// l, r: primitive face states [rho, v_n, v_t1, v_t2, p, B_n, B_t1, B_t2, psi].
// f: fluxes of [rho, m_n, m_t1, m_t2, E, B_n, B_t1, B_t2, psi].
//
// Every branch of the reference that has a left and a right variant is
// evaluated for the one side the interface needs: that side's inputs are
// selected first and the reference's formula is applied once (a - x*y is
// written a + ((-1)*x)*y, which rounds identically).  Only the two sm == 0
// averages, which need both sides, evaluate both.
PK_HD void flux_hlld(const double *l, const double *r, double gam, double ch, double *f)
{
    // Dedner et al. eq. 42 (hlld.F90:229-230)
    f[8] = ((ch * ch) * 0.5) * ((r[5] + l[5]) - qdiv(r[8] - l[8], ch));
    f[5] = 0.5 * ((r[8] + l[8]) - ch * (r[5] - l[5]));

    const double bxl = l[5], byl = l[6], bzl = l[7];
    const double bxr = r[5], byr = r[6], bzr = r[7];
    const double b2l = (bxl * bxl + byl * byl) + bzl * bzl;
    const double b2r = (bxr * bxr + byr * byr) + bzr * bzr;
    const double prl = l[4] + 0.5 * b2l;
    const double prr = r[4] + 0.5 * b2r;
    const double gpl = gam * l[4], gpr = gam * r[4];
    const double tl = gpl + b2l, tr = gpr + b2r;
    const double cl = qdiv(tl + qsqrt(tl * tl - (4.0 * gpl) * (bxl * bxl)), l[0]);
    const double cr = qdiv(tr + qsqrt(tr * tr - (4.0 * gpr) * (bxr * bxr)), r[0]);
    const double c = qsqrt(0.5 * fmx(cl, cr));
    const double sl = fmn(l[1], r[1]) - c;
    const double sr = fmx(l[1], r[1]) + c;

    if (sl >= 0.0 || sr <= 0.0) {
        // supersonic: the upwind physical flux (hlld.F90:312-319)
        const bool L = sl >= 0.0;
        HlldSide s;
        hlld_side(L ? l : r, L ? b2l : b2r, L ? prl : prr, gam, s);
#pragma unroll
        for (int v = 0; v < 5; v++) f[v] = s.f[v];
        f[6] = s.fby;
        f[7] = s.fbz;
        return;
    }

    const double auxr = (sr - r[1]) * r[0], auxl = (sl - l[1]) * l[0];
    const double sm = (fabs(auxr - auxl) < spacing(auxr + auxl))
                          ? (sl + sr) / 2.0
                          : qdiv((auxr * r[1] - prr) - (auxl * l[1] - prl), auxr - auxl);
    const double slsm = sl - sm, srsm = sr - sm;
    const double slvxl = sl - l[1], srvxr = sr - r[1];
    const double srmsl = sr - sl;
    const double dn_l = l[0] * slvxl, dn_r = r[0] * srvxr;
    const double b_lr = bxl * bxr;
    const double b_lrgam = qdiv(b_lr, gam);
    const double prt_star = 0.5 * ((prl + dn_l * (sm - l[1])) + (prr + dn_r * (sm - r[1])));
    const double fbyl = byl * l[1] - bxl * l[2], fbzl = bzl * l[1] - bxl * l[3];
    const double fbyr = byr * r[1] - bxr * r[2], fbzr = bzr * r[1] - bxr * r[3];

    // Transverse B of the star states with the HLL fallback (hlld.F90:362-378)
    double bsly, bslz, bsry, bsrz;
    {
        const double c1 = dn_l * slsm - b_lr;
        const bool okl = (fabs(c1) > spacing(b_lr)) && b_lrgam <= l[4];
        const double c2 = dn_r * srsm - b_lr;
        const bool okr = (fabs(c2) > spacing(b_lr)) && b_lrgam <= r[4];
        double hy = 0.0, hz = 0.0;
        if (!(okl && okr)) {
            hy = qdiv((sr * byr - sl * byl) - (fbyr - fbyl), srmsl);
            hz = qdiv((sr * bzr - sl * bzl) - (fbzr - fbzl), srmsl);
        }
        if (okl) {
            const double k = dn_l * slvxl - b_lr;
            bsly = qdiv(byl * k, c1);
            bslz = qdiv(bzl * k, c1);
        } else {
            bsly = hy;
            bslz = hz;
        }
        if (okr) {
            const double k = dn_r * srvxr - b_lr;
            bsry = qdiv(byr * k, c2);
            bsrz = qdiv(bzr * k, c2);
        } else {
            bsry = hy;
            bsrz = hz;
        }
    }
    const double vsly = l[2] + qdiv(bxl, dn_l) * (byl - bsly);
    const double vslz = l[3] + qdiv(bxl, dn_l) * (bzl - bslz);
    const double vsry = r[2] + qdiv(bxr, dn_r) * (byr - bsry);
    const double vsrz = r[3] + qdiv(bxr, dn_r) * (bzr - bsrz);
    const double dsl = qdiv(dn_l, slsm), dsr = qdiv(dn_r, srsm);

    // which region holds x/t = 0 (hlld.F90:410-599)
    const bool bxnz = fabs(bxl) > 0.0;
    double dn_lsqt = 0.0, dn_rsqt = 0.0, alfven_l = 0.0, alfven_r = 0.0;
    bool dbl, L;
    if (bxnz) {
        dn_lsqt = qsqrt(dsl);
        dn_rsqt = qsqrt(dsr);
        alfven_l = sm - qdiv(fabs(bxl), dn_lsqt);
        alfven_r = sm + qdiv(fabs(bxr), dn_rsqt);
        const bool ls = alfven_l > 0.0, rs = !ls && (alfven_r < 0.0);
        dbl = !ls && !rs;
        L = dbl ? (sm > 0.0) : ls;
    } else {
        dbl = false;
        L = sm > 0.0;
    }

    // double-star transverse v and B (hlld.F90:458-474), needed by both averages
    const double b_sig = (bxl >= 0.0) ? 1.0 : -1.0;
    double v2y = 0.0, v2z = 0.0, b2y = 0.0, b2z = 0.0, vb2 = 0.0;
    if (dbl) {
        const double add = dn_lsqt + dn_rsqt;
        const double mul = dn_lsqt * dn_rsqt;
        v2y = qdiv((dn_lsqt * vsly + dn_rsqt * vsry) + b_sig * (bsry - bsly), add);
        v2z = qdiv((dn_lsqt * vslz + dn_rsqt * vsrz) + b_sig * (bsrz - bslz), add);
        const double b2x = 0.5 * (bxl + bxr);
        b2y = qdiv((dn_lsqt * bsry + dn_rsqt * bsly) + (b_sig * mul) * (vsry - vsly), add);
        b2z = qdiv((dn_lsqt * bsrz + dn_rsqt * bslz) + (b_sig * mul) * (vsrz - vslz), add);
        vb2 = (sm * b2x + v2y * b2y) + v2z * b2z;
    }

    if (sm == 0.0 && (dbl || !bxnz)) {
        // both sides, averaged: hlld.F90:522-543 (Alfven) and :578-596 (B_n = 0)
        HlldSide a, b;
        hlld_side(l, b2l, prl, gam, a);
        hlld_side(r, b2r, prr, gam, b);
        const double vb_l = (l[1] * bxl + l[2] * byl) + l[3] * bzl;
        const double vb_r = (r[1] * bxr + r[2] * byr) + r[3] * bzr;
        const double vb_sl = (sm * bxl + vsly * bsly) + vslz * bslz;
        const double vb_sr = (sm * bxr + vsry * bsry) + vsrz * bsrz;
        const double usl[5] = {dsl, dsl * sm, dsl * vsly, dsl * vslz,
                               ((((slvxl * a.en) - (prl * l[1])) + (prt_star * sm)) + bxl * (vb_l - vb_sl)) / slsm};
        const double usr[5] = {dsr, dsr * sm, dsr * vsry, dsr * vsrz,
                               ((((srvxr * b.en) - (prr * r[1])) + (prt_star * sm)) + bxr * (vb_r - vb_sr)) / srsm};
        if (dbl) {
            const double u2l[5] = {dsl, dsl * sm, dsl * v2y, dsl * v2z, usl[4] - (b_sig * dn_lsqt) * (vb_sl - vb2)};
            const double u2r[5] = {dsr, dsr * sm, dsr * v2y, dsr * v2z, usr[4] + (b_sig * dn_rsqt) * (vb_sr - vb2)};
#pragma unroll
            for (int v = 0; v < 5; v++)
                f[v] = 0.5 * ((((a.f[v] + alfven_l * u2l[v]) - (alfven_l - sl) * usl[v]) - sl * a.U[v]) +
                              (((b.f[v] + alfven_r * u2r[v]) - (alfven_r - sr) * usr[v]) - sr * b.U[v]));
            f[6] = 0.5 * ((((fbyl + alfven_l * b2y) - (alfven_l - sl) * bsly) - sl * byl) +
                          (((fbyr + alfven_r * b2y) - (alfven_r - sr) * bsry) - sr * byr));
            f[7] = 0.5 * ((((fbzl + alfven_l * b2z) - (alfven_l - sl) * bslz) - sl * bzl) +
                          (((fbzr + alfven_r * b2z) - (alfven_r - sr) * bsrz) - sr * bzr));
        } else {
#pragma unroll
            for (int v = 0; v < 5; v++)
                f[v] = 0.5 * (((a.f[v] + sl * (usl[v] - a.U[v])) + b.f[v]) + sr * (usr[v] - b.U[v]));
            f[6] = 0.5 * (((fbyl + sl * (bsly - byl)) + fbyr) + sr * (bsry - byr));
            f[7] = 0.5 * (((fbzl + sl * (bslz - bzl)) + fbzr) + sr * (bsrz - bzr));
        }
        return;
    }

    // the one side the interface needs
    const double *q = L ? l : r;
    const double s_ = L ? sl : sr;
    const double svx = L ? slvxl : srvxr;
    const double ssm = L ? slsm : srsm;
    const double ds = L ? dsl : dsr;
    const double pr = L ? prl : prr;
    const double vsy = L ? vsly : vsry, vsz = L ? vslz : vsrz;
    const double bsy = L ? bsly : bsry, bsz = L ? bslz : bsrz;
    HlldSide o;
    hlld_side(q, L ? b2l : b2r, pr, gam, o);
    const double vb = (q[1] * q[5] + q[2] * q[6]) + q[3] * q[7];
    const double vbs = (sm * q[5] + vsy * bsy) + vsz * bsz;
    const double us[5] = {ds, ds * sm, ds * vsy, ds * vsz,
                          qdiv((((svx * o.en) - (pr * q[1])) + (prt_star * sm)) + q[5] * (vb - vbs), ssm)};
    if (dbl) {
        // Alfven intermediate flux, Eq. 65 (hlld.F90:496-520)
        const double al = L ? alfven_l : alfven_r;
        const double sq = L ? dn_lsqt : dn_rsqt;
        const double sg = L ? -1.0 : 1.0;
        const double u2[5] = {ds, ds * sm, ds * v2y, ds * v2z, us[4] + (sg * (b_sig * sq)) * (vbs - vb2)};
#pragma unroll
        for (int v = 0; v < 5; v++) f[v] = ((o.f[v] + al * u2[v]) - (al - s_) * us[v]) - s_ * o.U[v];
        f[6] = ((o.fby + al * b2y) - (al - s_) * bsy) - s_ * q[6];
        f[7] = ((o.fbz + al * b2z) - (al - s_) * bsz) - s_ * q[7];
    } else {
        // single-star flux, Eq. 64
#pragma unroll
        for (int v = 0; v < 5; v++) f[v] = o.f[v] + s_ * (us[v] - o.U[v]);
        f[6] = o.fby + s_ * (bsy - q[6]);
        f[7] = o.fbz + s_ * (bsz - q[7]);
    }
}

// Floors after the update (care_for_positives): density floor, then
// internal-energy floor with the unconditional energy re-sum.  The B used for
// the magnetic energy is the updated one, in sweep order.  Bit 0 of flags: a
// negative density was seen (sources.F90:349-350); bit 1: a negative internal
// energy was seen, recorded only when ei_flag is set (disallow_negatives,
// sources.F90:425).  dmass receives the mass the density floor added to the
// cell (max(rho, smalld) - rho, which is >= 0).
template <class P>
PK_HD void floors(double (&u)[P::NV], double smalld, double smallei, bool use_smalld, bool use_smallei,
                  bool ei_flag, unsigned &flags, double &dmass)
{
    if (u[0] < 0.0) flags |= 1u;
    if (use_smalld) {
        const double nd = fmx(u[0], smalld);
        dmass = nd - u[0];
        u[0] = nd;
    }
    const double kin = qdiv(0.5 * ((u[1] * u[1] + u[2] * u[2]) + u[3] * u[3]), u[0]);
    double ei, mag = 0.0;
    if constexpr (P::MAG) {
        mag = 0.5 * ((u[5] * u[5] + u[6] * u[6]) + u[7] * u[7]);
        ei = (u[4] - kin) - mag;
    } else {
        ei = u[4] - kin;
    }
    if (ei_flag && ei < 0.0) flags |= 2u;
    if (use_smallei) ei = fmx(ei, smallei);
    double e = ei + kin;
    if constexpr (P::MAG) e = e + mag;
    u[4] = e;
}

}  // namespace PK_KNS
