// pk_adjoint.hpp -- reverse mode (vector-Jacobian products) of the per-cell
// and per-interface arithmetic of the hydro sweep in pk_math.hpp:
//   prim_vjp_hd()   utoq            (interpolations.F90:69-106)
//   limit_vjp()     the limiters    (fluxlimiters.F90:109-216)
//   hllc_vjp()      riemann_hlld_u  (hlld.F90:615-841)
//   floors_vjp_hd() care_for_positives (sources.F90:299-433)
//   dt_vjp_hd()     timestep_fluid at one cell (timestep.F90:460-490, neu_cs)
//
// Each function re-evaluates the forward operations in the forward's order
// (so the linearisation point is the forward's value bit for bit) and then
// runs the chain rule backwards through the branch the forward took.
// Where a min or max has two exactly equal arguments the derivative is split
// evenly between them (the forward value does not depend on the choice).
#pragma once
#include "pk_math.hpp"

namespace PK_KNS {

// d min(a, b): weights of a and b
PK_HD void dmin(double a, double b, double &wa, double &wb)
{
    wa = (a < b) ? 1.0 : ((a > b) ? 0.0 : 0.5);
    wb = 1.0 - wa;
}
PK_HD void dmax(double a, double b, double &wa, double &wb)
{
    wa = (a > b) ? 1.0 : ((a < b) ? 0.0 : 0.5);
    wb = 1.0 - wa;
}

// ---- utoq (HD) -------------------------------------------------------------
// q = prim(u); given qb = dL/dq, add dL/du to ub.
PK_HD void prim_vjp_hd(const double (&u)[5], double gam_1, const double (&qb)[5], double (&ub)[5])
{
    const double d = u[0];
    const double q1 = u[1] / d, q2 = u[2] / d, q3 = u[3] / d;
    const double ek = (0.5 * ((u[1] * u[1] + u[2] * u[2]) + u[3] * u[3])) / d;
    // p = gam_1 * (E - ek)
    double db = qb[0];
    ub[4] += gam_1 * qb[4];
    const double ekb = -(gam_1 * qb[4]);
    // ek = S / d, S = 0.5*(m1^2 + m2^2 + m3^2)
    const double sb = ekb / d;
    db -= ekb * ek / d;
    ub[1] += sb * u[1];
    ub[2] += sb * u[2];
    ub[3] += sb * u[3];
    // v_k = m_k / d
    ub[1] += qb[1] / d;
    ub[2] += qb[2] / d;
    ub[3] += qb[3] / d;
    db -= (qb[1] * q1 + qb[2] * q2 + qb[3] * q3) / d;
    ub[0] += db;
}

// ---- limiters ----------------------------------------------------------------
// dq = limit(lim, dl, dr); given dqb, return dL/d(dl), dL/d(dr).
PK_HD void limit_vjp(int lim, double dl, double dr, double dqb, double &dlb, double &drb)
{
    dlb = 0.0;
    drb = 0.0;
    if (lim == LIM_VANLEER) {
        const double dcen = dl * dr;
        if (dcen > 0.0) {
            // 2*dl*dr/(dl+dr)
            const double s = dl + dr;
            const double s2 = s * s;
            dlb = dqb * (2.0 * dr * dr) / s2;
            drb = dqb * (2.0 * dl * dl) / s2;
        }
        return;
    }
    const double sgl = copysign(1.0, dl), sgr = copysign(1.0, dr);
    const double s = sgl + sgr;
    if (s == 0.0) return;  // opposite signs: zero slope
    const double a = fabs(dl), b = fabs(dr);
    if (lim == LIM_MINMOD) {
        // (s * min(a, b)) * 0.5
        double wa, wb;
        dmin(a, b, wa, wb);
        const double mb = (s * 0.5) * dqb;
        dlb = mb * wa * sgl;
        drb = mb * wb * sgr;
        return;
    }
    if (lim == LIM_MONCEN) {
        // (s * min(min(2a, 2b), 0.5*|dl+dr|)) * 0.5
        const double m1 = fmn(2.0 * a, 2.0 * b), m2 = 0.5 * fabs(dl + dr);
        double w1, w2, wa, wb;
        dmin(m1, m2, w1, w2);
        dmin(2.0 * a, 2.0 * b, wa, wb);
        const double mb = (s * 0.5) * dqb;
        const double sc = copysign(1.0, dl + dr);
        dlb = mb * (w1 * wa * 2.0 * sgl + w2 * 0.5 * sc);
        drb = mb * (w1 * wb * 2.0 * sgr + w2 * 0.5 * sc);
        return;
    }
    // superbee: (s*0.5) * (a > b ? min(a, |2 dr|) : min(|2 dl|, b))
    const double mb = (s * 0.5) * dqb;
    double w1, w2;
    if (a > b) {
        dmin(a, fabs(2.0 * dr), w1, w2);
        dlb = mb * w1 * sgl;
        drb = mb * w2 * 2.0 * sgr;
    } else {
        dmin(fabs(2.0 * dl), b, w1, w2);
        dlb = mb * w1 * 2.0 * sgl;
        drb = mb * w2 * sgr;
    }
}

// ---- HLLC --------------------------------------------------------------------

// Reverse of one side's star state and flux (hlld.F90:771-815):
//   f = f0 + s_ * (us - U) with weight w (1, or 1/2 when both sides are
//   averaged), ssm = s_ - sm, d = dn / ssm,
//   us = [d, d*sm, d*q2, d*q3, (((svx*en) - (pr*q1)) + (prt*sm)) / ssm],
//   U = [q0, q0*q1, q0*q2, q0*q3, en].
struct HllcSideBar {
    double s, svx, en, pr, dn, sm, prt;
    double q[5], f0[5];
};

PK_HD void hllc_side_vjp(const double *q, double s_, double svx, double en, double pr, double dn, double sm,
                         double prt, double w, const double *fb, HllcSideBar &o)
{
    const double ssm = s_ - sm;
    const double d = dn / ssm;
    const double num = (((svx * en) - (pr * q[1])) + (prt * sm));
    const double us[5] = {d, d * sm, d * q[2], d * q[3], num / ssm};
    const double U[5] = {q[0], q[0] * q[1], q[0] * q[2], q[0] * q[3], en};
    double usb[5], Ub[5];
    o.s = 0.0;
#pragma unroll
    for (int v = 0; v < 5; v++) {
        const double g = w * fb[v];
        o.f0[v] = g;
        o.s += g * (us[v] - U[v]);
        usb[v] = s_ * g;
        Ub[v] = -(s_ * g);
    }
    // U
    o.q[0] = Ub[0] + Ub[1] * q[1] + Ub[2] * q[2] + Ub[3] * q[3];
    o.q[1] = Ub[1] * q[0];
    o.q[2] = Ub[2] * q[0];
    o.q[3] = Ub[3] * q[0];
    o.q[4] = 0.0;
    o.en = Ub[4];
    // us
    double db = usb[0] + usb[1] * sm + usb[2] * q[2] + usb[3] * q[3];
    o.sm = usb[1] * d;
    o.q[2] += usb[2] * d;
    o.q[3] += usb[3] * d;
    const double numb = usb[4] / ssm;
    double ssmb = -(usb[4] * us[4]) / ssm;
    o.svx = numb * en;
    o.en += numb * svx;
    o.pr = -(numb * q[1]);
    o.q[1] += -(numb * pr);
    o.prt = numb * sm;
    o.sm += numb * prt;
    // d = dn / ssm
    o.dn = db / ssm;
    ssmb -= db * d / ssm;
    // ssm = s_ - sm
    o.s += ssmb;
    o.sm -= ssmb;
}

// f = flux_hllc(l, r); given fb, add dL/dl to lb and dL/dr to rb.
PK_HD void hllc_vjp(const double *l, const double *r, double gam, const double *fb, double *lb, double *rb)
{
    const double prl = l[4], prr = r[4];
    const double kl = (l[1] * l[1] + l[2] * l[2]) + l[3] * l[3];
    const double kr = (r[1] * r[1] + r[2] * r[2]) + r[3] * r[3];
    const double enl = (l[4] / (gam - 1.0)) + (0.5 * l[0]) * kl;
    const double enr = (r[4] / (gam - 1.0)) + (0.5 * r[0]) * kr;
    const double cl2 = (gam * l[4]) / l[0], cr2 = (gam * r[4]) / r[0];
    const double c = sqrt(fmx(cl2, cr2));
    const double sl = fmn(l[1], r[1]) - c;
    const double sr = fmx(l[1], r[1]) + c;

    double flb[5] = {0, 0, 0, 0, 0}, frb[5] = {0, 0, 0, 0, 0};
    double slb = 0.0, srb = 0.0;
    double enlb = 0.0, enrb = 0.0, prlb = 0.0, prrb = 0.0;
    double lq[5] = {0, 0, 0, 0, 0}, rq[5] = {0, 0, 0, 0, 0};

    if (sl >= 0.0) {
#pragma unroll
        for (int v = 0; v < 5; v++) flb[v] = fb[v];
    } else if (sr <= 0.0) {
#pragma unroll
        for (int v = 0; v < 5; v++) frb[v] = fb[v];
    } else {
        const double auxr = (sr - r[1]) * r[0], auxl = (sl - l[1]) * l[0];
        const bool degen = fabs(auxr - auxl) < spacing(auxr + auxl);
        const double sm = degen ? (sl + sr) / 2.0 : ((auxr * r[1] - prr) - (auxl * l[1] - prl)) / (auxr - auxl);
        const double slvxl = sl - l[1], srvxr = sr - r[1];
        const double dn_l = l[0] * slvxl, dn_r = r[0] * srvxr;
        const double prt = 0.5 * ((prl + dn_l * (sm - l[1])) + (prr + dn_r * (sm - r[1])));
        double smb = 0.0, prtb = 0.0, dnlb = 0.0, dnrb = 0.0, slvxlb = 0.0, srvxrb = 0.0;
        const bool useL = sm >= 0.0, useR = sm <= 0.0;  // both when sm == 0
        const double w = (useL && useR) ? 0.5 : 1.0;
        if (useL) {
            HllcSideBar o;
            hllc_side_vjp(l, sl, slvxl, enl, prl, dn_l, sm, prt, w, fb, o);
            slb += o.s;
            slvxlb += o.svx;
            enlb += o.en;
            prlb += o.pr;
            dnlb += o.dn;
            smb += o.sm;
            prtb += o.prt;
#pragma unroll
            for (int v = 0; v < 5; v++) {
                lq[v] += o.q[v];
                flb[v] += o.f0[v];
            }
        }
        if (useR) {
            HllcSideBar o;
            hllc_side_vjp(r, sr, srvxr, enr, prr, dn_r, sm, prt, w, fb, o);
            srb += o.s;
            srvxrb += o.svx;
            enrb += o.en;
            prrb += o.pr;
            dnrb += o.dn;
            smb += o.sm;
            prtb += o.prt;
#pragma unroll
            for (int v = 0; v < 5; v++) {
                rq[v] += o.q[v];
                frb[v] += o.f0[v];
            }
        }
        // prt = 0.5*((prl + dn_l*(sm - l1)) + (prr + dn_r*(sm - r1)))
        const double a = 0.5 * prtb;
        prlb += a;
        prrb += a;
        dnlb += a * (sm - l[1]);
        dnrb += a * (sm - r[1]);
        smb += a * dn_l + a * dn_r;
        lq[1] -= a * dn_l;
        rq[1] -= a * dn_r;
        // dn = q0 * svx
        lq[0] += dnlb * slvxl;
        slvxlb += dnlb * l[0];
        rq[0] += dnrb * srvxr;
        srvxrb += dnrb * r[0];
        // svx = s - q1
        slb += slvxlb;
        lq[1] -= slvxlb;
        srb += srvxrb;
        rq[1] -= srvxrb;
        // sm
        if (degen) {
            slb += 0.5 * smb;
            srb += 0.5 * smb;
        } else {
            const double den = auxr - auxl;
            const double nb = smb / den, denb = -(smb * sm) / den;
            double auxrb = nb * r[1] + denb;
            double auxlb = -(nb * l[1]) - denb;
            rq[1] += nb * auxr;
            lq[1] -= nb * auxl;
            prrb -= nb;
            prlb += nb;
            // aux = (s - q1) * q0
            srb += auxrb * r[0];
            rq[1] -= auxrb * r[0];
            rq[0] += auxrb * (sr - r[1]);
            slb += auxlb * l[0];
            lq[1] -= auxlb * l[0];
            lq[0] += auxlb * (sl - l[1]);
        }
        // sl = min(l1, r1) - c, sr = max(l1, r1) + c, c = sqrt(max(cl2, cr2))
        double wl, wr;
        dmin(l[1], r[1], wl, wr);
        lq[1] += slb * wl;
        rq[1] += slb * wr;
        dmax(l[1], r[1], wl, wr);
        lq[1] += srb * wl;
        rq[1] += srb * wr;
        const double cb = srb - slb;
        const double cmb = cb / (2.0 * c);
        dmax(cl2, cr2, wl, wr);
        const double cl2b = cmb * wl, cr2b = cmb * wr;
        lq[4] += cl2b * gam / l[0];
        lq[0] -= cl2b * cl2 / l[0];
        rq[4] += cr2b * gam / r[0];
        rq[0] -= cr2b * cr2 / r[0];
    }
    // physical fluxes of both sides
    const double *qs[2] = {l, r};
    const double *fbs[2] = {flb, frb};
    double *qb[2] = {lq, rq};
    const double ens[2] = {enl, enr};
    const double ks[2] = {kl, kr};
    double enb[2] = {enlb, enrb}, prb[2] = {prlb, prrb};
#pragma unroll
    for (int side = 0; side < 2; side++) {
        const double *q = qs[side];
        const double *g = fbs[side];
        double *o = qb[side];
        // f0 = q0*q1; f1 = q0*(q1*q1) + p; f2 = (q0*q2)*q1; f3 = (q0*q3)*q1; f4 = (en + p)*q1
        o[0] += g[0] * q[1] + g[1] * (q[1] * q[1]) + g[2] * q[2] * q[1] + g[3] * q[3] * q[1];
        o[1] += g[0] * q[0] + g[1] * 2.0 * q[0] * q[1] + g[2] * q[0] * q[2] + g[3] * q[0] * q[3] +
                g[4] * (ens[side] + q[4]);
        o[2] += g[2] * q[0] * q[1];
        o[3] += g[3] * q[0] * q[1];
        prb[side] += g[1] + g[4] * q[1];
        enb[side] += g[4] * q[1];
        // en = p/(gam-1) + (0.5*q0) * K, K = q1^2 + q2^2 + q3^2
        o[4] += enb[side] / (gam - 1.0) + prb[side];
        o[0] += enb[side] * 0.5 * ks[side];
        const double kb = enb[side] * 0.5 * q[0];
        o[1] += 2.0 * q[1] * kb;
        o[2] += 2.0 * q[2] * kb;
        o[3] += 2.0 * q[3] * kb;
    }
#pragma unroll
    for (int v = 0; v < 5; v++) {
        lb[v] += lq[v];
        rb[v] += rq[v];
    }
}

// ---- floors (HD) ---------------------------------------------------------------
// u_out = floors(u_in) (sources.F90:299-433): rho = max(rho, smalld); then
// e = E - ek, e = max(e, smallei), E = e + ek (always re-summed).
// Given ub = dL/du_out (in place), returns dL/du_in.
PK_HD void floors_vjp_hd(const double (&u)[5], double smalld, double smallei, int use_smalld, int use_smallei,
                         double (&ub)[5])
{
    double d = u[0];
    const bool dfl = use_smalld && !(d >= smalld);  // max(d, smalld) took smalld
    if (use_smalld) d = fmx(d, smalld);
    if (!use_smallei) {
        if (dfl) ub[0] = 0.0;
        return;
    }
    const double ek = (0.5 * ((u[1] * u[1] + u[2] * u[2]) + u[3] * u[3])) / d;
    const double e = u[4] - ek;
    const bool efl = !(e >= smallei);
    // E_out = max(e, smallei) + ek
    const double eb = efl ? 0.0 : ub[4];
    const double ekb = ub[4] - eb;  // through "+ ek" and "e = E - ek"
    double db = ub[0];
    ub[4] = eb;
    const double sb = ekb / d;
    db -= ekb * ek / d;
    ub[1] += sb * u[1];
    ub[2] += sb * u[2];
    ub[3] += sb * u[3];
    ub[0] = dfl ? 0.0 : db;
}

// ---- timestep_fluid at one cell, one direction (HD, neu_cs) ----------------------
// t = dl / max(|m_q/rho| + cs, small), cs = sqrt(|gam*p/rho|),
// p = (E - ek)*(gam-1).  Given tb = dL/dt, add dL/du to ub.
PK_HD void dt_vjp_hd(const double (&u)[5], int qdir, double dl, double gam, double gam_1, double tb,
                     double (&ub)[5])
{
    const double small = 1.17549435082228750797e-38;
    const double d = u[0];
    const double vq = (d > 0.0) ? fabs(u[1 + qdir] / d) : 0.0;
    const double ek = (0.5 * ((u[1] * u[1] + u[2] * u[2]) + u[3] * u[3])) / d;
    const double pr = (u[4] - ek) * gam_1;
    const double x = (gam * pr) / d;
    const double cs = sqrt(fabs(x));
    const double c = fmx(vq + cs, small);
    if (!(vq + cs > small)) return;
    const double cb = -(tb * dl) / (c * c);
    // vq = |m/d|
    if (d > 0.0) {
        const double vv = u[1 + qdir] / d;
        const double sg = (vv > 0.0) ? 1.0 : ((vv < 0.0) ? -1.0 : 0.0);
        ub[1 + qdir] += cb * sg / d;
        ub[0] -= cb * sg * vv / d;
    }
    // cs = sqrt(|x|)
    if (cs > 0.0) {
        const double sx = (x > 0.0) ? 1.0 : ((x < 0.0) ? -1.0 : 0.0);
        const double xb = cb * sx / (2.0 * cs);
        const double prb = xb * gam / d;
        ub[0] -= xb * x / d;
        ub[4] += prb * gam_1;
        const double ekb = -(prb * gam_1);
        const double sb = ekb / d;
        ub[0] -= ekb * ek / d;
        ub[1] += sb * u[1];
        ub[2] += sb * u[2];
        ub[3] += sb * u[3];
    }
}

}  // namespace PK_KNS
