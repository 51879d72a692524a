// pk_adjoint.hpp -- reverse mode (vector-Jacobian products) of the per-cell
// and per-interface arithmetic of the hydro sweep in pk_math.hpp:
//   prim_vjp_hd()   utoq            (interpolations.F90:69-106)
//   limit_vjp()     the limiters    (fluxlimiters.F90:109-216)
//   hllc_vjp()      riemann_hlld_u  (hlld.F90:615-841)
//   hlld_vjp()      riemann_hlld with the GLM psi flux (hlld.F90:161-609)
//   prim_vjp_mhd()  utoq with B
//   floors_vjp_hd(), floors_vjp_mhd()   care_for_positives (sources.F90:299-433)
//   dt_vjp_hd(), dt_vjp_mhd()           timestep_fluid at one cell
//                   (timestep.F90:460-490 with neu_cs / ion_cs)
//
// Each function re-evaluates the forward values that select a branch (wave
// speeds and their signs, the star-state tests, limiter cases, min/max
// choices, floors) in the forward's order, bit for bit, and then runs the
// chain rule backwards through the branch the forward took.  Quotients that
// only scale adjoints are formed with one reciprocal per divisor: they
// differ from a division by at most an ulp, far below the derivative's own
// truncation, and an IEEE division costs about ten FP64 instructions.
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
    const double rd = 1.0 / u[0];
    const double q1 = u[1] * rd, q2 = u[2] * rd, q3 = u[3] * rd;
    const double ek = (0.5 * ((u[1] * u[1] + u[2] * u[2]) + u[3] * u[3])) * rd;
    // p = gam_1 * (E - ek)
    double db = qb[0];
    ub[4] += gam_1 * qb[4];
    const double ekb = -(gam_1 * qb[4]);
    // ek = S / d, S = 0.5*(m1^2 + m2^2 + m3^2)
    const double sb = ekb * rd;
    db -= sb * ek;
    ub[1] += sb * u[1];
    ub[2] += sb * u[2];
    ub[3] += sb * u[3];
    // v_k = m_k / d
    ub[1] += qb[1] * rd;
    ub[2] += qb[2] * rd;
    ub[3] += qb[3] * rd;
    db -= (qb[1] * q1 + qb[2] * q2 + qb[3] * q3) * rd;
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
            const double g = dqb / (s * s);
            dlb = g * (2.0 * dr * dr);
            drb = g * (2.0 * dl * dl);
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
    const double rs = 1.0 / ssm;
    const double d = dn * rs;
    const double num = (((svx * en) - (pr * q[1])) + (prt * sm));
    const double us[5] = {d, d * sm, d * q[2], d * q[3], num * rs};
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
    const double numb = usb[4] * rs;
    double ssmb = -(numb * us[4]);
    o.svx = numb * en;
    o.en += numb * svx;
    o.pr = -(numb * q[1]);
    o.q[1] += -(numb * pr);
    o.prt = numb * sm;
    o.sm += numb * prt;
    // d = dn / ssm
    o.dn = db * rs;
    ssmb -= o.dn * d;
    // ssm = s_ - sm
    o.s += ssmb;
    o.sm -= ssmb;
}

// f = flux_hllc(l, r); given fb, add dL/dl to lb and dL/dr to rb.
PK_HD void hllc_vjp(const double *l, const double *r, double gam, const double *fb, double *lb, double *rb)
{
    const double prl = l[4], prr = r[4];
    const double rg1 = 1.0 / (gam - 1.0);
    const double kl = (l[1] * l[1] + l[2] * l[2]) + l[3] * l[3];
    const double kr = (r[1] * r[1] + r[2] * r[2]) + r[3] * r[3];
    const double enl = (l[4] * rg1) + (0.5 * l[0]) * kl;
    const double enr = (r[4] * rg1) + (0.5 * r[0]) * kr;
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
            const double nb = smb / (auxr - auxl), denb = -(nb * sm);
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
        const double gl = (cmb * wl) / l[0], gr = (cmb * wr) / r[0];
        lq[4] += gl * gam;
        lq[0] -= gl * cl2;
        rq[4] += gr * gam;
        rq[0] -= gr * cr2;
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
        o[4] += enb[side] * rg1 + prb[side];
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

// ---- HLLD + GLM ----------------------------------------------------------------

// Reverse of hlld_side(): q[0..7], b2 = |B|^2, pr = total pressure.  Given
// the adjoints of the side's flux (fb[5], fby, fbz), conserved state (Ub[5])
// and energy (enb, from uses outside the side), add to qb[0..7], b2b, prb.
PK_HD void hlld_side_vjp(const double *q, double b2, double pr, double gam, const double *fb, double fbyb,
                         double fbzb, const double *Ub, double enb, double *qb, double &b2b, double &prb)
{
    const double K = (q[1] * q[1] + q[2] * q[2]) + q[3] * q[3];
    const double rg1 = 1.0 / (gam - 1.0);
    const double en = ((q[4] * rg1) + (0.5 * q[0]) * K) + 0.5 * b2;
    const double W = (q[1] * q[5] + q[2] * q[6]) + q[3] * q[7];
    const double ent = enb + Ub[4] + fb[4] * q[1];
    const double Wb = -(fb[4] * q[5]);
    prb += fb[1] + fb[4] * q[1];
    qb[0] += fb[0] * q[1] + fb[1] * (q[1] * q[1]) + fb[2] * q[2] * q[1] + fb[3] * q[3] * q[1] + Ub[0] + Ub[1] * q[1] +
             Ub[2] * q[2] + Ub[3] * q[3] + ent * 0.5 * K;
    qb[1] += fb[0] * q[0] + fb[1] * 2.0 * q[0] * q[1] + fb[2] * q[0] * q[2] + fb[3] * q[0] * q[3] +
             fb[4] * (en + pr) + Wb * q[5] + fbyb * q[6] + fbzb * q[7] + Ub[1] * q[0] + ent * q[0] * q[1];
    qb[2] += fb[2] * q[0] * q[1] + Wb * q[6] - fbyb * q[5] + Ub[2] * q[0] + ent * q[0] * q[2];
    qb[3] += fb[3] * q[0] * q[1] + Wb * q[7] - fbzb * q[5] + Ub[3] * q[0] + ent * q[0] * q[3];
    qb[4] += ent * rg1;
    qb[5] += -(2.0 * q[5] * fb[1]) - q[6] * fb[2] - q[7] * fb[3] - fb[4] * W + Wb * q[1] - fbyb * q[2] - fbzb * q[3];
    qb[6] += -(q[5] * fb[2]) + Wb * q[2] + fbyb * q[1];
    qb[7] += -(q[5] * fb[3]) + Wb * q[3] + fbzb * q[1];
    b2b += 0.5 * ent;
}

// f = flux_hlld(l, r, gam, ch); given fb[9], add dL/dl to lb, dL/dr to rb
// and return dL/dch.
PK_HD double hlld_vjp(const double *l, const double *r, double gam, double ch, const double *fb, double *lb,
                      double *rb)
{
    // GLM fluxes (hlld.F90:229-230)
    const double kch = (ch * ch) * 0.5;
    const double rch = 1.0 / ch;
    const double A = (r[5] + l[5]) - (r[8] - l[8]) * rch;
    double chb = fb[8] * (ch * A + 0.5 * (r[8] - l[8])) - fb[5] * 0.5 * (r[5] - l[5]);
    lb[5] += fb[8] * kch + fb[5] * 0.5 * ch;
    rb[5] += fb[8] * kch - fb[5] * 0.5 * ch;
    const double g8 = fb[8] * kch * rch;
    lb[8] += g8 + fb[5] * 0.5;
    rb[8] += -g8 + fb[5] * 0.5;

    // ---- forward values ----
    const double bxl = l[5], byl = l[6], bzl = l[7];
    const double bxr = r[5], byr = r[6], bzr = r[7];
    const double b2l = (bxl * bxl + byl * byl) + bzl * bzl;
    const double b2r = (bxr * bxr + byr * byr) + bzr * bzr;
    const double prl = l[4] + 0.5 * b2l;
    const double prr = r[4] + 0.5 * b2r;
    const double gpl = gam * l[4], gpr = gam * r[4];
    const double tl = gpl + b2l, tr = gpr + b2r;
    const double sql = sqrt(tl * tl - (4.0 * gpl) * (bxl * bxl));
    const double sqr = sqrt(tr * tr - (4.0 * gpr) * (bxr * bxr));
    const double cl = (tl + sql) / l[0];
    const double cr = (tr + sqr) / r[0];
    const double c = sqrt(0.5 * fmx(cl, cr));
    const double sl = fmn(l[1], r[1]) - c;
    const double sr = fmx(l[1], r[1]) + c;

    // adjoint accumulators of the shared intermediates
    double b2lb = 0.0, b2rb = 0.0, prlb = 0.0, prrb = 0.0, slb = 0.0, srb = 0.0;
    double ql[8] = {0, 0, 0, 0, 0, 0, 0, 0}, qr[8] = {0, 0, 0, 0, 0, 0, 0, 0};

    if (sl >= 0.0 || sr <= 0.0) {
        const bool L = sl >= 0.0;
        const double Ub[5] = {0, 0, 0, 0, 0};
        hlld_side_vjp(L ? l : r, L ? b2l : b2r, L ? prl : prr, gam, fb, fb[6], fb[7], Ub, 0.0, L ? ql : qr,
                      L ? b2lb : b2rb, L ? prlb : prrb);
    } else {
        const double auxr = (sr - r[1]) * r[0], auxl = (sl - l[1]) * l[0];
        const bool degen = fabs(auxr - auxl) < spacing(auxr + auxl);
        const double sm = degen ? (sl + sr) / 2.0 : ((auxr * r[1] - prr) - (auxl * l[1] - prl)) / (auxr - auxl);
        const double slsm = sl - sm, srsm = sr - sm;
        const double slvxl = sl - l[1], srvxr = sr - r[1];
        const double srmsl = sr - sl;
        const double dn_l = l[0] * slvxl, dn_r = r[0] * srvxr;
        const double b_lr = bxl * bxr;
        const double b_lrgam = b_lr / gam;
        const double prt = 0.5 * ((prl + dn_l * (sm - l[1])) + (prr + dn_r * (sm - r[1])));
        const double fbyl = byl * l[1] - bxl * l[2], fbzl = bzl * l[1] - bxl * l[3];
        const double fbyr = byr * r[1] - bxr * r[2], fbzr = bzr * r[1] - bxr * r[3];
        const double c1 = dn_l * slsm - b_lr;
        const bool okl = (fabs(c1) > spacing(b_lr)) && b_lrgam <= l[4];
        const double c2 = dn_r * srsm - b_lr;
        const bool okr = (fabs(c2) > spacing(b_lr)) && b_lrgam <= r[4];
        const bool hll = !(okl && okr);
        const double hy = hll ? ((sr * byr - sl * byl) - (fbyr - fbyl)) / srmsl : 0.0;
        const double hz = hll ? ((sr * bzr - sl * bzl) - (fbzr - fbzl)) / srmsl : 0.0;
        const double kl = dn_l * slvxl - b_lr, kr = dn_r * srvxr - b_lr;
        const double bsly = okl ? (byl * kl) / c1 : hy, bslz = okl ? (bzl * kl) / c1 : hz;
        const double bsry = okr ? (byr * kr) / c2 : hy, bsrz = okr ? (bzr * kr) / c2 : hz;
        const double vsly = l[2] + (bxl / dn_l) * (byl - bsly);
        const double vslz = l[3] + (bxl / dn_l) * (bzl - bslz);
        const double vsry = r[2] + (bxr / dn_r) * (byr - bsry);
        const double vsrz = r[3] + (bxr / dn_r) * (bzr - bsrz);
        const double dsl = dn_l / slsm, dsr = dn_r / srsm;
        const bool bxnz = fabs(bxl) > 0.0;
        double dn_lsqt = 0.0, dn_rsqt = 0.0, alfven_l = 0.0, alfven_r = 0.0;
        bool dbl, L;
        if (bxnz) {
            dn_lsqt = sqrt(dsl);
            dn_rsqt = sqrt(dsr);
            alfven_l = sm - fabs(bxl) / dn_lsqt;
            alfven_r = sm + fabs(bxr) / dn_rsqt;
            const bool ls = alfven_l > 0.0, rs = !ls && (alfven_r < 0.0);
            dbl = !ls && !rs;
            L = dbl ? (sm > 0.0) : ls;
        } else {
            dbl = false;
            L = sm > 0.0;
        }
        const double b_sig = (bxl >= 0.0) ? 1.0 : -1.0;
        double add = 1.0, mul = 0.0, v2y = 0.0, v2z = 0.0, b2x = 0.0, b2y = 0.0, b2z = 0.0, vb2 = 0.0;
        if (dbl) {
            add = dn_lsqt + dn_rsqt;
            mul = dn_lsqt * dn_rsqt;
            v2y = ((dn_lsqt * vsly + dn_rsqt * vsry) + b_sig * (bsry - bsly)) / add;
            v2z = ((dn_lsqt * vslz + dn_rsqt * vsrz) + b_sig * (bsrz - bslz)) / add;
            b2x = 0.5 * (bxl + bxr);
            b2y = ((dn_lsqt * bsry + dn_rsqt * bsly) + (b_sig * mul) * (vsry - vsly)) / add;
            b2z = ((dn_lsqt * bsrz + dn_rsqt * bslz) + (b_sig * mul) * (vsrz - vslz)) / add;
            vb2 = (sm * b2x + v2y * b2y) + v2z * b2z;
        }
        const double vb_l = (l[1] * bxl + l[2] * byl) + l[3] * bzl;
        const double vb_r = (r[1] * bxr + r[2] * byr) + r[3] * bzr;
        const double vb_sl = (sm * bxl + vsly * bsly) + vslz * bslz;
        const double vb_sr = (sm * bxr + vsry * bsry) + vsrz * bsrz;
        const double rg1 = 1.0 / (gam - 1.0);
        const double enl = ((l[4] * rg1) + (0.5 * l[0]) * ((l[1] * l[1] + l[2] * l[2]) + l[3] * l[3])) + 0.5 * b2l;
        const double enr = ((r[4] * rg1) + (0.5 * r[0]) * ((r[1] * r[1] + r[2] * r[2]) + r[3] * r[3])) + 0.5 * b2r;
        const double nl = (((slvxl * enl) - (prl * l[1])) + (prt * sm)) + bxl * (vb_l - vb_sl);
        const double nr = (((srvxr * enr) - (prr * r[1])) + (prt * sm)) + bxr * (vb_r - vb_sr);
        const double usl4 = nl / slsm, usr4 = nr / srsm;
        const double usl[5] = {dsl, dsl * sm, dsl * vsly, dsl * vslz, usl4};
        const double usr[5] = {dsr, dsr * sm, dsr * vsry, dsr * vsrz, usr4};
        const double u2l4 = usl4 - (b_sig * dn_lsqt) * (vb_sl - vb2);
        const double u2r4 = usr4 + (b_sig * dn_rsqt) * (vb_sr - vb2);
        const double u2l[5] = {dsl, dsl * sm, dsl * v2y, dsl * v2z, u2l4};
        const double u2r[5] = {dsr, dsr * sm, dsr * v2y, dsr * v2z, u2r4};
        const double Ul[5] = {l[0], l[0] * l[1], l[0] * l[2], l[0] * l[3], enl};
        const double Ur[5] = {r[0], r[0] * r[1], r[0] * r[2], r[0] * r[3], enr};

        // ---- output adjoints of the branch taken ----
        // per side: weight of the side, its flux/state/energy adjoints,
        // and the adjoints of us, u2, the star B and the wave speeds
        double wL = 0.0, wR = 0.0;
        const bool both = sm == 0.0 && (dbl || !bxnz);
        if (both) {
            wL = wR = 0.5;
        } else if (L) {
            wL = 1.0;
        } else {
            wR = 1.0;
        }
        double fl_b[5] = {0, 0, 0, 0, 0}, fr_b[5] = {0, 0, 0, 0, 0};
        double Ul_b[5] = {0, 0, 0, 0, 0}, Ur_b[5] = {0, 0, 0, 0, 0};
        double usl_b[5] = {0, 0, 0, 0, 0}, usr_b[5] = {0, 0, 0, 0, 0};
        double u2l_b[5] = {0, 0, 0, 0, 0}, u2r_b[5] = {0, 0, 0, 0, 0};
        double fbyl_b = 0.0, fbzl_b = 0.0, fbyr_b = 0.0, fbzr_b = 0.0;
        double bsly_b = 0.0, bslz_b = 0.0, bsry_b = 0.0, bsrz_b = 0.0;
        double b2y_b = 0.0, b2z_b = 0.0, all_b = 0.0, alr_b = 0.0;
        const double *ys[2] = {l, r};
        for (int side = 0; side < 2; side++) {
            const double w = side == 0 ? wL : wR;
            if (w == 0.0) continue;
            const double *q = ys[side];
            const double s_ = side == 0 ? sl : sr;
            double *f_b = side == 0 ? fl_b : fr_b;
            double *U_b = side == 0 ? Ul_b : Ur_b;
            double *us_b = side == 0 ? usl_b : usr_b;
            double *u2_b = side == 0 ? u2l_b : u2r_b;
            const double *us = side == 0 ? usl : usr;
            const double *u2 = side == 0 ? u2l : u2r;
            const double *U = side == 0 ? Ul : Ur;
            const double bsy = side == 0 ? bsly : bsry, bsz = side == 0 ? bslz : bsrz;
            double &fby_b = side == 0 ? fbyl_b : fbyr_b;
            double &fbz_b = side == 0 ? fbzl_b : fbzr_b;
            double &bsy_b = side == 0 ? bsly_b : bsry_b;
            double &bsz_b = side == 0 ? bslz_b : bsrz_b;
            double &s_b = side == 0 ? slb : srb;
            double *qb = side == 0 ? ql : qr;
            if (dbl) {
                // f = ((o.f + al*u2) - (al - s_)*us) - s_*o.U
                const double al = side == 0 ? alfven_l : alfven_r;
                double &al_b = side == 0 ? all_b : alr_b;
#pragma unroll
                for (int v = 0; v < 5; v++) {
                    const double g = w * fb[v];
                    f_b[v] += g;
                    u2_b[v] += al * g;
                    us_b[v] -= (al - s_) * g;
                    U_b[v] -= s_ * g;
                    al_b += g * (u2[v] - us[v]);
                    s_b += g * (us[v] - U[v]);
                }
                // f6 = ((fby + al*b2y) - (al - s_)*bsy) - s_*q6, f7 alike
                const double g6 = w * fb[6], g7 = w * fb[7];
                fby_b += g6;
                fbz_b += g7;
                b2y_b += al * g6;
                b2z_b += al * g7;
                bsy_b -= (al - s_) * g6;
                bsz_b -= (al - s_) * g7;
                qb[6] -= s_ * g6;
                qb[7] -= s_ * g7;
                al_b += g6 * (b2y - bsy) + g7 * (b2z - bsz);
                s_b += g6 * (bsy - q[6]) + g7 * (bsz - q[7]);
            } else {
                // f = o.f + s_*(us - o.U); f6 = fby + s_*(bsy - q6)
#pragma unroll
                for (int v = 0; v < 5; v++) {
                    const double g = w * fb[v];
                    f_b[v] += g;
                    us_b[v] += s_ * g;
                    U_b[v] -= s_ * g;
                    s_b += g * (us[v] - U[v]);
                }
                const double g6 = w * fb[6], g7 = w * fb[7];
                fby_b += g6;
                fbz_b += g7;
                bsy_b += s_ * g6;
                bsz_b += s_ * g7;
                qb[6] -= s_ * g6;
                qb[7] -= s_ * g7;
                s_b += g6 * (bsy - q[6]) + g7 * (bsz - q[7]);
            }
        }

        // ---- backwards through the intermediates ----
        double smb = 0.0, prtb = 0.0, dslb = 0.0, dsrb = 0.0, enlb = 0.0, enrb = 0.0;
        double vslyb = 0.0, vslzb = 0.0, vsryb = 0.0, vsrzb = 0.0, slvxlb = 0.0, srvxrb = 0.0;
        double vb_lb = 0.0, vb_rb = 0.0, vb_slb = 0.0, vb_srb = 0.0, vb2b = 0.0;
        double v2yb = 0.0, v2zb = 0.0, dlsb = 0.0, drsb = 0.0;
        // u2 = [ds, ds*sm, ds*v2y, ds*v2z, us4 -/+ (b_sig*sq)*(vbs - vb2)]
        {
            dslb += u2l_b[0] + u2l_b[1] * sm + u2l_b[2] * v2y + u2l_b[3] * v2z;
            dsrb += u2r_b[0] + u2r_b[1] * sm + u2r_b[2] * v2y + u2r_b[3] * v2z;
            smb += u2l_b[1] * dsl + u2r_b[1] * dsr;
            v2yb += u2l_b[2] * dsl + u2r_b[2] * dsr;
            v2zb += u2l_b[3] * dsl + u2r_b[3] * dsr;
            usl_b[4] += u2l_b[4];
            usr_b[4] += u2r_b[4];
            dlsb -= u2l_b[4] * b_sig * (vb_sl - vb2);
            vb_slb -= u2l_b[4] * (b_sig * dn_lsqt);
            vb2b += u2l_b[4] * (b_sig * dn_lsqt);
            drsb += u2r_b[4] * b_sig * (vb_sr - vb2);
            vb_srb += u2r_b[4] * (b_sig * dn_rsqt);
            vb2b -= u2r_b[4] * (b_sig * dn_rsqt);
        }
        // us = [ds, ds*sm, ds*vsy, ds*vsz, n/ssm]
        double slsmb = 0.0, srsmb = 0.0;
        const double rsl = 1.0 / slsm, rsr = 1.0 / srsm;
        {
            dslb += usl_b[0] + usl_b[1] * sm + usl_b[2] * vsly + usl_b[3] * vslz;
            dsrb += usr_b[0] + usr_b[1] * sm + usr_b[2] * vsry + usr_b[3] * vsrz;
            smb += usl_b[1] * dsl + usr_b[1] * dsr;
            vslyb += usl_b[2] * dsl;
            vslzb += usl_b[3] * dsl;
            vsryb += usr_b[2] * dsr;
            vsrzb += usr_b[3] * dsr;
            const double nlb = usl_b[4] * rsl, nrb = usr_b[4] * rsr;
            slsmb -= nlb * usl4;
            srsmb -= nrb * usr4;
            // n = (((svx*en) - (pr*q1)) + (prt*sm)) + bx*(vb - vbs)
            slvxlb += nlb * enl;
            enlb += nlb * slvxl;
            prlb -= nlb * l[1];
            ql[1] -= nlb * prl;
            prtb += nlb * sm + nrb * sm;
            smb += (nlb + nrb) * prt;
            ql[5] += nlb * (vb_l - vb_sl);
            vb_lb += nlb * bxl;
            vb_slb -= nlb * bxl;
            srvxrb += nrb * enr;
            enrb += nrb * srvxr;
            prrb -= nrb * r[1];
            qr[1] -= nrb * prr;
            qr[5] += nrb * (vb_r - vb_sr);
            vb_rb += nrb * bxr;
            vb_srb -= nrb * bxr;
        }
        // the sides' physical fluxes, conserved states and energies
        hlld_side_vjp(l, b2l, prl, gam, fl_b, fbyl_b, fbzl_b, Ul_b, enlb, ql, b2lb, prlb);
        hlld_side_vjp(r, b2r, prr, gam, fr_b, fbyr_b, fbzr_b, Ur_b, enrb, qr, b2rb, prrb);
        // (the side VJP covers fby/fbz of the side's own formula; zero them so
        // the shared fby terms below are not counted twice)
        fbyl_b = fbzl_b = fbyr_b = fbzr_b = 0.0;
        // vb_s = (sm*bx + vsy*bsy) + vsz*bsz; vb = (q1*bx + q2*by) + q3*bz
        smb += vb_slb * bxl + vb_srb * bxr;
        ql[5] += vb_slb * sm;
        qr[5] += vb_srb * sm;
        vslyb += vb_slb * bsly;
        bsly_b += vb_slb * vsly;
        vslzb += vb_slb * bslz;
        bslz_b += vb_slb * vslz;
        vsryb += vb_srb * bsry;
        bsry_b += vb_srb * vsry;
        vsrzb += vb_srb * bsrz;
        bsrz_b += vb_srb * vsrz;
        ql[1] += vb_lb * bxl;
        ql[5] += vb_lb * l[1];
        ql[2] += vb_lb * byl;
        ql[6] += vb_lb * l[2];
        ql[3] += vb_lb * bzl;
        ql[7] += vb_lb * l[3];
        qr[1] += vb_rb * bxr;
        qr[5] += vb_rb * r[1];
        qr[2] += vb_rb * byr;
        qr[6] += vb_rb * r[2];
        qr[3] += vb_rb * bzr;
        qr[7] += vb_rb * r[3];
        // double-star state
        if (dbl) {
            // vb2 = (sm*b2x + v2y*b2y) + v2z*b2z
            smb += vb2b * b2x;
            const double b2xb = vb2b * sm;
            v2yb += vb2b * b2y;
            b2y_b += vb2b * v2y;
            v2zb += vb2b * b2z;
            b2z_b += vb2b * v2z;
            ql[5] += 0.5 * b2xb;
            qr[5] += 0.5 * b2xb;
            // b2y = ((sqL*bsry + sqR*bsly) + (b_sig*mul)*(vsry - vsly)) / add
            const double radd = 1.0 / add;
            double addb = -(b2y_b * b2y + b2z_b * b2z + v2yb * v2y + v2zb * v2z) * radd;
            const double ny = b2y_b * radd, nz = b2z_b * radd;
            dlsb += ny * bsry + nz * bsrz;
            bsry_b += ny * dn_lsqt;
            bsrz_b += nz * dn_lsqt;
            drsb += ny * bsly + nz * bslz;
            bsly_b += ny * dn_rsqt;
            bslz_b += nz * dn_rsqt;
            const double mulb = b_sig * (ny * (vsry - vsly) + nz * (vsrz - vslz));
            vsryb += ny * (b_sig * mul);
            vslyb -= ny * (b_sig * mul);
            vsrzb += nz * (b_sig * mul);
            vslzb -= nz * (b_sig * mul);
            // v2y = ((sqL*vsly + sqR*vsry) + b_sig*(bsry - bsly)) / add
            const double my = v2yb * radd, mz = v2zb * radd;
            dlsb += my * vsly + mz * vslz;
            vslyb += my * dn_lsqt;
            vslzb += mz * dn_lsqt;
            drsb += my * vsry + mz * vsrz;
            vsryb += my * dn_rsqt;
            vsrzb += mz * dn_rsqt;
            bsry_b += my * b_sig;
            bsly_b -= my * b_sig;
            bsrz_b += mz * b_sig;
            bslz_b -= mz * b_sig;
            // add = sqL + sqR, mul = sqL*sqR
            dlsb += addb + mulb * dn_rsqt;
            drsb += addb + mulb * dn_lsqt;
        }
        if (bxnz) {
            // alfven_l = sm - |bxl|/sqL, alfven_r = sm + |bxr|/sqR
            smb += all_b + alr_b;
            const double sgl = (bxl > 0.0) ? 1.0 : -1.0;
            const double sgr = (bxr > 0.0) ? 1.0 : ((bxr < 0.0) ? -1.0 : 0.0);
            const double rql = 1.0 / dn_lsqt, rqr = 1.0 / dn_rsqt;
            ql[5] -= all_b * sgl * rql;
            dlsb += all_b * fabs(bxl) * (rql * rql);
            qr[5] += alr_b * sgr * rqr;
            drsb -= alr_b * fabs(bxr) * (rqr * rqr);
            // sq = sqrt(ds)
            dslb += dlsb * (0.5 * rql);
            dsrb += drsb * (0.5 * rqr);
        }
        // ds = dn / ssm
        double dnlb = dslb * rsl, dnrb = dsrb * rsr;
        slsmb -= dnlb * dsl;
        srsmb -= dnrb * dsr;
        // vs = q2 + (bx/dn)*(by - bsy)
        {
            const double rdl = 1.0 / dn_l, rdr = 1.0 / dn_r;
            const double ol = bxl * rdl, orr = bxr * rdr;
            ql[2] += vslyb;
            ql[3] += vslzb;
            qr[2] += vsryb;
            qr[3] += vsrzb;
            const double olb = vslyb * (byl - bsly) + vslzb * (bzl - bslz);
            const double orb = vsryb * (byr - bsry) + vsrzb * (bzr - bsrz);
            ql[6] += vslyb * ol;
            bsly_b -= vslyb * ol;
            ql[7] += vslzb * ol;
            bslz_b -= vslzb * ol;
            qr[6] += vsryb * orr;
            bsry_b -= vsryb * orr;
            qr[7] += vsrzb * orr;
            bsrz_b -= vsrzb * orr;
            ql[5] += olb * rdl;
            dnlb -= olb * ol * rdl;
            qr[5] += orb * rdr;
            dnrb -= orb * orr * rdr;
        }
        // star transverse B
        double hyb = 0.0, hzb = 0.0, c1b = 0.0, c2b = 0.0, klb = 0.0, krb = 0.0, b_lrb = 0.0;
        if (okl) {
            // bsy = (by*k)/c1
            const double rc = 1.0 / c1;
            ql[6] += bsly_b * kl * rc;
            ql[7] += bslz_b * kl * rc;
            klb += (bsly_b * byl + bslz_b * bzl) * rc;
            c1b -= (bsly_b * bsly + bslz_b * bslz) * rc;
        } else {
            hyb += bsly_b;
            hzb += bslz_b;
        }
        if (okr) {
            const double rc = 1.0 / c2;
            qr[6] += bsry_b * kr * rc;
            qr[7] += bsrz_b * kr * rc;
            krb += (bsry_b * byr + bsrz_b * bzr) * rc;
            c2b -= (bsry_b * bsry + bsrz_b * bsrz) * rc;
        } else {
            hyb += bsry_b;
            hzb += bsrz_b;
        }
        // k = dn*svx - b_lr
        dnlb += klb * slvxl;
        slvxlb += klb * dn_l;
        dnrb += krb * srvxr;
        srvxrb += krb * dn_r;
        b_lrb -= klb + krb;
        // c1 = dn_l*slsm - b_lr, c2 = dn_r*srsm - b_lr
        dnlb += c1b * slsm;
        slsmb += c1b * dn_l;
        dnrb += c2b * srsm;
        srsmb += c2b * dn_r;
        b_lrb -= c1b + c2b;
        double srmslb = 0.0;
        if (hll) {
            // h = ((sr*byr - sl*byl) - (fbyr - fbyl)) / srmsl
            const double rh = 1.0 / srmsl;
            const double ny = hyb * rh, nz = hzb * rh;
            srmslb -= (hyb * hy + hzb * hz) * rh;
            srb += ny * byr + nz * bzr;
            qr[6] += ny * sr;
            qr[7] += nz * sr;
            slb -= ny * byl + nz * bzl;
            ql[6] -= ny * sl;
            ql[7] -= nz * sl;
            fbyr_b -= ny;
            fbyl_b += ny;
            fbzr_b -= nz;
            fbzl_b += nz;
        }
        // fby = by*q1 - bx*q2 (the shared ones used by h)
        ql[6] += fbyl_b * l[1];
        ql[1] += fbyl_b * byl + fbzl_b * bzl;
        ql[5] -= fbyl_b * l[2] + fbzl_b * l[3];
        ql[2] -= fbyl_b * bxl;
        ql[7] += fbzl_b * l[1];
        ql[3] -= fbzl_b * bxl;
        qr[6] += fbyr_b * r[1];
        qr[1] += fbyr_b * byr + fbzr_b * bzr;
        qr[5] -= fbyr_b * r[2] + fbzr_b * r[3];
        qr[2] -= fbyr_b * bxr;
        qr[7] += fbzr_b * r[1];
        qr[3] -= fbzr_b * bxr;
        // b_lr = bxl*bxr
        ql[5] += b_lrb * bxr;
        qr[5] += b_lrb * bxl;
        // prt = 0.5*((prl + dn_l*(sm - l1)) + (prr + dn_r*(sm - r1)))
        {
            const double a = 0.5 * prtb;
            prlb += a;
            prrb += a;
            dnlb += a * (sm - l[1]);
            dnrb += a * (sm - r[1]);
            smb += a * (dn_l + dn_r);
            ql[1] -= a * dn_l;
            qr[1] -= a * dn_r;
        }
        // dn = q0*svx
        ql[0] += dnlb * slvxl;
        slvxlb += dnlb * l[0];
        qr[0] += dnrb * srvxr;
        srvxrb += dnrb * r[0];
        // srmsl = sr - sl; svx = s - q1; ssm = s - sm
        srb += srmslb + srvxrb + srsmb;
        slb += -srmslb + slvxlb + slsmb;
        ql[1] -= slvxlb;
        qr[1] -= srvxrb;
        smb -= slsmb + srsmb;
        // sm
        if (degen) {
            slb += 0.5 * smb;
            srb += 0.5 * smb;
        } else {
            const double nb = smb / (auxr - auxl), denb = -(nb * sm);
            const double auxrb = nb * r[1] + denb, auxlb = -(nb * l[1]) - denb;
            qr[1] += nb * auxr;
            ql[1] -= nb * auxl;
            prrb -= nb;
            prlb += nb;
            srb += auxrb * r[0];
            qr[1] -= auxrb * r[0];
            qr[0] += auxrb * (sr - r[1]);
            slb += auxlb * l[0];
            ql[1] -= auxlb * l[0];
            ql[0] += auxlb * (sl - l[1]);
        }
    }
    // wave speeds: sl = min(l1, r1) - c, sr = max(l1, r1) + c, c = sqrt(0.5*max(cl, cr))
    {
        double wl, wr;
        dmin(l[1], r[1], wl, wr);
        ql[1] += slb * wl;
        qr[1] += slb * wr;
        dmax(l[1], r[1], wl, wr);
        ql[1] += srb * wl;
        qr[1] += srb * wr;
        const double cb = srb - slb;
        const double cmb = 0.5 * cb / (2.0 * c);
        dmax(cl, cr, wl, wr);
        const double clb = cmb * wl, crb = cmb * wr;
        // c_ = (t + sqrt(t*t - (4*gp)*bx^2)) / q0
        const double *qs_[2] = {l, r};
        double *qb_[2] = {ql, qr};
        const double cb_[2] = {clb, crb}, c_[2] = {cl, cr}, t_[2] = {tl, tr}, sq_[2] = {sql, sqr};
        const double gp_[2] = {gpl, gpr};
        double *b2b_[2] = {&b2lb, &b2rb};
        for (int side = 0; side < 2; side++) {
            const double *q = qs_[side];
            double *qb = qb_[side];
            if (cb_[side] == 0.0) continue;
            const double nb = cb_[side] / q[0];
            qb[0] -= nb * c_[side];
            const double sqb = nb;
            double tb = nb;
            // sq = sqrt(t^2 - 4 gp bx^2)
            const double xb = sqb / (2.0 * sq_[side]);
            tb += xb * 2.0 * t_[side];
            const double gpb = -(xb * 4.0 * (q[5] * q[5])) + tb;
            qb[5] -= xb * (4.0 * gp_[side]) * 2.0 * q[5];
            *b2b_[side] += tb;
            qb[4] += gpb * gam;
        }
    }
    // pr = p + 0.5*b2; b2 = |B|^2
    ql[4] += prlb;
    qr[4] += prrb;
    b2lb += 0.5 * prlb;
    b2rb += 0.5 * prrb;
    ql[5] += 2.0 * l[5] * b2lb;
    ql[6] += 2.0 * l[6] * b2lb;
    ql[7] += 2.0 * l[7] * b2lb;
    qr[5] += 2.0 * r[5] * b2rb;
    qr[6] += 2.0 * r[6] * b2rb;
    qr[7] += 2.0 * r[7] * b2rb;
#pragma unroll
    for (int v = 0; v < 8; v++) {
        lb[v] += ql[v];
        rb[v] += qr[v];
    }
    return chb;
}

// ---- utoq (MHD) ------------------------------------------------------------------
// q = prim<MHD>(u): HD part, then p -= (0.5*gam_1)*|B|^2, B and psi copied.
PK_HD void prim_vjp_mhd(const double (&u)[9], double gam_1, const double (&qb)[9], double (&ub)[9])
{
    const double a[5] = {u[0], u[1], u[2], u[3], u[4]};
    const double b[5] = {qb[0], qb[1], qb[2], qb[3], qb[4]};
    double o[5] = {0, 0, 0, 0, 0};
    prim_vjp_hd(a, gam_1, b, o);
#pragma unroll
    for (int v = 0; v < 5; v++) ub[v] += o[v];
    const double b2b = -((0.5 * gam_1) * qb[4]);
    ub[5] += qb[5] + 2.0 * u[5] * b2b;
    ub[6] += qb[6] + 2.0 * u[6] * b2b;
    ub[7] += qb[7] + 2.0 * u[7] * b2b;
    ub[8] += qb[8];
}

// floors<MHD>: as the HD floors with e = (E - kin) - mag, E = (e + kin) + mag.
PK_HD void floors_vjp_mhd(const double (&u)[9], double smalld, double smallei, int use_smalld, int use_smallei,
                          double (&ub)[9])
{
    double d = u[0];
    const bool dfl = use_smalld && !(d >= smalld);
    if (use_smalld) d = fmx(d, smalld);
    if (!use_smallei) {
        if (dfl) ub[0] = 0.0;
        return;
    }
    const double ek = (0.5 * ((u[1] * u[1] + u[2] * u[2]) + u[3] * u[3])) / d;
    const double mag = 0.5 * ((u[5] * u[5] + u[6] * u[6]) + u[7] * u[7]);
    const double e = (u[4] - ek) - mag;
    const bool efl = !(e >= smallei);
    const double eb = efl ? 0.0 : ub[4];
    const double kb = ub[4] - eb;  // adjoint reaching kin and mag (they cancel when e is not floored)
    double db = ub[0];
    ub[4] = eb;
    const double sb = kb / d;
    db -= sb * ek;
    ub[1] += sb * u[1];
    ub[2] += sb * u[2];
    ub[3] += sb * u[3];
    ub[5] += kb * u[5];
    ub[6] += kb * u[6];
    ub[7] += kb * u[7];
    ub[0] = dfl ? 0.0 : db;
}

// timestep_fluid at one cell, one direction (ion_cs, initionized.F90:137-141):
// t = dl / max(|m_q/rho| + cs, small), cs = sqrt(|(2 pmag + gam p)/rho|),
// p = ((E - ek)*(gam-1) + (2-gam)*pmag) - pmag.
PK_HD void dt_vjp_mhd(const double (&u)[9], int qdir, double dl, double gam, double gam_1, double tb,
                      double (&ub)[9])
{
    const double small = 1.17549435082228750797e-38;
    const double d = u[0];
    const double vq = (d > 0.0) ? fabs(u[1 + qdir] / d) : 0.0;
    const double ek = (0.5 * ((u[1] * u[1] + u[2] * u[2]) + u[3] * u[3])) / d;
    const double pmag = 0.5 * ((u[5] * u[5] + u[6] * u[6]) + u[7] * u[7]);
    const double ps = (u[4] - ek) * gam_1 + (2.0 - gam) * pmag;
    const double pr = ps - pmag;
    const double x = (2.0 * pmag + gam * pr) / d;
    const double cs = sqrt(fabs(x));
    const double c = fmx(vq + cs, small);
    if (!(vq + cs > small)) return;
    const double cb = -(tb * dl) / (c * c);
    if (d > 0.0) {
        const double vv = u[1 + qdir] / d;
        const double sg = (vv > 0.0) ? 1.0 : ((vv < 0.0) ? -1.0 : 0.0);
        ub[1 + qdir] += cb * sg / d;
        ub[0] -= cb * sg * vv / d;
    }
    if (cs > 0.0) {
        const double sx = (x > 0.0) ? 1.0 : ((x < 0.0) ? -1.0 : 0.0);
        const double xb = cb * sx / (2.0 * cs);
        ub[0] -= xb * x / d;
        const double nb = xb / d;
        double pmb = 2.0 * nb;
        const double prb = nb * gam;
        // pr = ps - pmag; ps = (E - ek)*gam_1 + (2-gam)*pmag
        pmb += -prb + prb * (2.0 - gam);
        ub[4] += prb * gam_1;
        const double ekb = -(prb * gam_1);
        const double sb = ekb / d;
        ub[0] -= ekb * ek / d;
        ub[1] += sb * u[1];
        ub[2] += sb * u[2];
        ub[3] += sb * u[3];
        ub[5] += pmb * u[5];
        ub[6] += pmb * u[6];
        ub[7] += pmb * u[7];
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
    db -= sb * ek;
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
