"""FP64 PyTorch transcription of the hydro Riemann step, for autograd.

Test infrastructure: the reference gradient for the hand-written adjoint in
piernik_amd.  Same scheme as the oracle (and the reference it restates):
  utoq                interpolations.F90:69-106
  limiters, PLM       fluxlimiters.F90:109-216, interpolations.F90:170-204
  HLLC                hlld.F90:615-841 (riemann_hlld_u)
  RK2 update          solve_cg_riemann.F90:366, rk_coef = [1/2, 1]
  floors              sources.F90:299-433
  timestep_fluid      timestep.F90:431-497 with neu_cs; caps :200-208
on one periodic block (interior only; torch.roll is the periodic boundary),
operations in the same order so the forward values match the oracle's.
Branches are torch.where selections, so autograd differentiates the branch
the forward took.
"""
import math

import torch

LIM = {"vanleer": 0, "minmod": 1, "moncen": 2, "superbee": 3}
SMALL = 1.17549435082228750797e-38
PERM = {0: [0, 1, 2, 3, 4], 1: [0, 2, 1, 3, 4], 2: [0, 3, 2, 1, 4]}


def spacing(x):
    """Fortran SPACING for real(8): 2**(exponent(x) - 53), at least 2**-1022."""
    _, e = torch.frexp(x)
    s = torch.ldexp(torch.ones_like(x), (e - 53).to(torch.int32))
    return torch.clamp(s, min=2.2250738585072014e-308)


def prim(u, gam_1):
    d = u[0]
    ek = (0.5 * ((u[1] * u[1] + u[2] * u[2]) + u[3] * u[3])) / d
    p = gam_1 * (u[4] - ek)
    return torch.stack([d, u[1] / d, u[2] / d, u[3] / d, p])


def limiter(lim, dl, dr):
    if lim == 0:
        dcen = dl * dr
        safe = torch.where(dcen > 0.0, dl + dr, torch.ones_like(dl))
        return torch.where(dcen > 0.0, (2.0 * dcen) / safe, torch.zeros_like(dl))
    s = torch.copysign(torch.ones_like(dl), dl) + torch.copysign(torch.ones_like(dr), dr)
    a, b = torch.abs(dl), torch.abs(dr)
    if lim == 1:
        return (s * torch.minimum(a, b)) * 0.5
    if lim == 2:
        return (s * torch.minimum(torch.minimum(2.0 * a, 2.0 * b), 0.5 * torch.abs(dl + dr))) * 0.5
    m = torch.where(a > b, torch.minimum(a, torch.abs(2.0 * dr)), torch.minimum(torch.abs(2.0 * dl), b))
    return (s * 0.5) * m


def hllc(l, r, gam):
    prl, prr = l[4], r[4]
    enl = (l[4] / (gam - 1.0)) + (0.5 * l[0]) * ((l[1] * l[1] + l[2] * l[2]) + l[3] * l[3])
    enr = (r[4] / (gam - 1.0)) + (0.5 * r[0]) * ((r[1] * r[1] + r[2] * r[2]) + r[3] * r[3])
    c = torch.sqrt(torch.maximum((gam * l[4]) / l[0], (gam * r[4]) / r[0]))
    sl = torch.minimum(l[1], r[1]) - c
    sr = torch.maximum(l[1], r[1]) + c
    fl = torch.stack([l[0] * l[1], l[0] * (l[1] * l[1]) + prl, (l[0] * l[2]) * l[1], (l[0] * l[3]) * l[1],
                      (enl + prl) * l[1]])
    fr = torch.stack([r[0] * r[1], r[0] * (r[1] * r[1]) + prr, (r[0] * r[2]) * r[1], (r[0] * r[3]) * r[1],
                      (enr + prr) * r[1]])
    auxr = (sr - r[1]) * r[0]
    auxl = (sl - l[1]) * l[0]
    degen = torch.abs(auxr - auxl) < spacing(auxr + auxl)
    den = torch.where(degen, torch.ones_like(auxr), auxr - auxl)
    sm = torch.where(degen, (sl + sr) / 2.0, ((auxr * r[1] - prr) - (auxl * l[1] - prl)) / den)
    slvxl, srvxr = sl - l[1], sr - r[1]
    dn_l, dn_r = l[0] * slvxl, r[0] * srvxr
    prt = 0.5 * ((prl + dn_l * (sm - l[1])) + (prr + dn_r * (sm - r[1])))

    def side(q, s_, svx, en, pr, dn, f0):
        ssm = s_ - sm
        ssm = torch.where(ssm == 0.0, torch.ones_like(ssm), ssm)  # unused where it is 0
        d = dn / ssm
        us = torch.stack([d, d * sm, d * q[2], d * q[3], (((svx * en) - (pr * q[1])) + (prt * sm)) / ssm])
        U = torch.stack([q[0], q[0] * q[1], q[0] * q[2], q[0] * q[3], en])
        return f0, s_, us - U

    f0l, s_l, dl_ = side(l, sl, slvxl, enl, prl, dn_l, fl)
    f0r, s_r, dr_ = side(r, sr, srvxr, enr, prr, dn_r, fr)
    fL = f0l + s_l * dl_
    fR = f0r + s_r * dr_
    fB = 0.5 * (((fl + sl * dl_) + fr) + sr * dr_)
    star = torch.where(sm > 0.0, fL, torch.where(sm < 0.0, fR, fB))
    return torch.where(sl >= 0.0, fl, torch.where(sr <= 0.0, fr, star))


def floors(u, smalld, smallei):
    d = torch.maximum(u[0], torch.full_like(u[0], smalld))
    kin = (0.5 * ((u[1] * u[1] + u[2] * u[2]) + u[3] * u[3])) / d
    ei = torch.maximum(u[4] - kin, torch.full_like(u[0], smallei))
    return torch.stack([d, u[1], u[2], u[3], ei + kin])


def stage(w, w0, coef, axis, lim, gam, smalld, smallei):
    """One RK stage along the pencil axis (array axis 1 + axis), periodic."""
    ax = 1 + axis
    q = prim(w, gam - 1.0)
    dl = q - torch.roll(q, 1, ax)
    dr = torch.roll(q, -1, ax) - q
    dq = limiter(lim, dl, dr)
    h = 0.5 * dq
    fp, fm = q + h, q - h
    F = hllc(fp, torch.roll(fm, -1, ax), gam)
    out = (w0 + coef * (torch.roll(F, 1, ax) - F)) + 0.0
    return floors(out, smalld, smallei)


def sweep(u, axis, dt, dl, lim, gam, smalld, smallei):
    """Both RK2 stages of one directional sweep; u in xyz variable order."""
    p = PERM[axis]
    w = u[p]
    c1 = (0.5 * dt) / dl
    c2 = (1.0 * dt) / dl
    ws = stage(w, w, c1, axis, lim, gam, smalld, smallei)
    wn = stage(ws, w, c2, axis, lim, gam, smalld, smallei)
    return wn[p]


def timestep(u, dl, gam, cfl):
    d = u[0]
    ek = (0.5 * ((u[1] * u[1] + u[2] * u[2]) + u[3] * u[3])) / d
    pr = (u[4] - ek) * (gam - 1.0)
    cs = torch.sqrt(torch.abs((gam * pr) / d))
    mn = None
    for q in range(3):
        c = torch.clamp(torch.abs(u[1 + q] / d) + cs, min=SMALL)
        m = torch.min(dl[q] / c)
        mn = m if mn is None else torch.minimum(mn, m)
    return cfl * mn


def run(u, nsteps, dl, gam=5.0 / 3.0, limiter_name="vanleer", cfl=0.5, smalld=1e-10, smallei=1e-10):
    """nsteps of time_step + fluid_update_full; u: (5, nx, ny, nz) interior."""
    lim = LIM[limiter_name]
    dt_old = None
    for s in range(nsteps):
        dt = timestep(u, dl, gam, cfl)
        if s >= 1:
            dt = torch.minimum(dt, dt_old * 2.0)
        for axis in (0, 1, 2, 2, 1, 0):
            u = sweep(u, axis, dt, dl[axis], lim, gam, smalld, smallei)
        dt_old = dt
    return u


def gaussian_density(n, seed, kmax=6.0, amp=0.1):
    """1 + amp*G(x): G a seeded unit-variance periodic Gaussian field with |k| <= kmax."""
    g = torch.Generator().manual_seed(seed)
    k1 = torch.fft.fftfreq(n, d=1.0 / n, dtype=torch.float64)
    kz = torch.fft.rfftfreq(n, d=1.0 / n, dtype=torch.float64)
    KX, KY, KZ = torch.meshgrid(k1, k1, kz, indexing="ij")
    band = (torch.sqrt(KX * KX + KY * KY + KZ * KZ) <= kmax) & (torch.sqrt(KX * KX + KY * KY + KZ * KZ) > 0)
    ph = torch.complex(torch.randn(KX.shape, generator=g, dtype=torch.float64),
                       torch.randn(KX.shape, generator=g, dtype=torch.float64))
    f = torch.fft.irfftn(ph * band, s=(n, n, n))
    f = f / math.sqrt(float(torch.mean(f * f)))
    return 1.0 + amp * f
