"""Reverse mode of the hydro sweep's per-cell arithmetic (pk_adjoint.hpp,
built for the host) against central finite differences of the forward
functions (pk_math.hpp, same build).

Each case draws generic states away from branch switches (so a central
difference of step h stays on one branch) and checks <fb, J e_j> for every
input j: a dropped or mis-signed term shows as an O(1) relative error, the
FD truncation error is ~1e-9.  The mirror-symmetric HLLC state (contact
speed exactly 0) checks the averaged derivative of the two-sided branch.
"""
import ctypes

import numpy as np
import pytest

P = ctypes.POINTER(ctypes.c_double)
GAM = 5.0 / 3.0


@pytest.fixture(scope="module")
def hm():
    from hostmath import loader
    return loader.load()


def dp(a):
    return a.ctypes.data_as(P)


def fd_vjp(f, x, fb, rel=1e-6):
    """sum_k fb_k * d f_k / d x_j by central differences."""
    g = np.zeros_like(x)
    for j in range(len(x)):
        h = rel * max(1.0, abs(x[j]))
        xp, xm = x.copy(), x.copy()
        xp[j] += h
        xm[j] -= h
        g[j] = np.dot(fb, (f(xp) - f(xm)) / (2.0 * h))
    return g


def close(a, b, tol=1e-7):
    scale = max(1.0, np.max(np.abs(b)))
    return np.max(np.abs(a - b)) <= tol * scale


def prim(hm, u):
    q = np.zeros(5)
    uu = np.asfortranarray(u.reshape(1, 5))
    qq = np.zeros((1, 5), order="F")
    hm.hm_prim(1, 0, GAM - 1.0, dp(uu), dp(qq))
    q[:] = qq[0]
    return q


def test_prim_vjp(hm):
    rng = np.random.default_rng(1)
    for _ in range(50):
        d = rng.uniform(0.2, 3.0)
        v = rng.normal(0, 1, 3)
        p = rng.uniform(0.1, 2.0)
        u = np.array([d, d * v[0], d * v[1], d * v[2], p / (GAM - 1) + 0.5 * d * v @ v])
        qb = rng.normal(0, 1, 5)
        ub = np.zeros(5)
        hm.hm_prim_vjp_hd(GAM - 1.0, dp(u), dp(qb), dp(ub))
        assert close(ub, fd_vjp(lambda x: prim(hm, x), u, qb))


@pytest.mark.parametrize("lim", [0, 1, 2, 3])
def test_limiter_vjp(hm, lim):
    rng = np.random.default_rng(2 + lim)
    n = 0
    while n < 200:
        dl, dr = rng.normal(0, 1, 2)
        a, b = abs(dl), abs(dr)
        # stay away from the kinks: sign changes, |dl| = |dr|, 2a = 2b, 2a or 2b = 0.5|dl+dr|,
        # a = 2|dr|, 2|dl| = b
        kinks = [a, b, a - b, 2 * a - 0.5 * abs(dl + dr), 2 * b - 0.5 * abs(dl + dr), a - 2 * b, 2 * a - b]
        if min(abs(k) for k in kinks) < 1e-3:
            continue
        n += 1
        out = np.zeros(2)
        hm.hm_limit_vjp(lim, dl, dr, 1.0, dp(out))
        f = lambda x: np.array([hm.hm_limit(lim, x[0], x[1])])
        assert close(out, fd_vjp(f, np.array([dl, dr]), np.array([1.0]))), (lim, dl, dr, out)


def hllc(hm, l, r):
    f = np.zeros((1, 5), order="F")
    hm.hm_hllc(1, dp(np.asfortranarray(l.reshape(1, 5))), dp(np.asfortranarray(r.reshape(1, 5))), GAM, dp(f))
    return f[0].copy()


def speeds(l, r):
    c = np.sqrt(max(GAM * l[4] / l[0], GAM * r[4] / r[0]))
    sl = min(l[1], r[1]) - c
    sr = max(l[1], r[1]) + c
    auxr = (sr - r[1]) * r[0]
    auxl = (sl - l[1]) * l[0]
    sm = ((auxr * r[1] - r[4]) - (auxl * l[1] - l[4])) / (auxr - auxl)
    return sl, sr, sm


def random_face(rng, shift=0.0):
    return np.array([rng.uniform(0.3, 2.0), rng.normal(shift, 0.8), rng.normal(0, 0.8), rng.normal(0, 0.8),
                     rng.uniform(0.2, 2.0)])


@pytest.mark.parametrize("case", ["subsonic", "supersonic_right", "supersonic_left"])
def test_hllc_vjp(hm, case):
    rng = np.random.default_rng({"subsonic": 5, "supersonic_right": 6, "supersonic_left": 7}[case])
    shift = {"subsonic": 0.0, "supersonic_right": 4.0, "supersonic_left": -4.0}[case]
    n = 0
    while n < 100:
        l, r = random_face(rng, shift), random_face(rng, shift)
        sl, sr, sm = speeds(l, r)
        if case == "subsonic" and not (sl < -1e-3 and sr > 1e-3 and abs(sm) > 1e-3):
            continue
        if case == "supersonic_right" and not sl > 1e-3:
            continue
        if case == "supersonic_left" and not sr < -1e-3:
            continue
        if abs(l[1] - r[1]) < 1e-3 or abs(GAM * l[4] / l[0] - GAM * r[4] / r[0]) < 1e-3:
            continue
        n += 1
        fb = rng.normal(0, 1, 5)
        lb, rb = np.zeros(5), np.zeros(5)
        hm.hm_hllc_vjp(dp(l), dp(r), GAM, dp(fb), dp(lb), dp(rb))
        x = np.concatenate([l, r])
        g = fd_vjp(lambda y: hllc(hm, y[:5], y[5:]), x, fb)
        assert close(np.concatenate([lb, rb]), g, 1e-6), (case, l, r, np.concatenate([lb, rb]) - g)


def test_hllc_vjp_mirror_state_averages_the_two_sides(hm):
    """l and r mirror images: the contact speed is exactly 0 and the flux
    averages both star states; the reverse mode weights both sides by 1/2,
    which is the central difference across sm = 0."""
    rng = np.random.default_rng(8)
    for _ in range(30):
        l = random_face(rng)
        r = l.copy()
        r[1] = -l[1]
        sl, sr, sm = speeds(l, r)
        assert sm == 0.0
        fb = rng.normal(0, 1, 5)
        lb, rb = np.zeros(5), np.zeros(5)
        hm.hm_hllc_vjp(dp(l), dp(r), GAM, dp(fb), dp(lb), dp(rb))
        # across the kink at sm = 0 a central difference errs by O(h), not O(h^2)
        g = fd_vjp(lambda y: hllc(hm, y[:5], y[5:]), np.concatenate([l, r]), fb, rel=1e-7)
        assert close(np.concatenate([lb, rb]), g, 1e-6)


@pytest.mark.parametrize("case", ["none", "density", "energy", "both"])
def test_floors_vjp(hm, case):
    rng = np.random.default_rng(11)
    smalld, smallei = 0.5, 0.3
    for _ in range(40):
        d = rng.uniform(0.7, 2.0) if case in ("none", "energy") else rng.uniform(0.1, 0.4)
        v = rng.normal(0, 1, 3)
        e = rng.uniform(0.5, 2.0) if case in ("none", "density") else rng.uniform(-1.0, 0.2)
        u = np.array([d, d * v[0], d * v[1], d * v[2], e + 0.5 * d * v @ v])
        if case == "both":
            # E - ek with the floored density must be below smallei too
            u[4] = 0.5 * (u[1:4] @ u[1:4]) / smalld + rng.uniform(-1.0, 0.2)

        def f(x):
            y = x.copy()
            hm.hm_floors_hd(smalld, smallei, 1, 1, dp(y))
            return y

        fb = rng.normal(0, 1, 5)
        ub = fb.copy()
        hm.hm_floors_vjp_hd(smalld, smallei, 1, 1, dp(u), dp(ub))
        assert close(ub, fd_vjp(f, u, fb))


@pytest.mark.parametrize("qdir", [0, 1, 2])
def test_timestep_vjp(hm, qdir):
    rng = np.random.default_rng(20 + qdir)
    dl = 1.0 / 32

    def t(x):
        d = x[0]
        vq = abs(x[1 + qdir] / d)
        ek = 0.5 * (x[1:4] @ x[1:4]) / d
        pr = (x[4] - ek) * (GAM - 1.0)
        cs = np.sqrt(abs(GAM * pr / d))
        return np.array([dl / max(vq + cs, 1.17549435082228750797e-38)])

    for _ in range(40):
        d = rng.uniform(0.3, 2.0)
        v = rng.normal(0, 1, 3)
        p = rng.uniform(0.2, 2.0)
        u = np.array([d, d * v[0], d * v[1], d * v[2], p / (GAM - 1) + 0.5 * d * v @ v])
        ub = np.zeros(5)
        hm.hm_dt_vjp_hd(dp(u), qdir, dl, GAM, 1.0, dp(ub))
        assert close(ub, fd_vjp(t, u, np.array([1.0])), 1e-6)
