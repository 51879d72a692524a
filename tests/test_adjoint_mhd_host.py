"""Reverse mode of the MHD sweep arithmetic (HLLD + GLM, utoq with B,
floors, ion timestep) built for the host, against central finite
differences of the forward functions.

flux_hlld has many branches; every one is reached: supersonic to either
side, single-star and double-star (Alfven) states on either side, the HLL
fallback for the star B, B_n = 0, and the two averaged sm = 0 cases.
Samples keep a margin from every branch switch so a central difference
stays on one branch; in the sm = 0 cases the derivative across the switch is
the average of the two sides, which is what the reverse mode returns.
"""
import ctypes

import numpy as np
import pytest

P = ctypes.POINTER(ctypes.c_double)
GAM = 5.0 / 3.0
CH = 1.3


@pytest.fixture(scope="module")
def hm():
    from hostmath import loader
    return loader.load()


def dp(a):
    return a.ctypes.data_as(P)


def fd_vjp(f, x, fb, rel=1e-6, skip=()):
    g = np.zeros_like(x)
    for j in range(len(x)):
        if j in skip:
            continue
        h = rel * max(1.0, abs(x[j]))
        xp, xm = x.copy(), x.copy()
        xp[j] += h
        xm[j] -= h
        g[j] = np.dot(fb, (f(xp) - f(xm)) / (2.0 * h))
    return g


def hlld(hm, l, r, ch=CH):
    f = np.zeros(9)
    hm.hm_hlld1(dp(np.ascontiguousarray(l)), dp(np.ascontiguousarray(r)), GAM, ch, dp(f))
    return f


def branch(hm, l, r):
    out = np.zeros(2)
    hm.hm_hlld_branch(dp(l), dp(r), GAM, dp(out))
    return int(out[0]), out[1]


def face(rng, vshift=0.0, bscale=1.0, pscale=1.0):
    return np.array([rng.uniform(0.3, 2.0), rng.normal(vshift, 0.7), rng.normal(0, 0.7), rng.normal(0, 0.7),
                     pscale * rng.uniform(0.2, 2.0), bscale * rng.normal(0, 1.0), bscale * rng.normal(0, 1.0),
                     bscale * rng.normal(0, 1.0), rng.normal(0, 0.3)])


def samples(hm, rng, want, n=25, tries=200000):
    """n states of branch code `want` with margin > 1e-3."""
    got = []
    for _ in range(tries):
        kind = rng.integers(0, 4)
        vshift = rng.choice([0.0, 0.0, 3.0, -3.0])
        bscale = [1.0, 0.2, 2.5, 1.0][kind]
        pscale = [1.0, 1.0, 0.05, 1.0][kind]
        l = face(rng, vshift, bscale, pscale)
        r = face(rng, vshift, bscale, pscale)
        r[5] = l[5]  # B_n is continuous in the sweep
        if want in (7, 8, 9, 23, 24, 25):
            l[5] = r[5] = 0.0
        if want in (6, 7, 22, 23):
            r = l.copy()
            r[1] = -l[1]
            r[8] = l[8]
        code, margin = branch(hm, l, r)
        if code == want and (margin > 1e-3 or want in (6, 7, 22, 23)):
            got.append((l, r))
            if len(got) == n:
                return got
    return got


CODES = {0: "supersonic L", 1: "supersonic R", 2: "single star L", 3: "single star R", 4: "double star L",
         5: "double star R", 6: "Alfven, sm = 0", 7: "B_n = 0, sm = 0", 8: "B_n = 0, L", 9: "B_n = 0, R",
         18: "single star L, HLL star B", 19: "single star R, HLL star B", 20: "double star L, HLL star B",
         21: "double star R, HLL star B"}


@pytest.mark.parametrize("code", sorted(CODES))
def test_hlld_vjp_branch(hm, code):
    rng = np.random.default_rng(100 + code)
    cases = samples(hm, rng, code)
    assert len(cases) >= 5, f"branch {CODES[code]} not reached"
    mirror = code in (6, 7)
    skip = (5, 14) if code in (7, 8, 9) else ()  # B_n = 0: no derivative across B_n
    for l, r in cases:
        fb = rng.normal(0, 1, 9)
        lb, rb, chb = np.zeros(9), np.zeros(9), np.zeros(1)
        hm.hm_hlld_vjp(dp(l), dp(r), GAM, CH, dp(fb), dp(lb), dp(rb), dp(chb))
        x = np.concatenate([l, r])
        g = fd_vjp(lambda y: hlld(hm, y[:9], y[9:]), x, fb, rel=1e-7 if mirror else 1e-6, skip=skip)
        a = np.concatenate([lb, rb])
        for j in skip:
            a[j] = 0.0
        scale = max(1.0, np.max(np.abs(g)))
        assert np.max(np.abs(a - g)) <= 1e-6 * scale, (CODES[code], l, r, a - g)
        h = 1e-6
        gch = np.dot(fb, (hlld(hm, l, r, CH + h) - hlld(hm, l, r, CH - h)) / (2 * h))
        assert abs(chb[0] - gch) <= 1e-6 * max(1.0, abs(gch))


def mhd_state(rng):
    d = rng.uniform(0.3, 2.0)
    v = rng.normal(0, 1, 3)
    b = rng.normal(0, 1, 3)
    p = rng.uniform(0.2, 2.0)
    e = p / (GAM - 1) + 0.5 * d * v @ v + 0.5 * b @ b
    return np.array([d, d * v[0], d * v[1], d * v[2], e, b[0], b[1], b[2], rng.normal(0, 0.3)])


def test_prim_mhd_vjp(hm):
    rng = np.random.default_rng(3)

    def prim(x):
        q = np.zeros((1, 9), order="F")
        hm.hm_prim(1, 1, GAM - 1.0, dp(np.asfortranarray(x.reshape(1, 9))), dp(q))
        return q[0].copy()

    for _ in range(40):
        u = mhd_state(rng)
        qb = rng.normal(0, 1, 9)
        ub = np.zeros(9)
        hm.hm_prim_vjp_mhd(GAM - 1.0, dp(u), dp(qb), dp(ub))
        g = fd_vjp(prim, u, qb)
        assert np.max(np.abs(ub - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))


@pytest.mark.parametrize("case", ["none", "energy"])
def test_floors_mhd_vjp(hm, case):
    rng = np.random.default_rng(7)
    smalld, smallei = 0.2, 0.3
    for _ in range(40):
        u = mhd_state(rng)
        if case == "energy":
            u[4] = 0.5 * (u[1:4] @ u[1:4]) / u[0] + 0.5 * (u[5:8] @ u[5:8]) + rng.uniform(-1.0, 0.2)

        def f(x):
            y = x.copy()
            hm.hm_floors_mhd(smalld, smallei, 1, 1, dp(y))
            return y

        fb = rng.normal(0, 1, 9)
        ub = fb.copy()
        hm.hm_floors_vjp_mhd(smalld, smallei, 1, 1, dp(u), dp(ub))
        g = fd_vjp(f, u, fb)
        assert np.max(np.abs(ub - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))


@pytest.mark.parametrize("qdir", [0, 1, 2])
def test_timestep_mhd_vjp(hm, qdir):
    rng = np.random.default_rng(30 + qdir)
    dl = 1.0 / 64

    def t(x):
        d = x[0]
        vq = abs(x[1 + qdir] / d)
        ek = 0.5 * (x[1:4] @ x[1:4]) / d
        pmag = 0.5 * (x[5:8] @ x[5:8])
        pr = ((x[4] - ek) * (GAM - 1.0) + (2.0 - GAM) * pmag) - pmag
        cs = np.sqrt(abs((2.0 * pmag + GAM * pr) / d))
        return np.array([dl / max(vq + cs, 1.17549435082228750797e-38)])

    for _ in range(40):
        u = mhd_state(rng)
        ub = np.zeros(9)
        hm.hm_dt_vjp_mhd(dp(u), qdir, dl, GAM, 1.0, dp(ub))
        g = fd_vjp(t, u, np.array([1.0]))
        assert np.max(np.abs(ub - g)) <= 1e-6 * max(1.0, np.max(np.abs(g)))
