"""The sweep's per-interface arithmetic (piernik_amd/csrc/pk_math.hpp, built
for the host) against the oracle, bitwise, on random and edge-case states.

A CPU-side filter for transcription mistakes in the device code; the GPU
parity tests check the device build of the same functions."""
import ctypes
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
P = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def hm():
    from hostmath import loader
    return loader.load()


def dp(a):
    return a.ctypes.data_as(P)


def random_prims(rng, m, mhd, kind="mixed"):
    """(m, 5|9) primitive states; a mix of generic, equal-state, B_n = 0,
    supersonic and mirror-symmetric pairs."""
    nv = 9 if mhd else 5
    q = np.empty((m, nv), order="F")
    q[:, 0] = rng.uniform(0.05, 3.0, m)
    q[:, 1:4] = rng.normal(0, 1.0, (m, 3))
    q[:, 4] = rng.uniform(0.01, 3.0, m)
    if mhd:
        q[:, 5:8] = rng.normal(0, 1.0, (m, 3))
        q[:, 8] = rng.normal(0, 0.3, m)
    return q


def pairs(rng, m, mhd):
    ql = random_prims(rng, m, mhd)
    qr = random_prims(rng, m, mhd)
    k = m // 8
    qr[:k] = ql[:k]                                   # equal states
    qr[k:2 * k, 1] = -ql[k:2 * k, 1]                  # mirror pairs (sm = 0 region)
    qr[k:2 * k, 0] = ql[k:2 * k, 0]
    qr[k:2 * k, 4] = ql[k:2 * k, 4]
    qr[k:2 * k, 2:4] = ql[k:2 * k, 2:4]
    ql[2 * k:3 * k, 1] += 8.0                         # supersonic to the right
    qr[2 * k:3 * k, 1] += 8.0
    ql[3 * k:4 * k, 1] -= 8.0                         # supersonic to the left
    qr[3 * k:4 * k, 1] -= 8.0
    if mhd:
        ql[4 * k:5 * k, 5] = 0.0                      # B_n = 0 on the left
        qr[4 * k:5 * k, 5] = 0.0
        qr[5 * k:6 * k, 5] = ql[5 * k:6 * k, 5]       # continuous B_n
        ql[6 * k:7 * k, 6:8] = 0.0                    # vanishing transverse B
        qr[6 * k:7 * k, 6:8] = 0.0
        # exact mirror images (v_n, B_n -> -v_n, -B_n): the contact speed is
        # exactly 0, which selects the two-sided averages of hlld.F90:522,578
        j = slice(7 * k, 7 * k + k // 2)
        qr[j] = ql[j]
        qr[j, 1] = -ql[j, 1]
        qr[j, 5] = -ql[j, 5]
        j2 = slice(7 * k + k // 2, 8 * k)
        ql[j2, 5] = 0.0
        qr[j2] = ql[j2]
        qr[j2, 1] = -ql[j2, 1]
    return ql, qr


def test_spacing(hm, oracle):
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.normal(0, 1, 1000) * 10.0 ** rng.integers(-300, 300, 1000),
                         [0.0, -0.0, 1.0, -1.0, 2.0, 0.75, 5e-324, 2.2250738585072014e-308, 1e-310, 1e300]])
    for x in xs:
        assert hm.hm_spacing(float(x)) == oracle.spacing(float(x)), x


def test_limiters(hm, oracle):
    rng = np.random.default_rng(1)
    for name, k in oracle.LIM.items():
        q = np.cumsum(rng.normal(0, 1, 400))
        q[50:60] = 3.0                                # flat stretch: zero differences
        dq = oracle.limiter(name, q)
        for i in range(1, len(q) - 1):
            assert hm.hm_limit(k, q[i] - q[i - 1], q[i + 1] - q[i]) == dq[i], (name, i)


@pytest.mark.parametrize("mhd", [False, True])
def test_prim(hm, oracle, mhd):
    rng = np.random.default_rng(2)
    m = 500
    nv = 9 if mhd else 5
    u = np.empty((m, nv), order="F")
    u[:, 0] = rng.uniform(0.1, 2, m)
    u[:, 1:4] = rng.normal(0, 1, (m, 3))
    u[:, 4] = rng.uniform(2, 5, m)
    if mhd:
        u[:, 5:9] = rng.normal(0, 1, (m, 4))
    q = np.zeros_like(u, order="F")
    hm.hm_prim(m, int(mhd), 5.0 / 3.0 - 1.0, dp(u), dp(q))
    p = oracle.params(magnetized=mhd)
    qo = oracle.utoq(p, u[:, :5], u[:, 5:9] if mhd else None)
    assert np.array_equal(q[:, :5], qo)


def test_hllc(hm, oracle):
    rng = np.random.default_rng(3)
    m = 20000
    ql, qr = pairs(rng, m, False)
    f = np.zeros((m, 5), order="F")
    hm.hm_hllc(m, dp(ql), dp(qr), 5.0 / 3.0, dp(f))
    fo = oracle.hllc(ql, qr, 5.0 / 3.0)
    bad = np.argwhere(~(f == fo))
    assert bad.size == 0, (bad[:5], f[tuple(bad[0])], fo[tuple(bad[0])])


@pytest.mark.parametrize("gam,ch", [(5.0 / 3.0, 1.7), (1.4, 0.3)])
def test_hlld(hm, oracle, gam, ch):
    rng = np.random.default_rng(4)
    m = 20000
    ql, qr = pairs(rng, m, True)
    f = np.zeros((m, 9), order="F")
    hm.hm_hlld(m, dp(ql), dp(qr), gam, ch, dp(f))
    fo, bfo = oracle.hlld(ql[:, :5], qr[:, :5], ql[:, 5:9], qr[:, 5:9], gam, ch)
    ref = np.concatenate([fo, bfo], axis=1)
    # the mirror pairs do reach the contact-speed-zero averages
    k = m // 8
    assert np.all(f[7 * k:8 * k, 0] == 0.0)
    bad = np.argwhere(~(f == ref))
    assert bad.size == 0, (bad[:5], f[tuple(bad[0])], ref[tuple(bad[0])])


@pytest.mark.parametrize("disallow", [False, True])
@pytest.mark.parametrize("mhd", [False, True])
def test_floors(hm, oracle, mhd, disallow):
    """Floors inside the full pencil stage: one pencil with cells that cross
    smalld and smallei, against the oracle's care_for_positives; the
    negative-internal-energy flag only with disallow_negatives."""
    rng = np.random.default_rng(5)
    m = 64
    nv = 9 if mhd else 5
    u = np.empty((m, nv), order="F")
    u[:, 0] = rng.uniform(-0.1, 1, m)
    u[:, 1:4] = rng.normal(0, 1, (m, 3))
    u[:, 4] = rng.uniform(-0.5, 3, m)
    if mhd:
        u[:, 5:9] = rng.normal(0, 1, (m, 4))
    a = u.copy(order="F")
    fl = ctypes.c_uint(0)
    hm.hm_floors(m, int(mhd), 0.05, 0.01, 1, 1, int(disallow), dp(a), ctypes.byref(fl))
    p = oracle.params(magnetized=mhd, smalld=0.05, smallei=0.01, disallow_negatives=disallow)
    u1 = np.asfortranarray(u[:, :5]).copy(order="F")
    b1 = np.asfortranarray(u[:, 5:8]) if mhd else None
    rep = oracle.Report()
    oracle.lib().or_floors(ctypes.byref(p), m, dp(u1), None if b1 is None else dp(b1), 0, 1.0,
                           ctypes.byref(rep))
    assert np.array_equal(a[:, :5], u1)
    assert rep.dn_negative == 1 and bool(fl.value & 1)
    assert bool(fl.value & 2) == bool(rep.ei_negative) == disallow


def test_floor_mass_counters_exact(hm):
    """The device's floor-mass counters (pk_mass.hpp): the total of values
    spread over 70 binary orders of magnitude, added digit by digit in any
    order, is the exact sum correctly rounded to FP64 (the same bits in
    every order); zeros and values below 2^-160 add nothing."""
    from fractions import Fraction

    rng = np.random.default_rng(21)
    for trial in range(20):
        n = int(rng.integers(1, 400))
        x = np.abs(rng.standard_normal(n)) * np.exp2(rng.integers(-60, 10, n).astype(np.float64))
        x[rng.random(n) < 0.1] = 0.0
        exact = float(sum(Fraction(float(v)) for v in x))
        a = np.ascontiguousarray(x)
        got = hm.hm_mass_sum(n, dp(a))
        assert got == exact, (trial, got, exact)
        b = np.ascontiguousarray(rng.permutation(x))
        assert hm.hm_mass_sum(n, dp(b)) == got
    tiny = np.array([2.0 ** -170, 0.0, 2.0 ** -200])
    assert hm.hm_mass_sum(3, dp(tiny)) == 0.0
    one = np.array([1.0, 2.0 ** -160, 2.0 ** 80])
    assert hm.hm_mass_sum(3, dp(one)) == float(Fraction(1) + Fraction(2) ** -160 + Fraction(2) ** 80)
