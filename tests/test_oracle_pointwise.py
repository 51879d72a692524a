"""Pins of the oracle's pointwise routines against what mathematics fixes.

The reference ships no golden vectors for this path and its Fortran cannot
be built here (its modules need MPI), so each oracle routine is checked
against an independent statement of the same mathematics:
  spacing     numpy.spacing (IEEE), Fortran's TINY below the normal range
  limiters    textbook minmod / monotonized-central / superbee (Roe) /
              harmonic van Leer, plus the reference's pencil-end rules
  PLM         exact face values on linear data
  utoq        inverse of the primitive -> conserved map
  HLLC/HLLD   consistency F(U,U) = F(U); upwinding; exact resolution of
              isolated contact, tangential and rotational (Alfven)
              discontinuities (Rankine-Hugoniot); mirror symmetry; HLLD with
              B = 0 against HLLC
  GLM flux    exact Godunov flux of the linear (B_n, psi) subsystem
"""
import numpy as np
import pytest

GAM = 5.0 / 3.0


# ---------------------------------------------------------------- helpers --

def phys_flux_mhd(q, gam=GAM):
    """Physical flux of ideal MHD from primitives [rho, vn, vt1, vt2, p, Bn, Bt1, Bt2]."""
    rho, vx, vy, vz, p, bx, by, bz = [q[..., i] for i in range(8)]
    b2 = bx * bx + by * by + bz * bz
    pt = p + 0.5 * b2
    e = p / (gam - 1) + 0.5 * rho * (vx * vx + vy * vy + vz * vz) + 0.5 * b2
    vb = vx * bx + vy * by + vz * bz
    f = np.stack([rho * vx, rho * vx * vx + pt - bx * bx, rho * vy * vx - bx * by, rho * vz * vx - bx * bz,
                  (e + pt) * vx - bx * vb], axis=-1)
    fb = np.stack([by * vx - bx * vy, bz * vx - bx * vz], axis=-1)
    return f, fb


def phys_flux_hd(q, gam=GAM):
    z = np.zeros(q.shape[:-1] + (3,))
    f, _ = phys_flux_mhd(np.concatenate([q, z], axis=-1), gam)
    return f


def hlld_full(oracle, ql, qr, gam=GAM, ch=1.0, psil=None, psir=None):
    m = ql.shape[0]
    bl = np.zeros((m, 4))
    br = np.zeros((m, 4))
    bl[:, :3], br[:, :3] = ql[:, 5:8], qr[:, 5:8]
    if psil is not None:
        bl[:, 3], br[:, 3] = psil, psir
    return oracle.hlld(ql[:, :5], qr[:, :5], bl, br, gam, ch)


def rand_prims(rng, m, mhd=False, vmax=1.0):
    q = np.empty((m, 8 if mhd else 5))
    q[:, 0] = rng.uniform(0.1, 3.0, m)
    q[:, 1:4] = rng.uniform(-vmax, vmax, (m, 3))
    q[:, 4] = rng.uniform(0.05, 3.0, m)
    if mhd:
        q[:, 5:8] = rng.uniform(-1.5, 1.5, (m, 3))
    return q


def rel(a, b):
    return np.max(np.abs(a - b) / (np.abs(b) + 1e-12))


# --------------------------------------------------------------- spacing --

def test_spacing_matches_ieee_spacing(oracle):
    rng = np.random.default_rng(0)
    xs = rng.normal(0, 1, 2000) * 10.0 ** rng.integers(-250, 250, 2000)
    for x in xs:
        assert oracle.spacing(x) == np.spacing(abs(x))
    tiny = np.finfo(np.float64).tiny
    for x in (0.0, -0.0, 5e-324, 1e-310, tiny):
        assert oracle.spacing(x) == tiny


# -------------------------------------------------------------- limiters --

def textbook(kind, a, b):
    if kind == "minmod":
        return np.where(a * b > 0, np.sign(a) * np.minimum(np.abs(a), np.abs(b)), 0.0)
    if kind == "moncen":
        m = np.minimum(np.minimum(2 * np.abs(a), 2 * np.abs(b)), 0.5 * np.abs(a + b))
        return np.where(a * b > 0, np.sign(a) * m, 0.0)
    if kind == "superbee":
        m = np.maximum(np.minimum(2 * np.abs(a), np.abs(b)), np.minimum(np.abs(a), 2 * np.abs(b)))
        return np.where(a * b > 0, np.sign(a) * m, 0.0)
    if kind == "vanleer":
        with np.errstate(invalid="ignore", divide="ignore"):
            return np.where(a * b > 0, 2 * a * b / (a + b), 0.0)
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["minmod", "moncen", "superbee", "vanleer"])
def test_limiters_match_textbook(oracle, kind):
    rng = np.random.default_rng(1)
    q = np.cumsum(rng.normal(0, 1, 3000) * (rng.random(3000) < 0.9))  # includes flat stretches
    dq = oracle.limiter(kind, q)
    a = q[1:-1] - q[:-2]
    b = q[2:] - q[1:-1]
    np.testing.assert_allclose(dq[1:-1], textbook(kind, a, b), rtol=1e-15, atol=0)


@pytest.mark.parametrize("kind", ["minmod", "moncen", "superbee", "vanleer"])
def test_limiter_pencil_ends(oracle, kind):
    """fluxlimiters.F90: van Leer zeroes both end slopes; the others reuse
    the one available difference (dlft(1)=dlft(2), drgt(n)=drgt(n-1))."""
    q = np.array([1.0, 1.5, 2.75, 2.0, 4.0, 3.0])
    dq = oracle.limiter(kind, q)
    if kind == "vanleer":
        assert dq[0] == 0.0 and dq[-1] == 0.0
    else:
        assert dq[0] == pytest.approx(q[1] - q[0], rel=1e-15)
        assert dq[-1] == pytest.approx(q[-1] - q[-2], rel=1e-15)


@pytest.mark.parametrize("kind", ["minmod", "moncen", "superbee", "vanleer"])
def test_plm_exact_on_linear_data(oracle, kind):
    x = np.arange(12.0)
    q = np.stack([0.5 + 0.25 * x, -2.0 + 0.125 * x], axis=1)  # exactly representable lines
    ql, qr = oracle.linear(kind, q)
    face = np.stack([0.5 + 0.25 * (x[:-1] + 0.5), -2.0 + 0.125 * (x[:-1] + 0.5)], axis=1)
    np.testing.assert_array_equal(ql[1:-1], face[1:-1])
    np.testing.assert_array_equal(qr[1:-1], face[1:-1])


@pytest.mark.parametrize("kind", ["minmod", "moncen", "superbee", "vanleer"])
def test_limiter_tvd_bounds(oracle, kind):
    rng = np.random.default_rng(2)
    q = rng.normal(0, 1, 2000)
    dq = oracle.limiter(kind, q)[1:-1]
    a = q[1:-1] - q[:-2]
    b = q[2:] - q[1:-1]
    ext = a * b <= 0
    assert np.all(dq[ext] == 0)
    assert np.all(np.abs(dq) <= 2 * np.minimum(np.abs(a), np.abs(b)) * (1 + 1e-15))
    assert np.all(dq[~ext] * a[~ext] > 0)


# ------------------------------------------------------------------ utoq --

@pytest.mark.parametrize("mhd", [False, True])
def test_utoq_inverts_prim_to_cons(oracle, mhd):
    rng = np.random.default_rng(3)
    m = 400
    q = rand_prims(rng, m, mhd)
    u = np.empty((m, 5))
    u[:, 0] = q[:, 0]
    u[:, 1:4] = q[:, :1] * q[:, 1:4]
    e = q[:, 4] / (GAM - 1) + 0.5 * q[:, 0] * np.sum(q[:, 1:4] ** 2, axis=1)
    b = None
    if mhd:
        b = np.zeros((m, 4))
        b[:, :3] = q[:, 5:8]
        e = e + 0.5 * np.sum(b[:, :3] ** 2, axis=1)
    u[:, 4] = e
    p = oracle.params(magnetized=mhd)
    qo = oracle.utoq(p, u, b)
    np.testing.assert_allclose(qo, q[:, :5], rtol=1e-12, atol=1e-13)


# ------------------------------------------------------------ consistency --

@pytest.mark.parametrize("vmax", [0.3, 5.0])
def test_hllc_consistency(oracle, vmax):
    rng = np.random.default_rng(4)
    q = rand_prims(rng, 2000, vmax=vmax)
    f = oracle.hllc(q, q, GAM)
    np.testing.assert_allclose(f, phys_flux_hd(q), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("vmax", [0.3, 5.0])
def test_hlld_consistency(oracle, vmax):
    rng = np.random.default_rng(5)
    q = rand_prims(rng, 2000, mhd=True, vmax=vmax)
    f, bf = hlld_full(oracle, q, q)
    fp, fbp = phys_flux_mhd(q)
    np.testing.assert_allclose(f, fp, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(bf[:, 1:3], fbp, rtol=1e-11, atol=1e-11)


def test_supersonic_upwinding(oracle):
    rng = np.random.default_rng(6)
    ql = rand_prims(rng, 500, mhd=True)
    qr = rand_prims(rng, 500, mhd=True)
    ql[:, 1] += 20.0
    qr[:, 1] += 20.0
    f, bf = hlld_full(oracle, ql, qr)
    fp, fbp = phys_flux_mhd(ql)
    np.testing.assert_allclose(f, fp, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(bf[:, 1:3], fbp, rtol=1e-13, atol=1e-13)
    g = oracle.hllc(ql[:, :5] - np.array([0, 40, 0, 0, 0]), qr[:, :5] - np.array([0, 40, 0, 0, 0]), GAM)
    np.testing.assert_allclose(g, phys_flux_hd(qr[:, :5] - np.array([0, 40, 0, 0, 0])), rtol=1e-13, atol=1e-13)


# ------------------------------------------------- isolated discontinuities --

@pytest.mark.parametrize("u0", [0.0, 0.4, -0.7])
def test_hllc_contact_resolved_exactly(oracle, u0):
    """An isolated contact (equal p and v, jump in rho and transverse v)
    moving at u0: the Godunov flux is the upwind physical flux."""
    ql = np.array([[1.0, u0, 0.3, -0.2, 0.8]])
    qr = np.array([[0.2, u0, -0.5, 0.9, 0.8]])
    f = oracle.hllc(ql, qr, GAM)[0]
    up = ql if u0 > 0 else qr
    expect = phys_flux_hd(up)[0]
    if u0 == 0.0:
        expect = np.array([0.0, 0.8, 0.0, 0.0, 0.0])
    np.testing.assert_allclose(f, expect, rtol=1e-14, atol=1e-15)


def test_hlld_tangential_discontinuity_stationary(oracle):
    """B_n = 0, v_n = 0, equal total pressure: a stationary tangential
    discontinuity; the flux is (0, p_tot, 0, 0, 0) and no B flux."""
    ql = np.array([[1.0, 0.0, 0.4, -0.1, 1.0, 0.0, 0.8, 0.2]])
    b2l = 0.8 ** 2 + 0.2 ** 2
    b2r = 0.3 ** 2 + 1.1 ** 2
    pr = 1.0 + 0.5 * b2l - 0.5 * b2r
    qr = np.array([[0.3, 0.0, -0.6, 0.7, pr, 0.0, 0.3, 1.1]])
    f, bf = hlld_full(oracle, ql, qr)
    np.testing.assert_allclose(f[0], [0.0, 1.0 + 0.5 * b2l, 0.0, 0.0, 0.0], atol=1e-14)
    np.testing.assert_allclose(bf[0, 1:3], [0.0, 0.0], atol=1e-14)


@pytest.mark.parametrize("bx,theta,p", [(1.0, 0.9, 0.9), (-0.7, 2.0, 0.9), (1.3, -1.2, 1.2)])
def test_hlld_rotational_discontinuity_resolved(oracle, bx, theta, p):
    """An isolated Alfven (rotational) discontinuity: same rho, p, |B|, v_n;
    B_t rotated and v_t jumping by -sign(B_n) dB_t / sqrt(rho).  It moves at
    v_n + |B_n|/sqrt(rho) (Rankine-Hugoniot), so the flux at the interface
    is the physical flux of the left state; HLLD resolves it exactly as long
    as B_n^2/gamma <= p (otherwise hlld.F90:364 takes the HLL fallback)."""
    rho, vn = 1.3, 0.1
    assert bx * bx / GAM <= p
    bt = 0.8
    byl, bzl = bt, 0.0
    byr, bzr = bt * np.cos(theta), bt * np.sin(theta)
    s = np.sign(bx)
    vyl, vzl = 0.2, -0.3
    vyr = vyl - s * (byr - byl) / np.sqrt(rho)
    vzr = vzl - s * (bzr - bzl) / np.sqrt(rho)
    ql = np.array([[rho, vn, vyl, vzl, p, bx, byl, bzl]])
    qr = np.array([[rho, vn, vyr, vzr, p, bx, byr, bzr]])
    # Rankine-Hugoniot check of the construction itself
    fl, fbl = phys_flux_mhd(ql)
    fr, fbr = phys_flux_mhd(qr)
    speed = vn + abs(bx) / np.sqrt(rho)

    def cons(q):
        e = q[0, 4] / (GAM - 1) + 0.5 * q[0, 0] * np.sum(q[0, 1:4] ** 2) + 0.5 * np.sum(q[0, 5:8] ** 2)
        return np.array([q[0, 0], *(q[0, 0] * q[0, 1:4]), e])
    np.testing.assert_allclose(fr[0] - fl[0], speed * (cons(qr) - cons(ql)), atol=1e-13)
    f, bf = hlld_full(oracle, ql, qr)
    np.testing.assert_allclose(f[0], fl[0], rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(bf[0, 1:3], fbl[0], rtol=1e-12, atol=1e-13)


# ----------------------------------------------------------- symmetries --

def mirror(q, mhd):
    """x -> -x combined with B -> -B (ideal MHD is even in B): v_n and B_n
    change sign, everything else is unchanged."""
    r = q.copy()
    r[:, 1] = -r[:, 1]
    if mhd:
        r[:, 5] = -r[:, 5]
    return r


def test_hllc_mirror_symmetry(oracle):
    rng = np.random.default_rng(7)
    ql, qr = rand_prims(rng, 1000), rand_prims(rng, 1000)
    f = oracle.hllc(ql, qr, GAM)
    g = oracle.hllc(mirror(qr, False), mirror(ql, False), GAM)
    sgn = np.array([-1, 1, -1, -1, -1])
    np.testing.assert_allclose(g, sgn * f, rtol=1e-12, atol=1e-12)


def test_hlld_mirror_symmetry(oracle):
    """With a continuous normal field (HLLD's premise: the solver branches on
    the left B_n alone, hlld.F90:410,452)."""
    rng = np.random.default_rng(8)
    ql, qr = rand_prims(rng, 1000, True), rand_prims(rng, 1000, True)
    qr[:, 5] = ql[:, 5]
    f, bf = hlld_full(oracle, ql, qr)
    g, bg = hlld_full(oracle, mirror(qr, True), mirror(ql, True))
    np.testing.assert_allclose(g, np.array([-1, 1, -1, -1, -1]) * f, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(bg[:, 1:3], -bf[:, 1:3], rtol=1e-10, atol=1e-10)


def test_hlld_without_field_equals_hllc(oracle):
    rng = np.random.default_rng(9)
    ql, qr = rand_prims(rng, 1000, True), rand_prims(rng, 1000, True)
    ql[:, 5:8] = 0.0
    qr[:, 5:8] = 0.0
    f, bf = hlld_full(oracle, ql, qr)
    g = oracle.hllc(ql[:, :5], qr[:, :5], GAM)
    np.testing.assert_allclose(f, g, rtol=1e-12, atol=1e-13)
    assert np.all(bf[:, 1:3] == 0.0)


# -------------------------------------------------------------- GLM flux --

@pytest.mark.parametrize("ch", [0.5, 2.3])
def test_glm_flux_is_exact_linear_godunov_flux(oracle, ch):
    """(B_n, psi) obey dB/dt + dpsi/dx = 0, dpsi/dt + ch^2 dB/dx = 0.  The
    exact flux at x/t = 0 follows from the characteristic decomposition."""
    rng = np.random.default_rng(10)
    m = 200
    ql, qr = rand_prims(rng, m, True), rand_prims(rng, m, True)
    psil, psir = rng.normal(0, 1, m), rng.normal(0, 1, m)
    _, bf = hlld_full(oracle, ql, qr, ch=ch, psil=psil, psir=psir)
    A = np.array([[0.0, 1.0], [ch * ch, 0.0]])
    lam, R = np.linalg.eig(A)
    Rinv = np.linalg.inv(R)
    for i in range(m):
        wl = Rinv @ np.array([ql[i, 5], psil[i]])
        wr = Rinv @ np.array([qr[i, 5], psir[i]])
        w = np.where(lam > 0, wl, wr)
        state = R @ w
        flux = A @ state
        np.testing.assert_allclose([bf[i, 0], bf[i, 3]], flux, rtol=1e-12, atol=1e-12)
