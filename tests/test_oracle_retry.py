"""Pins of the oracle's main loop with step retries (pyoracle.main_loop,
the reference's repeat_fluidstep / check_cfl_violation / cfl_warn /
cfl_auto control flow) on synthetic steps whose outcome is known in closed
form, and of the floor bookkeeping (care_for_positives) against the mass it
must account for.
"""
import ctypes

import numpy as np
import pytest

import floors_setup as fs


class Synthetic:
    """A 'state' whose c_max after a step is ratio(dt) times the c_max
    before it; time_step always proposes dt = 1.  neg_attempts: the update
    attempts (0-based) that raise ei_negative."""

    def __init__(self, ratio, neg_attempts=()):
        self.ratio, self.neg_attempts = ratio, set(neg_attempts)
        self.c, self.saved, self.n, self.flag = 1.0, None, 0, (0, 0)

    def time_step(self, s, dt_old):
        return 1.0, self.c

    def update(self, dt):
        self.flag = (0, 1) if self.n in self.neg_attempts else (0, 0)
        self.n += 1
        self.c = self.c * self.ratio(dt)

    def save(self):
        return self.c

    def restore(self, c):
        self.c = c

    def negatives(self):
        out, self.flag = self.flag, (0, 0)
        return out


def drive(oracle, syn, mode, nsteps=1, **kw):
    return oracle.main_loop(nsteps, syn.time_step, syn.update, syn.save, syn.restore, syn.negatives, 0.5,
                            mode, **kw)


def test_cfl_max_sanitized(oracle):
    # global.F90:304: at least min(1.1 cfl, cfl + 0.05, (1 + cfl)/2), at most 1
    assert oracle.cfl_max_sanitized(0.5, 0.6) == 0.6
    assert oracle.cfl_max_sanitized(0.5, 0.51) == 0.55
    assert oracle.cfl_max_sanitized(0.9, 2.0) == 1.0


def test_redo_shrinks_by_powers(oracle):
    """c_max quadruples in any step: stepcfl = 4*0.5*shrink > 0.6 until
    shrink = 1/4, so the step is redone twice, with dt 1, 1/2, 1/4."""
    syn = Synthetic(lambda dt: 4.0)
    t, redone, att = drive(oracle, syn, "redo", cfl_max=0.6, dt_shrink=0.5)
    assert redone == 2
    assert att == [(0, 1.0), (0, 0.5), (0, 0.25)]
    assert t == 0.5


def test_flexible_lands_on_cfl_max(oracle):
    """'flexible' shrinks by cfl_max/stepcfl: 0.6/2 = 0.3 makes stepcfl
    exactly cfl_max, which is kept; one redo."""
    syn = Synthetic(lambda dt: 4.0)
    t, redone, att = drive(oracle, syn, "flexible", cfl_max=0.6, dt_shrink=0.5)
    assert redone == 1
    assert att == [(0, 1.0), (0, 0.3)]


def test_flexible_second_attempt_also_shrinks(oracle):
    """From the second attempt on the flexible shrink is cfl_max/stepcfl
    times dt_shrink (cfl_warn, timestep.F90:349-351).  Attempt 0: stepcfl
    = 4*0.5 = 2 -> shrink 0.6/2 = 0.3.  Attempt 1 (stepcfl 0.6) raises
    ei_negative -> shrink 0.6/0.6*0.5 = 0.5.  Attempt 2: stepcfl = 1 ->
    shrink 0.6/1*0.5 = 0.3, kept: three redos, dt 1, 0.3, 0.5, 0.3."""
    syn = Synthetic(lambda dt: 4.0, neg_attempts=(1,))
    _, redone, att = drive(oracle, syn, "flexible", cfl_max=0.6, dt_shrink=0.5)
    assert redone == 3
    assert [dt for _, dt in att] == [1.0, 0.3, 0.5, 0.3]


@pytest.mark.parametrize("disallow", [False, True])
def test_negatives_redo_only_when_disallowed(oracle, disallow):
    """A step that raises ei_negative once is redone (dt halved) only with
    disallow_negatives (timestep.F90:291-298)."""
    syn = Synthetic(lambda dt: 1.0, neg_attempts=(0,))
    _, redone, att = drive(oracle, syn, "redo", cfl_max=0.9, disallow_negatives=disallow)
    assert redone == (1 if disallow else 0)
    assert att == ([(0, 1.0), (0, 0.5)] if disallow else [(0, 1.0)])


def test_warn_never_redoes(oracle):
    syn = Synthetic(lambda dt: 100.0, neg_attempts=(0, 1, 2))
    t, redone, att = drive(oracle, syn, "warn", nsteps=3)
    assert redone == 0 and len(att) == 3 and t == 6.0


def test_max_attempts(oracle):
    syn = Synthetic(lambda dt: 1e9)
    with pytest.raises(RuntimeError, match="max_redostep_attempts"):
        drive(oracle, syn, "redo", max_attempts=3)
    assert syn.n == 4


def test_auto_factor(oracle):
    """cfl_auto: constant c_max leaves dt alone; c_max doubling in step 1
    gives cfl_c = (1 + 1/2)/2 = 0.75 (timestep.F90:386-389)."""
    syn = Synthetic(lambda dt: 1.0)
    _, _, att = drive(oracle, syn, "auto", nsteps=3)
    assert [dt for _, dt in att] == [1.0, 1.0, 1.0]
    syn = Synthetic(lambda dt: 2.0)
    _, _, att = drive(oracle, syn, "auto", nsteps=2)
    assert [dt for _, dt in att] == [1.0, 0.75]


# ---- the floor bookkeeping -------------------------------------------------

def test_magic_mass_reference_form(oracle):
    """local_magic_mass takes each pencil's interior mass off and adds it
    back after the floor (sources.F90:356-358): the result equals the exact
    mass the floor added over the interior, sum(max(rho, smalld) - rho)*dvol,
    to within rounding at the scale of the pencil's mass; cells outside
    nb+1:n-nb do not count; a pencil the floor leaves alone adds exactly 0."""
    from fractions import Fraction

    rng = np.random.default_rng(11)
    n, nbp, dvol, smalld = 40, 4, 0.37, 0.5
    p = oracle.params(smalld=smalld, smallei=1e-12)
    u = np.empty((n, 5), order="F")
    u[:, 0] = rng.uniform(0.1, 2.0, n)
    u[:, 1:4] = rng.normal(0, 0.1, (n, 3))
    u[:, 4] = 10.0
    u[0, 0] = u[n - 1, 0] = 1e-3  # ghost cells below the floor: not counted
    exact = sum(Fraction(max(r, smalld)) - Fraction(r) for r in u[nbp:n - nbp, 0]) * Fraction(dvol)
    rep = oracle.Report()
    a = u.copy(order="F")
    oracle.lib().or_floors(ctypes.byref(p), n, oracle._dp(a), None, nbp, dvol, ctypes.byref(rep))
    scale = float(np.sum(a[nbp:n - nbp, 0])) * dvol
    assert abs(rep.magic_mass[0] - float(exact)) <= 4 * n * fs.EPS * scale
    assert float(exact) > 100 * n * fs.EPS * scale  # the check resolves the added mass
    assert np.all(a[:, 0] >= smalld)
    # a pencil above the floor: (0 - S*dvol) + S*dvol is exactly 0
    rep2 = oracle.Report()
    b = u.copy(order="F")
    b[:, 0] = rng.uniform(0.6, 2.0, n)
    oracle.lib().or_floors(ctypes.byref(p), n, oracle._dp(b), None, nbp, dvol, ctypes.byref(rep2))
    assert rep2.magic_mass[0] == 0.0


@pytest.mark.parametrize("mhd", [False, True])
def test_magic_mass_balances_block_mass(oracle, mhd):
    """On a periodic block the fluxes conserve mass, so with one stage per
    sweep (EULER) the mass gained over a run is the mass the density floor
    added: M(end) - M(start) equals the reported magic mass to within
    rounding (sources.F90:356-358, mass_defect.F90).  (With RK2 the first
    stage's floor changes only the intermediate state, whose fluxes
    conserve mass, yet its additions are counted as the reference counts
    them.)"""
    cg, cfg = fs.floor_block(magnetized=mhd)
    cfg.integration_order = 1
    p = oracle.params_from_config(cfg)
    ob = oracle.OBlock.from_cg(cg)
    dvol = float(np.prod(cg.dl))
    m0 = float(np.sum(cg.interior(ob.u)[0])) * dvol
    oracle.run(p, ob, 3)
    m1 = float(np.sum(cg.interior(ob.u)[0])) * dvol
    mm = ob.rep.magic_mass[0]
    assert mm > 1e-3
    assert abs((m1 - m0) - mm) <= fs.mass_tolerance(cg, 3 * 6) + 64 * fs.EPS * m1


@pytest.mark.parametrize("mhd", [False, True])
def test_floor_setup_fires_both_floors(oracle, mhd):
    """The GPU floor tests' state: with RK2 the density floor adds mass and
    the zero-energy cells raise ei_negative (with disallow_negatives) but no
    density goes negative; without disallow_negatives the state and the
    mass are the same and only the flag differs."""
    cg, cfg = fs.floor_block(magnetized=mhd)
    p = oracle.params_from_config(cfg)
    ob = oracle.OBlock.from_cg(cg)
    oracle.run(p, ob, 2)
    assert ob.rep.magic_mass[0] > 1e-3
    assert ob.rep.ei_negative == 1 and ob.rep.dn_negative == 0
    p2 = oracle.params_from_config(cfg, disallow_negatives=False)
    ob2 = oracle.OBlock.from_cg(cg)
    oracle.run(p2, ob2, 2)
    assert ob2.rep.ei_negative == 0 and ob2.rep.magic_mass[0] == ob.rep.magic_mass[0]
    assert np.array_equal(ob2.u, ob.u)
