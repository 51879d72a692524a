"""Step retries (cflcontrol = 'redo' / 'flexible', pk_set_cfl_control) on the
GPU against the reference's control flow driving the oracle
(oracle.pyoracle.main_loop, pinned in tests/test_oracle_retry.py).

The reference's main loop (piernik.F90:125-170) saves the state before each
step (repeat_fluidstep, timestep_retry.F90:86-273), and after the step
check_cfl_violation (timestep.F90:247-310) recomputes the timestep of the
new state; cfl_warn (:317-361) flags a redo when c_max(new)/c_max(old) * cfl
* dt_cur_shrink exceeds cfl_max, and the step is redone from the saved state
with dt = dt_full * dt_cur_shrink.  The Sedov blast at CFL 0.7 has a
stepcfl of 0.859 in its first step: with cfl_max = 0.8 that step is redone
once.  Everything else must stay bit-exact.
"""
import numpy as np
import pytest

import piernik_amd as pk
from piernik_amd import setups

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,driver", [("redo", "block"), ("flexible", "block"), ("redo", "list"),
                                         ("auto", "block"), ("auto", "list")])
def test_redo_matches_reference_control_flow(oracle, mode, driver):
    cg, cfg = setups.sedov_hd(n=24)
    p = oracle.params_from_config(cfg)
    ob = oracle.OBlock.from_cg(cg)
    t_o, redone_o = oracle.run_redo(p, ob, 6, mode, 0.8)
    assert redone_o >= 1 or mode == "auto"
    print(f"cflcontrol {mode} via {driver}: oracle redid {redone_o} step(s)", flush=True)
    pk.load_library()
    pk.finalize()
    pk.init(cfg)
    pk.set_cfl_control(mode, cfl_max=0.8)
    cg.upload()
    if driver == "block":
        t_g, _ = pk.run(cg, 6)
    else:
        t_g, _ = pk.run_all(6, pk.BATCH_GRAPH)
    assert pk.redone_steps() == redone_o
    assert t_g == t_o
    cg.download()
    bad = np.argwhere(~(cg.u == ob.u))
    assert bad.size == 0, f"{len(bad)} mismatches, first {tuple(bad[0])}"
    pk.finalize()


def test_warn_never_redoes():
    cg, cfg = setups.sedov_hd(n=16)
    pk.load_library()
    pk.finalize()
    pk.init(cfg)
    pk.set_cfl_control("warn", cfl_max=0.8)
    cg.upload()
    pk.run(cg, 4)
    assert pk.redone_steps() == 0
    pk.finalize()


@pytest.mark.parametrize("driver", ["block", "list"])
def test_time_control_matches_reference(oracle, driver):
    """dt_initial > 0, dt_max_grow, dt_max and tend applied as time_step does
    (timestep.F90:198-224), the loop ending at tend (piernik.F90:125).  The
    oracle's time_step applies the default caps (dt_initial = -1,
    dt_max_grow = 2); the tighter ones here are applied on top of them."""
    dt_initial, grow, dt_max, tend = 2.0e-4, 1.2, 1.5e-3, 0.012
    eps = np.finfo(np.float64).eps
    cg, cfg = setups.sedov_hd(n=24)
    p = oracle.params_from_config(cfg)
    ob = oracle.OBlock.from_cg(cg)
    oracle.fill_ghosts(p, ob)
    t, s, dt_old = 0.0, 0, -1.0
    while s < 100 and t < tend:
        dt, _ = oracle.time_step(p, ob, s, dt_old)
        dt = min(dt, dt_initial) if s < 1 else min(dt, dt_old * grow)
        dt = min(min(dt, dt_max), (0.5 * (tend - t)) + ((2.0 * eps) * (tend - t)))
        oracle.fluid_update(p, ob, dt)
        dt_old = dt
        t = t + dt
        t = t + dt
        s += 1
    assert 5 < s < 100
    print(f"time control via {driver}: {s} steps", flush=True)
    pk.load_library()
    pk.finalize()
    pk.init(cfg)
    pk.set_time_control(dt_initial=dt_initial, dt_max_grow=grow, dt_max=dt_max, tend=tend)
    cg.upload()
    t_g, _ = pk.run(cg, 100) if driver == "block" else pk.run_all(100, pk.BATCH_GRAPH)
    assert t_g == t and t_g >= tend
    cg.download()
    assert np.array_equal(cg.u, ob.u)
    with pytest.raises(pk.PiernikError, match="below dt_min"):
        pk.set_time_control(dt_min=1.0)
        pk.run(cg, 1)
    pk.finalize()


@pytest.mark.parametrize("disallow", [False, True])
def test_negatives_and_redo(oracle, disallow):
    """Floors that raise ei_negative in every attempt (cells with zero total
    energy): without disallow_negatives the steps are kept and the state
    matches the oracle bit for bit; with it every attempt is redone until
    max_redostep_attempts is exceeded, on the GPU as in the reference's
    control flow (which dies there)."""
    from floors_setup import floor_block

    cg, cfg = floor_block(magnetized=True)
    p = oracle.params_from_config(cfg, disallow_negatives=disallow)
    ob = oracle.OBlock.from_cg(cg)
    pk.load_library()
    pk.finalize()
    pk.init(cfg)
    pk.set_cfl_control("redo", cfl_max=0.95, max_redostep_attempts=2, disallow_negatives=disallow)
    cg.upload()
    print(f"negatives with cflcontrol redo, disallow_negatives={disallow}", flush=True)
    if disallow:
        with pytest.raises(RuntimeError, match="max_redostep_attempts"):
            oracle.run_redo(p, ob, 2, "redo", 0.95, max_attempts=2)
        with pytest.raises(pk.PiernikError, match="max_redostep_attempts"):
            pk.run(cg, 2)
    else:
        t_o, redone_o = oracle.run_redo(p, ob, 2, "redo", 0.95, max_attempts=2)
        t_g, _ = pk.run(cg, 2)
        assert t_g == t_o and pk.redone_steps() == redone_o
        cg.download()
        for name in ("u", "b", "psi"):
            assert np.array_equal(getattr(cg, name), getattr(ob, name)), name
        rep = pk.get_report(cg)
        assert rep["dn_negative"] == bool(ob.rep.dn_negative)
        assert rep["ei_negative"] == bool(ob.rep.ei_negative) == False  # noqa: E712
    pk.finalize()
