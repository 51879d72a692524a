"""The oracle's block-list driver and two-fluid blocks, pinned against the
single-block, single-fluid oracle (itself pinned by test_oracle_pointwise
and test_oracle_scheme):

- fluids do not interact (internal sources vanish, sources.F90:151-244), so
  a block carrying an ionized and a neutral fluid must evolve each fluid
  exactly as a one-fluid block does with the same dt;
- time_step over fluids and blocks is the min of the parts
  (timestep.F90:112-225); update_chspeed is the min over blocks and the GLM
  factor comes from the max over blocks (hdc.F90:112-140,313-331), so blocks
  of one size evolve in a list exactly as alone.
"""
import numpy as np
import pytest

from piernik_amd import setups

ION = (5.0 / 3.0, True)
NEU = (1.4, False)


def one_fluid_copy(po, cg2, f, mag):
    """The fluid f of a two-fluid block as a one-fluid oracle block."""
    ob = po.OBlock(cg2.n, cg2.nb, cg2.dl, mag, cg2.has_dir, nvar=5)
    ob.u[...] = cg2.u[5 * f:5 * f + 5]
    if mag:
        ob.b[...] = cg2.b
        ob.psi[...] = cg2.psi
    return ob


@pytest.mark.parametrize("limiter", ["vanleer", "minmod"])
def test_two_fluid_block_equals_separate_fluids(oracle, limiter):
    po = oracle
    cfg = setups.batch_config(fluids=(ION, NEU), limiter=limiter)
    cg = setups.batch_blocks([10], cfg, seed=4)[0]
    p2 = po.params_from_config(cfg)
    p_ion = po.params(gamma=ION[0], magnetized=True, limiter=limiter, smalld=cfg.smalld, smallei=cfg.smallei,
                      cfl=cfg.cfl, cfl_glm=cfg.cfl_glm, glm_alpha=cfg.glm_alpha)
    p_neu = po.params(gamma=NEU[0], magnetized=False, limiter=limiter, smalld=cfg.smalld, smallei=cfg.smallei,
                      cfl=cfg.cfl)
    ob2 = po.OBlock.from_cg(cg)
    ion = one_fluid_copy(po, cg, 0, True)
    neu = one_fluid_copy(po, cg, 1, False)
    for ob, p in ((ob2, p2), (ion, p_ion), (neu, p_neu)):
        po.fill_ghosts(p, ob)
    for step in range(2):
        dt2, _ = po.time_step(p2, ob2, step, -1.0)
        dti, _ = po.time_step(p_ion, ion, step, -1.0)
        dtn, _ = po.time_step(p_neu, neu, step, -1.0)
        assert dt2 == min(dti, dtn)
        po.fluid_update(p2, ob2, dt2)
        po.fluid_update(p_ion, ion, dt2)
        po.fluid_update(p_neu, neu, dt2)
        assert np.array_equal(ob2.u[0:5], ion.u)
        assert np.array_equal(ob2.u[5:10], neu.u)
        assert np.array_equal(ob2.b, ion.b) and np.array_equal(ob2.psi, ion.psi)


def test_list_of_one_block_equals_run(oracle):
    po = oracle
    cfg = setups.batch_config(fluids=(ION,))
    cg = setups.batch_blocks([12], cfg, seed=1)[0]
    p = po.params_from_config(cfg)
    a, b = po.OBlock.from_cg(cg), po.OBlock.from_cg(cg)
    ta = po.run(p, a, 3)
    tb = po.run_list(p, [b], 3)
    assert ta == tb
    assert np.array_equal(a.u, b.u) and np.array_equal(a.b, b.b) and np.array_equal(a.psi, b.psi)


def test_list_dt_is_min_over_blocks_hd(oracle):
    """HD blocks of different sizes: one list step == each block updated alone
    with the list's dt (the min over blocks)."""
    po = oracle
    cfg = setups.batch_config(fluids=(NEU,))
    cgs = setups.batch_blocks([8, 12, 10], cfg, seed=2)
    p = po.params_from_config(cfg)
    lst = [po.OBlock.from_cg(c) for c in cgs]
    alone = [po.OBlock.from_cg(c) for c in cgs]
    po.run_list(p, lst, 1)
    for ob in alone:
        po.fill_ghosts(p, ob)
    dts = [po.time_step(p, ob, 0, -1.0)[0] for ob in alone]
    assert len(set(dts)) > 1  # the blocks disagree, so the min matters
    for ob in alone:
        po.fluid_update(p, ob, min(dts))
    for a, b in zip(lst, alone):
        assert np.array_equal(a.u, b.u)


def test_list_mhd_same_size_blocks_evolve_as_alone(oracle):
    """MHD blocks of one size share ch and the GLM factor with their solo runs."""
    po = oracle
    cfg = setups.batch_config(fluids=(ION, NEU))
    cgs = setups.batch_blocks([10, 10], cfg, seed=5)
    p = po.params_from_config(cfg)
    lst = [po.OBlock.from_cg(c) for c in cgs]
    alone = [po.OBlock.from_cg(c) for c in cgs]
    po.run_list(p, lst, 1)
    for ob in alone:
        po.fill_ghosts(p, ob)
    dt = min(po.time_step(p, ob, 0, -1.0)[0] for ob in alone)
    for ob in alone:
        po.fluid_update(p, ob, dt)
    for a, b in zip(lst, alone):
        assert np.array_equal(a.u, b.u) and np.array_equal(a.b, b.b) and np.array_equal(a.psi, b.psi)


def test_list_mhd_chspeed_is_min_over_blocks(oracle):
    """Different block sizes: the list uses ch = cfl_glm*min(dl)/dt over all
    blocks, so a coarse block in the list differs from its solo run unless it
    is given that ch (planted check: passing the block's own ch must differ)."""
    po = oracle
    cfg = setups.batch_config(fluids=(ION,))
    cgs = setups.batch_blocks([8, 12], cfg, seed=6)
    p = po.params_from_config(cfg)
    lst = [po.OBlock.from_cg(c) for c in cgs]
    po.run_list(p, lst, 1)
    solo = po.OBlock.from_cg(cgs[0])  # the coarse block (dl = 1/8)
    po.fill_ghosts(p, solo)
    fine = po.OBlock.from_cg(cgs[1])
    po.fill_ghosts(p, fine)
    dt = min(po.time_step(p, solo, 0, -1.0)[0], po.time_step(p, fine, 0, -1.0)[0])
    po.fluid_update(p, solo, dt)  # uses the coarse block's own (larger) ch
    assert not np.array_equal(lst[0].psi, solo.psi)
    # restated with the list's ch and factor, stage by stage
    ch = min(po.update_chspeed(p, solo, dt), po.update_chspeed(p, fine, dt))
    fac = min(po.glm_factor(p, solo, ch, dt), po.glm_factor(p, fine, ch, dt))
    ref = po.OBlock.from_cg(cgs[0])
    po.fill_ghosts(p, ref)
    for half in range(2):
        for d in ((0, 1, 2) if half == 0 else (2, 1, 0)):
            for istep in (po.RK2_1, po.RK2_2):
                po.solve_cg(p, ref, d, istep, dt, ch)
                po.fill_ghosts(p, ref)
        ref.psi *= fac
    assert np.array_equal(lst[0].u, ref.u) and np.array_equal(lst[0].psi, ref.psi)
