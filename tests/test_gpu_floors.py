"""care_for_positives on the GPU (sources.F90:299-433): the density and
internal-energy floors fire, the state stays bit-exact with the oracle, the
flags agree, and the magic mass agrees within the rounding of the
reference's bookkeeping.

The reference accumulates local_magic_mass per pencil as
(mm - sum(rho_before)*dvol) + sum(rho_after)*dvol (sources.F90:356-358),
which rounds at the scale of each pencil's mass; the oracle does the same.
The GPU adds every floor event's max(rho, smalld) - rho exactly (integer
counters, pk_mass.hpp) and rounds once, so its value does not depend on the
kernel, the tile shapes, the launch mode or the run: those are compared
bit for bit.  Tolerance against the oracle: floors_setup.mass_tolerance.
"""
import numpy as np
import pytest

import floors_setup as fs
import piernik_amd as pk
from piernik_amd import setups

pytestmark = pytest.mark.gpu

MODES = {"graph": pk.BATCH_GRAPH, "persistent": pk.BATCH_PERSISTENT, "per_block": pk.BATCH_PER_BLOCK}


def start(cfg, cgs, disallow=True):
    pk.load_library()
    pk.finalize()
    pk.init(cfg)
    pk.set_cfl_control("warn", disallow_negatives=disallow)
    for cg in cgs:
        cg.upload()


def assert_same(cg, ob, where=""):
    for name in ("u", "b", "psi"):
        a, o = getattr(cg, name), getattr(ob, name)
        if a is None:
            continue
        bad = np.argwhere(~(a == o))
        assert bad.size == 0, (f"{where} {name}: {len(bad)} mismatches, first at {tuple(bad[0])}: "
                               f"gpu={a[tuple(bad[0])]!r} oracle={o[tuple(bad[0])]!r}")


def assert_report(cg, ob, stage_sweeps, where=""):
    rep = pk.get_report(cg)
    assert rep["dn_negative"] == bool(ob.rep.dn_negative), where
    assert rep["ei_negative"] == bool(ob.rep.ei_negative), where
    tol = fs.mass_tolerance(cg, stage_sweeps)
    for f in range(2):
        assert abs(rep["magic_mass"][f] - ob.rep.magic_mass[f]) <= tol, (where, f, rep["magic_mass"],
                                                                          list(ob.rep.magic_mass), tol)
    return rep


@pytest.mark.parametrize("mhd", [False, True])
@pytest.mark.parametrize("disallow", [True, False])
def test_floors_run_bit_exact(oracle, mhd, disallow):
    """Three whole steps of the fused stepping loop on a block where both
    floors fire: state bit-exact, flags equal, magic mass within tolerance."""
    cg, cfg = fs.floor_block(magnetized=mhd)
    p = oracle.params_from_config(cfg, disallow_negatives=disallow)
    ob = oracle.OBlock.from_cg(cg)
    print(f"floors run mhd={mhd} disallow={disallow}", flush=True)
    start(cfg, [cg], disallow)
    t_g, _ = pk.run(cg, 3)
    t_o = oracle.run(p, ob, 3)
    assert t_g == t_o
    cg.download()
    assert_same(cg, ob, "run")
    rep = assert_report(cg, ob, 3 * 12, "run")
    assert rep["magic_mass"][0] > 1e-3
    assert rep["ei_negative"] == disallow
    pk.finalize()


@pytest.mark.parametrize("mhd,neg_rho", [(False, False), (True, False), (False, True)])
def test_floors_stages_bit_exact(oracle, mhd, neg_rho):
    """The drop-in stage (pk_solve_cg, whole pencils) stage by stage as the
    adapter drives it; with neg_rho the negative densities raise dn_negative
    in both."""
    cg, cfg = fs.floor_block(magnetized=mhd, neg_rho=neg_rho)
    p = oracle.params_from_config(cfg)
    ob = oracle.OBlock.from_cg(cg)
    print(f"floors stages mhd={mhd} neg_rho={neg_rho}", flush=True)
    start(cfg, [cg])
    dt, ch = 2.0e-3, 1.1
    for ddim in (0, 1, 2):
        for istep in (pk.RK2_1, pk.RK2_2):
            pk.solve_cg(cg, ddim, istep, dt, ch)
            oracle.solve_cg(p, ob, ddim, istep, dt, ch)
            pk.update_boundaries(cg)
            oracle.fill_ghosts(p, ob)
    cg.download()
    assert_same(cg, ob, "stages")
    rep = assert_report(cg, ob, 6, "stages")
    if neg_rho:
        # the negative cells' images at the pencil ends stay negative in the
        # first stage; the floor there (a ghost cell) adds no interior mass
        assert rep["dn_negative"]
    else:
        assert rep["magic_mass"][0] > 0.0
    pk.finalize()


def floor_list(cfg_fluids, shapes, seed):
    cfg = setups.batch_config(fluids=cfg_fluids)
    cfg.smalld, cfg.smallei = 0.9, 1.5
    cgs = setups.batch_blocks(shapes, cfg, seed=seed)
    for cg in cgs:
        ui = cg.interior(cg.u)
        for (i, j, k) in [(1, 2, 3), (5, 0, 4)]:
            for f in range(cfg.nfluids):
                ui[5 * f + 4, i, j, k] = 0.0
        setups.fill_ghosts_periodic(cg)
    return cfg, cgs


def test_floors_list_all_modes(oracle):
    """A two-fluid block list (ionized HLLD + GLM, neutral HLLC) with both
    floors firing in both fluids, in graph, persistent and per-block modes:
    every block bit-exact with the oracle's list driver, flags equal, magic
    mass per fluid within tolerance of the oracle and identical bit for bit
    between the modes."""
    shapes = [(12, 10, 8), (16, 16, 16), (8, 12, 10), (20, 8, 12)]
    fluids = ((5.0 / 3.0, True), (1.4, False))
    ref = None
    for mode in MODES:
        cfg, cgs = floor_list(fluids, shapes, seed=2)
        obs = [oracle.OBlock.from_cg(c) for c in cgs]
        p = oracle.params_from_config(cfg)
        print(f"floors list mode {mode}", flush=True)
        start(cfg, cgs)
        t_g, _ = pk.run_all(2, MODES[mode])
        t_o = oracle.run_list(p, obs, 2)
        assert t_g == t_o
        reps = []
        for cg, ob in zip(cgs, obs):
            cg.download()
            assert_same(cg, ob, f"{mode} block {cg.cg_id}")
            reps.append(assert_report(cg, ob, 2 * 12, f"{mode} block {cg.cg_id}"))
        assert all(r["magic_mass"][0] > 0 and r["magic_mass"][1] > 0 for r in reps)
        if ref is None:
            ref = reps
        else:
            assert reps == ref, mode
        pk.finalize()


def test_floor_mass_independent_of_tiles(oracle):
    """The fused sweep with different tile shapes, the march kernel (one lane
    per pencil along y and z) and the stream kernel gives the same state and
    the same magic-mass bits."""
    out = []
    for tx, tyz in [((1, 64), (4, 64)), ((3, 40), (8, 24)), ((2, 72), (64, -1)), ((1, -4), (4, -3)), ((1, -8), (2, -5))]:
        cg, cfg = fs.floor_block(shape=(24, 20, 16), magnetized=True)
        cfg.tile_x, cfg.tile_yz = tx, tyz
        print(f"floors tiles {tx} {tyz}", flush=True)
        start(cfg, [cg])
        pk.run(cg, 2)
        cg.download()
        out.append((cg.u.copy(), pk.get_report(cg)))
        pk.finalize()
    for u, rep in out[1:]:
        assert np.array_equal(u, out[0][0])
        assert rep == out[0][1]
