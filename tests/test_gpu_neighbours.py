"""Blocks connected by neighbour links (pk_set_neighbour) on the GPU.

A periodic domain is cut into uneven blocks (setups.decompose), linked, and
advanced by the block-list functions, which exchange the ghost faces each
sweep reads before that sweep (k_exchange_faces) instead of writing periodic
images.  Decomposition invariance makes the oracle's run of the undivided
domain the reference: every interior cell of every block equals it bit for
bit in graph, persistent and per-block modes, and after a download (which
refreshes every ghost cell from the neighbours) every ghost cell does too.
The oracle's own exchange (or_fill_ghosts_list) is pinned the same way in
tests/test_oracle_neighbours.py.
"""
import numpy as np
import pytest

import piernik_amd as pk
from piernik_amd import setups

pytestmark = pytest.mark.gpu

MODES = {"graph": pk.BATCH_GRAPH, "persistent": pk.BATCH_PERSISTENT, "per_block": pk.BATCH_PER_BLOCK}
ION = (5.0 / 3.0, True)
NEU = (1.4, False)


def _state_of(cg, ob):
    out = pk.GridContainer(cg.n, nb=cg.nb, dl=cg.dl, magnetized=cg.magnetized, has_dir=cg.has_dir,
                           nfluids=cg.nfluids)
    out.u[...] = ob.u
    if cg.magnetized:
        out.b[...] = ob.b
        out.psi[...] = ob.psi
    return out


def make_case(name):
    if name == "mhd3d":
        cfg = pk.Config(magnetized=True, limiter="minmod", cfl=0.5, cfl_glm=0.5, smalld=1e-10, smallei=1e-10)
        cg = setups.smooth_random((24, 20, 16), seed=21, magnetized=True)
        cg.psi[...] = np.random.default_rng(3).normal(0, 0.01, cg.psi.shape)
        return cfg, cg, ([8, 14], [10], [8])
    if name == "hd3d":
        cfg = pk.Config(limiter="vanleer", cfl=0.7, smalld=1e-10, smallei=1e-10)
        cg = setups.smooth_random((20, 16, 24), seed=22)
        return cfg, cg, ([10], [6], [12])
    if name == "mhd2d":
        u, b, psi, cfg = setups.orszag_tang_2d(32, seed=4, perturb=0.1)
        cg = pk.GridContainer((32, 32, 1), nb=4, dl=(1.0 / 32, 1.0 / 32, 1.0), magnetized=True,
                              has_dir=cfg.has_dir)
        cg.interior(cg.u)[...] = u
        cg.interior(cg.b)[...] = b
        cg.interior(cg.psi)[...] = psi
        return cfg, cg, ([12], [8, 20], [])
    if name == "twofluid":
        cfg = setups.batch_config(fluids=(ION, NEU))
        cg = setups.batch_blocks([(20, 16, 12)], cfg, seed=5)[0]
        return cfg, cg, ([6, 12], [8], [6])
    raise ValueError(name)


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("name", ["mhd3d", "hd3d", "mhd2d", "twofluid"])
def test_decomposed_list_equals_undivided_domain(oracle, name, mode):
    steps = 3
    cfg, cg, cuts = make_case(name)
    setups.fill_ghosts_periodic(cg)
    blocks, nbr, boxes = setups.decompose(cg, cuts)
    print(f"{name}: {cg.n} cut into {len(blocks)} blocks {[b.n for b in blocks]}, mode {mode}", flush=True)
    p = oracle.params_from_config(cfg)
    whole = oracle.OBlock.from_cg(cg)
    t_o = oracle.run(p, whole, steps)

    pk.load_library()
    pk.finalize()
    pk.init(cfg)
    pk.connect(blocks, nbr)
    for b in blocks:
        b.upload()
    t_g, _ = pk.run_all(steps, MODES[mode])
    assert t_g == t_o
    for b in blocks:
        b.download()
    names = ["u"] + (["b", "psi"] if cg.magnetized else [])
    for nm in names:
        got = setups.assemble(blocks, boxes, nm, cg.n)
        ref = cg.interior(getattr(whole, nm))
        bad = np.argwhere(~(got == ref))
        assert bad.size == 0, f"{nm}: {len(bad)} cells differ, first {tuple(bad[0])}"
    # ghost cells, refreshed from the neighbours by the download
    refs = setups.decompose(_state_of(cg, whole), cuts)[0]
    for b, r in zip(blocks, refs):
        for nm in names:
            assert np.array_equal(getattr(b, nm), getattr(r, nm)), f"block {b.n} {nm} ghosts"
    pk.finalize()


def test_decomposed_step_by_step_and_timestep(oracle):
    """fluid_update_all / timestep_all over the linked list, one step at a
    time, against the oracle's decomposed driver (or_fluid_update_list_nb)."""
    cfg, cg, cuts = make_case("mhd3d")
    setups.fill_ghosts_periodic(cg)
    blocks, nbr, boxes = setups.decompose(cg, cuts)
    p = oracle.params_from_config(cfg)
    obs = [oracle.OBlock.from_cg(b) for b in blocks]
    oracle.fill_ghosts_list(p, obs, nbr)
    pk.load_library()
    pk.finalize()
    pk.init(cfg)
    pk.connect(blocks, nbr)
    for b in blocks:
        b.upload()
    for step in range(2):
        dt_g, _ = pk.timestep_all()
        dt_o = min(oracle.time_step(p, ob, 0, -1.0)[0] for ob in obs)
        assert dt_g == dt_o
        pk.fluid_update_all(dt_g, pk.BATCH_GRAPH)
        oracle.fluid_update_list_nb(p, obs, nbr, dt_o)
    for b, ob in zip(blocks, obs):
        b.download()
        assert np.array_equal(b.u, ob.u) and np.array_equal(b.b, ob.b) and np.array_equal(b.psi, ob.psi)
    pk.finalize()


def test_linked_blocks_refuse_single_block_stepping():
    cfg, cg, cuts = make_case("hd3d")
    blocks, nbr, boxes = setups.decompose(cg, cuts)
    pk.load_library()
    pk.finalize()
    pk.init(cfg)
    pk.connect(blocks, nbr)
    for b in blocks:
        b.upload()
    with pytest.raises(pk.PiernikError):
        pk.fluid_update(blocks[0], 1e-3)
    odd = pk.GridContainer((8, 8, 8), nb=4, dl=cg.dl)
    odd.register()
    with pytest.raises(pk.PiernikError):
        pk.set_neighbour(blocks[0], 0, 1, odd)  # faces differ (20 vs 16 ghosted cells along y)
    pk.finalize()
