"""The block-list path (pk_*_all) against the oracle's list driver.

Bit-exact for every array element: the persistent list kernel, the graph
replay and the per-block launches run the same arithmetic as the
single-block sweep, which is bit-exact against the oracle.  Blocks are
uneven (non-cubic shapes, different cell sizes) and carry an ionized MHD
fluid and a neutral HLLC fluid with different gammas.
"""
import numpy as np
import pytest

import piernik_amd as pk
from piernik_amd import setups

pytestmark = pytest.mark.gpu

MODES = {"graph": pk.BATCH_GRAPH, "persistent": pk.BATCH_PERSISTENT, "per_block": pk.BATCH_PER_BLOCK}
SHAPES = [(16, 16, 16), (8, 12, 10), (24, 8, 8), (12, 20, 16), (6, 6, 6)]
ION = (5.0 / 3.0, True)
NEU = (1.4, False)


def start(cfg, cgs):
    pk.load_library()
    pk.finalize()
    pk.init(cfg)
    for cg in cgs:
        cg.registered = False  # pk_finalize dropped every registration
        cg.upload()


def assert_same(cg, ob, where=""):
    for name in ("u", "b", "psi"):
        a, o = getattr(cg, name), getattr(ob, name)
        if a is None:
            continue
        bad = np.argwhere(~(a == o))
        assert bad.size == 0, (f"{where} block {cg.cg_id} {cg.n} {name}: {len(bad)} mismatches, first at "
                               f"{tuple(bad[0])}: gpu={a[tuple(bad[0])]!r} oracle={o[tuple(bad[0])]!r}")


@pytest.mark.parametrize("mode", list(MODES))
def test_two_fluid_list_bit_exact(oracle, mode):
    cfg = setups.batch_config(fluids=(ION, NEU))
    cgs = setups.batch_blocks(SHAPES, cfg, seed=9)
    obs = [oracle.OBlock.from_cg(c) for c in cgs]
    p = oracle.params_from_config(cfg)
    print(f"two-fluid list, mode {mode}, shapes {SHAPES}", flush=True)
    start(cfg, cgs)
    t_g, _ = pk.run_all(3, MODES[mode])
    t_o = oracle.run_list(p, obs, 3)
    assert t_g == t_o
    for cg, ob in zip(cgs, obs):
        cg.download()
        assert_same(cg, ob, mode)
        rep = pk.get_report(cg)
        assert rep["magic_mass"] == list(ob.rep.magic_mass)
    pk.finalize()


@pytest.mark.parametrize("fluid,limiter", [(ION, "minmod"), (NEU, "vanleer"), (ION, "superbee")])
def test_one_fluid_list_bit_exact(oracle, fluid, limiter):
    cfg = setups.batch_config(fluids=(fluid,), limiter=limiter)
    cgs = setups.batch_blocks(SHAPES[:4], cfg, seed=3)
    obs = [oracle.OBlock.from_cg(c) for c in cgs]
    p = oracle.params_from_config(cfg)
    print(f"one-fluid list {fluid} {limiter}", flush=True)
    start(cfg, cgs)
    t_g, _ = pk.run_all(2)
    t_o = oracle.run_list(p, obs, 2)
    assert t_g == t_o
    for cg, ob in zip(cgs, obs):
        cg.download()
        assert_same(cg, ob, f"{fluid} {limiter}")
    pk.finalize()


def test_2d_two_fluid_list_bit_exact(oracle):
    cfg = setups.batch_config(fluids=(ION, NEU), has_dir=(True, True, False))
    cgs = setups.batch_blocks([(16, 24, 1), (12, 8, 1), (32, 16, 1)], cfg, seed=4)
    obs = [oracle.OBlock.from_cg(c) for c in cgs]
    p = oracle.params_from_config(cfg)
    print("2d two-fluid list", flush=True)
    start(cfg, cgs)
    t_g, _ = pk.run_all(3)
    t_o = oracle.run_list(p, obs, 3)
    assert t_g == t_o
    for cg, ob in zip(cgs, obs):
        cg.download()
        assert_same(cg, ob, "2d")
    pk.finalize()


def test_list_calls_continue_and_timestep(oracle):
    """run_all twice (the second call starts from the dt and ghosts the first
    left on the device), then fluid_update_all with a given dt, and
    timestep_all, each against the oracle."""
    cfg = setups.batch_config(fluids=(ION, NEU))
    cgs = setups.batch_blocks(SHAPES[:3], cfg, seed=12)
    obs = [oracle.OBlock.from_cg(c) for c in cgs]
    p = oracle.params_from_config(cfg)
    start(cfg, cgs)
    dt0, _ = pk.timestep_all()
    for ob in obs:
        oracle.fill_ghosts(p, ob)
    assert dt0 == min(oracle.time_step(p, ob, 0, -1.0)[0] for ob in obs)
    pk.run_all(2)
    oracle.run_list(p, obs, 2)
    t_g, _ = pk.run_all(1)
    t_o = oracle.run_list(p, obs, 1)
    assert t_g == t_o
    dt = 0.5 * pk.timestep_all()[0]
    pk.fluid_update_all(dt, pk.BATCH_PERSISTENT)
    oracle.fluid_update_list(p, obs, dt)
    for cg, ob in zip(cgs, obs):
        cg.download()
        assert_same(cg, ob, "continued")
    pk.finalize()


def test_autotuned_list_bit_exact(oracle):
    cfg = setups.batch_config(fluids=(ION, NEU))
    cgs = setups.batch_blocks(SHAPES, cfg, seed=21)
    obs = [oracle.OBlock.from_cg(c) for c in cgs]
    p = oracle.params_from_config(cfg)
    start(cfg, cgs)
    pk.autotune_all()
    seen = set()
    for cg in cgs:
        for d in range(3):
            for f in range(2):
                lanes, tp, ts = pk.list_tiles(d, f, cg)
                if lanes == 64:  # the stream form: pencils per lane group, -iterations per item
                    assert tp == (1 if d == 0 else 8) and ts <= -2
                else:
                    assert lanes in (128, 256, 512) and 1 <= tp and tp * ts <= lanes
                seen.add(lanes)
    t_g, _ = pk.run_all(2)
    t_o = oracle.run_list(p, obs, 2)
    assert t_g == t_o
    for cg, ob in zip(cgs, obs):
        cg.download()
        assert_same(cg, ob, "autotuned")
    pk.finalize()


@pytest.mark.parametrize("ts", [-2, -5, -16])
def test_stream_list_bit_exact(oracle, ts):
    """The stream form of the list sweep (one-wave workgroups over ranges
    of each block's pencils, pinned with set_list_tiles for every direction
    and fluid) equals the oracle's list driver bit for bit after two steps,
    in graph, persistent and per-block modes; items that end inside pencils
    and blocks shorter than one item included."""
    cfg = setups.batch_config(fluids=(ION, NEU))
    cgs = setups.batch_blocks(SHAPES + [(32, 24, 20)], cfg, seed=33)
    init = [(cg.u.copy(order="F"), cg.b.copy(order="F"), cg.psi.copy(order="F")) for cg in cgs]
    obs = [oracle.OBlock.from_cg(c) for c in cgs]
    p = oracle.params_from_config(cfg)
    t_o = oracle.run_list(p, obs, 2)
    for mode in MODES:
        for cg, (u, b, psi) in zip(cgs, init):
            cg.u[...] = u
            cg.b[...] = b
            cg.psi[...] = psi
        print(f"stream list, {-ts} iterations per item, mode {mode}", flush=True)
        start(cfg, cgs)
        for d in range(3):
            for f in range(2):
                pk.set_list_tiles(d, f, 64, ts)
        assert pk.list_tiles(0, 0, cgs[0]) == (64, 1, ts)
        assert pk.list_tiles(2, 1, cgs[-1]) == (64, 8, ts)
        t_g, _ = pk.run_all(2, MODES[mode])
        assert t_g == t_o
        for cg, ob in zip(cgs, obs):
            cg.download()
            assert_same(cg, ob, f"stream list {mode}")
        pk.finalize()


def test_stream_list_calls_mix_modes(oracle):
    """Blocks whose every sweep is in the stream form read their ghost
    positions from the interior and leave the ghosts old between calls; the
    next call that reads ghosts refills them first.  Graph, per-block
    (which reads ghosts), persistent and single-step calls in a row, then a
    download, against the oracle's list driver, bit for bit."""
    cfg = setups.batch_config(fluids=(ION, NEU))
    cgs = setups.batch_blocks(SHAPES[:4], cfg, seed=14)
    obs = [oracle.OBlock.from_cg(c) for c in cgs]
    p = oracle.params_from_config(cfg)
    start(cfg, cgs)
    for d in range(3):
        for f in range(2):
            pk.set_list_tiles(d, f, 64, -4)
    t_g = t_o = 0.0
    for mode in (pk.BATCH_GRAPH, pk.BATCH_PER_BLOCK, pk.BATCH_PERSISTENT, pk.BATCH_GRAPH):
        print(f"stream list, one step in mode {mode}", flush=True)
        t_g += pk.run_all(1, mode)[0]
        t_o += oracle.run_list(p, obs, 1)
    assert t_g == t_o
    dt = 0.5 * pk.timestep_all()[0]
    pk.fluid_update_all(dt, pk.BATCH_GRAPH)
    oracle.fluid_update_list(p, obs, dt)
    for cg, ob in zip(cgs, obs):
        cg.download()
        assert_same(cg, ob, "mixed modes")
    pk.finalize()


def test_split_list_bit_exact(oracle):
    """A list split between two launches per sweep (set_list_split: the
    16^3 and 32x24x20 blocks by the stream form, the others by 256-lane
    tiles) equals the oracle's list driver bit for bit after two steps, in
    every mode."""
    cfg = setups.batch_config(fluids=(ION, NEU))
    cgs = setups.batch_blocks(SHAPES + [(32, 24, 20)], cfg, seed=35)
    init = [(cg.u.copy(order="F"), cg.b.copy(order="F"), cg.psi.copy(order="F")) for cg in cgs]
    obs = [oracle.OBlock.from_cg(c) for c in cgs]
    p = oracle.params_from_config(cfg)
    t_o = oracle.run_list(p, obs, 2)
    split = [(16, 16, 16), (32, 24, 20)]
    for mode in MODES:
        for cg, (u, b, psi) in zip(cgs, init):
            cg.u[...] = u
            cg.b[...] = b
            cg.psi[...] = psi
        print(f"split list, mode {mode}", flush=True)
        start(cfg, cgs)
        for d in range(3):
            for f in range(2):
                pk.set_list_tiles(d, f, 256, 0)
                pk.set_list_split(d, f, split, -4)
        for cg in cgs:
            lanes, tp, ts = pk.list_tiles(1, 0, cg)
            if tuple(cg.n) in split:
                assert (lanes, tp, ts) == (64, 8, -4)
            else:
                assert lanes == 256
        t_g, _ = pk.run_all(2, MODES[mode])
        assert t_g == t_o
        for cg, ob in zip(cgs, obs):
            cg.download()
            assert_same(cg, ob, f"split list {mode}")
        pk.finalize()


def test_set_list_tiles_rejects_bad_shapes():
    """set_list_tiles refuses workgroup sizes other than 64, 128, 256 and
    512, stream items shorter than two iterations and tiles longer than the
    workgroup."""
    cfg = setups.batch_config(fluids=(ION, NEU))
    cgs = setups.batch_blocks(SHAPES[:2], cfg, seed=3)
    start(cfg, cgs)
    for lanes, ts in ((96, 0), (64, -1), (64, 20), (128, 200), (256, 8)):
        with pytest.raises(pk.PiernikError):
            pk.set_list_tiles(1, 0, lanes, ts)
    pk.set_list_tiles(1, 0, 256, 0)
    assert pk.list_tiles(1, 0, cgs[0])[0] == 256
    pk.finalize()


def test_device_generated_state_matches_host():
    """batch_blocks_device fills the mirrors from GPU tensors through
    pk_upload_device; the state equals the host generator's to rounding of
    sin (torch vs numpy), and pk_download_device returns it unchanged."""
    import torch

    cfg = setups.batch_config(fluids=(ION, NEU))
    sizes = [16, 24]
    host = setups.batch_blocks(sizes, cfg, seed=5, first_id=100)
    pk.load_library()
    pk.finalize()
    pk.init(cfg)
    dev = setups.batch_blocks_device(sizes, cfg, seed=5, first_id=1)
    for h, d in zip(host, dev):
        ng = d.ng
        u = torch.empty((ng[2], ng[1], ng[0], 10), dtype=torch.float64, device="cuda")
        b = torch.empty((ng[2], ng[1], ng[0], 3), dtype=torch.float64, device="cuda")
        psi = torch.empty((ng[2], ng[1], ng[0]), dtype=torch.float64, device="cuda")
        d.download_device(u, b, psi)
        uh = u.cpu().numpy().transpose(3, 2, 1, 0)
        bh = b.cpu().numpy().transpose(3, 2, 1, 0)
        assert np.max(np.abs(uh - h.u)) <= 1e-13
        assert np.max(np.abs(bh - h.b)) <= 1e-13
        assert float(psi.abs().max()) == 0.0
    pk.finalize()


def test_config_block_sizes_all_modes_bit_exact(oracle):
    """A two-fluid list holding every block size of the batched workload
    (interiors 16^3, 24^3, 32^3, 48^3, 64^3, and a second 48^3), autotuned
    list tiles: graph, persistent and per-block modes each equal the
    oracle's list driver, bit for bit, after two steps."""
    cfg = setups.batch_config()
    shapes = [16, 24, 32, 48, 64, 48]
    cgs = setups.batch_blocks(shapes, cfg, seed=31)
    init = [(cg.u.copy(order="F"), cg.b.copy(order="F"), cg.psi.copy(order="F")) for cg in cgs]
    obs = [oracle.OBlock.from_cg(c) for c in cgs]
    p = oracle.params_from_config(cfg)
    t_o = oracle.run_list(p, obs, 2)
    for mode in MODES:
        for cg, (u, b, psi) in zip(cgs, init):
            cg.u[...] = u
            cg.b[...] = b
            cg.psi[...] = psi
        print(f"config block sizes {shapes}, mode {mode}", flush=True)
        start(cfg, cgs)
        pk.autotune_all()
        t_g, _ = pk.run_all(2, MODES[mode])
        assert t_g == t_o
        for cg, ob in zip(cgs, obs):
            cg.download()
            assert_same(cg, ob, mode)
            rep = pk.get_report(cg)
            assert rep["magic_mass"] == list(ob.rep.magic_mass)
        pk.finalize()


def _device_state(cg, nvar):
    import torch

    ng = cg.ng
    u = torch.empty((ng[2], ng[1], ng[0], nvar), dtype=torch.float64, device="cuda")
    b = torch.empty((ng[2], ng[1], ng[0], 3), dtype=torch.float64, device="cuda")
    psi = torch.empty((ng[2], ng[1], ng[0]), dtype=torch.float64, device="cuda")
    cg.download_device(u, b, psi)
    return u, b, psi


def test_config_list_512_blocks_modes_agree():
    """512 blocks drawn like the batched workload (edges 16..64, seeded, two
    fluids, generated on the GPU), autotuned: graph replay, persistent
    launches and one launch per block give the same bits after two steps,
    every block conserves its mass, momentum, energy and B to rounding and
    stays positive."""
    import torch

    cfg = setups.batch_config()
    sizes = setups.batch_sizes(512, seed=0)
    assert sorted(set(sizes)) == list(setups.BATCH_SIZES)
    print(f"512-block list, {sum(n ** 3 for n in sizes)} cells", flush=True)
    pk.load_library()
    outs = {}
    init_tot = None
    for name, mode in MODES.items():
        pk.finalize()
        pk.init(cfg)
        cgs = setups.batch_blocks_device(sizes, cfg, seed=0)
        if init_tot is None:
            init_tot = []
            for cg in cgs:
                u, b, _ = _device_state(cg, 10)
                ui, bi = setups.interior_of(u, 4, cfg.has_dir), setups.interior_of(b, 4, cfg.has_dir)
                # interior_of gives (component, x, y, z): one total per component
                init_tot.append((torch.cat([ui.sum(dim=(1, 2, 3)), bi.sum(dim=(1, 2, 3))]),
                                 torch.cat([ui.abs().sum(dim=(1, 2, 3)), bi.abs().sum(dim=(1, 2, 3))])))
        pk.autotune_all()
        t, dt = pk.run_all(2, mode)
        print(f"mode {name}: t = {t!r}", flush=True)
        outs[name] = (t, [_device_state(cg, 10) for cg in cgs])
    pk.finalize()
    t_ref, ref = outs["graph"]
    for name in ("persistent", "per_block"):
        t, st = outs[name]
        assert t == t_ref
        for i, (a, c) in enumerate(zip(st, ref)):
            for x, y in zip(a, c):
                assert torch.equal(x, y), f"block {i} ({sizes[i]}^3): {name} differs from graph"
    for i, (u, b, psi) in enumerate(ref):
        ui, bi = setups.interior_of(u, 4, cfg.has_dir), setups.interior_of(b, 4, cfg.has_dir)
        tot = torch.cat([ui.sum(dim=(1, 2, 3)), bi.sum(dim=(1, 2, 3))])
        t0, m0 = init_tot[i]
        err = float(torch.max(torch.abs(tot - t0) / m0))
        assert err <= 1e-13, f"block {i}: conservation {err:.3e}"
        assert float(ui[0].min()) > 0.0 and float(ui[5].min()) > 0.0
