"""Parity of the HIP sweep (through the C-ABI) with the oracle.

Bit-exact for every array element (floating-point ==, so +0 and -0 compare
equal): the GPU path keeps the reference's evaluation order and is built
without FMA contraction, and the oracle does the same on the CPU.  Sizes are
chosen so the oracle finishes in seconds; the full 256^3 size is covered by
size-independent properties in test_gpu_properties.py.
"""
import itertools

import numpy as np
import pytest

import piernik_amd as pk
from piernik_amd import setups

pytestmark = pytest.mark.gpu


def setup_lib(cfg):
    pk.load_library()
    pk.finalize()
    pk.init(cfg)


def rel_l1(a, b):
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


def assert_same(cg, ob, where=""):
    for name in ("u", "b", "psi"):
        a, o = getattr(cg, name), getattr(ob, name)
        if a is None:
            continue
        bad = np.argwhere(~(a == o))
        assert bad.size == 0, (f"{where} {name}: {len(bad)} mismatches, first at {tuple(bad[0])}: "
                               f"gpu={a[tuple(bad[0])]!r} oracle={o[tuple(bad[0])]!r} "
                               f"rel-L1={rel_l1(a, o):.3e}")


def make_pair(oracle, shape, magnetized, seed, cfg, nb=4, has_dir=(True, True, True)):
    cg = setups.smooth_random(shape, nb=nb, seed=seed, magnetized=magnetized, gamma=cfg.gamma, has_dir=has_dir)
    ob = oracle.OBlock.from_cg(cg)
    cg.upload()
    return cg, ob


CASES = [(mag, lim) for mag in (False, True) for lim in ("vanleer", "minmod", "moncen", "superbee")]


@pytest.mark.parametrize("magnetized,limiter", CASES)
def test_solve_cg_stages_bit_exact(oracle, magnetized, limiter):
    """pk_solve_cg (one RK stage, whole pencils incl. ghosts) == solve_cg_riemann."""
    cfg = pk.Config(magnetized=magnetized, limiter=limiter, limiter_b=("moncen" if limiter == "minmod" else limiter),
                    smalld=1e-3, smallei=1e-4)
    setup_lib(cfg)
    p = oracle.params_from_config(cfg)
    print(f"case mag={magnetized} lim={limiter}", flush=True)
    cg, ob = make_pair(oracle, (12, 10, 8), magnetized, seed=3, cfg=cfg)
    dt, ch = 2.0e-3, 1.7
    for ddim in (0, 1, 2):
        for istep in (pk.RK2_1, pk.RK2_2):
            pk.solve_cg(cg, ddim, istep, dt, ch)
            oracle.solve_cg(p, ob, ddim, istep, dt, ch)
            cg.download()
            assert_same(cg, ob, f"dir {ddim} stage {istep}")
            # ghost refresh between stages, as update_boundaries
            pk.update_boundaries(cg)
            oracle.fill_ghosts(p, ob)
    pk.finalize()


@pytest.mark.parametrize("magnetized,limiter", [(False, "vanleer"), (True, "minmod"), (True, "vanleer")])
def test_fused_sweep_equals_two_stages(oracle, magnetized, limiter):
    """pk_sweep (both RK2 stages in one pass) == stage 1, ghost refresh, stage 2."""
    cfg = pk.Config(magnetized=magnetized, limiter=limiter, smalld=1e-3, smallei=1e-4)
    setup_lib(cfg)
    p = oracle.params_from_config(cfg)
    cg, ob = make_pair(oracle, (20, 12, 9), magnetized, seed=5, cfg=cfg)
    dt, ch = 3.0e-3, 1.3
    for ddim in (0, 1, 2, 2, 1, 0):
        pk.sweep(cg, ddim, dt, ch)
        pk.update_boundaries(cg)
        for istep in (pk.RK2_1, pk.RK2_2):
            oracle.solve_cg(p, ob, ddim, istep, dt, ch)
            oracle.fill_ghosts(p, ob)
        cg.download()
        assert_same(cg, ob, f"fused dir {ddim}")
    pk.finalize()


@pytest.mark.parametrize("magnetized", [False, True])
def test_timestep_fluid_bit_exact(oracle, magnetized):
    cfg = pk.Config(magnetized=magnetized, cfl=0.6)
    setup_lib(cfg)
    p = oracle.params_from_config(cfg)
    cg, ob = make_pair(oracle, (16, 12, 10), magnetized, seed=7, cfg=cfg)
    dt_g, c_g = pk.timestep_fluid(cg)
    dt_o, c_o = oracle.timestep_fluid(p, ob)
    assert dt_g == dt_o and c_g == c_o
    pk.finalize()


def test_sedov_run_bit_exact(oracle):
    """Sedov blast (HLLC, van Leer, CFL 0.7) at 24^3, 6 steps, whole stepping loop."""
    cg, cfg = setups.sedov_hd(n=24)
    setup_lib(cfg)
    p = oracle.params_from_config(cfg)
    ob = oracle.OBlock.from_cg(cg)
    cg.upload()
    t_g, _ = pk.run(cg, 6)
    t_o = oracle.run(p, ob, 6)
    cg.download()
    assert t_g == t_o
    assert_same(cg, ob, "sedov 24^3")
    pk.finalize()


def test_mhd_turbulence_run_bit_exact(oracle):
    """MHD turbulence (HLLD + GLM, minmod, CFL 0.5) at 24^3, 3 steps."""
    cg, cfg = setups.mhd_turbulence(n=24, seed=11)
    setup_lib(cfg)
    p = oracle.params_from_config(cfg)
    ob = oracle.OBlock.from_cg(cg)
    cg.upload()
    t_g, _ = pk.run(cg, 3)
    t_o = oracle.run(p, ob, 3)
    cg.download()
    assert t_g == t_o
    assert_same(cg, ob, "mhd 24^3")
    pk.finalize()


def test_2d_mhd_bit_exact(oracle):
    """2D block (z inactive: no ghosts, no z sweep), HLLD + GLM."""
    cfg = pk.Config(magnetized=True, limiter="vanleer", has_dir=(True, True, False), smalld=1e-3, smallei=1e-4)
    setup_lib(cfg)
    p = oracle.params_from_config(cfg)
    cg, ob = make_pair(oracle, (16, 12, 1), True, seed=9, cfg=cfg, has_dir=(True, True, False))
    t_g, _ = pk.run(cg, 3)
    t_o = oracle.run(p, ob, 3)
    cg.download()
    assert t_g == t_o
    assert_same(cg, ob, "2d mhd")
    pk.finalize()


def test_autotuned_tiles_keep_bit_exactness(oracle):
    """Tile shapes change the schedule, never the arithmetic."""
    cg, cfg = setups.mhd_turbulence(n=24, seed=13)
    setup_lib(cfg)
    p = oracle.params_from_config(cfg)
    ob = oracle.OBlock.from_cg(cg)
    cg.upload()
    tiles = pk.autotune(cg)
    assert set(tiles) == {0, 1, 2} and all(tp * ts <= 512 for tp, ts in tiles.values())
    t_g, _ = pk.run(cg, 2)
    t_o = oracle.run(p, ob, 2)
    cg.download()
    assert t_g == t_o
    assert_same(cg, ob, "autotuned 24^3")
    pk.finalize()


MARCH = (64, -1)  # tile length -1: one lane per whole pencil (k_march)


@pytest.mark.parametrize("case,tile_x,tile_yz", [
    ("mhd", MARCH, MARCH), ("mhd", (0, 0), MARCH), ("sedov", MARCH, MARCH), ("mhd2d", MARCH, MARCH),
    ("mhd_vanleer_moncen", MARCH, MARCH), ("mhd_long", (0, 0), MARCH), ("sedov_long", MARCH, MARCH)])
def test_march_kernel_bit_exact(oracle, case, tile_x, tile_yz):
    """The march kernel (y and z, and x too when tile_x selects it) gives
    the oracle's bits over whole stepping loops, periodic images and the fused
    timestep included; also on blocks 64 cells long along y and z."""
    if case == "mhd":
        cg, cfg = setups.mhd_turbulence(n=24, seed=17)
        steps = 3
    elif case == "sedov":
        cg, cfg = setups.sedov_hd(n=24)
        steps = 6
    elif case == "mhd_vanleer_moncen":
        cg, cfg = setups.mhd_turbulence(n=20, seed=18)
        cfg.limiter, cfg.limiter_b = "vanleer", "moncen"
        steps = 3
    elif case.endswith("_long"):
        cfg = pk.Config(magnetized=case == "mhd_long", limiter="minmod" if case == "mhd_long" else "vanleer",
                        smalld=1e-3, smallei=1e-4)
        cg = None
        steps = 3
    else:
        cfg = pk.Config(magnetized=True, limiter="vanleer", has_dir=(True, True, False), smalld=1e-3, smallei=1e-4)
        cg = None
        steps = 3
    cfg.tile_x = tile_x
    cfg.tile_yz = tile_yz
    print(f"march {case}, tile_x {tile_x}, tile_yz {tile_yz}", flush=True)
    setup_lib(cfg)
    p = oracle.params_from_config(cfg)
    if cg is None and case.endswith("_long"):
        cg, ob = make_pair(oracle, (8, 64, 64), cfg.magnetized, seed=29, cfg=cfg)
    elif cg is None:
        cg, ob = make_pair(oracle, (16, 12, 1), True, seed=19, cfg=cfg, has_dir=(True, True, False))
    else:
        ob = oracle.OBlock.from_cg(cg)
        cg.upload()
    t_g, _ = pk.run(cg, steps)
    t_o = oracle.run(p, ob, steps)
    cg.download()
    assert t_g == t_o
    assert_same(cg, ob, f"march {case}")
    pk.finalize()


# the stream kernel: (pencils per lane group, -iterations per workgroup)
STREAM_CASES = [("mhd", (1, -8), (4, -8)), ("mhd", (1, -3), (2, -5)), ("mhd", (1, -16), (1, -16)),
                ("sedov", (1, -8), (4, -16)), ("sedov", (1, -2), (2, -2)), ("mhd2d", (1, -4), (4, -4)),
                ("mhd_vanleer_moncen", (1, -8), (4, -8)), ("mhd_odd", (1, -5), (4, -3)),
                ("mhd", (1, -4), (8, -8)), ("sedov", (1, -3), (8, -3)), ("mhd_odd", (1, -2), (8, -5)),
                ("mhd_rounds", (1, -2), (2, -2)), ("mhd_rounds", (1, -3), (8, -64)), ("sedov", (1, -64), (8, -64)),
                ("mhd", (4, -8), (8, -8)), ("mhd_odd", (2, -3), (4, -3)), ("sedov", (8, -4), (2, -4)),
                ("mhd_rounds", (8, -2), (8, -3))]


@pytest.mark.parametrize("case,tile_x,tile_yz", STREAM_CASES)
def test_stream_kernel_bit_exact(oracle, case, tile_x, tile_yz):
    """The stream kernel (one wave walking concatenated pencils, each cell's
    stages computed once) gives the oracle's bits over whole stepping loops,
    periodic images and the fused timestep included; workgroup ranges that
    end inside pencils and short pencils (shorter than a wave) included.
    Ranges are whole pencil groups; mhd_rounds: a 64 x 64 x 96 block with
    more ranges asked for than the device holds at once, so the launch is
    rounded to whole rounds (up along x, down along y) and the ranges differ
    by one group."""
    if case == "mhd":
        cg, cfg = setups.mhd_turbulence(n=24, seed=17)
        steps = 3
    elif case == "sedov":
        cg, cfg = setups.sedov_hd(n=24)
        steps = 6
    elif case == "mhd_vanleer_moncen":
        cg, cfg = setups.mhd_turbulence(n=20, seed=18)
        cfg.limiter, cfg.limiter_b = "vanleer", "moncen"
        steps = 3
    elif case in ("mhd_odd", "mhd_rounds"):
        cfg = pk.Config(magnetized=True, limiter="minmod", smalld=1e-3, smallei=1e-4)
        cg = None
        steps = 3 if case == "mhd_odd" else 2
    else:
        cfg = pk.Config(magnetized=True, limiter="vanleer", has_dir=(True, True, False), smalld=1e-3, smallei=1e-4)
        cg = None
        steps = 3
    cfg.tile_x = tile_x
    cfg.tile_yz = tile_yz
    print(f"stream {case}, tile_x {tile_x}, tile_yz {tile_yz}", flush=True)
    setup_lib(cfg)
    p = oracle.params_from_config(cfg)
    if cg is None and case == "mhd_odd":
        cg, ob = make_pair(oracle, (7, 13, 5), True, seed=23, cfg=cfg)
    elif cg is None and case == "mhd_rounds":
        cg, ob = make_pair(oracle, (64, 64, 96), True, seed=31, cfg=cfg)
    elif cg is None:
        cg, ob = make_pair(oracle, (16, 12, 1), True, seed=19, cfg=cfg, has_dir=(True, True, False))
    else:
        ob = oracle.OBlock.from_cg(cg)
        cg.upload()
    t_g, _ = pk.run(cg, steps)
    t_o = oracle.run(p, ob, steps)
    cg.download()
    assert t_g == t_o
    assert_same(cg, ob, f"stream {case}")
    pk.finalize()
