"""Every pattern of active directions the reference allows (dom%has_dir):
1D along x, y or z and 2D in the xz and yz planes (xy is in
test_gpu_parity.py), hydro and MHD.  Inactive directions have no ghost cells
and no sweep (make_3sweeps skips them, fluidupdate.F90:177-214).  The whole
stepping loop and the drop-in stages are bit-exact with the oracle."""
import numpy as np
import pytest

import piernik_amd as pk
from piernik_amd import setups

pytestmark = pytest.mark.gpu

PATTERNS = {
    "x": (True, False, False),
    "y": (False, True, False),
    "z": (False, False, True),
    "xz": (True, False, True),
    "yz": (False, True, True),
}


def shape_of(has_dir):
    n = (20, 18, 16)
    return tuple(n[d] if has_dir[d] else 1 for d in range(3))


def same(cg, ob, where):
    for name in ("u", "b", "psi"):
        a, o = getattr(cg, name), getattr(ob, name)
        if a is None:
            continue
        bad = np.argwhere(~(a == o))
        assert bad.size == 0, f"{where} {name}: {len(bad)} mismatches, first at {tuple(bad[0])}"


@pytest.mark.parametrize("mag", [False, True])
@pytest.mark.parametrize("pat", list(PATTERNS))
def test_run_bit_exact(oracle, pat, mag):
    has_dir = PATTERNS[pat]
    cfg = pk.Config(magnetized=mag, limiter="vanleer" if mag else "moncen", has_dir=has_dir, smalld=1e-3,
                    smallei=1e-4)
    print(f"dims {pat} mag={mag}", flush=True)
    pk.load_library()
    pk.finalize()
    pk.init(cfg)
    p = oracle.params_from_config(cfg)
    cg = setups.smooth_random(shape_of(has_dir), nb=4, seed=21, magnetized=mag, gamma=cfg.gamma, has_dir=has_dir)
    ob = oracle.OBlock.from_cg(cg)
    cg.upload()
    t_g, _ = pk.run(cg, 3)
    t_o = oracle.run(p, ob, 3)
    cg.download()
    assert t_g == t_o
    same(cg, ob, f"{pat} run")
    # the drop-in: one RK2 sweep of stages along each active direction
    dt, ch = 1e-3, 1.1
    for d in range(3):
        if not has_dir[d]:
            continue
        for istep in (pk.RK2_1, pk.RK2_2):
            pk.solve_cg(cg, d, istep, dt, ch)
            oracle.solve_cg(p, ob, d, istep, dt, ch)
            cg.download()
            same(cg, ob, f"{pat} stage {d}/{istep}")
            pk.update_boundaries(cg)
            oracle.fill_ghosts(p, ob)
    # an inactive direction is refused
    off = [d for d in range(3) if not has_dir[d]][0]
    with pytest.raises(pk.PiernikError, match="inactive"):
        pk.solve_cg(cg, off, pk.RK2_1, dt, ch)
    pk.finalize()
