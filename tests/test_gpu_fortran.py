"""The ISO_C_BINDING boundary driven from Fortran on the GPU.

tests/fortran/drive_solve_cg.F90 is a Fortran host program built with
amdflang against piernik_amd/fortran/piernik_hip.F90 (by build()).  It does
what the solve_cg_hip adapter does inside PIERNIK: registers the block's
Fortran arrays, then per call of the solve_cg_sub seam (sweeps.F90:140-151)
pk_upload -> pk_solve_cg -> pk_download, and refreshes the periodic ghosts
on the host after every stage as update_boundaries does (sweeps.F90:271).
Three directions x two RK2 stages of a 12x10x8 MHD block, and of a 20x12x10
MHD block on which both floors fire: every element of u, b and psi equals
the oracle's solve_cg/fill_ghosts sequence, bit for bit, and the floor
bookkeeping (flags, magic mass) agrees.
"""
import os
import subprocess

import numpy as np
import pytest

import piernik_amd as pk
from piernik_amd import setups

from floors_setup import floor_block, mass_tolerance

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "fortran", "drive_solve_cg")
LIM = {"vanleer": 0, "minmod": 1, "moncen": 2, "superbee": 3}


def run_driver(tmp_path, cg, cfg, dt, ch):
    if not os.path.exists(DRIVER):
        raise FileNotFoundError(f"{DRIVER} is not built: run __graft_entry__.build()")
    hdr = np.array(list(cg.n) + [cg.nb, LIM[cfg.limiter], LIM[cfg.limiter_b]], dtype=np.int32)
    par = np.array([cfg.gamma, cfg.smalld, cfg.smallei, dt, ch], dtype=np.float64)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(hdr.tobytes())
        f.write(par.tobytes())
        for a in (cg.u, cg.b, cg.psi):
            f.write(np.asarray(a, dtype=np.float64).tobytes(order="F"))
    r = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = fout.read_bytes()
    rcs = np.frombuffer(raw[:80], dtype=np.int32)
    flags = np.frombuffer(raw[80:88], dtype=np.int32)
    mm = np.frombuffer(raw[88:104], dtype=np.float64)
    data = np.frombuffer(raw[104:], dtype=np.float64)
    ng = cg.ng
    nu, nbb = 5 * int(np.prod(ng)), 3 * int(np.prod(ng))
    assert data.size == nu + nbb + int(np.prod(ng))
    u = data[:nu].reshape((5,) + ng, order="F")
    b = data[nu:nu + nbb].reshape((3,) + ng, order="F")
    psi = data[nu + nbb:].reshape(ng, order="F")
    return rcs, flags, mm, u, b, psi


def oracle_sequence(oracle, cg, cfg, dt, ch):
    p = oracle.params_from_config(cfg)
    ob = oracle.OBlock.from_cg(cg)
    for ddim in (0, 1, 2):
        for istep in (pk.RK2_1, pk.RK2_2):
            oracle.solve_cg(p, ob, ddim, istep, dt, ch)
            oracle.fill_ghosts(p, ob)
    return ob


def assert_equal(name, a, o):
    bad = np.argwhere(~(a == o))
    assert bad.size == 0, (f"{name}: {len(bad)} mismatches, first at {tuple(bad[0])}: "
                           f"fortran={a[tuple(bad[0])]!r} oracle={o[tuple(bad[0])]!r}")


@pytest.mark.parametrize("limiter,limiter_b", [("minmod", "moncen"), ("vanleer", "vanleer")])
def test_fortran_drop_in_stages_bit_exact(oracle, tmp_path, limiter, limiter_b):
    cfg = pk.Config(magnetized=True, limiter=limiter, limiter_b=limiter_b, smalld=1e-3, smallei=1e-4)
    cg = setups.smooth_random((12, 10, 8), nb=4, seed=11, magnetized=True, gamma=cfg.gamma)
    dt, ch = 2.0e-3, 1.7
    print(f"fortran driver, 12x10x8 MHD, {limiter}/{limiter_b}", flush=True)
    rcs, flags, mm, u, b, psi = run_driver(tmp_path, cg, cfg, dt, ch)
    assert list(rcs) == [0] * 20
    ob = oracle_sequence(oracle, cg, cfg, dt, ch)
    assert_equal("u", u, ob.u)
    assert_equal("b", b, ob.b)
    assert_equal("psi", psi, ob.psi)
    assert list(flags) == [0, 0] and list(mm) == [0.0, 0.0]


def test_fortran_drop_in_floors(oracle, tmp_path):
    """The same with both floors firing (smalld = 0.9, smallei = 1.5, three
    cells of zero total energy): state bit for bit, flags equal, magic mass
    within the rounding of the reference's per-pencil form."""
    cg, cfg = floor_block(shape=(20, 12, 10), magnetized=True, seed=2)
    dt, ch = 1.0e-3, 1.1
    print("fortran driver, floors", flush=True)
    rcs, flags, mm, u, b, psi = run_driver(tmp_path, cg, cfg, dt, ch)
    assert list(rcs) == [0] * 20
    ob = oracle_sequence(oracle, cg, cfg, dt, ch)
    assert_equal("u", u, ob.u)
    assert_equal("b", b, ob.b)
    assert_equal("psi", psi, ob.psi)
    assert int(flags[0]) == int(ob.rep.dn_negative) and int(flags[1]) == int(ob.rep.ei_negative)
    assert ob.rep.magic_mass[0] > 0.0
    assert abs(mm[0] - ob.rep.magic_mass[0]) <= mass_tolerance(cg, 6)
