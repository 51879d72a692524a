"""The relaxed-arithmetic build of the sweep (pk_config.fast_math: FMA
contraction, fast division and sqrt) against the oracle and the exact build.

Tolerances (relative L1 per conserved variable, sum|a-b| / sum|b|):
  <= 1e-12 against the oracle on small blocks after a few steps,
  <= 1e-10 against the exact GPU path (itself bit-identical to the oracle)
  on the full 256^3 MHD turbulence block after its 50 steps -- the bar the
  sweep must meet against PIERNIK's FP64 path.
"""
import numpy as np
import pytest

import piernik_amd as pk
from piernik_amd import setups

pytestmark = pytest.mark.gpu

VARS = ["rho", "m_x", "m_y", "m_z", "E", "B_x", "B_y", "B_z"]


def rel_l1(a, b):
    return float(np.abs(a - b).sum() / np.abs(b).sum())


def conserved(cg):
    ui = cg.interior(cg.u)
    out = [ui[k] for k in range(5)]
    if cg.magnetized:
        bi = cg.interior(cg.b)
        out += [bi[k] for k in range(3)]
    return out


def run_gpu(cg0, cfg, steps, fast):
    cfg.fast_math = fast
    pk.load_library()
    pk.finalize()
    pk.init(cfg)
    cg = pk.GridContainer(cg0.n, magnetized=cg0.magnetized)
    cg.u[...] = cg0.u
    if cg0.magnetized:
        cg.b[...] = cg0.b
        cg.psi[...] = cg0.psi
    cg.upload()
    t, _ = pk.run(cg, steps)
    cg.download()
    pk.finalize()
    return cg, t


@pytest.mark.parametrize("case", ["sedov24", "mhd24"])
def test_fast_math_vs_oracle(oracle, case):
    if case == "sedov24":
        cg0, cfg = setups.sedov_hd(n=24)
        steps = 6
    else:
        cg0, cfg = setups.mhd_turbulence(n=24, seed=11)
        steps = 3
    p = oracle.params_from_config(cfg)
    ob = oracle.OBlock.from_cg(cg0)
    t_o = oracle.run(p, ob, steps)
    cg, t_g = run_gpu(cg0, cfg, steps, fast=True)
    assert abs(t_g - t_o) <= 1e-12 * t_o
    ref = pk.GridContainer(cg0.n, magnetized=cg0.magnetized)
    ref.u[...] = ob.u
    if cg0.magnetized:
        ref.b[...] = ob.b
    for name, a, b in zip(VARS, conserved(cg), conserved(ref)):
        err = rel_l1(a, b)
        print(f"{case} {name} rel-L1 {err:.3e}")
        assert err <= 1e-12, (name, err)


def test_fast_math_mhd256_50_steps_within_1e10_of_exact():
    cg0, cfg = setups.mhd_turbulence(n=256, seed=1)
    print("mhd256 exact, 50 steps", flush=True)
    ex, t_e = run_gpu(cg0, cfg, 50, fast=False)
    print("mhd256 fast, 50 steps", flush=True)
    fa, t_f = run_gpu(cg0, cfg, 50, fast=True)
    assert abs(t_f - t_e) <= 1e-12 * t_e
    worst = 0.0
    for name, a, b in zip(VARS, conserved(fa), conserved(ex)):
        err = rel_l1(a, b)
        worst = max(worst, err)
        print(f"mhd256 {name} rel-L1 {err:.3e}", flush=True)
        assert err <= 1e-10, (name, err)
    ui = fa.interior(fa.u)
    assert ui[0].min() > 0.0
    print(f"worst rel-L1 {worst:.3e}")
