"""Fine/coarse flux points of the split solver in the oracle (set_fluxpointers
and save_outfluxes, grid_container_fcflx.F90:148-235; their use in solve,
solve_cg_riemann.F90:353-363).

Pinned by what the mathematics fixes:
- the flux saved at a face is the pencil's own interface flux there
  (or_pencil_flux on the same pencil);
- replacing a face's flux by the block's own flux there changes nothing, bit
  for bit;
- a coarse block partly covered by a fine block keeps the composite sum of
  every conserved component to rounding when the coarse fluxes at the
  fine/coarse faces are the restricted fine fluxes; without the replacement
  it does not.  A wrong face (row), side, pencil order or restriction weight
  breaks this.
"""
import numpy as np
import pytest

import piernik_amd as pk
import fcflx_setup as fx

EULER, RK2_1, RK2_2 = 10000, 10001, 10002

CASES = {
    "hd": dict(fluids=[(5.0 / 3.0, False)]),
    "mhd": dict(fluids=[(5.0 / 3.0, True)]),
    "ion+neutral": dict(fluids=[(5.0 / 3.0, True), (1.4, False)]),
}


def config(case, order, limiter="vanleer"):
    return pk.Config(limiter=limiter, integration_order=order, smalld=1e-6, smallei=1e-6, **CASES[case])


def run_composite(po, case, d, order, correct=True, limiter="vanleer"):
    cfg = config(case, order, limiter)
    p = po.params_from_config(cfg)
    C, F, fc_c, fc_f = fx.build(po, p, cfg, d)
    before = fx.composite_sums(d, C, F)
    scale = fx.composite_sums(d, fx.Absolute(C), fx.Absolute(F))
    dt, ch = fx.stage_dt(d), 1.0
    steps = (EULER,) if order == 1 else (RK2_1, RK2_2)
    for istep in steps:
        fx.oracle_stage(po, p, d, istep, dt, ch, C, F, fc_c, fc_f, correct)
    after = fx.composite_sums(d, C, F)
    return before, after, scale, C, F


@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("order", [1, 2])
def test_composite_conserves(oracle, d, case, order):
    before, after, scale, _, _ = run_composite(oracle, case, d, order)
    err = np.abs(after - before) / scale
    assert err.max() < 2e-15, f"relative changes {err}"


@pytest.mark.parametrize("d", [0, 1, 2])
def test_composite_without_correction_leaks(oracle, d):
    before, after, scale, _, _ = run_composite(oracle, "mhd", d, 2, correct=False)
    err = np.abs(after - before) / scale
    assert err[0] > 1e-8 and err[4] > 1e-8, f"relative changes {err}"


@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("case", ["hd", "ion+neutral"])
def test_own_flux_replacement_is_identity(oracle, d, case):
    """Save F's boundary fluxes, then redo the stage from the same state with
    those fluxes replacing the fluxes at the same faces (ri at F's first
    cell is the face lo saves, li at its last cell the face ro saves)."""
    po = oracle
    cfg = config(case, 1, "minmod")
    p = po.params_from_config(cfg)
    _, F, _, fc_f = fx.build(po, p, cfg, d)
    u0 = [a.copy() for a in (F.u, F.b, F.psi) if a is not None]
    dt = fx.stage_dt(d)
    po.solve_cg(p, F, d, EULER, dt, 1.0, fc=fc_f)
    ref = [a.copy() for a in (F.u, F.b, F.psi) if a is not None]
    nf = fx.shapes(d)[1]
    fc2 = po.FcFaces(nf, d, F.u.shape[0], cfg.magnetized, fx.FIRST_F)
    fc2.index["fine", 1][...] = fx.FIRST_F               # ri: left face of the first cell
    fc2.index["fine", 0][...] = fx.FIRST_F + nf[d] - 1   # li: right face of the last cell
    for s_own, s_in in ((0, 1), (1, 0)):
        fc2.uflx["fine", s_in][...] = fc_f.uflx["coarse", s_own]
        if cfg.magnetized:
            fc2.bflx["fine", s_in][...] = fc_f.bflx["coarse", s_own]
    for a, v in zip([a for a in (F.u, F.b, F.psi) if a is not None], u0):
        a[...] = v
    po.solve_cg(p, F, d, EULER, dt, 1.0, fc=fc2)
    for a, r in zip([a for a in (F.u, F.b, F.psi) if a is not None], ref):
        assert np.array_equal(a, r)


def test_saved_flux_is_the_pencil_flux(oracle):
    """x sweep (no permutation): lo saves the interface flux at the block's
    left boundary face, ro at its right one, for every pencil."""
    po = oracle
    d = 0
    cfg = config("mhd", 1)
    p = po.params_from_config(cfg)
    _, F, _, fc_f = fx.build(po, p, cfg, d)
    u, b, psi = F.u.copy(), F.b.copy(), F.psi.copy()
    po.solve_cg(p, F, d, EULER, fx.stage_dt(d), 1.0, fc=fc_f)
    nb, n = fx.NB, fx.shapes(d)[1]
    for j in range(n[1]):
        for k in range(n[2]):
            uu = u[:, :, nb + j, nb + k].T
            bb = np.concatenate([b[:, :, nb + j, nb + k], psi[None, :, nb + j, nb + k]]).T
            flx, bflx = po.pencil_flux(p, uu, bb, 1.0)
            lo, hi = nb - 1, nb + n[0] - 1
            assert np.array_equal(fc_f.uflx["coarse", 0][:, j, k], flx[lo])
            assert np.array_equal(fc_f.uflx["coarse", 1][:, j, k], flx[hi])
            assert np.array_equal(fc_f.bflx["coarse", 0][:, j, k], bflx[lo])
            assert np.array_equal(fc_f.bflx["coarse", 1][:, j, k], bflx[hi])


def test_no_faces_is_plain_solve(oracle):
    """All indices outside the interior: identical to or_solve_cg."""
    po = oracle
    d = 1
    cfg = config("mhd", 2)
    p = po.params_from_config(cfg)
    _, F, _, _ = fx.build(po, p, cfg, d)
    G = po.OBlock(F.n, F.nb, F.dl, True)
    for name in ("u", "b", "psi"):
        getattr(G, name)[...] = getattr(F, name)
    empty = po.FcFaces(F.n, d, 5, True, fx.FIRST_F)
    for istep in (RK2_1, RK2_2):
        po.solve_cg(p, F, d, istep, fx.stage_dt(d), 1.0, fc=empty)
        po.solve_cg(p, G, d, istep, fx.stage_dt(d), 1.0)
    for name in ("u", "b", "psi"):
        assert np.array_equal(getattr(F, name), getattr(G, name))
