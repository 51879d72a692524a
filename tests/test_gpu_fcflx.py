"""pk_solve_cg with fine/coarse flux points (pk_register_fcflx) on the GPU.

The two-level composite of fcflx_setup runs stage by stage as the drop-in
adapter would drive it: upload, pk_solve_cg, download per call, with the
reference's host parts (restriction of the fine fluxes, prolongation of the
fine block's ghosts) in numpy between the calls.  Every state, ghost cells
included, and every saved flux must equal the oracle's bit for bit, for the
three sweep directions (the y sweep's face arrays run (z, x)), hydro, MHD
with GLM and two fluids, one Euler stage and both RK2 stages.
"""
import numpy as np
import pytest

import piernik_amd as pk
import fcflx_setup as fx
from test_oracle_fcflx import CASES, EULER, RK2_1, RK2_2, config

pytestmark = pytest.mark.gpu


def gpu_twin(ob, cfg):
    cg = pk.GridContainer(ob.n, ob.nb, ob.dl, cfg.magnetized, nfluids=len(cfg.fluids))
    cg.u[...] = ob.u
    if cfg.magnetized:
        cg.b[...] = ob.b
        cg.psi[...] = ob.psi
    return cg


def faces_copy(po, fc, n, nvar, magnetized):
    out = po.FcFaces(n, fc.ddim, nvar, magnetized, fc.first)
    for key in out.index:
        out.index[key][...] = fc.index[key]
    return out


def register(cg, fc):
    for kind, k in (("fine", pk.FC_FINEBND), ("coarse", pk.FC_COARSEBND)):
        for s in (0, 1):
            pk.register_fcflx(cg, fc.ddim, s, k, fc.first, fc.index[kind, s], fc.uflx[kind, s], fc.bflx[kind, s])


def same(a, b, what):
    bad = np.argwhere(~(a == b))
    assert bad.size == 0, f"{what}: {len(bad)} mismatches, first at {tuple(bad[0])}: {a[tuple(bad[0])]!r} " \
                          f"vs {b[tuple(bad[0])]!r}"


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("d", [0, 1, 2])
def test_composite_stage_bit_exact(oracle, d, case, order):
    po = oracle
    cfg = config(case, order)
    p = po.params_from_config(cfg)
    print(f"fc composite d={d} case={case} order={order}", flush=True)
    oC, oF, ofc_c, ofc_f = fx.build(po, p, cfg, d)
    nc, nf, _, _ = fx.shapes(d)
    nvar = 5 * len(cfg.fluids)
    pk.load_library()
    pk.finalize()
    pk.init(cfg)
    C, F = gpu_twin(oC, cfg), gpu_twin(oF, cfg)
    fc_c = faces_copy(po, ofc_c, nc, nvar, cfg.magnetized)
    fc_f = faces_copy(po, ofc_f, nf, nvar, cfg.magnetized)
    register(C, fc_c)
    register(F, fc_f)
    C.upload()
    before = fx.composite_sums(d, C, F)
    dt, ch = fx.stage_dt(d), 1.0
    for istep in ((EULER,) if order == 1 else (RK2_1, RK2_2)):
        F.upload()
        pk.solve_cg(F, d, istep, dt, ch)
        F.download()
        fx.restrict(fc_f, fc_c)
        pk.solve_cg(C, d, istep, dt, ch)
        pk.update_boundaries(C)
        C.download()
        fx.prolong_ghosts(d, F, C)
        fx.oracle_stage(po, p, d, istep, dt, ch, oC, oF, ofc_c, ofc_f)
        for name in ("u", "b", "psi"):
            if getattr(C, name) is None:
                continue
            same(getattr(C, name), getattr(oC, name), f"coarse {name} after stage {istep}")
            same(getattr(F, name), getattr(oF, name), f"fine {name} after stage {istep}")
        for s in (0, 1):
            same(fc_f.uflx["coarse", s], ofc_f.uflx["coarse", s], f"saved u-fluxes side {s}")
            if cfg.magnetized:
                same(fc_f.bflx["coarse", s], ofc_f.bflx["coarse", s], f"saved [B, psi] fluxes side {s}")
    after = fx.composite_sums(d, C, F)
    scale = fx.composite_sums(d, fx.Absolute(C), fx.Absolute(F))
    assert (np.abs(after - before) / scale).max() < 2e-15
    pk.finalize()


def test_fused_paths_refuse_flux_points(oracle):
    po = oracle
    cfg = config("mhd", 2)
    p = po.params_from_config(cfg)
    _, oF, _, ofc_f = fx.build(po, p, cfg, 0)
    pk.load_library()
    pk.finalize()
    pk.init(cfg)
    F = gpu_twin(oF, cfg)
    register(F, faces_copy(po, ofc_f, fx.shapes(0)[1], 5, True))
    F.upload()
    with pytest.raises(RuntimeError, match="fine/coarse flux points"):
        pk.sweep(F, 0, 1e-3, 1.0)
    with pytest.raises(RuntimeError, match="fine/coarse flux points"):
        pk.run_all(1)
    for s in (0, 1):
        for k in (pk.FC_FINEBND, pk.FC_COARSEBND):
            pk.register_fcflx(F, 0, s, k, 0, None, None, None)
    pk.sweep(F, 0, 1e-3, 1.0)
    pk.finalize()
