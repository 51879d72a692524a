"""Ensembles of independent instances (pk_ensemble_*): every instance with
its own time step, all advanced by the same persistent launches.

- forward: each instance equals a single-block run of the same problem,
  bitwise (same kernels, same arithmetic);
- reverse: the ensemble adjoint (one persistent launch per sweep over all
  instances) equals the single-block adjoint of each instance (itself
  checked against finite differences) to 1e-12; only the order of the
  dL/d(dt) partial sums differs.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import piernik_amd as pk  # noqa: E402
from piernik_amd import adjoint, setups  # noqa: E402

pytestmark = pytest.mark.gpu


def ot_cases():
    cases = []
    for i, (n, amp) in enumerate([(24, 1.05), (32, 0.93), (16, 1.0)]):
        u, b, psi, cfg = setups.orszag_tang_2d(n, seed=10 + i, amp=amp, perturb=0.05)
        cases.append((n, u, b, psi))
    return cases, cfg


def dev(a, hd):
    return setups.to_device_layout(a, 4, hd)


@pytest.mark.parametrize("mode", [pk.BATCH_GRAPH, pk.BATCH_PER_BLOCK])
def test_ensemble_matches_single_instances(mode):
    cases, cfg = ot_cases()
    hd = cfg.has_dir
    steps, k = 5, 2
    # dL/db of 0.5*|B|^2 is B on the interior and 0 on ghost cells
    def grad_of(st):
        g = [torch.zeros_like(st[0]), torch.zeros_like(st[1]), torch.zeros_like(st[2])]
        gi = setups.interior_of(g[1], 4, hd)
        gi.copy_(setups.interior_of(st[1], 4, hd))
        return g

    print(f"ensemble of {len(cases)} Orszag-Tang instances, mode {mode}", flush=True)
    pk.load_library()
    pk.finalize()
    pk.init(cfg)
    ens = adjoint.EnsemblePlan(cfg, [(n, n, 1) for n, *_ in cases], steps, checkpoint_every=k)
    states = [(dev(u, hd), dev(b, hd), dev(psi, hd)) for _, u, b, psi in cases]
    fin = ens.run(states, mode=mode)
    grads = ens.backward([grad_of(f) for f in fin])
    ens_dts = ens.dts
    ens.close()
    for i, (n, u, b, psi) in enumerate(cases):
        pk.finalize()
        pk.init(cfg)
        plan = adjoint.RolloutPlan(cfg, (n, n, 1), steps, checkpoint_every=k)
        st = (dev(u, hd), dev(b, hd), dev(psi, hd))
        f1 = plan.run(st)
        assert [d[i] for d in ens_dts] == plan.dts
        for a, c in zip(f1, fin[i]):
            assert torch.equal(a, c), f"instance {i}: forward differs"
        g1 = plan.backward(grad_of(f1))
        for a, c in zip(g1, grads[i]):
            err = float(torch.max(torch.abs(a - c))) / max(float(torch.max(torch.abs(a))), 1e-300)
            print(f"instance {i} (n={n}): max rel gradient difference {err:.3e}", flush=True)
            assert err <= 1e-12
        plan.close()
    pk.finalize()
