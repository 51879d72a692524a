"""The FP64 PyTorch transcription (tests/torch_reference.py) against the
oracle: the autograd reference for the adjoint must compute the same forward
map.  Same operations in the same order, so the values agree to rounding of
the last bit (checked to 1e-14 relative) on a smooth moving state."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
import torch_reference as tr  # noqa: E402

import piernik_amd as pk  # noqa: E402


def smooth_state(n, seed, vamp=0.05, gam=5.0 / 3.0):
    rho = tr.gaussian_density(n, seed)
    vel = [vamp * (tr.gaussian_density(n, seed + 10 + d, amp=1.0) - 1.0) for d in range(3)]
    p = torch.ones_like(rho)
    e = p / (gam - 1.0) + 0.5 * rho * (vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)
    return torch.stack([rho, rho * vel[0], rho * vel[1], rho * vel[2], e])


@pytest.mark.parametrize("limiter", ["vanleer", "minmod"])
def test_transcription_matches_oracle(oracle, limiter):
    n, nb, steps = 10, 4, 3
    u = smooth_state(n, 3)
    cfg = pk.Config(limiter=limiter, cfl=0.5, smalld=1e-10, smallei=1e-10)
    p = oracle.params_from_config(cfg)
    ob = oracle.OBlock((n, n, n), nb, (1.0 / n,) * 3, False)
    ob.u[:, nb:-nb, nb:-nb, nb:-nb] = u.numpy()
    t_o = oracle.run(p, ob, steps)
    out = tr.run(u, steps, [1.0 / n] * 3, limiter_name=limiter, cfl=0.5)
    ref = ob.u[:, nb:-nb, nb:-nb, nb:-nb]
    err = np.max(np.abs(out.numpy() - ref)) / np.max(np.abs(ref))
    assert err <= 1e-14, err
    assert t_o > 0
