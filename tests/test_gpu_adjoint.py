"""The hand-written adjoint of the hydro rollout (piernik_amd.adjoint) on
the GPU, against

- autograd of the FP64 PyTorch transcription of the same scheme
  (tests/torch_reference.py, itself checked against the oracle), on a small
  grid: gradients agree to 1e-10 relative;
- FP64 central finite differences of the GPU forward map with eps = 1e-6 on
  64 random unit directions at 32^3, 32 steps, checkpoint every 8: each
  directional derivative agrees to 1e-6 relative.

The loss is 0.5*||rho_final - rho_target||^2 over interior cells.  The
finite-difference states are generic (v != 0): at v = 0 exactly the
limited slopes of the velocity sit at the kink of the limiter, where the
map has no derivative in momentum directions; the rest state is checked in
the directions of density and energy, where it is differentiable.
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import piernik_amd as pk  # noqa: E402
from piernik_amd import adjoint  # noqa: E402
import torch_reference as tr  # noqa: E402

pytestmark = pytest.mark.gpu
GAM = 5.0 / 3.0


def interior_state(n, seed, vamp):
    rho = tr.gaussian_density(n, seed)
    vel = [vamp * (tr.gaussian_density(n, seed + 10 + d, amp=1.0) - 1.0) for d in range(3)]
    p = torch.ones_like(rho)
    e = p / (GAM - 1.0) + 0.5 * rho * (vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)
    return torch.stack([rho, rho * vel[0], rho * vel[1], rho * vel[2], e])  # (5, nx, ny, nz)


def to_device_layout(u, nb):
    """(5, nx, ny, nz) interior -> (nzg, nyg, nxg, 5) with periodic ghosts."""
    full = u
    for ax in (1, 2, 3):
        full = torch.cat([full.narrow(ax, full.shape[ax] - nb, nb), full, full.narrow(ax, 0, nb)], dim=ax)
    return full.permute(3, 2, 1, 0).contiguous().cuda()


def interior_of(t, nb):
    """(nzg, nyg, nxg, 5) -> (5, nx, ny, nz) interior."""
    return t[nb:-nb, nb:-nb, nb:-nb, :].permute(3, 2, 1, 0)


def make_plan(n, nsteps, k, limiter="vanleer"):
    cfg = pk.Config(limiter=limiter, cfl=0.5, smalld=1e-10, smallei=1e-10)
    pk.load_library()
    pk.finalize()
    pk.init(cfg)
    return adjoint.RolloutPlan(cfg, n, nsteps, checkpoint_every=k)


def loss_grad(plan, u0, rho_t, nb):
    u0 = u0.clone().requires_grad_(True)
    uf = adjoint.rollout(u0, plan)
    rf = interior_of(uf, nb)[0]
    loss = 0.5 * torch.sum((rf - rho_t) ** 2)
    loss.backward()
    return float(loss), u0.grad


@pytest.mark.parametrize("limiter", ["vanleer", "minmod"])
def test_adjoint_matches_autograd_transcription(limiter):
    n, nb, steps = 10, 4, 3
    u = interior_state(n, 5, 0.05)
    rho_t = tr.gaussian_density(n, 6)
    print(f"adjoint vs autograd, {n}^3, {steps} steps, {limiter}", flush=True)
    plan = make_plan(n, steps, 2, limiter)
    _, g = loss_grad(plan, to_device_layout(u, nb), rho_t.cuda(), nb)
    g = interior_of(g, nb).cpu()
    ur = u.clone().requires_grad_(True)
    out = tr.run(ur, steps, [1.0 / n] * 3, limiter_name=limiter, cfl=0.5)
    (0.5 * torch.sum((out[0] - rho_t) ** 2)).backward()
    err = float(torch.max(torch.abs(g - ur.grad)) / torch.max(torch.abs(ur.grad)))
    print(f"max rel diff {err:.3e}")
    assert err <= 1e-10, err
    plan.close()
    pk.finalize()


def fd_check(n, steps, k, vamp, ndir, variables, seed):
    nb = 4
    u = interior_state(n, seed, vamp)
    rho_t = tr.gaussian_density(n, seed + 1).cuda()
    plan = make_plan(n, steps, k)
    u0 = to_device_layout(u, nb)
    _, g = loss_grad(plan, u0, rho_t, nb)
    g_int = interior_of(g, nb)
    assert torch.all(g[:nb] == 0) and torch.all(g[-nb:] == 0)  # ghost inputs are overwritten
    gen = torch.Generator().manual_seed(seed + 2)
    eps = 1e-6
    worst = 0.0
    for j in range(ndir):
        v = torch.zeros((5, n, n, n), dtype=torch.float64)
        v[variables] = torch.randn((len(variables), n, n, n), generator=gen, dtype=torch.float64)
        v /= torch.sqrt(torch.sum(v * v))
        vd = v.cuda()
        ana = float(torch.sum(g_int * vd))
        up = plan.run(to_device_layout(u + eps * v, nb), record=False)
        um = plan.run(to_device_layout(u - eps * v, nb), record=False)
        rp, rm = interior_of(up, nb)[0], interior_of(um, nb)[0]
        # L(+) - L(-) = 0.5 * sum((rp - rm) * (rp + rm - 2 t)), no cancellation of L itself
        fd = float(0.5 * torch.sum((rp - rm) * (rp + rm - 2.0 * rho_t))) / (2.0 * eps)
        rel = abs(fd - ana) / max(abs(ana), 1e-30)
        worst = max(worst, rel)
        if j < 4 or rel > 1e-6:
            print(f"dir {j}: adjoint {ana:+.12e} fd {fd:+.12e} rel {rel:.2e}", flush=True)
    print(f"worst relative difference over {ndir} directions: {worst:.3e}")
    plan.close()
    pk.finalize()
    return worst


def test_adjoint_finite_differences_32cubed():
    print("FD check 32^3, 32 steps, checkpoint every 8, generic state", flush=True)
    worst = fd_check(32, 32, 8, 0.05, 64, [0, 1, 2, 3, 4], seed=40)
    assert worst <= 1e-6


def test_adjoint_finite_differences_rest_state_density_energy():
    """rho = 1 + 0.1 G, p = 1, v = 0 (the rollout's initial state), 8 steps,
    directions in density and energy."""
    print("FD check 32^3, rest state, rho/E directions", flush=True)
    worst = fd_check(32, 8, 4, 0.0, 16, [0, 4], seed=50)
    assert worst <= 1e-6
