"""The hand-written adjoint of the hydro rollout (piernik_amd.adjoint) on
the GPU, against

- autograd of the FP64 PyTorch transcription of the same scheme
  (tests/torch_reference.py, itself checked against the oracle): on 10^3 for
  3 steps and on 16^3 for the rollout's 32 steps with a checkpoint every 8,
  gradients agree to 1e-10 / 1e-9 relative;
- FP64 central finite differences of the GPU forward map with eps = 1e-6 on
  64 random unit directions at 32^3 over 2 steps: each directional
  derivative agrees to 1e-6 relative.

Finite differences of a TVD scheme only agree with its derivative while no
limiter switches branch between u - eps*v and u + eps*v.  Near every
extremum a slope passes through zero, so on 32^3 over 32 steps dozens of
cells sit within the perturbation of a switch, and each shifts a directional
derivative by ~1e-2; the long rollout is therefore checked against autograd,
which differentiates the same branch-taken map exactly.

The loss is 0.5*||rho_final - rho_target||^2 over interior cells.  The
states are generic: at v = 0 exactly the limited slopes of the velocity sit
at the kink of the limiter, where the map has no derivative in momentum
directions, and with a uniform p the pressure slopes are rounding noise, at
the kink in every direction.  The state at rest (v = 0, p varying) is
checked in the directions of density and energy, which keep v = 0.
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
    """rho = 1 + 0.1 G, p = 1 + 0.1 G', v = vamp * G''.  A uniform p would
    leave the pressure slopes at rounding noise, where each limiter branch
    hangs on the last bits of the forward state."""
    rho = tr.gaussian_density(n, seed)
    vel = [vamp * (tr.gaussian_density(n, seed + 10 + d, amp=1.0) - 1.0) for d in range(3)]
    p = tr.gaussian_density(n, seed + 20)
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
    return float(loss.detach()), u0.grad


@pytest.mark.parametrize("limiter,n,steps,every,tol", [("vanleer", 10, 3, 2, 1e-10), ("minmod", 10, 3, 2, 1e-10),
                                                        ("vanleer", 16, 32, 8, 1e-9)])
def test_adjoint_matches_autograd_transcription(limiter, n, steps, every, tol):
    nb = 4
    u = interior_state(n, 5, 0.05)
    rho_t = tr.gaussian_density(n, 6)
    print(f"adjoint vs autograd, {n}^3, {steps} steps, checkpoint every {every}, {limiter}", flush=True)
    plan = make_plan(n, steps, every, limiter)
    _, g = loss_grad(plan, to_device_layout(u, nb), rho_t.cuda(), nb)
    g = interior_of(g, nb).cpu()
    ur = u.clone().requires_grad_(True)
    out = tr.run(ur, steps, [1.0 / n] * 3, limiter_name=limiter, cfl=0.5)
    (0.5 * torch.sum((out[0] - rho_t) ** 2)).backward()
    err = float(torch.max(torch.abs(g - ur.grad)) / torch.max(torch.abs(ur.grad)))
    print(f"max rel diff {err:.3e}")
    assert err <= tol, err
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
    pairs = []
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
        pairs.append((ana, fd))
    plan.close()
    pk.finalize()
    return report(pairs)


def report(pairs):
    """Largest |fd - adjoint| over the directions, relative to
    max(|adjoint|, rms over directions): the central difference carries a
    roundoff floor of ~1e-7 (dt is one scalar whose last-bit rounding is
    ~1e-7 of its change under an eps perturbation), so a direction whose
    derivative is small by chance is measured against the typical size."""
    rms = math.sqrt(sum(a * a for a, _ in pairs) / len(pairs))
    worst = 0.0
    for j, (ana, fd) in enumerate(pairs):
        rel = abs(fd - ana) / max(abs(ana), rms)
        worst = max(worst, rel)
        if j < 4 or rel > 1e-6:
            print(f"dir {j}: adjoint {ana:+.12e} fd {fd:+.12e} rel {rel:.2e}", flush=True)
    print(f"worst relative difference over {len(pairs)} directions: {worst:.3e} (rms derivative {rms:.3e})")
    return worst


def test_adjoint_finite_differences_32cubed():
    print("FD check 32^3, 2 steps, checkpoint every 1, generic state", flush=True)
    worst = fd_check(32, 2, 1, 0.05, 64, [0, 1, 2, 3, 4], seed=40)
    assert worst <= 1e-6


def test_adjoint_finite_differences_rest_state_density_energy():
    """rho = 1 + 0.1 G, p = 1 + 0.1 G', v = 0, 2 steps, directions in
    density and energy."""
    print("FD check 32^3, rest state, rho/E directions", flush=True)
    worst = fd_check(32, 2, 2, 0.0, 16, [0, 4], seed=50)
    assert worst <= 1e-6
