"""Ensembles of independent instances (pk_ensemble_*): every instance with
its own time step, all advanced by the same persistent launches.

- forward: each instance equals the oracle's main loop (or_run: time_step
  with its own dt, dt_max_grow cap, c_h and GLM damping, then
  fluid_update_full) run on that instance alone, bitwise, and a single-block
  GPU run of the same problem;
- the 256-instance Orszag-Tang ensemble at its sizes ({64..256}^2): one
  instance of every size against the oracle, bitwise; every instance
  conserves mass, momentum, energy and B to rounding, stays positive, and a
  second run gives the same bits;
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


def oracle_block(oracle, n, u, b, psi, hd):
    """The instance as an oracle block: periodic ghosts, reference layout."""
    ob = oracle.OBlock((n, n, 1), 4, (1.0 / n, 1.0 / n, 1.0), True, hd)
    ob.u[...] = dev(u, hd).cpu().numpy().transpose(3, 2, 1, 0)
    ob.b[...] = dev(b, hd).cpu().numpy().transpose(3, 2, 1, 0)
    ob.psi[...] = dev(psi, hd).cpu().numpy().transpose(2, 1, 0)
    return ob


def assert_state_equals_oracle(fin, ob, where):
    for name, t, o, perm in (("u", fin[0], ob.u, (3, 2, 1, 0)), ("b", fin[1], ob.b, (3, 2, 1, 0)),
                             ("psi", fin[2], ob.psi, (2, 1, 0))):
        a = t.cpu().numpy().transpose(perm)
        bad = np.argwhere(~(a == o))
        assert bad.size == 0, (f"{where} {name}: {len(bad)} mismatches, first at {tuple(bad[0])}: "
                               f"gpu={a[tuple(bad[0])]!r} oracle={o[tuple(bad[0])]!r}")


def sim_time(dts):
    t = 0.0
    for dt in dts:  # t = t + dt per dt-substep (fluidupdate.F90:120,129)
        t = t + dt
        t = t + dt
    return t


@pytest.mark.parametrize("mode,split", [(pk.BATCH_GRAPH, False), (pk.BATCH_PER_BLOCK, False),
                                        (pk.BATCH_GRAPH, True), (pk.BATCH_PER_BLOCK, True)])
def test_ensemble_instances_match_oracle(oracle, mode, split):
    """Every instance, with its own dt, c_h and GLM damping factor, equals
    the oracle's main loop on that instance alone, bit for bit; split: the
    24^2 instance is swept by a second launch of the stream form, so the
    launches index the instances' dt through their list positions."""
    cases, cfg = ot_cases()
    hd = cfg.has_dir
    steps = 6
    p = oracle.params_from_config(cfg)
    print(f"ensemble vs oracle, {len(cases)} instances, mode {mode}, split {split}", flush=True)
    pk.load_library()
    pk.finalize()
    pk.init(cfg)
    ens = adjoint.EnsemblePlan(cfg, [(n, n, 1) for n, *_ in cases], steps, checkpoint_every=steps)
    if split:
        for d in range(2):
            pk.set_list_tiles(d, 0, 256, 0)
            pk.set_list_split(d, 0, [(24, 24, 1)], -4)
    fin = ens.run([(dev(u, hd), dev(b, hd), dev(psi, hd)) for _, u, b, psi in cases], record=False, mode=mode)
    dts = ens.dts
    ens.close()
    pk.finalize()
    # the instances take different time steps (sizes and amplitudes differ)
    assert len({d[i] for d in dts[:1] for i in range(len(cases))}) == len(cases)
    for i, (n, u, b, psi) in enumerate(cases):
        ob = oracle_block(oracle, n, u, b, psi, hd)
        t_o = oracle.run(p, ob, steps)
        assert sim_time([d[i] for d in dts]) == t_o, f"instance {i}: time differs"
        assert_state_equals_oracle(fin[i], ob, f"instance {i} (n={n})")


def _totals(st, hd):
    """Interior sums of rho, m, E (u) and B: conserved on a periodic block."""
    u = setups.interior_of(st[0], 4, hd)
    b = setups.interior_of(st[1], 4, hd)
    # interior_of gives (component, x, y, z): one total per component
    tot = torch.cat([u.sum(dim=(1, 2, 3)), b.sum(dim=(1, 2, 3))])
    mag = torch.cat([u.abs().sum(dim=(1, 2, 3)), b.abs().sum(dim=(1, 2, 3))])
    return tot.cpu(), mag.cpu()


def test_ensemble256_config_sizes():
    """The 256-instance Orszag-Tang ensemble at its sizes (64^2 .. 256^2,
    amplitudes 1 +- 0.1): one instance of each size against the oracle,
    bit for bit; every instance conservative to rounding and positive; two
    runs give the same bits."""
    from oracle import pyoracle as oracle

    cases, cfg = setups.orszag_tang_ensemble(256, seed=0)
    hd = cfg.has_dir
    steps = 3
    print(f"ensemble256: {len(cases)} instances, {sum(n * n for n, *_ in cases)} cells, {steps} steps", flush=True)
    pk.load_library()
    pk.finalize()
    pk.init(cfg)
    ens = adjoint.EnsemblePlan(cfg, [(n, n, 1) for n, *_ in cases], steps, checkpoint_every=steps)
    init = [(dev(u, hd), dev(b, hd), dev(psi, hd)) for _, u, b, psi in cases]
    fin = ens.run(init, record=False)
    dts = ens.dts
    fin2 = ens.run(init, record=False)
    assert ens.dts == dts
    ens.close()
    pk.finalize()
    for i in range(len(cases)):
        for a, c in zip(fin[i], fin2[i]):
            assert torch.equal(a, c), f"instance {i}: second run differs"
        t0, m0 = _totals(init[i], hd)
        t1, _ = _totals(fin[i], hd)
        zero = m0 == 0.0  # B_z of the 2D state: zero throughout, and stays so
        assert torch.equal(t1[zero], t0[zero]), f"instance {i}: a zero component changed"
        err = float(torch.max(torch.abs(t1 - t0)[~zero] / m0[~zero]))
        assert err <= 1e-13, f"instance {i}: conservation {err:.3e}"
        u = setups.interior_of(fin[i][0], 4, hd)
        bb = setups.interior_of(fin[i][1], 4, hd)
        rho = u[0]
        ekin = 0.5 * (u[1] ** 2 + u[2] ** 2 + u[3] ** 2) / rho
        eint = u[4] - ekin - 0.5 * (bb ** 2).sum(dim=0)
        assert float(rho.min()) > 0.0 and float(eint.min()) > 0.0, f"instance {i}: not positive"
    p = oracle.params_from_config(cfg)
    first = {}
    for i, (n, *_) in enumerate(cases):
        first.setdefault(n, i)
    assert sorted(first) == list(setups.OT_SIZES)
    for n, i in sorted(first.items()):
        _, u, b, psi = cases[i]
        ob = oracle_block(oracle, n, u, b, psi, hd)
        t_o = oracle.run(p, ob, steps)
        print(f"instance {i} ({n}^2) against the oracle", flush=True)
        assert sim_time([d[i] for d in dts]) == t_o
        assert_state_equals_oracle(fin[i], ob, f"instance {i} (n={n})")


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
