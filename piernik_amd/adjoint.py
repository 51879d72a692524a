"""Differentiable rollout of the hydro sweep for PyTorch.

A rollout is `nsteps` iterations of PIERNIK's main loop on one periodic
hydro block (time_step then fluid_update_full; timestep.F90:112-240,
fluidupdate.F90:99-136) on the GPU.  `rollout(u0, plan)` is a
torch.autograd.Function: the forward pass keeps the state every
`checkpoint_every` steps, the backward pass recomputes each segment from its
checkpoint (keeping the state before every sweep) and runs the library's
hand-written adjoint sweeps backwards through it.

dt is part of the differentiated map: each step's dt = cfl * min(dl/c)
depends on the state at the start of the step (through the minimising
cell), except where the growth cap dt <= 2 dt_old binds, where it depends on
the previous dt.

The last segment's sweep inputs are recorded during the forward pass itself
(record_last): its backward needs no recomputation, and with
checkpoint_every >= nsteps the whole rollout is recorded and nothing is
recomputed.

A state is a tuple of tensors in the reference layout: u(5, nxg, nyg, nzg)
in Fortran order, i.e. torch shape (nzg, nyg, nxg, 5), and for MHD also
b (nzg, nyg, nxg, 3) and psi (nzg, nyg, nxg); float64, on the GPU, ghost
cells included (their input values are replaced by periodic images, so their
gradient is 0).  For MHD the GLM speed ch = cfl_glm*min(dl)/dt is part of the
differentiated map.
"""
import ctypes

import torch

from . import PiernikError, _call, _device_ptr, timestep_fluid, fluid_update, update_boundaries
from .setups import _DeviceBlock


class RolloutPlan:
    """One registered device block and the slots a checkpointed rollout uses."""

    def __init__(self, cfg, n, nsteps, checkpoint_every=8, nb=4, cg_id=900001, record_last=True):
        if cfg.nfluids != 1:
            raise PiernikError("the adjoint covers one fluid per block")
        self.cfg = cfg
        self.nsteps = int(nsteps)
        self.k = max(1, int(checkpoint_every))
        self.nsw = sum(1 for h in cfg.has_dir if h) * 2
        self.nck = (self.nsteps + self.k - 1) // self.k
        shape = tuple(int(n) if h else 1 for h in cfg.has_dir) if not isinstance(n, (tuple, list)) else tuple(n)
        self.cg = _DeviceBlock(shape, nb, cfg, cg_id)
        self.cg.register()
        _call("pk_adjoint_setup", self.cg.cg_id, self.nck + self.nsw * self.k)
        self.dts = []
        self.capped = []
        self.record_last = bool(record_last)
        self.last_recorded = False  # the slots hold the last segment's sweep inputs

    @property
    def shape(self):
        ng = self.cg.ng
        return (ng[2], ng[1], ng[0], 5)

    def _empty(self, device):
        ng = self.cg.ng
        u = torch.empty((ng[2], ng[1], ng[0], 5), dtype=torch.float64, device=device)
        if not self.cfg.magnetized:
            return (u,)
        return (u, torch.empty((ng[2], ng[1], ng[0], 3), dtype=torch.float64, device=device),
                torch.empty((ng[2], ng[1], ng[0]), dtype=torch.float64, device=device))

    def run(self, state, record=True):
        """Forward rollout from state (tensor u, or tuple (u, b, psi) for MHD);
        returns the final state in the same form."""
        cg = self.cg
        single = torch.is_tensor(state)
        state = (state,) if single else tuple(state)
        cg.upload_device(*[t.contiguous() for t in state])
        update_boundaries(cg)
        self.dts, self.capped = [], []
        self.last_recorded = False
        last0 = (self.nck - 1) * self.k  # first step of the last segment
        dt_old = -1.0
        for s in range(self.nsteps):
            if record and s % self.k == 0:
                _call("pk_state_save", cg.cg_id, s // self.k)
            dt, _ = timestep_fluid(cg)
            cap = False
            if s >= 1 and dt_old > 0.0 and dt_old * 2.0 < dt:  # timestep.F90:200-208
                dt, cap = dt_old * 2.0, True
            if record and self.record_last and s >= last0:
                # the same step, keeping the state before every sweep
                _call("pk_fluid_update_record", cg.cg_id, self.nck + (s - last0) * self.nsw, dt)
            else:
                fluid_update(cg, dt)
            self.dts.append(dt)
            self.capped.append(cap)
            dt_old = dt
        self.last_recorded = record and self.record_last
        out = self._empty(state[0].device)
        cg.download_device(*out)
        return out[0] if single else out

    def backward(self, g):
        """dL/d(initial state) for dL/d(final state) = g, in the same form."""
        cg = self.cg
        single = torch.is_tensor(g)
        g = (g,) if single else tuple(g)
        ptrs = [ctypes.c_void_p(_device_ptr(t.contiguous())) for t in g] + [None] * (3 - len(g))
        _call("pk_adjoint_set", cg.cg_id, *ptrs)
        carry = 0.0
        for seg in reversed(range(self.nck)):
            s0, s1 = seg * self.k, min(seg * self.k + self.k, self.nsteps)
            if not (seg == self.nck - 1 and self.last_recorded):
                _call("pk_state_load", cg.cg_id, seg)
                for s in range(s0, s1):
                    _call("pk_fluid_update_record", cg.cg_id, self.nck + (s - s0) * self.nsw, self.dts[s])
            for s in reversed(range(s0, s1)):
                slot = self.nck + (s - s0) * self.nsw
                dtb = ctypes.c_double(0.0)
                _call("pk_fluid_update_adjoint", cg.cg_id, slot, self.dts[s], ctypes.byref(dtb))
                total = dtb.value + carry
                if self.capped[s]:
                    carry = 2.0 * total  # dt_s = 2 dt_{s-1}
                else:
                    _call("pk_timestep_adjoint", cg.cg_id, slot, total)
                    carry = 0.0
        self.last_recorded = False  # the recomputed segments have overwritten the slots
        out = self._empty(g[0].device)
        ptrs = [ctypes.c_void_p(_device_ptr(t)) for t in out] + [None] * (3 - len(out))
        _call("pk_adjoint_get", cg.cg_id, *ptrs)
        return out[0] if single else out

    def close(self):
        self.cg.unregister()


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, plan, *state):
        ctx.plan = plan
        ctx.n = len(state)
        out = plan.run(tuple(t.detach() for t in state), record=True)
        return out

    @staticmethod
    def backward(ctx, *g):
        grads = ctx.plan.backward(tuple(t.detach().contiguous() for t in g))
        return (None,) + tuple(grads)


def rollout(state, plan):
    """Final state of `plan.nsteps` steps, differentiable w.r.t. the initial
    state (a tensor u for hydro, a tuple (u, b, psi) for MHD)."""
    if torch.is_tensor(state):
        return _Rollout.apply(plan, state)[0]
    return _Rollout.apply(plan, *state)


class EnsemblePlan:
    """Independent problems, one block each, advanced together.

    Every instance takes its own time_step (own dt, caps, ch, damping); the
    forward steps are the library's persistent list launches, the backward
    pass runs the adjoint of all instances in one persistent launch per
    sweep (pk_ensemble_*).  States are tuples (u, b, psi) for MHD or (u,)
    for hydro, per instance, in the reference layout (see RolloutPlan)."""

    def __init__(self, cfg, shapes, nsteps, checkpoint_every=8, nb=4, first_id=700001):
        if cfg.nfluids != 1:
            raise PiernikError("the adjoint covers one fluid per block")
        self.cfg = cfg
        self.nsteps = int(nsteps)
        self.k = max(1, int(checkpoint_every))
        self.nsw = sum(1 for h in cfg.has_dir if h) * 2
        self.nck = (self.nsteps + self.k - 1) // self.k
        self.blocks = []
        for i, shape in enumerate(shapes):
            cg = _DeviceBlock(tuple(shape), nb, cfg, first_id + i)
            cg.register()
            _call("pk_adjoint_setup", cg.cg_id, self.nck + self.nsw * self.k)
            self.blocks.append(cg)
        n = len(self.blocks)
        ids = (ctypes.c_int64 * n)()
        _call("pk_ensemble_ids", ids, n)
        if [int(x) for x in ids] != [cg.cg_id for cg in self.blocks]:
            raise PiernikError("the ensemble must hold exactly this plan's blocks")
        self.dts = []
        self.capped = []

    def _empty(self, cg, device):
        ng = cg.ng
        u = torch.empty((ng[2], ng[1], ng[0], 5), dtype=torch.float64, device=device)
        if not self.cfg.magnetized:
            return (u,)
        return (u, torch.empty((ng[2], ng[1], ng[0], 3), dtype=torch.float64, device=device),
                torch.empty((ng[2], ng[1], ng[0]), dtype=torch.float64, device=device))

    def run(self, states, record=True, mode=None):
        """Forward: every instance nsteps main-loop iterations; returns the
        final states."""
        from . import BATCH_GRAPH
        mode = BATCH_GRAPH if mode is None else mode
        n = len(self.blocks)
        for cg, st in zip(self.blocks, states):
            cg.upload_device(*[t.contiguous() for t in st])
            update_boundaries(cg)
        dt = (ctypes.c_double * n)()
        cap = (ctypes.c_int32 * n)()
        self.dts, self.capped = [], []
        for s in range(self.nsteps):
            if record and s % self.k == 0:
                _call("pk_ensemble_state_save", s // self.k)
            _call("pk_ensemble_step", int(mode), int(s == 0), dt, cap)
            self.dts.append([dt[i] for i in range(n)])
            self.capped.append([bool(cap[i]) for i in range(n)])
        out = []
        for cg, st in zip(self.blocks, states):
            o = self._empty(cg, st[0].device)
            cg.download_device(*o)
            out.append(o)
        return out

    def backward(self, grads):
        """dL/d(initial states) for dL/d(final states) = grads."""
        n = len(self.blocks)
        for cg, g in zip(self.blocks, grads):
            ptrs = [ctypes.c_void_p(_device_ptr(t.contiguous())) for t in g] + [None] * (3 - len(g))
            _call("pk_adjoint_set", cg.cg_id, *ptrs)
        carry = [0.0] * n
        for seg in reversed(range(self.nck)):
            s0, s1 = seg * self.k, min(seg * self.k + self.k, self.nsteps)
            _call("pk_ensemble_state_load", seg)
            for s in range(s0, s1):
                _call("pk_ensemble_record_step", (ctypes.c_double * n)(*self.dts[s]), self.nck + (s - s0) * self.nsw)
            for s in reversed(range(s0, s1)):
                slot = self.nck + (s - s0) * self.nsw
                dtb = (ctypes.c_double * n)()
                _call("pk_ensemble_step_adjoint", (ctypes.c_double * n)(*self.dts[s]), slot, dtb)
                tb = (ctypes.c_double * n)()
                for i in range(n):
                    total = dtb[i] + carry[i]
                    if self.capped[s][i]:
                        carry[i] = 2.0 * total  # dt_s = 2 dt_{s-1}
                    else:
                        tb[i] = total
                        carry[i] = 0.0
                _call("pk_ensemble_timestep_adjoint", slot, tb)
        out = []
        for cg, g in zip(self.blocks, grads):
            o = self._empty(cg, g[0].device)
            ptrs = [ctypes.c_void_p(_device_ptr(t)) for t in o] + [None] * (3 - len(o))
            _call("pk_adjoint_get", cg.cg_id, *ptrs)
            out.append(o)
        return out

    def close(self):
        for cg in self.blocks:
            cg.unregister()
