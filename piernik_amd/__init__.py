"""piernik_amd -- MI355X (gfx950) split Riemann sweep of PIERNIK.

Python host mirror of the reference operator interface over the C-ABI of
libpiernik_hip.so (include/piernik_hip.h).  Names follow the reference:

  Config                 &NUMERICAL_SETUP + fluid gamma (src/base/global.F90:205-262)
  GridContainer (cg)     one grid block, reference array layout (named_array.F90)
  solve_cg(cg, ddim, istep, dt, chspeed)
                         solve_cg_sub / solve_cg_riemann (sweeps.F90:140-151,
                         solve_cg_riemann.F90:68-107)
  sweep(cg, ddim, dt, chspeed)    both RK2 stages of one directional sweep
  update_boundaries(cg)  periodic ghost refresh (sweeps.F90:66-107)
  time_step(cg, ...)     timestep_fluid + caps (timestep.F90:112-240)
  fluid_update(cg, dt)   fluid_update_full (fluidupdate.F90:99-136)
  run(cg, nsteps)        the main loop (piernik.F90:124-170)
  run_all(nsteps, mode)  the main loop over every registered block (the block
                         loops of sweeps.F90:230-260), batched on the device

Errors raise PiernikError with the library's message, where the reference
calls die().  There is no CPU fallback: without the built library every call
raises.
"""
import ctypes
import math
import os

import numpy as np

__all__ = [
    "PiernikError", "Config", "GridContainer", "load_library", "init", "finalize",
    "solve_cg", "sweep", "update_boundaries", "glmdamping", "timestep_fluid", "time_step",
    "update_chspeed", "glm_factor", "fluid_update", "run", "get_report", "synchronize",
    "profile_begin", "profile_end", "autotune", "timestep_all", "fluid_update_all", "run_all",
    "autotune_all", "list_tiles", "set_list_tiles", "set_list_split",
    "EULER", "RK2_1", "RK2_2", "LIMITERS", "BATCH_GRAPH", "BATCH_PERSISTENT", "BATCH_PER_BLOCK",
]

EULER, RK2_1, RK2_2 = 10000, 10001, 10002          # src/base/constants.F90:230-232
LIMITERS = {"vanleer": 0, "minmod": 1, "moncen": 2, "superbee": 3}  # fluxlimiters.F90:90-103
BATCH_GRAPH, BATCH_PERSISTENT, BATCH_PER_BLOCK = 0, 1, 2

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpiernik_hip.so")


class PiernikError(RuntimeError):
    """An error reported by the library (the reference would call die())."""


class _Fluid(ctypes.Structure):
    _fields_ = [("gamma", ctypes.c_double), ("magnetized", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class _Config(ctypes.Structure):
    _fields_ = [
        ("nfluids", ctypes.c_int32), ("limiter", ctypes.c_int32), ("limiter_b", ctypes.c_int32),
        ("integration_order", ctypes.c_int32), ("fluid", _Fluid * 2),
        ("smalld", ctypes.c_double), ("smallei", ctypes.c_double),
        ("use_smalld", ctypes.c_int32), ("use_smallei", ctypes.c_int32),
        ("cfl", ctypes.c_double), ("cfl_glm", ctypes.c_double), ("glm_alpha", ctypes.c_double),
        ("has_dir", ctypes.c_int32 * 3), ("tile_x", ctypes.c_int32 * 2), ("tile_yz", ctypes.c_int32 * 2),
        ("fast_math", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


class _ProfStats(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_int64 * 4), ("ms", ctypes.c_double * 4), ("algo_bytes", ctypes.c_double * 4)]


class _Report(ctypes.Structure):
    _fields_ = [("dn_negative", ctypes.c_int32), ("ei_negative", ctypes.c_int32),
                ("magic_mass", ctypes.c_double * 2)]


_lib = None

_PROTOS = {
    "pk_init": [ctypes.POINTER(_Config)],
    "pk_finalize": [],
    "pk_register_block": [ctypes.c_int64, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32,
                          ctypes.POINTER(ctypes.c_double)] + [ctypes.c_void_p] * 6,
    "pk_unregister_block": [ctypes.c_int64],
    "pk_upload": [ctypes.c_int64],
    "pk_download": [ctypes.c_int64],
    "pk_upload_from": [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "pk_download_to": [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "pk_upload_device": [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "pk_download_device": [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "pk_solve_cg": [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double],
    "pk_sweep": [ctypes.c_int64, ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_double],
    "pk_fill_ghosts": [ctypes.c_int64],
    "pk_set_neighbour": [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64],
    "pk_datafield": [ctypes.c_int64, ctypes.c_char_p, ctypes.c_void_p],
    "pk_set_time_control": [ctypes.c_double] * 5,
    "pk_datafield_f32": [ctypes.c_int64, ctypes.c_char_p, ctypes.c_void_p],
    "pk_register_fcflx": [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "pk_unsplit_stage": [ctypes.c_int64, ctypes.c_int32, ctypes.c_double, ctypes.c_double],
    "pk_set_cfl_control": [ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_int32, ctypes.c_int32],
    "pk_redone_steps": [ctypes.POINTER(ctypes.c_int32)],
    "pk_fluid_update_unsplit": [ctypes.c_int64, ctypes.c_double],
    "pk_run_unsplit": [ctypes.c_int64, ctypes.c_double, ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                       ctypes.POINTER(ctypes.c_int32)],
    "pk_glm_damp": [ctypes.c_int64, ctypes.c_double],
    "pk_timestep_fluid": [ctypes.c_int64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
    "pk_fluid_update": [ctypes.c_int64, ctypes.c_double],
    "pk_run": [ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
    "pk_get_report": [ctypes.c_int64, ctypes.POINTER(_Report), ctypes.c_int32],
    "pk_synchronize": [],
    "pk_autotune": [ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_int32)],
    "pk_timestep_all": [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
    "pk_fluid_update_all": [ctypes.c_int32, ctypes.c_double],
    "pk_run_all": [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
    "pk_autotune_all": [ctypes.c_double, ctypes.c_double],
    "pk_list_tiles": [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32)],
    "pk_set_list_tiles": [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32],
    "pk_set_list_split": [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32,
                          ctypes.c_int32],
    "pk_adjoint_setup": [ctypes.c_int64, ctypes.c_int32],
    "pk_state_save": [ctypes.c_int64, ctypes.c_int32],
    "pk_state_load": [ctypes.c_int64, ctypes.c_int32],
    "pk_adjoint_set": [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "pk_adjoint_get": [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "pk_fluid_update_record": [ctypes.c_int64, ctypes.c_int32, ctypes.c_double],
    "pk_sweep_adjoint": [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                         ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
    "pk_fluid_update_adjoint": [ctypes.c_int64, ctypes.c_int32, ctypes.c_double, ctypes.POINTER(ctypes.c_double)],
    "pk_timestep_adjoint": [ctypes.c_int64, ctypes.c_int32, ctypes.c_double],
    "pk_ensemble_ids": [ctypes.POINTER(ctypes.c_int64), ctypes.c_int32],
    "pk_ensemble_step": [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                         ctypes.POINTER(ctypes.c_int32)],
    "pk_ensemble_run": [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                        ctypes.POINTER(ctypes.c_double)],
    "pk_ensemble_state_save": [ctypes.c_int32],
    "pk_ensemble_state_load": [ctypes.c_int32],
    "pk_ensemble_record_step": [ctypes.POINTER(ctypes.c_double), ctypes.c_int32],
    "pk_ensemble_step_adjoint": [ctypes.POINTER(ctypes.c_double), ctypes.c_int32, ctypes.POINTER(ctypes.c_double)],
    "pk_ensemble_timestep_adjoint": [ctypes.c_int32, ctypes.POINTER(ctypes.c_double)],
    "pk_profile_begin": [],
    "pk_profile_end": [ctypes.POINTER(_ProfStats)],
}


def _bind_torch_runtime():
    """Bind to PyTorch's HIP runtime when PyTorch is installed.

    PyTorch ships its own libamdhip64 (same soname as ROCm's).  If this
    library loaded ROCm's first, a later `import torch` would bring a second
    HIP runtime into the process and PyTorch would find no GPU.  Loading
    PyTorch's copy first (without importing torch) makes both use one
    runtime, so device tensors can be handed to pk_upload_device.  Set
    PIERNIK_HIP_RUNTIME=system to keep ROCm's runtime (no PyTorch in the
    process)."""
    if os.environ.get("PIERNIK_HIP_RUNTIME", "") == "system":
        return None
    try:
        import importlib.util
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return None
    if spec is None or not spec.submodule_search_locations:
        return None
    rt = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    if not os.path.exists(rt):
        return None
    ctypes.CDLL(rt, mode=ctypes.RTLD_GLOBAL)
    return rt


def load_library(path=None):
    """Load libpiernik_hip.so; raise PiernikError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise PiernikError(f"{p} not found: build it with `python -m piernik_amd.build`")
    _bind_torch_runtime()
    lib = ctypes.CDLL(p)
    for name, args in _PROTOS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.pk_last_error.argtypes = []
    lib.pk_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _call(name, *args):
    lib = load_library()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        raise PiernikError(f"{name}: {lib.pk_last_error().decode()} (code {rc})")


class Config:
    """Run configuration, named as the reference's namelists.

    Defaults follow src/base/global.F90:205-262 except the floors, which the
    reference derives from the initial state (initfluids.F90:212-320) and which
    must be given explicitly here (``sanitize_smallx`` reproduces that rule).
    """

    def __init__(self, gamma=5.0 / 3.0, magnetized=False, limiter="vanleer", limiter_b=None,
                 integration_order=2, cfl=0.7, cfl_glm=None, glm_alpha=0.1, smalld=1e-10, smallei=1e-10,
                 use_smalld=True, use_smallei=True, has_dir=(True, True, True), tile_x=(0, 0), tile_yz=(0, 0),
                 fast_math=False, fluids=None):
        # fluids: (gamma, magnetized) per fluid in flind order (at most two,
        # at most one magnetized); by default the one fluid (gamma, magnetized)
        if fluids is None:
            fluids = [(gamma, magnetized)]
        self.fluids = [(float(g), bool(m)) for g, m in fluids]
        self.gamma = self.fluids[0][0]
        self.magnetized = any(m for _, m in self.fluids)
        self.limiter = limiter
        self.limiter_b = limiter if limiter_b is None else limiter_b
        self.integration_order = int(integration_order)
        self.cfl = float(cfl)
        self.cfl_glm = self.cfl if cfl_glm is None else float(cfl_glm)
        self.glm_alpha = float(glm_alpha)
        self.smalld = float(smalld)
        self.smallei = float(smallei)
        self.use_smalld = bool(use_smalld)
        self.use_smallei = bool(use_smallei)
        self.has_dir = tuple(bool(h) for h in has_dir)
        self.tile_x = tuple(tile_x)
        self.tile_yz = tuple(tile_yz)
        self.fast_math = bool(fast_math)

    @property
    def nfluids(self):
        return len(self.fluids)

    @property
    def nvar(self):
        return 5 * self.nfluids

    def to_c(self):
        c = _Config()
        if not 1 <= self.nfluids <= 2:
            raise PiernikError(f"{self.nfluids} fluids: one or two are supported")
        c.nfluids = self.nfluids
        if self.limiter not in LIMITERS or self.limiter_b not in LIMITERS:
            raise PiernikError(f"unknown limiter {self.limiter!r}/{self.limiter_b!r}")
        c.limiter = LIMITERS[self.limiter]
        c.limiter_b = LIMITERS[self.limiter_b]
        c.integration_order = self.integration_order
        for f, (g, m) in enumerate(self.fluids):
            c.fluid[f].gamma = g
            c.fluid[f].magnetized = int(m)
        c.smalld, c.smallei = self.smalld, self.smallei
        c.use_smalld, c.use_smallei = int(self.use_smalld), int(self.use_smallei)
        c.cfl, c.cfl_glm, c.glm_alpha = self.cfl, self.cfl_glm, self.glm_alpha
        for d in range(3):
            c.has_dir[d] = int(self.has_dir[d])
        c.tile_x[0], c.tile_x[1] = self.tile_x
        c.tile_yz[0], c.tile_yz[1] = self.tile_yz
        c.fast_math = int(self.fast_math)
        return c


def sanitize_smallx(u, b=None, gamma=5.0 / 3.0, nb=4, has_dir=(True, True, True)):
    """The reference's default floors (initfluids.F90:240-312): smalld is the
    minimum initial density itself; smallp is 1e-4 of the minimum pressure."""
    sl = tuple(slice(nb, -nb) if h else slice(None) for h in has_dir)
    dn = u[(0,) + sl]
    e = u[(4,) + sl] - 0.5 * (u[(1,) + sl] ** 2 + u[(2,) + sl] ** 2 + u[(3,) + sl] ** 2) / dn
    if b is not None:
        e = e - 0.5 * (b[(0,) + sl] ** 2 + b[(1,) + sl] ** 2 + b[(2,) + sl] ** 2)
    return float(dn.min()), float((e / (gamma - 1.0)).min() * 1e-4)


class GridContainer:
    """One grid block ("cg") with host arrays in the reference layout.

    u(nvar, nxg, nyg, nzg), b(3, nxg, nyg, nzg), psi(nxg, nyg, nzg), Fortran
    order, ghost cells included (nxg = nx + 2*nb in active directions);
    nvar = 5 per fluid, fluid f at components 5f..5f+4.
    """

    _next_id = 1

    def __init__(self, n, nb=4, dl=None, magnetized=False, has_dir=(True, True, True), cg_id=None, nfluids=1):
        self.n = tuple(int(x) for x in n)
        self.nb = int(nb)
        self.has_dir = tuple(bool(h) for h in has_dir)
        self.g = tuple(self.nb if h else 0 for h in self.has_dir)
        self.ng = tuple(self.n[d] + 2 * self.g[d] for d in range(3))
        if dl is None:
            dl = tuple(1.0 / self.n[d] for d in range(3))
        self.dl = tuple(float(x) for x in dl)
        self.magnetized = bool(magnetized)
        self.nfluids = int(nfluids)
        self.nvar = 5 * self.nfluids
        self.u = np.zeros((self.nvar,) + self.ng, dtype=np.float64, order="F")
        self.uh = np.zeros_like(self.u)
        self.b = np.zeros((3,) + self.ng, dtype=np.float64, order="F") if magnetized else None
        self.bh = np.zeros_like(self.b) if magnetized else None
        self.psi = np.zeros(self.ng, dtype=np.float64, order="F") if magnetized else None
        self.psih = np.zeros_like(self.psi) if magnetized else None
        if cg_id is None:
            cg_id = GridContainer._next_id
            GridContainer._next_id += 1
        self.cg_id = int(cg_id)
        self.registered = False

    def interior(self, a):
        """View of the interior cells of a host array (leading axis kept)."""
        sl = tuple(slice(g, g + n) for g, n in zip(self.g, self.n))
        return a[(slice(None),) + sl] if a.ndim == 4 else a[sl]

    # -- device mirror ------------------------------------------------------
    def register(self):
        if self.registered:
            return
        n = (ctypes.c_int32 * 3)(*self.n)
        dl = (ctypes.c_double * 3)(*self.dl)
        ptr = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
        _call("pk_register_block", self.cg_id, n, self.nb, dl, ptr(self.u), ptr(self.uh), ptr(self.b),
              ptr(self.bh), ptr(self.psi), ptr(self.psih))
        self.registered = True

    def upload(self):
        self.register()
        _call("pk_upload", self.cg_id)

    def download(self):
        _call("pk_download", self.cg_id)

    def upload_device(self, u, b=None, psi=None):
        """Copy from device tensors in the reference layout (torch tensors of
        shape (nzg, nyg, nxg, nvar), (nzg, nyg, nxg, 3), (nzg, nyg, nxg),
        float64, contiguous, on this GPU) without host staging."""
        self.register()
        ptr = lambda t: None if t is None else ctypes.c_void_p(_device_ptr(t))
        _call("pk_upload_device", self.cg_id, ptr(u), ptr(b), ptr(psi))

    def download_device(self, u, b=None, psi=None):
        ptr = lambda t: None if t is None else ctypes.c_void_p(_device_ptr(t))
        _call("pk_download_device", self.cg_id, ptr(u), ptr(b), ptr(psi))

    def unregister(self):
        if self.registered:
            _call("pk_unregister_block", self.cg_id)
            self.registered = False


def _device_ptr(t):
    """Device address of a torch tensor handed to the library.  The library
    works on its own stream: torch's queued work on the tensor is finished
    first (the library's calls return with their own work finished)."""
    if not (t.is_cuda and t.is_contiguous() and str(t.dtype) == "torch.float64"):
        raise PiernikError("device arrays must be contiguous float64 tensors on the GPU")
    import torch
    torch.cuda.synchronize(t.device)
    return t.data_ptr()


_cfg = None


def init(cfg):
    """Configure the library (init_global / init_fluids / set_interpolations)."""
    global _cfg
    c = cfg.to_c()
    _call("pk_init", ctypes.byref(c))
    _cfg = cfg


def finalize():
    global _cfg
    if _lib is not None:
        _call("pk_finalize")
    _cfg = None


def solve_cg(cg, ddim, istep, dt, chspeed=0.0):
    """One RK stage of one block along ddim (0,1,2) -- solve_cg_sub."""
    _call("pk_solve_cg", cg.cg_id, int(ddim), int(istep), float(dt), float(chspeed))


def sweep(cg, ddim, dt, chspeed=0.0, psi_fac=1.0):
    """Both RK2 stages of one directional sweep in one fused pass."""
    _call("pk_sweep", cg.cg_id, int(ddim), float(dt), float(chspeed), float(psi_fac))


def update_boundaries(cg):
    """Refresh of every ghost cell of the block on the device: periodic on
    itself, or from its neighbours (set_neighbour)."""
    _call("pk_fill_ghosts", cg.cg_id)


CFL_CONTROL = {"warn": 0, "none": 0, "redo": 1, "repeat": 1, "flex": 2, "flexible": 2, "auto": 3, "adaptive": 3}


def set_cfl_control(cflcontrol="warn", cfl_max=0.9, dt_shrink=0.5, max_redostep_attempts=10,
                    disallow_negatives=True):
    """&NUMERICAL_SETUP's step-retry knobs for run / run_all (timestep.F90:73-90)."""
    _call("pk_set_cfl_control", CFL_CONTROL[cflcontrol], float(cfl_max), float(dt_shrink),
          int(max_redostep_attempts), int(bool(disallow_negatives)))


def set_time_control(dt_initial=-1.0, dt_max_grow=2.0, dt_min=2.2250738585072014e-308, dt_max=1.7976931348623157e308,
                     tend=1.7976931348623157e308):
    """&NUMERICAL_SETUP's dt_initial, dt_max_grow, dt_min, dt_max and the time
    left (tend) for run / run_all (pk_set_time_control)."""
    _call("pk_set_time_control", float(dt_initial), float(dt_max_grow), float(dt_min), float(dt_max), float(tend))


def redone_steps():
    """How many steps the last run / run_all redid."""
    n = ctypes.c_int32()
    _call("pk_redone_steps", ctypes.byref(n))
    return n.value


def unsplit_stage(cg, istep, dt, chspeed=0.0):
    """One RK stage of the unsplit solver on a periodic block
    (solve_cg_unsplit + update_boundaries)."""
    _call("pk_unsplit_stage", cg.cg_id, int(istep), float(dt), float(chspeed))


def fluid_update_unsplit(cg, dt):
    """fluid_update_unsplit on a periodic block."""
    _call("pk_fluid_update_unsplit", cg.cg_id, float(dt))


def run_unsplit(cg, tend, nmax=1 << 30):
    """The main loop with the unsplit solver until t >= tend (or nmax steps);
    returns (t, nsteps)."""
    t, ns = ctypes.c_double(), ctypes.c_int32()
    _call("pk_run_unsplit", cg.cg_id, float(tend), int(nmax), ctypes.byref(t), ctypes.byref(ns))
    return t.value, ns.value


def datafield(cg, name):
    """Output field `name` (the reference's short names: deni, vlxi, prei,
    magB, psi, ...) of cg's interior, computed on the device from its current
    state (pk_datafield); float64 array of shape n (x, y, z)."""
    out = np.zeros(cg.n, dtype=np.float64, order="F")
    _call("pk_datafield", cg.cg_id, name.encode(), out.ctypes.data_as(ctypes.c_void_p))
    return out


def datafield_f32(cg, name):
    """datafield rounded to nearest FP32 on the device (pk_datafield_f32)."""
    out = np.zeros(cg.n, dtype=np.float32, order="F")
    _call("pk_datafield_f32", cg.cg_id, name.encode(), out.ctypes.data_as(ctypes.c_void_p))
    return out


FC_FINEBND, FC_COARSEBND = 0, 1


def register_fcflx(cg, ddim, side, kind, first, index, uflx, bflx=None):
    """Fine/coarse flux points of cg along ddim (pk_register_fcflx): one side
    of cg%finebnd (kind FC_FINEBND) or cg%coarsebnd (FC_COARSEBND) in the
    reference layout, index (n1, n2) int32, uflx (nvar, n1, n2), bflx (4, n1,
    n2), Fortran order over pdims(ddim, ORTHO1:ORTHO2); first is
    cg%ijkse(ddim, LO).  pk_solve_cg reads finebnd and writes coarsebnd in
    place, so the arrays are kept referenced here.  index None removes."""
    cg.register()
    refs = cg.__dict__.setdefault("_fc_refs", {})
    if index is None:
        _call("pk_register_fcflx", cg.cg_id, int(ddim), int(side), int(kind), int(first), None, None, None)
        refs.pop((ddim, side, kind), None)
        return
    d1, d2 = ((1, 2), (2, 0), (0, 1))[ddim]
    n1, n2 = cg.n[d1], cg.n[d2]
    ok = (index.dtype == np.int32 and index.shape == (n1, n2) and index.flags.f_contiguous
          and uflx.dtype == np.float64 and uflx.shape == (cg.nvar, n1, n2) and uflx.flags.f_contiguous)
    if cg.magnetized:
        ok = ok and bflx is not None and bflx.dtype == np.float64 and bflx.shape == (4, n1, n2) \
            and bflx.flags.f_contiguous
    if not ok:
        raise ValueError("fine/coarse flux arrays must be Fortran-ordered index (n1, n2) int32, "
                         f"uflx ({cg.nvar}, n1, n2) and bflx (4, n1, n2) float64 with (n1, n2) = {(n1, n2)}")
    ptr = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
    _call("pk_register_fcflx", cg.cg_id, int(ddim), int(side), int(kind), int(first), ptr(index), ptr(uflx),
          ptr(bflx if cg.magnetized else None))
    refs[(ddim, side, kind)] = (index, uflx, bflx)


def set_neighbour(cg, direction, side, nb):
    """The ghost layers of cg on `side` (0 low, 1 high) of `direction` take the
    adjoining interior layers of block nb (pk_set_neighbour)."""
    _call("pk_set_neighbour", cg.cg_id, int(direction), int(side), nb.cg_id)


def connect(blocks, nbr):
    """Register the blocks and link them by a neighbour table nbr[i][d][s]
    (indices into blocks, as setups.decompose returns)."""
    for b in blocks:
        b.register()
    for i, b in enumerate(blocks):
        for d in range(3):
            if not b.has_dir[d]:
                continue
            for s in range(2):
                set_neighbour(b, d, s, blocks[nbr[i][d][s]])


def glmdamping(cg, fac):
    _call("pk_glm_damp", cg.cg_id, float(fac))


def timestep_fluid(cg):
    """(dt, c_max) of timestep_fluid: dt = cfl * min(dl/c)."""
    dt, c = ctypes.c_double(), ctypes.c_double()
    _call("pk_timestep_fluid", cg.cg_id, ctypes.byref(dt), ctypes.byref(c))
    return dt.value, c.value


def time_step(cg, nstep, dt_old):
    """time_step with dt_initial=-1, dt_max_grow=2, dt_max huge (timestep.F90:200-225)."""
    dt, _ = timestep_fluid(cg)
    if nstep >= 1 and dt_old > 0.0:
        dt = min(dt, dt_old * 2.0)
    return dt


def _min_dl(cg):
    return min(cg.dl[d] for d in range(3) if cg.has_dir[d])


def update_chspeed(cg, dt, cfg=None):
    """hdc.F90:112-122 with ch_grid: ch = cfl_glm*min(dl)/dt."""
    cfg = cfg or _cfg
    return min(1.7976931348623157e308, (cfg.cfl_glm * _min_dl(cg)) / dt)


def glm_factor(cg, chspeed, dt, cfg=None):
    """hdc.F90:313-331: exp(-max(0, glm_alpha*ch/(min(dl)/dt)))."""
    cfg = cfg or _cfg
    return math.exp(-max(0.0, (cfg.glm_alpha * chspeed) / (_min_dl(cg) / dt)))


def fluid_update(cg, dt):
    _call("pk_fluid_update", cg.cg_id, float(dt))


def run(cg, nsteps):
    """nsteps main-loop iterations on the device; returns (t, last dt)."""
    t, dt = ctypes.c_double(), ctypes.c_double()
    _call("pk_run", cg.cg_id, int(nsteps), ctypes.byref(t), ctypes.byref(dt))
    return t.value, dt.value


def get_report(cg, reset=False):
    r = _Report()
    _call("pk_get_report", cg.cg_id, ctypes.byref(r), int(bool(reset)))
    return {"dn_negative": bool(r.dn_negative), "ei_negative": bool(r.ei_negative),
            "magic_mass": list(r.magic_mass)}


def synchronize():
    _call("pk_synchronize")


def autotune(cg, dt=None, chspeed=None):
    """Pick the fastest tile shape of the fused sweep per direction on this
    block; returns {(dir): (pencils, cells)}."""
    if dt is None:
        dt = timestep_fluid(cg)[0]
    if chspeed is None:
        chspeed = update_chspeed(cg, dt)
    out = (ctypes.c_int32 * 6)()
    _call("pk_autotune", cg.cg_id, float(dt), float(chspeed), out)
    return {d: (out[2 * d], out[2 * d + 1]) for d in range(3) if cg.has_dir[d]}


def timestep_all():
    """(dt, c_max) of timestep_fluid over every registered block and fluid."""
    dt, c = ctypes.c_double(), ctypes.c_double()
    _call("pk_timestep_all", ctypes.byref(dt), ctypes.byref(c))
    return dt.value, c.value


def fluid_update_all(dt, mode=BATCH_GRAPH):
    """fluid_update_full over every registered block (each periodic on itself)."""
    _call("pk_fluid_update_all", int(mode), float(dt))


def run_all(nsteps, mode=BATCH_GRAPH):
    """nsteps main-loop iterations over every registered block with one dt;
    returns (t, last dt).  mode: BATCH_GRAPH (persistent launches replayed
    from a HIP graph), BATCH_PERSISTENT, or BATCH_PER_BLOCK (one launch per
    block); all give identical results."""
    t, dt = ctypes.c_double(), ctypes.c_double()
    _call("pk_run_all", int(mode), int(nsteps), ctypes.byref(t), ctypes.byref(dt))
    return t.value, dt.value


def autotune_all(dt=None, chspeed=0.0):
    """Tune the list sweep's workgroup size and per-shape tile lengths
    (chspeed 0: the list's own update_chspeed for this dt)."""
    if dt is None:
        dt = timestep_all()[0]
    _call("pk_autotune_all", float(dt), float(chspeed))


def set_list_tiles(ddim, ifluid, lanes, ts=0):
    """Pin the list sweep's shape for direction ddim and fluid ifluid's kind
    until the next autotune_all: lanes 64 (the stream form, ts = -iterations
    per item) or 128, 256, 512 (tiles of ts cells); ts 0: per-block default."""
    _call("pk_set_list_tiles", int(ddim), int(ifluid), int(lanes), int(ts))


def set_list_split(ddim, ifluid, shapes, ts=0):
    """Sweep the blocks of the given interior shapes ((nx, ny, nz) tuples) by
    a second launch of the stream form (ts = -iterations per item), the rest
    as before, until the next autotune_all or set_list_tiles."""
    flat = [int(v) for sh in shapes for v in sh]
    arr = (ctypes.c_int32 * max(1, len(flat)))(*flat)
    _call("pk_set_list_split", int(ddim), int(ifluid), arr, len(flat) // 3, int(ts))


def list_tiles(ddim, ifluid, cg):
    """(workgroup lanes, pencils, cells) of the list sweep for one block
    (stream form: (64, pencils per lane group, -iterations per item))."""
    out = (ctypes.c_int32 * 3)()
    _call("pk_list_tiles", int(ddim), int(ifluid), cg.cg_id, out)
    return tuple(out)


def profile_begin():
    """Start timing every sweep launch with HIP events on the library's stream."""
    _call("pk_profile_begin")


def profile_end():
    """Stop timing; per sweep direction (entries 0-2) and for the output-field
    kernel (entry 3): launches, total ms, algorithmic bytes."""
    st = _ProfStats()
    _call("pk_profile_end", ctypes.byref(st))
    return {"launches": list(st.launches), "ms": list(st.ms), "algo_bytes": list(st.algo_bytes)}
