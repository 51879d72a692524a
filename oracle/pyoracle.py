"""ctypes binding of oracle/liboracle.so.

TEST INFRASTRUCTURE ONLY: the checker for piernik_amd.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

LIM = {"vanleer": 0, "minmod": 1, "moncen": 2, "superbee": 3}
EULER, RK2_1, RK2_2 = 10000, 10001, 10002


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


class Fluid(ctypes.Structure):
    _fields_ = [("gam", ctypes.c_double), ("magnetized", ctypes.c_int32), ("beg", ctypes.c_int32)]


class Params(ctypes.Structure):
    _fields_ = [
        ("nfluids", ctypes.c_int32), ("fl", Fluid * 2), ("nvar", ctypes.c_int32),
        ("limiter", ctypes.c_int32), ("limiter_b", ctypes.c_int32), ("integration_order", ctypes.c_int32),
        ("use_smalld", ctypes.c_int32), ("use_smallei", ctypes.c_int32),
        ("smalld", ctypes.c_double), ("smallei", ctypes.c_double),
        ("cfl", ctypes.c_double), ("cfl_glm", ctypes.c_double), ("glm_alpha", ctypes.c_double),
        ("has_dir", ctypes.c_int32 * 3), ("disallow_negatives", ctypes.c_int32),
    ]


class Report(ctypes.Structure):
    _fields_ = [("dn_negative", ctypes.c_int32), ("ei_negative", ctypes.c_int32),
                ("magic_mass", ctypes.c_double * 2)]


class Block(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32 * 3), ("nb", ctypes.c_int32), ("dl", ctypes.c_double * 3)] + [
        (nm, ctypes.c_void_p) for nm in ("u", "uh", "b", "bh", "psi", "psih")]


class FcFace(ctypes.Structure):
    _fields_ = [("index", ctypes.c_void_p), ("uflx", ctypes.c_void_p), ("bflx", ctypes.c_void_p)]


class FcFlx(ctypes.Structure):
    _fields_ = [("first", ctypes.c_int64), ("fine", FcFace * 2), ("coarse", FcFace * 2)]


_lib = None
P = ctypes.POINTER


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        dp = P(ctypes.c_double)
        L.or_spacing.argtypes = [ctypes.c_double]
        L.or_spacing.restype = ctypes.c_double
        L.or_limiter.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        L.or_linear.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, dp]
        L.or_utoq.argtypes = [P(Params), ctypes.c_int, dp, dp, dp]
        L.or_hllc.argtypes = [ctypes.c_int, dp, dp, ctypes.c_double, dp]
        L.or_hlld.argtypes = [ctypes.c_int, dp, dp, dp, dp, ctypes.c_double, ctypes.c_double, dp, dp]
        L.or_floors.argtypes = [P(Params), ctypes.c_int, dp, dp, ctypes.c_int, ctypes.c_double, P(Report)]
        L.or_pencil_stage.argtypes = [P(Params), ctypes.c_int, ctypes.c_double, ctypes.c_double, dp, dp, dp, dp,
                                      dp, dp, ctypes.c_int, ctypes.c_double, P(Report)]
        L.or_solve_cg.argtypes = [P(Params), P(Block), ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_double, P(Report)]
        L.or_pencil_flux.argtypes = [P(Params), ctypes.c_int, ctypes.c_double, dp, dp, dp, dp]
        L.or_solve_cg_fc.argtypes = [P(Params), P(Block), ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_double, P(Report), P(FcFlx)]
        L.or_fill_ghosts_periodic.argtypes = [P(Params), P(Block)]
        L.or_update_chspeed.argtypes = [P(Params), P(Block), ctypes.c_double]
        L.or_update_chspeed.restype = ctypes.c_double
        L.or_glm_factor.argtypes = [P(Params), P(Block), ctypes.c_double, ctypes.c_double]
        L.or_glm_factor.restype = ctypes.c_double
        L.or_glmdamping.argtypes = [P(Params), P(Block), ctypes.c_double]
        L.or_timestep_fluid.argtypes = [P(Params), P(Block), ctypes.c_int, dp, dp]
        L.or_time_step.argtypes = [P(Params), P(Block), ctypes.c_int, ctypes.c_double, dp]
        L.or_time_step.restype = ctypes.c_double
        L.or_fluid_update.argtypes = [P(Params), P(Block), ctypes.c_double, P(Report)]
        L.or_run.argtypes = [P(Params), P(Block), ctypes.c_int, P(Report)]
        L.or_run.restype = ctypes.c_double
        L.or_fluid_update_list.argtypes = [P(Params), P(P(Block)), ctypes.c_int, ctypes.c_double, P(Report)]
        L.or_run_list.argtypes = [P(Params), P(P(Block)), ctypes.c_int, ctypes.c_int, P(Report)]
        L.or_run_list.restype = ctypes.c_double
        ip = P(ctypes.c_int32)
        L.or_fill_ghosts_list.argtypes = [P(Params), P(P(Block)), ip, ctypes.c_int]
        L.or_fluid_update_list_nb.argtypes = [P(Params), P(P(Block)), ip, ctypes.c_int, ctypes.c_double, P(Report)]
        L.or_run_list_nb.argtypes = [P(Params), P(P(Block)), ip, ctypes.c_int, ctypes.c_int, P(Report)]
        L.or_run_list_nb.restype = ctypes.c_double
        L.or_unsplit_stage.argtypes = [P(Params), P(Block), ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                       P(Report)]
        L.or_fluid_update_unsplit.argtypes = [P(Params), P(Block), ctypes.c_double, P(Report)]
        L.or_run_unsplit.argtypes = [P(Params), P(Block), ctypes.c_double, ctypes.c_int, P(ctypes.c_int32),
                                     P(Report)]
        L.or_run_unsplit.restype = ctypes.c_double
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(P(ctypes.c_double))


def params(gamma=5.0 / 3.0, magnetized=False, limiter="vanleer", limiter_b=None, integration_order=2,
           smalld=1e-10, smallei=1e-10, use_smalld=True, use_smallei=True, cfl=0.7, cfl_glm=None,
           glm_alpha=0.1, has_dir=(True, True, True), fluids=None, disallow_negatives=True):
    """fluids: (gamma, magnetized) per fluid; default the one fluid (gamma, magnetized).
    disallow_negatives: global.F90:237 (default .true.); the floors raise
    ei_negative only with it (sources.F90:425)."""
    if fluids is None:
        fluids = [(gamma, magnetized)]
    p = Params()
    p.nfluids = len(fluids)
    for f, (g, m) in enumerate(fluids):
        p.fl[f].gam = g
        p.fl[f].magnetized = int(m)
        p.fl[f].beg = 5 * f
    p.nvar = 5 * len(fluids)
    p.limiter = LIM[limiter]
    p.limiter_b = LIM[limiter if limiter_b is None else limiter_b]
    p.integration_order = integration_order
    p.use_smalld, p.use_smallei = int(use_smalld), int(use_smallei)
    p.smalld, p.smallei = smalld, smallei
    p.cfl = cfl
    p.cfl_glm = cfl if cfl_glm is None else cfl_glm
    p.glm_alpha = glm_alpha
    for d in range(3):
        p.has_dir[d] = int(has_dir[d])
    p.disallow_negatives = int(disallow_negatives)
    return p


def params_from_config(cfg, disallow_negatives=True):
    """Oracle parameters equal to a piernik_amd.Config (disallow_negatives is
    set on the GPU side by piernik_amd.set_cfl_control)."""
    return params(gamma=cfg.gamma, magnetized=cfg.magnetized, limiter=cfg.limiter, limiter_b=cfg.limiter_b,
                  integration_order=cfg.integration_order, smalld=cfg.smalld, smallei=cfg.smallei,
                  use_smalld=cfg.use_smalld, use_smallei=cfg.use_smallei, cfl=cfg.cfl, cfl_glm=cfg.cfl_glm,
                  glm_alpha=cfg.glm_alpha, has_dir=cfg.has_dir, fluids=cfg.fluids,
                  disallow_negatives=disallow_negatives)


class OBlock:
    """Host block in the reference layout, driven by the oracle."""

    def __init__(self, n, nb, dl, magnetized, has_dir=(True, True, True), nvar=5):
        self.n, self.nb, self.dl = tuple(n), nb, tuple(dl)
        g = tuple(nb if h else 0 for h in has_dir)
        self.ng = tuple(n[d] + 2 * g[d] for d in range(3))
        self.u = np.zeros((nvar,) + self.ng, order="F")
        self.uh = np.zeros_like(self.u)
        self.b = np.zeros((3,) + self.ng, order="F") if magnetized else None
        self.bh = np.zeros_like(self.b) if magnetized else None
        self.psi = np.zeros(self.ng, order="F") if magnetized else None
        self.psih = np.zeros_like(self.psi) if magnetized else None
        self.rep = Report()

    @classmethod
    def from_cg(cls, cg):
        o = cls(cg.n, cg.nb, cg.dl, cg.magnetized, cg.has_dir, nvar=cg.u.shape[0])
        o.u[...] = cg.u
        if cg.magnetized:
            o.b[...] = cg.b
            o.psi[...] = cg.psi
        return o

    def c(self):
        blk = Block()
        for d in range(3):
            blk.n[d] = self.n[d]
            blk.dl[d] = self.dl[d]
        blk.nb = self.nb
        for nm in ("u", "uh", "b", "bh", "psi", "psih"):
            a = getattr(self, nm)
            setattr(blk, nm, None if a is None else a.ctypes.data_as(ctypes.c_void_p))
        return blk


# pdims(ddim, ORTHO1:ORTHO2) (constants.F90:127-128)
ORTHO = ((1, 2), (2, 0), (0, 1))


class FcFaces:
    """cg%finebnd(ddim, LO:HI) and cg%coarsebnd(ddim, LO:HI) of one block and
    direction, in the reference layout (flx_arr.F90:38-42): index(n1, n2),
    uflx(nvar, n1, n2), bflx(4, n1, n2), Fortran order, (n1, n2) the interior
    extents along pdims(ddim, ORTHO1:ORTHO2).  `first` is cg%ijkse(ddim, LO);
    an index below it (LO) or above first + n - 1 (HI) marks no face, as
    init_gc_fcflx leaves them (grid_container_fcflx.F90:60-65)."""

    def __init__(self, n, ddim, nvar, magnetized, first=0):
        self.ddim, self.first = ddim, first
        n1, n2 = n[ORTHO[ddim][0]], n[ORTHO[ddim][1]]
        none = (first - 1, first + n[ddim])  # outside the interior on each side
        self.index = {(k, s): np.full((n1, n2), none[s], dtype=np.int32, order="F")
                      for k in ("fine", "coarse") for s in (0, 1)}
        self.uflx = {(k, s): np.zeros((nvar, n1, n2), order="F") for k in ("fine", "coarse") for s in (0, 1)}
        self.bflx = {(k, s): (np.zeros((4, n1, n2), order="F") if magnetized else None)
                     for k in ("fine", "coarse") for s in (0, 1)}

    def c(self):
        fc = FcFlx()
        fc.first = self.first
        for kind, arr in (("fine", fc.fine), ("coarse", fc.coarse)):
            for s in (0, 1):
                arr[s].index = self.index[kind, s].ctypes.data
                arr[s].uflx = self.uflx[kind, s].ctypes.data
                b = self.bflx[kind, s]
                arr[s].bflx = b.ctypes.data if b is not None else None
        return fc


def solve_cg(p, ob, ddim, istep, dt, chspeed=0.0, fc=None):
    blk = ob.c()
    if fc is None:
        lib().or_solve_cg(ctypes.byref(p), ctypes.byref(blk), ddim, istep, dt, chspeed, ctypes.byref(ob.rep))
    else:
        assert fc.ddim == ddim
        f = fc.c()
        lib().or_solve_cg_fc(ctypes.byref(p), ctypes.byref(blk), ddim, istep, dt, chspeed, ctypes.byref(ob.rep),
                             ctypes.byref(f))


def pencil_flux(p, u, b=None, chspeed=0.0):
    """Interface fluxes of one pencil: u (n, nvar) and b (n, 4) in sweep
    order; returns flx (n-1, nvar), bflx (n-1, 4) (row i: interface i+1/2)."""
    dp = P(ctypes.c_double)
    u = np.asfortranarray(u, dtype=np.float64)
    n, nv = u.shape
    flx = np.zeros((n - 1, nv), order="F")
    bflx = np.zeros((n - 1, 4), order="F")
    bb = np.asfortranarray(b, dtype=np.float64) if b is not None else None
    lib().or_pencil_flux(ctypes.byref(p), n, chspeed, u.ctypes.data_as(dp),
                         bb.ctypes.data_as(dp) if bb is not None else None, flx.ctypes.data_as(dp),
                         bflx.ctypes.data_as(dp))
    return flx, (bflx if b is not None else None)


def fill_ghosts(p, ob):
    blk = ob.c()
    lib().or_fill_ghosts_periodic(ctypes.byref(p), ctypes.byref(blk))


def time_step(p, ob, nstep, dt_old):
    blk = ob.c()
    c = ctypes.c_double()
    dt = lib().or_time_step(ctypes.byref(p), ctypes.byref(blk), nstep, dt_old, ctypes.byref(c))
    return dt, c.value


def timestep_fluid(p, ob):
    blk = ob.c()
    dt, c = ctypes.c_double(), ctypes.c_double()
    lib().or_timestep_fluid(ctypes.byref(p), ctypes.byref(blk), 0, ctypes.byref(dt), ctypes.byref(c))
    return dt.value, c.value


def update_chspeed(p, ob, dt):
    blk = ob.c()
    return lib().or_update_chspeed(ctypes.byref(p), ctypes.byref(blk), dt)


def glm_factor(p, ob, ch, dt):
    blk = ob.c()
    return lib().or_glm_factor(ctypes.byref(p), ctypes.byref(blk), ch, dt)


def fluid_update(p, ob, dt):
    blk = ob.c()
    lib().or_fluid_update(ctypes.byref(p), ctypes.byref(blk), dt, ctypes.byref(ob.rep))


def run(p, ob, nsteps):
    blk = ob.c()
    return lib().or_run(ctypes.byref(p), ctypes.byref(blk), nsteps, ctypes.byref(ob.rep))


def _block_array(obs):
    blks = [ob.c() for ob in obs]
    arr = (P(Block) * len(blks))(*[ctypes.pointer(b) for b in blks])
    reps = (Report * len(obs))()
    return blks, arr, reps


def _store_reports(obs, reps):
    for ob, r in zip(obs, reps):
        ob.rep.dn_negative |= r.dn_negative
        ob.rep.ei_negative |= r.ei_negative
        for f in range(2):
            ob.rep.magic_mass[f] += r.magic_mass[f]


def fluid_update_list(p, obs, dt):
    """fluid_update_full over a list of blocks (each periodic on itself)."""
    blks, arr, reps = _block_array(obs)
    lib().or_fluid_update_list(ctypes.byref(p), arr, len(obs), dt, reps)
    _store_reports(obs, reps)


def run_list(p, obs, nsteps, nbr=None):
    """nsteps main-loop iterations over a list of blocks with one dt; nbr
    (see fill_ghosts_list) connects them, None: each periodic on itself."""
    blks, arr, reps = _block_array(obs)
    if nbr is None:
        t = lib().or_run_list(ctypes.byref(p), arr, len(obs), nsteps, reps)
    else:
        nb = _nbr_array(nbr, len(obs))
        t = lib().or_run_list_nb(ctypes.byref(p), arr, nb.ctypes.data_as(P(ctypes.c_int32)), len(obs), nsteps, reps)
    _store_reports(obs, reps)
    return t


def _nbr_array(nbr, nblk):
    nb = np.ascontiguousarray(nbr, dtype=np.int32).reshape(nblk, 3, 2)
    if nb.min() < 0 or nb.max() >= nblk:
        raise ValueError("neighbour index outside the list")
    return nb


def fill_ghosts_list(p, obs, nbr):
    """Ghost exchange over blocks: nbr[i][d][s] is the index of the block whose
    interior adjoins block i's ghost layers on side s (0 low, 1 high) of
    direction d (a block periodic on itself names itself)."""
    blks, arr, reps = _block_array(obs)
    nb = _nbr_array(nbr, len(obs))
    lib().or_fill_ghosts_list(ctypes.byref(p), arr, nb.ctypes.data_as(P(ctypes.c_int32)), len(obs))


def fluid_update_list_nb(p, obs, nbr, dt):
    """fluid_update_full over connected blocks."""
    blks, arr, reps = _block_array(obs)
    nb = _nbr_array(nbr, len(obs))
    lib().or_fluid_update_list_nb(ctypes.byref(p), arr, nb.ctypes.data_as(P(ctypes.c_int32)), len(obs), dt, reps)
    _store_reports(obs, reps)


# ---- pencil-level helpers (column-major (n, nv) arrays) -------------------

def limiter(kind, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    dq = np.empty_like(q)
    lib().or_limiter(LIM[kind], len(q), _dp(q), _dp(dq))
    return dq


def linear(kind, q):
    """q: (n, nv) -> ql, qr: (n-1, nv)."""
    q = np.asfortranarray(q, dtype=np.float64)
    n, nv = q.shape
    ql = np.zeros((n - 1, nv), order="F")
    qr = np.zeros((n - 1, nv), order="F")
    lib().or_linear(LIM[kind], n, nv, _dp(q), _dp(ql), _dp(qr))
    return ql, qr


def utoq(p, u, b=None):
    u = np.asfortranarray(u, dtype=np.float64)
    q = np.zeros_like(u, order="F")
    bb = None if b is None else np.asfortranarray(b, dtype=np.float64)
    lib().or_utoq(ctypes.byref(p), u.shape[0], _dp(u), None if bb is None else _dp(bb), _dp(q))
    return q


def hllc(ql, qr, gam):
    ql = np.asfortranarray(ql, dtype=np.float64)
    qr = np.asfortranarray(qr, dtype=np.float64)
    f = np.zeros_like(ql, order="F")
    lib().or_hllc(ql.shape[0], _dp(ql), _dp(qr), gam, _dp(f))
    return f


def hlld(ql, qr, bl, br, gam, ch):
    """ql, qr: (m, 5) prims; bl, br: (m, 4) [Bn, Bt1, Bt2, psi] -> f (m, 5), bf (m, 4)."""
    ql = np.asfortranarray(ql, dtype=np.float64)
    qr = np.asfortranarray(qr, dtype=np.float64)
    bl = np.asfortranarray(bl, dtype=np.float64)
    br = np.asfortranarray(br, dtype=np.float64)
    f = np.zeros_like(ql, order="F")
    bf = np.zeros_like(bl, order="F")
    lib().or_hlld(ql.shape[0], _dp(ql), _dp(qr), _dp(bl), _dp(br), gam, ch, _dp(f), _dp(bf))
    return f, bf


def spacing(x):
    return lib().or_spacing(float(x))


def fluid_update_unsplit(p, ob, dt):
    """fluid_update_unsplit on one periodic block."""
    blk = ob.c()
    lib().or_fluid_update_unsplit(ctypes.byref(p), ctypes.byref(blk), dt, ctypes.byref(ob.rep))


def run_unsplit(p, ob, tend, nmax=1 << 30):
    """The main loop with the unsplit solver until t >= tend; returns (t, nsteps)."""
    blk = ob.c()
    ns = ctypes.c_int32()
    t = lib().or_run_unsplit(ctypes.byref(p), ctypes.byref(blk), tend, nmax, ctypes.byref(ns), ctypes.byref(ob.rep))
    return t, ns.value


# ---- the main loop with step retries (cflcontrol) -------------------------

def cfl_max_sanitized(cfl, cfl_max):
    """cfl_max as init_global sanitizes it (global.F90:304)."""
    return min(max(cfl_max, min(cfl * 1.1, cfl + 0.05, (1.0 + cfl) / 2.0)), 1.0)


def main_loop(nsteps, time_step, fluid_update, save, restore, negatives, cfl, mode="warn", cfl_max=0.9,
              dt_shrink=0.5, max_attempts=10, disallow_negatives=True):
    """The reference's main loop (piernik.F90:125-170) with its step controls,
    over callbacks, so that the same control flow drives the oracle's blocks
    and synthetic steps:

      time_step(s, dt_old) -> (dt, c_max)  time_step of the current state
      fluid_update(dt)                     one step (two dt-substeps)
      save() -> snapshot, restore(snap)    repeat_fluidstep's backup and restore
                                           (floor bookkeeping included)
      negatives() -> (dn, ei)              the flags the floors raised since the
                                           last call; reading clears them
                                           (check_cfl_violation, timestep.F90:291-300)

    mode 'redo' / 'flexible': after a step check_cfl_violation recomputes
    c_max; cfl_warn (timestep.F90:317-361) asks for a redo when
    stepcfl = c_new/c_old*cfl*dt_cur_shrink > cfl_max (sanitized) or, with
    disallow_negatives, when the step raised a flag; the step is redone from
    the backup with dt = dt_full*dt_cur_shrink, dt_cur_shrink = dt_shrink**attempt
    ('redo') or cfl_max/stepcfl, times dt_shrink from the second attempt on
    ('flexible') (repeat_fluidstep, timestep_retry.F90:86-146).  mode 'auto':
    cfl_auto (timestep.F90:368-405) scales each dt by cfl_c.  'warn': no retries.
    Returns (t, redone, attempts), attempts a list of (step, dt) of every
    update in order."""
    cmax = cfl_max_sanitized(cfl, cfl_max)
    redo = mode in ("redo", "flexible")
    t, dt_old, dt_full, c_old, shrink = 0.0, -1.0, 0.0, 0.0, 1.0
    s, attempt, redone, repeat = 0, 0, 0, False
    c_main_old, stepcfl, cfl_c = 0.0, 0.0, 1.0  # timestep_pub.F90:42
    snap = None
    attempts = []
    while s < nsteps:
        if not repeat:
            dt, c_old = time_step(s, dt_old)
            if mode == "auto":
                stepcfl_old, stepcfl = stepcfl, cfl
                if c_main_old > 0.0:
                    stepcfl = c_old / c_main_old * cfl
                else:
                    stepcfl_old = cfl
                if stepcfl > 0.0 and dt_old > 0.0:
                    cfl_c = min(1.0, 0.5 * (cfl_c + min(1.0, stepcfl_old / stepcfl * dt / dt_old)))
                    dt = dt * cfl_c
                else:
                    cfl_c = 1.0
            c_main_old = c_old
            dt_full = dt
            if redo:
                snap = save()
        else:
            restore(snap)
            dt = dt_full * shrink
        negatives()  # this step's flags only
        fluid_update(dt)
        attempts.append((s, dt))
        dt_old = dt
        if redo:
            _, c_new = time_step(s + 1, dt_old)
            dn, ei = negatives()
            stepcfl = cfl * shrink
            if c_old > 0.0:
                stepcfl = c_new / c_old * cfl * shrink
            repeat = stepcfl > cmax or (disallow_negatives and bool(dn or ei))
            if repeat:
                if mode == "flexible":
                    shrink = cmax / stepcfl
                    if attempt >= 1:
                        shrink = shrink * dt_shrink
                else:
                    shrink = dt_shrink ** (attempt + 1)
                attempt += 1
                redone += 1
                if attempt > max_attempts:
                    raise RuntimeError(f"step {s} redone more than max_redostep_attempts = {max_attempts} times")
                continue
            shrink, attempt = 1.0, 0
        t = t + dt
        t = t + dt
        s += 1
    return t, redone, attempts


def run_redo(p, ob, nsteps, mode, cfl_max=0.9, dt_shrink=0.5, max_attempts=10):
    """main_loop over one oracle block (periodic on itself).  A redone step
    starts from the backup's state and floor bookkeeping (downgrade_magic_mass);
    the block's report accumulates the flags of the steps that were kept.
    Returns (t, redone)."""
    fill_ghosts(p, ob)
    arrays = [a for a in (ob.u, ob.b, ob.psi) if a is not None]
    last = [(0, 0)]

    def save():
        return ([a.copy() for a in arrays], list(ob.rep.magic_mass),
                (ob.rep.dn_negative, ob.rep.ei_negative))

    def restore(snap):
        for a, sv in zip(arrays, snap[0]):
            a[...] = sv
        ob.rep.magic_mass[0], ob.rep.magic_mass[1] = snap[1]
        ob.rep.dn_negative, ob.rep.ei_negative = snap[2]

    def step(dt):
        keep = (ob.rep.dn_negative, ob.rep.ei_negative)
        ob.rep.dn_negative = ob.rep.ei_negative = 0
        fluid_update(p, ob, dt)
        last[0] = (ob.rep.dn_negative, ob.rep.ei_negative)
        ob.rep.dn_negative |= keep[0]
        ob.rep.ei_negative |= keep[1]

    def negatives():
        out, last[0] = last[0], (0, 0)
        return out

    t, redone, _ = main_loop(nsteps, lambda s, dt_old: time_step(p, ob, s, dt_old), step, save, restore,
                             negatives, p.cfl, mode, cfl_max, dt_shrink, max_attempts,
                             bool(p.disallow_negatives))
    return t, redone
