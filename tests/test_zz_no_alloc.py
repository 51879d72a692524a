"""pk_sweep_noalloc: the fused sweep over caller-owned device planes, with
its scratch in the library's static device array, against pk_sweep on a
registered block (same state, same result bits, same floor report and
timestep).  Two calls in a row, the second on a smaller block where no floor
fires, so a scratch left over from the first call would show."""
import ctypes

import numpy as np
import pytest

import floors_setup as fs
import piernik_amd as pk
from piernik_amd import setups

pytestmark = pytest.mark.gpu


def planes_of(cg):
    """Host arrays of cg -> (nplanes, nzg, nyg, nxg) float64, the library's
    device plane order: fluid planes, then B and psi."""
    pl = [cg.u[v].T for v in range(cg.nvar)]
    if cg.magnetized:
        pl += [cg.b[k].T for k in range(3)] + [cg.psi.T]
    return np.ascontiguousarray(np.stack(pl))


def sweep_noalloc(src, dst, cg, ddim, dt, ch, fac, report):
    lib = pk.load_library()
    fn = lib.pk_sweep_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                   ctypes.POINTER(ctypes.c_int32), ctypes.c_int32, ctypes.POINTER(ctypes.c_double), ctypes.c_int32,
                   ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
    nzg, nyg, nxg = src.shape[1:]
    n = (ctypes.c_int32 * 3)(*cg.n)
    dl = (ctypes.c_double * 3)(*cg.dl)
    rc = fn(ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()), nzg * nyg * nxg, nxg, nxg * nyg, n,
            cg.nb, dl, ddim, dt, ch, fac, ctypes.c_void_p(report.data_ptr()), None)
    if rc != 0:
        raise pk.PiernikError(f"pk_sweep_noalloc: {lib.pk_last_error().decode()} (code {rc})")


CASES = [  # (block, ddim, expect floors)
    ("floors_mhd", 0, True),
    ("smooth_mhd_small", 2, False),
]


def make(case):
    if case == "floors_mhd":
        return fs.floor_block(shape=(20, 12, 10), magnetized=True, seed=4)
    # density and pressure three times those of the smooth state: above the
    # floors' smalld = 0.9 and smallei = 1.5, so no floor fires
    cg = setups.smooth_random((12, 10, 8), seed=5, magnetized=True)
    cg.u *= 3.0
    setups.fill_ghosts_periodic(cg)
    return cg, None


def test_sweep_noalloc_matches_pk_sweep():
    import torch

    pk.load_library()
    pk.finalize()
    # one configuration for both calls (pk_init once, as a caller that
    # cannot allocate would do it once at set-up)
    cg0, cfg = make(CASES[0][0])
    pk.init(cfg)
    pk.set_cfl_control("warn", disallow_negatives=True)
    for case, ddim, floors in CASES:
        cg, _ = make(case)
        print(f"noalloc {case} ddim={ddim} shape={cg.n}", flush=True)
        src = torch.from_numpy(planes_of(cg)).cuda()
        dst = src.clone()
        report = torch.full((8,), -1.0, dtype=torch.float64, device="cuda")
        dt, ch, fac = 0.3 * min(cg.dl), 1.1, 0.97
        torch.cuda.synchronize()
        sweep_noalloc(src, dst, cg, ddim, dt, ch, fac, report)
        pk.synchronize()
        out = dst.cpu().numpy()
        rep_na = report.cpu().numpy()

        # the same sweep on a registered block
        cg.upload()
        pk.get_report(cg, reset=True)
        pk.sweep(cg, ddim, dt, ch, fac)
        rep = pk.get_report(cg)
        dt_ref, c_ref = pk.timestep_fluid(cg)
        cg.download()
        want = planes_of(cg)
        g = cg.g
        inner = (slice(None),) + tuple(slice(g[d], g[d] + cg.n[d]) for d in (2, 1, 0))
        bad = np.argwhere(~(out[inner] == want[inner]))
        assert bad.size == 0, f"{case}: {len(bad)} interior values differ, first at {tuple(bad[0])}"
        # ghost cells of dst are not written: still the copy of src
        assert np.array_equal(out[:, :g[2]], planes_of_src_ghost(src, g)), case

        flags = int(rep_na[4])
        assert bool(flags & 1) == rep["dn_negative"] and bool(flags & 2) == rep["ei_negative"], (case, flags, rep)
        dvol = 1.0
        for d in range(3):
            dvol *= cg.dl[d]
        assert rep_na[5] * dvol == rep["magic_mass"][0], (case, rep_na[5] * dvol, rep["magic_mass"])
        assert (rep["magic_mass"][0] > 0.0) == floors, case
        assert cfg.cfl * min(rep_na[0], rep_na[1], rep_na[2]) == dt_ref, case
        assert rep_na[3] == c_ref, case
        cg.unregister()
    pk.finalize()


def planes_of_src_ghost(src, g):
    return src.cpu().numpy()[:, :g[2]]
