"""Build (when stale) and load tests/hostmath/libhostmath.so: the sweep's
per-cell arithmetic and its reverse mode, compiled for the host."""
import ctypes
import os
import subprocess

HM = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HM, "..", "..", "piernik_amd", "csrc")
SRC = [os.path.join(HM, "hostmath.cpp"), os.path.join(CSRC, "pk_math.hpp"), os.path.join(CSRC, "pk_adjoint.hpp"),
       os.path.join(CSRC, "pk_mass.hpp")]
LIB = os.path.join(HM, "libhostmath.so")
P = ctypes.POINTER(ctypes.c_double)


def build():
    if os.path.exists(LIB) and all(os.path.getmtime(LIB) >= os.path.getmtime(s) for s in SRC):
        return LIB
    # built under a name of this process and renamed into place, so that
    # parallel test workers never load a half-written library
    tmp = f"{LIB}.tmp{os.getpid()}"
    subprocess.run(["hipcc", "-x", "hip", "--offload-host-only", "-O2", "-fPIC", "-shared", "-ffp-contract=off",
                    "-std=c++17", "-o", tmp, SRC[0]], check=True, capture_output=True)
    os.replace(tmp, LIB)
    return LIB


def load():
    build()
    L = ctypes.CDLL(LIB)
    d, i = ctypes.c_double, ctypes.c_int
    L.hm_spacing.argtypes = [d]
    L.hm_spacing.restype = d
    L.hm_mass_sum.argtypes = [i, P]
    L.hm_mass_sum.restype = d
    L.hm_limit.argtypes = [i, d, d]
    L.hm_limit.restype = d
    L.hm_prim.argtypes = [i, i, d, P, P]
    L.hm_hllc.argtypes = [i, P, P, d, P]
    L.hm_hlld.argtypes = [i, P, P, d, d, P]
    L.hm_floors.argtypes = [i, i, d, d, i, i, i, P, ctypes.POINTER(ctypes.c_uint)]
    L.hm_prim_vjp_hd.argtypes = [d, P, P, P]
    L.hm_limit_vjp.argtypes = [i, d, d, d, P]
    L.hm_hllc_vjp.argtypes = [P, P, d, P, P, P]
    L.hm_floors_hd.argtypes = [d, d, i, i, P]
    L.hm_floors_vjp_hd.argtypes = [d, d, i, i, P, P]
    L.hm_dt_vjp_hd.argtypes = [P, i, d, d, d, P]
    L.hm_hlld_vjp.argtypes = [P, P, d, d, P, P, P, P]
    L.hm_hlld1.argtypes = [P, P, d, d, P]
    L.hm_prim_vjp_mhd.argtypes = [d, P, P, P]
    L.hm_floors_mhd.argtypes = [d, d, i, i, P]
    L.hm_floors_vjp_mhd.argtypes = [d, d, i, i, P, P]
    L.hm_dt_vjp_mhd.argtypes = [P, i, d, d, d, P]
    L.hm_hlld_branch.argtypes = [P, P, d, P]
    return L
