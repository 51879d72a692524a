"""Build libpiernik_hip.so (gfx950) in-tree with hipcc.

Usage: ``python -m piernik_amd.build`` or ``piernik_amd.build.build()``.
Objects go to piernik_amd/_build/, the shared library to
piernik_amd/libpiernik_hip.so.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpiernik_hip.so")
GDF_LIB = os.path.join(HERE, "libpiernik_gdf.so")
HDF5 = os.environ.get("PK_HDF5_ROOT", "/opt/conda")  # HDF5 1.10 C library and headers of this image

SOURCES = ["pk_kernels.hip", "pk_kernels_fast.hip", "pk_adjoint.hip", "pk_unsplit.hip", "pk_unsplit_fast.hip",
           "pk_datafields.hip", "pk_capi.hip"]

# -ffp-contract=off: every a*b+c rounds twice, as the reference's FP64 code.
# -packed-fp32-ops off: no packed FP32 instructions are generated.
HIPFLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.isabs(c) and os.path.exists(c):
            return c
    return "hipcc"


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("command failed: " + " ".join(cmd))
    return r


def build(verbose=False):
    """Compile every HIP source for gfx950 and link libpiernik_hip.so."""
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    jobs = []
    for src in SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        flags = list(HIPFLAGS)
        if src.endswith("_fast.hip"):
            # the relaxed build of the same kernels: FMA contraction, fast division and sqrt
            flags = [f for f in flags if f not in ("-ffp-contract=off", "-fno-fast-math")]
            flags += ["-ffp-contract=fast", "-ffast-math"]
        if src == "pk_kernels_fast.hip":
            # schedule the sweeps for instruction-level parallelism: at two waves per SIMD
            # either way (227 -> 256 VGPRs), 256^3 MHD steps run 1-2 % faster (the
            # persistent list sweep 0.5 % slower), measured side by side on one GPU
            flags += ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]
        jobs.append([hipcc, *flags, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    build_gdf(verbose)
    return LIB


def build_gdf(verbose=False):
    """libpiernik_gdf.so: the GDF data files, host code over the HDF5 C
    library; pk_datafield comes from libpiernik_hip.so beside it."""
    if not os.path.exists(os.path.join(HDF5, "include", "hdf5.h")):
        sys.stderr.write("HDF5 headers not found under %s: libpiernik_gdf.so not built\n" % HDF5)
        return None
    # the run path lists the system's C++ runtime before the HDF5 directory,
    # which also holds an older libstdc++ that ROCm's libraries cannot use
    cxx = subprocess.run(["g++", "-print-file-name=libstdc++.so"], capture_output=True, text=True).stdout.strip()
    rpath = ":".join(["$ORIGIN", os.path.dirname(os.path.realpath(cxx)), os.path.join(HDF5, "lib")])
    tmp = GDF_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I" + os.path.join(HDF5, "include"),
          os.path.join(CSRC, "pk_gdf.cpp"), "-o", tmp, "-L" + HERE, "-l:libpiernik_hip.so",
          "-L" + os.path.join(HDF5, "lib"), "-lhdf5", "-Wl,-rpath," + rpath], verbose)
    os.replace(tmp, GDF_LIB)
    return GDF_LIB


if __name__ == "__main__":
    print(build(verbose=True))
