"""The threaded copy pool behind the library's host <-> device transfers
(piernik_amd/csrc/pk_xfer.hpp), built for the host and run without a GPU."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "xferpool", "pool_test.cpp")
HDR = os.path.join(HERE, "..", "piernik_amd", "csrc", "pk_xfer.hpp")
EXE = os.path.join(HERE, "xferpool", "pool_test.bin")


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    return None


@pytest.mark.skipif(_hipcc() is None, reason="hipcc not found")
def test_copy_pool_copies_every_byte():
    if not os.path.exists(EXE) or os.path.getmtime(EXE) < max(os.path.getmtime(SRC), os.path.getmtime(HDR)):
        subprocess.run([_hipcc(), "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-pthread", "-o", EXE, SRC],
                       check=True, capture_output=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
