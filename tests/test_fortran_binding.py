"""The ISO_C_BINDING module compiles with amdflang, agrees with the C and
Python views of the structs, and a Fortran program links against the
library and gets the library's error back (no GPU work is done)."""
import ctypes
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = shutil.which("amdflang") or ("/opt/rocm/llvm/bin/amdflang" if os.path.exists("/opt/rocm/llvm/bin/amdflang")
                                     else None)
LIBDIR = os.path.join(ROOT, "piernik_amd")

DRIVER = """
program drive
   use iso_c_binding
   use piernik_hip
   implicit none
   type(pk_config) :: cfg
   type(pk_stage_report) :: rep
   type(pk_profile_stats) :: st
   integer(c_int32_t) :: rc
   print '(a,3i6)', 'SIZES', c_sizeof(cfg), c_sizeof(rep), c_sizeof(st)
   rc = pk_solve_cg(1_c_int64_t, 0_c_int32_t, PK_RK2_1, 1d-3, 0d0)
   print '(a,i4,1x,a)', 'RC', rc, pk_error_message()
end program drive
"""


@pytest.mark.skipif(FLANG is None, reason="amdflang not available")
@pytest.mark.skipif(not os.path.exists(os.path.join(LIBDIR, "libpiernik_hip.so")), reason="library not built")
def test_fortran_module_links_and_matches_c_layout(tmp_path):
    import piernik_amd as pk
    src = os.path.join(ROOT, "piernik_amd", "fortran", "piernik_hip.F90")
    subprocess.run([FLANG, "-c", "-cpp", src, "-o", "piernik_hip.o"], cwd=tmp_path, check=True)
    (tmp_path / "drive.F90").write_text(DRIVER)
    subprocess.run([FLANG, "-c", "drive.F90", "-o", "drive.o"], cwd=tmp_path, check=True)
    subprocess.run([FLANG, "drive.o", "piernik_hip.o", f"-L{LIBDIR}", "-lpiernik_hip", f"-Wl,-rpath,{LIBDIR}",
                    "-o", "drive"], cwd=tmp_path, check=True)
    out = subprocess.run(["./drive"], cwd=tmp_path, capture_output=True, text=True, check=True).stdout
    sizes = [int(x) for x in out.split("SIZES")[1].split()[:3]]
    assert sizes == [ctypes.sizeof(pk._Config), ctypes.sizeof(pk._Report), ctypes.sizeof(pk._ProfStats)]
    line = [ln for ln in out.splitlines() if ln.startswith("RC")][0]
    assert int(line.split()[1]) == 4          # PK_ESTATE
    assert "pk_init" in line
