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


GDF_DRIVER = """
program drive_gdf
   use iso_c_binding
   use piernik_gdf
   implicit none
   type(pk_gdf_meta) :: meta
   type(pk_gdf_cg) :: cg(1)
   real(c_double), target :: den(4, 3, 2)
   character(kind=c_char, len=5), target :: nm
   type(c_ptr) :: fields(1), data(1)
   integer :: i, j, k
   integer(c_int32_t) :: rc
   print '(a,2i6)', 'SIZES', c_sizeof(meta), c_sizeof(cg(1))
   call pk_gdf_meta_default(meta)
   meta%n_d = [4, 3, 2]
   meta%time = 0.25d0
   cg(1)%cg_id = 0
   cg(1)%level = 0
   cg(1)%n_b = [4, 3, 2]
   cg(1)%off = 0
   cg(1)%left_edge = 0d0
   cg(1)%dl = [0.25d0, 1d0/3d0, 0.5d0]
   cg(1)%right_edge = 1d0
   do k = 1, 2
      do j = 1, 3
         do i = 1, 4
            den(i, j, k) = i + 10 * j + 100 * k
         enddo
      enddo
   enddo
   nm = 'deni' // c_null_char
   fields(1) = c_loc(nm)
   data(1) = c_loc(den)
   rc = pk_gdf_write_host('f.h5' // c_null_char, meta, 1_c_int32_t, cg, 1_c_int32_t, fields, data)
   print '(a,i4,1x,a)', 'RC', rc, pk_gdf_error_message()
end program drive_gdf
"""


@pytest.mark.skipif(FLANG is None, reason="amdflang not available")
@pytest.mark.skipif(not os.path.exists(os.path.join(LIBDIR, "libpiernik_gdf.so")), reason="libpiernik_gdf.so not built")
def test_fortran_gdf_module_writes_a_file(tmp_path):
    from piernik_amd import gdf
    import numpy as np
    src = os.path.join(ROOT, "piernik_amd", "fortran", "piernik_gdf.F90")
    subprocess.run([FLANG, "-c", "-cpp", src, "-o", "piernik_gdf.o"], cwd=tmp_path, check=True)
    (tmp_path / "drive.F90").write_text(GDF_DRIVER)
    subprocess.run([FLANG, "-c", "drive.F90", "-o", "drive.o"], cwd=tmp_path, check=True)
    subprocess.run([FLANG, "drive.o", "piernik_gdf.o", f"-L{LIBDIR}", "-l:libpiernik_gdf.so", f"-Wl,-rpath,{LIBDIR}",
                    "-o", "drive"], cwd=tmp_path, check=True)
    out = subprocess.run(["./drive"], cwd=tmp_path, capture_output=True, text=True, check=True).stdout
    sizes = [int(x) for x in out.split("SIZES")[1].split()[:2]]
    assert sizes == [ctypes.sizeof(gdf._Meta), ctypes.sizeof(gdf._Cg)]
    assert int([ln for ln in out.splitlines() if ln.startswith("RC")][0].split()[1]) == 0
    got = gdf.read(str(tmp_path / "f.h5"), 0, "density", (4, 3, 2))
    i, j, k = np.meshgrid(np.arange(1, 5), np.arange(1, 4), np.arange(1, 3), indexing="ij")
    assert np.array_equal(got, (i + 10 * j + 100 * k).astype(np.float64))
