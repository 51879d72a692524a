"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "piernik_amd", "libpiernik_hip.so")


GDF_LIB = os.path.join(ROOT, "piernik_amd", "libpiernik_gdf.so")
# the library that exports each header's entry points
LIB_OF = {"piernik_hip.h": LIB, "piernik_gdf.h": GDF_LIB}


def declared_symbols(header=None):
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        if header and os.path.basename(h) != header:
            continue
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pk_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_every_header_has_a_library():
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        assert os.path.basename(h) in LIB_OF, h


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("pk_init", "pk_solve_cg", "pk_sweep", "pk_fluid_update", "pk_last_error"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built (python -m piernik_amd.build)")
def test_library_exports_every_declared_symbol():
    import piernik_amd as pk
    lib = pk.load_library()
    missing = [n for n in declared_symbols("piernik_hip.h") if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(GDF_LIB), reason="libpiernik_gdf.so not built")
def test_gdf_library_exports_every_declared_symbol():
    from piernik_amd import gdf
    lib = gdf.lib()
    missing = [n for n in declared_symbols("piernik_gdf.h") if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_calls_before_init_fail_cleanly():
    import piernik_amd as pk
    lib = pk.load_library()
    rc = lib.pk_solve_cg(1, 0, pk.RK2_1, 1e-3, 0.0)
    assert rc != 0
    assert b"pk_init" in lib.pk_last_error()
