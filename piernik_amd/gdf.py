"""GDF-v2 data files (PIERNIK's v2 HDF5 output) and bin/gdf_distance, over
libpiernik_gdf.so.

    meta = gdf.Meta(time=t, n_d=(64, 64, 64), edge=((0, 1), (0, 1), (0, 1)))
    gdf.write("run_0001.h5", [(cg, gdf.Placement(...))], ["deni", "vlxi", "prei"], meta)
    norms, total = gdf.distance("a.h5", "b.h5")

``python -m piernik_amd.gdf a.h5 b.h5`` prints the differences as
bin/gdf_distance does.
"""
import ctypes
import os
import sys

import numpy as np

from . import load_library

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libpiernik_gdf.so")


class _Meta(ctypes.Structure):
    _fields_ = [
        ("time", ctypes.c_double), ("timestep", ctypes.c_double),
        ("nstep", ctypes.c_int32), ("nres", ctypes.c_int32), ("nhdf", ctypes.c_int32),
        ("require_problem_IC", ctypes.c_int32),
        ("last_log_time", ctypes.c_double), ("last_tsl_time", ctypes.c_double),
        ("last_hdf_time", ctypes.c_double), ("last_res_time", ctypes.c_double),
        ("c_all_old", ctypes.c_double), ("stepcfl", ctypes.c_double), ("cfl_c", ctypes.c_double),
        ("nfluids", ctypes.c_int32), ("magic_mass", ctypes.c_double * 2),
        ("n_d", ctypes.c_int32 * 3), ("edge", (ctypes.c_double * 2) * 3), ("bnd", (ctypes.c_int32 * 2) * 3),
        ("eff_dim", ctypes.c_int32), ("h5_64bit", ctypes.c_int32), ("gdf_strict", ctypes.c_int32),
        ("problem_name", ctypes.c_char * 64), ("run_id", ctypes.c_char * 16),
        ("cm", ctypes.c_double), ("gram", ctypes.c_double), ("sek", ctypes.c_double), ("miu0", ctypes.c_double),
        ("lmtvB", ctypes.c_double * 5), ("s_lmtvB", (ctypes.c_char * 16) * 7),
    ]


class _Cg(ctypes.Structure):
    _fields_ = [
        ("cg_id", ctypes.c_int64), ("level", ctypes.c_int32), ("n_b", ctypes.c_int32 * 3),
        ("off", ctypes.c_int64 * 3), ("left_edge", ctypes.c_double * 3), ("right_edge", ctypes.c_double * 3),
        ("dl", ctypes.c_double * 3),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        load_library()  # libpiernik_hip.so first: pk_datafield
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built (python -m piernik_amd.build; needs the HDF5 C library)")
        L = ctypes.CDLL(LIB)
        vp, dp, i32 = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int32
        L.pk_gdf_meta_default.argtypes = [ctypes.POINTER(_Meta)]
        L.pk_gdf_meta_default.restype = None
        L.pk_gdf_write.argtypes = [ctypes.c_char_p, ctypes.POINTER(_Meta), i32, ctypes.POINTER(_Cg), i32, vp]
        L.pk_gdf_write_host.argtypes = [ctypes.c_char_p, ctypes.POINTER(_Meta), i32, ctypes.POINTER(_Cg), i32, vp,
                                        vp]
        L.pk_gdf_translate.argtypes = [ctypes.c_char_p, ctypes.c_char_p, i32]
        L.pk_gdf_read.argtypes = [ctypes.c_char_p, i32, ctypes.c_char_p, dp, ctypes.c_int64]
        L.pk_gdf_distance.argtypes = [ctypes.c_char_p, ctypes.c_char_p, i32, i32, i32, dp,
                                      ctypes.POINTER(ctypes.c_int32), ctypes.c_char_p, ctypes.c_int64, dp]
        L.pk_gdf_last_error.restype = ctypes.c_char_p
        for f in ("pk_gdf_write", "pk_gdf_write_host", "pk_gdf_translate", "pk_gdf_read", "pk_gdf_distance"):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


class GdfError(RuntimeError):
    pass


def _check(name, rc):
    if rc != 0:
        raise GdfError(f"{name}: {lib().pk_gdf_last_error().decode()} (code {rc})")


class Meta:
    """The run's description for a data file (pk_gdf_meta); defaults as
    pk_gdf_meta_default (CGS units, one fluid, FP32 datasets, GDF names)."""

    def __init__(self, **kw):
        self.c = _Meta()
        lib().pk_gdf_meta_default(ctypes.byref(self.c))
        for k, v in kw.items():
            self.set(k, v)

    def set(self, key, value):
        c = self.c
        if key in ("n_d", "magic_mass", "lmtvB"):
            arr = getattr(c, key)
            for i, x in enumerate(value):
                arr[i] = x
        elif key in ("edge", "bnd"):
            arr = getattr(c, key)
            for d in range(3):
                arr[d][0], arr[d][1] = value[d]
        elif key in ("problem_name", "run_id"):
            setattr(c, key, value.encode())
        elif key == "s_lmtvB":
            for i, s in enumerate(value):
                c.s_lmtvB[i].value = s.encode()
        else:
            setattr(c, key, value)


class Placement:
    """Where a block sits: level, offset of its first interior cell on that
    level, its edges and cell size (the reference's cg%l%id, cg%my_se, cg%fbnd, cg%dl)."""

    def __init__(self, n_b, off=(0, 0, 0), level=0, left_edge=None, right_edge=None, dl=None):
        self.n_b = tuple(int(x) for x in n_b)
        self.off = tuple(int(x) for x in off)
        self.level = int(level)
        self.dl = tuple(float(x) for x in (dl if dl is not None else [1.0 / n for n in self.n_b]))
        self.left_edge = tuple(float(x) for x in (left_edge if left_edge is not None else
                                                  [o * h for o, h in zip(self.off, self.dl)]))
        self.right_edge = tuple(float(x) for x in (right_edge if right_edge is not None else
                                                   [l + n * h for l, n, h in zip(self.left_edge, self.n_b, self.dl)]))

    def c(self, cg_id=0):
        r = _Cg()
        r.cg_id, r.level = cg_id, self.level
        for d in range(3):
            r.n_b[d], r.off[d] = self.n_b[d], self.off[d]
            r.left_edge[d], r.right_edge[d], r.dl[d] = self.left_edge[d], self.right_edge[d], self.dl[d]
        return r


def _names(fields):
    arr = (ctypes.c_char_p * len(fields))(*[f.encode() for f in fields])
    return arr


def write(path, blocks, fields, meta):
    """Data file of device-resident blocks: blocks = [(GridContainer, Placement)];
    every field is computed on the device (pk_datafield)."""
    cgs = (_Cg * len(blocks))(*[p.c(cg.cg_id) for cg, p in blocks])
    names = _names(fields)
    _check("pk_gdf_write", lib().pk_gdf_write(os.fsencode(path), ctypes.byref(meta.c), len(blocks), cgs,
                                              len(fields), ctypes.cast(names, ctypes.c_void_p)))


def write_host(path, placements, fields, data, meta):
    """Data file from host arrays: data[c][f] is field f of block c, shape n_b
    (x, y, z), any float64 array (copied to Fortran order)."""
    cgs = (_Cg * len(placements))(*[p.c() for p in placements])
    names = _names(fields)
    keep = []
    ptrs = (ctypes.c_void_p * (len(placements) * len(fields)))()
    for c, p in enumerate(placements):
        for f in range(len(fields)):
            a = np.asfortranarray(data[c][f], dtype=np.float64)
            assert a.shape == p.n_b, (a.shape, p.n_b)
            keep.append(a)
            ptrs[c * len(fields) + f] = a.ctypes.data
    _check("pk_gdf_write_host", lib().pk_gdf_write_host(os.fsencode(path), ctypes.byref(meta.c), len(placements),
                                                        cgs, len(fields), ctypes.cast(names, ctypes.c_void_p),
                                                        ctypes.cast(ptrs, ctypes.c_void_p)))


def translate(field):
    out = ctypes.create_string_buffer(64)
    _check("pk_gdf_translate", lib().pk_gdf_translate(field.encode(), out, 64))
    return out.value.decode()


def read(path, grid, name, n_b):
    """Dataset `name` of block `grid` as float64, shape n_b (x, y, z)."""
    out = np.zeros(int(np.prod(n_b)), dtype=np.float64)
    _check("pk_gdf_read", lib().pk_gdf_read(os.fsencode(path), int(grid), name.encode(),
                                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size))
    return out.reshape(tuple(n_b), order="F")


def distance(path1, path2, levels=None, cap=256):
    """bin/gdf_distance: ({(measure, level): value}, total).  level is -1 for
    the measures of the whole file; levels=(l1, l2) compares two given levels."""
    l1, l2 = (-1, -1) if levels is None else levels
    norms = np.zeros(cap)
    lev = (ctypes.c_int32 * cap)()
    names = ctypes.create_string_buffer(64 * cap)
    total = ctypes.c_double()
    n = lib().pk_gdf_distance(os.fsencode(path1), os.fsencode(path2), int(l1), int(l2), cap,
                              norms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), lev, names, 64 * cap,
                              ctypes.byref(total))
    if n < 0:
        _check("pk_gdf_distance", -n)
    keys = names.raw.split(b"\0")[:min(n, cap)]
    return {(k.decode(), int(lev[i])): float(norms[i]) for i, k in enumerate(keys)}, total.value


def main(argv):
    if len(argv) not in (2, 4):
        sys.stderr.write("usage: python -m piernik_amd.gdf file1 file2 [level1 level2]\n")
        return 1
    levels = (int(argv[2]), int(argv[3])) if len(argv) == 4 else None
    try:
        norms, total = distance(argv[0], argv[1], levels)
    except GdfError as e:
        print(f"Cannot compare files `{argv[0]}' and `{argv[1]}': {e}")
        return 2
    width = max(len(k) for k, _ in norms) + 1
    for (k, lv), v in norms.items():
        where = "" if lv < 0 else f" (level {lv})"
        print("Difference of %-*s: %14g%s" % (width, k, v, where))
    print("Total difference between `%s' and `%s': %g" % (argv[0], argv[1], total))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
