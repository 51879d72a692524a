"""GDF-v2 data files and the gdf_distance comparison (libpiernik_gdf.so), on
the CPU: files written from host arrays.

Pinned by:
- the oracle's output fields against closed forms: a state built from known
  primitives gives them back (velocity exactly, pressure and specific
  energy to rounding, with the magnetic energy taken out only for the
  ionized fluid);
- read-back: FP64 datasets return the values written bit for bit, FP32
  datasets the values rounded to nearest;
- the comparison's closed forms (bin/gdf_distance:215-244): identical files
  0, a field doubled 1/3, negated 1, time |t1 - t2| / (t1 + t2), field lists
  1 - |common| / |union|, the total 1 - prod(1 - n); vectors grouped as the
  tool groups them; a domain written as one block and as 2 x 2 x 1 blocks
  compares as identical;
- h5dump (the HDF5 distribution's own tool) for the layout the reference's
  readers expect: groups, attribute names and types, FP32 field datasets.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import datafields as df

gdf = pytest.importorskip("piernik_amd.gdf")
if not os.path.exists(gdf.LIB):
    pytest.skip("libpiernik_gdf.so not built", allow_module_level=True)

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


# ----------------------------------------------------- oracle output fields --

def state_from_prims(rho, v, p, gam, B=None, nb=2, n=(3, 2, 2)):
    ng = tuple(x + 2 * nb for x in n)
    u = np.zeros((5,) + ng, order="F")
    u[0] = rho
    for c in range(3):
        u[1 + c] = rho * v[c]
    e = p / (gam - 1.0) + 0.5 * rho * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2)
    b = None
    if B is not None:
        b = np.zeros((3,) + ng, order="F")
        for c in range(3):
            b[c] = B[c]
        e = e + 0.5 * (B[0] ** 2 + B[1] ** 2 + B[2] ** 2)
    u[4] = e
    return u, b


@pytest.mark.parametrize("mag", [False, True])
def test_oracle_fields_closed_forms(mag):
    gam, rho, v, p = 5.0 / 3.0, 2.0, (0.5, -0.25, 0.125), 0.75
    B = (0.5, 0.25, -1.0) if mag else None
    u, b = state_from_prims(rho, v, p, gam, B)
    psi = np.full(u.shape[1:], 0.125, order="F")
    fl = [(gam, mag)]
    letter = "i" if mag else "n"
    for c, ax in enumerate("xyz"):
        assert np.all(df.datafield(f"vl{ax}{letter}", u, b, psi, fl, 2) == v[c])
        assert np.all(df.datafield(f"mom{ax}{letter}", u, b, psi, fl, 2) == rho * v[c])
    assert np.allclose(df.datafield(f"pre{letter}", u, b, psi, fl, 2), p, rtol=4e-16, atol=0)
    assert np.allclose(df.datafield(f"eth{letter}", u, b, psi, fl, 2), p / ((gam - 1.0) * rho), rtol=4e-16, atol=0)
    assert np.all(df.datafield(f"den{letter}", u, b, psi, fl, 2) == rho)
    assert np.allclose(df.datafield("v", u, b, psi, fl, 2), np.sqrt(sum(x * x for x in v)), rtol=2e-16, atol=0)
    if mag:
        assert np.allclose(df.datafield("magB", u, b, psi, fl, 2), np.sqrt(sum(x * x for x in B)), rtol=2e-16)
        assert np.all(df.datafield("magy", u, b, psi, fl, 2) == B[1])
        assert np.all(df.datafield("psi", u, b, psi, fl, 2) == 0.125)
        with pytest.raises(KeyError):
            df.datafield("pren", u, b, psi, fl, 2)  # a magnetized lone fluid is ionized


def test_oracle_two_fluid_letters():
    """Ion (magnetized, first) and neutral (second): the neutral pressure
    keeps the magnetic energy out, the ion pressure takes it out."""
    gi, gn = 5.0 / 3.0, 1.4
    ui, b = state_from_prims(1.0, (0.5, 0.0, 0.0), 1.0, gi, (1.0, 0.0, 0.0))
    un, _ = state_from_prims(0.5, (0.0, 0.25, 0.0), 2.0, gn)
    u = np.concatenate([ui, un])
    fl = [(gi, True), (gn, False)]
    assert np.allclose(df.datafield("prei", u, b, None, fl, 2), 1.0, rtol=4e-16)
    assert np.allclose(df.datafield("pren", u, b, None, fl, 2), 2.0, rtol=4e-16)
    assert np.all(df.datafield("vlyn", u, b, None, fl, 2) == 0.25)
    assert np.all(df.datafield("denn", u, b, None, fl, 2) == 0.5)


# ---------------------------------------------------------------- files ----

N = (8, 6, 4)
FIELDS = ["deni", "vlxi", "vlyi", "vlzi", "prei", "magx", "magy", "magz"]


def fields_of(seed, n=N):
    rng = np.random.default_rng(seed)
    return [rng.normal(size=n) + (2.0 if f in ("deni", "prei") else 0.0) for f in FIELDS]


def meta(**kw):
    m = gdf.Meta(n_d=N, edge=((0.0, 1.0), (0.0, 0.75), (0.0, 0.5)), **kw)
    return m


def one_block(path, data, **kw):
    gdf.write_host(path, [gdf.Placement(N, dl=(0.125, 0.125, 0.125))], FIELDS, [data], meta(**kw))


def four_blocks(path, data, **kw):
    """The same domain as 2 x 2 x 1 blocks, written in a scrambled order."""
    pl, dd = [], []
    for bx, by in ((1, 0), (0, 1), (1, 1), (0, 0)):
        o = (4 * bx, 3 * by, 0)
        pl.append(gdf.Placement((4, 3, 4), off=o, dl=(0.125, 0.125, 0.125)))
        dd.append([a[o[0]:o[0] + 4, o[1]:o[1] + 3, :] for a in data])
    gdf.write_host(path, pl, FIELDS, dd, meta(**kw))


def test_read_back_fp64_exact_fp32_rounded(tmp_path):
    data = fields_of(1)
    f64, f32 = str(tmp_path / "a64.h5"), str(tmp_path / "a32.h5")
    one_block(f64, data, h5_64bit=1)
    one_block(f32, data)
    for name, a in zip(FIELDS, data):
        dn = gdf.translate(name)
        assert np.array_equal(gdf.read(f64, 0, dn, N), a)
        assert np.array_equal(gdf.read(f32, 0, dn, N), a.astype(np.float32).astype(np.float64))


def test_translate_gdf_names():
    assert [gdf.translate(f) for f in FIELDS] == ["density", "velocity_x", "velocity_y", "velocity_z", "pressure",
                                                  "mag_field_x", "mag_field_y", "mag_field_z"]
    assert gdf.translate("momzn") == "momentum_z" and gdf.translate("ethi") == "specific_energy"
    assert gdf.translate("magB") == "magnetic_field_magnitude" and gdf.translate("psi") == "psi"


def test_identical_and_decomposed_files_are_equal(tmp_path):
    data = fields_of(2)
    a, b = str(tmp_path / "a.h5"), str(tmp_path / "b.h5")
    one_block(a, data, time=0.5)
    four_blocks(b, data, time=0.5)
    norms, total = gdf.distance(a, b)
    assert total == 0.0 and all(v == 0.0 for v in norms.values())
    names = {k for k, _ in norms}
    assert {"datafield `velocity_[x..z]'", "datafield `mag_field_[x..z]'", "datafield `density'",
            "datafield `pressure'", "time", "domains", "resolution", "datafield lists"} <= names


def test_closed_forms(tmp_path):
    data = fields_of(3)
    a, b, c = (str(tmp_path / f"{x}.h5") for x in "abc")
    one_block(a, data, time=1.0, h5_64bit=1)
    doubled = [2.0 * x if f in ("deni",) else x for f, x in zip(FIELDS, data)]
    negated = [-x if f.startswith("mag") else x for f, x in zip(FIELDS, data)]
    mixed = [n if f.startswith("mag") else d for f, d, n in zip(FIELDS, doubled, negated)]
    one_block(b, mixed, time=3.0, h5_64bit=1)
    norms, total = gdf.distance(a, b)
    assert norms[("time", -1)] == 0.5
    assert abs(norms[("datafield `density'", 0)] - 1.0 / 3.0) < 1e-15
    assert norms[("datafield `mag_field_[x..z]'", 0)] == 1.0
    assert norms[("datafield `velocity_[x..z]'", 0)] == 0.0
    prod = 1.0
    for v in norms.values():
        prod *= 1.0 - v
    assert abs(total - (1.0 - prod)) < 1e-15 and total == 1.0  # a measure of 1 makes the total 1
    # fewer fields: the list measure and only the common fields compared
    gdf.write_host(c, [gdf.Placement(N, dl=(0.125,) * 3)], FIELDS[:5], [data[:5]], meta(time=1.0, h5_64bit=1))
    norms, _ = gdf.distance(a, c)
    assert abs(norms[("datafield lists", -1)] - (1.0 - 5.0 / 8.0)) < 1e-15
    assert ("datafield `mag_field_[x..z]'", 0) not in norms


def test_levels_compared_separately(tmp_path):
    """A base level and a refined block at level 1: each level is a measure."""
    data = fields_of(4)
    fine = fields_of(5, n=(4, 4, 4))
    pl = [gdf.Placement(N, dl=(0.125,) * 3), gdf.Placement((4, 4, 4), off=(2, 2, 0), level=1, dl=(0.0625,) * 3)]
    a, b = str(tmp_path / "a.h5"), str(tmp_path / "b.h5")
    gdf.write_host(a, pl, FIELDS, [data, fine], meta(h5_64bit=1))
    gdf.write_host(b, pl, FIELDS, [data, [2.0 * x for x in fine]], meta(h5_64bit=1))
    norms, _ = gdf.distance(a, b)
    assert norms[("datafield `density'", 0)] == 0.0
    assert abs(norms[("datafield `density'", 1)] - 1.0 / 3.0) < 1e-15
    norms01, _ = gdf.distance(a, b, levels=(1, 1))
    assert abs(norms01[("datafield `pressure'", 1)] - 1.0 / 3.0) < 1e-15


def test_refuses_what_the_tool_refuses(tmp_path):
    data = fields_of(6)
    a, b = str(tmp_path / "a.h5"), str(tmp_path / "b.h5")
    one_block(a, data)
    m = gdf.Meta(n_d=N, edge=((0.0, 2.0), (0.0, 0.75), (0.0, 0.5)))
    gdf.write_host(b, [gdf.Placement(N, dl=(0.25, 0.125, 0.125))], FIELDS, [data], m)
    with pytest.raises(gdf.GdfError, match="different size or resolution"):
        gdf.distance(a, b)
    with pytest.raises(gdf.GdfError, match="written as dataset .density."):
        gdf.write_host(str(tmp_path / "c.h5"), [gdf.Placement(N)], ["deni", "denn"], [[data[0], data[0]]], meta())


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_layout_by_h5dump(tmp_path):
    data = fields_of(7)
    a = str(tmp_path / "a.h5")
    one_block(a, data, time=0.25, nstep=7)
    out = subprocess.run([H5DUMP, "-H", a], capture_output=True, text=True, check=True).stdout
    for group in ("gridded_data_format", "simulation_parameters", "data", "grid_0000000000", "domains", "base",
                  "field_types", "dataset_units"):
        assert f'GROUP "{group}"' in out, group
    for ds in ("grid_dimensions", "grid_left_index", "grid_level", "grid_parent_id", "grid_particle_count",
               "density", "velocity_x", "mag_field_z", "length_unit"):
        assert f'DATASET "{ds}"' in out, ds
    for at in ("piernik", "time", "nstep", "cg_count", "n_b", "off", "level", "left_edge", "dl", "n_d",
               "x-edge_position", "refine_by", "dimensionality", "domain_dimensions", "field_to_cgs",
               "field_units", "unique_identifier", "data_software"):
        assert f'ATTRIBUTE "{at}"' in out, at
    # field datasets are FP32 (h5_64bit = .false.), shape (nz, ny, nx)
    dens = out[out.index('DATASET "density"'):].split("DATASET")[1]
    assert "H5T_IEEE_F32LE" in dens and f"( {N[2]}, {N[1]}, {N[0]} )" in dens
    ver = subprocess.run([H5DUMP, "-a", "/piernik", a], capture_output=True, text=True, check=True).stdout
    assert "2.03" in ver and "H5T_IEEE_F64LE" in ver
