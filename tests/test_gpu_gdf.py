"""Output fields computed on the device (pk_datafield) and GDF files written
from device-resident blocks (pk_gdf_write).

After a few solver steps, every output field of the configuration equals the
oracle's restatement of datafields_hdf5 bit for bit.  A file written from the
device equals, by gdf_distance and value by value, one written from the
oracle's fields of the same state.
"""
import os

import numpy as np
import pytest

import piernik_amd as pk
from piernik_amd import setups
from oracle import datafields as df

pytestmark = pytest.mark.gpu

ION = (5.0 / 3.0, True)
NEU = (1.4, False)


def fields_for(fluids):
    mag = any(m for _, m in fluids)
    names = []
    letters = ["i", "n"] if len(fluids) == 2 else (["i"] if mag else ["n"])
    for c in letters:
        names += [f"den{c}", f"vlx{c}", f"vly{c}", f"vlz{c}", f"momx{c}", f"momy{c}", f"momz{c}", f"ene{c}",
                  f"pre{c}", f"eth{c}"]
    names.append("v")
    if mag:
        names += ["magx", "magy", "magz", "magB", "psi"]
    return names


def cases():
    return ["sedov", "mhd", "ion+neutral"]


def start(case):
    pk.load_library()
    pk.finalize()
    if case == "sedov":
        cg, cfg = setups.sedov_hd(n=16)
        pk.init(cfg)
        cg.upload()
        pk.run(cg, 3)
        return [cg], cfg
    if case == "mhd":
        cg, cfg = setups.mhd_turbulence(n=16, seed=5)
        pk.init(cfg)
        cg.upload()
        pk.run(cg, 2)
        return [cg], cfg
    cfg = setups.batch_config(fluids=(ION, NEU))
    cgs = setups.batch_blocks([(12, 10, 8), (8, 8, 8)], cfg, seed=3)
    pk.init(cfg)
    for cg in cgs:
        cg.upload()
    pk.run_all(2)
    return cgs, cfg


@pytest.mark.parametrize("case", cases())
def test_datafields_bit_exact(case):
    cgs, cfg = start(case)
    print(f"datafields {case}", flush=True)
    for cg in cgs:
        cg.download()
        for name in fields_for(cfg.fluids):
            got = pk.datafield(cg, name)
            want = df.datafield(name, cg.u, cg.b, cg.psi, cfg.fluids, cg.nb, cg.has_dir)
            bad = np.argwhere(~(got == want))
            assert bad.size == 0, f"{case} {name}: {len(bad)} cells differ, first {tuple(bad[0])}"
            # FP32 on the device: the round-to-nearest of the FP64 value
            assert np.array_equal(pk.datafield_f32(cg, name), want.astype(np.float32)), name
    with pytest.raises(pk.PiernikError, match="not an output field"):
        pk.datafield(cgs[0], "nonsense")
    pk.finalize()


@pytest.mark.parametrize("case", cases())
def test_device_file_equals_oracle_file(case, tmp_path):
    gdf = pytest.importorskip("piernik_amd.gdf")
    if not os.path.exists(gdf.LIB):
        pytest.skip("libpiernik_gdf.so not built")
    cgs, cfg = start(case)
    print(f"gdf file {case}", flush=True)
    names = fields_for(cfg.fluids)
    strict = len(cfg.fluids) == 1  # gdf_strict names do not tell two fluids apart
    for i, cg in enumerate(cgs):  # one file per block, the block filling its domain
        place = gdf.Placement(cg.n, dl=cg.dl)
        meta = gdf.Meta(time=0.5, n_d=cg.n, nfluids=len(cfg.fluids), gdf_strict=int(strict),
                        edge=tuple((0.0, cg.n[d] * cg.dl[d]) for d in range(3)))
        a, b = str(tmp_path / f"dev{i}.h5"), str(tmp_path / f"ora{i}.h5")
        gdf.write(a, [(cg, place)], names, meta)
        cg.download()
        data = [df.datafield(nm, cg.u, cg.b, cg.psi, cfg.fluids, cg.nb, cg.has_dir) for nm in names]
        gdf.write_host(b, [place], names, [data], meta)
        norms, total = gdf.distance(a, b)
        assert total == 0.0, {k: v for k, v in norms.items() if v}
        for nm, want in zip(names, data):
            got = gdf.read(a, 0, gdf.translate(nm) if strict else nm, cg.n)
            assert np.array_equal(got, want.astype(np.float32).astype(np.float64)), nm
    pk.finalize()
