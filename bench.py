"""Benchmark of the split Riemann sweep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload mhd256|sedov64|batch4096]

A step is one PIERNIK step (time_step + fluid_update_full: two dt-substeps,
X->Y->Z then Z->Y->X, both RK2 stages per sweep) over the whole block, with
the state resident in HBM.  The metric is FP64 cell-updates/s, where one
cell-update advances one cell by one dt (two per step).

The sweep runs the FMA-contracted build (pk_config.fast_math), which agrees
with the bit-exact build to rel-L1 ~1e-15 after 50 steps; --exact times the
bit-exact build.

Default workload: 3D ideal MHD, HLLD + GLM, minmod, CFL 0.5, single periodic
256^3 block with 4 ghost layers, rho=1, p=0.6, gamma=5/3, seeded Gaussian
random v and B = curl A (beta = 1).

--workload batch4096 times the batched block list instead: 4096 periodic
blocks with edges drawn from 16, 24, 32, 48, 64 (seeded), an ionized MHD
fluid (HLLD + GLM) and a neutral fluid (HLLC) per block, smooth seeded
fields; one persistent launch per sweep and fluid replayed from a HIP graph,
against the same launches issued directly and against one launch per block.

Prints one JSON line with the measured throughput, the roofline of the
dominant kernel (the fused sweep; HIP events on the library's stream) and a
CPU baseline timed on this host's cores (the oracle, a C restatement of the
reference, on a bounded sample of the same workload).
"""
import argparse
import contextlib
import json
import multiprocessing as mp
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = "FP64 MHD cell-updates/s at 256^3"
UNSPLIT_CFL = 0.3  # problems/cpaw/problem.par: 0.41 is near the unsplit solver's limit
# the data-file fields of an MHD run (the reference's short names, hdf_vars)
GDF_FIELDS_MHD = ["deni", "vlxi", "vlyi", "vlzi", "momxi", "momyi", "momzi", "enei", "prei", "ethi", "v",
                  "magx", "magy", "magz", "magB", "psi"]


def _cpu_worker(args):
    """One CPU worker: the oracle on its own block of the workload for ~budget s."""
    workload, n, seed, budget = args
    from oracle import pyoracle as po
    from piernik_amd import setups
    if workload in ("mhd256", "unsplit256", "stage256", "gdf256"):
        cg, cfg = setups.mhd_turbulence(n=n, seed=seed)
    elif workload == "batch4096":
        cfg = setups.batch_config()
        cg = setups.batch_blocks([n], cfg, seed=seed)[0]
    else:
        cg, cfg = setups.sedov_hd(n=n)
    p = po.params_from_config(cfg)
    ob = po.OBlock.from_cg(cg)
    t0 = time.perf_counter()
    steps = 0
    if workload == "stage256":
        # the drop-in operator: RK2 stages of solve_cg along x, y, z
        while True:
            for d in range(3):
                for istep in (10001, 10002):
                    po.solve_cg(p, ob, d, istep, 1e-4, 1.0)
                    steps += 1
            if time.perf_counter() - t0 >= budget:
                break
        return steps * n ** 3, time.perf_counter() - t0
    if workload == "gdf256":
        # output fields in numpy (the oracle's datafields_hdf5) and the file
        # from host arrays, into a scratch directory of this host
        import tempfile
        from oracle import datafields as df
        from piernik_amd import gdf
        written = 0
        with tempfile.TemporaryDirectory() as tmp:
            meta = gdf.Meta(n_d=(n, n, n), nfluids=1)
            place = gdf.Placement((n, n, n), dl=cg.dl)
            while True:
                data = [df.datafield(nm, ob.u, ob.b, ob.psi, cfg.fluids, cg.nb) for nm in GDF_FIELDS_MHD]
                path = os.path.join(tmp, f"f{steps}.h5")
                gdf.write_host(path, [place], GDF_FIELDS_MHD, [data], meta)
                written += os.path.getsize(path)
                os.remove(path)
                steps += 1
                if time.perf_counter() - t0 >= budget:
                    break
        return written, time.perf_counter() - t0
    if workload == "unsplit256":
        p.cfl = UNSPLIT_CFL
        tt = 0.0
        while True:
            dt, _ = po.time_step(p, ob, steps, -1.0)
            po.fluid_update_unsplit(p, ob, dt)
            tt += dt
            steps += 1
            if time.perf_counter() - t0 >= budget:
                break
        return steps * n ** 3, time.perf_counter() - t0
    while True:
        po.run(p, ob, 1)
        steps += 1
        if time.perf_counter() - t0 >= budget:
            break
    el = time.perf_counter() - t0
    return steps * n ** 3 * 2, el


def cpu_baseline(workload, budget=10.0):
    """The oracle on all host cores: one process per core, each advancing its
    own block of the same problem (64^3; 32^3 two-fluid blocks for the
    batched list) for about `budget` seconds."""
    ncpu = os.cpu_count() or 1
    cores = min(ncpu, int(os.environ.get("OMP_NUM_THREADS", "16")), 16)
    nblk = 32 if workload == "batch4096" else 64
    ctx = mp.get_context("fork")
    if workload == "gdf256":
        # one writer process and one file, as the device path writes it:
        # the whole 256^3 block, numpy fields and the HDF5 write on one core
        with ctx.Pool(1) as pool:
            written, el = pool.map(_cpu_worker, [(workload, 256, 100, budget)])[0]
        return {
            "value": written / el / 1e9,
            "unit": "GB/s of data file",
            "cores": 1,
            "kind": "port",
            "sample": (f"oracle output fields (numpy restatement of datafields_hdf5) of the {len(GDF_FIELDS_MHD)} "
                       f"fields of a 256^3 MHD block and its GDF-v2 file written from host arrays, one process, "
                       f"files written until ~{budget:.0f} s had passed; {written} file bytes in {el:.2f} s"),
        }
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        res = pool.map(_cpu_worker, [(workload, nblk, 100 + i, budget) for i in range(cores)])
    wall = time.perf_counter() - t0
    updates = sum(r[0] for r in res)
    slowest = max(r[1] for r in res)
    return {
        "value": updates / slowest,
        "unit": "cell-stages/s" if workload == "stage256" else "cell-updates/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"oracle (C restatement of the reference sweep built with the reference's benchmark flags "
                   f"-O3 -funroll-loops, no allocation per pencil, one thread per process) on {cores} "
                   f"independent periodic {nblk}^3 blocks of the same "
                   f"{ {'mhd256': 'MHD HLLD+GLM', 'unsplit256': 'MHD HLLD+GLM, unsplit solver', 'stage256': 'MHD HLLD+GLM, solve_cg stages', 'batch4096': 'two-fluid (MHD HLLD+GLM + HD HLLC)'}.get(workload, 'HD HLLC')} "
                   f"problem, {'RK2 stages along x, y, z' if workload == 'stage256' else 'whole PIERNIK steps'} for ~{budget:.0f} s each; {updates} cell-updates in {slowest:.2f} s "
                   f"(wall {wall:.2f} s)"),
    }


class TimedRegion:
    """Marks the timed steps for rocprofv3 (rocprofiler-sdk's roctx): a range
    named TIMED_RANGE, which --marker-trace records with its timestamps, so
    that trace_summary() keeps exactly the timed launches of a kernel trace
    (no warm-up, no autotuning, no other workload); and
    roctxProfilerResume(0) / roctxProfilerPause(0) around it for
    --selected-regions.  Without a profiler attached the calls do nothing;
    without the library the region is not marked."""

    def __init__(self):
        import ctypes
        self.lib = None
        for path in ("/opt/rocm/lib/librocprofiler-sdk-roctx.so", "librocprofiler-sdk-roctx.so"):
            try:
                self.lib = ctypes.CDLL(path)
                break
            except OSError:
                continue

    def __enter__(self):
        if self.lib is not None:
            self.lib.roctxRangePushA(TIMED_RANGE.encode())
            self.lib.roctxProfilerResume(0)
        return self

    def __exit__(self, *exc):
        if self.lib is not None:
            self.lib.roctxProfilerPause(0)
            self.lib.roctxRangePop()
        return False


# the roctx range around the timed steps (rocprofv3 --marker-trace records it)
TIMED_RANGE = "bench timed steps"


def trace_summary(d):
    """Per-kernel statistics of the dispatches inside the timed range, from
    a rocprofv3 run with --kernel-trace --marker-trace -f csv (output under
    directory d): calls, total ms, average, min and max us, share."""
    import csv
    import glob

    def rows(pattern):
        out = []
        for path in glob.glob(os.path.join(d, "**", pattern), recursive=True):
            with open(path, newline="") as f:
                out.extend(csv.DictReader(f))
        return out

    marks = [r for r in rows("*marker_api_trace.csv") if TIMED_RANGE in " ".join(str(v) for v in r.values())]
    if not marks:
        raise SystemExit(f"no '{TIMED_RANGE}' range in the marker trace under {d}")
    t0 = min(int(r["Start_Timestamp"]) for r in marks)
    t1 = max(int(r["End_Timestamp"]) for r in marks)
    stats = {}
    for r in rows("*kernel_trace.csv"):
        a, b = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        if a >= t0 and b <= t1:
            stats.setdefault(r["Kernel_Name"], []).append((b - a) * 1e-3)
    total = sum(sum(v) for v in stats.values())
    lines = [f"# timed range: {(t1 - t0) * 1e-6:.3f} ms, kernels inside it: {total * 1e-3:.3f} ms",
             "# columns: calls, total ms, average us, min us, max us, share %, kernel"]
    for name, v in sorted(stats.items(), key=lambda kv: -sum(kv[1])):
        lines.append(f"{len(v):6d} {sum(v) * 1e-3:10.3f} {sum(v) / len(v):12.2f} {min(v):12.2f} {max(v):12.2f} "
                     f"{100.0 * sum(v) / total:7.2f}  {name}")
    return "\n".join(lines)


def counter_summary(d):
    """Per-kernel means over dispatches of the counters in a rocprofv3 --pmc
    -f csv output directory d (one row per dispatch and counter): kernel,
    dispatches, then counter = mean value."""
    import csv
    import glob

    vals = {}
    for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                k = vals.setdefault(r["Kernel_Name"], {})
                k.setdefault(r["Counter_Name"], {})[r.get("Dispatch_Id", len(k))] = float(r["Counter_Value"])
    lines = []
    for name, counters in sorted(vals.items()):
        nd = max(len(v) for v in counters.values())
        lines.append(f"{name}  ({nd} dispatches)")
        for c, v in sorted(counters.items()):
            lines.append(f"   {c:28s} {sum(v.values()) / len(v):14.6e}")
    return "\n".join(lines)


def load_traffic(name):
    """HBM bytes per sweep launch from committed rocprofv3 --pmc runs, if any."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("traffic_bytes_per_launch")


FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector (and matrix) peak
SIMDS, CLOCK_HZ = 1024, 2.4e9  # 256 CUs x 4 SIMDs, peak engine clock


def fp64_view(prof, cells, name="sweep_mhd256_fp64_flops.json"):
    """The sweep seen as the FP64 arithmetic it issues: committed per-cell
    instruction counts (rocprofv3 --pmc SQ_INSTS_VALU*, one launch per
    direction) times the cells of each timed launch, over the launches' HIP
    event time.  valu_issue_frac: wave-level VALU instructions at 4 cycles
    each (64 lanes on a 16-lane SIMD) over the SIMD cycles of that time."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        per = json.load(f)["per_dir"]
    flops = sum(per[d]["fp64_flops_per_cell"] * cells * prof["launches"][d] for d in range(3))
    insts = sum(per[d]["valu_insts_per_cell"] * cells / 64.0 * prof["launches"][d] for d in range(3))
    sec = sum(prof["ms"][d] for d in range(3)) * 1e-3
    if sec <= 0:
        return None
    achieved = flops / sec / 1e12
    with open(path) as f:
        busy = json.load(f).get("valu_busy_counters")
    return {
        "achieved_tflops": achieved,
        "peak_tflops": FP64_PEAK_TFLOPS,
        "frac": achieved / FP64_PEAK_TFLOPS,
        # at the 2.4 GHz peak engine clock; the clock under this load is lower
        "valu_issue_frac": insts * 4.0 / (sec * SIMDS * CLOCK_HZ),
        # SQ_ACTIVE_INST_VALU over the GPU's own cycles (GRBM_GUI_ACTIVE), per direction
        "valu_busy_counters": busy,
        "fp64_flops_per_cell": [per[d]["fp64_flops_per_cell"] for d in range(3)],
        "counts": f"profiles/{name}",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["mhd256", "sedov64", "batch4096", "adjoint128", "ensemble256", "copy",
                                           "unsplit256", "stage256", "gdf256", "variants"],
                    default="mhd256")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost measurement")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched-list speed-up (4096 blocks, graph vs per-block launches)")
    ap.add_argument("--nblocks", type=int, default=4096, help="blocks of the batched list")
    ap.add_argument("--n", type=int, default=None, help="override the block size (cells per side)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--tile-x", type=int, nargs=2, default=(0, 0))
    ap.add_argument("--tile-yz", type=int, nargs=2, default=(0, 0))
    ap.add_argument("--no-autotune", action="store_true", help="keep the default tile shapes")
    ap.add_argument("--blocks", type=int, nargs=3, default=None,
                    help="mhd256: cut the domain into BX x BY x BZ linked blocks advanced as one list")
    ap.add_argument("--no-fc", action="store_true",
                    help="stage256: time only the plain stages (no flux points, no PCIe round trips)")
    ap.add_argument("--exact", action="store_true",
                    help="time the bit-exact build of the sweep instead of the FMA-contracted one")
    ap.add_argument("--trace-summary", metavar="DIR", default=None,
                    help="print per-kernel statistics of the timed steps from a rocprofv3 "
                         "--kernel-trace --marker-trace -f csv output directory, then exit")
    ap.add_argument("--counter-summary", metavar="DIR", default=None,
                    help="print per-kernel mean counter values from a rocprofv3 --pmc -f csv output directory, "
                         "then exit")
    args = ap.parse_args()
    if args.trace_summary:
        print(trace_summary(args.trace_summary))
        return
    if args.counter_summary:
        print(counter_summary(args.counter_summary))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: there is no multi-GPU path to run "
              "(the domain would shard as z slabs exchanging nb ghost layers); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    # CPU baseline first, before this process touches the GPU
    cpu = None if args.no_cpu else cpu_baseline(args.workload, args.cpu_budget)
    print(f"[bench] cpu baseline done: {cpu and round(cpu['value'])} cell-updates/s", file=sys.stderr, flush=True)
    if args.workload == "copy":
        gbps = copy_ceiling(steps=args.steps)
        print(json.dumps({
            "metric": "device-to-device copy bandwidth", "value": gbps, "unit": "GB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": 1, "ms_per_step": 2 * COPY_BYTES / gbps / 1e6, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"torch copy_ of {COPY_BYTES >> 30} GiB between two device tensors "
                                   "(read + write counted)"}}))
        return
    if args.workload == "batch4096":
        print(json.dumps(run_batch(args, cpu)))
        return
    if args.workload == "unsplit256":
        print(json.dumps(run_unsplit(args, cpu)))
        return
    if args.workload == "stage256":
        print(json.dumps(run_stage(args, cpu)))
        return
    if args.workload == "gdf256":
        print(json.dumps(run_gdf(args, cpu)))
        return
    if args.workload == "variants":
        print(json.dumps(run_variants(args)))
        return
    if args.workload == "mhd256" and args.blocks:
        print(json.dumps(run_blocks(args)))
        return
    if args.workload == "ensemble256":
        print(json.dumps(run_ensemble(args)))
        return
    if args.workload == "adjoint128":
        adj = adjoint_cost()
        print(json.dumps({
            "metric": "adjoint/forward cost ratio, 128^3 hydro rollout", "value": adj["ratio"], "unit": "x",
            "n_gpus": 1, "steps": adj["steps"], "warmup": 1, "ms_per_step": adj["forward_ms"] / adj["steps"],
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic", "config": {"workload": adj["workload"]}, "adjoint": adj}))
        return

    import piernik_amd as pk
    from piernik_amd import setups

    if args.workload == "mhd256":
        n = args.n or 256
        cg, cfg = setups.mhd_turbulence(n=n, seed=1)
        nv = 9
    else:
        n = args.n or 64
        cg, cfg = setups.sedov_hd(n=n)
        nv = 5
    cfg.tile_x = tuple(args.tile_x)
    cfg.tile_yz = tuple(args.tile_yz)
    cfg.fast_math = not args.exact
    print(f"[bench] initial state {n}^3 ready", file=sys.stderr, flush=True)

    pk.init(cfg)
    cg.upload()
    pk.run(cg, args.warmup)
    tiles = None
    if not args.no_autotune and args.tile_x == (0, 0) and args.tile_yz == (0, 0):
        tiles = pk.autotune(cg)
    pk.synchronize()
    print(f"[bench] warmup done, tiles {tiles}", file=sys.stderr, flush=True)

    pk.profile_begin()
    pk.synchronize()
    with TimedRegion():
        t0 = time.perf_counter()
        t_sim, dt_last = pk.run(cg, args.steps)
        pk.synchronize()
        t1 = time.perf_counter()
    prof = pk.profile_end()
    el = t1 - t0
    cells = n ** 3
    value = cells * 2 * args.steps / el

    launches = sum(prof["launches"])
    ms = sum(prof["ms"])
    algo = sum(prof["algo_bytes"])
    achieved = (algo / launches) / (ms / launches * 1e-3) / 1e9 if launches else 0.0
    per_dir = [(prof["algo_bytes"][d] / (prof["ms"][d] * 1e-3) / 1e9) if prof["ms"][d] > 0 else None
               for d in range(3)]

    rep = pk.get_report(cg)
    pk.finalize()
    copy_gbps = copy_ceiling()
    adj = None
    if not args.no_adjoint:
        print("[bench] adjoint cost ratio (128^3 hydro rollout)", file=sys.stderr, flush=True)
        adj = adjoint_cost()
    batching = None
    if not args.no_batch and args.workload == "mhd256":
        print(f"[bench] batched list: {args.nblocks} blocks, graph vs per-block launches", file=sys.stderr, flush=True)
        batching = batch_speedup(args)
    line = {
        "metric": METRIC if args.workload == "mhd256" else "FP64 HD cell-updates/s at 64^3",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"3D ideal MHD, HLLD+GLM, minmod, RK2, CFL 0.5, single periodic {n}^3 block, nb=4, "
                         "Gaussian random v and B=curl A, beta=1" if args.workload == "mhd256" else
                         f"3D Sedov blast, HD, HLLC, van Leer, RK2, CFL 0.7, single periodic {n}^3 block, nb=4"),
            "cells": cells,
            "nvar": nv,
            "step": "time_step + fluid_update_full (2 dt-substeps x 3 fused RK2 sweeps)",
            "sim_time": t_sim,
            "last_dt": dt_last,
            "parallelism": "1 GPU",
            "arithmetic": ("FP64 with FMA contraction and fast division/sqrt; rel-L1 vs the bit-exact path "
                           "8.4e-16 after 50 steps at 256^3 (bar 1e-10, tests/test_gpu_fast_math.py)"
                           if not args.exact else "FP64, every operation rounded as the reference's code (bit-exact)"),
            "tiles_pencils_x_cells": {"xyz"[d]: list(t) for d, t in (tiles or {}).items()},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": load_traffic(f"sweep_{args.workload}_traffic.json") if n in (256, 64) else None,
            "kernel": (f"fused RK2 sweep ({'HLLD+GLM' if nv == 9 else 'HLLC'}), x/y/z: "
                       + ", ".join(kernel_name((tiles or {}).get(d, (0, 0))[1]) + " " + "xyz"[d] for d in range(3))),
            "algo_bytes_per_launch": algo / launches if launches else None,
            "avg_launch_ms": ms / launches if launches else None,
            "per_dir_gbps": per_dir,
            "measured_copy_gbps": copy_gbps,
            "frac_of_measured_copy": achieved / copy_gbps,
        },
        # the same launches as FP64 arithmetic: the bound the sweep actually meets
        "fp64_valu": fp64_view(prof, cells) if (n == 256 and nv == 9 and not args.exact) else None,
        "cpu_baseline": cpu,
        "floors": rep,
        "adjoint_forward_cost_ratio": adj and adj["ratio"],
        "adjoint": adj,
        "batching": batching,
    }
    print(json.dumps(line))


# Sweep kernel shapes compared by --workload variants: (pencils, cells) as in
# pk_config tile_x / tile_yz; cells > 0 the tile kernel (TP x TS lanes),
# -1 the march kernel, -m <= -2 the stream kernel (m iterations per workgroup).
VARIANTS_X = [(1, 64), (1, -64), (2, -32), (4, -32), (8, -32)]
VARIANTS_YZ = [(4, 64), (64, -1), (4, -16), (8, -16), (8, -32)]


def run_variants(args, reps=10):
    """Each sweep kernel shape of VARIANTS_X / VARIANTS_YZ on the 256^3 MHD
    block (relaxed build unless --exact): `reps` fused sweeps along each
    direction after one untimed, HIP events around every launch; GB/s of the
    algorithmic bytes (144 B per cell and sweep) and the same as a fraction
    of 8 TB/s."""
    import piernik_amd as pk
    from piernik_amd import setups

    n = args.n or 256
    cg, cfg = setups.mhd_turbulence(n=n, seed=1)
    cfg.fast_math = not args.exact
    pk.init(cfg)
    cg.upload()
    dt = pk.timestep_fluid(cg)[0]  # the CFL step of the state: no floor fires
    pk.finalize()
    out = {}
    for d in range(3):
        rows = []
        for shape in (VARIANTS_X if d == 0 else VARIANTS_YZ):
            if d == 0:
                cfg.tile_x, cfg.tile_yz = shape, (0, 0)
            else:
                cfg.tile_x, cfg.tile_yz = (0, 0), shape
            pk.init(cfg)
            cg.registered = False
            cg.upload()
            ch = pk.update_chspeed(cg, dt, cfg)
            pk.sweep(cg, d, dt, ch)
            pk.synchronize()
            pk.profile_begin()
            for _ in range(reps):
                pk.sweep(cg, d, dt, ch)
            prof = pk.profile_end()
            ms = prof["ms"][d] / prof["launches"][d]
            gbps = prof["algo_bytes"][d] / prof["launches"][d] / (ms * 1e-3) / 1e9
            rows.append({"shape": list(shape), "kernel": kernel_name(shape[1]), "ms": ms, "gbps": gbps,
                         "frac": gbps / HBM_PEAK_GBPS})
            print(f"[bench] {'xyz'[d]} {shape}: {ms:.3f} ms, {gbps:.0f} GB/s", file=sys.stderr, flush=True)
            pk.finalize()
        out["xyz"[d]] = rows
    best = {k: max(v, key=lambda r: r["gbps"]) for k, v in out.items()}
    return {
        "metric": "fused sweep GB/s per kernel shape, 256^3 MHD", "value": max(r["gbps"] for r in best.values()),
        "unit": "GB/s", "n_gpus": 1, "steps": reps, "warmup": 1, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3D ideal MHD, HLLD+GLM, minmod, single periodic {n}^3 block, one fused sweep per "
                               f"launch, {'bit-exact' if args.exact else 'relaxed'} build"},
        "variants": out, "best": best,
    }


def batch_speedup(args, steps=3):
    """The batched two-fluid block list (args.nblocks
    blocks with edges from {16, 24, 32, 48, 64}, seed 0, ionized MHD +
    neutral HD per block), generated on the GPU: ms per step of the
    graph-captured persistent launches and of one launch per block, after an
    untimed warm-up step each; the same state and bit-identical results."""
    import piernik_amd as pk
    from piernik_amd import setups

    sizes = setups.batch_sizes(args.nblocks, seed=0)
    cfg = setups.batch_config()
    cfg.fast_math = not args.exact
    pk.init(cfg)
    setups.batch_blocks_device(sizes, cfg, seed=0)
    pk.synchronize()
    t0 = time.perf_counter()
    if not args.no_autotune:
        pk.autotune_all()
    t_tune = time.perf_counter() - t0
    ms = {}
    for name, mode in (("graph", pk.BATCH_GRAPH), ("per_block", pk.BATCH_PER_BLOCK)):
        pk.run_all(1, mode)
        pk.synchronize()
        t0 = time.perf_counter()
        pk.run_all(steps, mode)
        pk.synchronize()
        ms[name] = (time.perf_counter() - t0) / steps * 1e3
        print(f"[bench] batched list {name}: {ms[name]:.1f} ms/step", file=sys.stderr, flush=True)
    pk.finalize()
    cells = sum(n ** 3 for n in sizes)
    return {
        "workload": (f"{len(sizes)} periodic blocks, edges {({str(n): sizes.count(n) for n in sorted(set(sizes))})} "
                     "(seed 0), nb=4, ionized MHD (HLLD+GLM) + neutral HD (HLLC) per block, van Leer, RK2, CFL 0.7"),
        "cells": cells,
        "steps": steps,
        "ms_per_step": ms,
        "speedup_graph_vs_per_block": ms["per_block"] / ms["graph"],
        "autotune_s": t_tune,
    }


def kernel_name(ts):
    """The sweep kernel a tile length selects (pk_config tile_x / tile_yz)."""
    return "k_march" if ts == -1 else ("k_stream" if ts <= -2 else "k_sweep")


def adjoint_cost(n=128, steps=32, every=8):
    """Cost of the gradient of a hydro rollout: t(forward with checkpoints +
    backward with segment recompute) / t(forward), on the 128^3 HD problem
    (rho = 1 + 0.1 G, p = 1, v = 0, HLLC, van Leer, CFL 0.5, 32 steps,
    checkpoint every 8, loss 0.5*||rho - rho_target||^2)."""
    import torch
    import piernik_amd as pk
    from piernik_amd import adjoint, setups

    u, target, cfg = setups.adjoint_hydro(n, seed=0)
    pk.init(cfg)
    plan = adjoint.RolloutPlan(cfg, n, steps, checkpoint_every=every)
    u0 = setups.to_device_layout(u)
    tgt = torch.as_tensor(target).cuda()

    def grad_of(uf):
        g = torch.zeros_like(uf)
        g[4:-4, 4:-4, 4:-4, 0] = (uf[4:-4, 4:-4, 4:-4, 0] - tgt.permute(2, 1, 0))
        return g

    def timed(plan):
        uf = plan.run(u0)  # warm-up
        plan.backward(grad_of(uf))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.run(u0, record=False)
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t0
        t0 = time.perf_counter()
        uf = plan.run(u0, record=True)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        g = grad_of(uf)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.backward(g)
        torch.cuda.synchronize()
        t_bwd = time.perf_counter() - t0
        plan.close()
        return t_plain, t_fwd, t_bwd

    t_plain, t_fwd, t_bwd = timed(plan)
    # every sweep's input recorded during the forward pass (one segment, no
    # recomputation): what the reverse sweeps themselves cost
    t_plain_r, t_fwd_r, t_bwd_r = timed(adjoint.RolloutPlan(cfg, n, steps, checkpoint_every=steps))
    pk.finalize()
    return {
        "ratio": (t_fwd + t_bwd) / t_plain,
        "forward_ms": t_plain * 1e3,
        "forward_checkpointed_ms": t_fwd * 1e3,
        "backward_ms": t_bwd * 1e3,
        "ratio_all_sweeps_recorded": (t_fwd_r + t_bwd_r) / t_plain_r,
        "all_sweeps_recorded_ms": {"forward": t_plain_r * 1e3, "forward_recorded": t_fwd_r * 1e3,
                                   "backward": t_bwd_r * 1e3},
        "steps": steps,
        "checkpoint_every": every,
        "workload": (f"3D hydro, HLLC, van Leer, RK2, CFL 0.5, periodic {n}^3, rho = 1 + 0.1 G (|k| <= 6), "
                     f"p = 1, v = 0, {steps} steps, checkpoint every {every}, loss 0.5*||rho - rho_target||^2"),
        "definition": ("(forward with checkpoints + backward with segment recompute) / plain forward; the last "
                       "segment's sweeps are recorded by the forward pass, the others recomputed"),
    }


def run_ensemble(args, ninst=256, steps=40, every=8):
    """Gradient of the total magnetic energy of 256 independent 2D
    Orszag-Tang instances (sizes 64^2 .. 256^2, amplitudes 1 +- 0.1, HLLD +
    GLM, van Leer, CFL 0.5) after 40 steps w.r.t. their initial fields: one
    persistent launch per sweep for the forward steps of all instances and
    one per sweep for their adjoint; checkpoint every 8 steps."""
    import torch
    import piernik_amd as pk
    from piernik_amd import adjoint, setups

    cases, cfg = setups.orszag_tang_ensemble(ninst, seed=0)
    hd = cfg.has_dir
    pk.init(cfg)
    ens = adjoint.EnsemblePlan(cfg, [(n, n, 1) for n, *_ in cases], steps, checkpoint_every=every)
    states = [tuple(setups.to_device_layout(a, 4, hd) for a in (u, b, psi)) for _, u, b, psi in cases]
    cells = sum(n * n for n, *_ in cases)

    def grad_of(st):
        g = [torch.zeros_like(st[0]), torch.zeros_like(st[1]), torch.zeros_like(st[2])]
        setups.interior_of(g[1], 4, hd).copy_(setups.interior_of(st[1], 4, hd))
        return g

    print(f"[bench] {ninst} instances, {cells} cells", file=sys.stderr, flush=True)
    fin = ens.run(states)  # warm-up
    ens.backward([grad_of(f) for f in fin])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.run(states, record=False)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    fin = ens.run(states, record=True)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    g = [grad_of(f) for f in fin]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    grads = ens.backward(g)
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    energy = sum(0.5 * float(torch.sum(setups.interior_of(f[1], 4, hd) ** 2)) for f in fin)
    gnorm = sum(float(torch.sum(x * x)) for gr in grads for x in gr) ** 0.5
    ens.close()
    pk.finalize()
    ratio = (t_fwd + t_bwd) / t_plain
    sizes = {}
    for n, *_ in cases:
        sizes[str(n)] = sizes.get(str(n), 0) + 1
    return {
        "metric": "adjoint/forward cost ratio, 256-instance 2D MHD ensemble", "value": ratio, "unit": "x",
        "n_gpus": 1, "steps": steps, "warmup": 1, "ms_per_step": t_plain / steps * 1e3, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": (f"{ninst} independent 2D Orszag-Tang instances, sizes {sizes}, amplitudes 1 +- 0.1 "
                                f"(seed 0), HLLD+GLM, van Leer, RK2, CFL 0.5, {steps} steps, checkpoint every {every}, "
                                "gradient of the total magnetic energy w.r.t. the initial u, b, psi"),
                   "cells": cells},
        "ensemble": {"forward_ms": t_plain * 1e3, "forward_checkpointed_ms": t_fwd * 1e3, "backward_ms": t_bwd * 1e3,
                     "cell_updates_per_s": cells * 2 * steps / t_plain, "magnetic_energy": energy,
                     "gradient_norm": gnorm},
    }


COPY_BYTES = 1 << 30


def copy_ceiling(steps=20):
    """Measured HBM ceiling: GB/s of a plain device-to-device copy of 1 GiB
    (bytes read + bytes written over the kernel time).  The same command
    under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE gives the counters' scale
    for a kernel of known traffic."""
    import torch
    a = torch.full((COPY_BYTES // 8,), 1.0, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(max(1, steps)):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / max(1, steps)
    del a, b
    return 2 * COPY_BYTES / (ms * 1e-3) / 1e9


def run_unsplit(args, cpu):
    """The 256^3 MHD turbulence block advanced by the unsplit solver
    (RIEMANN_UNSPLIT): one step is one dt, so one cell-update per cell and
    step.  Algorithmic bytes: each of the two RK stages reads the state and
    writes it once, 2 x 144 B per cell-update."""
    import piernik_amd as pk
    from piernik_amd import setups

    n = args.n or 256
    cg, cfg = setups.mhd_turbulence(n=n, seed=1)
    cfg.cfl = UNSPLIT_CFL
    cfg.fast_math = not args.exact
    pk.init(cfg)
    cg.upload()
    pk.run_unsplit(cg, 1e30, nmax=max(1, args.warmup))
    pk.synchronize()
    t0 = time.perf_counter()
    t_sim, nst = pk.run_unsplit(cg, 1e30, nmax=args.steps)
    pk.synchronize()
    el = time.perf_counter() - t0
    rep = pk.get_report(cg)
    pk.finalize()
    cells = n ** 3
    algo = 2 * 144.0 * cells * args.steps
    achieved = algo / el / 1e9
    return {
        "metric": "FP64 MHD cell-updates/s at 256^3, unsplit solver",
        "value": cells * args.steps / el,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"3D ideal MHD, HLLD+GLM, minmod, unsplit RK2 (solver_str = 'unsplit'), CFL {UNSPLIT_CFL}, "
                         f"single periodic {n}^3 block, nb=4, Gaussian random v and B=curl A, beta=1"),
            "cells": cells,
            "step": "time_step + fluid_update_unsplit (2 stages: fluxes of x, y, z, then the update)",
            "sim_time": t_sim,
            "arithmetic": ("FP64 with FMA contraction and fast division/sqrt (bar rel-L1 1e-10 against the "
                           "bit-exact build, tests/test_gpu_unsplit.py)" if not args.exact else
                           "FP64, every operation rounded as the reference's code (bit-exact)"),
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": None,
            "kernel": "whole step (k_unsplit_fused_tile x3 per stage)",
            "algo_bytes_per_step": algo / args.steps,
        },
        "cpu_baseline": cpu,
        "floors": rep,
    }


def run_stage(args, cpu):
    """The drop-in operator: pk_solve_cg (one RK stage of one direction, the
    solve_cg_sub call of sweeps.F90:245) on the 256^3 MHD block, RK2 stages
    along x, y and z with the periodic ghost refresh between them, as sweep()
    runs them.  Timed three ways: device-resident without fine/coarse flux
    points; device-resident with flux points on every pencil end (every
    pencil has li, lo and ro: the worst case of an AMR level); and as the
    Fortran adapter drives it, with the block uploaded and downloaded around
    every call.  Algorithmic bytes of a stage: the state read and written
    once per interior cell (2 x 144 B; the second stage also reads U0, which
    is not counted)."""
    import numpy as np
    import piernik_amd as pk
    from piernik_amd import setups

    n = args.n or 256
    cg, cfg = setups.mhd_turbulence(n=n, seed=1)
    pk.init(cfg)
    cg.upload()
    pk.run(cg, 1)
    dt, ch = 1e-4, 1.0
    print(f"[bench] stage {n}^3 ready", file=sys.stderr, flush=True)

    def stages(k):
        for _ in range(k):
            for d in range(3):
                for istep in (pk.RK2_1, pk.RK2_2):
                    pk.solve_cg(cg, d, istep, dt, ch)
                    pk.update_boundaries(cg)

    def timed(k):
        stages(1)
        pk.synchronize()
        pk.profile_begin()
        t0 = time.perf_counter()
        stages(k)
        pk.synchronize()
        el = time.perf_counter() - t0
        return el, pk.profile_end()

    k = max(1, args.steps)
    el_plain, prof = timed(k)

    def with_points(every):
        """Flux points (lo, ro and li) on the pencils whose transverse
        indices are both multiples of `every` (1: every pencil, the worst case
        of a refined level), every direction; li at the last cell replaces its
        right face by the flux ro saved there in the previous call."""
        refs = []
        for d in range(3):
            o1, o2 = ((1, 2), (2, 0), (0, 1))[d]
            n1, n2 = cg.n[o1], cg.n[o2]
            on = np.zeros((n1, n2), dtype=bool)
            on[::every, ::every] = True
            saved = [np.zeros((5, n1, n2), order="F") for _ in range(2)]
            bsaved = [np.zeros((4, n1, n2), order="F") for _ in range(2)]
            first = np.asfortranarray(np.where(on, 0, -1).astype(np.int32))
            last = np.asfortranarray(np.where(on, n - 1, n).astype(np.int32))
            last_lo = np.asfortranarray(np.where(on, n - 1, -1).astype(np.int32))
            pk.register_fcflx(cg, d, 0, pk.FC_COARSEBND, 0, first, saved[0], bsaved[0])
            pk.register_fcflx(cg, d, 1, pk.FC_COARSEBND, 0, last, saved[1], bsaved[1])
            pk.register_fcflx(cg, d, 0, pk.FC_FINEBND, 0, last_lo, saved[1], bsaved[1])
            refs.append((saved, bsaved, first, last, last_lo))
        el, pr = timed(k)
        for d in range(3):
            for s_ in (0, 1):
                for kind in (pk.FC_FINEBND, pk.FC_COARSEBND):
                    pk.register_fcflx(cg, d, s_, kind, 0, None, None, None)
        return el, pr

    el_fc = el_fc16 = None
    prof_fc = {"algo_bytes": [0.0] * 4, "ms": [1.0] * 4}
    if not args.no_fc:
        el_fc, prof_fc = with_points(1)
        el_fc16, _ = with_points(4)
    # as the adapter drives it: upload, one stage, download
    kp = 0 if args.no_fc else max(1, k // 4)
    t0 = time.perf_counter()
    for _ in range(kp):
        for d in range(3):
            for istep in (pk.RK2_1, pk.RK2_2):
                cg.upload()
                pk.solve_cg(cg, d, istep, dt, ch)
                cg.download()
    el_pcie = time.perf_counter() - t0
    pk.finalize()

    cells = n ** 3
    nst = 6 * k
    algo = sum(prof["algo_bytes"][:3])
    kms = sum(prof["ms"][:3])
    launches = sum(prof["launches"][:3])
    achieved = algo / (kms * 1e-3) / 1e9
    return {
        "metric": "FP64 MHD cell-stages/s at 256^3, drop-in solve_cg (one RK stage of one direction per call)",
        "value": cells * nst / el_plain,
        "unit": "cell-stages/s",
        "n_gpus": 1,
        "steps": k,
        "warmup": 1,
        "ms_per_step": el_plain / k * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"3D ideal MHD, HLLD+GLM, minmod, single periodic {n}^3 block, nb=4; a step is the 6 "
                         "solve_cg calls of one dt-substep (RK2_1, RK2_2 for x, y, z) with update_boundaries"),
            "cells": cells,
            "with_fc_points_cell_stages_per_s": cells * nst / el_fc if el_fc else None,
            "fc_overhead": el_fc / el_plain if el_fc else None,
            "fc_overhead_1_in_16_pencils": el_fc16 / el_plain if el_fc16 else None,
            "fc_kernel_gbps": (sum(prof_fc["algo_bytes"][:3]) / (sum(prof_fc["ms"][:3]) * 1e-3) / 1e9
                               if el_fc else None),
            "pcie_inclusive_cell_stages_per_s": cells * 6 * kp / el_pcie if kp else None,
            "arithmetic": "FP64, every operation rounded as the reference's code (bit-exact)",
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": None,
            "kernel": "k_sweep<dir, MHD, false, minmod> (stage mode, whole pencils), x/y/z",
            "algo_bytes_per_launch": algo / launches if launches else None,
            "avg_launch_ms": kms / launches if launches else None,
        },
        "cpu_baseline": cpu,
    }


def run_gdf(args, cpu):
    """Data files of the 256^3 MHD block: the 16 output fields of an MHD run
    derived on the device (k_datafield, one launch per field; the field then
    copied to the host) and the GDF-v2 file written from them
    (libpiernik_gdf), FP32 datasets as the reference writes by default, the
    rounding done on the device.  value: file bytes written per second.
    Algorithmic bytes of a k_datafield launch: the planes the field reads
    (8 B each) plus the FP32 field written (4 B), per interior cell."""
    import tempfile
    import piernik_amd as pk
    from piernik_amd import gdf, setups

    n = args.n or 256
    cg, cfg = setups.mhd_turbulence(n=n, seed=1)
    pk.init(cfg)
    cg.upload()
    pk.run(cg, 1)
    names = GDF_FIELDS_MHD
    for nm in names:
        pk.datafield_f32(cg, nm)
    pk.synchronize()
    print(f"[bench] gdf {n}^3 ready", file=sys.stderr, flush=True)
    k = max(1, args.steps // 4)
    # the fields as the file takes them: FP32, rounded on the device
    pk.profile_begin()
    t0 = time.perf_counter()
    for _ in range(k):
        for nm in names:
            pk.datafield_f32(cg, nm)
    el_fields = time.perf_counter() - t0
    prof = pk.profile_end()
    meta = gdf.Meta(time=0.0, n_d=(n, n, n), nfluids=1)
    place = gdf.Placement((n, n, n), dl=cg.dl)
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "gdf.h5")
        gdf.write(path, [(cg, place)], names, meta)  # warm the page cache and the writer
        t0 = time.perf_counter()
        for _ in range(k):
            gdf.write(path, [(cg, place)], names, meta)
        el_file = time.perf_counter() - t0
        size = os.path.getsize(path)
    pk.finalize()
    launches = prof["launches"][3]
    kms = prof["ms"][3]
    algo = prof["algo_bytes"][3]
    achieved = algo / (kms * 1e-3) / 1e9
    field_bytes = 4.0 * n ** 3 * len(names)
    return {
        "metric": "GDF-v2 data file of the 256^3 MHD block (16 fields, FP32) written per second",
        "value": size * k / el_file / 1e9,
        "unit": "GB/s of data file",
        "n_gpus": 1,
        "steps": k,
        "warmup": 1,
        "ms_per_step": el_file / k * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"output fields {','.join(names)} of a 3D MHD {n}^3 block (one solver step from the "
                         "Gaussian random state) derived on the device, then one GDF-v2 file per step into a "
                         "local scratch directory"),
            "file_bytes": size,
            "fields_to_host_gbps": field_bytes * k / el_fields / 1e9,
            "arithmetic": "FP64 fields, bit-exact with datafields_hdf5; FP32 datasets (h5_64bit = .false.)",
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": None,
            "kernel": "k_datafield (one launch per field, FP32 output)",
            "algo_bytes_per_launch": algo / launches if launches else None,
            "avg_launch_ms": kms / launches if launches else None,
        },
        "cpu_baseline": cpu,
    }


def run_blocks(args):
    """The 256^3 MHD domain cut into linked blocks (setups.decompose,
    pk_set_neighbour) and advanced as one list: the same cells and
    arithmetic as the single block (bit-identical results), plus one ghost
    exchange launch before every sweep.  Each block stands for one GPU's
    share of a decomposed domain."""
    import piernik_amd as pk
    from piernik_amd import setups

    n = args.n or 256
    cg, cfg = setups.mhd_turbulence(n=n, seed=1)
    cfg.fast_math = not args.exact
    cuts = [[n * k // args.blocks[d] for k in range(1, args.blocks[d])] for d in range(3)]
    blocks, nbr, boxes = setups.decompose(cg, cuts)
    del cg
    pk.init(cfg)
    pk.connect(blocks, nbr)
    for b in blocks:
        b.upload()
        b.u = b.uh = b.b = b.bh = b.psi = b.psih = None  # host copies no longer needed
    print(f"[bench] {len(blocks)} linked blocks of {blocks[0].n} uploaded", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    if not args.no_autotune:
        pk.autotune_all()
    t_tune = time.perf_counter() - t0
    tiles = {"xyz"[d]: pk.list_tiles(d, 0, blocks[0]) for d in range(3)}
    pk.run_all(max(1, args.warmup), pk.BATCH_GRAPH)
    pk.synchronize()
    t0 = time.perf_counter()
    t_sim, dt_last = pk.run_all(args.steps, pk.BATCH_GRAPH)
    pk.synchronize()
    el = time.perf_counter() - t0
    pk.profile_begin()
    pk.run_all(1, pk.BATCH_PERSISTENT)
    prof = pk.profile_end()
    kms = sum(prof["ms"])
    algo = sum(prof["algo_bytes"])
    achieved = (algo / (kms * 1e-3)) / 1e9 if kms > 0 else 0.0
    pk.finalize()
    cells = n ** 3
    return {
        "metric": METRIC + f", as {len(blocks)} linked blocks",
        "value": cells * 2 * args.steps / el,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"3D ideal MHD, HLLD+GLM, minmod, RK2, CFL 0.5, periodic {n}^3 cut into "
                         f"{args.blocks[0]}x{args.blocks[1]}x{args.blocks[2]} blocks of {blocks[0].n}, nb=4, "
                         "ghost faces exchanged before every sweep, Gaussian random v and B=curl A, beta=1"),
            "cells": cells,
            "blocks": len(blocks),
            "sim_time": t_sim,
            "last_dt": dt_last,
            "autotune_s": t_tune,
            "list_tiles_lanes_pencils_cells": tiles,
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": None,
            "kernel": ("persistent fused RK2 list sweep (HLLD+GLM), x/y/z: " +
                       ", ".join(("k_stream_list" if tiles["xyz"[d]][0] == 64 else "k_sweep_list") for d in range(3))),
            "per_dir_ms": prof["ms"],
        },
        "cpu_baseline": None,
    }


def run_batch(args, cpu):
    """The batched list: graph replay vs direct persistent launches vs one
    launch per block, same state, same arithmetic (bit-identical results)."""
    import piernik_amd as pk
    from piernik_amd import setups

    sizes = setups.batch_sizes(args.nblocks, seed=0)
    cfg = setups.batch_config()
    cfg.fast_math = not args.exact
    pk.init(cfg)
    t0 = time.perf_counter()
    cgs = setups.batch_blocks_device(sizes, cfg, seed=0)
    pk.synchronize()
    t_setup = time.perf_counter() - t0
    cells = sum(n ** 3 for n in sizes)
    print(f"[bench] {len(sizes)} blocks, {cells} cells generated on the GPU in {t_setup:.1f} s", file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    if not args.no_autotune:
        pk.autotune_all()
    t_tune = time.perf_counter() - t0
    # the list shapes chosen: (workgroup lanes, pencils, cells) per direction,
    # fluid and block edge; lanes 64 with cells -m is the stream form
    first_of = {}
    for i, nn in enumerate(sizes):
        first_of.setdefault(nn, i)
    list_shapes = {"xyz"[d]: {fl: {str(nn): list(pk.list_tiles(d, f, cgs[i])) for nn, i in sorted(first_of.items())}
                              for f, fl in enumerate(("ionized", "neutral"))} for d in range(3)}
    print(f"[bench] autotune {t_tune:.1f} s, list shapes {list_shapes}", file=sys.stderr, flush=True)
    modes = [("graph", pk.BATCH_GRAPH), ("persistent", pk.BATCH_PERSISTENT), ("per_block", pk.BATCH_PER_BLOCK)]
    for name, mode in modes:
        pk.run_all(max(1, args.warmup), mode)
    ms = {}
    t_sim = dt_last = None
    for name, mode in modes:
        pk.synchronize()
        # the graph-mode steps (the reported value) are the bench's timed range
        with (TimedRegion() if name == "graph" else contextlib.nullcontext()):
            t0 = time.perf_counter()
            t_sim, dt_last = pk.run_all(args.steps, mode)
            pk.synchronize()
            t1 = time.perf_counter()
        ms[name] = (t1 - t0) / args.steps * 1e3
        print(f"[bench] {name}: {ms[name]:.2f} ms/step", file=sys.stderr, flush=True)
    pk.profile_begin()
    pk.run_all(1, pk.BATCH_PERSISTENT)
    prof = pk.profile_end()
    launches = sum(prof["launches"])
    kms = sum(prof["ms"])
    algo = sum(prof["algo_bytes"])
    achieved = (algo / (kms * 1e-3)) / 1e9 if kms > 0 else 0.0
    value = cells * 2 / (ms["graph"] * 1e-3)
    rep_sizes = {str(n): sizes.count(n) for n in sorted(set(sizes))}
    return {
        "metric": "FP64 two-fluid cell-updates/s, batched list of 4096 uneven blocks",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms["graph"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{len(sizes)} periodic blocks, edges {rep_sizes} (seed 0), nb=4, ionized MHD (HLLD+GLM) "
                         "+ neutral HD (HLLC) per block, van Leer, RK2, CFL 0.7, rho, p = 1 +- 0.2, |v| <= 0.5"),
            "cells": cells,
            "nvar": 14,
            "step": "time_step + fluid_update_full over the list (6 sweeps x 2 fluids)",
            "sim_time": t_sim,
            "last_dt": dt_last,
            "parallelism": "1 GPU",
            "arithmetic": ("FP64 with FMA contraction and fast division/sqrt" if not args.exact else
                           "FP64, every operation rounded as the reference's code (bit-exact)"),
            "setup_s": t_setup,
            "autotune_s": t_tune,
            "list_shapes_lanes_pencils_cells": list_shapes,
        },
        "batching": {
            "ms_per_step": ms,
            "speedup_graph_vs_per_block": ms["per_block"] / ms["graph"],
            "speedup_persistent_vs_per_block": ms["per_block"] / ms["persistent"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": None,
            "kernel": ("persistent fused RK2 list sweeps (MHD and HD; k_stream_list for the block shapes where "
                       "the stream form is faster, k_sweep_list tiles for the rest), x/y/z"),
            "algo_bytes_per_launch": algo / launches if launches else None,
            "avg_launch_ms": kms / launches if launches else None,
            "per_dir_gbps": [(prof["algo_bytes"][d] / (prof["ms"][d] * 1e-3) / 1e9) if prof["ms"][d] > 0 else None
                             for d in range(3)],
        },
        "cpu_baseline": cpu,
    }


if __name__ == "__main__":
    main()
