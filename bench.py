"""Benchmark: the vesicle-step pipeline at 256^3, FP32 bulk beside plain FP64,
on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--main-only]

Workload (one "step", config 3 of the metric): WENO5 + TVD-RK3 advection of F
and psi, setDistance, normal velocity and its 40-iteration WENO5-RK3
extension, every 5th step a 100-iteration WENO5-RK3 reinitialization plus a
(10,5) closest-point surface redistance, and the cosine delta/H area and
volume; 256^3 on [-1.2,1.2]^3, ellipsoid (0.8,0.5,0.5), seeded 32-mode curl
velocity, CFL 0.4 (2018_7_vesicle_amd/pipeline.py).  The line times it with
FP32 bulk fields (value) and with FP64 fields (fp64.value), each after its
own set-up and warm-up steps.  Inputs are resident in HBM before timing.

value = WENO grid-point updates per second: (operator-stage passes counted by
pipeline.updates_per_step) x points / wall time of the K timed steps.

roofline: the dominant kernel of the FP32 pipeline is the fused WENO5-RK3
extension stage (k_extend_stage<R, float>: interface-list blocks and
z-marching bulk blocks in one launch).  Its average duration is measured with
HIP events around the extension calls inside the timed steps (on the
pipeline's stream, torch's current stream), minus the extension's per-call
set-up (timed separately with 0 iterations), divided by the stage launches.
achieved = 22.7 B/point (the algorithmic bytes of one FP32-bulk extension
stage) x points / that duration.  traffic: HBM bytes per launch from
rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of `bench.py --probe stages`,
calibrated on a 256 MiB copy (profiles/stage_counters.json);
roofline.valu: that launch's VALU instructions (SQ_INSTS_VALU of the same
runs) per second against the issue rate this run measures
(vls_probe_valu_rate).

cpu_baseline: the CPU oracle (OpenMP, FP64) on a bounded sample of the same
workload on this host: one step without reinitialization at 256^3.

Beside the main line: config2_128 (the FP64 pipeline at 128^3, the config-2
workload), config1 (64^3 ENO2 forward-Euler reinitialization with its error
measures), config3 (the 256^3 pipeline for 200 steps, FP32 bulk vs FP64, with
the FP32 run's errors), config4 (512 vesicles of uneven sizes batched, FP32
bulk vs FP64 batched vs the plain per-vesicle FP64 path), config5 (the 96^3
adjoint gradient against finite differences and tangent-mode derivatives),
variants (the WENO-RK3 / ENO-RK2 extensions and WENO-RK3 reinitialization at
96^3), toolbox (the calculus toolboxes of F and of an auxiliary field A at
128^3), relax (the single-phase vesicle relaxation step at 96^3).
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

B_ALG_EXTEND_FP64 = 37.3  # bytes per point per FP64 WENO5 extension stage
B_ALG_EXTEND_FP32 = 22.7  # the same with float C and velocity, FP64 phi
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5,
                    help="untimed steps first (5 covers one reinitialization step, so every graph is captured)")
    ap.add_argument("--n", type=int, default=256, help="grid edge of the main line")
    ap.add_argument("--modes", type=int, default=32, help="velocity modes of the main line")
    ap.add_argument("--pmc-summary", nargs="+", default=None, metavar="DB",
                    help="print per-kernel averages of the counters in rocprofv3 --pmc result databases "
                         "(the last three dispatches of each kernel) and exit")
    ap.add_argument("--trace-csv", default=None, metavar="CSV",
                    help="with --pmc-summary: kernel trace (rocprofv3 --kernel-trace --output-format csv) of a "
                         "run of the same probe, for launch times and VALU/HBM fractions")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--precision-steps", type=int, default=200,
                    help="steps of the 256^3 FP32-bulk vs FP64 comparison (0 skips it)")
    ap.add_argument("--batch-steps", type=int, default=100,
                    help="timed steps of the 512-vesicle batch comparison (the workload's 100; 0 skips it)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the 96^3 adjoint gradient check")
    ap.add_argument("--probe", choices=["stages", "valu-rate", "batch", "variants", "toolbox", "relax"], default=None,
                    help="valu-rate: the VALU issue rate (FP32, FP64) of this GPU and the HBM rate of a five-read, "
                         "one-write float stream; "
                         "stages: 3 launches each of the extension, reinitialization and advection stage "
                         "kernels at 128^3 FP64 and 256^3 FP32 (for rocprofv3 --pmc); "
                         "batch: a few steps of the 512-vesicle FP32 batch (for rocprofv3 --kernel-trace); "
                         "variants: the WENO-RK3 / ENO-RK2 calls of the variants line (for rocprofv3 --kernel-trace); "
                         "toolbox: the calculus-toolbox calls of the toolbox line (for rocprofv3 --kernel-trace); "
                         "relax: the vesicle relaxation steps of the relax line")
    ap.add_argument("--main-only", action="store_true",
                    help="only the main line: no config1/3/4/5 and no CPU baseline")
    return ap.parse_args()


def probe_stages(args):
    """A short run for hardware counters of the three fused stage kernels: the
    config-3 pipeline at 256^3 (FP32 bulk) and the config-2 pipeline at 128^3
    (FP64) are both set up first (initial reinitialization, a fixed launch
    shape: k_*_stage<32, float> at 256^3 with 16 pages per block, <32, double>
    at 128^3 with 8), then each runs one extension iteration, one
    reinitialization iteration and one advection step (3 launches of each
    stage kernel), each twice, with a synchronization between the phases; so
    the last three launches of every stage kernel are the probe's own, on warm
    buffers.  Last, a 256 MiB float64 copy whose source was pushed out of the
    caches by a 2 GiB fill: FETCH_SIZE and WRITE_SIZE of a known number of
    bytes, to calibrate them."""
    import importlib

    import torch

    vls = importlib.import_module("2018_7_vesicle_amd")
    pl = vls.pipeline
    out = {"probe": "stages"}
    lines = []
    # a fixed launch shape (the one the autotuner picks most often for these
    # grids), so that every profiled run launches the same kernels
    for n, prec, modes, tr, zd in ((256, "fp32", 32, 32, 16), (128, "fp64", 16, 32, 8)):
        p = pl.Pipeline(pl.PipelineConfig(n=n, precision=prec, modes=modes, initial_reinit=20, autotune=False))
        p.map.setTileRows(tr)
        p.map.setZmarchDepth(zd)
        p.map.setDistance()
        p.map.normalVelocity(*p.u, p.Vn)
        lines.append((f"{prec}_{n}", p))
    torch.cuda.synchronize()
    for key, p in lines:
        m = p.map
        # each phase twice: the first launches after set-up run on cold
        # buffers (the first extension stage took twice its time), so the
        # profiled launches are those of the second pass
        for _ in range(2):
            m.extend_(p.Vn, 1)
            torch.cuda.synchronize()
        for _ in range(2):
            m.reinitialize_(m.F, 1)
            torch.cuda.synchronize()
        for _ in range(2):
            m.advect_(m.F, *p.u, p.dt, 1)
            torch.cuda.synchronize()
        tr, zd = m.launch()
        out[key] = {"tile_rows": tr, "zmarch_depth": zd, "points": p.grid.NumElt}
    del lines, p, m
    torch.cuda.empty_cache()
    src = torch.ones(1 << 25, dtype=torch.float64, device="cuda")
    flush = torch.empty(1 << 28, dtype=torch.float64, device="cuda")
    flush.fill_(2.0)
    torch.cuda.synchronize()
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    out["copy_bytes"] = 8 << 25
    print(json.dumps(out), flush=True)
    return 0


def probe_batch(args):
    """A short batch run for kernel traces: the 512-vesicle FP32 workload
    (autotuned), both step graphs captured, then 3 steps."""
    import importlib

    import torch

    bm = importlib.import_module("2018_7_vesicle_amd.batch")
    p = bm.BatchPipeline(bm.BatchConfig())
    p.warm()
    upd = p.run(3)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "batch", "points": p.batch.total, "updates": upd, "tile_rows": p.batch.tile_rows,
                      "blocks": p.batch.blocks}), flush=True)
    return 0


def stage_counters(key):
    """Counters of one stage launch recorded with rocprofv3 --pmc runs of
    `bench.py --probe stages` (profiles/stage_counters.json): HBM bytes per
    launch (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration)
    and VALU wave-instructions per launch, if recorded."""
    path = os.path.join(ROOT, "profiles", "stage_counters.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path))["stages"].get(key)
    except Exception:
        return None


def valu_rate(vls, fp64):
    """VALU wave-instructions per second this GPU issues (vls_probe_valu_rate)."""
    import ctypes

    v = ctypes.c_double()
    vls._lib.call("vls_probe_valu_rate", 1 if fp64 else 0, ctypes.byref(v), None)
    return v.value


# algorithmic bytes per point of one stage launch (reads and writes of the
# operator's fields), FP64 and FP32 bulk
B_ALG_STAGE = {("extend", "fp64"): B_ALG_EXTEND_FP64, ("extend", "fp32"): B_ALG_EXTEND_FP32,
               ("reinit", "fp64"): 29.3, ("reinit", "fp32"): 18.7,
               ("advect", "fp64"): 45.3, ("advect", "fp32"): 22.7}
CLOCK_HZ = 2.4e9  # MI355X peak engine clock
CUS = 256


def trace_durations(path):
    """kernel name -> durations (us) of its dispatches in dispatch order, from a
    rocprofv3 --kernel-trace --output-format csv kernel trace"""
    import collections
    import csv

    rows = list(csv.DictReader(open(path)))
    out = collections.defaultdict(list)
    for r in sorted(rows, key=lambda r: int(r["Dispatch_Id"])):
        name = r["Kernel_Name"].split("(")[0].replace("void ", "")
        out[name].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    return out


def stream_rate(vls, points):
    """HBM bytes per second of a five-read, one-write float stream (vls_probe_stream_rate)."""
    import ctypes

    v = ctypes.c_double()
    vls._lib.call("vls_probe_stream_rate", ctypes.c_int64(points), ctypes.byref(v), None)
    return v.value


def pmc_summary(paths, trace=None):
    """Per kernel and counter, the average over the last three dispatches of
    that kernel in rocprofv3 --pmc result databases of `bench.py --probe
    stages` (the probe's own launches come after set-up).  Prints one JSON
    object: the raw table, and per stage kernel (extension, reinitialization,
    advection at 128^3 FP64 and 256^3 FP32) its VALU, SALU, LDS and memory
    instructions per launch and its HBM bytes per launch, FETCH_SIZE and
    WRITE_SIZE (KiB) scaled by what they read for the probe's 256 MiB copy
    (the run's last copy kernel).  With `trace` (the kernel trace of a run of
    the same probe without counters) each stage also gets its launch time
    (last three dispatches), the VALU issue fraction (VALU wave-instructions
    over one per CU per clock at 2.4 GHz) and the HBM fractions of its measured
    traffic and of its algorithmic bytes (8 TB/s)."""
    import collections
    import sqlite3

    per = collections.defaultdict(lambda: collections.defaultdict(lambda: collections.defaultdict(float)))
    for path in paths:
        con = sqlite3.connect(path)
        rows = con.execute("select kernel_name, counter_name, value, dispatch_id from counters_collection").fetchall()
        for name, counter, value, dispatch in rows:
            per[name.split("(")[0].replace("void ", "")][counter][(path, dispatch)] += float(value)
    table = {}
    for kname, counters in per.items():
        table[kname] = {}
        for counter, by_dispatch in counters.items():
            keys = sorted(by_dispatch, key=lambda k: k[1])
            last = [by_dispatch[k] for k in keys[-3:]]
            table[kname][counter] = {"dispatches": len(keys), "last3_avg": sum(last) / len(last),
                                     "max": max(by_dispatch.values()), "last_dispatch": keys[-1][1]}
    copy_bytes = 256 * 1024.0  # KiB read and written by the calibration copy, the run's last copy kernel
    calib = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        last = None
        for kname, counters in per.items():
            if "copy" in kname and c in counters:
                for key, value in counters[c].items():
                    if last is None or key[1] > last[0]:
                        last = (key[1], value)
        calib[c] = copy_bytes / last[1] if last and last[1] > 0 else None
    stages = {}
    for kname, v in table.items():
        for stage in ("extend", "reinit", "advect"):
            if f"k_{stage}_stage<" not in kname:
                continue
            prec = "fp32" if "float" in kname else "fp64"
            key = f"{stage}_{prec}_{256 if prec == 'fp32' else 128}"
            last = max(c["last_dispatch"] for c in v.values())
            if key in stages and stages[key]["last_dispatch"] > last:
                continue  # another launch shape of this kernel ran later (the probe's own)
            rec = {"kernel": kname, "last_dispatch": last}
            for c, name in (("SQ_INSTS_VALU", "valu_insts"), ("SQ_INSTS_SALU", "salu_insts"),
                            ("SQ_INSTS_LDS", "lds_insts"), ("SQ_INSTS_VMEM_RD", "vmem_rd_insts"),
                            ("SQ_INSTS_VMEM_WR", "vmem_wr_insts"), ("SQ_WAVES", "waves"),
                            ("FETCH_SIZE", "fetch_kib"), ("WRITE_SIZE", "write_kib")):
                if c in v:
                    rec[name] = v[c]["last3_avg"]
            if "fetch_kib" in rec and "write_kib" in rec and calib["FETCH_SIZE"] and calib["WRITE_SIZE"]:
                rec["traffic_bytes"] = 1024.0 * (rec["fetch_kib"] * calib["FETCH_SIZE"] +
                                                 rec["write_kib"] * calib["WRITE_SIZE"])
            stages[key] = rec
    if trace:
        durations = trace_durations(trace)
        for key, rec in stages.items():
            stage, prec, n = key.split("_")
            d = durations.get(rec["kernel"])
            if not d:
                continue
            us = sum(d[-3:]) / len(d[-3:])
            pts = int(n) ** 3
            rec["avg_us"] = us
            rec["points"] = pts
            rec["alg_bytes_per_point"] = B_ALG_STAGE[(stage, prec)]
            rec["alg_hbm_frac"] = rec["alg_bytes_per_point"] * pts / (us * 1e-6) / (HBM_PEAK_GBS * 1e9)
            if "valu_insts" in rec:
                rec["valu_per_point"] = rec["valu_insts"] * 64 / pts
                rec["valu_frac"] = rec["valu_insts"] / (CUS * CLOCK_HZ * us * 1e-6)
            if "salu_insts" in rec:
                rec["salu_frac"] = rec["salu_insts"] / (CUS * CLOCK_HZ * us * 1e-6)
            if "traffic_bytes" in rec:
                rec["traffic_bytes_per_point"] = rec["traffic_bytes"] / pts
                rec["traffic_hbm_frac"] = rec["traffic_bytes"] / (us * 1e-6) / (HBM_PEAK_GBS * 1e9)
    print(json.dumps({"dbs": paths, "trace": trace, "calibration": calib, "stages": stages, "kernels": table}),
          flush=True)
    return 0


def pipeline_line(vls, n, modes, precision, steps, warmup):
    """The vesicle-step pipeline at n^3 in one precision: set-up (initial
    reinitialization, launch-shape autotuning), `warmup` untimed steps (every
    graph captured), then `steps` timed steps with the extension calls
    bracketed by HIP events; returns the rate and the extension stage's
    average launch time."""
    import torch

    pl = vls.pipeline
    cfg = pl.PipelineConfig(n=n, modes=modes, precision=precision)
    p = pl.Pipeline(cfg)
    if warmup:
        p.run(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    updates = p.run(steps, record_extension=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ext_ms = [a.elapsed_time(b) for a, b in p.ext_events]
    # the extension call's own set-up (distances, extension velocity), timed
    # as the call with no iterations; and, as a cross-check, the stage time
    # from the difference of 12- and 2-iteration calls.  Medians: a call that
    # captures a graph takes milliseconds more than a replay.
    scratch = p.Vn.clone()
    call_ms = {}
    for k in (0, 2, 12):
        p.map.extend_(scratch, k)  # capture
        ts = []
        for _ in range(5):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            p.map.extend_(scratch, k)
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        call_ms[k] = statistics.median(ts)
    stage_us = (statistics.median(ext_ms) - call_ms[0]) / (3 * cfg.ext_iters) * 1e3
    stage_us_diff = (call_ms[12] - call_ms[2]) / 30 * 1e3
    av = p.area_volume.cpu().numpy()
    n_pts = p.grid.NumElt
    out = {"cfg": cfg, "points": n_pts, "value": updates / wall, "ms_per_step": wall / steps * 1e3,
           "updates_per_step_avg": updates / steps / n_pts, "stage_us": stage_us,
           "stage_us_by_difference": stage_us_diff, "extension_call_ms": statistics.median(ext_ms),
           "extension_setup_ms": call_ms[0],
           "area": [float(av[0, 0]), float(av[-1, 0])], "volume": [float(av[0, 1]), float(av[-1, 1])],
           "dt": p.dt, "launch": p.map.launch()}
    del p
    torch.cuda.empty_cache()
    return out


def config2_line(vls, steps, warmup):
    """Config 2 beside the main line: the FP64 pipeline at 128^3 (16-mode
    velocity), with its extension stage against the FP64 algorithmic bytes."""
    r = pipeline_line(vls, 128, 16, "fp64", steps, warmup)
    achieved = B_ALG_EXTEND_FP64 * r["points"] / (r["stage_us"] * 1e-6) / 1e9
    counters = stage_counters("extend_fp64_128")
    return {"value": r["value"], "unit": "grid-pt updates/s", "ms_per_step": r["ms_per_step"], "dtype": "f64",
            "workload": "config-2 vesicle step, 128^3, 16-mode velocity, FP64",
            "roofline": {"kernel": "FP64 WENO5-RK3 extension stage (k_extend_stage<R, double>)", "achieved": achieved,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "avg_launch_us": r["stage_us"], "bytes_per_point": B_ALG_EXTEND_FP64,
                         "traffic": counters.get("traffic_bytes") if counters else None}}


def cpu_baseline(cfg):
    """The CPU oracle (OpenMP, FP64, all host threads) on a bounded sample of
    the main line's workload: one step without reinitialization at its grid
    (advection of F and psi, setDistance, normal velocity, the 40-iteration
    extension, area/volume), on the same inputs, timed as the GPU's line (the
    rate per stage pass does not depend on which passes a step holds)."""
    import importlib

    import numpy as np

    from oracle import pyoracle as po

    pl = importlib.import_module("2018_7_vesicle_amd.pipeline")
    xv, F0, psi0, u = pl.make_inputs(cfg)
    dx = float(xv[1] - xv[0])
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    po.set_num_threads(threads)
    umax = max(float(np.abs(c).max()) for c in u)
    dt = cfg.cfl * dx / umax
    step_index = 0  # not a reinitializing step
    t0 = time.perf_counter()
    po.pipeline(F0, psi0, u, dt, cfg.eps_factor * dx, (dx, dx, dx), 1, cfg.ext_iters, cfg.reinit_every,
                cfg.reinit_iters, cfg.redist_iters, 0, first_step=step_index)
    t1 = time.perf_counter()
    passes = pl.updates_per_step(cfg, step_index)
    return {
        "value": passes * F0.size / (t1 - t0),
        "unit": "grid-pt updates/s",
        "cores": po.num_threads(),
        "kind": "port",
        "sample": f"one pipeline step without reinitialization at {cfg.n}^3 FP64 ({cfg.modes}-mode velocity): "
                  f"{passes} WENO stage passes in {t1 - t0:.1f} s",
    }


def config1(vls, with_cpu):
    """Config 1 beside the main line: 64^3 ENO2 forward-Euler reinitialization,
    CFL 0.5, 100 iterations, on the GPU (and the CPU oracle), with its error
    measures."""
    import numpy as np
    import torch

    diag = __import__("importlib").import_module("2018_7_vesicle_amd.diagnostics")
    xv = np.linspace(-1.0, 1.0, 64)
    Z, X, Y = np.meshgrid(xv, xv, xv, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    theta = np.arctan2(np.sqrt(X ** 2 + Y ** 2), Z)
    varphi = np.arctan2(Y, X)
    F0 = (r - 0.6) * (1 + 0.3 * np.sin(3 * theta) * np.cos(2 * varphi)) * np.exp(0.5 * X)
    G = vls.GD3(xv, xv, xv)
    F0t = torch.from_numpy(np.ascontiguousarray(F0)).cuda()
    m = vls.SDF3(G, F0t)
    m.ENORK2Reinitialization(F0t, 100, rk2=False, cfl=0.5)  # warm-up
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    F = m.ENORK2Reinitialization(F0t, 100, rk2=False, cfl=0.5)
    e1.record()
    torch.cuda.synchronize()
    gpu_ms = e0.elapsed_time(e1)
    e_max, e_mean, _ = diag.grad_error(G, F)
    disp, ncross = diag.sphere_displacement(G, F, 0.6)
    out = {"grid": 64, "iterations": 100, "gpu_ms": gpu_ms,
           "gpu_updates_per_s": 100 * 64 ** 3 / (gpu_ms * 1e-3),
           "grad_err_max": e_max, "grad_err_mean": e_mean, "displacement_dx": disp / G.Dx, "crossings": ncross}
    if with_cpu:
        from oracle import pyoracle as po

        t0 = time.perf_counter()
        ref = po.eno2_reinit(F0, 100, G.Dx, G.Dy, G.Dz, rk2=False, cfl=0.5)
        out["cpu_ms"] = (time.perf_counter() - t0) * 1e3
        out["gpu_equals_cpu_oracle"] = bool(np.array_equal(ref, F.cpu().numpy()))
    return out


def config3(steps):
    """Config 3 beside the main line: the same vesicle-step pipeline at 256^3
    with a 32-mode velocity, FP32 bulk vs FP64 from one seed and initial
    state, `steps` steps each; the FP32 run's errors against the FP64 run
    (which reproduces the CPU oracle bit for bit) and the speedup."""
    import importlib

    diag = importlib.import_module("2018_7_vesicle_amd.diagnostics")
    r = diag.precision_comparison(256, 32, steps, conditioning=True)
    r["workload"] = (f"256^3 vesicle step, 32-mode curl velocity (seed 2018), {steps} steps, FP32 bulk vs FP64; "
                     "fp64_rounded_start: FP64 from the FP32-rounded initial state (conditioning floor)")
    return r


def config4(steps):
    """Config 4 beside the main line: 512 vesicles on cubes of edge
    {32,48,64,80,96} (seeded), one 8-mode curl velocity each, advanced
    together; FP32 bulk and FP64 through the batched persistent kernels
    (graph-captured steps, autotuned tiles), `steps` timed steps each (by
    default the workload's 100), and the plain FP64 path (one object per
    vesicle) on every 16th vesicle for the same steps; the FP32 errors against
    FP64 over all 512 vesicles, beside those of FP64 started from the
    FP32-rounded initial state (the problem's own sensitivity to one rounding
    of the data: over 100 steps the vesicles double their area and move a
    fraction of a cell under it)."""
    import importlib

    bm = importlib.import_module("2018_7_vesicle_amd.batch")
    r = bm.batch_comparison(bm.BatchConfig(), steps=steps, sample_every=16, log=progress)
    r["workload"] = (f"512 vesicles, edges {{32,48,64,80,96}} on [-1,1]^3, ellipsoid semi-axes U[0.3,0.8], "
                     f"8-mode curl velocity (seed 2018+v), WENO5-RK3 advection CFL 0.4, 40-iteration extension, "
                     f"reinit 100 iterations every 5 steps; {steps} steps timed")
    return r


def variants(with_cpu, iterations=100):
    """Beside the main line: the WENO-RK3 and ENO-RK2 extensions and the
    WENO-RK3 reinitialization that the vesicle drivers call every step
    (src/7_0_paper/2_0_Vesicle/p00_test_singlephase_relax.m:76,84,115) on the drivers'
    96^3 grid: an ellipsoid level set on [-1,1]^3, a smooth field extended off
    it; ms per call (100 iterations) and stage updates per second, on the GPU
    and (10 iterations, scaled) on the CPU oracle."""
    import numpy as np
    import torch

    vls = __import__("importlib").import_module("2018_7_vesicle_amd")
    n = 96
    xv = np.linspace(-1.0, 1.0, n)
    Z, X, Y = np.meshgrid(xv, xv, xv, indexing="ij")
    F0 = np.ascontiguousarray(np.sqrt(X ** 2 / 0.7 ** 2 + Y ** 2 / 0.55 ** 2 + Z ** 2 / 0.45 ** 2) - 1.0)
    C0 = np.ascontiguousarray(0.5 * X + Y * Z + 0.2 * np.sin(3.0 * Y))
    G = vls.GD3(xv, xv, xv)
    F = torch.from_numpy(F0).cuda()
    C = torch.from_numpy(C0).cuda()
    m = vls.SDF3(G, F)
    calls = (("WENORK3Extend", 3, lambda k: m.WENORK3Extend(C, k), "wenork3"),
             ("ENORK2Extend", 2, lambda k: m.ENORK2Extend(C, k), "enork2"),
             ("WENORK3Reinitialization", 3, lambda k: m.WENORK3Reinitialization(F, k), None))
    out = {"grid": n, "iterations": iterations}
    for name, per_it, fn, scheme in calls:
        fn(iterations)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(3):
            fn(iterations)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 3
        rec = {"gpu_ms": ms, "gpu_updates_per_s": per_it * iterations * n ** 3 / (ms * 1e-3)}
        if with_cpu:
            from oracle import pyoracle as po

            k = 10
            t0 = time.perf_counter()
            if scheme is None:
                po.wenork3_reinit(F0, k, G.Dx, G.Dy, G.Dz)
            else:
                po.variant_extend(C0, k, F0, G.Dx, G.Dy, G.Dz, scheme)
            cpu_s = time.perf_counter() - t0
            rec["cpu_updates_per_s"] = per_it * k * n ** 3 / cpu_s
            rec["cpu_threads"] = po.num_threads()
        out[name] = rec
    out["batched"] = variants_batched(iterations)
    return out


def variants_batched(iterations, vesicles=8):
    """The same three calls on `vesicles` vesicles of 96^3 (ellipsoids with
    seeded semi-axes in [0.35, 0.75]) at once through the batch entry points,
    against one SDF3 object per vesicle called in turn: ms per call for all
    vesicles, and the speedup."""
    import numpy as np
    import torch

    vls = __import__("importlib").import_module("2018_7_vesicle_amd")
    bm = __import__("importlib").import_module("2018_7_vesicle_amd.batch")
    n = 96
    xv = np.linspace(-1.0, 1.0, n)
    Z, X, Y = np.meshgrid(xv, xv, xv, indexing="ij")
    rng = np.random.default_rng(2018)
    axes = rng.uniform(0.35, 0.75, size=(vesicles, 3))
    Fs = [torch.from_numpy(np.ascontiguousarray(np.sqrt(X ** 2 / a ** 2 + Y ** 2 / b ** 2 + Z ** 2 / c ** 2) - 1.0))
          .cuda() for a, b, c in axes]
    C0 = torch.from_numpy(np.ascontiguousarray(0.5 * X + Y * Z + 0.2 * np.sin(3.0 * Y))).cuda()
    G = vls.GD3(xv, xv, xv)
    objs = [vls.SDF3(G, F) for F in Fs]
    b = bm.Batch([(n, n, n)] * vesicles, [(G.Dx, G.Dy, G.Dz)] * vesicles, dtype=torch.float64)
    F = b.pack(Fs)
    C = b.pack([C0] * vesicles)
    Fw = F.clone()
    calls = (("WENORK3Extend", 3, lambda k: b.wenork3_extend(F, C, k), lambda m, k: m.WENORK3Extend(C0, k)),
             ("ENORK2Extend", 2, lambda k: b.enork2_extend(F, C, k), lambda m, k: m.ENORK2Extend(C0, k)),
             ("WENORK3Reinitialization", 3, lambda k: b.wenork3_reinitialize(Fw, k),
              lambda m, k: m.WENORK3Reinitialization(m.F, k)))

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(3):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 3

    out = {"vesicles": vesicles, "grid": n, "iterations": iterations}
    for name, per_it, batched, single in calls:
        Fw.copy_(F)
        ms_b = timed(lambda: batched(iterations))
        ms_s = timed(lambda: [single(m, iterations) for m in objs])
        out[name] = {"batched_ms": ms_b, "per_object_ms": ms_s, "speedup": ms_s / ms_b,
                     "batched_updates_per_s": per_it * iterations * vesicles * n ** 3 / (ms_b * 1e-3)}
    return out


B_ALG_TOOLBOX = 160.0  # set_calculus_toolbox: F and HPrimal in, 18 fields out (FP64)
B_ALG_AUX = 440.0  # GPUAsetCalculusToolBox: max(A,0) 16 + kernel (15 in, 27 out) 336 + tail (9 in, 2 out) 88


def toolbox(with_cpu, n=128, reps=20):
    """Beside the main line: the calculus toolboxes of a surface F and of an
    auxiliary field A on it (GPUsetCalculusToolBox.m, GPUAsetCalculusToolBox.m)
    at 128^3, on caller buffers (no allocation in the timed region): ms per call
    and the algorithmic bytes per call over that time against HBM peak; the CPU
    oracle on the same grid."""
    import numpy as np
    import torch

    vls = __import__("importlib").import_module("2018_7_vesicle_amd")
    xv = np.linspace(-1.0, 1.0, n)
    Z, X, Y = np.meshgrid(xv, xv, xv, indexing="ij")
    F0 = np.ascontiguousarray(np.sqrt(X ** 2 / 0.7 ** 2 + Y ** 2 / 0.55 ** 2 + Z ** 2 / 0.45 ** 2) - 1.0)
    A0 = np.ascontiguousarray(0.6 * Z + 0.25 * X * Y - 0.15 * Y + 0.05)
    dx = float(xv[1] - xv[0])
    F = torch.from_numpy(F0).cuda()
    A = torch.from_numpy(A0).cuda()
    hp = torch.clamp(F, min=0.0).contiguous()
    ft = [torch.empty_like(F) for _ in range(18)]
    aux = [torch.empty_like(F) for _ in range(30)]
    fin = ft[:13] + [ft[16]]
    args = (n, n, n, dx, dx, dx, dx)
    calls = (("set_calculus_toolbox", B_ALG_TOOLBOX, lambda: vls.feval_set_calculus_toolbox(ft, F, hp, *args)),
             ("aux_calculus_toolbox", B_ALG_AUX, lambda: vls.feval_aux_calculus_toolbox(aux, A, fin, *args)))
    out = {"grid": n, "reps": reps}
    for name, b, fn in calls:
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        gbs = b * n ** 3 / (ms * 1e-3) / 1e9
        out[name] = {"gpu_ms": ms, "bytes_per_point": b, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS}
    if with_cpu:
        from oracle import pyoracle as po

        t0 = time.perf_counter()
        ftc = po.set_calculus_toolbox(F0, np.maximum(F0, 0), dx, dx, dx, dx)
        t1 = time.perf_counter()
        po.aux_calculus_toolbox(A0, ftc, dx, dx, dx, dx)
        t2 = time.perf_counter()
        out["set_calculus_toolbox"]["cpu_ms"] = (t1 - t0) * 1e3
        out["aux_calculus_toolbox"]["cpu_ms"] = (t2 - t1) * 1e3
        out["cpu_threads"] = po.num_threads()
    return out


def relax(with_cpu, steps=5):
    """Beside the main line: the single-phase vesicle relaxation step of
    p00_test_singlephase_relax.m at its 96^3 setting (oblate ellipsoid of
    reduced volume 0.6, CFL 0.5, 100 extension and reinitialization
    iterations, reinitialization every step): toolbox, WENORK3Extend,
    bending speed, ENORK2Extend, tension/pressure, smoothGMRES (rocFFT),
    ENORK2Extend, update, setDistance, WENO5RK3Reinitialization.  ms per step
    on the GPU (one warm-up step, then `steps`); the same for eight vesicles
    of reduced volumes 0.60 .. 0.88 stepped together (BatchRelax); and one
    step of the CPU oracle's composition of the same operators."""
    import numpy as np
    import torch

    vls = __import__("importlib").import_module("2018_7_vesicle_amd")
    n = 96
    xv, yv, zv, F0 = vls.relax.oblate_ellipsoid(n)
    G = vls.GD3(xv, yv, zv)
    m0 = vls.SDF3(G, torch.from_numpy(F0).cuda())
    m0.setDistance()
    F = m0.WENO5RK3Reinitialization(m0.F, 500)
    cfg = vls.relax.RelaxConfig()
    r = vls.relax.VesicleRelax(G, F, cfg)
    r.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r.step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    info = dict(zip(vls.TENSION_PRESSURE_FIELDS, r.info.cpu().tolist()))
    # pseudo-time stage updates per step: WENO-RK3 (3 stages) and two ENO-RK2
    # (2 stages) extensions and the WENO5-RK3 reinitialization (3 stages)
    upd = (3 + 2 + 2 + 3) * cfg.ext_iters * n ** 3
    out = {"grid": n, "steps": steps, "gpu_ms_per_step": ms, "gpu_updates_per_s": upd / (ms * 1e-3),
           "area_drift": info["area"] / r.InitialArea - 1, "volume_drift": info["volume"] / r.InitialVolume - 1,
           "Tension": info["Tension"], "Pressure": info["Pressure"], "Dt": info["Dt"]}
    a0, v0 = r.InitialArea, r.InitialVolume
    del r
    # eight vesicles of reduced volumes 0.60 .. 0.88 on 96^3 grids stepped
    # together (BatchRelax: batched extensions, setDistance and reinitialization)
    nb = 8
    Fs = []
    for k in range(nb):
        mk = vls.SDF3(G, torch.from_numpy(vls.relax.oblate_ellipsoid(n, reduced_volume=0.6 + 0.04 * k)[3]).cuda())
        mk.setDistance()
        Fs.append(mk.WENO5RK3Reinitialization(mk.F, 500))
        del mk
    br = vls.relax.BatchRelax([G] * nb, Fs, cfg)
    del Fs
    br.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        br.step()
    torch.cuda.synchronize()
    bms = (time.perf_counter() - t0) * 1e3 / steps
    out["batch"] = {"vesicles": nb, "grid": n, "gpu_ms_per_step": bms, "gpu_updates_per_s": nb * upd / (bms * 1e-3),
                    "speedup_vs_one_at_a_time": nb * ms / bms}
    del br
    if with_cpu:
        from oracle import pyoracle as po
        from oracle import relax_oracle as ro

        Fc = F.cpu().numpy()
        t0 = time.perf_counter()
        ro.relax_step(Fc, a0, v0, G.Dx, G.Dy, G.Dz, cfl=cfg.cfl)
        cpu_s = time.perf_counter() - t0
        out["cpu_ms_per_step"] = cpu_s * 1e3
        out["cpu_threads"] = po.num_threads()
    return out


def config5():
    """Config 5 beside the main line: the 96^3 shape-matching problem (200
    WENO5-RK3 advection steps, reinitialization every 10, velocity from 3x64
    Fourier coefficients, cosine-Heaviside mismatch to an ellipsoid); the
    adjoint gradient (checkpoint every 20 steps) against FP64 central
    differences on all 192 coefficients (h = 1e-6, and h = 1e-8, a step that
    stays clear of the scheme's branch thresholds) and tangent-mode
    directional derivatives on 8 directions."""
    import importlib

    adj = importlib.import_module("2018_7_vesicle_amd.adjoint")
    cfg = adj.AdjointConfig()
    r = adj.check_gradient(adj.ShapeMatching(cfg), hs=(1e-6, 1e-8), directions=8)
    r["workload"] = (f"{cfg.n}^3 on [-1,1]^3, {cfg.steps} steps, reinit ({cfg.reinit_iters} iterations) every "
                     f"{cfg.reinit_every}, 3x{cfg.modes} coefficients ~ N(0,{cfg.coef_std}^2), checkpoint every "
                     f"{cfg.ckpt_every} steps, FP64")
    return r


def progress(msg):
    """One line on stderr as each part of the run ends (the JSON line alone
    goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.probe == "valu-rate":
        import importlib

        vls = importlib.import_module("2018_7_vesicle_amd")
        print(json.dumps({"probe": "valu-rate", "fp32_insts_per_s": valu_rate(vls, False),
                          "fp64_insts_per_s": valu_rate(vls, True),
                          "stream5_bytes_per_s": {str(n): stream_rate(vls, n) for n in (1 << 22, 1 << 24, 1 << 26)}}),
              flush=True)
        return 0
    if args.probe == "batch":
        return probe_batch(args)
    if args.probe == "stages":
        return probe_stages(args)
    if args.probe == "variants":
        print(json.dumps({"probe": "variants", **variants(False, iterations=20)}), flush=True)
        return 0
    if args.probe == "relax":
        print(json.dumps({"probe": "relax", **relax(not args.no_cpu, steps=3)}), flush=True)
        return 0
    if args.probe == "toolbox":
        print(json.dumps({"probe": "toolbox", **toolbox(False, reps=5)}), flush=True)
        # a plain 256 MiB float64 copy last: calibrates FETCH_SIZE / WRITE_SIZE
        import torch

        src = torch.ones(1 << 25, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        return 0
    if args.pmc_summary:
        return pmc_summary(args.pmc_summary, args.trace_csv)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(a multi-GPU z-slab decomposition of the grid is not implemented)", file=sys.stderr)
        return 2
    import importlib

    import torch

    vls = importlib.import_module("2018_7_vesicle_amd")
    torch.cuda.set_device(0)
    # the metric's configuration: the vesicle step at 256^3 (32-mode velocity),
    # FP32 bulk, beside the plain FP64 path on the same workload
    r32 = pipeline_line(vls, args.n, args.modes, "fp32", args.steps, args.warmup)
    progress(f"FP32 pipeline at {args.n}^3 done")
    r64 = pipeline_line(vls, args.n, args.modes, "fp64", args.steps, args.warmup)
    cfg = r32["cfg"]
    stage_us = r32["stage_us"]
    achieved = B_ALG_EXTEND_FP32 * r32["points"] / (stage_us * 1e-6) / 1e9
    counters = stage_counters(f"extend_fp32_{args.n}")
    roofline = {
        "kernel": "FP32-bulk WENO5-RK3 extension stage (k_extend_stage<R, float> launches)",
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": counters.get("traffic_bytes") if counters else None,
        "avg_launch_us": stage_us,
        "avg_launch_us_by_difference": r32["stage_us_by_difference"],
        "extension_call_ms": r32["extension_call_ms"],
        "extension_setup_ms": r32["extension_setup_ms"],
        "bytes_per_point": B_ALG_EXTEND_FP32,
    }
    # what a plain stream with the stage's access mix reaches on this GPU
    # (five float arrays in, one out, 16.8 M points each: 403 MB per pass)
    achievable = stream_rate(vls, 1 << 24) / 1e9
    roofline["stream_GBs"] = achievable
    roofline["frac_of_stream"] = achieved / achievable
    if counters and counters.get("valu_insts"):
        # VALU wave-instructions of one launch over the launch time, against the
        # FP32 issue rate this GPU shows (vls_probe_valu_rate)
        peak = valu_rate(vls, False)
        rate = counters["valu_insts"] / (stage_us * 1e-6)
        roofline["valu"] = {"insts_per_launch": counters["valu_insts"],
                            "insts_per_point": counters["valu_insts"] * 64 / r32["points"],
                            "achieved_per_s": rate, "peak_per_s": peak, "frac": rate / peak,
                            "issue_frac": rate / (CUS * CLOCK_HZ)}
    out = {
        "metric": "WENO grid-pt updates/s (FP32-bulk vesicle-step pipeline at 256^3, beside plain FP64)",
        "value": r32["value"],
        "unit": "grid-pt updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (bulk stencils; f64 subcell root solve, interface lanes and area/volume sums)",
        "data": "synthetic",
        "config": {
            "workload": f"config-3 vesicle step: {cfg.n}^3 on [{cfg.lo},{cfg.hi}]^3, ellipsoid {cfg.axes}, "
                        f"{cfg.modes}-mode curl velocity (seed {cfg.seed}), WENO5-RK3 advection CFL {cfg.cfl}, "
                        f"{cfg.ext_iters}-iteration extension, reinit {cfg.reinit_iters} iterations every "
                        f"{cfg.reinit_every} steps, surface redistance {cfg.redist_iters}, cosine delta/H 1.5dx; "
                        "FP32 bulk",
            "grid": [cfg.n, cfg.n, cfg.n],
            "points": r32["points"],
            "updates_per_step_avg": r32["updates_per_step_avg"],
        },
        "fp64": {"value": r64["value"], "ms_per_step": r64["ms_per_step"], "extension_stage_us": r64["stage_us"],
                 "dtype": "f64", "note": "the plain FP64 GPU path on the same workload and seed"},
        "fp32_over_fp64": r32["value"] / r64["value"],
        "roofline": roofline,
        "diagnostics": {"area_first_last": r32["area"], "volume_first_last": r32["volume"], "dt": r32["dt"],
                        "fp32_vs_fp64_area_rel_last": abs(r32["area"][1] / r64["area"][1] - 1),
                        "fp32_vs_fp64_volume_rel_last": abs(r32["volume"][1] / r64["volume"][1] - 1)},
    }
    del r32, r64
    torch.cuda.empty_cache()
    progress(f"main line: FP32 {out['value'] / 1e9:.1f} G, FP64 {out['fp64']['value'] / 1e9:.1f} G updates/s")
    if not args.main_only:
        out["config2_128"] = config2_line(vls, args.steps, args.warmup)
        progress("config2_128 done")
        out["config1"] = config1(vls, not args.no_cpu)
        progress("config1 done")
        if args.precision_steps > 0:
            out["config3"] = config3(args.precision_steps)
            progress("config3 done")
        if args.batch_steps > 0:
            out["config4"] = config4(args.batch_steps)
            progress("config4 done")
        if not args.no_adjoint:
            out["config5"] = config5()
            progress("config5 done")
        out["variants"] = variants(not args.no_cpu)
        progress("variants done")
        out["toolbox"] = toolbox(not args.no_cpu)
        progress("toolbox done")
        out["relax"] = relax(not args.no_cpu)
        progress("relax done")
        if not args.no_cpu:
            out["cpu_baseline"] = cpu_baseline(cfg)
            progress("cpu_baseline done")
    print(json.dumps(out), flush=True)
    return 0

if __name__ == "__main__":
    sys.exit(main())
