#!/usr/bin/env python3
"""Benchmark of the explicit-step hot path on one MI355X.

Prints ONE JSON line.

Headline (value): the 1M-point hybrid scene -- the MPM column of 1,000,000
points on 2000 x 1000 cells (h = 1 mm, 4 points per cell) with a 0.1 m deep
discrete zone of about 48k circles at its free surface -- stepped with the
node-node coupled step (MPM + weighted DEM + exchange), FP32 bulk with FP64
positions and return map; the grain branch runs on its own stream beside the
continuum (the engine's hg_hybrid_step).  value = (points + circles) x steps
/ wall time, inputs resident in HBM.  The same steps as hipGraph replays (one
stream) are reported as graph_ms_per_step.

Beside it: the same scene on the plain FP64 GPU path (fp64_plain), the
batched FP32 form (64 independent MPM scenes, 5k-500k points each, in one
launch per stage), a served stream of scene requests (continuous batching,
p50/p99 completion latency), the 20k-circle DEM column, and the CPU
restatement on the host cores (cpu_baseline).

roofline: the MPM transfer kernels of the headline step (mpm_p2g + mpm_g2p):
106 B per particle-step in FP32 (reads x, y as FP64, v, the three b_bar
entries, J, m, V0: 48 B; writes x, y, v, b_bar, J: 40 B; 72 B per active
node at 0.25 active nodes per point) over their summed average launch
duration, measured with HIP events on the engine's stream.  traffic: the
HBM bytes of one P2G + G2P pair from rocprofv3 FETCH_SIZE / WRITE_SIZE,
scaled by what the counters read for the library's copy kernel.
"""
from __future__ import annotations

import argparse
import csv
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "MPM+DEM particle-steps/s @1M: batched FP32 vs plain FP64 GPU; served p99 latency"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per particle-step of the MPM step (P2G + grid + G2P)
ALG_BYTES = {"fp32": 106.0, "fp64": 172.0}
DT_1M = 5.0e-6


def _time_hybrid(m, d, w, precision, steps, warmup, dt):
    """Eager steps (the grain branch on its own stream, overlapping the
    continuum) and hipGraph steps (one stream); returns both walls."""
    from libhybridgrains_amd import GpuHybrid

    h = GpuHybrid(m, d, w, precision)
    tune = h.continuum.autotune(dt)
    walls = {}
    for graph in (True, False):
        h.step(dt, warmup, graph=graph)
        h.sync()
        t0 = time.perf_counter()
        h.step(dt, steps, graph=graph)
        h.sync()
        walls["graph" if graph else "eager"] = time.perf_counter() - t0
    wall = walls["eager"]
    # per-kernel durations over the same number of steps (events on the
    # stream); the steps queue up behind a hold so that they run back to back
    h.continuum.set_profiling(True)
    h.discrete.set_profiling(True)
    h.continuum.hold(3e5)
    h.step(dt, steps)
    h.sync()
    kt = dict(h.continuum.kernel_times())
    kt.update({"dem:" + k: v for k, v in h.discrete.kernel_times().items()})
    h.close()
    kt["_tune"] = tune
    kt["_graph_wall"] = walls["graph"]
    return wall, kt


def _time_batched(args, dt=1e-5):
    """64 independent column collapses (5k-500k points, log-uniform) stepped
    by the same launches, FP32, hipGraph step."""
    from libhybridgrains_amd import HG_FP32, GpuMpm, scenes

    ss = scenes.mpm_batch_scenes(nscenes=64, seed=0)
    N = sum(s.npoints for s in ss)
    g = GpuMpm(ss, HG_FP32)
    tune = g.autotune(dt)
    g.flow(dt, args.warmup, graph=True)
    g.sync()
    steps = max(10, args.steps // 2)
    t0 = time.perf_counter()
    g.flow(dt, steps, graph=True)
    g.sync()
    wall = time.perf_counter() - t0
    g.close()
    return {"value": N * steps / wall, "unit": "particle-steps/s", "scenes": len(ss),
            "points": N, "steps": steps, "ms_per_step": 1e3 * wall / steps,
            "autotune": tune}


def _time_dem(args, dt=1e-5):
    """Circle DEM column (20k bodies, penalty contacts with friction): body-steps/s."""
    from libhybridgrains_amd import HG_FP32, HG_FP64, GpuDem, scenes

    s = scenes.dem_column()
    out = {"bodies": s.nbodies}
    for name, prec in (("fp64", HG_FP64), ("fp32", HG_FP32)):
        g = GpuDem(s, prec)
        g.flow(dt, max(args.warmup, 50), graph=True)  # settle into contact
        g.sync()
        t0 = time.perf_counter()
        g.flow(dt, args.steps, graph=True)
        g.sync()
        wall = time.perf_counter() - t0
        g.set_profiling(True)
        g.hold(2e5)
        g.flow(dt, args.steps)
        g.sync()
        kt = g.kernel_times()
        g.close()
        out[name] = {"body_steps_per_s": s.nbodies * args.steps / wall,
                     "ms_per_step": 1e3 * wall / args.steps,
                     "kernels_ms": {k: v[0] / v[1] for k, v in kt.items()}}
    return out


def _served(args):
    """A Poisson stream of MPM / DEM / hybrid requests (50/30/20) through the
    scene server: requests arrive while others run and are admitted as memory
    and slots free up.  Sizes log-uniform in 10k-200k particles, horizons
    200-1000 steps (the full-size stream, 10k-2M and 500-5000 steps, runs with
    --served-full)."""
    from libhybridgrains_amd import HG_FP32, scenes
    from libhybridgrains_amd.server import DONE, SceneServer

    if args.served_full:
        reqs = scenes.served_requests(args.served_requests, seed=1, rate=args.served_rate,
                                      n_min=1e4, n_max=2e6, steps_min=500, steps_max=5000)
    else:
        reqs = scenes.served_requests(args.served_requests, seed=1, rate=args.served_rate)
    srv = SceneServer(budget_bytes=240 << 30, slice_steps=50)
    t0 = time.perf_counter()
    ids = []
    for arrival, kind, payload, dt, nsteps in reqs:
        now = time.perf_counter() - t0
        if arrival > now:
            time.sleep(arrival - now)
        if kind == "mpm":
            ids.append(srv.submit_mpm(payload, dt, nsteps, precision=HG_FP32))
        elif kind == "dem":
            ids.append(srv.submit_dem(payload, dt, nsteps, precision=HG_FP32))
        else:
            m, d, w = payload
            ids.append(srv.submit_hybrid(m, d, dt, nsteps, weights=w, precision=HG_FP32))
    srv.drain(600)
    st = srv.stats()
    ok = sum(1 for r in ids if srv.poll(r)[0] == DONE)
    srv.close()
    return {
        "requests": len(ids), "completed": ok, "rate_req_per_s": args.served_rate,
        "p50_latency_s": st["p50_latency_s"], "p99_latency_s": st["p99_latency_s"],
        "mean_latency_s": st["mean_latency_s"],
        "particle_steps_per_s": st["work_steps"] / st["span_s"] if st["span_s"] > 0 else None,
        "sizes": "10k-2M particles, 500-5000 steps" if args.served_full
        else "10k-200k particles, 200-1000 steps",
        "precision": "fp32",
    }


def _copy_gbs(_lib, nbytes=1 << 30, iters=20):
    """Measured streaming-copy bandwidth (read + write) of this GPU."""
    import ctypes as C

    ctx = _lib.default_context()
    out = C.c_double()
    ctx.check(_lib.lib().hg_bench_copy(ctx.handle, nbytes, iters, C.byref(out)))
    return out.value


def _pmc(counter, mode, extra=()):
    """Runs this script under rocprofv3 --pmc <counter> in a child process and
    returns {kernel: (mean counter value per launch, launches)}; None if the
    profiler is unavailable or fails."""
    prof = shutil.which("rocprofv3")
    if not prof:
        return None
    d = tempfile.mkdtemp(prefix="hg_pmc_")
    cmd = ["timeout", "-s", "KILL", "90", prof, "--pmc", counter, "--output-format", "csv", "-d", d,
           "-o", "run", "--", sys.executable, str(ROOT / "bench.py"), "--mode", mode,
           "--steps", "3", "--warmup", "1", *extra]
    env = dict(os.environ, TMPDIR=d)
    try:
        subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL, timeout=120, check=True)
        path = next(Path(d).rglob("*counter_collection.csv"))
        agg = {}
        with open(path) as f:
            for r in csv.DictReader(f):
                agg.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
        return {k: (sum(v) / len(v), len(v)) for k, v in agg.items()}
    except Exception:
        return None
    finally:
        shutil.rmtree(d, ignore_errors=True)


def _traffic(tile):
    """HBM bytes per P2G + G2P pair of the headline step (P2G tile as chosen
    by the headline run), from FETCH_SIZE and WRITE_SIZE (KiB per dispatch),
    each scaled by its reading for the copy kernel that reads and writes 1 GiB."""
    t = ("--tile", "%dx%d" % tuple(tile))
    copy_f, copy_w = _pmc("FETCH_SIZE", "copy"), _pmc("WRITE_SIZE", "copy")
    run_f, run_w = _pmc("FETCH_SIZE", "hyb32", t), _pmc("WRITE_SIZE", "hyb32", t)
    if not (copy_f and copy_w and run_f and run_w):
        return None, None

    def pick(d, key):
        vals = sorted((n, v) for k, (v, n) in d.items() if key in k)
        return vals[-1][1] if vals else None

    gib_kib = float(1 << 20)
    cf, cw = pick(copy_f, "hg_copy_kernel"), pick(copy_w, "hg_copy_kernel")
    if not cf or not cw:
        return None, None
    fscale, wscale = gib_kib / cf, gib_kib / cw
    by = {}
    for name in ("mpm_p2g", "mpm_g2p"):
        f, w = pick(run_f, name), pick(run_w, name)
        if f is None or w is None:
            return None, None
        by[name] = {"read_bytes": f * 1024.0 * fscale, "write_bytes": w * 1024.0 * wscale}
    total = sum(v["read_bytes"] + v["write_bytes"] for v in by.values())
    return total, {"per_kernel": by, "fetch_scale": fscale, "write_scale": wscale}


def _cpu_baseline(m, d, w, dt, budget_s):
    """The CPU restatement of the hybrid step (serial FP64, the reference's
    operation order) on a bounded sample of the headline scene, and the MPM
    part alone with the reference's OpenMP structure (per-thread grids)."""
    import oracle

    o = oracle.OracleHybrid(m, d, w)
    o.step(dt, 1)
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.step(dt, 1)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 200:
            break
    work = m.npoints + d.nbodies
    nthreads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    nthreads = max(1, min(nthreads, 16))  # the host share of one GPU
    om = oracle.OracleMpm(m)
    om.flow_omp(dt, 1, nthreads)
    k = 0
    t1 = time.perf_counter()
    while True:
        om.flow_omp(dt, 1, nthreads)
        k += 1
        el2 = time.perf_counter() - t1
        if el2 > budget_s / 2 or k >= 200:
            break
    return {
        "value": work * steps / el,
        "unit": "particle-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{steps} hybrid steps of the headline scene ({m.npoints} points + "
                  f"{d.nbodies} circles), FP64, serial",
        "mpm_openmp": {"value": m.npoints * k / el2, "cores": nthreads,
                       "sample": f"{k} MPM steps of the 1M-point column, FP64, OpenMP x{nthreads}"},
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 traffic runs")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--served-requests", type=int, default=40)
    ap.add_argument("--served-rate", type=float, default=8.0)
    ap.add_argument("--served-full", action="store_true")
    ap.add_argument("--tile", default="", help="P2G tile WxH for the profiling modes "
                                               "(default: autotune)")
    ap.add_argument("--mode", default="full",
                    choices=["full", "hyb32", "hyb64", "mpm32", "mpm64", "copy"],
                    help="full: the JSON line; hyb32/hyb64/mpm32/mpm64: only those steps "
                         "(for profilers); copy: the streaming-copy reference kernel")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus > 1 is not supported "
                                   "(independent scenes shard across GPUs by placement)"}))
        return 2

    from libhybridgrains_amd import HG_FP32, HG_FP64, GpuHybrid, GpuMpm, scenes
    from libhybridgrains_amd import _lib

    if args.mode == "copy":
        print(json.dumps({"copy_gbs": _copy_gbs(_lib, iters=3)}), flush=True)
        return 0
    if args.mode in ("mpm32", "mpm64", "hyb32", "hyb64"):
        prec = HG_FP32 if args.mode.endswith("32") else HG_FP64
        # the same set-up as the full run: the P2G tile it chose (--tile) or
        # the autotuner's choice, hipGraph steps
        if args.mode.startswith("mpm"):
            g = GpuMpm(scenes.mpm_1m(), prec)
            cont = g
            step = lambda n: g.flow(DT_1M, n, graph=True)  # noqa: E731
        else:
            m, d, w = scenes.hybrid_1m()
            g = GpuHybrid(m, d, w, prec)
            cont = g.continuum
            step = lambda n: g.step(DT_1M, n, graph=True)  # noqa: E731
        if args.tile:
            tw, th = (int(v) for v in args.tile.split("x"))
            cont.ctx.check(cont.L.hg_mpm_set_tile(cont._h, tw, th))
        else:
            cont.autotune(DT_1M)
        step(args.warmup)
        g.sync()
        t0 = time.perf_counter()
        step(args.steps)
        g.sync()
        print(json.dumps({"mode": args.mode,
                          "ms_per_step": 1e3 * (time.perf_counter() - t0) / args.steps}),
              flush=True)
        return 0

    m, d, w = scenes.hybrid_1m()
    N, B = m.npoints, d.nbodies
    steps = args.steps
    legs = {}
    for name, prec in (("fp32", HG_FP32), ("fp64", HG_FP64)):
        print(f"[bench] hybrid 1M {name}", file=sys.stderr, flush=True)
        wall, kt = _time_hybrid(m, d, w, prec, steps, args.warmup, DT_1M)
        tune = kt.pop("_tune")
        graph_wall = kt.pop("_graph_wall")
        transfer_ms = sum(kt[k][0] for k in ("mpm_p2g", "mpm_g2p")) / steps
        alg = ALG_BYTES[name] * N
        achieved = alg / (transfer_ms * 1e-3) / 1e9
        legs[name] = {
            "value": (N + B) * steps / wall,
            "ms_per_step": 1e3 * wall / steps,
            "graph_ms_per_step": 1e3 * graph_wall / steps,
            "kernels_ms": {k: v[0] / v[1] for k, v in kt.items()},
            "autotune": tune,
            "roofline": {
                "bound": "hbm",
                "kernel": "mpm_p2g+mpm_g2p",
                "achieved": achieved,
                "peak": HBM_PEAK_GBS,
                "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "traffic": None,
                "alg_bytes_per_launch": alg,
                "launch_ms": transfer_ms,
            },
        }
    f32, f64 = legs["fp32"], legs["fp64"]
    print("[bench] batched", file=sys.stderr, flush=True)
    batched = _time_batched(args)
    print("[bench] dem", file=sys.stderr, flush=True)
    dem = _time_dem(args)
    served = None
    if not args.no_served:
        print("[bench] served", file=sys.stderr, flush=True)
        served = _served(args)
    copy_gbs = _copy_gbs(_lib)
    traffic_detail = None
    if not args.no_traffic:
        print("[bench] traffic (rocprofv3)", file=sys.stderr, flush=True)
        traffic, traffic_detail = _traffic(f32["autotune"]["tile"])
        f32["roofline"]["traffic"] = traffic
    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = _cpu_baseline(m, d, w, DT_1M, args.cpu_budget)

    line = {
        "metric": METRIC,
        "value": f32["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": f32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "hybrid_column_1M: 1,000,000 MPM points (2000x1000 cells, h=1mm, 4 ppc, "
                        "0.5x0.5 m column) + 47,700 circles (r=0.5mm+-10%, 0.1 m discrete zone), "
                        "node-node coupled step, dt=5e-6, grain branch on a second stream",
            "graph_ms_per_step": f32["graph_ms_per_step"],
            "points": N,
            "circles": B,
            "precision": "fp32 bulk, fp64 positions and return map",
        },
        "roofline": f32["roofline"],
        "cpu_baseline": cpu,
        "fp64_plain": {"value": f64["value"], "ms_per_step": f64["ms_per_step"],
                       "graph_ms_per_step": f64["graph_ms_per_step"],
                       "roofline": f64["roofline"], "autotune": f64["autotune"]},
        "autotune": f32["autotune"],
        "fp32_over_fp64": f32["value"] / f64["value"],
        "batched_fp32": batched,
        "batched_fp32_over_plain_fp64": batched["value"] / f64["value"],
        "served": served,
        "dem_column_20k": dem,
        "hbm_copy_gbs_measured": copy_gbs,
        "traffic_detail": traffic_detail,
        "kernels_ms_fp32": f32["kernels_ms"],
        "kernels_ms_fp64": f64["kernels_ms"],
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
