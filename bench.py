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

Beside it:
  fp64_plain      the same scene on the FP64 path;
  batched         64 independent MPM column collapses (5k-500k points, 6.66M
                  in all) in one launch per stage, FP32 and FP64, and the
                  same 64 scenes stepped one handle at a time in FP64 (the
                  plain FP64 path); batched_fp32_over_plain_fp64 is the ratio
                  of the first to the last, on the same scenes;
  dem_column_20k  the 20k-circle DEM column, FP64 and FP32;
  hybrid_with_resampling
                  config 2 with its zone resampling every 100 steps, through
                  the C++ driver (HybridGrains2DSim): step time, pass time
                  and the amortised rate;
  served          Poisson streams of full-size scene requests (10k-2M
                  particles, 500-5000 steps, 50/30/20 MPM/DEM/hybrid) through
                  the scene server at three arrival rates: p50/p99 completion
                  latency and aggregate particle-steps/s;
  kernels         every kernel of the FP32 hybrid step: average duration
                  (rocprofv3 kernel trace) and measured HBM bytes (FETCH_SIZE,
                  WRITE_SIZE, scaled by their readings for a copy kernel of
                  known size), as GB/s and as a fraction of 8 TB/s;
  cpu_baseline    the CPU restatement (FP64, the reference's operation order,
                  its OpenMP structure for the MPM part) on the host cores:
                  the 6.4k-point MPM column, the 20k-circle DEM column, the 1M
                  hybrid scene and the largest batched scene, each at the
                  host's thread share and at one thread.

roofline: the MPM transfer kernels of the headline step.  In FP32 these are
mpm_fused (G2P of one step and the rasterisation of the next, per tile) and
mpm_grid (node sums to the grid update); in FP64 mpm_p2g and mpm_g2p.
Algorithmic bytes: 106 B per particle-step in FP32 (reads x, y as FP64, v,
the three b_bar entries, J, m, V0: 48 B; writes x, y, v, b_bar, J: 40 B;
72 B per active node at 0.25 active nodes per point), 172 B in FP64, over the
pair's summed average launch duration, measured with HIP events on the
engine's stream.  traffic: the pair's HBM bytes per step from the counters.
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

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "MPM+DEM particle-steps/s @1M: batched FP32 vs plain FP64 GPU; served p99 latency"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per particle-step of the MPM step (P2G + grid + G2P)
ALG_BYTES = {"fp32": 106.0, "fp64": 172.0}
TRANSFER_KERNELS = {"fp32": ("mpm_fused", "mpm_grid"), "fp64": ("mpm_p2g", "mpm_g2p")}
# the same kernels under their names in rocprofv3's records
TRANSFER_KERNELS_PROF = {"fp32": ("mpm_fused_kernel", "mpm_grid_kernel")}
DT_1M = 5.0e-6


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _time_hybrid(m, d, w, precision, steps, warmup, dt):
    """Eager steps (the grain branch on its own stream, overlapping the
    continuum) and hipGraph steps (six steps captured on both streams);
    returns both walls and the
    per-kernel event times."""
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
    return walls, kt, tune


def _time_mpm(scenes_, precision, steps, warmup, dt, tune=True):
    """hipGraph steps of one handle holding scenes_; returns seconds per step."""
    from libhybridgrains_amd import GpuMpm

    g = GpuMpm(scenes_, precision)
    t = g.autotune(dt) if tune else None
    g.flow(dt, warmup, graph=True)
    g.sync()
    t0 = time.perf_counter()
    g.flow(dt, steps, graph=True)
    g.sync()
    wall = time.perf_counter() - t0
    g.close()
    return wall / steps, t


def _time_batched(args, dt=1e-5):
    """The 64 independent column collapses (5k-500k points, log-uniform):
    batched FP32, batched FP64, and plain FP64 (one handle per scene, stepped
    one after the other, as a caller without batching runs them)."""
    from libhybridgrains_amd import HG_FP32, HG_FP64, scenes

    ss = scenes.mpm_batch_scenes(nscenes=64, seed=0)
    N = sum(s.npoints for s in ss)
    steps = max(20, args.steps // 2)
    out = {"scenes": len(ss), "points": N, "steps": steps}
    for name, prec in (("fp32", HG_FP32), ("fp64", HG_FP64)):
        log(f"batched {name}")
        sps, tune = _time_mpm(ss, prec, steps, args.warmup, dt)
        out["batched_" + name] = {"value": N / sps, "ms_per_step": 1e3 * sps, "autotune": tune}
    log("plain fp64, one scene per handle")
    total = 0.0
    for s in ss:
        sps, _ = _time_mpm([s], HG_FP64, steps, 3, dt, tune=False)
        total += sps
    out["plain_fp64"] = {"value": N / total, "ms_per_step": 1e3 * total,
                         "note": "64 handles, each stepped with hipGraph, one after another"}
    out["unit"] = "particle-steps/s"
    return out


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


def _served_one(reqs, rate, timeout_s):
    """One Poisson stream through a fresh scene server; progress on stderr."""
    from libhybridgrains_amd import HG_FP32
    from libhybridgrains_amd.server import DONE, FAILED, CANCELLED, SceneServer

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
    last, states = -1, {}
    while True:
        states = {r: srv.poll(r)[0] for r in ids}
        ended = sum(1 for s in states.values() if s in (DONE, FAILED, CANCELLED))
        if ended != last:
            log(f"served {rate:g} req/s: {ended}/{len(ids)} ended "
                f"({time.perf_counter() - t0:.1f} s)")
            last = ended
        if ended == len(ids) or time.perf_counter() - t0 > timeout_s:
            break
        time.sleep(1.0)
    st = srv.stats()
    done = [r for r in ids if states[r] == DONE]
    failed = {r: srv.error(r) for r in ids if states[r] == FAILED}
    srv.close()
    return {
        "rate_req_per_s": rate, "requests": len(ids), "completed": len(done),
        "failed": list(failed.values())[:3],
        "p50_latency_s": st["p50_latency_s"], "p99_latency_s": st["p99_latency_s"],
        "mean_latency_s": st["mean_latency_s"],
        "particle_steps_per_s": st["work_steps"] / st["span_s"] if st["span_s"] > 0 else None,
    }


def _served(args):
    """Poisson streams of MPM / DEM / hybrid requests (50/30/20) at each rate:
    requests arrive while others run and are admitted as memory and slots
    free up.  Full sizes: 10k-2M particles, 500-5000 steps (--served-small:
    10k-200k particles, 200-1000 steps)."""
    from libhybridgrains_amd import scenes

    rates = [float(r) for r in args.served_rates.split(",") if r]
    if args.served_small:
        size = dict(n_min=1e4, n_max=2e5, steps_min=200, steps_max=1000)
        sizes = "10k-200k particles, 200-1000 steps"
    else:
        size = dict(n_min=1e4, n_max=2e6, steps_min=500, steps_max=5000)
        sizes = "10k-2M particles, 500-5000 steps"
    out = {"sizes": sizes, "precision": "fp32", "mix": "50% MPM, 30% DEM, 20% hybrid",
           "streams": []}
    for k, rate in enumerate(rates):
        log(f"served: generating {args.served_requests} requests at {rate:g} req/s")
        reqs = scenes.served_requests(args.served_requests, seed=1 + k, rate=rate, **size)
        out["streams"].append(_served_one(reqs, rate, args.served_timeout))
    return out


def _resampling(steps=200, every=100):
    """Config 2 with its zone resampling, through the reference-named C++
    driver (hybridgrains2d_cli, HybridGrains2DSim): the hybrid column written
    as the reference's scene files (the continuum and a grain packing over the
    whole 0.5 m column; the resampler keeps the grains of the zones within
    2 cm of the free sides and seeds the continuum), FP32, `steps` steps with
    a resampling pass every `every` steps, and the same steps with none.
    The difference per pass is the pass's cost: device state to the host,
    the host zone tools, new device handles."""
    import re

    from libhybridgrains_amd import scenes

    exe = ROOT / "libhybridgrains_amd" / "bin" / "hybridgrains2d_cli"
    d = Path(tempfile.mkdtemp(prefix="hg_resample_"))
    runs = {}
    try:
        for name, freq in (("with", every), ("without", 10 ** 9)):
            log(f"resampling: {name} passes, {steps} steps")
            xml = scenes.write_hybrid_column_xml(d / name, level=0.02, resample_every=freq)
            r = subprocess.run([str(exe), "-n", str(steps), "-f", str(steps), "-p", "fp32", str(xml)],
                               cwd=str(d / name), capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                return {"error": (r.stderr or r.stdout)[-500:]}
            st = re.findall(r"grains (\d+) points (\d+) resamplings (\d+)", r.stdout)
            done = re.search(r"Simulation complete: (\d+) steps in ([0-9.]+) s", r.stdout)
            g, p, n = (int(v) for v in st[-1])
            runs[name] = {"seconds": float(done.group(2)), "grains": g, "points": p, "passes": n}
    finally:
        shutil.rmtree(d, ignore_errors=True)
    w, wo = runs["with"], runs["without"]
    npass = w["passes"] - wo["passes"]
    pass_s = (w["seconds"] - wo["seconds"]) / npass if npass > 0 else None
    step_s = wo["seconds"] / steps
    amort = step_s + (pass_s or 0.0) / every
    return {"steps": steps, "resample_every": every, "runs": runs,
            "pass_s": pass_s, "step_ms": 1e3 * step_s, "amortised_ms_per_step": 1e3 * amort,
            "amortised_particle_steps_per_s": (w["grains"] + w["points"]) / amort,
            "note": "steps and passes through hybridgrains2d_cli, FP32; one status download "
                    "at the end of each run"}


def _copy_gbs(_lib, nbytes=1 << 30, iters=20):
    """Measured streaming-copy bandwidth (read + write) of this GPU."""
    import ctypes as C

    ctx = _lib.default_context()
    out = C.c_double()
    ctx.check(_lib.lib().hg_bench_copy(ctx.handle, nbytes, iters, C.byref(out)))
    return out.value


def _rocprof(opts, mode, extra=()):
    """Runs this script under rocprofv3 <opts> in a child process and returns
    the path of its output folder (None if the profiler is missing or fails)."""
    prof = shutil.which("rocprofv3")
    if not prof:
        return None
    d = tempfile.mkdtemp(prefix="hg_prof_")
    cmd = ["timeout", "-s", "KILL", "90", prof, *opts, "--output-format", "csv", "-d", d,
           "-o", "run", "--", sys.executable, str(ROOT / "bench.py"), "--mode", mode,
           "--steps", "3", "--warmup", "1", *extra]
    env = dict(os.environ, TMPDIR=d)
    try:
        subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL, timeout=120, check=True)
        return Path(d)
    except Exception:
        shutil.rmtree(d, ignore_errors=True)
        return None


def _pmc(counter, mode, extra=()):
    """{kernel: (mean counter value per launch, launches)} for one counter."""
    d = _rocprof(["--pmc", counter], mode, extra)
    if d is None:
        return None
    try:
        path = next(d.rglob("*counter_collection.csv"))
        agg = {}
        with open(path) as f:
            for r in csv.DictReader(f):
                agg.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
        return {k: (sum(v) / len(v), len(v)) for k, v in agg.items()}
    except Exception:
        return None
    finally:
        shutil.rmtree(d, ignore_errors=True)


def _trace(mode, extra=()):
    """{kernel: (average duration in us, launches)} from a kernel trace."""
    d = _rocprof(["--kernel-trace", "--stats"], mode, extra)
    if d is None:
        return None
    try:
        path = next(d.rglob("*kernel_stats.csv"))
        out = {}
        with open(path) as f:
            for r in csv.DictReader(f):
                out[r["Name"]] = (float(r["AverageNs"]) * 1e-3, int(r["Calls"]))
        return out
    except Exception:
        return None
    finally:
        shutil.rmtree(d, ignore_errors=True)


def _short(name: str) -> str:
    """Kernel name without return type, namespaces, template arguments or
    parameters: 'void (anonymous namespace)::mpm_grid_kernel<1>(Args)' ->
    'mpm_grid_kernel'."""
    base = name.replace("(anonymous namespace)::", "")
    depth, out = 0, []
    for ch in base:
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif depth == 0:
            out.append(ch)
    base = "".join(out).split("(")[0].strip()
    return base.split()[-1].split("::")[-1] if base else name


def _kernel_traffic(tile):
    """Per kernel of the FP32 hybrid step: average duration, HBM bytes read
    and written per launch (FETCH_SIZE and WRITE_SIZE, each scaled by its
    reading for the copy kernel that reads and writes 1 GiB), GB/s."""
    t = ("--tile", "%dx%d" % tuple(tile))
    copy_f, copy_w = _pmc("FETCH_SIZE", "copy"), _pmc("WRITE_SIZE", "copy")
    run_f, run_w = _pmc("FETCH_SIZE", "hyb32", t), _pmc("WRITE_SIZE", "hyb32", t)
    trace = _trace("hyb32", t)
    if not (copy_f and copy_w and run_f and run_w and trace):
        return None

    def pick(d, key):
        # the kernel launched most often under that name (the step's, not set-up)
        vals = sorted((n, v) for k, (v, n) in d.items() if _short(k) == key)
        return vals[-1][1] if vals else None

    gib_kib = float(1 << 20)
    cf, cw = pick(copy_f, "hg_copy_kernel"), pick(copy_w, "hg_copy_kernel")
    if not cf or not cw:
        return None
    fscale, wscale = gib_kib / cf, gib_kib / cw
    names = sorted({_short(k) for k in trace})
    kernels = {}
    for nm in names:
        us = pick(trace, nm)
        calls = max(n for k, (v, n) in trace.items() if _short(k) == nm)
        f, w = pick(run_f, nm), pick(run_w, nm)
        if us is None or f is None or w is None or calls < 3:
            continue  # set-up kernels
        rb, wb = f * 1024.0 * fscale, w * 1024.0 * wscale
        gbs = (rb + wb) / (us * 1e-6) / 1e9
        kernels[nm] = {"avg_us": us, "launches": calls, "read_MB": rb / 1e6,
                       "write_MB": wb / 1e6, "GB_per_s": gbs, "frac_of_peak": gbs / HBM_PEAK_GBS}
    return {"kernels": kernels, "fetch_scale": fscale, "write_scale": wscale}


def _cpu_info():
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                model = ln.split(":", 1)[1].strip()
    except Exception:
        pass
    # one GPU's share of the host: the threads OpenMP is given (16 on the
    # GPU hosts), at most the cores this process may run on
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or nproc
    return {"nproc": nproc, "os_cpu_count": os.cpu_count(), "model": model,
            "threads": max(1, min(nproc, share))}


def _timed(step, budget_s, warm=2, max_steps=200):
    """warm untimed steps, then steps until budget_s or max_steps; steps/s."""
    step(warm)
    k = 0
    t0 = time.perf_counter()
    while True:
        step(1)
        k += 1
        el = time.perf_counter() - t0
        if el > budget_s or k >= max_steps:
            return k, el


def _cpu_baseline(budget_s):
    """The CPU restatement (FP64) on bounded samples of four workloads, at the
    host's thread share and at one thread (DEM: one thread, as the reference's
    penalty step is serial)."""
    import oracle
    from libhybridgrains_amd import scenes

    info = _cpu_info()
    T = info["threads"]
    legs = {}

    def mpm_leg(name, s, dt, share):
        for thr in (T, 1):
            o = oracle.OracleMpm(s)
            k, el = _timed(lambda n: o.flow_omp(dt, n, thr), budget_s * share,
                           warm=2 if s.npoints > 1e5 else 20)
            legs[f"{name}_t{thr}"] = {"value": s.npoints * k / el, "unit": "particle-steps/s",
                                      "threads": thr, "steps": k, "points": s.npoints}
            del o

    log("cpu: MPM column 6.4k")
    mpm_leg("mpm_column_6k", scenes.mpm_config0(), 1e-5, 0.06)
    log("cpu: DEM column 20k")
    s = scenes.dem_column()
    od = oracle.OracleDem(s)
    k, el = _timed(lambda n: od.flow(1e-5, n), budget_s * 0.1, warm=20)
    legs["dem_column_20k_t1"] = {"value": s.nbodies * k / el, "unit": "body-steps/s",
                                 "threads": 1, "steps": k, "bodies": s.nbodies}
    del od
    log("cpu: largest batched scene")
    big = max(scenes.mpm_batch_scenes(nscenes=64, seed=0), key=lambda x: x.npoints)
    mpm_leg("batched_largest", big, 1e-5, 0.15)
    log("cpu: hybrid 1M")
    m, d, w = scenes.hybrid_1m()
    for thr in (T, 1):
        o = oracle.OracleHybrid(m, d, w)
        k, el = _timed(lambda n: o.step_omp(DT_1M, n, thr), budget_s * 0.27, warm=1)
        legs[f"hybrid_1m_t{thr}"] = {"value": (m.npoints + d.nbodies) * k / el,
                                     "unit": "particle-steps/s", "threads": thr, "steps": k}
        del o
    head = legs[f"hybrid_1m_t{T}"]
    return {
        "value": head["value"],
        "unit": "particle-steps/s",
        "cores": T,
        "kind": "port",
        "sample": f"{head['steps']} steps of the 1M hybrid scene ({m.npoints} points + "
                  f"{d.nbodies} circles), FP64, MPM part on {T} OpenMP threads (per-thread "
                  f"grids), grains serial",
        "host": info,
        "legs": legs,
    }


def _profile_mode(args):
    """Only the steps of one workload, for profilers."""
    from libhybridgrains_amd import HG_FP32, HG_FP64, GpuDem, GpuHybrid, GpuMpm, scenes

    prec = HG_FP32 if args.mode.endswith("32") else HG_FP64
    dt = DT_1M
    if args.mode.startswith("mpm"):
        g = cont = GpuMpm(scenes.mpm_1m(), prec)
        step = lambda n: g.flow(dt, n, graph=True)  # noqa: E731
    elif args.mode.startswith("batch"):
        g = cont = GpuMpm(scenes.mpm_batch_scenes(nscenes=64, seed=0), prec)
        dt = 1e-5
        step = lambda n: g.flow(dt, n, graph=True)  # noqa: E731
    elif args.mode.startswith("dem"):
        g, cont = GpuDem(scenes.dem_column(), prec), None
        dt = 1e-5
        step = lambda n: g.flow(dt, n, graph=True)  # noqa: E731
    else:
        m, d, w = scenes.hybrid_1m()
        g = GpuHybrid(m, d, w, prec)
        cont = g.continuum
        step = lambda n: g.step(dt, n, graph=not args.eager)  # noqa: E731
    if cont is not None:
        if args.keep_grid:
            cont.set_keep_grid(True)  # FP32: the separate P2G and G2P kernels
        if args.tile:
            tw, th = (int(v) for v in args.tile.split("x"))
            cont.ctx.check(cont.L.hg_mpm_set_tile(cont._h, tw, th))
        else:
            cont.autotune(dt)
    step(args.warmup)
    g.sync()
    t0 = time.perf_counter()
    step(args.steps)
    g.sync()
    print(json.dumps({"mode": args.mode,
                      "ms_per_step": 1e3 * (time.perf_counter() - t0) / args.steps}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=30.0,
                    help="seconds of CPU work over all CPU legs")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 runs")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-resampling", action="store_true",
                    help="skip the config-2 run with zone resampling")
    ap.add_argument("--served-requests", type=int, default=16, help="requests per rate")
    ap.add_argument("--served-rates", default="0.5,4,20")
    ap.add_argument("--served-small", action="store_true")
    ap.add_argument("--served-timeout", type=float, default=240.0)
    ap.add_argument("--tile", default="", help="P2G tile WxH for the profiling modes "
                                               "(default: autotune)")
    ap.add_argument("--eager", action="store_true",
                    help="hybrid profiling modes: eager steps (the grain branch on its own "
                         "stream) instead of hipGraph replays")
    ap.add_argument("--keep-grid", action="store_true",
                    help="profiling modes: keep every grid field (FP32 then runs the separate "
                         "P2G and G2P kernels instead of the fused step)")
    ap.add_argument("--mode", default="full",
                    choices=["full", "hyb32", "hyb64", "mpm32", "mpm64", "batch32", "batch64",
                             "dem32", "dem64", "copy"],
                    help="full: the JSON line; the others run only those steps (for "
                         "profilers); copy: the streaming-copy reference kernel")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus > 1 is not supported "
                                   "(independent scenes shard across GPUs by placement)"}))
        return 2

    from libhybridgrains_amd import HG_FP32, HG_FP64, scenes
    from libhybridgrains_amd import _lib

    if args.mode == "copy":
        print(json.dumps({"copy_gbs": _copy_gbs(_lib, iters=3)}), flush=True)
        return 0
    if args.mode != "full":
        return _profile_mode(args)

    m, d, w = scenes.hybrid_1m()
    N, B = m.npoints, d.nbodies
    steps = args.steps
    legs = {}
    for name, prec in (("fp32", HG_FP32), ("fp64", HG_FP64)):
        log(f"hybrid 1M {name}")
        walls, kt, tune = _time_hybrid(m, d, w, prec, steps, args.warmup, DT_1M)
        per = {k: v[0] / v[1] for k, v in kt.items()}
        transfer_ms = sum(per[k] for k in TRANSFER_KERNELS[name])
        alg = ALG_BYTES[name] * N
        achieved = alg / (transfer_ms * 1e-3) / 1e9
        legs[name] = {
            "value": (N + B) * steps / walls["eager"],
            "ms_per_step": 1e3 * walls["eager"] / steps,
            "graph_ms_per_step": 1e3 * walls["graph"] / steps,
            "kernels_ms": per,
            "autotune": tune,
            "roofline": {
                "bound": "hbm",
                "kernel": "+".join(TRANSFER_KERNELS[name]),
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
    batched = None
    if not args.no_batched:
        batched = _time_batched(args)
    log("dem")
    dem = _time_dem(args)
    resampling = None if args.no_resampling else _resampling()
    served = None
    if not args.no_served:
        served = _served(args)
    copy_gbs = _copy_gbs(_lib)
    kernels = None
    if not args.no_traffic:
        log("per-kernel traffic (rocprofv3)")
        kernels = _kernel_traffic(f32["autotune"]["tile"])
        if kernels:
            ks = kernels["kernels"]
            tk = TRANSFER_KERNELS_PROF["fp32"]
            if all(k in ks for k in tk):
                f32["roofline"]["traffic"] = sum(1e6 * (ks[k]["read_MB"] + ks[k]["write_MB"])
                                                 for k in tk)
    cpu = None
    if not args.no_cpu:
        cpu = _cpu_baseline(args.cpu_budget)

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
        "batched": batched,
        "batched_fp32_over_plain_fp64": (batched["batched_fp32"]["value"] /
                                         batched["plain_fp64"]["value"]) if batched else None,
        "batched_fp32_over_batched_fp64": (batched["batched_fp32"]["value"] /
                                           batched["batched_fp64"]["value"]) if batched else None,
        "served": served,
        "dem_column_20k": dem,
        "hybrid_with_resampling": resampling,
        "hbm_copy_gbs_measured": copy_gbs,
        "kernels_fp32_rocprof": kernels,
        "kernels_ms_fp32": f32["kernels_ms"],
        "kernels_ms_fp64": f64["kernels_ms"],
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
