"""Writes longrun.json: long-run observables of the granular collapses,
computed with the CPU restatement of the reference (oracle/, FP64, the
reference's operation order), for the GPU tests in test_longrun_gpu.py.

Each case records, from the state of the run:
  runout   the largest x of the material minus the left wall's x;
  profile  the largest y of the material in each x-bin of one cell width
           from x = 0 (null where a bin is empty), at the end;
  ke       the kinetic energy sum(m |v|^2) / 2 (translational for grains)
           every `every` steps, from step 0.

Cases:
  mpm_column_6k   the 6.4k-point MPM column (1.0 x 0.5 m, h = 5 mm, 0.2 m
                  column), dt = 1e-5, 20,000 steps;
  dem_column_4k   a 4,000-circle column of the 20k-circle generator (0.08 m
                  wide, radii 0.9-1.1 mm), dt = 1e-5, 20,000 steps;
  batch_<k>       three of the 64 batched column collapses (seed 0; the
                  three with the fewest points), dt = 1e-5, 5,000 steps.

Run from the repository root (a few minutes on 8 cores):
    python tests/golden/make_longrun.py
"""
from __future__ import annotations

import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from libhybridgrains_amd import scenes  # noqa: E402

OUT = Path(__file__).resolve().parent / "longrun.json"


def profile(x, y, width, nbins):
    b = np.floor(x / width).astype(np.int64)
    ok = (b >= 0) & (b < nbins)
    h = np.full(nbins, -np.inf)
    np.maximum.at(h, b[ok], y[ok])
    return [None if not np.isfinite(v) else float(v) for v in h]


def mpm_case(s, dt, steps, every, wall_x):
    o = oracle.OracleMpm(s)
    m = np.asarray(s.m, float)
    ke = []
    for k in range(0, steps + 1, every):
        if k:
            o.flow(dt, every)
        p = o.points()
        ke.append(float(0.5 * np.sum(m * np.sum(p["v"] ** 2, axis=1))))
    p = o.points()
    nb = int(s.cell_count[0])
    return {"dt": dt, "steps": steps, "every": every, "wall_x": wall_x,
            "bin_width": float(s.cell_width), "nbins": nb,
            "runout": float(p["q"][:, 0].max() - wall_x),
            "profile": profile(p["q"][:, 0], p["q"][:, 1], float(s.cell_width), nb),
            "ke": ke, "points": int(s.npoints)}


def dem_case(s, dt, steps, every, bin_width, nbins):
    o = oracle.OracleDem(s)
    m = np.asarray(s.mass, float)[:, 0]
    ke = []
    for k in range(0, steps + 1, every):
        if k:
            o.flow(dt, every)
        st = o.state()
        ke.append(float(0.5 * np.sum(m * np.sum(st["v"][:, :2] ** 2, axis=1))))
    q = o.state()["q"]
    return {"dt": dt, "steps": steps, "every": every, "wall_x": 0.0, "bin_width": bin_width,
            "nbins": nbins, "runout": float(q[:, 0].max()),
            "profile": profile(q[:, 0], q[:, 1], bin_width, nbins), "ke": ke,
            "bodies": int(s.nbodies)}


def main() -> int:
    out = {}
    t0 = time.time()
    s = scenes.mpm_config0()
    out["mpm_column_6k"] = mpm_case(s, 1e-5, 20000, 500, wall_x=2 * float(s.cell_width))
    print(f"mpm_column_6k {time.time() - t0:.0f} s", flush=True)
    d = scenes.dem_column(n_bodies=4000, width=0.08)
    out["dem_column_4k"] = dem_case(d, 1e-5, 20000, 500, bin_width=2e-3, nbins=600)
    print(f"dem_column_4k {time.time() - t0:.0f} s", flush=True)
    batch = scenes.mpm_batch_scenes(nscenes=64, seed=0)
    small = sorted(range(len(batch)), key=lambda k: batch[k].npoints)[:3]
    for k in small:
        s = batch[k]
        c = mpm_case(s, 1e-5, 5000, 250, wall_x=2 * float(s.cell_width))
        c["scene"] = k
        out[f"batch_{k}"] = c
        print(f"batch_{k} ({s.npoints} points) {time.time() - t0:.0f} s", flush=True)
    OUT.write_text(json.dumps(out, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
