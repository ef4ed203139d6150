"""Parity of the benchmark configurations and the shipped scene on the GPU.

  * Config 0 (the 6,400-point MPM column on the 200 x 100 grid) for 100 steps
    in FP64 against the restatement, with the floor and wall sliding (the
    generator's choice) and with both sticking (the shipped scenes' choice,
    PhysicsGrid.cpp:611-673 and MaterialPoints.cpp:371-423): <= 1e-10 per
    point field; one step of every grid field with sticking planes: <= 1e-12.
    FP32 bulk for 100 steps: positions within 1e-8 m, velocities 2e-3 of the
    largest, J and b_bar within 2e-6 (FP32 storage).
  * The reference's shipped column-collapse continuum file
    (assets/hybridgrains2dnew/column-collapse/mpm.xml: CGS units, sticking
    planes, 7,936 points) read through the XML readers and stepped in FP64:
    one step's grid <= 1e-12, 10 steps <= 1e-11 per point field.
  * The 1M-point hybrid scene (the headline workload) in FP32 bulk for 3
    steps against the FP64 restatement: the coupled grain set may differ only
    at grains whose stencil holds a node within FP32 rounding of the 1e-6
    coupling threshold (the count is stated and bounded); positions within
    1e-9 m, J and b_bar within 2e-6; point velocities within 2e-3 of the
    largest (the FP32 MPM bound: early in the collapse the strain is near
    FP32's resolution of J, so the stress carries that error), grain
    velocities within 1e-4.
  * Polydisperse DEM clusters whose large circles touch 15 partners each,
    against the restatement (which keeps unbounded per-body contact lists):
    40 steps; FP64 contact set exact, q <= 1e-12, v <= 1e-10; FP32 q within
    1e-9 m and v within 1e-4 of the largest.
  * The served stream (config 4) drawn over its full size range (10k-2M
    particles) with a bounded request count and horizon: each request's
    result in the mix equals the same request run alone (MPM and DEM bit
    for bit; hybrid <= 1e-12, its grain-side node sums are FP64 atomics).
"""
import numpy as np
import pytest

import oracle
from libhybridgrains_amd import (HG_FP32, HG_FP64, HG_PLANE_STICKING, GpuDem, GpuHybrid, GpuMpm,
                                 scenes, xmlscene)
from libhybridgrains_amd.server import DONE, SceneServer

from ._util import assert_close, rel_err

pytestmark = pytest.mark.gpu

ASSETS = __import__("pathlib").Path(__file__).parent / "golden" / "assets"


def _print(msg):
    print(msg, flush=True)


def _sticking(s):
    s.plane_behavior = np.full(len(s.plane_behavior), HG_PLANE_STICKING, np.int32)
    return s


@pytest.mark.parametrize("planes", ["sliding", "sticking"])
def test_config0_fp64_100_steps(hg_ctx, planes):
    s = scenes.mpm_config0(seed=0)
    assert s.npoints == 6400 and list(s.cell_count) == [200, 100]
    if planes == "sticking":
        _sticking(s)
    g = GpuMpm(s, HG_FP64, ctx=hg_ctx)
    o = oracle.OracleMpm(s)
    g.flow(1e-5, 100)
    o.flow(1e-5, 100)
    p, op = g.points(), o.points()
    for k in ("q", "v", "J", "be_bar"):
        e = assert_close(p[k], op[k], 1e-10, k)
        _print(f"config 0 {planes} {k}: {e:.2e}")
    # the sticking walls hold the points that reach them
    if planes == "sticking":
        assert np.abs(op["v"]).max() > 0.0


def test_config0_sticking_grid_one_step(hg_ctx):
    s = _sticking(scenes.mpm_config0(seed=1))
    rng = np.random.default_rng(1)
    s.v = rng.normal(scale=0.05, size=s.v.shape)
    g = GpuMpm(s, HG_FP64, ctx=hg_ctx)
    g.set_keep_grid(True)
    o = oracle.OracleMpm(s)
    g.flow(1e-5, 1)
    o.flow(1e-5, 1)
    grid, og = g.grid(), o.grid()
    for k in ("mass", "mom", "mom_new", "force", "vel", "acc"):
        assert_close(grid[k], og[k], 1e-12, k)
    # nodes on or behind a sticking wall end the step with zero momentum
    nx = s.cell_count[0] + 1
    xs = s.grid_min[0] + s.cell_width * (np.arange(len(og["mass"])) % nx)
    behind = (xs <= s.plane_x[0][0]) & (og["mass"] > 0)
    assert behind.sum() > 0
    assert np.all(og["mom_new"][behind] == 0.0) and np.all(grid["mom_new"][behind] == 0.0)


def test_config0_fp32_100_steps(hg_ctx):
    s = scenes.mpm_config0(seed=2)
    g = GpuMpm(s, HG_FP32, ctx=hg_ctx)
    o = oracle.OracleMpm(s)
    g.flow(1e-5, 100)
    o.flow(1e-5, 100)
    p, op = g.points(), o.points()
    dq = np.max(np.abs(p["q"] - op["q"]))
    ev = rel_err(p["v"], op["v"])
    dJ = np.max(np.abs(p["J"] - op["J"]))
    db = np.max(np.abs(p["be_bar"] - op["be_bar"]))
    _print(f"config 0 fp32: q {dq:.2e} m, v {ev:.2e}, J {dJ:.2e}, b {db:.2e}")
    assert dq <= 1e-8 and ev <= 2e-3 and dJ <= 2e-6 and db <= 2e-6


def test_shipped_column_collapse_xml(hg_ctx):
    x = xmlscene.load_mpm(ASSETS / "column-collapse" / "mpm.xml")
    s = x.scene
    assert s.npoints == 7936 and list(s.plane_behavior) == [HG_PLANE_STICKING] * 2
    g = GpuMpm(s, HG_FP64, ctx=hg_ctx)
    g.set_keep_grid(True)
    o = oracle.OracleMpm(s)
    g.flow(x.dt, 1)
    o.flow(x.dt, 1)
    grid, og = g.grid(), o.grid()
    for k in ("mass", "mom", "mom_new", "force", "vel", "acc"):
        assert_close(grid[k], og[k], 1e-12, "grid " + k)
    g.flow(x.dt, 9)
    o.flow(x.dt, 9)
    p, op = g.points(), o.points()
    for k in ("q", "v", "J", "be_bar"):
        e = assert_close(p[k], op[k], 1e-11, k)
        _print(f"shipped column-collapse 10 steps {k}: {e:.2e}")


def test_hybrid_1m_fp32_matches_oracle(hg_ctx):
    _print("hybrid 1M + 47.7k circles fp32, 3 steps")
    m, d, w = scenes.hybrid_1m(seed=8)
    g = GpuHybrid(m, d, w, HG_FP32, ctx=hg_ctx)
    o = oracle.OracleHybrid(m, d, w)
    g.step(5e-6, 3)
    o.step(5e-6, 3)
    c, oc = g.coupled(), o.coupled
    mism = int((c != oc).sum())
    _print(f"coupled: {int(oc.sum())} of {len(oc)} grains, {mism} differ")
    assert oc.sum() > 1000 and mism <= max(5, int(0.002 * oc.sum()))
    p, op = g.points(), o.mpm.points()
    dq = np.max(np.abs(p["q"] - op["q"]))
    ev = rel_err(p["v"], op["v"])
    dJ = np.max(np.abs(p["J"] - op["J"]))
    db = np.max(np.abs(p["be_bar"] - op["be_bar"]))
    b, ob = g.bodies(), o.dem.state()
    bq = np.max(np.abs(b["q"][:, :2] - ob["q"][:, :2]))
    bv = rel_err(b["v"][:, :2], ob["v"][:, :2])
    _print(f"points: q {dq:.2e} m, v {ev:.2e}, J {dJ:.2e}, b {db:.2e}; "
           f"grains: q {bq:.2e} m, v {bv:.2e}")
    assert dq <= 1e-9 and ev <= 2e-3 and dJ <= 2e-6 and db <= 2e-6
    assert bq <= 1e-9 and bv <= 1e-4


@pytest.mark.parametrize("precision", [HG_FP64, HG_FP32])
def test_dem_more_than_eight_contacts(hg_ctx, precision):
    # large circles ringed by 15 small ones: the handle sizes its per-body
    # contact capacity from the radii (32 here) and follows the restatement
    s = scenes.dem_rosettes(seed=11)
    g = GpuDem(s, precision, ctx=hg_ctx)
    o = oracle.OracleDem(s)
    g.flow(1e-5, 40)
    o.flow(1e-5, 40)
    cc = o.circle_contacts()
    per_body = np.bincount(np.concatenate([cc["i"], cc["j"]]), minlength=s.nbodies)
    assert per_body.max() == 15
    gc, oc = g.contacts(), o.contacts()
    assert len(gc["i"]) == len(oc["i"])
    gs, os_ = g.state(), o.state()
    if precision == HG_FP64:
        assert np.array_equal(gc["i"], oc["i"]) and np.array_equal(gc["j"], oc["j"])
        assert_close(gs["q"], os_["q"], 1e-12, "q")
        assert_close(gs["v"], os_["v"], 1e-10, "v")
    else:
        assert np.abs(gs["q"][:, :2] - os_["q"][:, :2]).max() <= 1e-9
        assert rel_err(gs["v"][:, :2], os_["v"][:, :2]) <= 1e-4


def test_served_full_size_range_alone_equals_mix(hg_ctx):
    reqs = scenes.served_requests(8, seed=5, rate=8.0, n_min=1e4, n_max=2e6, steps_min=100,
                                  steps_max=300)
    _print("served: " + ", ".join(f"{k} {(p.npoints if k == 'mpm' else p.nbodies if k == 'dem' else p[0].npoints + p[1].nbodies)}"
                                 for _, k, p, _, _ in reqs))
    srv = SceneServer(ctx=hg_ctx, slice_steps=50)
    ids = []
    for _, kind, p, dt, n in reqs:
        if kind == "mpm":
            ids.append(srv.submit_mpm(p, dt, n, precision=HG_FP32))
        elif kind == "dem":
            ids.append(srv.submit_dem(p, dt, n, precision=HG_FP32))
        else:
            ids.append(srv.submit_hybrid(p[0], p[1], dt, n, weights=p[2], precision=HG_FP64))
    srv.drain(300)
    for rid in ids:
        assert srv.poll(rid)[0] == DONE, srv.error(rid)
    for rid, (_, kind, p, dt, n) in zip(ids, reqs):
        if kind == "mpm":
            a = GpuMpm(p, HG_FP32, ctx=hg_ctx)
            a.flow(dt, n, graph=True)
            pa, ps = a.points(), srv.points(rid)
            for f in pa:
                assert np.array_equal(pa[f], ps[f]), (rid, f)
            a.close()
        elif kind == "dem":
            a = GpuDem(p, HG_FP32, ctx=hg_ctx)
            a.flow(dt, n, graph=True)
            sa, ss = a.state(), srv.bodies(rid)
            assert np.array_equal(sa["q"], ss["q"]) and np.array_equal(sa["v"], ss["v"]), rid
            a.close()
        else:
            a = GpuHybrid(p[0], p[1], p[2], HG_FP64, ctx=hg_ctx)
            a.step(dt, n, graph=True)
            assert_close(srv.points(rid)["q"], a.points()["q"], 1e-12, "hybrid points")
            assert_close(srv.bodies(rid)["v"], a.bodies()["v"], 1e-12, "hybrid grains")
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    srv.close()


def _host_observables(q, v, m, x0, w, nbins):
    ke = float(np.sum(0.5 * m * (v[:, 0] ** 2 + v[:, 1] ** 2)))
    prof = np.full(nbins, np.nan)
    k = np.floor((q[:, 0] - x0) / w)
    ok = (k >= 0) & (k < nbins)
    for b in np.unique(k[ok]).astype(int):
        prof[b] = q[ok & (k == b), 1].max()
    return ke, q[:, 0].max(), q[:, 1].max(), prof


@pytest.mark.parametrize("precision", [HG_FP64, HG_FP32])
def test_observables_match_the_downloaded_state(hg_ctx, precision):
    s = scenes.mpm_config0(seed=4)
    g = GpuMpm(s, precision, ctx=hg_ctx)
    g.flow(1e-5, 300)
    p = g.points()
    # the masses as the handle stores them (FP32 mode rounds them)
    cast = np.float32 if precision == HG_FP32 else np.float64
    m = np.asarray(s.m).astype(cast).astype(float)
    ob = g.observables(x0=0.0, bin_width=0.01, nbins=100)
    ke, fx, ty, prof = _host_observables(p["q"], p["v"], m, 0.0, 0.01, 100)
    assert ke > 0 and abs(ob["ke"] - ke) <= 1e-12 * ke
    assert ob["front"] == fx and ob["top"] == ty
    assert np.array_equal(np.isnan(ob["profile"]), np.isnan(prof))
    assert np.array_equal(ob["profile"][~np.isnan(prof)], prof[~np.isnan(prof)])
    # reproducible, and fewer bins reuse the same scratch
    again = g.observables(x0=0.0, bin_width=0.01, nbins=100)
    assert again["ke"] == ob["ke"] and np.array_equal(again["profile"], ob["profile"],
                                                      equal_nan=True)
    few = g.observables(x0=0.05, bin_width=0.02, nbins=7)
    _, _, _, prof7 = _host_observables(p["q"], p["v"], m, 0.05, 0.02, 7)
    assert np.array_equal(few["profile"], prof7, equal_nan=True)
    d = scenes.dem_column(n_bodies=3000, seed=4)
    gd = GpuDem(d, precision, ctx=hg_ctx)
    gd.flow(1e-5, 200)
    st = gd.state()
    od = gd.observables(x0=0.0, bin_width=0.005, nbins=64)
    ke, fx, ty, prof = _host_observables(st["q"][:, :2], st["v"][:, :2],
                                         d.mass[:, 0].astype(cast).astype(float), 0.0, 0.005, 64)
    assert abs(od["ke"] - ke) <= 1e-12 * ke and od["front"] == fx and od["top"] == ty
    assert np.array_equal(od["profile"], prof, equal_nan=True)
