"""The scene server on the GPU: every request's result is the same when it
runs alone as in a mix of concurrent requests (MPM and DEM bit for bit; the
hybrid's grain-side node sums are FP64 atomics, so <= 1e-12 relative), a
request whose state leaves the grid fails alone, cancellation and the memory
budget work, and the footprint estimate bounds what a handle allocates."""
import ctypes as C

import numpy as np
import pytest

from libhybridgrains_amd import HG_FP32, HG_FP64, GpuDem, GpuHybrid, GpuMpm, scenes
from libhybridgrains_amd import _lib
from libhybridgrains_amd.server import CANCELLED, DONE, FAILED, SceneServer, estimate_bytes

from ._util import assert_close

pytestmark = pytest.mark.gpu

DT = 1e-5


def _mpm(seed, h=0.005, col=(0.1, 0.1)):
    return scenes.column_collapse(domain=(0.4, 0.3), h=h, column=col, seed=seed)


def _dem(seed, n=1500):
    s = scenes.dem_column(n_bodies=n, seed=seed)
    s.v[:, :2] = np.random.default_rng(seed).normal(scale=0.05, size=(n, 2))
    return s


def test_results_alone_equal_results_in_the_mix(hg_ctx):
    srv = SceneServer(ctx=hg_ctx, slice_steps=7)
    m1, m2 = _mpm(1), _mpm(2, h=0.004, col=(0.08, 0.12))
    d1 = _dem(3)
    hm, hd, hw = scenes.hybrid_column(seed=4)
    ids = {
        "m1": srv.submit_mpm(m1, DT, 40),
        "d1": srv.submit_dem(d1, DT, 33),
        "h": srv.submit_hybrid(hm, hd, DT, 25, weights=hw),
        "m2": srv.submit_mpm(m2, DT, 61, precision=HG_FP32),
    }
    srv.drain(180)
    for k, rid in ids.items():
        assert srv.poll(rid)[0] == DONE, (k, srv.error(rid))
    a = GpuMpm(m1, HG_FP64, ctx=hg_ctx)
    a.flow(DT, 40)
    pa, ps = a.points(), srv.points(ids["m1"])
    for f in pa:
        assert np.array_equal(pa[f], ps[f]), f
    b = GpuMpm(m2, HG_FP32, ctx=hg_ctx)
    b.flow(DT, 61)
    pb, ps = b.points(), srv.points(ids["m2"])
    for f in pb:
        assert np.array_equal(pb[f], ps[f]), f
    c = GpuDem(d1, HG_FP64, ctx=hg_ctx)
    c.flow(DT, 33)
    sc, ss = c.state(), srv.bodies(ids["d1"])
    assert np.array_equal(sc["q"], ss["q"]) and np.array_equal(sc["v"], ss["v"])
    h = GpuHybrid(hm, hd, hw, HG_FP64, ctx=hg_ctx)
    h.step(DT, 25)
    assert_close(srv.points(ids["h"])["q"], h.points()["q"], 1e-12, "hybrid points")
    assert_close(srv.bodies(ids["h"])["v"], h.bodies()["v"], 1e-12, "hybrid grains")
    st = srv.stats()
    assert st["completed"] == 4 and st["failed"] == 0
    assert st["p99_latency_s"] >= st["p50_latency_s"] > 0
    assert st["bytes_in_use"] == 0
    srv.close()


def test_failing_request_is_retired_alone(hg_ctx):
    srv = SceneServer(ctx=hg_ctx, slice_steps=10)
    bad = _mpm(5)
    bad.plane_x = np.zeros((0, 2))
    bad.plane_n = np.zeros((0, 2))
    bad.plane_lower = np.zeros(0)
    bad.plane_behavior = np.zeros(0, np.int32)
    bad.v[:] = [-50.0, 0.0]  # leaves the grid within 40 steps
    good = _mpm(6)
    rb = srv.submit_mpm(bad, DT, 60)
    rg = srv.submit_mpm(good, DT, 60)
    srv.drain(120)
    assert srv.poll(rb)[0] == FAILED
    assert "valid domain" in srv.error(rb) or "grid" in srv.error(rb)
    assert srv.poll(rg)[0] == DONE
    a = GpuMpm(good, HG_FP64, ctx=hg_ctx)
    a.flow(DT, 60)
    assert np.array_equal(a.points()["q"], srv.points(rg)["q"])
    srv.close()


def test_budget_and_cancellation(hg_ctx):
    m = _mpm(7)
    need = estimate_bytes("mpm", mpm=m)
    # room for one scene at a time
    srv = SceneServer(ctx=hg_ctx, budget_bytes=int(need * 1.5), slice_steps=5)
    r1 = srv.submit_mpm(m, DT, 200)
    r2 = srv.submit_mpm(m, DT, 10)
    r3 = srv.submit_mpm(m, DT, 10)
    st, _ = srv.poll(r2)
    srv.cancel(r3)
    srv.drain(120)
    assert srv.poll(r1)[0] == DONE and srv.poll(r2)[0] == DONE
    assert srv.poll(r3)[0] == CANCELLED
    # r2 could only start once r1 had released its memory
    assert srv.latency(r2) > 0
    too_big = SceneServer(ctx=hg_ctx, budget_bytes=need // 2)
    rx = too_big.submit_mpm(m, DT, 10)
    assert too_big.poll(rx)[0] == FAILED
    assert "budget" in too_big.error(rx)
    too_big.close()
    srv.close()


def test_footprint_estimate_bounds_the_allocation(hg_ctx):
    L = _lib.lib()
    for prec in (HG_FP64, HG_FP32):
        m = _mpm(8)
        g = GpuMpm(m, prec, ctx=hg_ctx)
        got = C.c_uint64()
        hg_ctx.check(L.hg_mpm_device_bytes(g._h, C.byref(got)))
        est = estimate_bytes("mpm", mpm=m, precision=prec)
        assert got.value <= est <= got.value * 1.5 + (2 << 20), (prec, got.value, est)
        d = _dem(9)
        gd = GpuDem(d, prec, ctx=hg_ctx)
        hg_ctx.check(L.hg_dem_device_bytes(gd._h, C.byref(got)))
        est = estimate_bytes("dem", dem=d, precision=prec)
        assert got.value <= est <= got.value * 1.5 + (2 << 20), (prec, got.value, est)
