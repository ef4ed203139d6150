"""Parity at the benchmark's full sizes: the 1M-point column (the continuum
half of the headline scene), the 1M-point hybrid scene with its 47,700
circles, the 20k-circle DEM column and the 64-scene batch (6.66M points).

The FP64 restatement finishes two steps of the 1M scenes in about a second,
so those are compared directly (FP64: relative error <= 1e-12 per field, the
coupled set exactly; FP32 bulk: the tolerances of the small FP32 tests).
The batch is checked through a size-independent property: every scene of
the full batch evolves bit for bit as it does alone.  Mass and momentum
are checked on the whole 1M grid: P2G conserves them exactly up to
rounding (PhysicsGrid.cpp:207, 316-327).
"""
import numpy as np
import pytest

import oracle
from libhybridgrains_amd import HG_FP32, HG_FP64, GpuDem, GpuHybrid, GpuMpm, scenes

from ._util import assert_close, rel_err

pytestmark = pytest.mark.gpu

DT = 5e-6


def _print(msg):
    print(msg, flush=True)


def test_mpm_1m_fp64_matches_oracle(hg_ctx):
    _print("mpm 1M fp64, 2 steps")
    s = scenes.mpm_1m(seed=3)
    g = GpuMpm(s, HG_FP64, ctx=hg_ctx)
    o = oracle.OracleMpm(s)
    g.flow(DT, 2)
    o.flow(DT, 2)
    p, op = g.points(), o.points()
    assert len(p["q"]) == 1_000_000
    for k in ("q", "v", "J", "be_bar"):
        assert_close(p[k], op[k], 1e-12, k)


def test_mpm_1m_fp32_graph_matches_oracle(hg_ctx):
    _print("mpm 1M fp32 graph, 4 steps, 32x8 tiles")
    s = scenes.mpm_1m(seed=4)
    g = GpuMpm(s, HG_FP32, ctx=hg_ctx, tile=(32, 8))
    o = oracle.OracleMpm(s)
    g.flow(DT, 4, graph=True)
    o.flow(DT, 4)
    p, op = g.points(), o.points()
    assert np.max(np.abs(p["q"] - op["q"])) <= 1e-8
    assert rel_err(p["v"], op["v"]) <= 2e-3
    assert np.max(np.abs(p["J"] - op["J"])) <= 2e-6
    assert np.max(np.abs(p["be_bar"] - op["be_bar"])) <= 2e-6


def test_mpm_1m_grid_matches_oracle_and_conserves(hg_ctx):
    _print("mpm 1M fp64 grid, 1 step")
    s = scenes.mpm_1m(seed=5)
    rng = np.random.default_rng(5)
    s.v = rng.normal(scale=0.1, size=s.v.shape)
    g = GpuMpm(s, HG_FP64, ctx=hg_ctx)
    g.set_keep_grid(True)
    o = oracle.OracleMpm(s)
    g.flow(DT, 1)
    o.flow(DT, 1)
    grid, ogrid = g.grid(), o.grid()
    for k in ("mass", "mom", "mom_new", "force", "vel", "acc"):
        assert_close(grid[k], ogrid[k], 1e-12, k)
    m = np.asarray(s.m, float)
    assert abs(grid["mass"].sum() - m.sum()) <= 1e-12 * m.sum()
    mv = m[:, None] * np.asarray(s.v, float)
    got = grid["mom"].sum(axis=0)
    assert np.all(np.abs(got - mv.sum(axis=0)) <= 1e-12 * np.abs(mv).sum(axis=0))


def test_hybrid_1m_fp64_matches_oracle(hg_ctx):
    _print("hybrid 1M + 47.7k circles fp64, 2 steps")
    m, d, w = scenes.hybrid_1m(seed=6)
    g = GpuHybrid(m, d, w, HG_FP64, ctx=hg_ctx)
    o = oracle.OracleHybrid(m, d, w)
    g.step(DT, 2)
    o.step(DT, 2)
    assert o.coupled.sum() > 1000
    assert np.array_equal(g.coupled(), o.coupled)
    p, op = g.points(), o.mpm.points()
    for k in ("q", "v", "J", "be_bar"):
        assert_close(p[k], op[k], 1e-12, "mpm " + k)
    b, ob = g.bodies(), o.dem.state()
    assert_close(b["q"], ob["q"], 1e-12, "dem q")
    assert_close(b["v"], ob["v"], 1e-12, "dem v")


def test_dem_20k_fp64_matches_oracle(hg_ctx):
    _print("dem 20k fp64, 20 steps")
    s = scenes.dem_column(n_bodies=20000, seed=7)
    rng = np.random.default_rng(7)
    s.v[:, :2] = rng.normal(scale=0.05, size=(s.nbodies, 2))
    g = GpuDem(s, HG_FP64, ctx=hg_ctx)
    o = oracle.OracleDem(s)
    g.flow(1e-5, 20)
    o.flow(1e-5, 20)
    st, ost = g.state(), o.state()
    assert_close(st["q"], ost["q"], 1e-12, "q")
    assert_close(st["v"], ost["v"], 1e-10, "v")


def test_batched_64_scenes_match_alone_bitwise(hg_ctx):
    _print("batch of 64 scenes fp32 graph vs alone")
    ss = scenes.mpm_batch_scenes(64)
    assert sum(len(x.q) for x in ss) > 6_000_000
    batch = GpuMpm(ss, HG_FP32, ctx=hg_ctx, tile=(32, 8))
    batch.flow(DT, 3, graph=True)
    for k in range(64):
        alone = GpuMpm(ss[k], HG_FP32, ctx=hg_ctx)
        alone.flow(DT, 3)
        pa, pb = alone.points(), batch.points(k)
        alone.close()
        for f in pa:
            assert np.array_equal(pa[f], pb[f]), (k, f)
