"""GPU parity of the MPM explicit step against the CPU FP64 restatement.

Tolerances (relative to the largest magnitude of each field):
  * FP64 path, one step: <= 1e-12 for every grid field and point field.  The
    only difference from the restatement is the summation order of the node
    sums (the engine sums each node over its stencil cells, points in id
    order; the reference adds point by point), which changes results by a
    few ulps.
  * FP64 path, 100 steps: <= 1e-10.
  * FP32 path, 100 steps: positions <= 1e-6 of the domain, velocities and
    J within the FP32 storage error (stated per field below).
"""
import numpy as np
import pytest

import oracle
from libhybridgrains_amd import HG_BASIS_CUBIC, HG_BASIS_LINEAR, HG_FP32, HG_FP64, GpuMpm, HgError
from libhybridgrains_amd import scenes

from ._util import assert_close, rel_err

pytestmark = pytest.mark.gpu

DT = 1.0e-5


def _small_scene(seed=0, basis=0):
    # config-0 geometry at a coarser column so the oracle stays fast
    return scenes.column_collapse(domain=(0.4, 0.3), h=0.005, column=(0.1, 0.1), seed=seed,
                                  basis=basis)


def _grid_fields_close(g, o, tol):
    for k in ("mass", "mom", "mom_new", "force", "vel", "acc"):
        assert_close(g[k], o[k], tol, k)


def _point_fields_close(p, o, tol, keys=("q", "v", "be_bar", "J")):
    for k in keys:
        assert_close(p[k], o[k], tol, k)


@pytest.mark.parametrize("basis", [0, HG_BASIS_LINEAR, HG_BASIS_CUBIC])
def test_one_step_fp64_grid_and_points(hg_ctx, basis):
    s = _small_scene(basis=basis)
    # give the column some motion so every term is exercised
    rng = np.random.default_rng(3)
    s.v = rng.normal(scale=0.05, size=s.v.shape)
    gpu = GpuMpm(s, HG_FP64, ctx=hg_ctx)
    gpu.set_keep_grid(True)
    ref = oracle.OracleMpm(s)
    gpu.flow(DT, 1)
    ref.flow(DT, 1)
    _grid_fields_close(gpu.grid(), ref.grid(), 1e-12)
    _point_fields_close(gpu.points(), ref.points(), 1e-12)


def test_one_step_fp64_deformed_state(hg_ctx):
    """Pre-strained points with J above and below 1 and yielding shear, so the
    stress, the no-tension branch and the return map all run."""
    s = _small_scene()
    rng = np.random.default_rng(5)
    N = s.npoints
    J = 1.0 + rng.uniform(-0.02, 0.01, N)
    a = rng.uniform(-0.05, 0.05, N)
    b = np.stack([1 + a, rng.uniform(-0.03, 0.03, N)], 1)
    be = np.zeros((N, 4))
    be[:, 0] = b[:, 0]
    be[:, 1] = b[:, 1]
    be[:, 2] = b[:, 1]
    be[:, 3] = (1 + b[:, 1] ** 2) / b[:, 0]  # det = 1
    s.be_bar = be
    s.J = J
    s.v = rng.normal(scale=0.2, size=s.v.shape)
    gpu = GpuMpm(s, HG_FP64, ctx=hg_ctx)
    gpu.set_keep_grid(True)
    ref = oracle.OracleMpm(s)
    gpu.flow(DT, 1)
    ref.flow(DT, 1)
    _grid_fields_close(gpu.grid(), ref.grid(), 1e-12)
    _point_fields_close(gpu.points(), ref.points(), 1e-12)


def test_hundred_steps_fp64(hg_ctx):
    s = _small_scene(seed=1)
    gpu = GpuMpm(s, HG_FP64, ctx=hg_ctx)
    ref = oracle.OracleMpm(s)
    gpu.flow(DT, 100)
    ref.flow(DT, 100)
    _point_fields_close(gpu.points(), ref.points(), 1e-10)


def test_phases_equal_flow(hg_ctx):
    """flow_zero/first/second_phase give bit-identical results to flow()."""
    s = _small_scene(seed=2)
    a = GpuMpm(s, HG_FP64, ctx=hg_ctx)
    b = GpuMpm(s, HG_FP64, ctx=hg_ctx)
    for _ in range(5):
        a.flow(DT, 1)
        b.flow_zero_phase(DT)
        b.flow_first_phase(DT)
        b.flow_second_phase(DT)
    pa, pb = a.points(), b.points()
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k


def test_graph_equals_stream_and_is_deterministic(hg_ctx):
    s = _small_scene(seed=4)
    a = GpuMpm(s, HG_FP64, ctx=hg_ctx)
    b = GpuMpm(s, HG_FP64, ctx=hg_ctx)
    c = GpuMpm(s, HG_FP64, ctx=hg_ctx)
    a.flow(DT, 7)
    b.flow(DT, 7, graph=True)
    c.flow(DT, 3, graph=True)
    c.flow(DT, 4, graph=True)
    pa, pb, pc = a.points(), b.points(), c.points()
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k
        assert np.array_equal(pa[k], pc[k]), k


def test_batch_invariance_bitwise(hg_ctx):
    """Each scene of a batch evolves bit-identically to the scene run alone."""
    ss = [_small_scene(seed=10 + k) for k in range(3)]
    ss[1] = scenes.column_collapse(domain=(0.5, 0.3), h=0.004, column=(0.08, 0.16), seed=11)
    batch = GpuMpm(ss, HG_FP64, ctx=hg_ctx)
    batch.flow(DT, 20)
    for k, s in enumerate(ss):
        alone = GpuMpm(s, HG_FP64, ctx=hg_ctx)
        alone.flow(DT, 20)
        pa, pb = alone.points(), batch.points(k)
        for f in pa:
            assert np.array_equal(pa[f], pb[f]), (k, f)


@pytest.mark.parametrize("tile", [(16, 16), (16, 8), (32, 8), (64, 4)])
def test_tile_shapes_agree(hg_ctx, tile):
    s = _small_scene(seed=6)
    ref = oracle.OracleMpm(s)
    gpu = GpuMpm(s, HG_FP64, ctx=hg_ctx, tile=tile)
    gpu.flow(DT, 10)
    ref.flow(DT, 10)
    _point_fields_close(gpu.points(), ref.points(), 1e-11)


def test_fp32_tile_shapes_agree_bitwise(hg_ctx):
    # FP32 mode sums a cell's terms in FP32 before the FP64 node sums; the
    # cells, not the tiles or chunks, set that order
    s = _small_scene(seed=10)
    out = []
    for tile in ((16, 16), (16, 8), (32, 8), (64, 4)):
        g = GpuMpm(s, HG_FP32, ctx=hg_ctx, tile=tile)
        g.flow(DT, 10)
        out.append(g.points())
    for p in out[1:]:
        for k in p:
            assert np.array_equal(p[k], out[0][k]), k


def test_fp32_hundred_steps(hg_ctx):
    s = _small_scene(seed=7)
    gpu = GpuMpm(s, HG_FP32, ctx=hg_ctx)
    ref = oracle.OracleMpm(s)
    gpu.flow(DT, 100)
    ref.flow(DT, 100)
    p, o = gpu.points(), ref.points()
    dq = np.max(np.abs(p["q"] - o["q"]))
    assert dq <= 1e-8, dq  # metres, on a 0.1 m column
    assert rel_err(p["v"], o["v"]) <= 2e-3
    assert np.max(np.abs(p["J"] - o["J"])) <= 2e-6
    assert np.max(np.abs(p["be_bar"] - o["be_bar"])) <= 2e-6


def test_point_leaving_grid_is_reported(hg_ctx):
    s = _small_scene(seed=8)
    s.plane_x = np.zeros((0, 2))
    s.plane_n = np.zeros((0, 2))
    s.plane_lower = np.zeros(0)
    s.plane_behavior = np.zeros(0, np.int32)
    s.v[:] = [-50.0, 0.0]  # 0.5 mm per step towards the grid edge 10 mm away
    gpu = GpuMpm(s, HG_FP64, ctx=hg_ctx)
    gpu.flow(DT, 40)
    with pytest.raises(HgError) as ei:
        gpu.sync()
    assert ei.value.code == -4
    assert gpu.scene_error(0) & 1


def test_hold_occupies_the_stream(hg_ctx):
    import time
    g = GpuMpm(_small_scene(seed=12), HG_FP64, ctx=hg_ctx)
    g.sync()
    t0 = time.perf_counter()
    g.hold(30000.0)
    g.flow(DT, 1)
    g.sync()
    assert time.perf_counter() - t0 >= 0.025
    with pytest.raises(HgError):
        g.hold(-1.0)


@pytest.mark.parametrize("tile", [(16, 16), (32, 8), (64, 4)])
def test_fp32_points_per_pass_change_no_result(hg_ctx, tile):
    # one or two threads per cell take a tile's points in different passes and
    # lanes; the exact node sums make the steps bit-identical
    s = _small_scene(seed=11)
    runs = []
    for pp in (1, 2):
        g = GpuMpm(s, HG_FP32, ctx=hg_ctx, tile=tile)
        g.set_points_per_pass(pp)
        assert g.tuning() == {"tile": tile, "points_per_pass": pp}
        g.flow(DT, 4)
        g.flow(DT, 4, graph=True)
        runs.append(g.points())
    for k in runs[0]:
        assert np.array_equal(runs[0][k], runs[1][k]), k


@pytest.mark.parametrize("prec", [HG_FP64, HG_FP32])
def test_autotune_picks_a_shape_and_changes_no_result(hg_ctx, prec):
    # FP32 times whole fused steps and restores the saved state afterwards
    s = _small_scene(seed=9)
    a = GpuMpm(s, prec, ctx=hg_ctx)
    b = GpuMpm(s, prec, ctx=hg_ctx)
    a.flow(DT, 3)
    b.flow(DT, 3)
    tune = b.autotune(DT, reps=2)
    assert tune["tile"] in ((16, 16), (16, 8), (32, 8), (64, 4))
    assert all(v > 0 for v in tune["us"].values())
    a.flow(DT, 5, graph=True)
    b.flow(DT, 5, graph=True)
    pa, pb = a.points(), b.points()
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k
