"""GPU parity of the hybrid (MPM + circle DEM) coupled step against the CPU
FP64 restatement (oracle/hybrid_oracle.c).

The grain-side node sums (grain mass and momentum) are FP64 atomics on the
GPU, so their summation order varies from run to run by a few ulps; every
other sum follows the restatement's order or the MPM engine's fixed order.
Tolerances (relative to each field's largest magnitude): one step <= 1e-12,
twenty steps <= 1e-10.  The coupled set is compared exactly: it depends only
on the continuum node masses, which are deterministic.
"""
import numpy as np
import pytest

import oracle
from libhybridgrains_amd import HG_FP32, HG_FP64, GpuDem, GpuHybrid, GpuMpm, scenes

from ._util import assert_close

pytestmark = pytest.mark.gpu

DT = 1e-5


def _compare(g: GpuHybrid, o: oracle.OracleHybrid, tol_q, tol_v):
    assert np.array_equal(g.coupled(), o.coupled)
    p, op = g.points(), o.mpm.points()
    assert_close(p["q"], op["q"], tol_q, "mpm q")
    assert_close(p["v"], op["v"], tol_v, "mpm v")
    assert_close(p["J"], op["J"], tol_v, "mpm J")
    assert_close(p["be_bar"], op["be_bar"], tol_v, "mpm be_bar")
    b, ob = g.bodies(), o.dem.state()
    assert_close(b["q"], ob["q"], tol_q, "dem q")
    assert_close(b["v"], ob["v"], tol_v, "dem v")


def test_one_step_fp64_matches_oracle(hg_ctx):
    m, d, w = scenes.hybrid_column(seed=1)
    rng = np.random.default_rng(1)
    d.v[:, :2] = rng.normal(scale=0.05, size=(d.nbodies, 2))
    m.v = rng.normal(scale=0.05, size=m.v.shape)
    g = GpuHybrid(m, d, w, HG_FP64, ctx=hg_ctx)
    o = oracle.OracleHybrid(m, d, w)
    g.step(DT, 1)
    o.step(DT, 1)
    assert o.coupled.sum() > 50
    _compare(g, o, 1e-12, 1e-12)


def test_twenty_steps_fp64_matches_oracle(hg_ctx):
    m, d, w = scenes.hybrid_column(seed=2)
    g = GpuHybrid(m, d, w, HG_FP64, ctx=hg_ctx)
    o = oracle.OracleHybrid(m, d, w)
    g.step(DT, 20)
    o.step(DT, 20)
    assert o.coupled.sum() > 50
    _compare(g, o, 1e-11, 1e-10)


@pytest.mark.parametrize("prec", [HG_FP64, HG_FP32])
def test_graph_matches_stream(hg_ctx, prec):
    """Graphs span six steps (the solvers' buffers alternate, the DEM's bucket
    tables rotate by three): the first call captures and replays one period
    and steps the rest eagerly; the second aligns to the period eagerly,
    replays it and ends eagerly."""
    m, d, w = scenes.hybrid_column(seed=3)
    a = GpuHybrid(m, d, w, prec, ctx=hg_ctx)
    b = GpuHybrid(m, d, w, prec, ctx=hg_ctx)
    a.step(DT, 20)
    b.step(DT, 7, graph=True)
    b.step(DT, 13, graph=True)
    # the grain side's node sums are FP64 atomics (reproducible to a few
    # ulps), which FP32 storage can round to a neighbouring value
    tol = 1e-12 if prec == HG_FP64 else 1e-5
    pa, pb = a.points(), b.points()
    for k in pa:
        assert_close(pb[k], pa[k], tol, k)
    qa, qb = a.bodies(), b.bodies()
    assert_close(qb["q"], qa["q"], tol, "q")
    assert_close(qb["v"], qa["v"], tol, "v")
    assert np.array_equal(a.coupled(), b.coupled())


def test_no_overlap_equals_separate_solvers_bitwise(hg_ctx):
    """Grains lifted clear of the continuum: nothing couples and the hybrid
    step reduces bit for bit to an MPM step and a weighted DEM step."""
    m, d, w = scenes.hybrid_column(seed=4)
    d.q[:, 1] += 0.1
    h = GpuHybrid(m, d, w, HG_FP64, ctx=hg_ctx)
    a = GpuMpm(m, HG_FP64, ctx=hg_ctx)
    b = GpuDem(d, HG_FP64, ctx=hg_ctx)
    b.set_weights(w)
    h.step(DT, 6)
    a.flow(DT, 6)
    b.flow(DT, 6)
    assert h.coupled().sum() == 0
    pa, ph = a.points(), h.points()
    for k in pa:
        assert np.array_equal(pa[k], ph[k]), k
    sb, sh = b.state(), h.bodies()
    assert np.array_equal(sb["q"], sh["q"]) and np.array_equal(sb["v"], sh["v"])


def test_coupled_free_fall_is_exact(hg_ctx):
    m, d, w = scenes.hybrid_column(seed=5)
    for s in (m, d):
        s.plane_x = np.zeros((0, 2))
        s.plane_n = np.zeros((0, 2))
    m.plane_lower = np.zeros(0)
    m.plane_behavior = np.zeros(0, np.int32)
    d.radius = d.radius * 0.5
    m.q[:, 1] += 0.05
    d.q[:, 1] += 0.05
    g = GpuHybrid(m, d, w, HG_FP64, ctx=hg_ctx)
    dt, n = 1e-4, 5
    g.step(dt, n)
    assert g.coupled().sum() > 20
    gv = np.asarray(m.gravity)
    assert np.abs(g.points()["v"] - n * dt * gv).max() < 1e-13
    assert np.abs(g.bodies()["v"][:, :2] - n * dt * gv).max() < 1e-13


def test_fp32_tracks_fp64(hg_ctx):
    m, d, w = scenes.hybrid_column(seed=6)
    g = GpuHybrid(m, d, w, HG_FP32, ctx=hg_ctx)
    o = oracle.OracleHybrid(m, d, w)
    g.step(DT, 20)
    o.step(DT, 20)
    c = g.coupled()
    assert (c == o.coupled).mean() > 0.99
    assert np.abs(g.points()["q"] - o.mpm.points()["q"]).max() < 1e-8
    assert np.abs(g.bodies()["q"][:, :2] - o.dem.state()["q"][:, :2]).max() < 1e-8


def _agitated(seed):
    m, d, w = scenes.hybrid_column(seed=seed)
    d.v[:, :2] = np.random.default_rng(seed).normal(scale=0.3, size=(d.nbodies, 2))
    return m, d, w


def test_homogenization_matches_oracle(hg_ctx):
    """Per-step homogenised stress and velocity (ResamplingTools.cpp:744-884)
    accumulated over 30 steps against the restatement; the node sums are FP64
    atomics (order varies by a few ulps)."""
    m, d, w = _agitated(7)
    g = GpuHybrid(m, d, w, HG_FP64, ctx=hg_ctx)
    g.set_homogenize(True, True)
    o = oracle.OracleHybrid(m, d, w, homogenize=True)
    g.step(DT, 30)
    o.step(DT, 30)
    assert len(o.dem.circle_contacts()["i"]) > 100
    gh, oh = g.homogenized(), o.homogenized()
    assert np.abs(oh["stress"]).max() > 0 and np.abs(oh["velocity"]).max() > 0
    assert_close(gh["stress"], oh["stress"], 1e-10, "homogenized stress")
    assert_close(gh["velocity"], oh["velocity"], 1e-10, "homogenized velocity")
    _compare(g, o, 1e-11, 1e-10)


def test_coupling_weights_match_oracle(hg_ctx):
    """PhysicsGrid::homog_factor as the node-node step saves it
    (HybridGrains2DSim.cpp:169-197): 1 with continuum mass, 0.5 with grain
    mass, 0 elsewhere.  Both sets depend only on which weights are nonzero,
    so the values match the restatement exactly, in eager and graph steps;
    recording them leaves the step unchanged."""
    from libhybridgrains_amd._lib import HgError
    m, d, w = _agitated(9)
    g = GpuHybrid(m, d, w, HG_FP64, ctx=hg_ctx)
    gg = GpuHybrid(m, d, w, HG_FP64, ctx=hg_ctx)
    c = GpuHybrid(m, d, w, HG_FP64, ctx=hg_ctx)
    o = oracle.OracleHybrid(m, d, w)
    with pytest.raises(HgError):
        g.homog_factor()
    g.set_homog_factor(True)
    gg.set_homog_factor(True)
    assert np.all(g.homog_factor() == 1.0)
    for n in (1, 5):
        g.step(DT, n)
        gg.step(DT, n, graph=True)
        c.step(DT, n)
        o.step(DT, n)
        hf = g.homog_factor()
        assert np.array_equal(hf, o.homog_factor)
        assert np.array_equal(gg.homog_factor(), o.homog_factor)
    assert (hf == 0.5).sum() > 50 and (hf == 1.0).sum() > 50 and (hf == 0.0).sum() > 50
    pg, pc = g.points(), c.points()
    for k in pg:
        assert_close(pg[k], pc[k], 1e-12, k)
    g.set_homog_factor(False)
    g.step(DT, 2)
    assert np.array_equal(g.homog_factor(), hf)


def test_homogenization_graph_clear_and_off(hg_ctx):
    m, d, w = _agitated(8)
    a = GpuHybrid(m, d, w, HG_FP64, ctx=hg_ctx)
    b = GpuHybrid(m, d, w, HG_FP64, ctx=hg_ctx)
    c = GpuHybrid(m, d, w, HG_FP64, ctx=hg_ctx)
    a.set_homogenize(True, True)
    b.set_homogenize(True, True)
    a.step(DT, 12)
    b.step(DT, 12, graph=True)
    c.step(DT, 12, graph=True)
    ha, hb = a.homogenized(), b.homogenized()
    for k in ha:
        assert np.abs(ha[k]).max() > 0
        assert_close(hb[k], ha[k], 1e-12, k)
    # homogenisation only reads the state: the step itself is unchanged
    pa, pc = a.points(), c.points()
    for k in pa:
        assert_close(pc[k], pa[k], 1e-12, k)
    hc = c.homogenized()
    assert not hc["stress"].any() and not hc["velocity"].any()
    b.clear_homogenized()
    hb = b.homogenized()
    assert not hb["stress"].any() and not hb["velocity"].any()
    b.step(DT, 4, graph=True)
    assert np.abs(b.homogenized()["velocity"]).max() > 0
