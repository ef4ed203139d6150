"""GPU parity of the circle DEM step against the CPU FP64 restatement.

FP64 path: the pair set is bit-exact (same FP64 predicate), and because the
engine gathers each body's forces in the reference's summation order, body
states and springs agree to the last bit for as long as the runs are compared
here (checked with a 1e-13 relative bound to be robust to libm differences).
FP32 path: FP64 positions and pair tests, FP32 velocities and forces; the
pair set over the first steps is identical, trajectories agree within the
stated tolerance.
"""
import numpy as np
import pytest

import oracle
from libhybridgrains_amd import HG_FP32, HG_FP64, GpuDem, scenes

from ._util import assert_close, rel_err

pytestmark = pytest.mark.gpu

DT = 1e-5


def _scene(n=3000, seed=0, fixed_every=0):
    s = scenes.dem_column(n_bodies=n, seed=seed)
    rng = np.random.default_rng(seed)
    s.v[:, :2] = rng.normal(scale=0.05, size=(n, 2))
    s.v[:, 2] = rng.normal(scale=5.0, size=n)
    if fixed_every:
        s.fixed[::fixed_every] = 1
    return s


def _pairset(c):
    return {(int(t), int(i), int(j)) for t, i, j in zip(c["type"], c["i"], c["j"])}


def test_pair_set_bit_exact_against_bruteforce(hg_ctx):
    s = _scene(n=2500, seed=1, fixed_every=29)
    s.q[:, :2] += np.random.default_rng(1).normal(scale=1.5e-4, size=(s.nbodies, 2))
    o = oracle.OracleDem(s)
    bf = {tuple(p) for p in o.pairs_bruteforce().tolist()}
    g = GpuDem(s, HG_FP64, ctx=hg_ctx)
    g.flow(DT, 1)
    c = g.contacts()
    got = {(min(i, j), max(i, j)) for t, i, j in zip(c["type"], c["i"], c["j"]) if t in (1, 2)}
    assert got == bf
    assert len(got) > 1000


def test_wrapped_table_keeps_pairs_and_results(hg_ctx):
    # a table of 4 x 4 buckets over a scene hundreds of cells wide: cells a
    # table width apart share buckets; the pair set is still the brute-force
    # one, and 20 steps agree bit for bit with the default table
    s = _scene(n=1500, seed=7, fixed_every=31)
    cw = 2.0 * float(s.radius.max()) * 1.01
    x0, y0 = float(s.q[:, 0].min()), float(s.q[:, 1].min())
    tiny = GpuDem(s, HG_FP64, ctx=hg_ctx, cell_width=cw, domain=((x0, y0), (x0 + cw, y0 + cw)))
    wide = GpuDem(s, HG_FP64, ctx=hg_ctx, cell_width=cw)
    tiny.flow(DT, 1)
    bf = {tuple(p) for p in oracle.OracleDem(s).pairs_bruteforce().tolist()}
    c = tiny.contacts()
    assert {(min(i, j), max(i, j)) for t, i, j in zip(c["type"], c["i"], c["j"]) if t in (1, 2)} == bf
    wide.flow(DT, 1)
    tiny.flow(DT, 19)
    wide.flow(DT, 19)
    a, b = tiny.state(), wide.state()
    assert np.array_equal(a["q"], b["q"]) and np.array_equal(a["v"], b["v"])
    assert _pairset(tiny.contacts()) == _pairset(wide.contacts())


def test_one_step_fp64_matches_oracle(hg_ctx):
    s = _scene(seed=2, fixed_every=41)
    g = GpuDem(s, HG_FP64, ctx=hg_ctx)
    g.keep_force(True)
    o = oracle.OracleDem(s)
    g.flow(DT, 1)
    o.flow(DT, 1)
    assert _pairset(g.contacts()) == _pairset(o.contacts())
    assert_close(g.force(), o.last_force(), 1e-13, "force")
    gs, os_ = g.state(), o.state()
    assert_close(gs["q"], os_["q"], 1e-13, "q")
    assert_close(gs["v"], os_["v"], 1e-13, "v")


def test_hundred_steps_fp64_matches_oracle(hg_ctx):
    s = _scene(seed=3, fixed_every=53)
    g = GpuDem(s, HG_FP64, ctx=hg_ctx)
    o = oracle.OracleDem(s)
    g.flow(DT, 100)
    o.flow(DT, 100)
    gc, oc = g.contacts(), o.contacts()
    assert _pairset(gc) == _pairset(oc)
    assert_close(gc["delta_s"], oc["delta_s"], 1e-11, "delta_s")
    gs, os_ = g.state(), o.state()
    assert_close(gs["q"], os_["q"], 1e-12, "q")
    assert_close(gs["v"], os_["v"], 1e-10, "v")


def test_weighted_flow_matches_oracle(hg_ctx):
    s = _scene(n=2000, seed=4)
    w = np.random.default_rng(4).uniform(0.2, 1.0, s.nbodies)
    g = GpuDem(s, HG_FP64, ctx=hg_ctx)
    g.set_weights(w)
    o = oracle.OracleDem(s, weights=w)
    g.flow(DT, 30)
    o.flow(DT, 30)
    gs, os_ = g.state(), o.state()
    assert _pairset(g.contacts()) == _pairset(o.contacts())
    assert_close(gs["v"], os_["v"], 1e-10, "v")


def test_graph_and_batch_are_bitwise_invariant(hg_ctx):
    ss = [_scene(n=1500, seed=10 + k) for k in range(3)]
    batch = GpuDem(ss, HG_FP64, ctx=hg_ctx)
    batch.flow(DT, 9, graph=True)
    for k, s in enumerate(ss):
        alone = GpuDem(s, HG_FP64, ctx=hg_ctx)
        alone.flow(DT, 9)
        a, b = alone.state(), batch.state(k)
        assert np.array_equal(a["q"], b["q"]) and np.array_equal(a["v"], b["v"]), k


@pytest.mark.parametrize("prec", [HG_FP64, HG_FP32])
def test_graph_periods_match_eager_steps_bitwise(hg_ctx, prec):
    """A step is one launch that reads one state buffer and one bucket table
    and writes the others, so graphs span six steps (buffers alternate,
    tables rotate by three).  Calls that start at any phase, capture, replay
    and end between periods give the eager steps' results bit for bit."""
    s = _scene(n=1500, seed=21)
    a = GpuDem(s, prec, ctx=hg_ctx)
    b = GpuDem(s, prec, ctx=hg_ctx)
    a.flow(DT, 31)
    for n in (4, 9, 18):
        b.flow(DT, n, graph=True)
    sa, sb = a.state(), b.state()
    assert np.array_equal(sa["q"], sb["q"]) and np.array_equal(sa["v"], sb["v"])
    assert _pairset(a.contacts()) == _pairset(b.contacts())


def test_fp32_tracks_fp64(hg_ctx):
    s = _scene(seed=5)
    g32 = GpuDem(s, HG_FP32, ctx=hg_ctx)
    o = oracle.OracleDem(s)
    g32.flow(DT, 1)
    o.flow(DT, 1)
    assert _pairset(g32.contacts()) == _pairset(o.contacts())
    g32.flow(DT, 99)
    o.flow(DT, 99)
    a, b = g32.state(), o.state()
    assert np.abs(a["q"][:, :2] - b["q"][:, :2]).max() < 1e-8  # metres
    assert rel_err(a["v"][:, :2], b["v"][:, :2]) < 1e-3
