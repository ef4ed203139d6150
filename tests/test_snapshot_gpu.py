"""Snapshot and resume of the device state (hg_*_snapshot / hg_*_restore).

A run that is snapshotted after k steps and resumed in a fresh handle made
from the same scene continues bit for bit like the original run (MPM in both
precisions, DEM with its contact history); the hybrid, whose grain-side node
sums are FP64 atomics, continues within 1e-12.  Blobs of another kind,
precision or size are refused.
"""
import numpy as np
import pytest

from libhybridgrains_amd import HG_FP32, HG_FP64, GpuDem, GpuHybrid, GpuMpm, HgError, scenes

from ._util import assert_close

pytestmark = pytest.mark.gpu

DT = 1e-5


def _mpm_scene(seed):
    s = scenes.column_collapse(domain=(0.4, 0.3), h=0.005, column=(0.1, 0.1), seed=seed)
    s.v = np.random.default_rng(seed).normal(scale=0.1, size=s.v.shape)
    return s


@pytest.mark.parametrize("prec", [HG_FP64, HG_FP32])
def test_mpm_resume_is_bitwise(hg_ctx, prec):
    s = _mpm_scene(1)
    a = GpuMpm(s, prec, ctx=hg_ctx)
    a.flow(DT, 13)
    blob = a.snapshot()
    a.flow(DT, 17, graph=True)
    b = GpuMpm(s, prec, ctx=hg_ctx)
    b.restore(blob)
    b.flow(DT, 17, graph=True)
    pa, pb = a.points(), b.points()
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k


def test_dem_resume_keeps_the_contact_history(hg_ctx):
    s = scenes.dem_column(n_bodies=2000, seed=2)
    s.v[:, :2] = np.random.default_rng(2).normal(scale=0.2, size=(s.nbodies, 2))
    a = GpuDem(s, HG_FP64, ctx=hg_ctx)
    a.flow(DT, 60)
    blob = a.snapshot()
    a.flow(DT, 25)
    b = GpuDem(s, HG_FP64, ctx=hg_ctx)
    b.restore(blob)
    b.flow(DT, 25)
    sa, sb = a.state(), b.state()
    assert np.array_equal(sa["q"], sb["q"]) and np.array_equal(sa["v"], sb["v"])
    ca, cb = a.contacts(), b.contacts()
    assert len(ca["i"]) > 300
    for k in ca:
        assert np.array_equal(ca[k], cb[k]), k
    # without the history the springs restart and the runs part
    c = GpuDem(s, HG_FP64, ctx=hg_ctx)
    c.flow(DT, 60)
    st = c.state()
    qv = s.copy()
    qv.q, qv.v = st["q"], st["v"]
    fresh = GpuDem(qv, HG_FP64, ctx=hg_ctx)
    fresh.flow(DT, 25)
    assert not np.array_equal(fresh.state()["v"], sa["v"])


def test_uploaded_contact_history_continues_the_run(hg_ctx):
    # hg_dem_upload_contacts: a handle made from another run's positions and
    # velocities, given that run's contact history, continues bit for bit
    s = scenes.dem_column(n_bodies=2000, seed=5)
    s.v[:, :2] = np.random.default_rng(5).normal(scale=0.2, size=(s.nbodies, 2))
    a = GpuDem(s, HG_FP64, ctx=hg_ctx)
    a.flow(DT, 60)
    st, ca = a.state(), a.contacts()
    assert len(ca["i"]) > 300 and (ca["type"] == 0).any() and (ca["type"] == 1).any()
    qv = s.copy()
    qv.q, qv.v = st["q"], st["v"]
    b = GpuDem(qv, HG_FP64, ctx=hg_ctx)
    b.upload_contacts(ca)
    cb = b.contacts()
    for k in ca:
        assert np.array_equal(ca[k], cb[k]), k
    a.flow(DT, 25)
    b.flow(DT, 25)
    sa, sb = a.state(), b.state()
    assert np.array_equal(sa["q"], sb["q"]) and np.array_equal(sa["v"], sb["v"])
    # the owners' forces are those of the contacts' last step
    f = a.contact_forces()
    assert f["force"].shape == (len(a.contacts()["i"]), 2)
    assert np.abs(f["force"][a.contacts()["type"] == 1]).max() > 0
    # ill-formed histories are refused
    bad = {k: v.copy() for k, v in ca.items()}
    bad["j"][np.nonzero(bad["type"] == 1)[0][0]] = -1
    with pytest.raises(HgError):
        b.upload_contacts(bad)


def test_hybrid_resume_with_homogenization(hg_ctx):
    m, d, w = scenes.hybrid_column(seed=3)
    d.v[:, :2] = np.random.default_rng(3).normal(scale=0.3, size=(d.nbodies, 2))
    a = GpuHybrid(m, d, w, HG_FP64, ctx=hg_ctx)
    a.set_homogenize(True, True)
    a.step(DT, 10)
    blob = a.snapshot()
    a.step(DT, 12, graph=True)
    b = GpuHybrid(m, d, w, HG_FP64, ctx=hg_ctx)
    b.restore(blob)
    b.step(DT, 12, graph=True)
    pa, pb = a.points(), b.points()
    for k in pa:
        assert_close(pb[k], pa[k], 1e-12, k)
    qa, qb = a.bodies(), b.bodies()
    assert_close(qb["q"], qa["q"], 1e-12, "q")
    assert_close(qb["v"], qa["v"], 1e-12, "v")
    ha, hb = a.homogenized(), b.homogenized()
    for k in ha:
        assert np.abs(ha[k]).max() > 0
        assert_close(hb[k], ha[k], 1e-12, k)


def test_foreign_blobs_are_refused(hg_ctx):
    s = _mpm_scene(4)
    a = GpuMpm(s, HG_FP64, ctx=hg_ctx)
    blob = a.snapshot()
    with pytest.raises(HgError):
        GpuMpm(s, HG_FP32, ctx=hg_ctx).restore(blob)  # precision
    with pytest.raises(HgError):
        GpuMpm(s, HG_FP64, ctx=hg_ctx).restore(blob[:-8])  # length
    d = scenes.dem_column(n_bodies=500, seed=4)
    with pytest.raises(HgError):
        GpuDem(d, HG_FP64, ctx=hg_ctx).restore(blob)  # kind
    bad = bytearray(blob)
    bad[0:8] = b"NOTASNAP"
    with pytest.raises(HgError):
        GpuMpm(s, HG_FP64, ctx=hg_ctx).restore(bytes(bad))
