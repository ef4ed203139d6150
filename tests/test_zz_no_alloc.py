"""The allocation-free observables entry point (hg_mpm_observables_noalloc)
against hg_mpm_observables, through the C-ABI: bitwise equal results on two
calls in a row, the second on a smaller scene with fewer bins (the static
scratch keeps the first call's contents, which must not leak into the
second), in both precisions; more bins than the static scratch holds are
refused with HG_ERR_INVALID."""
import ctypes as C

import numpy as np
import pytest

from libhybridgrains_amd import HG_FP32, HG_FP64, GpuMpm, HgError, scenes
from libhybridgrains_amd import _lib

pytestmark = pytest.mark.gpu


def _call(fn, g, x0, w, nbins):
    out = np.zeros(3 + nbins)
    g.ctx.check(fn(g._h, 0, float(x0), float(w), int(nbins), _lib.ptr(out)))
    return out


@pytest.mark.parametrize("precision", [HG_FP64, HG_FP32])
def test_noalloc_observables_equal_the_allocating_ones(hg_ctx, precision):
    L = _lib.lib()
    big = GpuMpm(scenes.mpm_config0(seed=21), precision, ctx=hg_ctx)
    big.flow(1e-5, 200)
    small = GpuMpm(scenes.column_collapse(domain=(0.4, 0.3), h=0.005, column=(0.06, 0.05),
                                          seed=22), precision, ctx=hg_ctx)
    small.flow(1e-5, 150)
    calls = [(big, 0.0, 0.005, 200), (small, 0.01, 0.01, 12)]
    for g, x0, w, nb in calls:
        na = _call(L.hg_mpm_observables_noalloc, g, x0, w, nb)
        al = _call(L.hg_mpm_observables, g, x0, w, nb)
        assert na[0] > 0.0 and np.isfinite(na[1]) and np.isfinite(na[2])
        assert np.array_equal(na, al, equal_nan=True), (nb, na[:3], al[:3])
        # bins past the smaller scene's points stay empty: nothing left over
        p = g.points()["q"]
        k = np.floor((p[:, 0] - x0) / w)
        filled = np.zeros(nb, bool)
        filled[np.unique(k[(k >= 0) & (k < nb)]).astype(int)] = True
        assert np.array_equal(~np.isnan(na[3:]), filled)
    # beyond the static scratch: refused, and nothing written
    out = np.full(3 + 8193, 7.0)
    rc = L.hg_mpm_observables_noalloc(big._h, 0, 0.0, 0.001, 8193, _lib.ptr(out))
    assert rc == _lib.HG_ERR_INVALID
    assert "8192" in L.hg_last_error(hg_ctx.handle).decode()
    assert np.all(out == 7.0)
    with pytest.raises(HgError):
        hg_ctx.check(rc)
