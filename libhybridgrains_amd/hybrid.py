"""The hybridgrains2dnew coupled step: one continuum (MPM) scene and one grain
(circle DEM) scene exchanging momentum on the MPM grid.

``GpuHybrid`` mirrors HybridGrains2DSim (hybridgrains2dnew/
HybridGrains2DSim.h:15-48): ``step`` is stepSystem with the
PREDICTION_CORRECTION_ADVECTION_NODE_NODE integrator style
(HybridGrains2DSim.cpp:116-252); ``discrete`` and ``continuum`` expose the two
states like discreteSim() / continuumState().
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import HG_FP64, ptr
from .dem import DemScene, GpuDem, _DemHold
from .mpm import GpuMpm, MpmScene, _DescHold


class GpuHybrid:
    def __init__(self, mpm_scene: MpmScene, dem_scene: DemScene, weights=None,
                 precision: int = HG_FP64, ctx=None, dem_cell_width: float = 0.0,
                 dem_domain=None):
        self.ctx = ctx or _lib.default_context()
        self.L = _lib.lib()
        self.mpm_scene, self.dem_scene = mpm_scene, dem_scene
        mh = _DescHold(mpm_scene, precision)
        dh = _DemHold(dem_scene, precision, dem_cell_width, dem_domain)
        w = None if weights is None else np.ascontiguousarray(weights, np.float64)
        if w is not None and w.shape != (dem_scene.nbodies,):
            raise ValueError("weights must have one entry per grain")
        h = C.c_void_p()
        self.ctx.check(self.L.hg_hybrid_create(self.ctx.handle, C.byref(mh.desc),
                                               C.byref(dh.desc), ptr(w), C.byref(h)))
        self._h = h
        hm, hd = C.c_void_p(), C.c_void_p()
        self.ctx.check(self.L.hg_hybrid_mpm(h, C.byref(hm)))
        self.ctx.check(self.L.hg_hybrid_dem(h, C.byref(hd)))
        self.continuum = GpuMpm._borrow(hm, [mpm_scene], precision, self.ctx)
        self.discrete = GpuDem._borrow(hd, [dem_scene], self.ctx)

    def step(self, dt: float, nsteps: int = 1, graph: bool = False) -> None:
        f = self.L.hg_hybrid_step_graph if graph else self.L.hg_hybrid_step
        self.ctx.check(f(self._h, float(dt), int(nsteps)))

    def sync(self) -> None:
        self.ctx.check(self.L.hg_hybrid_sync(self._h))

    def coupled(self) -> np.ndarray:
        c = np.zeros(self.dem_scene.nbodies, np.uint8)
        self.ctx.check(self.L.hg_hybrid_download_coupling(self._h, ptr(c, C.c_uint8)))
        return c

    # -- per-step homogenisation (ResamplingTools.cpp:744-884) -----------------
    def set_homogenize(self, stress: bool = True, velocity: bool = True) -> None:
        """Accumulate the grains' homogenised stress and velocity on the grid
        after every step, as updateHybridStates_poorman does when the scene
        sets homogenize_* with grid_smoothing_homogenized_*."""
        self.ctx.check(self.L.hg_hybrid_set_homogenize(self._h, int(bool(stress)),
                                                       int(bool(velocity))))

    def homogenized(self) -> dict:
        """The accumulators: stress (G x 4, column-major 2x2) and velocity
        (G x 2), nodes in the reference's flat order."""
        s = self.mpm_scene
        G = (int(s.cell_count[0]) + 1) * (int(s.cell_count[1]) + 1)
        st, ve = np.zeros((G, 4)), np.zeros((G, 2))
        self.ctx.check(self.L.hg_hybrid_download_homogenized(self._h, ptr(st), ptr(ve)))
        return {"stress": st, "velocity": ve}

    def set_homog_factor(self, on: bool = True) -> None:
        """Record the continuum grid's coupling weights in every later step,
        as HybridGrains2DSim's node-node step saves them
        (HybridGrains2DSim.cpp:169-197)."""
        self.ctx.check(self.L.hg_hybrid_set_homog_factor(self._h, int(bool(on))))

    def homog_factor(self) -> np.ndarray:
        """The last recorded coupling weights (G values in the reference's
        flat node order): 1 with continuum mass, 0.5 with grain mass, else 0."""
        s = self.mpm_scene
        G = (int(s.cell_count[0]) + 1) * (int(s.cell_count[1]) + 1)
        out = np.zeros(G)
        self.ctx.check(self.L.hg_hybrid_download_homog_factor(self._h, ptr(out)))
        return out

    def clear_homogenized(self) -> None:
        self.ctx.check(self.L.hg_hybrid_clear_homogenized(self._h))

    def snapshot(self) -> bytes:
        """Continuum, grains (with contact history) and the homogenisation
        accumulators as a blob (hg_hybrid_snapshot)."""
        return _lib.snapshot_blob(self.ctx, self.L.hg_hybrid_snapshot, self._h)

    def restore(self, blob: bytes) -> None:
        _lib.restore_blob(self.ctx, self.L.hg_hybrid_restore, self._h, blob)

    def points(self) -> dict:
        return self.continuum.points()

    def bodies(self) -> dict:
        return self.discrete.state()

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.continuum._h = None
            self.discrete._h = None
            self.L.hg_hybrid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
