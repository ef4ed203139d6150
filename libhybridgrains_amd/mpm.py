"""MPM scenes on the GPU: the mpmgrains2d SimulationState / ExplicitIntegrator
surface over the hg C-ABI.

``MpmScene`` holds what ``InitialSimulationState::generateSimulationState``
(mpmgrains2d/InitialSimulationState.cpp:147-228) produces; ``GpuMpm`` keeps
one or more scenes resident on the GPU and steps them with
``ExplicitIntegrator::flow`` semantics (mpmgrains2d/ExplicitIntegrator.cpp:63-67).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import HG_BASIS_UGIMP_LINEAR, HG_FP64, HG_PLANE_SLIDING, MpmDesc, ptr


@dataclass
class MpmScene:
    """Host copy of one MPM scene (reference field names, SI or CGS as given).

    Layouts follow the reference's Eigen storage: ``q`` and ``v`` are (N, 2),
    ``be_bar`` is (N, 4) column-major (b00, b10, b01, b11).
    """

    grid_min: np.ndarray
    cell_width: float
    cell_count: np.ndarray
    shear_modulus: float
    bulk_modulus: float
    dp_alpha: float
    pic_alpha: float
    gravity: np.ndarray
    q: np.ndarray
    v: np.ndarray
    m: np.ndarray
    volume: np.ndarray
    hl: np.ndarray
    be_bar: np.ndarray
    J: np.ndarray
    plane_x: np.ndarray = field(default_factory=lambda: np.zeros((0, 2)))
    plane_n: np.ndarray = field(default_factory=lambda: np.zeros((0, 2)))
    plane_lower: np.ndarray = field(default_factory=lambda: np.zeros(0))
    plane_behavior: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    basis: int = HG_BASIS_UGIMP_LINEAR
    kinematically_scripted: np.ndarray | None = None

    @property
    def npoints(self) -> int:
        return int(self.q.shape[0])

    @property
    def nnodes(self) -> int:
        return int((self.cell_count[0] + 1) * (self.cell_count[1] + 1))

    def copy(self) -> "MpmScene":
        kw = {}
        for k, v in self.__dict__.items():
            kw[k] = v.copy() if isinstance(v, np.ndarray) else v
        return MpmScene(**kw)

    def add_plane(self, x, n, behavior=HG_PLANE_SLIDING, lower=-np.inf) -> None:
        n = np.asarray(n, float)
        self.plane_x = np.vstack([self.plane_x, np.asarray(x, float)[None]])
        self.plane_n = np.vstack([self.plane_n, (n / np.linalg.norm(n))[None]])
        self.plane_lower = np.append(self.plane_lower, float(lower))
        self.plane_behavior = np.append(self.plane_behavior, np.int32(behavior)).astype(np.int32)


class _DescHold:
    """Keeps the numpy buffers an MpmDesc points into alive."""

    def __init__(self, s: MpmScene, precision: int):
        self.arrs = {}

        def c(name, a, dt=np.float64):
            a = np.ascontiguousarray(a, dtype=dt)
            self.arrs[name] = a
            return a

        d = MpmDesc()
        d.precision = precision
        d.basis = int(s.basis)
        d.grid_min[:] = [float(s.grid_min[0]), float(s.grid_min[1])]
        d.cell_width = float(s.cell_width)
        d.cell_count[:] = [int(s.cell_count[0]), int(s.cell_count[1])]
        d.shear_modulus = float(s.shear_modulus)
        d.bulk_modulus = float(s.bulk_modulus)
        d.dp_alpha = float(s.dp_alpha)
        d.pic_alpha = float(s.pic_alpha)
        d.gravity[:] = [float(s.gravity[0]), float(s.gravity[1])]
        d.nplanes = len(s.plane_lower)
        d.plane_x = ptr(c("px", s.plane_x.reshape(-1)))
        d.plane_n = ptr(c("pn", s.plane_n.reshape(-1)))
        d.plane_lower = ptr(c("pl", s.plane_lower))
        d.plane_behavior = ptr(c("pb", s.plane_behavior, np.int32), C.c_int32)
        d.npoints = s.npoints
        d.q = ptr(c("q", s.q.reshape(-1)))
        d.v = ptr(c("v", s.v.reshape(-1)))
        d.m = ptr(c("m", s.m))
        d.volume = ptr(c("vol", s.volume))
        d.hl = ptr(c("hl", s.hl))
        d.be_bar = ptr(c("b", s.be_bar.reshape(-1)))
        d.J = ptr(c("J", s.J))
        if s.kinematically_scripted is not None:
            d.kinematically_scripted = ptr(c("ks", s.kinematically_scripted, np.uint8), C.c_uint8)
        else:
            d.kinematically_scripted = None
        self.desc = d


class GpuMpm:
    """One or more MPM scenes resident on a GPU (a batch of one by default)."""

    def __init__(self, scenes, precision: int = HG_FP64, ctx: _lib.Context | None = None,
                 tile: tuple[int, int] | None = None):
        if isinstance(scenes, MpmScene):
            scenes = [scenes]
        self.ctx = ctx or _lib.default_context()
        self.L = _lib.lib()
        self.scenes = list(scenes)
        self.precision = precision
        holds = [_DescHold(s, precision) for s in self.scenes]
        arr = (MpmDesc * len(holds))(*[h.desc for h in holds])
        h = C.c_void_p()
        self.ctx.check(self.L.hg_mpm_create_batch(self.ctx.handle, arr, len(holds), C.byref(h)))
        self._h = h
        if tile is not None:
            self.ctx.check(self.L.hg_mpm_set_tile(self._h, int(tile[0]), int(tile[1])))

    # -- stepping (ExplicitIntegrator) -----------------------------------------
    def flow(self, dt: float, nsteps: int = 1, graph: bool = False) -> None:
        f = self.L.hg_mpm_step_graph if graph else self.L.hg_mpm_step
        self.ctx.check(f(self._h, float(dt), int(nsteps)))

    def flow_zero_phase(self, dt: float = 0.0) -> None:
        self.ctx.check(self.L.hg_mpm_phase(self._h, 0, float(dt)))

    def flow_first_phase(self, dt: float) -> None:
        self.ctx.check(self.L.hg_mpm_phase(self._h, 1, float(dt)))

    def flow_second_phase(self, dt: float) -> None:
        self.ctx.check(self.L.hg_mpm_phase(self._h, 2, float(dt)))

    def sync(self) -> None:
        self.ctx.check(self.L.hg_mpm_sync(self._h))

    def autotune(self, dt: float, reps: int = 5) -> dict:
        """Picks the fastest tile shape (and, in FP32, points per tile pass)
        for these scenes (results do not depend on them); returns the choice
        and the timings."""
        aw, ah = C.c_int32(), C.c_int32()
        us = np.zeros(4)
        self.ctx.check(self.L.hg_mpm_autotune(self._h, float(dt), int(reps), C.byref(aw),
                                              C.byref(ah), ptr(us)))
        return {"tile": (aw.value, ah.value), "points_per_pass": self.tuning()["points_per_pass"],
                "us": {"16x16": us[0], "16x8": us[1], "32x8": us[2], "64x4": us[3]}}

    def set_points_per_pass(self, n: int) -> None:
        """FP32: points per cell per tile pass of the fused step (1 or 2)."""
        self.ctx.check(self.L.hg_mpm_set_points_per_pass(self._h, int(n)))

    def tuning(self) -> dict:
        """The tile shape and points per tile pass (threads per cell) in use."""
        aw, ah, pp = C.c_int32(), C.c_int32(), C.c_int32()
        self.ctx.check(self.L.hg_mpm_tuning(self._h, C.byref(aw), C.byref(ah), C.byref(pp)))
        return {"tile": (aw.value, ah.value), "points_per_pass": pp.value}

    def scene_error(self, scene: int) -> int:
        f = C.c_int32()
        self.ctx.check(self.L.hg_mpm_scene_error(self._h, scene, C.byref(f)))
        return int(f.value)

    # -- state ------------------------------------------------------------------
    def points(self, scene: int = 0) -> dict:
        n = self.scenes[scene].npoints
        q = np.zeros((n, 2))
        v = np.zeros((n, 2))
        b = np.zeros((n, 4))
        J = np.zeros(n)
        self.ctx.check(self.L.hg_mpm_download_points_scene(
            self._h, scene, ptr(q), ptr(v), ptr(b), ptr(J)))
        return {"q": q, "v": v, "be_bar": b, "J": J}

    def set_keep_grid(self, keep: bool) -> None:
        self.ctx.check(self.L.hg_mpm_set_keep_grid(self._h, int(bool(keep))))

    def grid(self, scene: int = 0) -> dict:
        G = self.scenes[scene].nnodes
        out = {"mass": np.zeros(G)}
        for k in ("mom", "mom_new", "force", "vel", "acc"):
            out[k] = np.zeros((G, 2))
        self.ctx.check(self.L.hg_mpm_download_grid_scene(
            self._h, scene, ptr(out["mass"]), ptr(out["mom"]), ptr(out["mom_new"]),
            ptr(out["force"]), ptr(out["vel"]), ptr(out["acc"])))
        return out

    # -- timing -----------------------------------------------------------------
    def set_profiling(self, on: bool) -> None:
        self.ctx.check(self.L.hg_mpm_set_profiling(self._h, int(bool(on))))

    def snapshot(self) -> bytes:
        """The evolving device state as a blob (hg_mpm_snapshot)."""
        return _lib.snapshot_blob(self.ctx, self.L.hg_mpm_snapshot, self._h)

    def restore(self, blob: bytes) -> None:
        """Loads a snapshot of a handle made from the same scenes."""
        _lib.restore_blob(self.ctx, self.L.hg_mpm_restore, self._h, blob)

    def hold(self, us: float) -> None:
        """Keeps the stream busy for about us microseconds, so that the steps
        submitted next run back to back (clean per-kernel event times)."""
        self.ctx.check(self.L.hg_mpm_hold(self._h, float(us)))

    def observables(self, x0: float = 0.0, bin_width: float = 1.0, nbins: int = 0,
                    scene: int = 0) -> dict:
        """Kinetic energy, front (largest x), top (largest y) and the height
        profile over nbins x-bins from x0 (NaN where a bin is empty), computed
        on the device from the resident state."""
        out = np.zeros(3 + int(nbins))
        self.ctx.check(self.L.hg_mpm_observables(self._h, scene, float(x0), float(bin_width), int(nbins),
                                              ptr(out)))
        return {"ke": float(out[0]), "front": float(out[1]), "top": float(out[2]),
                "profile": out[3:].copy()}

    def kernel_times(self) -> dict:
        maxk = 32
        names = C.create_string_buffer(32 * maxk)
        ms = np.zeros(maxk)
        launches = np.zeros(maxk, np.int64)
        nk = C.c_int32()
        self.ctx.check(self.L.hg_mpm_kernel_times(
            self._h, maxk, names, ptr(ms), launches.ctypes.data_as(C.POINTER(C.c_int64)),
            C.byref(nk)))
        out = {}
        raw = names.raw
        for i in range(min(nk.value, maxk)):
            nm = raw[32 * i: 32 * i + 32].split(b"\0", 1)[0].decode()
            out[nm] = (float(ms[i]), int(launches[i]))
        return out

    @classmethod
    def _borrow(cls, handle, scenes, precision, ctx) -> "GpuMpm":
        """A view of a handle another object owns (the hybrid's continuum)."""
        o = cls.__new__(cls)
        o.ctx, o.L, o.scenes, o.precision = ctx, _lib.lib(), list(scenes), precision
        o._h = handle
        o._borrowed = True
        return o

    def close(self) -> None:
        if getattr(self, "_h", None):
            if not getattr(self, "_borrowed", False):
                self.L.hg_mpm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
