"""Circle DEM scenes: the rigidbody2d RigidBody2DState (circles) and the
PenaltyImpactFrictionMap parameters, and their GPU stepping.

``DemScene`` holds what RigidBody2DSceneParser::parseXMLSceneFile
(rigidbody2dutils/RigidBody2DSceneParser.cpp:1127-1276) builds for a circle
scene; ``GpuDem`` keeps scenes resident on the GPU and steps them with
``PenaltyImpactFrictionMap::flow`` / ``flowWithWeights`` semantics
(rigidbody2d/PenaltyImpactFrictionMap.cpp:1083-1149).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import HG_FP64, DemDesc, ptr


@dataclass
class DemScene:
    """Host copy of one circle DEM scene.  ``q``/``v`` are (B, 3) rows of
    (x, y, theta); ``mass`` is (B, 3): the diagonal of M (m, m, I)."""

    q: np.ndarray
    v: np.ndarray
    radius: np.ndarray
    mass: np.ndarray
    fixed: np.ndarray
    gravity: np.ndarray
    kn: float
    kt: float
    gamman: float
    gammat: float
    mu: float
    plane_x: np.ndarray = field(default_factory=lambda: np.zeros((0, 2)))
    plane_n: np.ndarray = field(default_factory=lambda: np.zeros((0, 2)))

    @property
    def nbodies(self) -> int:
        return int(self.q.shape[0])

    def copy(self) -> "DemScene":
        kw = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in self.__dict__.items()}
        return DemScene(**kw)

    def add_plane(self, x, n) -> None:
        n = np.asarray(n, float)
        self.plane_x = np.vstack([self.plane_x, np.asarray(x, float)[None]])
        self.plane_n = np.vstack([self.plane_n, (n / np.linalg.norm(n))[None]])


def circle_mass(rho: float, r: np.ndarray) -> np.ndarray:
    """CircleGeometry::massAndInertia (rigidbody2d/CircleGeometry.cpp:44-48):
    m = rho pi r^2, I = m r^2 / 2; returned as (B, 3) rows (m, m, I)."""
    r = np.asarray(r, float)
    m = rho * np.pi * r * r
    I = 0.5 * m * r * r
    return np.stack([m, m, I], 1)


def _kernel_times(ctx, fn, h) -> dict:
    maxk = 32
    names = C.create_string_buffer(32 * maxk)
    ms = np.zeros(maxk)
    launches = np.zeros(maxk, np.int64)
    nk = C.c_int32()
    ctx.check(fn(h, maxk, names, ptr(ms), launches.ctypes.data_as(C.POINTER(C.c_int64)),
                 C.byref(nk)))
    out = {}
    raw = names.raw
    for i in range(min(nk.value, maxk)):
        nm = raw[32 * i: 32 * i + 32].split(b"\0", 1)[0].decode()
        out[nm] = (float(ms[i]), int(launches[i]))
    return out


class _DemHold:
    def __init__(self, s: DemScene, precision: int, cell_width=0.0, domain=None):
        self.arrs = {}

        def c(name, a, dt=np.float64):
            a = np.ascontiguousarray(a, dtype=dt)
            self.arrs[name] = a
            return a

        d = DemDesc()
        d.precision = precision
        d.nbodies = s.nbodies
        d.q = ptr(c("q", s.q.reshape(-1)))
        d.v = ptr(c("v", s.v.reshape(-1)))
        d.mass = ptr(c("m", s.mass.reshape(-1)))
        d.radius = ptr(c("r", s.radius))
        d.fixed = ptr(c("f", s.fixed, np.uint8), C.c_uint8)
        d.gravity[:] = [float(s.gravity[0]), float(s.gravity[1])]
        d.nplanes = int(s.plane_x.shape[0])
        d.plane_x = ptr(c("px", s.plane_x.reshape(-1)))
        d.plane_n = ptr(c("pn", s.plane_n.reshape(-1)))
        d.kn, d.kt, d.gamman, d.gammat, d.mu = (float(s.kn), float(s.kt), float(s.gamman),
                                                float(s.gammat), float(s.mu))
        d.cell_width = float(cell_width)
        if domain is not None:
            d.domain_min[:] = [float(domain[0][0]), float(domain[0][1])]
            d.domain_max[:] = [float(domain[1][0]), float(domain[1][1])]
        self.desc = d


class GpuDem:
    """One or more circle DEM scenes resident on a GPU."""

    def __init__(self, scenes, precision: int = HG_FP64, ctx=None, cell_width: float = 0.0,
                 domain=None):
        if isinstance(scenes, DemScene):
            scenes = [scenes]
        self.ctx = ctx or _lib.default_context()
        self.L = _lib.lib()
        self.scenes = list(scenes)
        holds = [_DemHold(s, precision, cell_width, domain) for s in self.scenes]
        arr = (DemDesc * len(holds))(*[h.desc for h in holds])
        h = C.c_void_p()
        self.ctx.check(self.L.hg_dem_create_batch(self.ctx.handle, arr, len(holds), C.byref(h)))
        self._h = h

    def flow(self, dt: float, nsteps: int = 1, graph: bool = False) -> None:
        f = self.L.hg_dem_step_graph if graph else self.L.hg_dem_step
        self.ctx.check(f(self._h, float(dt), int(nsteps)))

    def set_weights(self, weights, scene: int = 0) -> None:
        w = None if weights is None else np.ascontiguousarray(weights, np.float64)
        self.ctx.check(self.L.hg_dem_set_weights(self._h, scene, ptr(w)))

    def sync(self) -> None:
        self.ctx.check(self.L.hg_dem_sync(self._h))

    def state(self, scene: int = 0) -> dict:
        n = self.scenes[scene].nbodies
        q = np.zeros((n, 3))
        v = np.zeros((n, 3))
        self.ctx.check(self.L.hg_dem_download_bodies(self._h, scene, ptr(q), ptr(v)))
        return {"q": q, "v": v}

    def contacts(self, scene: int = 0) -> dict:
        cnt = C.c_uint32()
        self.ctx.check(self.L.hg_dem_download_contacts(self._h, scene, 0, None, None, None, None,
                                                       C.byref(cnt)))
        n = cnt.value
        t = np.zeros(n, np.int32)
        i = np.zeros(n, np.int32)
        j = np.zeros(n, np.int32)
        ds = np.zeros((n, 2))
        self.ctx.check(self.L.hg_dem_download_contacts(
            self._h, scene, n, ptr(t, C.c_int32), ptr(i, C.c_int32), ptr(j, C.c_int32), ptr(ds),
            C.byref(cnt)))
        return {"type": t, "i": i, "j": j, "delta_s": ds}

    def upload_contacts(self, c: dict, scene: int = 0) -> None:
        """Replace the contact history with contacts in contacts()' form
        (hg_dem_upload_contacts)."""
        t = np.ascontiguousarray(c["type"], np.int32)
        i = np.ascontiguousarray(c["i"], np.int32)
        j = np.ascontiguousarray(c["j"], np.int32)
        ds = np.ascontiguousarray(c["delta_s"], np.float64).reshape(-1, 2)
        self.ctx.check(self.L.hg_dem_upload_contacts(self._h, scene, len(t), ptr(t, C.c_int32),
                                                      ptr(i, C.c_int32), ptr(j, C.c_int32),
                                                      ptr(ds)))

    def contact_forces(self, scene: int = 0) -> dict:
        """The last step's force on each contact's owner and its contact
        point, in contacts()' order (hg_dem_download_contact_forces)."""
        cnt = C.c_uint32()
        self.ctx.check(self.L.hg_dem_download_contact_forces(self._h, scene, 0, None, None,
                                                             C.byref(cnt)))
        n = cnt.value
        f = np.zeros((n, 2))
        p = np.zeros((n, 2))
        self.ctx.check(self.L.hg_dem_download_contact_forces(self._h, scene, n, ptr(f), ptr(p),
                                                             C.byref(cnt)))
        return {"force": f, "point": p}

    def keep_force(self, keep: bool) -> None:
        self.ctx.check(self.L.hg_dem_keep_force(self._h, int(bool(keep))))

    def force(self, scene: int = 0) -> np.ndarray:
        F = np.zeros((self.scenes[scene].nbodies, 3))
        self.ctx.check(self.L.hg_dem_download_force(self._h, scene, ptr(F)))
        return F

    def set_profiling(self, on: bool) -> None:
        self.ctx.check(self.L.hg_dem_set_profiling(self._h, int(bool(on))))

    def snapshot(self) -> bytes:
        """Bodies and contact history as a blob (hg_dem_snapshot)."""
        return _lib.snapshot_blob(self.ctx, self.L.hg_dem_snapshot, self._h)

    def restore(self, blob: bytes) -> None:
        _lib.restore_blob(self.ctx, self.L.hg_dem_restore, self._h, blob)

    def hold(self, us: float) -> None:
        """As GpuMpm.hold."""
        self.ctx.check(self.L.hg_dem_hold(self._h, float(us)))

    def observables(self, x0: float = 0.0, bin_width: float = 1.0, nbins: int = 0,
                    scene: int = 0) -> dict:
        """Kinetic energy, front (largest x), top (largest y) and the height
        profile over nbins x-bins from x0 (NaN where a bin is empty), computed
        on the device from the resident state."""
        out = np.zeros(3 + int(nbins))
        self.ctx.check(self.L.hg_dem_observables(self._h, scene, float(x0), float(bin_width), int(nbins),
                                              ptr(out)))
        return {"ke": float(out[0]), "front": float(out[1]), "top": float(out[2]),
                "profile": out[3:].copy()}

    def kernel_times(self) -> dict:
        return _kernel_times(self.ctx, self.L.hg_dem_kernel_times, self._h)

    @classmethod
    def _borrow(cls, handle, scenes, ctx) -> "GpuDem":
        """A view of a handle another object owns (the hybrid's grains)."""
        o = cls.__new__(cls)
        o.ctx, o.L, o.scenes = ctx, _lib.lib(), list(scenes)
        o._h = handle
        o._borrowed = True
        return o

    def close(self) -> None:
        if getattr(self, "_h", None):
            if not getattr(self, "_borrowed", False):
                self.L.hg_dem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
