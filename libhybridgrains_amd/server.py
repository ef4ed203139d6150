"""Scene server: many MPM, DEM and hybrid scenes resident on one GPU, admitted
by device-memory footprint and stepped side by side (continuous batching),
over the hg_server C-ABI (include/hg_server.h)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import HG_FP64, ServerConfig, ServerStats, ptr
from .dem import DemScene, _DemHold
from .mpm import MpmScene, _DescHold

QUEUED, RUNNING, DONE, FAILED, CANCELLED = range(5)
STATE_NAMES = {QUEUED: "queued", RUNNING: "running", DONE: "done", FAILED: "failed",
               CANCELLED: "cancelled"}


@dataclass
class Request:
    id: int
    kind: str
    npoints: int
    nbodies: int


class SceneServer:
    def __init__(self, ctx=None, budget_bytes: int = 0, max_active: int = 0,
                 slice_steps: int = 0, max_inflight: int = 0, max_bypass: int = 0):
        self.ctx = ctx or _lib.default_context()
        self.L = _lib.lib()
        cfg = ServerConfig(budget_bytes, max_active, slice_steps, max_inflight, max_bypass)
        h = C.c_void_p()
        self.ctx.check(self.L.hg_server_create(self.ctx.handle, C.byref(cfg), C.byref(h)))
        self._h = h
        self.requests: dict[int, Request] = {}

    # -- submission -------------------------------------------------------------
    def submit_mpm(self, scene: MpmScene, dt: float, nsteps: int,
                   precision: int = HG_FP64) -> int:
        hold = _DescHold(scene, precision)
        rid = C.c_uint64()
        self.ctx.check(self.L.hg_server_submit_mpm(self._h, C.byref(hold.desc), float(dt),
                                                   int(nsteps), C.byref(rid)))
        self.requests[rid.value] = Request(rid.value, "mpm", scene.npoints, 0)
        return rid.value

    def submit_dem(self, scene: DemScene, dt: float, nsteps: int, weights=None,
                   precision: int = HG_FP64) -> int:
        hold = _DemHold(scene, precision)
        w = None if weights is None else np.ascontiguousarray(weights, np.float64)
        rid = C.c_uint64()
        self.ctx.check(self.L.hg_server_submit_dem(self._h, C.byref(hold.desc), ptr(w), float(dt),
                                                   int(nsteps), C.byref(rid)))
        self.requests[rid.value] = Request(rid.value, "dem", 0, scene.nbodies)
        return rid.value

    def submit_hybrid(self, mpm: MpmScene, dem: DemScene, dt: float, nsteps: int, weights=None,
                      precision: int = HG_FP64) -> int:
        mh = _DescHold(mpm, precision)
        dh = _DemHold(dem, precision)
        w = None if weights is None else np.ascontiguousarray(weights, np.float64)
        rid = C.c_uint64()
        self.ctx.check(self.L.hg_server_submit_hybrid(
            self._h, C.byref(mh.desc), C.byref(dh.desc), ptr(w), float(dt), int(nsteps),
            C.byref(rid)))
        self.requests[rid.value] = Request(rid.value, "hybrid", mpm.npoints, dem.nbodies)
        return rid.value

    # -- control ----------------------------------------------------------------
    def cancel(self, rid: int) -> None:
        self.ctx.check(self.L.hg_server_cancel(self._h, rid))

    def poll(self, rid: int) -> tuple[int, int]:
        st, done = C.c_int32(), C.c_uint64()
        self.ctx.check(self.L.hg_server_poll(self._h, rid, C.byref(st), C.byref(done)))
        return st.value, done.value

    def wait(self, rid: int, timeout: float = 60.0) -> int:
        st = C.c_int32()
        self.ctx.check(self.L.hg_server_wait(self._h, rid, float(timeout), C.byref(st)))
        return st.value

    def drain(self, timeout: float = 120.0) -> None:
        self.ctx.check(self.L.hg_server_drain(self._h, float(timeout)))

    # -- results ----------------------------------------------------------------
    def points(self, rid: int) -> dict:
        n = self.requests[rid].npoints
        q, v, b, J = np.zeros((n, 2)), np.zeros((n, 2)), np.zeros((n, 4)), np.zeros(n)
        self.ctx.check(self.L.hg_server_result_points(self._h, rid, ptr(q), ptr(v), ptr(b),
                                                      ptr(J)))
        return {"q": q, "v": v, "be_bar": b, "J": J}

    def bodies(self, rid: int) -> dict:
        n = self.requests[rid].nbodies
        q, v = np.zeros((n, 3)), np.zeros((n, 3))
        self.ctx.check(self.L.hg_server_result_bodies(self._h, rid, ptr(q), ptr(v)))
        return {"q": q, "v": v}

    def latency(self, rid: int) -> float:
        t = C.c_double()
        self.ctx.check(self.L.hg_server_latency(self._h, rid, C.byref(t)))
        return t.value

    def error(self, rid: int) -> str:
        buf = C.create_string_buffer(1024)
        self.ctx.check(self.L.hg_server_error(self._h, rid, buf, len(buf)))
        return buf.value.decode(errors="replace")

    def release(self, rid: int) -> None:
        self.ctx.check(self.L.hg_server_release(self._h, rid))

    def stats(self) -> dict:
        s = ServerStats()
        self.ctx.check(self.L.hg_server_stats_get(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.L.hg_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def estimate_bytes(kind: str, mpm: MpmScene | None = None, dem: DemScene | None = None,
                   precision: int = HG_FP64) -> int:
    """The engine's device-memory bound for one scene (hg_*_estimate_bytes)."""
    L = _lib.lib()
    out = C.c_uint64()
    if kind == "mpm":
        h = _DescHold(mpm, precision)
        rc = L.hg_mpm_estimate_bytes(C.byref(h.desc), 1, C.byref(out))
    elif kind == "dem":
        h = _DemHold(dem, precision)
        rc = L.hg_dem_estimate_bytes(C.byref(h.desc), 1, C.byref(out))
    else:
        a, b = _DescHold(mpm, precision), _DemHold(dem, precision)
        rc = L.hg_hybrid_estimate_bytes(C.byref(a.desc), C.byref(b.desc), C.byref(out))
    if rc != 0:
        raise ValueError("estimate failed")
    return out.value
