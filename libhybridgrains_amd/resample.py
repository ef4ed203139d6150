"""The hybrid zone resampling's host pieces (libhg_host.so, hg_resample.hpp):
the 1998 MT19937 stream of random.c, the circle-rectangle intersection area,
ZoneTools on given circles and points, and resampling passes on a hybrid
scene with no device work (the pass HybridGrains2DSim runs between device
steps; hybridgrains2dnew/HybridGrains2DSim.cpp:301-559).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import HgError
from .xmlscene import host_lib

ZT_CONTINUUM, ZT_DISCRETE, ZT_HYBRID = 0, 1, 2

_ready = False


def _lib():
    global _ready
    L = host_lib()
    if _ready:
        return L
    vp, dp = C.c_void_p, C.POINTER(C.c_double)
    u32p, i32p = C.POINTER(C.c_uint32), C.POINTER(C.c_int32)
    sigs = {
        "hgx_mt1998": ([C.c_uint32, C.c_uint32, u32p, C.POINTER(C.c_float)], C.c_int32),
        "hgx_circle_rect_area": ([C.c_double] * 7, C.c_double),
        "hgx_zone_tools": ([dp, C.c_double, u32p, C.c_double, C.c_uint32, dp, dp, C.c_uint32, dp,
                            C.c_double, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int32,
                            dp, dp, i32p, i32p], C.c_int32),
        "hgx_resample_create": ([vp, C.POINTER(vp)], C.c_int32),
        "hgx_resample_run": ([vp, C.POINTER(C.c_int64)], C.c_int32),
        "hgx_resample_counts": ([vp, u32p, u32p], C.c_int32),
        "hgx_resample_points": ([vp] + [dp] * 10, C.c_int32),
        "hgx_resample_bodies": ([vp, dp, dp, dp, dp, dp, dp, C.POINTER(C.c_uint8)], C.c_int32),
        "hgx_resample_indicators": ([vp, i32p], C.c_int32),
        "hgx_resample_free": ([vp], None),
    }
    for name, (args, res) in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _ready = True
    return L


def _p(a, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype))


def mt1998(seed: int, n: int) -> tuple[np.ndarray, np.ndarray]:
    """randomMT_i() words and randomMT() reals of two streams seeded with seed."""
    w = np.zeros(n, np.uint32)
    r = np.zeros(n, np.float32)
    _lib().hgx_mt1998(seed, n, _p(w, C.c_uint32), _p(r, C.c_float))
    return w, r


def circle_rect_area(x0, x1, y0, y1, cx, cy, r) -> float:
    return float(_lib().hgx_circle_rect_area(x0, x1, y0, y1, cx, cy, r))


def zone_tools(gmin, h, cells, ls_h, circle_pos, circle_r, points, window, spc, phi_threshold,
               level_set, half_thickness, allow_direct=False) -> dict:
    """ZoneTools::updatePackingFraction, updateDistanceField and computeRZones
    (ZoneTools.cpp:185-685) on one grid; returns the packing fraction per
    level-set cell, the signed distance per level-set node and both indicator
    sets (ZT_CONTINUUM 0, ZT_DISCRETE 1, ZT_HYBRID 2)."""
    cells = np.asarray(cells, np.uint32)
    nls = np.ceil(cells * h / ls_h).astype(np.int64)
    cp = np.ascontiguousarray(circle_pos, np.float64).reshape(-1)
    cr = np.ascontiguousarray(circle_r, np.float64).reshape(-1)
    pts = np.ascontiguousarray(points, np.float64).reshape(-1)
    phi = np.zeros(int(nls[0] * nls[1]))
    dist = np.zeros(int((nls[0] + 1) * (nls[1] + 1)))
    ls = np.zeros(int(nls[0] * nls[1]), np.int32)
    mi = np.zeros(int(cells[0]) * int(cells[1]), np.int32)
    g = np.asarray(gmin, np.float64)
    st = _lib().hgx_zone_tools(_p(g), h, _p(cells, C.c_uint32), ls_h, len(cr), _p(cp), _p(cr),
                               len(pts) // 2, _p(pts), window, spc, phi_threshold, level_set,
                               half_thickness, 1 if allow_direct else 0, _p(phi), _p(dist),
                               _p(ls, C.c_int32), _p(mi, C.c_int32))
    if st != 0:
        raise HgError(st, "hgx_zone_tools")
    ny, nx = int(nls[1]), int(nls[0])
    return {"phi": phi.reshape(ny, nx), "dist": dist.reshape(ny + 1, nx + 1),
            "ls_indicator": ls.reshape(ny, nx), "mpm_indicator": mi.reshape(int(cells[1]), int(cells[0]))}


class HostResampler:
    """Resampling passes on a hybrid scene as HybridGrains2DSim starts it (no
    continuum points, the grains as read, no contacts), on the host only."""

    def __init__(self, scene_path: str):
        L = _lib()
        self._sc = C.c_void_p()
        err = C.create_string_buffer(512)
        if L.hgx_hybrid_scene_read(str(scene_path).encode(), C.byref(self._sc), err, 512) != 0:
            raise HgError(-1, err.value.decode())
        self._h = C.c_void_p()
        st = L.hgx_resample_create(self._sc, C.byref(self._h))
        if st != 0:
            L.hgx_hybrid_scene_free(self._sc)
            raise HgError(st, "hgx_resample_create (does the scene have a coupling_force?)")

    def close(self):
        L = _lib()
        if self._h:
            L.hgx_resample_free(self._h)
            self._h = C.c_void_p()
        if self._sc:
            L.hgx_hybrid_scene_free(self._sc)
            self._sc = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self) -> dict:
        rep = (C.c_int64 * 7)()
        st = _lib().hgx_resample_run(self._h, rep)
        if st != 0:
            raise HgError(st, "hgx_resample_run")
        r = list(rep)
        return {"dem": r[0:3], "mpm": r[3:6], "rzones": r[6]}

    def counts(self) -> tuple[int, int]:
        n, b = C.c_uint32(), C.c_uint32()
        _lib().hgx_resample_counts(self._h, C.byref(n), C.byref(b))
        return n.value, b.value

    def points(self) -> dict:
        n, _ = self.counts()
        f = {k: np.zeros(n * w) for k, w in (("q", 2), ("v", 2), ("m", 1), ("volume", 1), ("m0", 1),
                                             ("volume0", 1), ("weight", 1), ("J", 1),
                                             ("be_bar", 4), ("hl", 1))}
        _lib().hgx_resample_points(self._h, *[_p(f[k]) for k in ("q", "v", "m", "volume", "m0",
                                                                 "volume0", "weight", "J",
                                                                 "be_bar", "hl")])
        for k, w in (("q", 2), ("v", 2), ("be_bar", 4)):
            f[k] = f[k].reshape(n, w)
        return f

    def bodies(self) -> dict:
        _, b = self.counts()
        f = {k: np.zeros(b * w) for k, w in (("q", 3), ("v", 3), ("M", 3), ("M0", 3), ("weight", 1),
                                             ("radius", 1))}
        fixed = np.zeros(b, np.uint8)
        _lib().hgx_resample_bodies(self._h, *[_p(f[k]) for k in ("q", "v", "M", "M0", "weight",
                                                                 "radius")],
                                   _p(fixed, C.c_uint8))
        for k in ("q", "v", "M", "M0"):
            f[k] = f[k].reshape(b, 3)
        f["fixed"] = fixed
        return f

    def indicators(self, cell_count) -> np.ndarray:
        nx, ny = int(cell_count[0]), int(cell_count[1])
        out = np.zeros(nx * ny, np.int32)
        st = _lib().hgx_resample_indicators(self._h, _p(out, C.c_int32))
        if st != 0:
            raise HgError(st, "hgx_resample_indicators")
        return out.reshape(ny, nx)
