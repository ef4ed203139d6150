"""Scene files: the <mpmgrains2d>, <rigidbody2d_scene> and
<hybridgrains2d_scene> XML readers of libhg_host.so, returning the package's
MpmScene / DemScene.

The readers follow MPMGrains2DParser (mpmgrains2dutils/MPMGrains2DParser.cpp),
RigidBody2DSceneParser (rigidbody2dutils/RigidBody2DSceneParser.cpp, circle
scenes with a penalty_impact_map) and HybridGrains2DSceneParser
(hybridgrains2dnewutils/HybridGrains2DSceneParser.cpp), including the
rectangle seeding of InitialSimulationState::generateSimulationState.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from ._lib import DemDesc, HgError, MpmDesc, lib as _hg_lib
from .dem import DemScene
from .mpm import MpmScene

_HOST = Path(__file__).resolve().parent / "libhg_host.so"
_L = None


def host_lib() -> C.CDLL:
    global _L
    if _L is not None:
        return _L
    if not _HOST.exists():
        raise ImportError(f"{_HOST} is missing: build it with `python -m libhybridgrains_amd.build`")
    _hg_lib()  # libhg.so first, so the host library's dependency resolves in-tree
    L = C.CDLL(str(_HOST))
    vp, cp, sz = C.c_void_p, C.c_char_p, C.c_size_t
    pvp = C.POINTER(vp)
    dp = C.POINTER(C.c_double)
    sigs = {
        "hgx_mpm_scene_read": ([cp, pvp, cp, sz], C.c_int32),
        "hgx_mpm_scene_read_text": ([cp, pvp, cp, sz], C.c_int32),
        "hgx_mpm_scene_desc": ([vp, C.c_int32, C.POINTER(MpmDesc)], C.c_int32),
        "hgx_mpm_scene_times": ([vp, dp, dp], C.c_int32),
        "hgx_mpm_scene_free": ([vp], None),
        "hgx_dem_scene_read": ([cp, pvp, cp, sz], C.c_int32),
        "hgx_dem_scene_read_text": ([cp, pvp, cp, sz], C.c_int32),
        "hgx_dem_scene_desc": ([vp, C.c_int32, C.c_double, C.POINTER(DemDesc)], C.c_int32),
        "hgx_dem_scene_times": ([vp, dp, dp], C.c_int32),
        "hgx_dem_scene_free": ([vp], None),
        "hgx_hybrid_scene_read": ([cp, pvp, cp, sz], C.c_int32),
        "hgx_hybrid_scene_parts": ([vp, pvp, pvp], C.c_int32),
        "hgx_hybrid_scene_times": ([vp, dp, dp], C.c_int32),
        "hgx_hybrid_scene_style": ([vp], cp),
        "hgx_hybrid_scene_free": ([vp], None),
    }
    for name, (args, res) in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _L = L
    return L


def _arr(p, n, dtype=np.float64):
    if n == 0 or not p:
        return np.zeros(n, dtype)
    return np.ctypeslib.as_array(p, shape=(n,)).astype(dtype, copy=True)


def _mpm_from_handle(L, h) -> tuple[MpmScene, float, float]:
    d = MpmDesc()
    if L.hgx_mpm_scene_desc(h, 0, C.byref(d)) != 0:
        raise HgError(-1, "hgx_mpm_scene_desc")
    N, P = int(d.npoints), int(d.nplanes)
    s = MpmScene(
        grid_min=np.array(d.grid_min[:]), cell_width=float(d.cell_width),
        cell_count=np.array(d.cell_count[:], np.int64), shear_modulus=float(d.shear_modulus),
        bulk_modulus=float(d.bulk_modulus), dp_alpha=float(d.dp_alpha),
        pic_alpha=float(d.pic_alpha), gravity=np.array(d.gravity[:]),
        q=_arr(d.q, 2 * N).reshape(N, 2), v=_arr(d.v, 2 * N).reshape(N, 2), m=_arr(d.m, N),
        volume=_arr(d.volume, N), hl=_arr(d.hl, N), be_bar=_arr(d.be_bar, 4 * N).reshape(N, 4),
        J=_arr(d.J, N), plane_x=_arr(d.plane_x, 2 * P).reshape(P, 2),
        plane_n=_arr(d.plane_n, 2 * P).reshape(P, 2), plane_lower=_arr(d.plane_lower, P),
        plane_behavior=_arr(d.plane_behavior, P, np.int32), basis=int(d.basis),
        kinematically_scripted=_arr(d.kinematically_scripted, N, np.uint8),
    )
    dt, et = C.c_double(), C.c_double()
    L.hgx_mpm_scene_times(h, C.byref(dt), C.byref(et))
    return s, dt.value, et.value


def _dem_from_handle(L, h, cell_scale: float) -> tuple[DemScene, float, float, float]:
    d = DemDesc()
    if L.hgx_dem_scene_desc(h, 0, float(cell_scale), C.byref(d)) != 0:
        raise HgError(-1, "hgx_dem_scene_desc")
    B, P = int(d.nbodies), int(d.nplanes)
    s = DemScene(
        q=_arr(d.q, 3 * B).reshape(B, 3), v=_arr(d.v, 3 * B).reshape(B, 3),
        radius=_arr(d.radius, B), mass=_arr(d.mass, 3 * B).reshape(B, 3),
        fixed=_arr(d.fixed, B, np.uint8), gravity=np.array(d.gravity[:]), kn=float(d.kn),
        kt=float(d.kt), gamman=float(d.gamman), gammat=float(d.gammat), mu=float(d.mu),
        plane_x=_arr(d.plane_x, 2 * P).reshape(P, 2), plane_n=_arr(d.plane_n, 2 * P).reshape(P, 2),
    )
    dt, et = C.c_double(), C.c_double()
    L.hgx_dem_scene_times(h, C.byref(dt), C.byref(et))
    return s, dt.value, et.value, float(d.cell_width)


def _read(L, fn, src: str):
    h = C.c_void_p()
    err = C.create_string_buffer(1024)
    rc = fn(src.encode(), C.byref(h), err, len(err))
    if rc != 0:
        raise ValueError(err.value.decode(errors="replace"))
    return h


@dataclass
class MpmXml:
    scene: MpmScene
    dt: float
    end_time: float


@dataclass
class DemXml:
    scene: DemScene
    dt: float
    end_time: float
    cell_width: float


@dataclass
class HybridXml:
    continuum: MpmScene
    discrete: DemScene
    dt: float
    end_time: float
    style: str
    cell_width: float


def load_mpm(path=None, text: str | None = None) -> MpmXml:
    L = host_lib()
    h = _read(L, L.hgx_mpm_scene_read_text if text is not None else L.hgx_mpm_scene_read,
              text if text is not None else str(path))
    try:
        return MpmXml(*_mpm_from_handle(L, h))
    finally:
        L.hgx_mpm_scene_free(h)


def load_dem(path=None, text: str | None = None, cell_scale: float = 0.0) -> DemXml:
    """cell_scale <= 0 takes the scene's collision_detection cell_scale."""
    L = host_lib()
    h = _read(L, L.hgx_dem_scene_read_text if text is not None else L.hgx_dem_scene_read,
              text if text is not None else str(path))
    try:
        return DemXml(*_dem_from_handle(L, h, cell_scale))
    finally:
        L.hgx_dem_scene_free(h)


def load_hybrid(path) -> HybridXml:
    """The hybrid scene and the two scenes it names; the grains' broad phase
    uses scale 1 as the hybrid simulation does (RigidBody2DSim.cpp:41-42)."""
    L = host_lib()
    h = _read(L, L.hgx_hybrid_scene_read, str(path))
    try:
        mh, dh = C.c_void_p(), C.c_void_p()
        L.hgx_hybrid_scene_parts(h, C.byref(mh), C.byref(dh))
        m, _, _ = _mpm_from_handle(L, mh)
        d, _, _, cw = _dem_from_handle(L, dh, 1.0)
        dt, et = C.c_double(), C.c_double()
        L.hgx_hybrid_scene_times(h, C.byref(dt), C.byref(et))
        style = L.hgx_hybrid_scene_style(h).decode()
        return HybridXml(m, d, dt.value, et.value, style, cw)
    finally:
        L.hgx_hybrid_scene_free(h)
