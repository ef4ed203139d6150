"""ctypes loader for libhg.so, the gfx950 engine.

The library is built in-tree (``python -m libhybridgrains_amd.build``).  There
is no fallback: if the shared object is missing, importing the engine raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "libhg.so"
_lib = None

HG_OK = 0
HG_ERR_INVALID = -1
HG_ERR_HIP = -2
HG_ERR_NOMEM = -3
HG_ERR_STATE = -4
HG_ERR_UNSUPPORTED = -5

HG_FP64 = 0
HG_FP32 = 1
HG_BASIS_UGIMP_LINEAR = 0
HG_BASIS_LINEAR = 1
HG_BASIS_CUBIC = 3
HG_PLANE_SLIDING = 0
HG_PLANE_STICKING = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)


class HgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"hg error {code}: {msg}")
        self.code = code


class MpmDesc(C.Structure):
    _fields_ = [
        ("precision", C.c_int32),
        ("basis", C.c_int32),
        ("grid_min", C.c_double * 2),
        ("cell_width", C.c_double),
        ("cell_count", C.c_uint32 * 2),
        ("shear_modulus", C.c_double),
        ("bulk_modulus", C.c_double),
        ("dp_alpha", C.c_double),
        ("pic_alpha", C.c_double),
        ("gravity", C.c_double * 2),
        ("nplanes", C.c_uint32),
        ("plane_x", _dp),
        ("plane_n", _dp),
        ("plane_lower", _dp),
        ("plane_behavior", _ip),
        ("npoints", C.c_uint32),
        ("q", _dp),
        ("v", _dp),
        ("m", _dp),
        ("volume", _dp),
        ("hl", _dp),
        ("be_bar", _dp),
        ("J", _dp),
        ("kinematically_scripted", _u8p),
    ]


class DemDesc(C.Structure):
    _fields_ = [
        ("precision", C.c_int32),
        ("nbodies", C.c_uint32),
        ("q", _dp),
        ("v", _dp),
        ("mass", _dp),
        ("radius", _dp),
        ("fixed", _u8p),
        ("gravity", C.c_double * 2),
        ("nplanes", C.c_uint32),
        ("plane_x", _dp),
        ("plane_n", _dp),
        ("kn", C.c_double),
        ("kt", C.c_double),
        ("gamman", C.c_double),
        ("gammat", C.c_double),
        ("mu", C.c_double),
        ("cell_width", C.c_double),
        ("domain_min", C.c_double * 2),
        ("domain_max", C.c_double * 2),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("budget_bytes", C.c_uint64),
        ("max_active", C.c_uint32),
        ("slice_steps", C.c_uint32),
        ("max_inflight", C.c_uint32),
        ("max_bypass", C.c_uint32),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
        ("cancelled", C.c_uint64), ("active", C.c_uint64), ("queued", C.c_uint64),
        ("bytes_in_use", C.c_uint64), ("p50_latency_s", C.c_double),
        ("p99_latency_s", C.c_double), ("mean_latency_s", C.c_double),
        ("max_latency_s", C.c_double), ("work_steps", C.c_double), ("span_s", C.c_double),
    ]


def lib() -> C.CDLL:
    """Loads libhg.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise ImportError(
            f"{_LIB_PATH} is missing: build it with `python -m libhybridgrains_amd.build`"
        )
    L = C.CDLL(str(_LIB_PATH))
    vp = C.c_void_p
    sigs = {
        "hg_ctx_create": ([C.c_int32, C.POINTER(vp)], C.c_int32),
        "hg_ctx_destroy": ([vp], C.c_int32),
        "hg_last_error": ([vp], C.c_char_p),
        "hg_version": ([], C.c_char_p),
        "hg_mpm_create": ([vp, C.POINTER(MpmDesc), C.POINTER(vp)], C.c_int32),
        "hg_mpm_create_batch": ([vp, C.POINTER(MpmDesc), C.c_uint32, C.POINTER(vp)], C.c_int32),
        "hg_mpm_destroy": ([vp], C.c_int32),
        "hg_mpm_step": ([vp, C.c_double, C.c_uint32], C.c_int32),
        "hg_mpm_step_graph": ([vp, C.c_double, C.c_uint32], C.c_int32),
        "hg_mpm_phase": ([vp, C.c_int32, C.c_double], C.c_int32),
        "hg_mpm_sync": ([vp], C.c_int32),
        "hg_mpm_scene_error": ([vp, C.c_uint32, _ip], C.c_int32),
        "hg_mpm_download_points": ([vp, _dp, _dp, _dp, _dp], C.c_int32),
        "hg_mpm_download_points_scene": ([vp, C.c_uint32, _dp, _dp, _dp, _dp], C.c_int32),
        "hg_mpm_set_keep_grid": ([vp, C.c_int32], C.c_int32),
        "hg_mpm_download_grid": ([vp, _dp, _dp, _dp, _dp, _dp, _dp], C.c_int32),
        "hg_mpm_download_grid_scene": ([vp, C.c_uint32, _dp, _dp, _dp, _dp, _dp, _dp], C.c_int32),
        "hg_mpm_counts": ([vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)], C.c_int32),
        "hg_mpm_set_profiling": ([vp, C.c_int32], C.c_int32),
        "hg_mpm_hold": ([vp, C.c_double], C.c_int32),
        "hg_mpm_kernel_times": (
            [vp, C.c_int32, C.c_char_p, _dp, C.POINTER(C.c_int64), _ip],
            C.c_int32,
        ),
        "hg_mpm_set_tile": ([vp, C.c_int32, C.c_int32], C.c_int32),
        "hg_mpm_autotune": ([vp, C.c_double, C.c_int32, _ip, _ip, _dp], C.c_int32),
        "hg_mpm_tuning": ([vp, _ip, _ip, _ip], C.c_int32),
        "hg_mpm_set_points_per_pass": ([vp, C.c_int32], C.c_int32),
        "hg_bench_copy": ([vp, C.c_uint64, C.c_int32, _dp], C.c_int32),
        "hg_dem_create": ([vp, C.POINTER(DemDesc), C.POINTER(vp)], C.c_int32),
        "hg_dem_create_batch": ([vp, C.POINTER(DemDesc), C.c_uint32, C.POINTER(vp)], C.c_int32),
        "hg_dem_destroy": ([vp], C.c_int32),
        "hg_dem_step": ([vp, C.c_double, C.c_uint32], C.c_int32),
        "hg_dem_step_graph": ([vp, C.c_double, C.c_uint32], C.c_int32),
        "hg_dem_set_weights": ([vp, C.c_uint32, _dp], C.c_int32),
        "hg_dem_sync": ([vp], C.c_int32),
        "hg_dem_download_bodies": ([vp, C.c_uint32, _dp, _dp], C.c_int32),
        "hg_dem_download_contacts": (
            [vp, C.c_uint32, C.c_uint32, _ip, _ip, _ip, _dp, C.POINTER(C.c_uint32)], C.c_int32),
        "hg_dem_download_contact_forces": (
            [vp, C.c_uint32, C.c_uint32, _dp, _dp, C.POINTER(C.c_uint32)], C.c_int32),
        "hg_dem_upload_contacts": ([vp, C.c_uint32, C.c_uint32, _ip, _ip, _ip, _dp], C.c_int32),
        "hg_dem_keep_force": ([vp, C.c_int32], C.c_int32),
        "hg_dem_download_force": ([vp, C.c_uint32, _dp], C.c_int32),
        "hg_dem_set_profiling": ([vp, C.c_int32], C.c_int32),
        "hg_dem_hold": ([vp, C.c_double], C.c_int32),
        "hg_dem_kernel_times": (
            [vp, C.c_int32, C.c_char_p, _dp, C.POINTER(C.c_int64), _ip], C.c_int32),
        "hg_hybrid_create": (
            [vp, C.POINTER(MpmDesc), C.POINTER(DemDesc), _dp, C.POINTER(vp)], C.c_int32),
        "hg_hybrid_destroy": ([vp], C.c_int32),
        "hg_hybrid_step": ([vp, C.c_double, C.c_uint32], C.c_int32),
        "hg_hybrid_step_graph": ([vp, C.c_double, C.c_uint32], C.c_int32),
        "hg_hybrid_sync": ([vp], C.c_int32),
        "hg_hybrid_mpm": ([vp, C.POINTER(vp)], C.c_int32),
        "hg_hybrid_dem": ([vp, C.POINTER(vp)], C.c_int32),
        "hg_hybrid_download_coupling": ([vp, C.POINTER(C.c_uint8)], C.c_int32),
        "hg_hybrid_set_homogenize": ([vp, C.c_int32, C.c_int32], C.c_int32),
        "hg_hybrid_download_homogenized": ([vp, _dp, _dp], C.c_int32),
        "hg_hybrid_set_homog_factor": ([vp, C.c_int32], C.c_int32),
        "hg_hybrid_download_homog_factor": ([vp, _dp], C.c_int32),
        "hg_hybrid_clear_homogenized": ([vp], C.c_int32),
        "hg_mpm_snapshot": ([vp, C.c_void_p, C.POINTER(C.c_uint64)], C.c_int32),
        "hg_mpm_restore": ([vp, C.c_void_p, C.c_uint64], C.c_int32),
        "hg_dem_snapshot": ([vp, C.c_void_p, C.POINTER(C.c_uint64)], C.c_int32),
        "hg_dem_restore": ([vp, C.c_void_p, C.c_uint64], C.c_int32),
        "hg_hybrid_snapshot": ([vp, C.c_void_p, C.POINTER(C.c_uint64)], C.c_int32),
        "hg_hybrid_restore": ([vp, C.c_void_p, C.c_uint64], C.c_int32),
        "hg_mpm_estimate_bytes": ([C.POINTER(MpmDesc), C.c_uint32, C.POINTER(C.c_uint64)], C.c_int32),
        "hg_mpm_device_bytes": ([vp, C.POINTER(C.c_uint64)], C.c_int32),
        "hg_dem_estimate_bytes": ([C.POINTER(DemDesc), C.c_uint32, C.POINTER(C.c_uint64)], C.c_int32),
        "hg_dem_device_bytes": ([vp, C.POINTER(C.c_uint64)], C.c_int32),
        "hg_hybrid_estimate_bytes": (
            [C.POINTER(MpmDesc), C.POINTER(DemDesc), C.POINTER(C.c_uint64)], C.c_int32),
        "hg_hybrid_device_bytes": ([vp, C.POINTER(C.c_uint64)], C.c_int32),
        "hg_server_create": ([vp, C.POINTER(ServerConfig), C.POINTER(vp)], C.c_int32),
        "hg_server_destroy": ([vp], C.c_int32),
        "hg_server_submit_mpm": (
            [vp, C.POINTER(MpmDesc), C.c_double, C.c_uint64, C.POINTER(C.c_uint64)], C.c_int32),
        "hg_server_submit_dem": (
            [vp, C.POINTER(DemDesc), _dp, C.c_double, C.c_uint64, C.POINTER(C.c_uint64)],
            C.c_int32),
        "hg_server_submit_hybrid": (
            [vp, C.POINTER(MpmDesc), C.POINTER(DemDesc), _dp, C.c_double, C.c_uint64,
             C.POINTER(C.c_uint64)], C.c_int32),
        "hg_server_cancel": ([vp, C.c_uint64], C.c_int32),
        "hg_server_poll": ([vp, C.c_uint64, _ip, C.POINTER(C.c_uint64)], C.c_int32),
        "hg_server_wait": ([vp, C.c_uint64, C.c_double, _ip], C.c_int32),
        "hg_server_drain": ([vp, C.c_double], C.c_int32),
        "hg_server_result_points": ([vp, C.c_uint64, _dp, _dp, _dp, _dp], C.c_int32),
        "hg_server_result_bodies": ([vp, C.c_uint64, _dp, _dp], C.c_int32),
        "hg_server_latency": ([vp, C.c_uint64, _dp], C.c_int32),
        "hg_server_error": ([vp, C.c_uint64, C.c_char_p, C.c_size_t], C.c_int32),
        "hg_server_release": ([vp, C.c_uint64], C.c_int32),
        "hg_server_stats_get": ([vp, C.POINTER(ServerStats)], C.c_int32),
        "hg_mpm_observables": ([vp, C.c_uint32, C.c_double, C.c_double, C.c_uint32, _dp],
                               C.c_int32),
        "hg_dem_observables": ([vp, C.c_uint32, C.c_double, C.c_double, C.c_uint32, _dp],
                               C.c_int32),
        "hg_mpm_observables_noalloc": ([vp, C.c_uint32, C.c_double, C.c_double, C.c_uint32,
                                        _dp], C.c_int32),
    }
    for name, (args, res) in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def ptr(a: np.ndarray | None, ctype=C.c_double):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ctype))


class Context:
    """One GPU (hg_ctx)."""

    def __init__(self, device: int = 0):
        L = lib()
        h = C.c_void_p()
        rc = L.hg_ctx_create(device, C.byref(h))
        if rc != HG_OK:
            raise HgError(rc, f"hg_ctx_create({device}) failed: is a GPU visible?")
        self._h = h
        self._lib = L

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def check(self, rc: int) -> None:
        if rc != HG_OK:
            msg = self._lib.hg_last_error(self._h)
            raise HgError(rc, msg.decode() if msg else "")

    def close(self) -> None:
        if self._h:
            self._lib.hg_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Context | None = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def snapshot_blob(ctx, fn, handle) -> bytes:
    """Calls an hg_*_snapshot entry point twice (size, then data)."""
    n = C.c_uint64(0)
    ctx.check(fn(handle, None, C.byref(n)))
    buf = C.create_string_buffer(n.value)
    ctx.check(fn(handle, buf, C.byref(n)))
    return buf.raw[: n.value]


def restore_blob(ctx, fn, handle, blob: bytes) -> None:
    """Calls an hg_*_restore entry point with a snapshot blob."""
    buf = C.create_string_buffer(bytes(blob), len(blob))
    ctx.check(fn(handle, buf, len(blob)))
