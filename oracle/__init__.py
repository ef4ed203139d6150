"""ctypes binding of the CPU FP64 restatement (TEST INFRASTRUCTURE ONLY).

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg,
only ever as the checker or the timed CPU baseline.  The product path
(libhybridgrains_amd) never imports this package.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)


class OracleMpmC(C.Structure):
    _fields_ = [
        ("grid_min", C.c_double * 2),
        ("cell_width", C.c_double),
        ("cell_count", C.c_uint32 * 2),
        ("mass", _dp), ("mom", _dp), ("mom_new", _dp), ("force", _dp),
        ("vel", _dp), ("acc", _dp), ("coupling_force", _dp), ("hybridized", _ip),
        ("npoints", C.c_uint32),
        ("q", _dp), ("v", _dp), ("m", _dp), ("volume", _dp), ("hl", _dp),
        ("be_bar", _dp), ("J", _dp), ("sigma", _dp), ("vel_grad", _dp),
        ("just_homogenized", _ip), ("kinematically_scripted", _u8p),
        ("shear_modulus", C.c_double), ("bulk_modulus", C.c_double),
        ("dp_alpha", C.c_double), ("pic_alpha", C.c_double),
        ("gravity", C.c_double * 2),
        ("nplanes", C.c_uint32),
        ("plane_x", _dp), ("plane_n", _dp), ("plane_lower", _dp), ("plane_behavior", _ip),
        ("basis", C.c_int32),
    ]


class OracleDemC(C.Structure):
    _fields_ = [
        ("nbodies", C.c_uint32),
        ("q", _dp), ("v", _dp), ("mass", _dp), ("minv", _dp), ("radius", _dp),
        ("fixed", _u8p), ("weights", _dp),
        ("gravity", C.c_double * 2),
        ("nplanes", C.c_uint32),
        ("plane_x", _dp), ("plane_n", _dp),
        ("kn", C.c_double), ("kt", C.c_double), ("gamman", C.c_double),
        ("gammat", C.c_double), ("mu", C.c_double),
        ("hist", C.c_void_p),
    ]


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    L = C.CDLL(str(LIB_PATH))
    P = C.POINTER(OracleMpmC)
    for name in [
        "oracle_mpm_clear_rasterized_data", "oracle_mpm_rasterize_point_masses",
        "oracle_mpm_rasterize_point_momentum", "oracle_mpm_compute_cauchy_stress",
        "oracle_mpm_compute_forces", "oracle_mpm_grid_plane_collisions",
        "oracle_mpm_compute_vel_grad", "oracle_mpm_plastic_correction",
        "oracle_mpm_point_plane_collisions", "oracle_mpm_flow_zero_phase",
    ]:
        getattr(L, name).argtypes = [P]
        getattr(L, name).restype = None
    for name in [
        "oracle_mpm_update_momentum", "oracle_mpm_velocity_acceleration_update",
        "oracle_mpm_elastic_prediction", "oracle_mpm_update_velocities",
        "oracle_mpm_update_positions", "oracle_mpm_flow_first_phase",
        "oracle_mpm_flow_second_phase", "oracle_mpm_flow",
    ]:
        getattr(L, name).argtypes = [P, C.c_double]
        getattr(L, name).restype = None
    L.oracle_mpm_flow_n.argtypes = [P, C.c_double, C.c_uint32]
    L.oracle_mpm_flow_n.restype = None
    L.oracle_mpm_flow_n_omp.argtypes = [P, C.c_double, C.c_uint32, C.c_int]
    L.oracle_mpm_flow_n_omp.restype = None
    L.oracle_mpm_is_consistent.argtypes = [P]
    L.oracle_mpm_is_consistent.restype = C.c_int
    L.oracle_ugimp_linear_integral.argtypes = [C.c_double] * 4
    L.oracle_ugimp_linear_integral.restype = C.c_double
    L.oracle_ugimp_linear_integral_grad.argtypes = [C.c_double] * 4
    L.oracle_ugimp_linear_integral_grad.restype = C.c_double
    L.oracle_weight.argtypes = [P, _dp, C.c_double, C.c_int64, C.c_int64]
    L.oracle_weight.restype = C.c_double
    L.oracle_weight_grad.argtypes = [P, _dp, C.c_double, C.c_int64, C.c_int64, _dp]
    L.oracle_weight_grad.restype = None
    L.oracle_stencil.argtypes = [P, _dp, C.c_double, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.oracle_stencil.restype = None
    L.oracle_elastic_prediction.argtypes = [_dp, _dp, C.c_double, _dp]
    L.oracle_elastic_prediction.restype = None
    L.oracle_plastic_correction.argtypes = [_dp, C.c_double, C.c_double, C.c_double, _ip, _dp]
    L.oracle_plastic_correction.restype = None
    L.oracle_compute_tau.argtypes = [_dp, C.c_double, C.c_double, C.c_int32, _dp]
    L.oracle_compute_tau.restype = None
    PD = C.POINTER(OracleDemC)
    L.oracle_dem_init.argtypes = [PD]
    L.oracle_dem_init.restype = C.c_int
    L.oracle_dem_free.argtypes = [PD]
    L.oracle_dem_free.restype = None
    L.oracle_dem_flow.argtypes = [PD, C.c_double, C.c_uint32]
    L.oracle_dem_flow.restype = C.c_int
    L.oracle_dem_contacts.argtypes = [PD, _ip, _ip, _ip, _dp, C.c_uint32]
    L.oracle_dem_contacts.restype = C.c_uint32
    L.oracle_dem_pairs_bruteforce.argtypes = [PD, _ip, _ip, C.c_uint32]
    L.oracle_dem_pairs_bruteforce.restype = C.c_uint32
    L.oracle_dem_last_force.argtypes = [PD, _dp]
    L.oracle_dem_last_force.restype = None
    L.oracle_dem_circle_contacts.argtypes = [PD, _ip, _ip, _dp, _dp, C.c_uint32]
    L.oracle_dem_circle_contacts.restype = C.c_uint32
    _lib = L
    return L


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None else None


class OracleMpm:
    """A SimulationState on the host, stepped by the C restatement.

    ``scene`` is any object with the ``MpmScene`` fields (q (N,2), be_bar (N,4)
    column-major, planes, material).
    """

    def __init__(self, scene):
        self.L = lib()
        s = scene
        N = int(s.q.shape[0])
        G = int((s.cell_count[0] + 1) * (s.cell_count[1] + 1))
        self.N, self.G = N, G
        f8 = lambda a: np.ascontiguousarray(np.array(a, dtype=np.float64))  # noqa: E731
        self.mass = np.zeros(G)
        self.mom = np.zeros(2 * G)
        self.mom_new = np.zeros(2 * G)
        self.force = np.zeros(2 * G)
        self.vel = np.zeros(2 * G)
        self.acc = np.zeros(2 * G)
        self.coupling_force = np.zeros(2 * G)
        self.hybridized = np.zeros(G, np.int32)
        self.q = f8(s.q).reshape(-1).copy()
        self.v = f8(s.v).reshape(-1).copy()
        self.m = f8(s.m).copy()
        self.volume = f8(s.volume).copy()
        self.hl = f8(s.hl).copy()
        self.be_bar = f8(s.be_bar).reshape(-1).copy()
        self.J = f8(s.J).copy()
        self.sigma = np.zeros(4 * N)
        self.vel_grad = np.zeros(4 * N)
        self.just_homogenized = np.zeros(N, np.int32)
        ks = getattr(s, "kinematically_scripted", None)
        self.ks = np.ascontiguousarray(ks, np.uint8) if ks is not None else None
        self.plane_x = f8(s.plane_x).reshape(-1).copy()
        self.plane_n = f8(s.plane_n).reshape(-1).copy()
        self.plane_lower = f8(s.plane_lower).copy()
        self.plane_behavior = np.ascontiguousarray(s.plane_behavior, np.int32)
        c = OracleMpmC()
        c.grid_min[:] = [float(s.grid_min[0]), float(s.grid_min[1])]
        c.cell_width = float(s.cell_width)
        c.cell_count[:] = [int(s.cell_count[0]), int(s.cell_count[1])]
        for k in ("mass", "mom", "mom_new", "force", "vel", "acc", "coupling_force",
                  "q", "v", "m", "volume", "hl", "be_bar", "J", "sigma", "vel_grad",
                  "plane_x", "plane_n", "plane_lower"):
            setattr(c, k, _p(getattr(self, k)))
        c.hybridized = _p(self.hybridized, C.c_int32)
        c.just_homogenized = _p(self.just_homogenized, C.c_int32)
        c.plane_behavior = _p(self.plane_behavior, C.c_int32)
        c.kinematically_scripted = _p(self.ks, C.c_uint8) if self.ks is not None else None
        c.npoints = N
        c.shear_modulus = float(s.shear_modulus)
        c.bulk_modulus = float(s.bulk_modulus)
        c.dp_alpha = float(s.dp_alpha)
        c.pic_alpha = float(s.pic_alpha)
        c.gravity[:] = [float(s.gravity[0]), float(s.gravity[1])]
        c.nplanes = len(self.plane_lower)
        c.basis = int(s.basis)
        self.c = c

    @property
    def ref(self):
        return C.byref(self.c)

    def flow(self, dt: float, nsteps: int = 1) -> None:
        self.L.oracle_mpm_flow_n(self.ref, float(dt), int(nsteps))

    def flow_omp(self, dt: float, nsteps: int, nthreads: int) -> None:
        self.L.oracle_mpm_flow_n_omp(self.ref, float(dt), int(nsteps), int(nthreads))

    def call(self, name: str, *args) -> None:
        getattr(self.L, "oracle_mpm_" + name)(self.ref, *args)

    def points(self) -> dict:
        return {"q": self.q.reshape(-1, 2).copy(), "v": self.v.reshape(-1, 2).copy(),
                "be_bar": self.be_bar.reshape(-1, 4).copy(), "J": self.J.copy()}

    def grid(self) -> dict:
        return {"mass": self.mass.copy(), "mom": self.mom.reshape(-1, 2).copy(),
                "mom_new": self.mom_new.reshape(-1, 2).copy(),
                "force": self.force.reshape(-1, 2).copy(),
                "vel": self.vel.reshape(-1, 2).copy(), "acc": self.acc.reshape(-1, 2).copy()}

    def is_consistent(self) -> int:
        return int(self.L.oracle_mpm_is_consistent(self.ref))


class OracleDem:
    """A RigidBody2DState of circles stepped by the C restatement of the
    penalty map.  ``scene`` has the ``DemScene`` fields."""

    def __init__(self, scene, weights=None):
        self.L = lib()
        s = scene
        B = int(s.q.shape[0])
        self.B = B
        f8 = lambda a: np.ascontiguousarray(np.array(a, dtype=np.float64))  # noqa: E731
        self.q = f8(s.q).reshape(-1).copy()
        self.v = f8(s.v).reshape(-1).copy()
        self.mass = f8(s.mass).reshape(-1).copy()
        self.minv = 1.0 / self.mass
        self.radius = f8(s.radius).copy()
        self.fixed = np.ascontiguousarray(s.fixed, np.uint8).copy()
        self.weights = f8(weights).copy() if weights is not None else None
        self.plane_x = f8(s.plane_x).reshape(-1).copy()
        self.plane_n = f8(s.plane_n).reshape(-1).copy()
        c = OracleDemC()
        c.nbodies = B
        for k in ("q", "v", "mass", "minv", "radius", "plane_x", "plane_n"):
            setattr(c, k, _p(getattr(self, k)))
        c.fixed = _p(self.fixed, C.c_uint8)
        c.weights = _p(self.weights) if self.weights is not None else None
        c.gravity[:] = [float(s.gravity[0]), float(s.gravity[1])]
        c.nplanes = int(s.plane_x.shape[0])
        c.kn, c.kt, c.gamman, c.gammat, c.mu = (float(s.kn), float(s.kt), float(s.gamman),
                                                float(s.gammat), float(s.mu))
        c.hist = None
        self.c = c
        if self.L.oracle_dem_init(C.byref(c)) != 0:
            raise MemoryError("oracle_dem_init")

    def __del__(self):
        try:
            self.L.oracle_dem_free(C.byref(self.c))
        except Exception:
            pass

    def flow(self, dt: float, nsteps: int = 1) -> None:
        if self.L.oracle_dem_flow(C.byref(self.c), float(dt), int(nsteps)) != 0:
            raise MemoryError("oracle_dem_flow")

    def contacts(self) -> dict:
        n = self.L.oracle_dem_contacts(C.byref(self.c), None, None, None, None, 0)
        t = np.zeros(n, np.int32)
        i = np.zeros(n, np.int32)
        j = np.zeros(n, np.int32)
        ds = np.zeros((n, 2))
        self.L.oracle_dem_contacts(C.byref(self.c), _p(t, C.c_int32), _p(i, C.c_int32),
                                   _p(j, C.c_int32), _p(ds), n)
        return {"type": t, "i": i, "j": j, "delta_s": ds}

    def circle_contacts(self) -> dict:
        """Circle-circle contacts of the last step: owner i, partner j, contact
        point p and the normal + friction force f on i."""
        n = self.L.oracle_dem_circle_contacts(C.byref(self.c), None, None, None, None, 0)
        i = np.zeros(n, np.int32)
        j = np.zeros(n, np.int32)
        p = np.zeros((n, 2))
        f = np.zeros((n, 2))
        self.L.oracle_dem_circle_contacts(C.byref(self.c), _p(i, C.c_int32), _p(j, C.c_int32),
                                          _p(p), _p(f), n)
        return {"i": i, "j": j, "p": p, "f": f}

    def pairs_bruteforce(self) -> np.ndarray:
        n = self.L.oracle_dem_pairs_bruteforce(C.byref(self.c), None, None, 0)
        i = np.zeros(n, np.int32)
        j = np.zeros(n, np.int32)
        self.L.oracle_dem_pairs_bruteforce(C.byref(self.c), _p(i, C.c_int32), _p(j, C.c_int32), n)
        return np.stack([i, j], 1)

    def last_force(self) -> np.ndarray:
        F = np.zeros(3 * self.B)
        self.L.oracle_dem_last_force(C.byref(self.c), _p(F))
        return F.reshape(-1, 3)

    def state(self) -> dict:
        return {"q": self.q.reshape(-1, 3).copy(), "v": self.v.reshape(-1, 3).copy()}


class OracleHybridC(C.Structure):
    _fields_ = [
        ("mpm", C.POINTER(OracleMpmC)), ("dem", C.POINTER(OracleDemC)),
        ("dem_pos_before", _dp), ("dem_mass_grid", _dp), ("dem_mom_grid", _dp),
        ("v_after", _dp), ("at_after", _dp), ("homog_factor", _dp),
        ("coupled", _u8p), ("homog_stress", _dp), ("homog_velocity", _dp),
    ]


class OracleHybrid:
    """HybridGrains2DSim's coupled step (no zone resampling) on the host:
    an OracleMpm and an OracleDem (with mass weights) plus the coupling
    buffers."""

    def __init__(self, mpm_scene, dem_scene, weights, homogenize: bool = False):
        self.mpm = OracleMpm(mpm_scene)
        self.dem = OracleDem(dem_scene, weights=weights)
        L = self.mpm.L
        G, B = self.mpm.G, self.dem.B
        self.dem_pos_before = np.zeros(3 * B)
        self.dem_mass_grid = np.zeros(G)
        self.dem_mom_grid = np.zeros(2 * G)
        self.v_after = np.zeros(2 * G)
        self.at_after = np.zeros(2 * G)
        self.homog_factor = np.zeros(G)
        self.coupled = np.zeros(B, np.uint8)
        c = OracleHybridC()
        c.mpm = C.pointer(self.mpm.c)
        c.dem = C.pointer(self.dem.c)
        for k in ("dem_pos_before", "dem_mass_grid", "dem_mom_grid", "v_after", "at_after",
                  "homog_factor"):
            setattr(c, k, _p(getattr(self, k)))
        c.coupled = _p(self.coupled, C.c_uint8)
        self.homog_stress = np.zeros(4 * G) if homogenize else None
        self.homog_velocity = np.zeros(2 * G) if homogenize else None
        if homogenize:
            c.homog_stress = _p(self.homog_stress)
            c.homog_velocity = _p(self.homog_velocity)
        self.c = c
        P = C.POINTER(OracleHybridC)
        L.oracle_hybrid_step.argtypes = [P, C.c_double, C.c_uint32]
        L.oracle_hybrid_step.restype = C.c_int
        L.oracle_hybrid_step_omp.argtypes = [P, C.c_double, C.c_uint32, C.c_int]
        L.oracle_hybrid_step_omp.restype = C.c_int
        self.L = L

    def step(self, dt: float, nsteps: int = 1) -> None:
        if self.L.oracle_hybrid_step(C.byref(self.c), float(dt), int(nsteps)) != 0:
            raise MemoryError("oracle_hybrid_step")

    def step_omp(self, dt: float, nsteps: int, nthreads: int) -> None:
        """The same step with the MPM part on nthreads OpenMP threads (the
        reference's per-thread grid buffers); no homogenisation."""
        if self.L.oracle_hybrid_step_omp(C.byref(self.c), float(dt), int(nsteps),
                                         int(nthreads)) != 0:
            raise MemoryError("oracle_hybrid_step_omp")

    def homogenized(self) -> dict:
        """The accumulated homogenised stress (G x 4, column-major 2x2) and
        velocity (G x 2), nodes in the reference's order."""
        return {"stress": self.homog_stress.reshape(-1, 4).copy(),
                "velocity": self.homog_velocity.reshape(-1, 2).copy()}
