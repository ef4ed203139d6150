"""Pins of the MPM CPU restatement (oracle/mpm_oracle.c) against mathematics
the reference's algorithm fixes, independent of the oracle's own formulas:
quadrature of the uGIMP overlap integral, partition of unity and linear
reproduction of every basis, the P2G conservation asserts of the reference
(PhysicsGrid.cpp:201, :207, :316-327, :545), exact free fall, exact
velocity-gradient recovery of a linear velocity field (pins the v (x) grad w
orientation), plane closed forms, and the return-map invariants the reference
self-checks (ConstitutiveModel.cpp:62-77).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from libhybridgrains_amd import HG_BASIS_CUBIC, HG_BASIS_LINEAR, HG_BASIS_UGIMP_LINEAR, scenes

BASES = [HG_BASIS_UGIMP_LINEAR, HG_BASIS_LINEAR, HG_BASIS_CUBIC]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _scene(basis=HG_BASIS_UGIMP_LINEAR, seed=0, **kw):
    kw.setdefault("domain", (0.3, 0.25))
    kw.setdefault("column", (0.08, 0.08))
    return scenes.column_collapse(h=0.005, seed=seed, basis=basis, **kw)


# ---- basis functions ---------------------------------------------------------

def _hat_integral(xp, hl, xi, w, n=20001):
    # composite Simpson of the linear hat over the point's extent
    x = np.linspace(xp - hl, xp + hl, n)
    f = np.maximum(0.0, 1.0 - np.abs(x - xi) / w)
    hstep = (x[-1] - x[0]) / (n - 1)
    return hstep / 3 * (f[0] + f[-1] + 4 * f[1:-1:2].sum() + 2 * f[2:-1:2].sum())


def test_ugimp_integral_matches_quadrature():
    L = oracle.lib()
    rng = np.random.default_rng(0)
    w = 1.0
    for _ in range(200):
        hl = rng.uniform(0.05, 0.45)
        xp = rng.uniform(-2.0, 2.0)
        ref = _hat_integral(xp, hl, 0.0, w)
        got = L.oracle_ugimp_linear_integral(xp, hl, 0.0, w)
        assert abs(got - ref) < 1e-9, (xp, hl, got, ref)


def test_ugimp_integral_grad_is_derivative():
    # d/dxp of the overlap integral = hat(xp + hl) - hat(xp - hl)
    L = oracle.lib()
    rng = np.random.default_rng(1)
    w = 1.0
    hat = lambda r: max(0.0, 1.0 - abs(r) / w)  # noqa: E731
    for _ in range(200):
        hl = rng.uniform(0.05, 0.45)
        xp = rng.uniform(-2.0, 2.0)
        got = L.oracle_ugimp_linear_integral_grad(xp, hl, 0.0, w)
        assert abs(got - (hat(xp + hl) - hat(xp - hl))) < 1e-12


@pytest.mark.parametrize("basis", BASES)
def test_partition_of_unity_and_linear_reproduction(basis):
    s = _scene(basis)
    o = oracle.OracleMpm(s)
    L = o.L
    rng = np.random.default_rng(2)
    h = s.cell_width
    for _ in range(100):
        xp = np.array([rng.uniform(4 * h, 0.2), rng.uniform(4 * h, 0.2)])
        hl = h / 4
        lo = (C.c_int64 * 2)()
        hi = (C.c_int64 * 2)()
        L.oracle_stencil(o.ref, _dp(xp), hl, lo, hi)
        sw, sg, sx = 0.0, np.zeros(2), np.zeros(2)
        gx_outer = np.zeros((2, 2))
        for iy in range(lo[1], hi[1] + 1):
            for ix in range(lo[0], hi[0] + 1):
                w = L.oracle_weight(o.ref, _dp(xp), hl, ix, iy)
                g = np.zeros(2)
                L.oracle_weight_grad(o.ref, _dp(xp), hl, ix, iy, _dp(g))
                xi = np.array([ix * h, iy * h])
                sw += w
                sg += g
                sx += w * xi
                gx_outer += np.outer(xi, g)
        assert abs(sw - 1.0) < 1e-12
        assert np.abs(sg).max() < 1e-9
        assert np.abs(sx - xp).max() < 1e-14
        # grad of sum_i x_i w_i = identity
        assert np.abs(gx_outer - np.eye(2)).max() < 1e-10


# ---- P2G conservation ----------------------------------------------------------

@pytest.mark.parametrize("basis", BASES)
def test_p2g_conserves_mass_momentum_angular_momentum(basis):
    s = _scene(basis, seed=3)
    rng = np.random.default_rng(3)
    s.v = rng.normal(size=s.v.shape)
    o = oracle.OracleMpm(s)
    o.call("flow_zero_phase")
    g = o.grid()
    h = s.cell_width
    nx = s.cell_count[0] + 1
    idx = np.arange(g["mass"].size)
    X = np.stack([(idx % nx) * h, (idx // nx) * h], 1)
    assert abs(g["mass"].sum() - s.m.sum()) <= 1e-12 * s.m.sum()
    P = (s.m[:, None] * s.v).sum(0)
    assert np.abs(g["mom"].sum(0) - P).max() <= 1e-12 * np.abs(s.m[:, None] * s.v).sum()
    Lp = (s.m * (s.q[:, 0] * s.v[:, 1] - s.q[:, 1] * s.v[:, 0])).sum()
    Lg = (X[:, 0] * g["mom"][:, 1] - X[:, 1] * g["mom"][:, 0]).sum()
    assert abs(Lg - Lp) <= 1e-10 * np.abs(s.m[:, None] * s.v).sum()


def test_internal_forces_sum_to_zero():
    s = _scene(seed=4)
    rng = np.random.default_rng(4)
    N = s.npoints
    s.J = 1.0 + rng.uniform(-0.05, -0.001, N)
    s.gravity = np.zeros(2)
    o = oracle.OracleMpm(s)
    o.call("flow_zero_phase")
    o.call("compute_cauchy_stress")
    o.call("compute_forces")
    f = o.force.reshape(-1, 2)
    assert np.abs(f).max() > 0
    assert np.abs(f.sum(0)).max() <= 1e-11 * np.abs(f).sum()


# ---- whole-step closed forms ---------------------------------------------------

def test_free_fall_is_exact():
    """Stress-free points with no planes: every point gains exactly dt*g per
    step and moves by dt*(v + dt*g)."""
    s = _scene(seed=5)
    s.plane_x = np.zeros((0, 2))
    s.plane_n = np.zeros((0, 2))
    s.plane_lower = np.zeros(0)
    s.plane_behavior = np.zeros(0, np.int32)
    s.q[:, 1] += 0.1
    v0 = np.array([0.3, -0.2])
    s.v[:] = v0
    g = s.gravity
    dt = 1e-4
    o = oracle.OracleMpm(s)
    q0 = s.q.copy()
    n = 10
    o.flow(dt, n)
    p = o.points()
    k = np.arange(1, n + 1)
    dq = dt * (n * v0[None] + dt * k.sum() * g[None])
    assert np.abs(p["v"] - (v0 + n * dt * g)).max() < 1e-13
    assert np.abs(p["q"] - (q0 + dq)).max() < 1e-13
    assert np.abs(p["J"] - 1).max() < 1e-13


@pytest.mark.parametrize("basis", BASES)
def test_velocity_gradient_of_linear_field_is_exact(basis):
    s = _scene(basis, seed=6)
    o = oracle.OracleMpm(s)
    A = np.array([[0.3, -1.1], [0.7, 0.2]])
    c = np.array([0.05, -0.02])
    h = s.cell_width
    nx = s.cell_count[0] + 1
    idx = np.arange(o.G)
    X = np.stack([(idx % nx) * h, (idx // nx) * h], 1)
    o.vel[:] = (X @ A.T + c).reshape(-1)
    o.call("compute_vel_grad")
    Lp = o.vel_grad.reshape(-1, 4)
    # column-major 2x2: (L00, L10, L01, L11); L = dv/dx = A
    assert np.abs(Lp[:, 0] - A[0, 0]).max() < 1e-9
    assert np.abs(Lp[:, 1] - A[1, 0]).max() < 1e-9
    assert np.abs(Lp[:, 2] - A[0, 1]).max() < 1e-9
    assert np.abs(Lp[:, 3] - A[1, 1]).max() < 1e-9


# ---- planes ----------------------------------------------------------------------

@pytest.mark.parametrize("behavior", [0, 1])
def test_point_plane_closed_form(behavior):
    s = _scene(seed=7)
    s.plane_behavior[:] = behavior
    o = oracle.OracleMpm(s)
    n = o.N
    rng = np.random.default_rng(7)
    # floor plane y = 0.01 with n = (0, 1): put points below it, moving down
    q = o.q.reshape(-1, 2)
    v = o.v.reshape(-1, 2)
    q[:, 1] = 0.01 - rng.uniform(0, 1e-3, n)
    q[:, 0] = rng.uniform(0.02, 0.2, n)
    v[:, 0] = rng.normal(size=n)
    v[:, 1] = -np.abs(rng.normal(size=n))
    vx = v[:, 0].copy()
    o.call("point_plane_collisions")
    assert np.abs(q[:, 1] - 0.01).max() < 1e-15
    if behavior == 0:
        assert np.all(v[:, 1] == 0.0) and np.array_equal(v[:, 0], vx)
    else:
        assert np.all(v == 0.0)


def test_grid_plane_sliding_removes_inward_normal_velocity():
    s = _scene(seed=8)
    o = oracle.OracleMpm(s)
    rng = np.random.default_rng(8)
    G = o.G
    o.mass[:] = rng.uniform(0.5, 1.5, G)
    mn = o.mom_new.reshape(-1, 2)
    mn[:] = rng.normal(size=(G, 2))
    before = mn.copy()
    o.call("grid_plane_collisions")
    h = s.cell_width
    nx = s.cell_count[0] + 1
    idx = np.arange(G)
    X = np.stack([(idx % nx) * h, (idx // nx) * h], 1)
    floor = X[:, 1] <= 0.01 + 1e-15
    wall = X[:, 0] <= 0.01 + 1e-15
    free = ~(floor | wall)
    assert np.array_equal(mn[free], before[free])
    only_floor = floor & ~wall
    down = only_floor & (before[:, 1] < 0)
    assert np.all(np.abs(mn[down, 1]) < 1e-15)
    assert np.allclose(mn[down, 0], before[down, 0], rtol=1e-15, atol=1e-15)
    up = only_floor & (before[:, 1] > 0)
    assert np.array_equal(mn[up], before[up])


# ---- constitutive model ----------------------------------------------------------

def _det(b):
    return b[0] * b[3] - b[1] * b[2]


def test_return_map_invariants():
    L = oracle.lib()
    rng = np.random.default_rng(9)
    kappa, mu, alpha = 8.3e5, 3.8e5, 0.33
    yielded = 0
    for _ in range(2000):
        J = 1.0 + rng.uniform(-0.03, 0.01)
        a, c = rng.uniform(-0.06, 0.06, 2)
        bbar = np.array([1 + a, c, c, (1 + c * c) / (1 + a)])
        bs = bbar * J
        out = np.zeros(4)
        L.oracle_plastic_correction(_dp(bs), kappa, mu, alpha, None, _dp(out))
        Jr = np.sqrt(_det(bs))
        ys = max(0.0, 0.5 * alpha * kappa * (1 + Jr) * (1 - Jr))
        tr = 0.5 * (out[0] + out[3])
        dev = out - np.array([tr, 0, 0, tr])
        assert out[1] == out[2]
        assert abs(_det(out) - _det(bs)) <= 1e-12 * _det(bs)
        dev_s = bs - np.array([0.5 * (bs[0] + bs[3]), 0, 0, 0.5 * (bs[0] + bs[3])])
        if mu * np.linalg.norm(dev_s) / Jr > ys:
            yielded += 1
            # on the yield surface after the return
            assert abs(mu * np.linalg.norm(dev) / Jr - ys) <= 1e-9 * max(ys, 1.0)
            if J > 1.0:
                assert np.abs(dev).max() < 1e-15  # no tension: isotropic
        else:
            assert np.array_equal(out, bs)
    assert yielded > 100


def test_tau_closed_form_and_no_tension():
    L = oracle.lib()
    rng = np.random.default_rng(10)
    kappa, mu = 8.3e5, 3.8e5
    for _ in range(200):
        J = 1.0 + rng.uniform(-0.05, 0.0)
        a, c = rng.uniform(-0.05, 0.05, 2)
        bbar = np.array([1 + a, c, c, (1 + c * c) / (1 + a)])
        tau = np.zeros(4)
        L.oracle_compute_tau(_dp(bbar * J), kappa, mu, 0, _dp(tau))
        tr = 0.5 * (bbar[0] + bbar[3])
        expect = 0.5 * kappa * (J * J - 1) * np.array([1, 0, 0, 1]) + mu * (bbar - [tr, 0, 0, tr])
        assert np.abs(tau - expect).max() <= 1e-9 * np.abs(expect).max()
    # the point-level stress is zero for J > 1
    s = _scene(seed=11)
    s.J[:] = 1.001
    o = oracle.OracleMpm(s)
    o.call("compute_cauchy_stress")
    assert np.all(o.sigma == 0.0)


def test_elastic_prediction_matrix_form():
    L = oracle.lib()
    rng = np.random.default_rng(12)
    for _ in range(100):
        b = rng.normal(size=(2, 2))
        b = b @ b.T + np.eye(2)
        Lm = rng.normal(size=(2, 2))
        dt = 1e-3
        out = np.zeros(4)
        L.oracle_elastic_prediction(_dp(np.asfortranarray(b).reshape(-1, order="F").copy()),
                                    _dp(Lm.reshape(-1, order="F").copy()), dt, _dp(out))
        expect = b + dt * (Lm @ b + b @ Lm.T)
        assert np.abs(out.reshape(2, 2, order="F") - expect).max() < 1e-14
    # a pure spin leaves the identity unchanged
    W = np.array([[0.0, -2.0], [2.0, 0.0]])
    out = np.zeros(4)
    L.oracle_elastic_prediction(_dp(np.array([1.0, 0, 0, 1.0])), _dp(W.reshape(-1, order="F").copy()),
                                1e-2, _dp(out))
    assert np.array_equal(out, [1.0, 0.0, 0.0, 1.0])
