"""Seeded synthetic scene generators for the benchmark and test workloads.

Parameter mappings (the reference takes shear/bulk moduli and a
Drucker-Prager alpha, not E, nu and a friction angle):
  mu = E / (2 (1 + nu)),  kappa = E / (3 (1 - 2 nu))
  alpha_DP = sqrt(2/3) * 2 sin(phi) / (3 - sin(phi))
The reference's yield test is ||dev tau|| <= alpha kappa (1 - J^2) / 2
(mpmgrains2d/ConstitutiveModel.cpp:46-55), i.e. alpha times the Kirchhoff
pressure; the mapping above is the Drucker-Prager cone matched to the
Mohr-Coulomb angle phi in plane strain.
"""
from __future__ import annotations

import math

import numpy as np

from ._lib import HG_BASIS_UGIMP_LINEAR, HG_PLANE_SLIDING
from .mpm import MpmScene


def lame_from_young(E: float, nu: float) -> tuple[float, float]:
    return E / (2.0 * (1.0 + nu)), E / (3.0 * (1.0 - 2.0 * nu))


def dp_alpha_from_friction_angle(phi_deg: float) -> float:
    s = math.sin(math.radians(phi_deg))
    return math.sqrt(2.0 / 3.0) * 2.0 * s / (3.0 - s)


def column_collapse(
    *,
    domain=(1.0, 0.5),
    h: float = 0.005,
    column=(0.2, 0.2),
    ppc_per_dim: int = 2,
    jitter: float = 0.25,
    rho: float = 2500.0,
    E: float = 1.0e6,
    nu: float = 0.3,
    phi_deg: float = 30.0,
    pic_alpha: float = 0.999,
    gravity=(0.0, -9.81),
    wall_cells: int = 2,
    seed: int = 0,
    basis: int = HG_BASIS_UGIMP_LINEAR,
) -> MpmScene:
    """Granular column against the left wall, on the floor.

    The grid spans [0, domain]; the floor and left wall are sliding static
    planes placed ``wall_cells`` cells inside it so the stencil never leaves
    the grid.  Points sit on the reference's rectangle lattice (offsets
    (k + 0.5) h / n inside each cell, InitialSimulationState.cpp:94-99) plus a
    seeded uniform jitter of +-``jitter`` * h; hl = h / (2 n) and the volume
    is (2 hl)^2, as rasterizeRectangle sets them (:130-142).
    """
    rng = np.random.default_rng(seed)
    ncx = int(round(domain[0] / h))
    ncy = int(round(domain[1] / h))
    x0 = wall_cells * h
    y0 = wall_cells * h
    n = ppc_per_dim
    cw = int(round(column[0] / h))
    ch = int(round(column[1] / h))
    ii = np.arange(cw * n)
    jj = np.arange(ch * n)
    X, Y = np.meshgrid(x0 + (ii + 0.5) * h / n, y0 + (jj + 0.5) * h / n)
    q = np.stack([X.ravel(), Y.ravel()], axis=1)
    if jitter:
        q = q + rng.uniform(-jitter * h, jitter * h, size=q.shape)
        q[:, 0] = np.maximum(q[:, 0], x0)
        q[:, 1] = np.maximum(q[:, 1], y0)
    N = q.shape[0]
    hl = 0.5 * h / n
    vol = hl * hl * 4.0
    mu, kappa = lame_from_young(E, nu)
    s = MpmScene(
        grid_min=np.array([0.0, 0.0]),
        cell_width=h,
        cell_count=np.array([ncx, ncy]),
        shear_modulus=mu,
        bulk_modulus=kappa,
        dp_alpha=dp_alpha_from_friction_angle(phi_deg),
        pic_alpha=pic_alpha,
        gravity=np.asarray(gravity, float),
        q=q,
        v=np.zeros((N, 2)),
        m=np.full(N, rho * vol),
        volume=np.full(N, vol),
        hl=np.full(N, hl),
        be_bar=np.tile(np.array([1.0, 0.0, 0.0, 1.0]), (N, 1)),
        J=np.ones(N),
        basis=basis,
    )
    s.add_plane([x0, 0.0], [1.0, 0.0], HG_PLANE_SLIDING)
    s.add_plane([0.0, y0], [0.0, 1.0], HG_PLANE_SLIDING)
    return s


def mpm_config0(seed: int = 0) -> MpmScene:
    """Plain MPM column collapse: 1.0 x 0.5 m, h = 5 mm, 0.2 x 0.2 m column,
    2 x 2 points per cell, 6,400 points, E = 1e6, nu = 0.3, phi = 30 deg."""
    return column_collapse(seed=seed)


def mpm_1m(seed: int = 0) -> MpmScene:
    """MPM part of the 1M-point hybrid scene: 2000 x 1000 cells at h = 1 mm,
    0.5 x 0.5 m column at 4 points per cell (1,000,000 points)."""
    return column_collapse(domain=(2.0, 1.0), h=0.001, column=(0.5, 0.5), seed=seed)


def mpm_batch_scenes(nscenes: int = 64, seed: int = 0, n_min: float = 5e3, n_max: float = 5e5,
                     h: float = 0.001) -> list[MpmScene]:
    """Independent column collapses of uneven size: point count log-uniform in
    [n_min, n_max], aspect (height / width) in [0.5, 4], phi in [20, 40] deg,
    E log-uniform in [1e5, 1e7] Pa, nu = 0.3, 4 points per cell."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nscenes):
        N = math.exp(rng.uniform(math.log(n_min), math.log(n_max)))
        aspect = rng.uniform(0.5, 4.0)
        phi = rng.uniform(20.0, 40.0)
        E = math.exp(rng.uniform(math.log(1e5), math.log(1e7)))
        cells = N / 4.0
        wc = max(4, int(round(math.sqrt(cells / aspect))))
        hc = max(4, int(round(wc * aspect)))
        # room for the collapse: run-out of a few column heights
        ncx = wc + 3 * hc + 8
        ncy = hc + 8
        out.append(column_collapse(domain=(ncx * h, ncy * h), h=h, column=(wc * h, hc * h),
                                   E=E, phi_deg=phi, seed=seed * 1000 + k))
    return out


# ---- DEM ---------------------------------------------------------------------

def damping_from_restitution(e: float, kn: float, m_eff: float) -> float:
    """gamma_n for the reference's normal law F = -kn d - gamma_n v_n / 2
    (PenaltyImpactFrictionMap.cpp:650-651): with zeta = -ln e / sqrt(pi^2 +
    ln^2 e), the linear spring-dashpot m* x'' = -kn x - (gamma_n/2) x' has
    restitution e when gamma_n = 4 zeta sqrt(kn m*)."""
    le = math.log(e)
    zeta = -le / math.sqrt(math.pi ** 2 + le * le)
    return 4.0 * zeta * math.sqrt(kn * m_eff)


def dem_column(
    *,
    n_bodies: int = 20000,
    width: float = 0.15,
    spacing: float = 2.2e-3,
    r_mean: float = 1.0e-3,
    r_spread: float = 0.1e-3,
    jitter: float = 0.05e-3,
    rho: float = 2500.0,
    kn: float = 1.0e5,
    restitution: float = 0.3,
    mu: float = 0.3,
    box_width: float = 1.2,
    gravity=(0.0, -9.81),
    seed: int = 0,
):
    """Circle column collapse: radii uniform in r_mean +- r_spread on a
    hexagonal lattice of the given spacing with seeded jitter, filled row by
    row over ``width`` until ``n_bodies`` circles are placed (a 0.15 m wide
    column of 20k circles is about 0.56 m tall).  Floor and two walls; the
    right wall stands ``box_width`` from the left one so the column can
    collapse.  kt = kn and gamma_t = gamma_n; gamma_n from the restitution of
    a pair of mean circles."""
    from .dem import DemScene, circle_mass

    rng = np.random.default_rng(seed)
    per_row = int(width / spacing)
    dy = spacing * math.sqrt(3.0) / 2.0
    rows = (n_bodies + per_row - 1) // per_row
    if spacing * (per_row + 0.5) + r_mean + r_spread > box_width:
        raise ValueError("the column is wider than the box: bodies would start behind the right wall")
    # lattice sites row by row, the first n_bodies of them
    k = np.arange(n_bodies)
    j, i = k // per_row, k % per_row
    pts = np.stack([spacing * (i + 0.5) + 0.5 * spacing * (j % 2),
                    r_mean + r_spread + 1e-4 + j * dy], axis=1)
    q = np.zeros((n_bodies, 3))
    q[:, :2] = pts + rng.uniform(-jitter, jitter, size=(n_bodies, 2))
    q[:, 1] = np.maximum(q[:, 1], r_mean + r_spread)
    q[:, 0] = np.maximum(q[:, 0], r_mean + r_spread)
    radius = rng.uniform(r_mean - r_spread, r_mean + r_spread, n_bodies)
    mass = circle_mass(rho, radius)
    m_pair = 0.5 * rho * math.pi * r_mean * r_mean
    gn = damping_from_restitution(restitution, kn, m_pair)
    s = DemScene(q=q, v=np.zeros((n_bodies, 3)), radius=radius, mass=mass,
                 fixed=np.zeros(n_bodies, np.uint8), gravity=np.asarray(gravity, float),
                 kn=kn, kt=kn, gamman=gn, gammat=gn, mu=mu)
    s.add_plane([0.0, 0.0], [0.0, 1.0])
    s.add_plane([0.0, 0.0], [1.0, 0.0])
    s.add_plane([box_width, 0.0], [-1.0, 0.0])
    return s


def dem_rosettes(nx: int = 6, ny: int = 4, R: float = 4e-3, r: float = 1e-3,
                 overlap: float = 0.05e-3, kn: float = 1.0e5, seed: int = 0):
    """Polydisperse contact clusters: nx x ny large circles of radius R, each
    ringed by as many circles of radius r as fit around it (15 for the
    defaults), every one overlapping the large circle by ``overlap``, in the
    air above a floor and beside a wall.  Each large circle starts with more
    than eight partners; the ring's small circles do not touch each other."""
    from .dem import DemScene, circle_mass

    rng = np.random.default_rng(seed)
    n = int(2.0 * math.pi * (R + r) / (2.0 * r))
    pitch = 2.0 * (R + 2.0 * r) + 2e-3
    pts, rad = [], []
    for a in range(nx):
        for b in range(ny):
            c = np.array([R + 2 * r + 1e-3 + a * pitch, R + 2 * r + 5e-3 + b * pitch])
            pts.append(c)
            rad.append(R)
            th0 = rng.uniform(0.0, 2.0 * math.pi)
            for i in range(n):
                th = th0 + 2.0 * math.pi * i / n
                pts.append(c + (R + r - overlap) * np.array([math.cos(th), math.sin(th)]))
                rad.append(r)
    q = np.zeros((len(pts), 3))
    q[:, :2] = np.asarray(pts)
    radius = np.asarray(rad)
    gn = damping_from_restitution(0.3, kn, 0.5 * 2500.0 * math.pi * r * r)
    s = DemScene(q=q, v=np.zeros_like(q), radius=radius, mass=circle_mass(2500.0, radius),
                 fixed=np.zeros(len(pts), np.uint8), gravity=np.array([0.0, -9.81]), kn=kn,
                 kt=kn, gamman=gn, gammat=gn, mu=0.3)
    s.add_plane([0.0, 0.0], [0.0, 1.0])
    s.add_plane([0.0, 0.0], [1.0, 0.0])
    return s


# ---- hybrid --------------------------------------------------------------------

def hybrid_column(
    *,
    domain=(0.4, 0.3),
    h: float = 0.005,
    column=(0.1, 0.08),
    band_cells: int = 2,
    grain_layer: float = 0.04,
    r_mean: float = 1.25e-3,
    r_spread: float = 0.1,
    rho: float = 2500.0,
    E: float = 1.0e6,
    phi_deg: float = 30.0,
    kn: float = 1.0e4,
    restitution: float = 0.3,
    mu: float = 0.3,
    seed: int = 0,
):
    """A continuum column with a layer of grains on top.

    The MPM column occupies the grid's floor corner up to height
    ``column[1]``; circles fill the band from ``band_cells`` cells below the
    column's free surface up to ``grain_layer`` above it.  Where the two
    overlap (the coupling band) the grains and the points carry mass weight
    1/2, as the reference's zone resampling assigns in a transition band
    (HybridGrains2DSim.cpp:512-531); grains whose stencil nodes all carry
    continuum mass are coupled each step.  Returns (mpm_scene, dem_scene,
    weights), the DEM masses already scaled by the weights
    (RigidBody2DState::updateCurrentMassUsingMassWeights).
    """
    from .dem import DemScene, circle_mass

    rng = np.random.default_rng(seed)
    mpm = column_collapse(domain=domain, h=h, column=column, seed=seed, E=E, phi_deg=phi_deg)
    y_top = 2 * h + column[1]
    y_band = y_top - band_cells * h
    inband = mpm.q[:, 1] > y_band
    mpm.m[inband] *= 0.5
    mpm.volume[inband] *= 0.5
    spacing = 2.0 * r_mean * (1.0 + r_spread) + 1e-5
    dy = spacing * math.sqrt(3.0) / 2.0
    x0 = 2 * h + r_mean * (1 + r_spread)
    pts = []
    j = 0
    y = y_band + r_mean
    while y < y_top + grain_layer:
        off = 0.5 * spacing if (j % 2) else 0.0
        x = x0 + off
        while x < 2 * h + column[0] - r_mean:
            pts.append((x, y))
            x += spacing
        y += dy
        j += 1
    B = len(pts)
    q = np.zeros((B, 3))
    q[:, :2] = np.asarray(pts) + rng.uniform(-0.02 * r_mean, 0.02 * r_mean, size=(B, 2))
    radius = rng.uniform(r_mean * (1 - r_spread), r_mean * (1 + r_spread), B)
    mass0 = circle_mass(rho, radius)
    weights = np.where(q[:, 1] < y_top, 0.5, 1.0)
    m_pair = 0.5 * rho * math.pi * r_mean * r_mean
    gn = damping_from_restitution(restitution, kn, m_pair)
    dem = DemScene(q=q, v=np.zeros((B, 3)), radius=radius, mass=mass0 * weights[:, None],
                   fixed=np.zeros(B, np.uint8), gravity=mpm.gravity.copy(), kn=kn, kt=kn,
                   gamman=gn, gammat=gn, mu=mu)
    dem.add_plane([0.0, 2 * h], [0.0, 1.0])
    dem.add_plane([2 * h, 0.0], [1.0, 0.0])
    return mpm, dem, weights


def hybrid_1m(seed: int = 0):
    """The 1M-point hybrid scene: the 1M-point MPM column (2000 x 1000 cells
    at h = 1 mm, 0.5 x 0.5 m at 4 points per cell) with a discrete zone of
    circles (radius 0.5 mm +- 10 %) 0.1 m deep at its free surface,
    overlapping the continuum by two cells.  A 10-cell-deep zone holds about
    5k such circles and 200k would need a zone most of the column deep, so
    the zone is fixed by depth (0.1 m, about 48k circles).  kn = 1e5 N/m,
    restitution 0.3, mu = 0.3."""
    return hybrid_column(domain=(2.0, 1.0), h=0.001, column=(0.5, 0.5), grain_layer=0.1,
                         r_mean=0.5e-3, r_spread=0.1, kn=1.0e5, seed=seed)


def served_requests(n: int, seed: int = 0, rate: float = 8.0, n_min: float = 1e4,
                    n_max: float = 2e5, steps_min: int = 200, steps_max: int = 1000,
                    mix=(0.5, 0.3, 0.2), dem_rows: int = 300):
    """A Poisson stream of scene requests: (arrival_s, kind, payload, dt,
    nsteps), kind 'mpm' (payload MpmScene), 'dem' (DemScene) or 'hybrid'
    ((MpmScene, DemScene, weights)).  Sizes are log-uniform in [n_min, n_max]
    particles (points + bodies) and horizons uniform in [steps_min,
    steps_max]; each request is generated from its own seed.  DEM columns
    are at most dem_rows rows tall and as wide as their count needs."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n):
        t += rng.exponential(1.0 / rate)
        u = rng.uniform()
        kind = "mpm" if u < mix[0] else ("dem" if u < mix[0] + mix[1] else "hybrid")
        N = math.exp(rng.uniform(math.log(n_min), math.log(n_max)))
        nsteps = int(rng.integers(steps_min, steps_max + 1))
        rs = seed * 100003 + k
        if kind == "mpm":
            h = 0.002
            wc = max(8, int(round(math.sqrt(N / 4.0))))
            hc = max(8, int(round(N / 4.0 / wc)))
            s = column_collapse(domain=((wc + 3 * hc + 8) * h, (hc + 8) * h), h=h,
                                column=(wc * h, hc * h), seed=rs)
            out.append((t, kind, s, 1e-5, nsteps))
        elif kind == "dem":
            # at most dem_rows rows: a taller column of 1 mm circles at
            # kn = 1e5 N/m overlaps at its foot by more than a radius; the box
            # leaves room to run out four column heights past the column
            spacing = 2.2e-3
            per_row = max(8, -(-int(N) // dem_rows))
            rows = -(-int(N) // per_row)
            width = spacing * (per_row + 0.25)
            s = dem_column(n_bodies=int(N), width=width,
                           box_width=width + spacing + 4.0 * rows * spacing * math.sqrt(3.0) / 2.0,
                           seed=rs)
            out.append((t, kind, s, 1e-5, nsteps))
        else:
            h = 0.002
            wc = max(8, int(round(math.sqrt(N / 4.0))))
            hc = max(8, int(round(0.8 * N / 4.0 / wc)))
            m, d, w = hybrid_column(domain=((wc + 3 * hc + 8) * h, (hc + 40) * h), h=h,
                                    column=(wc * h, hc * h), grain_layer=0.1 * hc * h,
                                    r_mean=2.5e-3 * h / 0.005 * 1.0, seed=rs)
            out.append((t, kind, (m, d, w), 1e-5, nsteps))
    return out


def write_hybrid_column_xml(folder, *, domain=(2.0, 1.0), h: float = 0.001, column=(0.5, 0.5),
                            r_mean: float = 0.5e-3, r_spread: float = 0.1, rho: float = 2500.0,
                            E: float = 1.0e6, nu: float = 0.3, phi_deg: float = 30.0,
                            kn: float = 1.0e5, restitution: float = 0.3, mu: float = 0.3,
                            level: float = 0.1, half_thickness_cells: int = 2,
                            dt_den: int = 200000, resample_every: int = 100, seed: int = 0):
    """The hybrid column collapse as the reference's three scene files
    (hybrid.xml, mpm.xml, dem.xml) in folder, for hybridgrains2d_cli: the
    continuum rectangle over the whole column at 2 x 2 points per cell, and
    circles of radius r_mean +- r_spread on a hexagonal packing over the whole
    column, from which the zone resampling keeps those of the discrete and
    hybrid zones (the zone boundary `level` below the free surface, a hybrid
    band of half_thickness_cells cells).  The resampling runs every
    resample_every steps (avoidAVoidFreq = resample_every - 1); dt = 1 /
    dt_den.  At the defaults: 1,000,000 points and about 180k circles before
    the first resampling.  Returns the path of hybrid.xml."""
    from pathlib import Path

    folder = Path(folder)
    folder.mkdir(parents=True, exist_ok=True)
    rng = np.random.default_rng(seed)
    x0 = y0 = 2 * h
    cw, chh = column
    mu_s, kappa = lame_from_young(E, nu)
    alpha = dp_alpha_from_friction_angle(phi_deg)
    dt = f"1 / {int(dt_den)}"
    (folder / "mpm.xml").write_text(f"""<mpmgrains2d>
  <end_time t="1.0"/>
  <grid min="0.0 0.0" max="{domain[0]!r} {domain[1]!r}" cell_width="{h!r}"/>
  <integrator dt="{dt}" alpha="0.999" shape_function_order="1" shape_function_category="uGIMP"/>
  <material shear_modulus="{mu_s!r}" bulk_modulus="{kappa!r}" alpha="{alpha!r}"/>
  <rectangle min="{x0!r} {y0!r}" max="{x0 + cw!r} {y0 + chh!r}" density="{rho!r}" cell_samples_per_dim="2" vel="0.0 0.0" omega="0.0 0.0"/>
  <near_earth_gravity f="0.0 -9.81"/>
  <static_plane x="{x0!r} 0.0" n="1.0 0.0" boundary_behavior="sliding"/>
  <static_plane x="0.0 {y0!r}" n="0.0 1.0" boundary_behavior="sliding"/>
</mpmgrains2d>
""")
    spacing = 2.0 * r_mean * (1.0 + r_spread) + 1e-5
    dy = spacing * math.sqrt(3.0) / 2.0
    xs, ys = [], []
    j = 0
    y = y0 + r_mean * (1 + r_spread)
    while y < y0 + chh - r_mean:
        x = x0 + r_mean * (1 + r_spread) + (0.5 * spacing if j % 2 else 0.0)
        row = np.arange(x, x0 + cw - r_mean, spacing)
        xs.append(row)
        ys.append(np.full(row.size, y))
        y += dy
        j += 1
    px, py = np.concatenate(xs), np.concatenate(ys)
    radius = rng.uniform(r_mean * (1 - r_spread), r_mean * (1 + r_spread), px.size)
    m_pair = 0.5 * rho * math.pi * r_mean * r_mean
    gn = damping_from_restitution(restitution, kn, m_pair)
    lines = ["<rigidbody2d_scene>",
             '  <collision_detection broadphase="spatial_grid" cell_scale="1.0" bandwidth_reduction="none"/>',
             f'  <penalty_impact_map dt="{dt}" kn="{kn!r}" gamman="{gn!r}" kt="{kn!r}" '
             f'gammat="{gn!r}" mu="{mu!r}"/>',
             '  <end_time t="1.0"/>',
             '  <near_earth_gravity f="0.0 -9.81"/>',
             f'  <static_plane x="0.0 {y0!r}" n="0 1"/>',
             f'  <static_plane x="{x0!r} 0.0" n="1 0"/>']
    lines += [f'  <geometry type="circle" r="{r!r}"/>' for r in radius.tolist()]
    lines += [f'  <rigid_body x="{a!r} {b!r}" theta="0" v="0 0" omega="0" rho="{rho!r}" '
              f'geo_idx="{k}" fixed="0"/>' for k, (a, b) in enumerate(zip(px.tolist(), py.tolist()))]
    lines.append("</rigidbody2d_scene>")
    (folder / "dem.xml").write_text("\n".join(lines) + "\n")
    (folder / "hybrid.xml").write_text(f"""<hybridgrains2d_scene>
  <discrete_sim file_name="dem.xml"/>
  <continuum_sim file_name="mpm.xml"/>
  <time_step dt="{dt}"/>
  <end_time t="1.0"/>
  <integrator_type type="coupled"/>
  <integrator_style style="prediction_correction_advection_node_node"/>
  <post_stabilization_settings stabilization_order="after_advection" stabilization_type="fix_penetration" velocity_blending="0.0" spring_constant="0.0" tolerance="1.0e-5" max_iter="1000"/>
  <hybrid_kinematic_script type="none"/>
  <coupling_force k="0.0" hw="1.0" order="1" test_number="6" enrichment_flags="00000" fixed_rzone="0 0 0 0" lmcoupling="1" LMEnableFixedRZone="0" rzoneHalfThickness="{half_thickness_cells * h!r}" demRMean="{r_mean!r}" demRStd="{r_spread * r_mean / math.sqrt(3.0)!r}" avoidAVoidFreq="{int(resample_every) - 1}" phiThreshold="0.3" rzoneLevelSet="{level!r}" phi_window_size="{1.25 * h!r}" phi_samples_per_cell_side="1" level_set_cell_width="{h!r}" rho_dem="{rho!r}" homogenize_stress="0" grid_smoothing_homogenized_stress="0" homogenize_velocity="0" grid_smoothing_homogenized_velocity="0"/>
</hybridgrains2d_scene>
""")
    return folder / "hybrid.xml"
