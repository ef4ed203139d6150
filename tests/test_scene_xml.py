"""The XML scene readers (libhg_host.so) on the shipped scene files.

Fixtures under tests/golden/assets: the column-collapse and hourglass
<mpmgrains2d> and <hybridgrains2d_scene> files of the reference's assets, and
the first 40 circles of the column-collapse <rigidbody2d_scene> (the full
file is 1.7 MB).  Expected values are worked out by hand from the schema's
rules (MPMGrains2DParser.cpp, InitialSimulationState.cpp:45-145,
PhysicsGrid.cpp:14-55, RigidBody2DSceneParser.cpp, Rational.h:208-291), not
by rerunning the reader's arithmetic.
"""
import math
import os
from pathlib import Path

import numpy as np
import pytest

from libhybridgrains_amd import HG_BASIS_CUBIC, HG_BASIS_LINEAR, HG_BASIS_UGIMP_LINEAR
from libhybridgrains_amd import HG_PLANE_STICKING, xmlscene

ASSETS = Path(__file__).parent / "golden" / "assets"
REF_ASSETS = Path("/root/reference/assets/hybridgrains2dnew")


def test_column_collapse_mpm_scene():
    s = xmlscene.load_mpm(ASSETS / "column-collapse" / "mpm.xml")
    m = s.scene
    # dt="1 / 1000080" is the rational 1/1000080
    assert s.dt == 1.0 / 1000080.0
    assert s.end_time == 2.0
    # grid 40.4 x 7.4 at 0.2 -> 202 x 37 cells, re-centred on the same centre
    assert list(m.cell_count) == [202, 37]
    assert abs(m.grid_min[0] + 0.4) < 1e-12 and abs(m.grid_min[1] + 0.4) < 1e-12
    assert m.basis == HG_BASIS_UGIMP_LINEAR
    assert (m.shear_modulus, m.bulk_modulus, m.dp_alpha, m.pic_alpha) == (1.9e6, 4.15e6, 0.37, 0.999)
    assert list(m.gravity) == [0.0, -981.0]
    assert list(m.plane_behavior) == [HG_PLANE_STICKING] * 2
    assert np.all(np.isinf(m.plane_lower))
    # rectangle 12.4 x 6.4 at 2 samples per 0.2 cell: 62 x 32 cells x 4 points
    assert m.npoints == 62 * 32 * 4
    assert np.allclose(m.hl, 0.05) and np.allclose(m.volume, 0.01)
    assert np.allclose(m.m, 0.96 * 0.01)
    assert np.array_equal(m.be_bar, np.tile([1.0, 0.0, 0.0, 1.0], (m.npoints, 1)))
    assert np.all(m.J == 1.0) and np.all(m.v == 0.0)
    xs = np.unique(np.round(m.q[:, 0], 9))
    ys = np.unique(np.round(m.q[:, 1], 9))
    assert np.allclose(xs, 0.05 + 0.1 * np.arange(124))
    assert np.allclose(ys, 0.05 + 0.1 * np.arange(64))


def test_hourglass_mpm_seeding_order():
    """Cells then samples, row by row: the four points of the small
    hourglass block in the order rasterizeRectangle emits them."""
    s = xmlscene.load_mpm(ASSETS / "hourglass" / "mpm.xml")
    m = s.scene
    assert list(m.cell_count) == [34, 176]
    assert np.allclose(m.grid_min, [-4.25, -2.0])
    assert s.dt == 1e-6
    expect = [(-0.0625, 24.0625), (-0.0625, 24.1875), (0.0625, 24.0625), (0.0625, 24.1875)]
    assert np.allclose(m.q, expect, atol=1e-12)
    assert np.allclose(m.m, 0.96 * 4 * 0.0625 ** 2)
    assert m.plane_x.shape == (3, 2)


def test_column_collapse_dem_subset():
    s = xmlscene.load_dem(ASSETS / "column-collapse" / "dem.xml")
    d = s.scene
    assert d.nbodies == 40
    assert s.dt == 1.0 / 1000080.0 and s.end_time == 10.0
    assert (d.kn, d.kt, d.gamman, d.gammat, d.mu) == (17720938.760736626, 478465346.5398889,
                                                      4618.02, 4618.02, 0.5)
    assert list(d.gravity) == [0.0, -981.0]
    assert d.plane_x.shape == (1, 2) and list(d.plane_n[0]) == [1.0, 0.0]
    r0 = 0.05392151360279802
    assert d.radius[0] == r0
    assert d.q[0, 0] == 7.63844359474847234 and d.q[0, 2] == 1.10810381326437324
    m0 = 2.52999999999999936 * math.pi * r0 * r0
    assert d.mass[0, 0] == m0 and d.mass[0, 1] == m0
    assert d.mass[0, 2] == 0.5 * m0 * r0 * r0
    assert not d.fixed.any()
    # SpatialGrid cell: mean AABB extent (2 r) x cell_scale 5.5
    assert abs(s.cell_width - 5.5 * 2.0 * d.radius.mean()) < 1e-12


def test_hybrid_scene(tmp_path):
    # hybrid.xml names dem.xml and mpm.xml in its own folder
    d = tmp_path / "cc"
    d.mkdir()
    for f in ("hybrid.xml", "mpm.xml", "dem.xml"):
        (d / f).write_bytes((ASSETS / "column-collapse" / f).read_bytes())
    h = xmlscene.load_hybrid(d / "hybrid.xml")
    assert h.style == "prediction_correction_advection_node_node"
    assert h.dt == 1.0 / 1000080.0 and h.end_time == 2.0
    assert h.continuum.npoints == 7936 and h.discrete.nbodies == 40
    # the hybrid broad phase ignores cell_scale: mean extent x 1, at least 2 r_max
    assert abs(h.cell_width - max(2.0 * h.discrete.radius.mean(), 2.0 * h.discrete.radius.max())) < 1e-12


@pytest.mark.parametrize("text,dt", [
    ("0.00001", 1e-5), ("1/100000", 1e-5), ("1 / 1000080", 1.0 / 1000080.0), (".5", 0.5),
    ("3", 3.0), ("2/6", 1.0 / 3.0),
])
def test_rational_dt(text, dt):
    xml = f'''<mpmgrains2d><grid min="0 0" max="1 1" cell_width="0.25"/>
      <integrator dt="{text}" alpha="1" shape_function_order="1"/>
      <material shear_modulus="1" bulk_modulus="1" alpha="0"/>
      <rectangle min="0.25 0.25" max="0.5 0.5" density="1" cell_samples_per_dim="1" vel="0 0" omega="0"/>
    </mpmgrains2d>'''
    s = xmlscene.load_mpm(text=xml)
    assert s.dt == dt
    assert s.scene.basis == HG_BASIS_LINEAR  # Standard category by default


def test_basis_selection_and_errors():
    base = '''<mpmgrains2d><grid min="0 0" max="1 1" cell_width="0.25"/>
      <integrator dt="1e-3" alpha="1" shape_function_order="{o}" {cat}/>
      <material shear_modulus="1" bulk_modulus="1" alpha="0"/>
      <rectangle min="0.25 0.25" max="0.5 0.5" density="1" cell_samples_per_dim="2" vel="1 2" omega="0"/>
      {extra}
    </mpmgrains2d>'''
    s = xmlscene.load_mpm(text=base.format(o=3, cat='shape_function_category="Standard"', extra=""))
    assert s.scene.basis == HG_BASIS_CUBIC
    assert np.allclose(s.scene.v, [1.0, 2.0])
    with pytest.raises(ValueError, match="uGIMP"):
        xmlscene.load_mpm(text=base.format(o=3, cat='shape_function_category="uGIMP"', extra=""))
    with pytest.raises(ValueError, match="shape_function_order"):
        xmlscene.load_mpm(text=base.format(o=2, cat="", extra=""))
    with pytest.raises(ValueError, match="boundary_behavior"):
        xmlscene.load_mpm(text=base.format(o=1, cat="", extra='<static_plane x="0 0" n="0 1" boundary_behavior="slippery"/>'))
    # samples at 0.3125 and 0.4375 per axis; a plane keeps the points on the
    # side its normal points to and drops the ones behind it
    s = xmlscene.load_mpm(text=base.format(o=1, cat="", extra='<static_plane x="0 0.75" n="0 -1" boundary_behavior="sliding"/>'))
    assert s.scene.npoints == 4
    s = xmlscene.load_mpm(text=base.format(o=1, cat="", extra='<static_plane x="0 0.4" n="0 1" boundary_behavior="sliding"/>'))
    assert s.scene.npoints == 2 and np.allclose(s.scene.q[:, 1], 0.4375)
    # below a plane's lower bound the test stops at that plane: nothing dropped
    s = xmlscene.load_mpm(text=base.format(o=1, cat="", extra='<static_plane x="0 0.4" n="0 1" boundary_behavior="sliding" lower_bound="0.5"/>'))
    assert s.scene.npoints == 4


def test_unsupported_dem_geometry_is_rejected():
    xml = '''<rigidbody2d_scene>
      <collision_detection broadphase="spatial_grid" cell_scale="1" bandwidth_reduction="none"/>
      <penalty_impact_map dt="1e-4" kn="1" kt="1" gamman="0" gammat="0" mu="0"/>
      <geometry type="box" r="1 1"/>
    </rigidbody2d_scene>'''
    with pytest.raises(ValueError, match="circles only"):
        xmlscene.load_dem(text=xml)


@pytest.mark.skipif(not REF_ASSETS.exists(), reason="reference assets not mounted")
def test_full_shipped_column_collapse_dem():
    s = xmlscene.load_dem(REF_ASSETS / "column-collapse" / "dem.xml")
    assert s.scene.nbodies == 6694  # 6695 rigid_body lines, one inside a comment
    r = s.scene.radius
    # the hybrid scene's demRMean / demRStd describe these grains
    assert abs(r.mean() - 0.0574865702199) < 0.01 * 0.0575
    assert abs(r.std() - 0.00574890375453) < 0.1 * 0.00575
    assert 0 < s.scene.fixed.sum() < s.scene.nbodies
