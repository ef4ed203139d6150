"""libhybridgrains_amd: MI355X (gfx950) explicit-step engine for libHybridGrains
scenes -- MPM (mpmgrains2d ExplicitIntegrator), circle DEM (rigidbody2d
PenaltyImpactFrictionMap) and the hybridgrains2dnew coupling.

The compute path is libhg.so (hand-written HIP, built in-tree); this package is
its Python surface.  There is no CPU fallback.
"""
from ._lib import (  # noqa: F401
    HG_BASIS_CUBIC,
    HG_BASIS_LINEAR,
    HG_BASIS_UGIMP_LINEAR,
    HG_FP32,
    HG_FP64,
    HG_PLANE_SLIDING,
    HG_PLANE_STICKING,
    Context,
    HgError,
)
from .mpm import GpuMpm, MpmScene  # noqa: F401
from .dem import DemScene, GpuDem  # noqa: F401
from .hybrid import GpuHybrid  # noqa: F401

__version__ = "0.1.0"
