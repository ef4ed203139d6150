"""Circle DEM scenes: the rigidbody2d RigidBody2DState (circles) and the
PenaltyImpactFrictionMap parameters, and their GPU stepping.

``DemScene`` holds what RigidBody2DSceneParser::parseXMLSceneFile
(rigidbody2dutils/RigidBody2DSceneParser.cpp:1127-1276) builds for a circle
scene; ``GpuDem`` keeps scenes resident on the GPU and steps them with
``PenaltyImpactFrictionMap::flow`` / ``flowWithWeights`` semantics
(rigidbody2d/PenaltyImpactFrictionMap.cpp:1083-1149).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


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
