"""Two touching identical spheres, mirroring theory/Dimer.py:7-78.

The spheres sit at +- a_p / magnification pixels from the cluster centre
along the axis (cos theta cos phi, cos theta sin phi, sin theta).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ..lib import LMObject
from .Cluster import Cluster
from .Sphere import Sphere


@dataclass
class Dimer(Cluster):
    a_p: float = 0.75
    n_p: float = 1.45
    k_p: float = 0.
    theta: float = np.pi / 4.
    phi: float = 0.
    magnification: float = 0.048

    __hash__ = None
    __eq__ = LMObject.__eq__

    def __post_init__(self) -> None:
        p = dict(a_p=self.a_p, n_p=self.n_p, k_p=self.k_p)
        self.particles = [Sphere(**p), Sphere(**p)]
        self.update_positions()

    def __setattr__(self, key, value) -> None:
        super().__setattr__(key, value)
        if key in ('a_p', 'n_p', 'k_p') and hasattr(self, 'particles'):
            for p in self.particles:
                setattr(p, key, value)
        if key in ('a_p', 'theta', 'phi', 'magnification'):
            self.update_positions()

    @property
    def properties(self) -> dict:
        return {'x_p': self.x_p, 'y_p': self.y_p, 'z_p': self.z_p,
                'a_p': self.a_p, 'n_p': self.n_p, 'k_p': self.k_p,
                'theta': self.theta, 'phi': self.phi}

    @properties.setter
    def properties(self, properties: dict) -> None:
        LMObject.properties.fset(self, properties)

    def update_positions(self) -> None:
        """Sphere positions from the orientation and radius
        (theory/Dimer.py:64-78)."""
        if len(getattr(self, 'particles', [])) != 2:
            return
        axis = np.array([np.cos(self.theta) * np.cos(self.phi),
                         np.cos(self.theta) * np.sin(self.phi),
                         np.sin(self.theta)])
        offset = (self.a_p / self.magnification) * axis
        self.particles[0].r_p = offset
        self.particles[1].r_p = -offset
