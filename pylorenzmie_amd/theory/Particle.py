"""Particle position container, mirroring theory/Particle.py:8-119."""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from ..lib import LMObject


@dataclass
class Particle(LMObject):
    """Position (x_p, y_p, z_p) [px] and local origin r_0 (excluded from
    properties).  A single particle iterates as a one-element sequence so
    that lists and single particles are interchangeable
    (theory/Particle.py:50-59)."""

    x_p: float = 0.
    y_p: float = 0.
    z_p: float = 100.
    x_0: float = field(repr=False, default=0.)
    y_0: float = field(repr=False, default=0.)
    z_0: float = field(repr=False, default=0.)

    __hash__ = None
    __eq__ = LMObject.__eq__

    def __len__(self) -> int:
        return 1

    def __iter__(self):
        return iter([self])

    def __getitem__(self, index: int) -> 'Particle':
        if index != 0:
            raise IndexError('Particle index out of range')
        return self

    @property
    def r_p(self) -> np.ndarray:
        return np.asarray([self.x_p, self.y_p, self.z_p])

    @r_p.setter
    def r_p(self, r_p) -> None:
        self.x_p, self.y_p, self.z_p = r_p

    @property
    def r_0(self) -> np.ndarray:
        return np.asarray([self.x_0, self.y_0, self.z_0])

    @r_0.setter
    def r_0(self, r_0) -> None:
        self.x_0, self.y_0, self.z_0 = r_0

    @property
    def properties(self) -> dict:
        return {'x_p': self.x_p, 'y_p': self.y_p, 'z_p': self.z_p}

    @properties.setter
    def properties(self, properties: dict) -> None:
        LMObject.properties.fset(self, properties)

    def ab(self, n_m=1. + 0.j, wavelength: float = 0.532) -> np.ndarray:
        """Placeholder coefficients ones((1, 2)), theory/Particle.py:97-119."""
        return np.ones((1, 2), dtype=complex)
