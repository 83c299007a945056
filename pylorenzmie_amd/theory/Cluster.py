"""A group of particles moved as one, mirroring theory/Cluster.py:7-59.

The cluster's position is the origin (``r_0``) of every constituent
particle; LorenzMie superposes their fields in one kernel pass
(lmz_hologram_superposed).
"""
from __future__ import annotations

from dataclasses import dataclass, field

from ..lib import LMObject
from .Particle import Particle


@dataclass
class Cluster(Particle):
    """Constituent ``particles``; setting x_p, y_p, z_p or particles
    moves every constituent's r_0 to the cluster position."""

    particles: list = field(repr=False, default_factory=list)

    __hash__ = None
    __eq__ = LMObject.__eq__

    def __post_init__(self) -> None:
        self.update()

    def __setattr__(self, key, value) -> None:
        super().__setattr__(key, value)
        if key in ('x_p', 'y_p', 'z_p', 'particles'):
            self.update()

    def __len__(self) -> int:
        return len(self.particles)

    def __iter__(self):
        return iter(self.particles)

    def __getitem__(self, index: int) -> Particle:
        return self.particles[index]

    @property
    def properties(self) -> dict:
        return {'x_p': self.x_p, 'y_p': self.y_p, 'z_p': self.z_p}

    @properties.setter
    def properties(self, properties: dict) -> None:
        LMObject.properties.fset(self, properties)

    def update(self) -> None:
        """Propagate the cluster position to each constituent's r_0."""
        try:
            for particle in self.particles:
                particle.r_0 = self.r_p
        except AttributeError:      # during dataclass initialization
            pass
