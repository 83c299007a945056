"""Spherical aberration, mirroring theory/Aberrated.py:7-98.

``Aberrated(base)`` returns a subclass of a LorenzMie calculator with a
``spherical`` coefficient [wavelengths]; each particle's scattered field
is multiplied by exp(i (4 n_m^3 / k) s r^4 / (r^2 + z_p^2)^2) inside the
superposition kernel (lmz_hologram_superposed), not by a separate pass.
"""
from __future__ import annotations

from .LorenzMie import LorenzMie


def Aberrated(base_class: type) -> type:
    class AberratedLorenzMie(base_class):
        def __init__(self, *args, spherical: float = 0., **kwargs) -> None:
            super().__init__(*args, **kwargs)
            self.spherical = float(spherical)

        @property
        def properties(self) -> dict:
            return {**base_class.properties.fget(self),
                    'spherical': self.spherical}

        @properties.setter
        def properties(self, properties: dict) -> None:
            base_class.properties.fset(self, properties)

    AberratedLorenzMie.__qualname__ = f'Aberrated{base_class.__name__}'
    return AberratedLorenzMie


AberratedLorenzMie = Aberrated(LorenzMie)
