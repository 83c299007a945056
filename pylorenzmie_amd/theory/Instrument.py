"""Instrument parameters, mirroring theory/Instrument.py:8-70."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ..lib import LMObject


@dataclass
class Instrument(LMObject):
    """In-line holographic microscope: wavelength [um], magnification
    [um/px], numerical aperture, noise (residual scale) and medium index.
    Defaults and property order as theory/Instrument.py:32-44."""

    wavelength: float = 0.447
    magnification: float = 0.048
    numerical_aperture: float = 1.45
    noise: float = 0.05
    n_m: float = 1.340

    __hash__ = None
    __eq__ = LMObject.__eq__

    @property
    def properties(self) -> dict:
        return {'n_m': self.n_m,
                'wavelength': self.wavelength,
                'magnification': self.magnification,
                'numerical_aperture': self.numerical_aperture,
                'noise': self.noise}

    @properties.setter
    def properties(self, properties: dict) -> None:
        LMObject.properties.fset(self, properties)

    def wavenumber(self, in_medium: bool = True, scaled: bool = True) -> float:
        """2 pi / lambda [* n_m] [* magnification]; Instrument.py:46-70."""
        k = 2. * np.pi / self.wavelength
        if in_medium:
            k *= self.n_m
        if scaled:
            k *= self.magnification
        return k
