"""Random pixel subset for fitting, mirroring analysis/Mask.py:10-105."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ..lib import LMObject
from .Hologram import Hologram


@dataclass
class Mask(LMObject):
    """Selects a random ``fraction`` of the pixels of ``shape``, minus the
    ``exclude`` pixels; regenerated whenever one of them changes.  ``apply``
    returns (data, coordinates) of the selected pixels and draws a fresh
    subset each call (analysis/Mask.py:82-105)."""

    shape: tuple | None = None
    fraction: float = 0.1
    exclude: np.ndarray | None = None

    __hash__ = None

    def __post_init__(self) -> None:
        self._mask = np.empty((0, 0), dtype=bool)
        self.update()

    def __setattr__(self, prop, value) -> None:
        super().__setattr__(prop, value)
        if prop in ('shape', 'fraction', 'exclude') and hasattr(self, '_mask'):
            self.update()

    def __call__(self) -> np.ndarray:
        return self._mask

    @property
    def properties(self) -> dict:
        return dict(fraction=self.fraction)

    @properties.setter
    def properties(self, properties: dict) -> None:
        LMObject.properties.fset(self, properties)

    def _select(self) -> None:
        if self.shape is None:
            return
        self._mask = np.random.rand(*self.shape) < self.fraction

    def update(self) -> None:
        if self.shape is None:
            return
        self._select()
        if self.exclude is not None and self.exclude.shape == self._mask.shape:
            self._mask[self.exclude] = False

    def apply(self, hologram: Hologram):
        self.shape = tuple(hologram.shape)
        m = self._mask
        return host_data(hologram.data)[m], hologram.coordinates[:, m]


def host_data(data) -> np.ndarray:
    """The pixels of a crop as a numpy array (a copy when they live in
    device memory, for the single-feature fits and estimates)."""
    if hasattr(data, 'is_cuda') and data.is_cuda:
        return data.cpu().numpy()
    return np.asarray(data)
