"""Normalized hologram with pixel coordinates, mirroring
analysis/Hologram.py:11-108."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ..lib import meshgrid


@dataclass(eq=False)
class Hologram:
    """Image ``data`` (ny, nx) normalized to background ~1, with the
    coordinates of its pixels from ``corner`` (left, top).  ``flat_data`` and
    ``flat_coordinates`` are views; ``hologram[sy, sx]`` is a
    coordinate-aware crop."""

    data: np.ndarray
    corner: tuple = (0., 0.)

    def __post_init__(self) -> None:
        self._coordinates = meshgrid(self.data.shape, corner=self.corner,
                                     flatten=False)

    def __eq__(self, other) -> bool:
        if not isinstance(other, Hologram):
            return NotImplemented
        return (np.array_equal(self.data, other.data) and
                self.corner == other.corner)

    @property
    def shape(self):
        return self.data.shape

    @property
    def coordinates(self) -> np.ndarray:
        return self._coordinates

    @property
    def flat_data(self) -> np.ndarray:
        return self.data.ravel()

    @property
    def flat_coordinates(self) -> np.ndarray:
        return self._coordinates.reshape(2, -1)

    @staticmethod
    def _from_slice(data, coordinates) -> 'Hologram':
        h = object.__new__(Hologram)
        h.data = data
        h.corner = (float(coordinates[0, 0, 0]), float(coordinates[1, 0, 0]))
        h._coordinates = coordinates
        return h

    def __getitem__(self, key) -> 'Hologram':
        sy, sx = key
        return Hologram._from_slice(self.data[key], self._coordinates[:, sy, sx])
