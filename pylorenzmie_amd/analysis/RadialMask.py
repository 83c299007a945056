"""Pixel subset drawn with a probability that grows away from the crop's
centre, mirroring analysis/RadialMask.py:5-33."""
from __future__ import annotations

import numpy as np

from .Mask import Mask


class RadialMask(Mask):
    """A Mask whose pixel (row, col) is kept when hypot(a + b u, a + b v)
    is at least a uniform random number, with u, v the column and row
    positions scaled to [-1, 1] and (a, b) = (0, 2f) for fraction f < 1/2,
    else (2f - 1, 2 (1 - f)) (analysis/RadialMask.py:21-33).  Below
    f = 1/2 the centre is never kept and the outer fringes are favoured;
    towards f = 1 the draw becomes uniform.  ``exclude``, ``update`` and
    ``apply`` are the Mask's."""

    def _select(self) -> None:
        if self.shape is None:
            return
        f = self.fraction
        a, b = (0., 2. * f) if f < 0.5 else (2. * f - 1., 2. * (1. - f))
        ny, nx = self.shape
        u = a + b * np.linspace(-1., 1., nx)
        v = a + b * np.linspace(-1., 1., ny)
        keep = np.hypot(v[:, None], u[None, :])
        self._mask = keep >= np.random.random_sample(self.shape)
