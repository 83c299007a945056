"""One particle's crop of a hologram, mirroring analysis/Feature.py:9-140.

A Feature is a Hologram with a LorenzMie model, a Mask that excludes bad
pixels (saturated, NaN, inf; analysis/Feature.py:36-43) and an Optimizer
on that mask, and a DEEstimator for starting values.  ``optimize()`` fits
this one feature; ``Frame.optimize`` fits all features of a frame in one
device batch.
"""
from __future__ import annotations

import numpy as np
import pandas as pd

from ..theory import LorenzMie
from .DEEstimator import DEEstimator
from .Hologram import Hologram
from .Mask import Mask, host_data
from .Optimizer import Optimizer


class Feature(Hologram):
    """Crop ``hologram`` with its generative ``model`` and fitter."""

    def __init__(self, hologram: Hologram, model: LorenzMie | None = None,
                 fixed: list[str] | None = None) -> None:
        # Hologram state copied as analysis/Feature.py:29-33 does
        self.data = hologram.data
        self.corner = hologram.corner
        self._coordinates = hologram._coordinates
        self.model = model or LorenzMie()
        m = Mask()
        m.shape = tuple(self.data.shape)
        m.exclude = _bad_pixels(self.data)
        self.estimator = DEEstimator(model=self.model, exclude=m.exclude)
        self.optimizer = Optimizer(model=self.model, mask=m)
        if fixed is not None:
            self.optimizer.fixed = fixed

    @property
    def mask(self) -> Mask:
        """Pixel mask shared with the optimizer (analysis/Feature.py:49-61)."""
        return self.optimizer.mask

    @mask.setter
    def mask(self, mask: Mask) -> None:
        self.estimator.exclude = mask.exclude
        self.optimizer.mask = mask

    @property
    def particle(self):
        return self.model.particle

    @particle.setter
    def particle(self, particle) -> None:
        self.model.particle = particle

    def estimate(self) -> pd.Series:
        """Starting parameters by differential evolution on the GPU
        (DEEstimator, population-batched), as analysis/Feature.py:76-85."""
        return self.estimator.estimate(self)

    def optimize(self) -> pd.Series:
        """Fit this feature (analysis/Feature.py:87-100); a fresh random
        pixel subset is drawn on every call when the mask's fraction < 1."""
        return self.optimizer.optimize(self)

    def predicted(self) -> np.ndarray:
        """Model hologram over all pixels of the crop."""
        self.model.coordinates = self.flat_coordinates
        return self.model.hologram().reshape(self.shape)

    def residuals(self) -> np.ndarray:
        return self.predicted() - host_data(self.data)


def _bad_pixels(data) -> np.ndarray:
    """Saturated (== max), NaN and inf pixels (analysis/Feature.py:36-43);
    for a crop in device memory the test runs there and only the boolean
    mask comes back."""
    if hasattr(data, 'is_cuda') and data.is_cuda:
        import torch
        return ((data == data.max()) | ~torch.isfinite(data)).cpu().numpy()
    return (data == np.max(data)) | np.isnan(data) | np.isinf(data)
