"""Abstract base of the parameter estimators, mirroring
analysis/BaseEstimator.py:9-35: an LMObject with an ``estimate(hologram)``
that returns the estimated particle properties as a pandas Series."""
from __future__ import annotations

from abc import abstractmethod

from ..lib import LMObject


class BaseEstimator(LMObject):
    """Estimators implement ``estimate`` and the ``properties`` protocol."""

    @abstractmethod
    def estimate(self, hologram):
        """Estimated particle properties of a normalized hologram crop."""
