"""Abstract base of the parameter estimators, mirroring
analysis/BaseEstimator.py:9-35: an LMObject with an ``estimate(hologram)``
that returns the estimated particle properties as a pandas Series."""
from __future__ import annotations

from abc import ABC, abstractmethod

from ..lib import LMObject


class BaseEstimator(LMObject, ABC):
    """Estimators implement ``estimate`` and the ``properties`` getter; a
    class missing either cannot be instantiated."""

    @property
    @abstractmethod
    def properties(self) -> dict:
        """The estimator's settings."""

    @properties.setter
    def properties(self, properties: dict) -> None:
        LMObject.properties.fset(self, properties)

    @abstractmethod
    def estimate(self, hologram):
        """Estimated particle properties of a normalized hologram crop."""
