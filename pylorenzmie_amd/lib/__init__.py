"""Base protocol and pixel grid, mirroring pylorenzmie.lib.

``LMObject`` keeps the reference's properties protocol (lib/LMObject.py:
15-171): a ``properties`` dict used for serialization and by the fitter,
JSON and pandas round trips, and comparison by properties.  ``meshgrid``
follows lib/meshgrid.py:10-40 (x fastest, shape (2, ny*nx)).
"""
from __future__ import annotations

import json
import logging

import numpy as np

__all__ = ['LMObject', 'meshgrid', 'azimuthal']


def meshgrid(shape, corner=(0., 0.), flatten=True, dtype=float):
    """Pixel coordinate grid; lib/meshgrid.py:10-40."""
    ny, nx = shape
    left, top = corner
    x = np.arange(left, left + nx, dtype=dtype)
    y = np.arange(top, top + ny, dtype=dtype)
    xy = np.array(np.meshgrid(x, y))
    return xy.reshape((2, -1)) if flatten else xy


class LMObject:
    """Properties protocol of lib/LMObject.py:15-171."""

    meshgrid = staticmethod(meshgrid)
    __hash__ = None

    @property
    def logger(self) -> logging.Logger:
        return logging.getLogger(
            f'{self.__class__.__module__}.{self.__class__.__qualname__}')

    def __eq__(self, other):
        if not isinstance(other, self.__class__):
            return NotImplemented
        return self.properties == other.properties

    @property
    def properties(self) -> dict:
        raise NotImplementedError

    @properties.setter
    def properties(self, properties: dict) -> None:
        for name, value in properties.items():
            if hasattr(self, name):
                setattr(self, name, value)
            else:
                self.logger.debug(f'Ignoring unknown property {name}')

    def to_json(self, **kwargs) -> str:
        """JSON of the properties; numpy scalars become Python numbers
        (lib/LMObject.py:110-127)."""
        def encode(obj):
            if isinstance(obj, np.generic):
                return obj.item()
            raise TypeError(f'{type(obj).__name__} is not JSON serializable')
        return json.dumps(self.properties, default=encode, **kwargs)

    def from_json(self, s: str) -> None:
        self.properties = json.loads(s)

    def to_pandas(self):
        import pandas as pd
        return pd.Series(self.properties)

    def from_pandas(self, series) -> None:
        self.properties = dict(series)
