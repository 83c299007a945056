"""Global estimate of z_p, a_p, n_p from the azimuthal profile of a crop,
mirroring analysis/RadialEstimator.py:18-166, with every generation of the
search evaluated on the GPU in one pass.

The reference compares the model on a radial spoke (x_p + r, y_p),
r = 0 .. n - 1, with the crop's azimuthal average about (x_p, y_p) (the
crop's mean coordinates), one scipy differential-evolution candidate at a
time (RadialEstimator.py:18-40).  Here the spoke is the explicit pixel list
of a PopulationChi2 batch, one feature per candidate: a generation is one
Mie launch and one residual launch, as in DEEstimator.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import pandas as pd

from .. import _lmz
from ..lib import LMObject, azimuthal
from ..theory import LorenzMie
from ..theory.LorenzMie import PARAMS
from .DEEstimator import DEFAULT_BOUNDS, PopulationChi2
from .BaseEstimator import BaseEstimator
from .Hologram import Hologram
from .Mask import host_data


@dataclass
class RadialEstimator(BaseEstimator):
    """Estimate the ``bounds`` keys (default z_p, a_p, n_p) by differential
    evolution on the azimuthal profile (analysis/RadialEstimator.py:43-166);
    x_p and y_p are pinned to the crop's mean coordinates."""

    model: LorenzMie
    bounds: dict = field(default_factory=lambda: DEFAULT_BOUNDS.copy())
    popsize: int = 10
    seed: int | None = None
    settings: dict = field(default_factory=lambda: {'tol': 0.01,
                                                     'polish': False,
                                                     'updating': 'deferred',
                                                     'workers': 1})

    __hash__ = None

    @property
    def properties(self) -> dict:
        return dict(popsize=self.popsize, bounds=self.bounds,
                    settings=self.settings)

    @properties.setter
    def properties(self, properties: dict) -> None:
        LMObject.properties.fset(self, properties)

    def profile(self, hologram: Hologram):
        """(radii, azimuthal average) about the crop's mean coordinates
        (RadialEstimator.py:134-146)."""
        x_p = float(hologram.coordinates[0].mean())
        y_p = float(hologram.coordinates[1].mean())
        cx, cy = hologram.corner
        prof = azimuthal.avg(host_data(hologram.data), center=(x_p - cx, y_p - cy))
        return x_p, y_p, prof

    def estimate(self, hologram: Hologram) -> pd.Series:
        from scipy.optimize import differential_evolution
        model = self.model
        x_p, y_p, prof = self.profile(hologram)
        model.particle.x_p = x_p
        model.particle.y_p = y_p
        radii = np.arange(prof.size, dtype=float)
        spoke = np.vstack([x_p + radii, np.full(prof.size, y_p)])
        variables = list(self.bounds.keys())
        unknown = set(variables) - set(PARAMS)
        if unknown:
            raise ValueError(f'cannot search {sorted(unknown)}')
        props = model.properties
        base = np.array([props[k] for k in PARAMS], float)
        base[:3] += model.particle.r_0
        inst = model.instrument
        hi = {k: v[1] for k, v in self.bounds.items()}
        cap = _lmz.nmax_cap(hi.get('a_p', base[3]), hi.get('n_p', base[4]),
                            hi.get('k_p', base[5]), inst.n_m, inst.wavelength)
        good = np.isfinite(prof)
        chi2 = PopulationChi2(spoke[:, good], prof[good], inst, base, cap)
        cols = [PARAMS.index(v) for v in variables]

        def objective(x):
            x = np.atleast_2d(np.asarray(x, float).T)      # (S, nvar)
            p = np.repeat(base[None, :], x.shape[0], axis=0)
            p[:, cols] = x
            return chi2(p)

        settings = {k: v for k, v in self.settings.items() if k != 'workers'}
        settings['updating'] = 'deferred'
        result = differential_evolution(objective, list(self.bounds.values()),
                                        popsize=self.popsize, seed=self.seed,
                                        vectorized=True, **settings)
        found = dict(zip(variables, result.x))
        for k in ('x_p', 'y_p', 'z_p'):
            if k in found:
                found[k] -= model.particle.r_0['xyz'.index(k[0])]
        model.properties = found
        return pd.Series(model.particle.properties)
