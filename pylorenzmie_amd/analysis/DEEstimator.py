"""Global parameter estimate by differential evolution, mirroring
analysis/DEEstimator.py:13-180, with the whole population evaluated on
the GPU in one pass.

The reference evaluates scipy's differential_evolution objective one
candidate at a time in CPU worker processes (analysis/DEEstimator.py:
20-42).  Here the objective is vectorized: every candidate of a
generation becomes one feature of a device batch over the same masked
pixels, so a generation costs one Mie launch (a wavefront per
candidate) and one residual launch whose per-tile chi2 partials are
summed per candidate.  The residuals are FP32 (LMZ_FP32, 1.6x the FP64
rate): the search only ranks candidates, and their chi2 agree with FP64
to ~1e-6 relative.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import pandas as pd

from .. import _lmz
from ..lib import LMObject
from ..theory import LorenzMie
from ..theory.LorenzMie import PARAMS
from .BaseEstimator import BaseEstimator
from .Hologram import Hologram
from .Mask import Mask

DEFAULT_BOUNDS = {'z_p': (10., 600.), 'a_p': (0.25, 10.0), 'n_p': (1.0, 3.0)}


class PopulationChi2:
    """chi2 = sum ((I - d) / noise)^2 of many candidate spheres on the same
    pixels, one device launch per call.  ``coords`` (2 or 3, n) absolute
    pixel coordinates, ``data`` (n,), ``base`` the seven parameters that
    candidates do not set, ``caps`` the largest order count any candidate
    can need."""

    def __init__(self, coords, data, instrument, base, cap: int):
        t = _lmz.torch()
        c = np.atleast_2d(np.asarray(coords, float))
        if c.shape[0] == 2:
            c = np.vstack([c, np.zeros((1, c.shape[1]))])
        self.origin = (float(np.round(c[0].mean())), float(np.round(c[1].mean())), 0.)
        self.rel = (c - np.asarray(self.origin)[:, None]).astype(np.float32)
        self.data = np.asarray(data, np.float32).ravel()
        self.instrument = instrument
        self.base = np.asarray(base, float)
        self.cap = int(cap)
        self.dev = _lmz.device()
        self._size = 0
        self._t = t

    def _allocate(self, size: int) -> None:
        t, n = self._t, self.data.size
        feats = [_lmz.make_feature(n, 0, self.origin, self.instrument, self.cap,
                                   pix_offset=i * n) for i in range(size)]
        self.batch = _lmz.Batch(feats, 1)
        b = self.batch
        self.coords = t.from_numpy(np.tile(self.rel, (1, size))).to(self.dev)
        self.d = t.from_numpy(np.tile(self.data, size)).to(self.dev)
        self.params = t.zeros((size, _lmz.NPARAM), dtype=t.float64, device=self.dev)
        self.nmax = t.zeros(size, dtype=t.int32, device=self.dev)
        self.coef = t.zeros(b.coef_floats(), dtype=t.float32, device=self.dev)
        self.work = t.zeros(b.work_doubles, dtype=t.float64, device=self.dev)
        self.chi2 = t.zeros(b.ntiles, dtype=t.float64, device=self.dev)
        self.tiles = (n + _lmz.TILE - 1) // _lmz.TILE
        self._size = size

    def __call__(self, params: np.ndarray) -> np.ndarray:
        """chi2 of each row of ``params`` ((S, 7): x_p y_p z_p a_p n_p k_p
        alpha)."""
        p = np.atleast_2d(np.asarray(params, float))
        size = p.shape[0]
        if size != self._size:
            self._allocate(size)
        host = np.zeros((size, _lmz.NPARAM))
        host[:, :7] = p
        self.params.copy_(self._t.from_numpy(host))
        lib = _lmz.load_library()
        s = _lmz.stream_handle()
        b = self.batch
        _lmz.check(lib.lmz_mie_ab(b.handle, _lmz.ptr(self.params), None,
                                  _lmz.ptr(self.nmax), _lmz.ptr(self.coef), None,
                                  None, _lmz.ptr(self.work), s))
        flags = _lmz.CARTESIAN | _lmz.BOHREN | _lmz.FP32
        _lmz.check(lib.lmz_hologram(b.handle, _lmz.ptr(self.params), None,
                                    _lmz.ptr(self.nmax), _lmz.ptr(self.coef),
                                    _lmz.ptr(self.coords), flags,
                                    _lmz.ptr(self.d), None, _lmz.ptr(self.chi2),
                                    None, s))
        chi2 = self.chi2.view(size, self.tiles).sum(1).cpu().numpy()
        return np.where(np.isfinite(chi2), chi2, np.inf)


@dataclass
class DEEstimator(BaseEstimator):
    """Estimate z_p, a_p, n_p (the ``bounds`` keys) by differential
    evolution on a ``fraction`` of the pixels; x_p, y_p pinned to the
    crop's centre (analysis/DEEstimator.py:124-180).  The population is
    evaluated on the GPU (scipy's ``vectorized`` objective); the
    ``workers`` setting of the reference has no use here and is ignored."""

    model: LorenzMie
    bounds: dict = field(default_factory=lambda: DEFAULT_BOUNDS.copy())
    fraction: float = 0.02
    exclude: np.ndarray | None = None
    popsize: int = 10
    seed: int | None = None
    settings: dict = field(default_factory=lambda: {'tol': 0.01,
                                                     'polish': False,
                                                     'updating': 'deferred'})

    __hash__ = None

    @property
    def properties(self) -> dict:
        return dict(fraction=self.fraction, popsize=self.popsize,
                    bounds=self.bounds, settings=self.settings)

    @properties.setter
    def properties(self, properties: dict) -> None:
        LMObject.properties.fset(self, properties)

    def estimate(self, hologram: Hologram) -> pd.Series:
        from scipy.optimize import differential_evolution
        prob = self._problem(hologram)
        chi2 = PopulationChi2(prob['coords'], prob['data'],
                              self.model.instrument, prob['base'], prob['cap'])
        result = differential_evolution(_objective(prob, chi2),
                                        list(self.bounds.values()),
                                        **self._de_settings())
        return self._finish(result.x)

    def _problem(self, hologram: Hologram) -> dict:
        """Pin x_p, y_p to the crop's mean coordinates, draw the mask and
        collect what the objective needs (analysis/DEEstimator.py:124-150)."""
        model = self.model
        model.particle.x_p = float(hologram.coordinates[0].mean())
        model.particle.y_p = float(hologram.coordinates[1].mean())
        mask = Mask(fraction=self.fraction)
        if self.exclude is not None:
            mask.exclude = self.exclude
        data, coords = mask.apply(hologram)
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
        return dict(data=data, coords=coords, base=base, cap=cap,
                    cols=[PARAMS.index(v) for v in variables], instrument=inst)

    def _de_settings(self, seed=None) -> dict:
        settings = {k: v for k, v in self.settings.items() if k != 'workers'}
        settings['updating'] = 'deferred'
        return dict(popsize=self.popsize,
                    seed=self.seed if seed is None else seed,
                    vectorized=True, **settings)

    def _finish(self, x) -> pd.Series:
        model = self.model
        found = dict(zip(self.bounds.keys(), x))
        for k in ('x_p', 'y_p', 'z_p'):
            if k in found:
                found[k] -= model.particle.r_0['xyz'.index(k[0])]
        model.properties = found
        return pd.Series(model.particle.properties)


def _objective(prob: dict, chi2):
    """scipy's vectorized objective: columns of x are candidates."""
    base, cols = prob['base'], prob['cols']

    def objective(x):
        x = np.atleast_2d(np.asarray(x, float).T)      # (S, nvar)
        p = np.repeat(base[None, :], x.shape[0], axis=0)
        p[:, cols] = x
        return chi2(p)
    return objective
