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


class _FramePopulation:
    """chi2 of every candidate of every feature in one device launch: the
    batch holds NP copies of each feature's masked pixels, feature-major."""

    def __init__(self, probs: list, npop: int):
        t = _lmz.torch()
        self._t = t
        dev = _lmz.device()
        inst = probs[0]['instrument']
        cap = max(p['cap'] for p in probs)
        feats, coords, data, owner = [], [], [], []
        off = 0
        for p in probs:
            c = np.atleast_2d(np.asarray(p['coords'], float))
            if c.shape[0] == 2:
                c = np.vstack([c, np.zeros((1, c.shape[1]))])
            origin = (float(np.round(c[0].mean())), float(np.round(c[1].mean())), 0.)
            rel = (c - np.asarray(origin)[:, None]).astype(np.float32)
            d = np.asarray(p['data'], np.float32).ravel()
            for _ in range(npop):
                feats.append(_lmz.make_feature(d.size, 0, origin, inst, cap,
                                               pix_offset=off))
                coords.append(rel)
                data.append(d)
                off += d.size
        self.batch = _lmz.Batch(feats, 1)
        b = self.batch
        self.coords = t.from_numpy(np.ascontiguousarray(
            np.concatenate(coords, axis=1))).to(dev)
        self.data = t.from_numpy(np.concatenate(data)).to(dev)
        self.params = t.zeros((len(feats), _lmz.NPARAM), dtype=t.float64, device=dev)
        self.nmax = t.zeros(len(feats), dtype=t.int32, device=dev)
        self.coef = t.zeros(b.coef_floats(), dtype=t.float32, device=dev)
        self.work = t.zeros(b.work_doubles, dtype=t.float64, device=dev)
        self.chi2 = t.zeros(b.ntiles, dtype=t.float64, device=dev)
        ntiles = [f.ntiles for f in b.features]
        self.tile_owner = t.from_numpy(np.repeat(np.arange(len(feats)),
                                                 ntiles)).to(dev)
        self.n = len(feats)

    def __call__(self, rows: np.ndarray) -> np.ndarray:
        t = self._t
        host = np.zeros((self.n, _lmz.NPARAM))
        host[:, :7] = rows
        self.params.copy_(t.from_numpy(host))
        lib = _lmz.load_library()
        s = _lmz.stream_handle()
        b = self.batch
        _lmz.check(lib.lmz_mie_ab(b.handle, _lmz.ptr(self.params), None,
                                  _lmz.ptr(self.nmax), _lmz.ptr(self.coef), None,
                                  None, _lmz.ptr(self.work), s))
        _lmz.check(lib.lmz_hologram(b.handle, _lmz.ptr(self.params), None,
                                    _lmz.ptr(self.nmax), _lmz.ptr(self.coef),
                                    _lmz.ptr(self.coords),
                                    _lmz.CARTESIAN | _lmz.BOHREN | _lmz.FP32,
                                    _lmz.ptr(self.data), None,
                                    _lmz.ptr(self.chi2), None, s))
        out = t.zeros(self.n, dtype=t.float64, device=self.chi2.device)
        out.index_add_(0, self.tile_owner, self.chi2)
        chi2 = out.cpu().numpy()
        return np.where(np.isfinite(chi2), chi2, np.inf)


_BATCH_SETTINGS = {'tol', 'atol', 'maxiter', 'mutation', 'recombination',
                   'strategy', 'polish', 'updating', 'workers', 'init'}


def estimate_batch(estimators: list, holograms: list, maxiter: int | None = None
                   ) -> list:
    """One differential-evolution search per (estimator, hologram) pair,
    all advanced together: every generation's candidates of every search in
    one device launch, and scipy's best1bin step (scale in the ``mutation``
    range drawn per generation, binomial crossover at ``recombination`` with
    one forced dimension, out-of-range values redrawn, deferred selection,
    stop when std(energies) <= atol + tol |mean(energies)|, at most
    ``maxiter`` generations) applied to all of them at once as array
    operations; scipy's population of max(5, popsize * D) members.  The
    estimators must be DEEstimators with the same bounds keys, popsize,
    settings and instrument; the searches draw from one random stream
    seeded by the first estimator's seed.  Returns each estimator's
    _finish (its model updated)."""
    if not estimators:
        return []
    e0 = estimators[0]
    for e in estimators:
        if not isinstance(e, DEEstimator):
            raise TypeError('estimate_batch needs DEEstimators, got '
                            f'{type(e).__name__}')
    names = list(e0.bounds.keys())
    D, F = len(names), len(estimators)
    npop = max(5, e0.popsize * D)
    settings = dict(e0.settings)
    unknown = set(settings) - _BATCH_SETTINGS
    if unknown:
        raise ValueError(f'estimate_batch does not support {sorted(unknown)}')
    if settings.get('strategy', 'best1bin') != 'best1bin':
        raise ValueError('estimate_batch runs the best1bin strategy only')
    if settings.get('polish', False):
        raise ValueError('estimate_batch does not polish (polish=False)')
    if settings.get('init', 'latinhypercube') != 'latinhypercube':
        raise ValueError('estimate_batch starts from a Latin hypercube')
    tol = float(settings.get('tol', 0.01))
    atol = float(settings.get('atol', 0.))
    mutation = settings.get('mutation', (0.5, 1.))
    mlo, mhi = ((float(mutation), float(mutation)) if np.isscalar(mutation)
                else (float(mutation[0]), float(mutation[1])))
    recombination = float(settings.get('recombination', 0.7))
    if maxiter is None:
        maxiter = int(settings.get('maxiter', 1000))
    inst0 = e0.model.instrument
    for e in estimators:
        if list(e.bounds.keys()) != names or e.popsize != e0.popsize:
            raise ValueError('estimators differ in bounds or popsize')
        if {k: v for k, v in e.settings.items() if k != 'workers'} != \
                {k: v for k, v in settings.items() if k != 'workers'}:
            raise ValueError('estimators differ in settings')
        inst = e.model.instrument
        if (inst.n_m, inst.wavelength, inst.magnification) != \
                (inst0.n_m, inst0.wavelength, inst0.magnification):
            raise ValueError('estimators differ in instrument')
    probs = [e._problem(h) for e, h in zip(estimators, holograms)]
    lo = np.array([[e.bounds[k][0] for k in names] for e in estimators])
    hi = np.array([[e.bounds[k][1] for k in names] for e in estimators])
    base = np.stack([p['base'] for p in probs])
    cols = probs[0]['cols']
    pop_chi2 = _FramePopulation(probs, npop)
    rng = np.random.default_rng(e0.seed)

    def energies(unit):
        rows = np.repeat(base, npop, axis=0)
        rows[:, cols] = (lo[:, None, :] + unit * (hi - lo)[:, None, :]).reshape(-1, D)
        return pop_chi2(rows).reshape(F, npop)

    # Latin hypercube start, per search and dimension
    unit = (np.argsort(rng.random((F, npop, D)), axis=1) + rng.random((F, npop, D))) / npop
    e = energies(unit)
    done = np.zeros(F, bool)
    idx = np.arange(npop)
    for _ in range(maxiter):
        best = unit[np.arange(F), np.argmin(e, axis=1)]             # (F, D)
        scale = rng.uniform(mlo, mhi, (F, 1, 1))
        k0 = rng.integers(0, npop - 1, (F, npop))
        r0 = k0 + (k0 >= idx)
        a, b = np.minimum(idx, r0), np.maximum(idx, r0)
        r1 = rng.integers(0, npop - 2, (F, npop))
        r1 = r1 + (r1 >= a)
        r1 = r1 + (r1 >= b)
        take = np.arange(F)[:, None]
        mutant = best[:, None, :] + scale * (unit[take, r0] - unit[take, r1])
        cross = rng.random((F, npop, D)) < recombination
        cross[take, idx[None, :], rng.integers(0, D, (F, npop))] = True
        trial = np.where(cross, mutant, unit)
        bad = (trial < 0.) | (trial > 1.)
        trial[bad] = rng.random(int(bad.sum()))
        et = energies(trial)
        better = (et < e) & ~done[:, None]
        unit = np.where(better[..., None], trial, unit)
        e = np.where(better, et, e)
        with np.errstate(invalid='ignore'):     # a search holding inf waits
            done |= np.std(e, axis=1) <= atol + tol * np.abs(np.mean(e, axis=1))
        if done.all():
            break
    x = lo + unit[np.arange(F), np.argmin(e, axis=1)] * (hi - lo)
    return [est._finish(xi) for est, xi in zip(estimators, x)]
