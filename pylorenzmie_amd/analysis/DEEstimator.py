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
        self.rel = np.ascontiguousarray(
            (c - np.asarray(self.origin)[:, None]).astype(np.float32))
        self.data = np.ascontiguousarray(np.asarray(data, np.float32).ravel())
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
        # np.tile with unit repetitions copies in the input's memory order;
        # the kernels read [3][npix] row-major
        self.coords = t.from_numpy(np.ascontiguousarray(
            np.tile(self.rel, (1, size)))).to(self.dev)
        self.d = t.from_numpy(np.ascontiguousarray(np.tile(self.data, size))).to(self.dev)
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


class _FrameSearch:
    """All searches of a frame as one device batch (npop copies of each
    feature's masked pixels, feature-major) run by lmz_de_run: the whole
    differential evolution stays on the device."""

    def __init__(self, probs: list, npop: int, lo: np.ndarray, hi: np.ndarray):
        t = _lmz.torch()
        self._t = t
        dev = _lmz.device()
        inst = probs[0]['instrument']
        cap = max(p['cap'] for p in probs)
        feats, coords, data = [], [], []
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
        self.nsearch, self.npop = len(probs), npop
        self.ndim = lo.shape[1]
        self.coords = t.from_numpy(np.ascontiguousarray(
            np.concatenate(coords, axis=1))).to(dev)
        self.data = t.from_numpy(np.concatenate(data)).to(dev)
        base = np.zeros((len(probs), _lmz.NPARAM))
        base[:, :7] = np.stack([p['base'] for p in probs])
        self.base = t.from_numpy(base).to(dev)
        self.bounds = t.from_numpy(np.ascontiguousarray(
            np.stack([lo, hi], axis=-1))).to(dev)
        self.cols = probs[0]['cols']
        self.best = t.zeros((self.nsearch, self.ndim), dtype=t.float64, device=dev)
        self.energy = t.zeros(self.nsearch, dtype=t.float64, device=dev)
        self.gens = t.zeros(self.nsearch, dtype=t.int32, device=dev)

    def settings(self, *, maxiter, tol, atol, mutation, recombination, seed,
                 check_every=4) -> _lmz.DESettings:
        st = _lmz.DESettings()
        st.ndim = self.ndim
        for j, c in enumerate(self.cols):
            st.cols[j] = c
        st.npop = self.npop
        st.maxiter = int(maxiter)
        st.check_every = int(check_every)
        st.tol, st.atol = float(tol), float(atol)
        st.mutation_lo, st.mutation_hi = float(mutation[0]), float(mutation[1])
        st.recombination = float(recombination)
        st.seed = int(seed) & 0xffffffffffffffff
        return st

    def run(self, st: _lmz.DESettings):
        """(best (nsearch, ndim) in parameter units, energy (nsearch,),
        generations (nsearch,)) on the host."""
        import ctypes
        t = self._t
        lib = _lmz.load_library()
        b = self.batch
        nbytes = ctypes.c_size_t()
        _lmz.check(lib.lmz_de_workspace_bytes(b.handle, ctypes.byref(st),
                                              ctypes.byref(nbytes)))
        ws = t.empty(nbytes.value + 256, dtype=t.uint8, device=self.data.device)
        wptr = (ws.data_ptr() + 255) & ~255
        flags = _lmz.CARTESIAN | _lmz.BOHREN | _lmz.FP32
        _lmz.check(lib.lmz_de_run(b.handle, ctypes.byref(st), _lmz.ptr(self.base),
                                  _lmz.ptr(self.bounds), _lmz.ptr(self.coords),
                                  flags, _lmz.ptr(self.data), _lmz.ptr(self.best),
                                  _lmz.ptr(self.energy), _lmz.ptr(self.gens),
                                  wptr, _lmz.stream_handle()))
        out = (self.best.cpu().numpy(), self.energy.cpu().numpy(),
               self.gens.cpu().numpy())
        del ws
        return out


_BATCH_SETTINGS = {'tol', 'atol', 'maxiter', 'mutation', 'recombination',
                   'strategy', 'polish', 'updating', 'workers', 'init'}


def estimate_batch(estimators: list, holograms: list, maxiter: int | None = None
                   ) -> list:
    """One differential-evolution search per (estimator, hologram) pair,
    all run together on the device (lmz_de_run): every generation's
    trials, their Mie coefficients and residuals for every unconverged
    search in three launches, and the selection and stopping test in a
    fourth; nothing returns to the host until the searches end.  Each is
    scipy's search as the reference's DEEstimator runs it
    (analysis/DEEstimator.py:119-173): best1bin with the scale drawn per
    generation from the ``mutation`` range, binomial crossover at
    ``recombination`` with one forced dimension, coordinates outside the
    bounds redrawn, deferred selection, stop when std(energies) <= atol +
    tol |mean(energies)|, at most ``maxiter`` generations, a Latin-
    hypercube start of max(5, popsize * D) members.  The draws come from a
    counter-based generator seeded by the first estimator's seed (None:
    fresh entropy, as scipy takes it), not from scipy's stream.  The estimators must be DEEstimators
    with the same bounds keys, popsize, settings and instrument.  Returns
    each estimator's _finish (its model updated)."""
    if not estimators:
        return []
    e0 = estimators[0]
    for e in estimators:
        if not isinstance(e, DEEstimator):
            raise TypeError('estimate_batch needs DEEstimators, got '
                            f'{type(e).__name__}')
    names = list(e0.bounds.keys())
    D = len(names)
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
    if D > 7 or npop > 1024:
        raise ValueError('estimate_batch searches at most 7 parameters with '
                         'at most 1024 members')
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
    lo = np.array([[e.bounds[k][0] for k in names] for e in estimators], float)
    hi = np.array([[e.bounds[k][1] for k in names] for e in estimators], float)
    search = _FrameSearch(probs, npop, lo, hi)
    st = search.settings(maxiter=maxiter, tol=tol, atol=atol, mutation=(mlo, mhi),
                         recombination=recombination,
                         seed=(np.random.SeedSequence().entropy
                               if e0.seed is None else e0.seed))
    x, _, _ = search.run(st)
    return [est._finish(xi) for est, xi in zip(estimators, x)]
