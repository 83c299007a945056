"""Homogeneous sphere, mirroring theory/Sphere.py:7-226.

``ab()`` runs the FP64 Mie recursion on the GPU (lmz_mie_ab).  The
order-count rules ``wiscombe_yang`` and ``neves_pisignano`` are closed
forms evaluated on the host, as the reference does.
"""
from __future__ import annotations

import threading
from dataclasses import dataclass

import numpy as np

from .. import _lmz
from .Particle import Particle
from ..lib import LMObject


@dataclass
class Sphere(Particle):
    """Sphere of radius a_p [um] and index n_p + i k_p; defaults as
    theory/Sphere.py:43-45."""

    a_p: float = 1.
    n_p: float = 1.5
    k_p: float = 0.

    __hash__ = None
    __eq__ = LMObject.__eq__

    @property
    def d_p(self) -> float:
        return 2. * self.a_p

    @d_p.setter
    def d_p(self, d_p: float) -> None:
        self.a_p = d_p / 2.

    @property
    def properties(self) -> dict:
        return {'x_p': self.x_p, 'y_p': self.y_p, 'z_p': self.z_p,
                'a_p': self.a_p, 'n_p': self.n_p, 'k_p': self.k_p}

    @properties.setter
    def properties(self, properties: dict) -> None:
        LMObject.properties.fset(self, properties)

    def ab(self, n_m=1.42, wavelength: float = 0.532) -> np.ndarray:
        """Mie coefficients, (n_max + 1, 2) complex128 with row 0 zero."""
        return Sphere.mie_coefficients(self.a_p, self.n_p, self.k_p,
                                       n_m, wavelength)

    @staticmethod
    def wiscombe_yang(x: float, m) -> int:
        """Wiscombe (1980) / Yang (2003) Eq. 30; theory/Sphere.py:91-123."""
        x = np.atleast_1d(np.asarray(x, float))
        m = np.atleast_1d(np.asarray(m, complex))
        xl = abs(x[-1])
        if xl <= 8.:
            ns = np.floor(xl + 4. * np.cbrt(xl) + 1.)
        elif xl <= 4200.:
            ns = np.floor(xl + 4.05 * np.cbrt(xl) + 2.)
        else:
            ns = np.floor(xl + 4. * np.cbrt(xl) + 2.)
        xm = np.abs(x * m)
        return int(max(ns, np.max(xm)))

    @staticmethod
    def neves_pisignano(x: float, precision: float = 6.) -> int:
        """Neves & Pisignano (2012); theory/Sphere.py:125-145."""
        return int(x + 0.76 * np.cbrt(precision * precision * x) - 4.1)

    @staticmethod
    def mie_coefficients(a_p, n_p, k_p, n_m, wavelength) -> np.ndarray:
        """theory/Sphere.py:147-226, computed by the FP64 GPU kernel."""
        return _mie_ab_gpu(float(a_p), float(n_p), float(k_p),
                           float(np.real(n_m)), float(wavelength))


_ab_cache: dict = {}
_ab_lock = threading.Lock()


def _mie_ab_gpu(a_p, n_p, k_p, n_m, wavelength) -> np.ndarray:
    t = _lmz.torch()
    dev = _lmz.device()
    k = 2. * np.pi / wavelength * n_m
    cap = Sphere.wiscombe_yang(k * a_p, (n_p + 1j * k_p) / n_m) + 2
    with _ab_lock:
        ent = _ab_cache.get(cap)
        if ent is None:
            f = _lmz.Feature()
            f.npix, f.width = 1, 1
            f.n_m, f.wavelength, f.magnification, f.noise = 1., 1., 1., 1.
            f.nmax_cap = cap
            batch = _lmz.Batch([f], 1)
            ent = dict(batch=batch,
                       params=t.zeros(_lmz.NPARAM, dtype=t.float64, device=dev),
                       nmax=t.zeros(1, dtype=t.int32, device=dev),
                       coef=t.zeros(batch.coef_floats(), dtype=t.float32,
                                    device=dev),
                       ab=t.zeros(batch.coef_slots * 4, dtype=t.float64,
                                  device=dev),
                       work=t.zeros(batch.work_doubles, dtype=t.float64,
                                    device=dev))
            _ab_cache[cap] = ent
        batch = ent['batch']
        # per-call instrument values live in the feature table
        feat = batch.features[0]
        if (feat.n_m, feat.wavelength) != (n_m, wavelength):
            feat.n_m, feat.wavelength = n_m, wavelength
            batch = _lmz.Batch([feat], 1)
            ent['batch'] = batch
        p = np.zeros(_lmz.NPARAM)
        p[3:6] = a_p, n_p, k_p
        ent['params'].copy_(t.from_numpy(p))
        lib = _lmz.load_library()
        _lmz.check(lib.lmz_mie_ab(batch.handle, _lmz.ptr(ent['params']), None,
                                  _lmz.ptr(ent['nmax']), _lmz.ptr(ent['coef']),
                                  None, _lmz.ptr(ent['ab']),
                                  _lmz.ptr(ent['work']),
                                  _lmz.stream_handle()))
        n = int(ent['nmax'].cpu()[0])
        ab4 = ent['ab'][:4 * n].cpu().numpy().reshape(n, 4)
    out = np.zeros((n + 1, 2), np.complex128)
    out[1:, 0] = ab4[:, 0] + 1j * ab4[:, 1]
    out[1:, 1] = ab4[:, 2] + 1j * ab4[:, 3]
    return out
