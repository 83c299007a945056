"""Sphere of concentric layers, the layered counterpart of Sphere.

The reference computes layered coefficients only in its obsolete
lib/obsolete/sphere_coefficients.py:126-215 (Yang's recursion), which fails
for list input; here ``ab()`` runs Yang's recursion in FP64 on the GPU
(lmz_mie_ab_layered) with n_max from Yang's Eq. 30.  A one-layer
LayeredSphere has exactly the coefficients of Sphere.
"""
from __future__ import annotations

import threading
from dataclasses import dataclass, field

import numpy as np

from .. import _lmz
from ..lib import LMObject
from .Particle import Particle


@dataclass
class LayeredSphere(Particle):
    """Outer radii ``a_p`` [um] of each layer (innermost first, increasing)
    and the layers' indices ``n_p`` + i ``k_p``."""

    a_p: list = field(default_factory=lambda: [0.5, 1.0])
    n_p: list = field(default_factory=lambda: [1.6, 1.4])
    k_p: list = field(default_factory=lambda: [0., 0.])

    __hash__ = None
    __eq__ = LMObject.__eq__

    @property
    def layers(self) -> np.ndarray:
        """(L, 3) rows of outer radius, n, k; validated."""
        a = np.atleast_1d(np.asarray(self.a_p, float))
        n = np.atleast_1d(np.asarray(self.n_p, float))
        k = np.atleast_1d(np.asarray(self.k_p, float))
        if not (a.size == n.size == k.size) or a.size < 1:
            raise ValueError('a_p, n_p and k_p need one entry per layer')
        if a.size > _lmz.MAX_LAYERS:
            raise ValueError(f'at most {_lmz.MAX_LAYERS} layers')
        if np.any(a <= 0) or np.any(np.diff(a) <= 0):
            raise ValueError('layer radii must be positive and increasing')
        return np.stack([a, n, k], 1)

    @property
    def properties(self) -> dict:
        return {'x_p': self.x_p, 'y_p': self.y_p, 'z_p': self.z_p,
                'a_p': list(np.atleast_1d(self.a_p)),
                'n_p': list(np.atleast_1d(self.n_p)),
                'k_p': list(np.atleast_1d(self.k_p))}

    @properties.setter
    def properties(self, properties: dict) -> None:
        LMObject.properties.fset(self, properties)

    def ab(self, n_m=1.42, wavelength: float = 0.532) -> np.ndarray:
        """Mie coefficients, (n_max + 1, 2) complex128 with row 0 zero."""
        return layered_ab([self.layers], float(np.real(n_m)),
                          float(wavelength))[0]

    @staticmethod
    def wiscombe_yang(x, m) -> int:
        """Wiscombe (1980) with Yang (2003) Eq. 30 over the layers:
        max(n_s(x_L), |m_l x_l|, |m_l x_{l-1}|)."""
        x = np.atleast_1d(np.asarray(x, float))
        m = np.atleast_1d(np.asarray(m, complex))
        xl = abs(x[-1])
        if xl <= 8.:
            ns = np.floor(xl + 4. * np.cbrt(xl) + 1.)
        elif xl <= 4200.:
            ns = np.floor(xl + 4.05 * np.cbrt(xl) + 2.)
        else:
            ns = np.floor(xl + 4. * np.cbrt(xl) + 2.)
        nstop = max(ns, np.max(np.abs(x * m)))
        if x.size > 1:
            nstop = max(nstop, np.max(np.abs(x[:-1] * m[1:])))
        return int(nstop)


def layer_table(layers_list) -> tuple[np.ndarray, np.ndarray]:
    """(nfeat, MAX_LAYERS, 3) table and layer counts for lmz_mie_ab_layered."""
    nfeat = len(layers_list)
    table = np.zeros((nfeat, _lmz.MAX_LAYERS, 3))
    counts = np.zeros(nfeat, np.int32)
    for i, lay in enumerate(layers_list):
        lay = np.asarray(lay, float).reshape(-1, 3)
        table[i, :len(lay)] = lay
        counts[i] = len(lay)
    return table, counts


def layered_cap(layers, n_m: float, wavelength: float,
                headroom: float = 0.) -> int:
    """Orders to reserve for one layered particle ((L, 3) rows); with
    ``headroom`` for a fit whose radii may grow by that fraction and whose
    outer index may grow by 0.1."""
    lay = np.asarray(layers, float).reshape(-1, 3).copy()
    if headroom:
        lay[:, 0] *= 1. + headroom
        lay[-1, 1] += 0.1
    k = 2. * np.pi / wavelength * n_m
    return LayeredSphere.wiscombe_yang(k * lay[:, 0],
                                       (lay[:, 1] + 1j * lay[:, 2]) / n_m) + 2


_lock = threading.Lock()


def layered_ab(layers_list, n_m: float, wavelength: float) -> list:
    """Coefficients of many layered particles in one launch; list of
    (n_max + 1, 2) complex128 arrays."""
    t = _lmz.torch()
    dev = _lmz.device()
    feats = []
    for i, lay in enumerate(layers_list):
        f = _lmz.Feature()
        f.npix, f.width, f.pix_offset = 1, 1, i
        f.n_m, f.wavelength, f.magnification, f.noise = n_m, wavelength, 1., 1.
        f.nmax_cap = layered_cap(lay, n_m, wavelength)
        feats.append(f)
    with _lock:
        batch = _lmz.Batch(feats, 1)
        table, counts = layer_table(layers_list)
        lt = t.from_numpy(table).to(dev)
        nl = t.from_numpy(counts).to(dev)
        nmax = t.zeros(len(feats), dtype=t.int32, device=dev)
        ab = t.zeros(batch.coef_slots * 4, dtype=t.float64, device=dev)
        work = t.zeros(batch.work_doubles, dtype=t.float64, device=dev)
        lib = _lmz.load_library()
        _lmz.check(lib.lmz_mie_ab_layered(batch.handle, _lmz.ptr(lt), _lmz.ptr(nl),
                                          _lmz.ptr(nmax), None, _lmz.ptr(ab),
                                          _lmz.ptr(work), _lmz.stream_handle()))
        nm = nmax.cpu().numpy()
        ab4 = ab.cpu().numpy().reshape(-1, 4)
        offsets = [batch.features[i].coef_offset for i in range(len(feats))]
        batch.close()
    out = []
    for off, n in zip(offsets, nm):
        if n < 0:
            raise ValueError('invalid layer table')
        r = ab4[off:off + n]
        a = np.zeros((n + 1, 2), np.complex128)
        a[1:, 0] = r[:, 0] + 1j * r[:, 1]
        a[1:, 1] = r[:, 2] + 1j * r[:, 3]
        out.append(a)
    return out
