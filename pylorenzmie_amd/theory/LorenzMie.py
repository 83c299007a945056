"""Lorenz-Mie field, hologram and Jacobian on the GPU.

Drop-in for ``LorenzMie(coordinates, particle, instrument)`` of
theory/LorenzMie.py:8-342: same constructor, ``properties`` protocol,
coordinate handling and ``field()`` / ``hologram()`` results, computed by
the gfx950 kernels of liblmz.so.  Adds the analytic-Jacobian hook of
theory/jaxLorenzMie.py:321, 354-395 (``jac_params`` / ``jac()``), extended
to k_p and the scattered-field amplitude ``alpha``
(I = |x + alpha E_s|^2, theory/obsolete/pycudaLMHologram.py:48-49).
"""
from __future__ import annotations

import numpy as np

from .. import _lmz
from ..lib import LMObject
from .Instrument import Instrument
from .Particle import Particle
from .Sphere import Sphere

PARAMS = ('x_p', 'y_p', 'z_p', 'a_p', 'n_p', 'k_p', 'alpha')


class _GivenCoefficients(Particle):
    """A particle at the origin whose Mie coefficients are given
    (LorenzMie.lorenzmie evaluates them at arbitrary displacements)."""

    def __init__(self, ab) -> None:
        super().__init__(x_p=0., y_p=0., z_p=0.)
        self._ab = np.asarray(ab, complex)

    def ab(self, n_m=None, wavelength=None) -> np.ndarray:
        return self._ab


def dense_layout(c: np.ndarray):
    """(width, origin) if (3, P) coordinates are a full meshgrid at one z."""
    z = c[2]
    if c.shape[1] == 0 or not np.all(z == z[0]):
        return None
    x, y = c[0], c[1]
    npts = x.size
    change = np.flatnonzero(y != y[0])
    w = int(change[0]) if change.size else npts
    if w == 0 or npts % w:
        return None
    h = npts // w
    gx = x[0] + np.arange(w, dtype=float)
    gy = y[0] + np.arange(h, dtype=float)
    if (np.array_equal(x.reshape(h, w), np.broadcast_to(gx, (h, w))) and
            np.array_equal(y.reshape(h, w),
                           np.broadcast_to(gy[:, None], (h, w)))):
        return w, (float(x[0]), float(y[0]), float(z[0]))
    return None


class LorenzMie(LMObject):
    """Scattered field of one or more particles, computed on the GPU.

    ``method = 'hip numpy'`` makes the class acceptable to fitters whose
    ``method`` is 'hip' or 'numpy' (analysis/Optimizer.py:64-65).
    """

    method: str = 'hip numpy'
    jac_params: frozenset = frozenset(PARAMS)

    def __init__(self, coordinates=None, particle=None, instrument=None,
                 alpha: float = 1.0, double_precision: bool = True) -> None:
        self._batch = None
        self._batch_key = None
        self.alpha = float(alpha)
        # cupyLorenzMie's switch (theory/cupyLorenzMie.py:46-65): False
        # computes hologram() of one particle in FP32 arithmetic (errors up
        # to 1e-4 + 2e-5 I, Cartesian only); the field stays FP64
        self.double_precision = bool(double_precision)
        self.coordinates = coordinates
        self.particle = particle or Sphere()
        self.instrument = instrument or Instrument()

    def __repr__(self) -> str:
        return (f'{self.__class__.__qualname__}(instrument={self.instrument!r}, '
                f'particle={self.particle!r}, alpha={self.alpha})')

    # ---------------------------------------------------------- properties
    @property
    def properties(self) -> dict:
        """Particle, then instrument properties (theory/LorenzMie.py:68-71),
        then alpha."""
        return {**self.particle.properties, **self.instrument.properties,
                'alpha': self.alpha}

    @properties.setter
    def properties(self, properties: dict) -> None:
        """Set on the particle, else the instrument, else the model
        (theory/LorenzMie.py:73-81)."""
        for name, value in properties.items():
            if hasattr(self.particle, name):
                setattr(self.particle, name, value)
            elif hasattr(self.instrument, name):
                setattr(self.instrument, name, value)
            elif hasattr(self, name):
                setattr(self, name, value)

    # --------------------------------------------------------- coordinates
    @property
    def coordinates(self) -> np.ndarray:
        return self._coordinates

    @coordinates.setter
    def coordinates(self, coordinates) -> None:
        """(1..3, npts) -> (3, npts) with z = 0 padding; None -> 201x201
        grid; more than 3 rows -> ValueError (theory/LorenzMie.py:93-104)."""
        if coordinates is None:
            c = self.meshgrid((201, 201))
        else:
            c = np.atleast_2d(np.asarray(coordinates, dtype=float))
        ndim, npts = c.shape
        if ndim > 3:
            raise ValueError(f'Incompatible shape: {c.shape=}')
        self._coordinates = np.vstack([c, np.zeros((3 - ndim, npts))])
        self._allocate()

    def _allocate(self) -> None:
        """Describe the pixels for the kernels: a dense crop (implicit
        coordinates) when the points form a full grid, else explicit FP32
        coordinates relative to a rounded origin."""
        c = self._coordinates
        self._npts = c.shape[1]
        layout = dense_layout(c)
        if layout is not None:
            self._width, self._origin = layout
            self._rel = None
        else:
            self._width = 0
            self._origin = tuple(float(np.round(v)) if np.all(np.isfinite(v))
                                 else 0. for v in c.mean(axis=1))
            self._rel = np.ascontiguousarray(
                (c - np.asarray(self._origin)[:, None]).astype(np.float32))
        self._field = np.empty((3, self._npts), dtype=complex)
        self._dev = None
        self._batch = None
        self._batch_key = None

    # ------------------------------------------------------------- device
    def _device_buffers(self, cap: int):
        t = _lmz.torch()
        dev = _lmz.device()
        inst = self.instrument
        key = (float(np.real(inst.n_m)), float(inst.wavelength),
               float(inst.magnification), float(inst.noise), cap)
        if self._dev is None:
            P = max(self._npts, 1)
            d = dict(
                params=t.zeros(_lmz.NPARAM, dtype=t.float64, device=dev),
                holo=t.empty(P, dtype=t.float32, device=dev),
                field=t.empty(3 * P * 2, dtype=t.float32, device=dev),
                nmax=t.zeros(1, dtype=t.int32, device=dev),
                coords=(t.from_numpy(self._rel).to(dev)
                        if self._rel is not None else None))
            self._dev = d
        if self._batch is None or self._batch_key != key:
            feat = _lmz.make_feature(self._npts, self._width, self._origin,
                                     inst, cap)
            self._batch = _lmz.Batch([feat], 1)
            self._batch_key = key
            d = self._dev
            b = self._batch
            d['coef'] = t.zeros(b.coef_floats(), dtype=t.float32, device=dev)
            d['dcoef'] = t.zeros(b.dcoef_floats(), dtype=t.float32, device=dev)
            d['ab'] = t.zeros(b.coef_slots * 4, dtype=t.float64, device=dev)
            d['nmax_in'] = t.zeros(1, dtype=t.int32, device=dev)
            d['work'] = t.zeros(self._batch.work_doubles, dtype=t.float64,
                                device=dev)
        return self._dev

    def _particle_cap(self, particle) -> int:
        inst = self.instrument
        if isinstance(particle, Sphere):
            k = inst.wavenumber(scaled=False)
            m = (particle.n_p + 1j * particle.k_p) / np.real(inst.n_m)
            return Sphere.wiscombe_yang(k * particle.a_p, m) + 2
        return max(particle.ab(inst.n_m, inst.wavelength).shape[0] - 1, 1)

    def _cap(self) -> int:
        need = max(self._particle_cap(p) for p in self.particle)
        cur = self._batch_key[4] if self._batch_key else 0
        if need <= cur:
            return cur
        return need + 8

    def _load_particle(self, particle, d, tangents: bool = False) -> None:
        """Parameters and per-order tables of one particle on the device."""
        t = _lmz.torch()
        lib = _lmz.load_library()
        b = self._batch
        p = np.zeros(_lmz.NPARAM)
        p[:3] = particle.r_p + particle.r_0
        p[6] = self.alpha
        stream = _lmz.stream_handle()
        if isinstance(particle, Sphere):
            p[3:6] = particle.a_p, particle.n_p, particle.k_p
            d['params'].copy_(t.from_numpy(p))
            _lmz.check(lib.lmz_mie_ab(
                b.handle, _lmz.ptr(d['params']), None, _lmz.ptr(d['nmax']),
                _lmz.ptr(d['coef']), _lmz.ptr(d['dcoef']) if tangents else None,
                None, _lmz.ptr(d['work']), stream))
        else:
            if tangents:
                raise TypeError('jac() supports single Sphere particles only')
            ab = np.asarray(particle.ab(self.instrument.n_m,
                                        self.instrument.wavelength), complex)
            n = ab.shape[0] - 1
            rows = np.zeros((max(n, 1), 4))
            if n > 0:
                rows[:n, 0], rows[:n, 1] = ab[1:, 0].real, ab[1:, 0].imag
                rows[:n, 2], rows[:n, 3] = ab[1:, 1].real, ab[1:, 1].imag
            d['params'].copy_(t.from_numpy(p))
            d['ab'][:rows.size].copy_(t.from_numpy(rows.ravel()))
            d['nmax_in'].fill_(n)
            _lmz.check(lib.lmz_coef_from_ab(
                b.handle, _lmz.ptr(d['ab']), _lmz.ptr(d['nmax_in']),
                _lmz.ptr(d['nmax']), _lmz.ptr(d['coef']), stream))

    @staticmethod
    def _flags(cartesian: bool, bohren: bool) -> int:
        return ((_lmz.CARTESIAN if cartesian else 0) |
                (_lmz.BOHREN if bohren else 0))

    # ------------------------------------------------------------ compute
    def _superposed(self) -> bool:
        """Several particles, or an aberration mask: one pass of the
        superposition kernel instead of the single-particle kernels."""
        return len(self.particle) != 1 or getattr(self, 'spherical', 0.) != 0.

    def _particle_table(self, d):
        """Per-particle parameters and coefficient tables of every
        constituent (lmz_hologram_superposed's particle arrays)."""
        t = _lmz.torch()
        dev = _lmz.device()
        lib = _lmz.load_library()
        inst = self.instrument
        parts = list(self.particle)
        abs_ = [np.asarray(p.ab(inst.n_m, inst.wavelength), complex) for p in parts]
        cap = max(max(a.shape[0] - 1 for a in abs_), 1)
        feats = [_lmz.make_feature(1, 1, (0., 0., 0.), inst, cap, i)
                 for i in range(len(parts))]
        pb = _lmz.Batch(feats, 1)
        rows = np.zeros((len(parts) * cap, 4))
        pnmax = np.zeros(len(parts), np.int32)
        pp = np.zeros((len(parts), _lmz.NPARAM))
        for i, (p, ab) in enumerate(zip(parts, abs_)):
            n = ab.shape[0] - 1
            pnmax[i] = n
            r = rows[i * cap:i * cap + n]
            r[:, 0], r[:, 1] = ab[1:, 0].real, ab[1:, 0].imag
            r[:, 2], r[:, 3] = ab[1:, 1].real, ab[1:, 1].imag
            pp[i, :3] = np.asarray(p.r_p, float) + np.asarray(p.r_0, float)
        ab_d = t.from_numpy(rows.ravel()).to(dev)
        nin = t.from_numpy(pnmax).to(dev)
        pn = t.zeros(len(parts), dtype=t.int32, device=dev)
        coef = t.zeros(pb.coef_floats(), dtype=t.float32, device=dev)
        _lmz.check(lib.lmz_coef_from_ab(pb.handle, _lmz.ptr(ab_d), _lmz.ptr(nin),
                                        _lmz.ptr(pn), _lmz.ptr(coef),
                                        _lmz.stream_handle()))
        d['superposed'] = dict(batch=pb, params=t.from_numpy(pp).to(dev), nmax=pn,
                               coef=coef, cap=cap, n=len(parts),
                               offset=t.tensor([0, len(parts)], dtype=t.int32,
                                               device=dev),
                               alpha=t.tensor([self.alpha], dtype=t.float64,
                                              device=dev))
        return d['superposed']

    def _launch_superposed(self, d, flags, out=None, field=None) -> None:
        sp = self._particle_table(d)
        _lmz.check(_lmz.load_library().lmz_hologram_superposed(
            self._batch.handle, sp['n'], _lmz.ptr(sp['offset']),
            _lmz.ptr(sp['params']), _lmz.ptr(sp['nmax']), _lmz.ptr(sp['coef']),
            sp['cap'], _lmz.ptr(d['coords']), flags,
            float(getattr(self, 'spherical', 0.)), _lmz.ptr(sp['alpha']),
            _lmz.ptr(out), _lmz.ptr(field), _lmz.stream_handle()))

    def field(self, cartesian: bool = True, bohren: bool = True,
              device: bool = False):
        """Scattered field alpha E_s, (3, npts) complex.

        Returns the model's own buffer, as theory/LorenzMie.py:160-181 does;
        ``device=True`` returns a complex64 tensor on the GPU instead.
        """
        if self._npts == 0:
            return self._field
        d = self._device_buffers(self._cap())
        lib = _lmz.load_library()
        flags = self._flags(cartesian, bohren)
        stream = _lmz.stream_handle()
        if self._superposed():
            self._launch_superposed(d, flags, field=d['field'])
        else:
            self._load_particle(self.particle[0], d)
            _lmz.check(lib.lmz_field(
                self._batch.handle, _lmz.ptr(d['params']), None,
                _lmz.ptr(d['nmax']), _lmz.ptr(d['coef']), _lmz.ptr(d['coords']),
                flags, _lmz.ptr(d['field']), stream))
        t = _lmz.torch()
        fld = t.view_as_complex(d['field'].view(3, self._npts, 2))
        if device:
            return fld
        self._field[...] = fld.cpu().numpy()
        return self._field

    def scattered_field(self, particle, cartesian: bool = True,
                        bohren: bool = True) -> np.ndarray:
        """Field of ``particle`` alone at the coordinates, the propagation
        phase exp(-i k z_p) included and no alpha, (3, npts) complex, a
        new array (theory/LorenzMie.py:135-158); times the particle's
        aberration mask when ``spherical`` is not 0 (theory/Aberrated.py:
        78-84)."""
        if self._npts == 0:
            return np.zeros((3, 0), complex)
        if getattr(self, 'spherical', 0.) != 0.:
            own, alpha = self.particle, self.alpha
            try:
                self.particle, self.alpha = particle, 1.
                return self.field(cartesian, bohren).copy()
            finally:
                self.particle, self.alpha = own, alpha
        need = self._particle_cap(particle)
        cur = self._batch_key[4] if self._batch_key else 0
        d = self._device_buffers(cur if need <= cur else need + 8)
        alpha = self.alpha
        try:
            self.alpha = 1.
            self._load_particle(particle, d)
        finally:
            self.alpha = alpha
        _lmz.check(_lmz.load_library().lmz_field(
            self._batch.handle, _lmz.ptr(d['params']), None,
            _lmz.ptr(d['nmax']), _lmz.ptr(d['coef']), _lmz.ptr(d['coords']),
            self._flags(cartesian, bohren), _lmz.ptr(d['field']),
            _lmz.stream_handle()))
        t = _lmz.torch()
        fld = t.view_as_complex(d['field'].view(3, self._npts, 2))
        return fld.cpu().numpy().astype(complex)

    def lorenzmie(self, ab: np.ndarray, kdr: np.ndarray,
                  cartesian: bool = True, bohren: bool = True) -> np.ndarray:
        """Field of coefficients ``ab`` ((N + 1, 2), row 0 unused) at the
        wavenumber-scaled displacements ``kdr`` (3, npts) from the particle,
        without the propagation phase (theory/LorenzMie.py:183-342): the
        per-order sum evaluated on the GPU for a particle at the origin."""
        kdr = np.atleast_2d(np.asarray(kdr, float))
        k = self.instrument.wavenumber()
        m = LorenzMie(coordinates=kdr / k, particle=_GivenCoefficients(ab),
                      instrument=self.instrument)
        return m.scattered_field(m.particle, cartesian, bohren)

    def hologram(self, cartesian: bool = True, bohren: bool = True,
                 device: bool = False):
        """|x + alpha E_s|^2 at the coordinates, (npts,) float64
        (theory/LorenzMie.py:113-133); ``device=True`` returns the float32
        tensor on the GPU."""
        if self._npts == 0:
            return np.zeros(0)
        d = self._device_buffers(self._cap())
        lib = _lmz.load_library()
        flags = self._flags(cartesian, bohren)
        stream = _lmz.stream_handle()
        if self._superposed():
            self._launch_superposed(d, flags, out=d['holo'])
        else:
            self._load_particle(self.particle[0], d)
            if not self.double_precision and cartesian:
                flags |= _lmz.FP32
            _lmz.check(lib.lmz_hologram(
                self._batch.handle, _lmz.ptr(d['params']), None,
                _lmz.ptr(d['nmax']), _lmz.ptr(d['coef']), _lmz.ptr(d['coords']),
                flags, None, _lmz.ptr(d['holo']), None, None, stream))
        holo = d['holo'][:self._npts]
        if device:
            return holo
        return holo.cpu().numpy().astype(np.float64)

    def residuals(self, data):
        """(I - d) / noise of a single particle on the GPU (float32 tensor),
        the residual of analysis/Optimizer.py:257-260; ``data`` is a float32
        tensor on the device with one value per coordinate."""
        if self._superposed():
            raise TypeError('residuals() supports a single unaberrated particle')
        d = self._device_buffers(self._cap())
        self._load_particle(self.particle[0], d)
        lib = _lmz.load_library()
        _lmz.check(lib.lmz_hologram(
            self._batch.handle, _lmz.ptr(d['params']), None, _lmz.ptr(d['nmax']),
            _lmz.ptr(d['coef']), _lmz.ptr(d['coords']),
            _lmz.CARTESIAN | _lmz.BOHREN, _lmz.ptr(data), _lmz.ptr(d['holo']),
            None, None, _lmz.stream_handle()))
        return d['holo'][:self._npts]

    def jac(self) -> dict:
        """dI/dp for p in jac_params at the current point, (npts,) float64
        each, from the forward-mode Jacobian kernel.  n_max is that of the
        current point (theory/jaxLorenzMie.py:359-361).  TypeError unless
        the particle is a single Sphere (theory/jaxLorenzMie.py:373-374)."""
        if not isinstance(self.particle, Sphere):
            raise TypeError('jac() supports single Sphere particles only')
        J = self.jac_device()
        Jh = J.cpu().numpy().astype(np.float64)
        return {name: Jh[i] for i, name in enumerate(PARAMS)}

    def jac_device(self):
        """(7, npts) float32 Jacobian tensor on the GPU."""
        t = _lmz.torch()
        d = self._device_buffers(self._cap())
        if d.get('jac') is None or d['jac'].numel() != 7 * max(self._npts, 1):
            d['jac'] = t.empty(7 * max(self._npts, 1), dtype=t.float32,
                               device=_lmz.device())
        self._load_particle(self.particle, d, tangents=True)
        lib = _lmz.load_library()
        _lmz.check(lib.lmz_jacobian(
            self._batch.handle, _lmz.ptr(d['params']), None, _lmz.ptr(d['nmax']),
            _lmz.ptr(d['coef']), _lmz.ptr(d['dcoef']), _lmz.ptr(d['coords']),
            _lmz.CARTESIAN | _lmz.BOHREN, None, _lmz.ptr(d['jac']), None, None,
            None, _lmz.stream_handle()))
        return d['jac'][:7 * self._npts].view(7, self._npts)
