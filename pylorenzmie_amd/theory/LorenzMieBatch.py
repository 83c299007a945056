"""B holograms at one set of coordinates, each of its own particles, in one
pass over the pixels: the batched path of the reference's
theory/torchLorenzMieBatch.py:218-431 (TorchLorenzMieBatch.batch_hologram
and batch_field).

All B x P constituents go through one launch of the superposition kernel
(lmz_hologram_superposed, FP64), the spheres among them through one batched
Mie launch (lmz_mie_ab); other particle types bring their own a_n, b_n
(Particle.ab) through lmz_coef_from_ab.  The reference instead loops over
particle slots in Python with one Triton launch per slot and pads short
lists with zero-coefficient dummies (:316-373); here a feature's particle
range simply ends.
"""
from __future__ import annotations

import numpy as np

from .. import _lmz
from .LorenzMie import LorenzMie
from .Sphere import Sphere


class LorenzMieBatch(LorenzMie):
    """LorenzMie with TorchLorenzMieBatch's batched methods
    (theory/torchLorenzMieBatch.py:218-245); the single-hologram API is
    inherited.  A Cluster or Dimer in a particle list contributes its
    members."""

    def batch_hologram(self, particle_lists: list, cartesian: bool = True,
                       bohren: bool = True, device: bool = False):
        """(B, npts) float64 holograms |x + alpha E_s|^2, one per particle
        list (theory/torchLorenzMieBatch.py:258-283); ``device=True``
        returns the float32 tensor on the GPU instead, as hologram() does
        (no copy to the host)."""
        B, P = len(particle_lists), self._npts
        out = self._run_batch(particle_lists, cartesian, bohren, hologram=True)
        if out is None:
            if device:
                t = _lmz.torch()
                return t.ones(B, P, dtype=t.float32, device=_lmz.device())
            return np.ones((B, P))
        if device:
            return out.view(B, P)
        # a fresh float64 array in pinned host memory (torch recycles the
        # blocks once the array is dropped): a pageable copy of this size
        # runs several times slower
        t = _lmz.torch()
        host = t.empty((B, P), dtype=t.float64, pin_memory=True)
        host.copy_(out.view(B, P).double(), non_blocking=True)
        t.cuda.current_stream().synchronize()
        return host.numpy()

    def batch_field(self, particle_lists: list, cartesian: bool = True,
                    bohren: bool = True):
        """(B, 3, npts) complex64 tensor on the GPU: the scattered field
        alpha E_s of each particle list (theory/torchLorenzMieBatch.py:
        285-373)."""
        t = _lmz.torch()
        B, P = len(particle_lists), self._npts
        out = self._run_batch(particle_lists, cartesian, bohren, hologram=False)
        if out is None:
            return t.zeros(B, 3, P, dtype=t.complex64, device=_lmz.device())
        fld = t.view_as_complex(out.view(3, B, P, 2))
        return fld.permute(1, 0, 2).contiguous()

    # ------------------------------------------------------------ device
    def _allocate(self) -> None:
        super()._allocate()
        self._cache = {}          # batches kept between calls

    def _cached(self, name, key, make):
        """The object ``make()`` builds for ``key``, kept until the key or
        the coordinates change."""
        hit = self._cache.get(name)
        if hit is None or hit[0] != key:
            hit = (key, make())
            self._cache[name] = hit
        return hit[1]

    def _run_batch(self, particle_lists, cartesian, bohren, hologram):
        """The superposition launch over B features; None when there is
        nothing to scatter (no lists, no pixels or no particles)."""
        members = [[q for p in plist for q in p] for plist in particle_lists]
        flat = [q for m in members for q in m]
        B, P = len(members), self._npts
        if B == 0 or P == 0 or not flat:
            return None
        t = _lmz.torch()
        dev = _lmz.device()
        lib = _lmz.load_library()
        stream = _lmz.stream_handle()
        inst = self.instrument
        ikey = (float(np.real(inst.n_m)), float(inst.wavelength),
                float(inst.magnification), float(inst.noise))
        n = len(flat)
        pcap = max(self._particle_cap(q) for q in flat)
        pb = self._cached('particles', (n, pcap, ikey), lambda: _lmz.Batch(
            [_lmz.make_feature(1, 1, (0., 0., 0.), inst, pcap, i)
             for i in range(n)], 1))
        pp = np.zeros((n, _lmz.NPARAM))
        pp[:, 6] = 1.
        for i, q in enumerate(flat):
            pp[i, :3] = np.asarray(q.r_p, float) + np.asarray(q.r_0, float)
            if isinstance(q, Sphere):
                pp[i, 3:6] = q.a_p, q.n_p, q.k_p
        params = t.from_numpy(pp).to(dev)
        pnmax = t.zeros(n, dtype=t.int32, device=dev)
        coef = t.zeros(pb.coef_floats(), dtype=t.float32, device=dev)
        if all(isinstance(q, Sphere) for q in flat):
            work = t.zeros(pb.work_doubles, dtype=t.float64, device=dev)
            _lmz.check(lib.lmz_mie_ab(pb.handle, _lmz.ptr(params), None,
                                      _lmz.ptr(pnmax), _lmz.ptr(coef), None,
                                      None, _lmz.ptr(work), stream))
        else:
            rows = np.zeros((n * pcap, 4))
            nin = np.zeros(n, np.int32)
            for i, q in enumerate(flat):
                ab = np.asarray(q.ab(inst.n_m, inst.wavelength), complex)
                m = min(ab.shape[0] - 1, pcap)
                nin[i] = m
                r = rows[i * pcap:i * pcap + m]
                r[:, 0], r[:, 1] = ab[1:m + 1, 0].real, ab[1:m + 1, 0].imag
                r[:, 2], r[:, 3] = ab[1:m + 1, 1].real, ab[1:m + 1, 1].imag
            ab_d = t.from_numpy(rows.ravel()).to(dev)
            nin_d = t.from_numpy(nin).to(dev)
            _lmz.check(lib.lmz_coef_from_ab(pb.handle, _lmz.ptr(ab_d),
                                            _lmz.ptr(nin_d), _lmz.ptr(pnmax),
                                            _lmz.ptr(coef), stream))
        # the B features: the model's pixels, once per particle list
        fb = self._cached('pixels', (B, ikey), lambda: _lmz.Batch(
            [_lmz.make_feature(P, self._width, self._origin, inst, 1, b * P)
             for b in range(B)], 1))
        coords = None
        if self._rel is not None:
            coords = self._cached('coords', B, lambda: t.from_numpy(
                np.ascontiguousarray(np.tile(self._rel.reshape(3, 1, P),
                                             (1, B, 1)).reshape(3, B * P))).to(dev))
        offset = t.from_numpy(np.concatenate(
            [[0], np.cumsum([len(m) for m in members])]).astype(np.int32)).to(dev)
        alpha = t.full((B,), float(self.alpha), dtype=t.float64, device=dev)
        if hologram:
            out = t.empty(B * P, dtype=t.float32, device=dev)
            holo, fld = out, None
        else:
            out = t.empty(3 * B * P * 2, dtype=t.float32, device=dev)
            holo, fld = None, out
        _lmz.check(lib.lmz_hologram_superposed(
            fb.handle, n, _lmz.ptr(offset), _lmz.ptr(params), _lmz.ptr(pnmax),
            _lmz.ptr(coef), pcap, _lmz.ptr(coords), self._flags(cartesian, bohren),
            float(getattr(self, 'spherical', 0.)), _lmz.ptr(alpha),
            _lmz.ptr(holo), _lmz.ptr(fld), stream))
        return out
