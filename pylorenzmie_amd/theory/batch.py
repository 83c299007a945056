"""Holograms of many particles with uneven crops in one launch.

The batched counterpart of theory/torchLorenzMieBatch.py:258-431
(``batch_hologram``), without its shared-``npts`` restriction: each
feature brings its own crop (dense grid) or coordinate list, its own
Mie coefficients (FP64 kernel, one lane per feature; homogeneous or
layered spheres) and one tile range of the FP64 forward kernel.
"""
from __future__ import annotations

import numpy as np

from .. import _lmz
from .Instrument import Instrument


class HologramBatch:
    """Device-resident description of many features' pixels.

    ``shapes``: list of (ny, nx) crops (dense grids at ``corners``), and
    ``caps``: n_max capacity per feature (orders reserved).
    """

    def __init__(self, shapes, caps, corners=None,
                 instrument: Instrument | None = None, flags: int = 3):
        t = _lmz.torch()
        dev = _lmz.device()
        self.instrument = instrument or Instrument()
        corners = corners or [(0., 0.)] * len(shapes)
        feats, off = [], 0
        self.offsets, self.shapes = [], []
        for (ny, nx), cap, (x0, y0) in zip(shapes, caps, corners):
            n = int(ny) * int(nx)
            feats.append(_lmz.make_feature(n, nx, (x0, y0, 0.), self.instrument,
                                           cap, off))
            self.offsets.append(off)
            self.shapes.append((int(ny), int(nx)))
            off += n
        self.batch = _lmz.Batch(feats, 1)
        self.nfeat = len(feats)
        self.npix_total = off
        b = self.batch
        self.params = t.zeros((self.nfeat, _lmz.NPARAM), dtype=t.float64,
                              device=dev)
        self.nmax = t.zeros(self.nfeat, dtype=t.int32, device=dev)
        self.coef = t.zeros(b.coef_floats(), dtype=t.float32, device=dev)
        self.work = t.zeros(b.work_doubles, dtype=t.float64, device=dev)
        self.out = t.empty(off, dtype=t.float32, device=dev)
        self.flags = flags

    def load(self, params: np.ndarray) -> None:
        """(nfeat, 7) rows of x_p y_p z_p (absolute) a_p n_p k_p alpha."""
        t = _lmz.torch()
        p = np.zeros((self.nfeat, _lmz.NPARAM))
        p[:, :7] = np.asarray(params, float)[:, :7]
        self.params.copy_(t.from_numpy(p))

    def load_layers(self, layers_list) -> None:
        """Layered particles: per feature an (L, 3) array of outer radius
        [um], n, k, innermost first (one row for a homogeneous sphere).
        Their coefficients then come from lmz_mie_ab_layered and do not
        depend on params' a_p, n_p, k_p columns."""
        from .LayeredSphere import layer_table
        t = _lmz.torch()
        if len(layers_list) != self.nfeat:
            raise ValueError(f'need {self.nfeat} layer tables')
        table, counts = layer_table(layers_list)
        dev = self.params.device
        self.layers = (t.from_numpy(table).to(dev), t.from_numpy(counts).to(dev))

    def coefficients(self, stream=None):
        """Mie coefficient tables of every feature (homogeneous spheres from
        params, or the loaded layer tables); returns nmax per feature."""
        lib = _lmz.load_library()
        s = _lmz.stream_handle(stream)
        b = self.batch
        layers = getattr(self, 'layers', None)
        if layers is None:
            _lmz.check(lib.lmz_mie_ab(b.handle, _lmz.ptr(self.params), None,
                                      _lmz.ptr(self.nmax), _lmz.ptr(self.coef),
                                      None, None, _lmz.ptr(self.work), s))
        else:
            _lmz.check(lib.lmz_mie_ab_layered(
                b.handle, _lmz.ptr(layers[0]), _lmz.ptr(layers[1]),
                _lmz.ptr(self.nmax), _lmz.ptr(self.coef), None,
                _lmz.ptr(self.work), s))
        return self.nmax

    def compute(self, stream=None):
        """Mie coefficients then all holograms; returns the packed tensor."""
        lib = _lmz.load_library()
        s = _lmz.stream_handle(stream)
        b = self.batch
        self.coefficients(stream)
        _lmz.check(lib.lmz_hologram(b.handle, _lmz.ptr(self.params), None,
                                    _lmz.ptr(self.nmax), _lmz.ptr(self.coef),
                                    None, self.flags, None, _lmz.ptr(self.out),
                                    None, None, s))
        return self.out

    def holograms(self) -> list[np.ndarray]:
        """The float32 holograms on the host, one (ny, nx) view per crop of
        one fresh array in pinned memory (torch recycles it once every view
        is dropped); a pageable copy of a large batch runs several times
        slower."""
        t = _lmz.torch()
        host = t.empty(self.out.shape, dtype=self.out.dtype, pin_memory=True)
        host.copy_(self.out, non_blocking=True)
        t.cuda.current_stream().synchronize()
        host = host.numpy()
        return [host[o:o + ny * nx].reshape(ny, nx)
                for o, (ny, nx) in zip(self.offsets, self.shapes)]


def batch_hologram(shapes, params, instrument: Instrument | None = None,
                   corners=None) -> list[np.ndarray]:
    """Float32 holograms of spheres ``params`` (nfeat, 7) on crops
    ``shapes``; one Mie launch and one pixel launch for the whole batch."""
    inst = instrument or Instrument()
    params = np.asarray(params, float)
    caps = [_lmz.nmax_cap(p[3], p[4], p[5], inst.n_m, inst.wavelength, 0.)
            for p in params]
    hb = HologramBatch(shapes, caps, corners, inst)
    hb.load(params)
    hb.compute()
    return hb.holograms()
