"""A full camera frame and its features, mirroring analysis/Frame.py:9-190.

``Frame.optimize`` replaces the reference's loop over features
(analysis/Frame.py:163-174, one scipy fit after another) with one
device-resident batched LM over every feature: each feature's Mask picks
its pixels (a fresh random subset per call, as Feature.optimize draws),
the picked pixels travel as an explicit coordinate list, and all features
share one HIP graph per LM iteration.  Features are grouped by their
optimizer's variables and settings; each group is one batch.  Fitted
values are written back to every feature's model and optimizer, so
``feature.optimizer.result`` and ``report()`` work as after a single fit.

The frame may live in device memory (a torch tensor on the GPU, for
instance the output of ``utilities.Normalizer`` on a camera frame already
there): its features' pixels then go from the frame to the batch on the
device (lmz_gather_frame), and only the boxes and the picked pixel indices
come from the host.

Feature detection (the reference's Localizer, which needs trackpy) is
outside this library: set ``bboxes`` to place features.
"""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pandas as pd

from .. import _lmz
from ..theory import Instrument, LorenzMie, Sphere
from ..theory.LorenzMie import PARAMS
from ..lib import meshgrid
from ..utilities.normalizer import is_device_tensor
from .Feature import Feature
from .Hologram import Hologram
from .lm import LMBatch


class Frame(Hologram):
    """Normalized hologram of a camera field with features in it."""

    def __init__(self, data=None, corner=(0., 0.),
                 instrument: Instrument | None = None, localizer=None) -> None:
        if localizer is not None:
            raise NotImplementedError('feature detection (Localizer) is not '
                                      'part of this library: set bboxes')
        self.instrument = instrument or Instrument()
        self._features: list[Feature] = []
        self._bboxes: list = []
        self._results = pd.DataFrame()
        self.corner = corner
        self._coordinates = None
        self._data = None
        if data is not None:
            self.data = data

    # ------------------------------------------------------------ state
    @property
    def data(self):
        return self._data

    @data.setter
    def data(self, data) -> None:
        """Setting data clears features, boxes and results
        (analysis/Frame.py:63-80).  A device tensor stays on the device as
        float32 (the type the fitter reads)."""
        self._features = []
        self._bboxes = []
        self._results = pd.DataFrame()
        if data is None:
            self._data = None
            self._coordinates = None
            return
        if is_device_tensor(data):
            import torch
            if data.dim() != 2:
                raise ValueError('frame data must be 2-D')
            self._data = data.to(torch.float32).contiguous()
        else:
            self._data = np.asarray(data)
        self._coordinates = meshgrid(tuple(self._data.shape), corner=self.corner,
                                     flatten=False)

    @property
    def shape(self) -> tuple[int, int]:
        """(ny, nx), or (0, 0) without data (analysis/Frame.py:81-86)."""
        return (0, 0) if self._data is None else tuple(self._data.shape)

    @property
    def on_device(self) -> bool:
        return is_device_tensor(self._data)

    @property
    def results(self) -> pd.DataFrame:
        return self._results

    @property
    def features(self) -> list[Feature]:
        return self._features

    @property
    def bboxes(self) -> list:
        return self._bboxes

    @bboxes.setter
    def bboxes(self, bboxes) -> None:
        """Square crops ((x0, y0), w, h) with a sphere at their centre
        (analysis/Frame.py:103-115); a single box is taken as a list of
        one."""
        if (isinstance(bboxes, tuple) and len(bboxes) == 3
                and isinstance(bboxes[0], tuple)):
            bboxes = [bboxes]
        self._features = []
        self._bboxes = list(bboxes)
        for (x0, y0), w, h in self._bboxes:
            dim = int(min(w, h))
            x0, y0 = int(x0), int(y0)
            feature = self[y0:y0 + dim, x0:x0 + dim]
            feature.particle.x_p = x0 + dim / 2.
            feature.particle.y_p = y0 + dim / 2.
            self._features.append(feature)

    def __getitem__(self, key) -> Feature:
        """A Feature on the crop ``key`` with this frame's instrument; it
        remembers its box in the frame."""
        sy, sx = key
        ny, nx = self.shape
        y0, y1, ystep = sy.indices(ny)
        x0, x1, xstep = sx.indices(nx)
        crop = Hologram.__getitem__(self, key)
        model = LorenzMie(particle=Sphere(), instrument=self.instrument)
        feature = Feature(crop, model=model)
        if ystep == 1 and xstep == 1 and y1 > y0 and x1 > x0:
            feature._box = (x0, y0, x1 - x0, y1 - y0)
        return feature

    # ------------------------------------------------------------ pipeline
    def detect(self) -> int:
        """Without data: no features, returns 0 (analysis/Frame.py:133-145).
        Detection itself (the Localizer) is not part of this library."""
        if self._data is None:
            self._features = []
            self._bboxes = []
            return 0
        raise NotImplementedError('feature detection (Localizer) is not part '
                                  'of this library: set bboxes')

    def estimate(self, batched: bool = False) -> None:
        """Starting values of every feature: one DE search after another
        (analysis/Frame.py:158-161), each generation's population on the
        GPU; or, with ``batched``, all searches run together on the device
        (DEEstimator.estimate_batch, lmz_de_run: every generation of every
        feature's search in four launches, no host work between
        generations; the features' estimators must share bounds keys,
        popsize and settings)."""
        if batched and self.features:
            from .DEEstimator import estimate_batch
            estimate_batch([f.estimator for f in self.features], self.features)
            return
        for feature in self.features:
            feature.estimate()

    def optimize(self) -> pd.DataFrame:
        """Fit every feature; one device batch per group of features with
        the same variables and settings.  Returns one row per feature with
        the keys of Optimizer.result."""
        groups: dict = {}
        for i, f in enumerate(self._features):
            opt = f.optimizer
            key = (tuple(opt.variables), tuple(sorted(opt.settings.items())))
            groups.setdefault(key, []).append(i)
        rows: list = [None] * len(self._features)
        source = self if self.on_device else None
        for idx in groups.values():
            group = [self._features[i] for i in idx]
            for i, row in zip(idx, _fit_batch(group, source)):
                rows[i] = row
        self._results = pd.DataFrame(rows)
        return self._results

    def analyze(self, data=None) -> pd.DataFrame:
        if data is not None:
            self.data = data
        self.detect()
        self.estimate()
        return self.optimize()


def _select_on_frame(feature: Feature) -> dict:
    """The pixels a feature's optimizer fits, as a box and crop-local
    indices in the frame: the same mask draw and model coordinates as
    Optimizer._select (analysis/Optimizer.py:211-215), without the data."""
    opt, model = feature.optimizer, feature.model
    box = getattr(feature, '_box', None)
    if box is None:
        raise ValueError('a feature of a device frame must be a box slice of it')
    if opt.mask is None:
        model.coordinates = feature.flat_coordinates
        index = None
    else:
        m = opt.mask
        m.shape = tuple(feature.shape)          # draws a fresh subset
        picked = m()
        index = np.flatnonzero(picked).astype(np.int32)
        model.coordinates = feature.coordinates[:, picked]
    npix = model._npts
    if model._width > 0:
        # the picked pixels form a full grid (all of the box, or a
        # rectangle inside it): a dense crop of that rectangle
        i0 = int(index[0]) if index is not None else 0
        x0, y0, w, _ = box
        box = (x0 + i0 % w, y0 + i0 // w, model._width, npix // model._width)
        index = None
    elif index is None:
        index = np.arange(npix, dtype=np.int32)
    opt._data = None                # the pixels stay in device memory
    opt._count = npix
    return dict(box=box, npix=npix, index=index)


def _fit_batch(features: list[Feature], frame: Frame | None = None
               ) -> list[pd.Series]:
    """One LMBatch over the features' masked pixels; results written back
    to each feature's model and optimizer.  With ``frame`` (in device
    memory) the pixels are gathered there."""
    opt0 = features[0].optimizer
    if not opt0._use_device_lm():
        return [f.optimize() for f in features]
    crops, p0 = [], []
    for f in features:
        opt, model = f.optimizer, f.model
        if frame is None:
            opt._select(f)                    # mask pick, model.coordinates
            crop = dict(data=np.asarray(opt._data, np.float32),
                        coords=model._rel)
        else:
            crop = _select_on_frame(f)
        sph, inst = model.particle, model.instrument
        crop.update(width=model._width, origin=model._origin, noise=inst.noise,
                    nmax_cap=_lmz.nmax_cap(sph.a_p, sph.n_p, sph.k_p,
                                           inst.n_m, inst.wavelength))
        crops.append(crop)
        props = model.properties
        p = np.array([props[k] for k in PARAMS], float)
        p[:3] += sph.r_0
        p0.append(p)
    source = None
    if frame is not None:
        xy = frame._coordinates
        source = (frame.data, xy[0, 0, :], xy[1, :, 0])
    lm = LMBatch(crops, features[0].model.instrument, frame=source)
    lm.load(np.array(p0))
    lm.run(opt0._lm_settings(), opt0.variables)
    params, res = lm.fetch()
    lm.close()
    vary = [PARAMS.index(v) for v in opt0.variables]
    out = []
    for i, f in enumerate(features):
        opt, model = f.optimizer, f.model
        x = params[i, :7].copy()
        x[:3] -= model.particle.r_0
        model.properties = dict(zip(PARAMS, x))
        status = int(res[i, 38])
        ok = (status & (_lmz.LM_FTOL | _lmz.LM_XTOL | _lmz.LM_GTOL)) != 0 and \
            (status & (_lmz.LM_STALLED | _lmz.LM_NONFINITE)) == 0
        opt._result = SimpleNamespace(
            x=np.array([x[PARAMS.index(v)] for v in opt.variables]),
            success=bool(ok), cost=0.5 * res[i, 35], status=status,
            nit=int(res[i, 37]), njev=int(res[i, 39]),
            jtj=res[i, :28].copy(), vary=vary, jac=None)
        out.append(opt.result)
    return out
