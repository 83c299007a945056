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
        (analysis/Frame.py:63-80)."""
        self._features = []
        self._bboxes = []
        self._results = pd.DataFrame()
        self._data = None if data is None else np.asarray(data)
        if self._data is not None:
            h = Hologram(self._data, corner=self.corner)
            self._coordinates = h._coordinates

    @property
    def shape(self):
        return None if self._data is None else self._data.shape

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
        (analysis/Frame.py:103-115)."""
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
        """A Feature on the crop ``key`` with this frame's instrument."""
        crop = Hologram.__getitem__(self, key)
        model = LorenzMie(particle=Sphere(), instrument=self.instrument)
        return Feature(crop, model=model)

    # ------------------------------------------------------------ pipeline
    def detect(self) -> int:
        raise NotImplementedError('feature detection (Localizer) is not part '
                                  'of this library: set bboxes')

    def estimate(self) -> None:
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
        for idx in groups.values():
            for i, row in zip(idx, _fit_batch([self._features[i] for i in idx])):
                rows[i] = row
        self._results = pd.DataFrame(rows)
        return self._results

    def analyze(self, data=None) -> pd.DataFrame:
        if data is not None:
            self.data = data
        self.detect()
        self.estimate()
        return self.optimize()


def _fit_batch(features: list[Feature]) -> list[pd.Series]:
    """One LMBatch over the features' masked pixels; results written back
    to each feature's model and optimizer."""
    opt0 = features[0].optimizer
    if not opt0._use_device_lm():
        return [f.optimize() for f in features]
    crops, p0 = [], []
    for f in features:
        opt, model = f.optimizer, f.model
        opt._select(f)                        # mask pick, model.coordinates
        sph, inst = model.particle, model.instrument
        crops.append(dict(data=np.asarray(opt._data, np.float32),
                          width=model._width, origin=model._origin,
                          coords=model._rel, noise=inst.noise,
                          nmax_cap=_lmz.nmax_cap(sph.a_p, sph.n_p, sph.k_p,
                                                 inst.n_m, inst.wavelength)))
        props = model.properties
        p = np.array([props[k] for k in PARAMS], float)
        p[:3] += sph.r_0
        p0.append(p)
    lm = LMBatch(crops, features[0].model.instrument)
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
