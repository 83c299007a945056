"""Fit many holograms at once on the GPU.

The reference fits the features of a frame one at a time
(analysis/Frame.py:163-174 -> Feature.optimize -> Optimizer.optimize);
``BatchOptimizer.optimize`` takes all crops and starting spheres of a
batch and runs one device-resident LM over them (analysis/lm.py).  Rows
of the returned DataFrame carry the same keys as Optimizer.result.
"""
from __future__ import annotations

import numpy as np
import pandas as pd

from .. import _lmz
from ..theory import Instrument
from ..theory.LorenzMie import PARAMS
from .Hologram import Hologram
from .lm import LMBatch, LMSettings, stats_from_jtj


class BatchOptimizer:
    """Batched Levenberg-Marquardt for single-sphere holograms."""

    def __init__(self, instrument: Instrument | None = None,
                 variables=PARAMS, settings: LMSettings | None = None):
        self.instrument = instrument or Instrument()
        self.variables = list(variables)
        self.settings = settings or LMSettings()

    def prepare(self, holograms: list[Hologram], particles, noise=None,
                alpha=None) -> LMBatch:
        """Device batch for these crops and starting spheres."""
        inst = self.instrument
        crops, p0 = [], []
        for i, (h, s) in enumerate(zip(holograms, particles)):
            ny, nx = h.shape
            cap = _lmz.nmax_cap(s.a_p, s.n_p, s.k_p, inst.n_m, inst.wavelength)
            crops.append(dict(data=h.data, width=nx,
                              origin=(h.corner[0], h.corner[1], 0.),
                              nmax_cap=cap,
                              noise=None if noise is None else np.atleast_1d(noise)[
                                  i % np.size(noise)]))
            a = 1.0 if alpha is None else np.atleast_1d(alpha)[i % np.size(alpha)]
            p0.append([s.x_p + s.x_0, s.y_p + s.y_0, s.z_p + s.z_0, s.a_p,
                       s.n_p, s.k_p, a])
        lm = LMBatch(crops, inst)
        lm.load(np.array(p0))
        return lm

    def results(self, lm: LMBatch) -> pd.DataFrame:
        params, res = lm.fetch()
        vary = [PARAMS.index(v) for v in self.variables]
        rows = []
        for f in range(lm.nfeat):
            status = int(res[f, 38])
            ok = (status & (_lmz.LM_FTOL | _lmz.LM_XTOL | _lmz.LM_GTOL)) != 0 and \
                (status & (_lmz.LM_STALLED | _lmz.LM_NONFINITE)) == 0
            redchi, unc = stats_from_jtj(res[f, :28], res[f, 35], lm.npix[f], vary)
            row = {}
            for v, i, d in zip(self.variables, vary, unc):
                row[v] = params[f, i]
                row['d' + v] = d
            row.update(success=bool(ok), npix=lm.npix[f], redchi=redchi,
                       iterations=int(res[f, 37]), status=status)
            rows.append(row)
        return pd.DataFrame(rows)

    def optimize(self, holograms: list[Hologram], particles, noise=None,
                 alpha=None) -> pd.DataFrame:
        lm = self.prepare(holograms, particles, noise, alpha)
        lm.run(self.settings, self.variables)
        out = self.results(lm)
        lm.close()
        return out
