"""Fit the Lorenz-Mie model to holograms, mirroring analysis/Optimizer.py.

``Optimizer`` keeps the reference's constructor, settings dict, fixed /
variables bookkeeping, ``optimize(hologram) -> pandas.Series`` contract
(keys [p, dp] per variable, success, npix, redchi), ``metadata`` and
``report()`` (analysis/Optimizer.py:13-278).  The least-squares step runs
on the GPU: batched Levenberg-Marquardt with the forward-mode Jacobian
kernel (analysis/lm.py).  Robust losses and variables outside the seven
Jacobian parameters are driven by scipy on the host with GPU residuals and
Jacobians, as the reference drives MINPACK.

``hipOptimizer`` is the reference's cupyOptimizer pattern
(analysis/cupyOptimizer.py:47-52): GPU residuals copied to the host at
every evaluation and scipy's MINPACK forward-difference LM.  It is the
plain GPU finite-difference form the batched fitter is measured against.
"""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pandas as pd

from .. import _lmz
from ..lib import LMObject
from ..theory import LorenzMie, Sphere
from ..theory.LorenzMie import PARAMS
from .Hologram import Hologram
from .Mask import Mask, host_data
from .lm import LMBatch, LMSettings, stats_from_jtj

_DEFAULT_FIXED = ('wavelength magnification numerical_aperture n_m noise '
                  'k_p alpha').split()


def _default_settings() -> dict:
    """analysis/Optimizer.py:104-112."""
    return {'method': 'lm', 'ftol': 1e-4, 'xtol': 1e-6, 'gtol': 1e-6,
            'loss': 'linear', 'max_nfev': 2000, 'diff_step': 1e-5,
            'x_scale': 'jac'}


class Optimizer(LMObject):
    """Fit a LorenzMie model to a normalized Hologram on the GPU."""

    method: str = 'hip'

    def __init__(self, model=None, mask: Mask | None = None,
                 robust: bool = False, fixed: list[str] | None = None,
                 settings: dict | None = None, **kwargs) -> None:
        self.model = model or LorenzMie(**kwargs)
        if self.method not in self.model.method:
            raise TypeError('Model not compatible with Optimizer')
        self.mask = mask
        self.settings = settings
        self.fixed = fixed if fixed is not None else list(_DEFAULT_FIXED)
        self.robust = robust
        self._data = None
        self._count = 0             # pixels of the last fit
        self._result = None
        self._lm_cache = None

    # -------------------------------------------------------- configuration
    @property
    def properties(self) -> dict:
        return dict(settings=self.settings, fixed=self.fixed)

    @properties.setter
    def properties(self, properties: dict) -> None:
        for name, value in properties.items():
            if hasattr(self, name):
                setattr(self, name, value)

    @property
    def settings(self) -> dict:
        return self._settings

    @settings.setter
    def settings(self, settings: dict | None) -> None:
        self._settings = settings if settings is not None else _default_settings()

    @property
    def robust(self) -> bool:
        return self.settings['loss'] != 'linear'

    @robust.setter
    def robust(self, robust: bool) -> None:
        if robust:
            self.settings['method'] = 'trf'
            self.settings['loss'] = 'cauchy'
        else:
            self.settings['method'] = 'lm'
            self.settings['loss'] = 'linear'

    @property
    def fraction(self) -> float:
        return 1.0 if self.mask is None else self.mask.fraction

    @fraction.setter
    def fraction(self, fraction: float) -> None:
        if self.mask is None:
            self.mask = Mask()
        self.mask.fraction = fraction

    @property
    def fixed(self) -> list[str]:
        return self._fixed

    @fixed.setter
    def fixed(self, fixed: list[str]) -> None:
        self._fixed = list(fixed)
        self._variables = [p for p in self.model.properties if p not in fixed]

    @property
    def variables(self) -> list[str]:
        return self._variables

    @variables.setter
    def variables(self, variables: list[str]) -> None:
        self._variables = list(variables)
        self._fixed = [p for p in self.model.properties if p not in variables]

    # -------------------------------------------------------------- results
    @property
    def result(self) -> pd.Series | None:
        if self._result is None:
            return None
        a = self.variables
        b = ['d' + v for v in a]
        keys = list(sum(zip(a, b), ()))
        keys.extend('success npix redchi'.split())
        redchi, unc = self._statistics()
        values = list(sum(zip(list(self._result.x), unc), ()))
        values.extend([self._result.success, self._count, redchi])
        return pd.Series(dict(zip(keys, values)))

    @property
    def metadata(self) -> pd.Series:
        props = self.model.properties
        metadata = {key: props[key] for key in self.fixed if key in props}
        metadata.update(self.settings)
        return pd.Series(metadata)

    def report(self) -> str:
        result = self.result
        if result is None:
            raise RuntimeError('optimize() must be called before report()')
        units = {'x_p': 'pixels', 'y_p': 'pixels', 'z_p': 'pixels', 'a_p': 'μm'}
        fmt = {'x_p': '.2f', 'y_p': '.2f', 'z_p': '.2f', 'a_p': '.3f'}
        lines = []
        for p in self.variables:
            f = fmt.get(p, '.4f')
            suffix = f' {units[p]}' if p in units else ''
            lines.append(f'{p} = {result[p]:{f}} ± {result["d" + p]:{f}}{suffix}')
        lines += [f'npixels = {result.npix:.0f}', f'χ² = {result.redchi:.2f}']
        return '\n'.join(lines)

    # ------------------------------------------------------------- fitting
    def _select(self, hologram: Hologram) -> None:
        """analysis/Optimizer.py:211-215."""
        if self.mask is None:
            self._data = host_data(hologram.data).ravel()
            self.model.coordinates = hologram.flat_coordinates
        else:
            self._data, self.model.coordinates = self.mask.apply(hologram)
        self._count = self._data.size

    def _use_device_lm(self) -> bool:
        s = self.settings
        return (s.get('method', 'lm') == 'lm' and s.get('loss', 'linear') == 'linear'
                and isinstance(self.model.particle, Sphere)
                and set(self.variables) <= set(PARAMS)
                and len(self.variables) > 0)

    def optimize(self, hologram: Hologram) -> pd.Series:
        """Fit the model to ``hologram``; returns ``self.result``."""
        self._select(hologram)
        if self._use_device_lm():
            self._optimize_device()
        else:
            self._optimize_host()
        return self.result

    def _lm_settings(self) -> LMSettings:
        s = self.settings
        max_nfev = s.get('max_nfev') or 2000
        return LMSettings(max_iter=int(max(1, min(max_nfev, 500))),
                          ftol=float(s.get('ftol') or 0.),
                          xtol=float(s.get('xtol') or 0.),
                          gtol=float(s.get('gtol') or 0.))

    def _optimize_device(self) -> None:
        model = self.model
        sph = model.particle
        inst = model.instrument
        props = model.properties
        cap = _lmz.nmax_cap(sph.a_p, sph.n_p, sph.k_p, inst.n_m, inst.wavelength)
        crop = dict(data=np.asarray(self._data, np.float32), nmax_cap=cap,
                    width=model._width, origin=model._origin,
                    coords=model._rel)
        lm = LMBatch([crop], inst)
        p0 = np.array([[props[k] for k in PARAMS]])
        p0[0, :3] += sph.r_0
        lm.load(p0)
        lm.run(self._lm_settings(), self.variables)
        params, res = lm.fetch()
        lm.close()
        x = params[0, :7].copy()
        x[:3] -= sph.r_0
        model.properties = dict(zip(PARAMS, x))
        status = int(res[0, 38])
        ok = (status & (_lmz.LM_FTOL | _lmz.LM_XTOL | _lmz.LM_GTOL)) != 0 and \
            (status & (_lmz.LM_STALLED | _lmz.LM_NONFINITE)) == 0
        vary = [PARAMS.index(v) for v in self.variables]
        self._result = SimpleNamespace(
            x=np.array([x[PARAMS.index(v)] for v in self.variables]),
            success=bool(ok), cost=0.5 * res[0, 35], status=status,
            nit=int(res[0, 37]), njev=int(res[0, 39]),
            jtj=res[0, :28].copy(), vary=vary, jac=None)

    # scipy-driven path: residuals and Jacobians from the GPU
    def _residuals(self, values) -> np.ndarray:
        self.model.properties = dict(zip(self.variables, values))
        noise = self.model.instrument.noise
        return (self.model.hologram() - self._data) / noise

    def _analytical_jac(self, values) -> np.ndarray:
        self.model.properties = dict(zip(self.variables, values))
        J = self.model.jac()
        noise = self.model.instrument.noise
        return np.column_stack([J[p] for p in self.variables]) / noise

    def _optimize_host(self) -> None:
        from scipy.optimize import least_squares
        p0 = np.array([self.model.properties[p] for p in self.variables])
        settings = dict(self.settings)
        jac_params = getattr(self.model, 'jac_params', None)
        if (jac_params is not None and set(self.variables) <= jac_params and
                isinstance(self.model.particle, Sphere)):
            settings['jac'] = self._analytical_jac
        self._result = least_squares(self._residuals, p0, **settings)

    def _statistics(self):
        res = self._result
        npix = self._count
        if getattr(res, 'jtj', None) is not None:
            return stats_from_jtj(res.jtj, 2. * res.cost, npix, res.vary)
        # analysis/Optimizer.py:262-278
        from scipy.linalg import svd
        ndeg = npix - res.x.size
        redchi = 2. * res.cost / ndeg
        _, s, VT = svd(res.jac, full_matrices=False)
        threshold = np.finfo(float).eps * max(res.jac.shape) * s[0]
        s = s[s > threshold]
        VT = VT[:s.size]
        pcov = np.dot(VT.T / s ** 2, VT)
        return redchi, np.sqrt(redchi * np.diag(pcov))


class hipOptimizer(Optimizer):
    """Plain GPU form: GPU residuals, host MINPACK with forward differences.

    Every residual evaluation launches the Mie and residual kernels and
    copies the residual vector to the host, as cupyOptimizer._residuals
    does (analysis/cupyOptimizer.py:47-52); scipy estimates the Jacobian
    by forward differences with step diff_step * |x| (Optimizer settings).
    """

    method: str = 'hip'

    def optimize(self, hologram: Hologram) -> pd.Series:
        from scipy.optimize import least_squares
        self._select(hologram)
        t = _lmz.torch()
        self._devdata = t.from_numpy(
            np.ascontiguousarray(self._data, np.float32)).to(_lmz.device())
        p0 = np.array([self.model.properties[p] for p in self.variables])
        self._result = least_squares(self._residuals, p0, **self.settings)
        return self.result

    def _residuals(self, values) -> np.ndarray:
        self.model.properties = dict(zip(self.variables, values))
        if len(self.model.particle) != 1:
            return super()._residuals(values)
        r = self.model.residuals(self._devdata)
        return r.cpu().numpy().astype(np.float64)
