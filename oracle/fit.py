"""The reference's fitter in float64: scipy MINPACK Levenberg-Marquardt on
the oracle's hologram, with the reference's settings, residual scaling and
SVD uncertainties.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): the checker for the GPU
Levenberg-Marquardt fitter and the CPU timing leg of the benchmark, never
called by the product path.

The reference's numpy LorenzMie has no analytical Jacobian (jac_params is
defined only on theory/jaxLorenzMie.py:321), so analysis/Optimizer.py:216-221
hands scipy.optimize.least_squares the residuals alone and MINPACK's lmdif
builds the Jacobian by forward differences with step diff_step * |x|.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import svd
from scipy.optimize import least_squares

from . import lorenzmie as lm

__all__ = ['default_settings', 'residuals', 'statistics', 'fit']

VARS6 = ('x_p', 'y_p', 'z_p', 'a_p', 'n_p', 'k_p')


def default_settings() -> dict:
    """analysis/Optimizer.py:104-112."""
    return {'method': 'lm', 'ftol': 1e-4, 'xtol': 1e-6, 'gtol': 1e-6,
            'loss': 'linear', 'max_nfev': 2000, 'diff_step': 1e-5,
            'x_scale': 'jac'}


def layered_properties(p, layers, instrument):
    """Properties of a layered sphere in the fitting parametrization:
    ``layers`` (L, 3) rows of outer radius, n, k fix the radius fractions
    r_l / r_L and the inner indices; a_p is the outer radius (all radii
    scale with it) and n_p + i k_p the outer layer's index.  The
    coefficients come from Yang's recursion with its own n_max.  The
    reference fits only homogeneous spheres; this model extends its
    residuals (analysis/Optimizer.py:257-260) to the layered particle."""
    lay = np.asarray(layers, float).reshape(-1, 3)
    r = lay[:, 0] * (p['a_p'] / lay[-1, 0])
    n = lay[:, 1].copy()
    k = lay[:, 2].copy()
    n[-1], k[-1] = p['n_p'], p['k_p']
    q = dict(p)
    q['ab'] = lm.mie_coefficients_layered(r, n, k, instrument.get('n_m', 1.340),
                                          instrument.get('wavelength', 0.447))
    return q


def residuals(values, variables, fixed, data, coordinates, instrument,
              noise, layers=None):
    """(hologram - data) / noise at the properties ``fixed`` updated with
    ``values`` of ``variables``; analysis/Optimizer.py:257-260.  With
    ``layers`` the particle is layered (:func:`layered_properties`)."""
    p = dict(fixed)
    p.update(zip(variables, values))
    if layers is not None:
        p = layered_properties(p, layers, instrument)
    holo = lm.hologram(coordinates, p, alpha=p.get('alpha', 1.),
                       **instrument)
    return (holo - data) / noise


def statistics(jac, cost, npix, nvar):
    """Reduced chi-squared and parameter uncertainties from the Jacobian
    by Moore-Penrose SVD with small singular values dropped;
    analysis/Optimizer.py:262-278."""
    redchi = 2. * cost / (npix - nvar)
    _, s, VT = svd(jac, full_matrices=False)
    threshold = np.finfo(float).eps * max(jac.shape) * s[0]
    s = s[s > threshold]
    VT = VT[:s.size]
    pcov = np.dot(VT.T / s ** 2, VT)
    return redchi, np.sqrt(redchi * np.diag(pcov))


def fit(data, coordinates, start, variables=VARS6, instrument=None,
        noise=0.05, settings=None, layers=None):
    """Fit the hologram model to ``data`` (npts,) at ``coordinates``.

    ``start`` maps property names to starting values (x_p .. k_p, and
    alpha, which defaults to 1); ``variables`` are the ones fitted, the
    rest stay at their start.  ``settings`` update the reference's
    defaults.  Returns the reference's result record
    (analysis/Optimizer.py:166-186): value and 'd'-prefixed uncertainty
    per variable, success, npix, redchi; plus nfev.  ``layers`` fits a
    layered sphere (:func:`layered_properties`).
    """
    instrument = dict(instrument or {})
    s = default_settings()
    s.update(settings or {})
    data = np.asarray(data, float).ravel()
    fixed = dict(start)
    variables = list(variables)
    p0 = np.array([fixed.get(v, 1. if v == 'alpha' else 0.)
                   for v in variables], float)
    res = least_squares(residuals, p0, args=(variables, fixed, data,
                                             coordinates, instrument, noise,
                                             layers),
                        **s)
    redchi, unc = statistics(res.jac, res.cost, data.size, len(variables))
    out = {}
    for v, x, d in zip(variables, res.x, unc):
        out[v] = float(x)
        out['d' + v] = float(d)
    out.update(success=bool(res.success), npix=int(data.size),
               redchi=float(redchi), nfev=int(res.nfev))
    return out
