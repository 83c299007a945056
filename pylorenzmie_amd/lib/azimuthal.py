"""Azimuthal statistics of an image about a centre, mirroring
lib/Azimuthal.py:8-160 (avg, std, med, mad with integer radius bins).

Pixel (row i, column j) falls in bin int(hypot(i - y_c, j - x_c)); the bin
statistics are those of the reference (bincount means, their standard
deviation, medians with NaN for empty bins, median absolute deviations).
Host numpy: a profile is computed once per crop, outside any fit loop.
"""
from __future__ import annotations

from functools import lru_cache

import numpy as np

__all__ = ['avg', 'std', 'med', 'mad']


@lru_cache(maxsize=16)
def _bins(ny: int, nx: int, xc: float, yc: float):
    """Integer radius of every pixel, the radius-ordered pixel order and the
    bin edges in it (lib/Azimuthal.py:23-32)."""
    r = np.hypot.outer(np.arange(ny) - yc, np.arange(nx) - xc).astype(np.int32).ravel()
    order = np.argsort(r, kind='stable')
    edges = np.searchsorted(r[order], np.arange(int(r.max()) + 2))
    return r, order, edges


def _prepare(data, center):
    data = np.asarray(data)
    ny, nx = data.shape
    xc, yc = (nx / 2., ny / 2.) if center is None else center
    return data.ravel(), _bins(ny, nx, float(xc), float(yc))


def avg(data, center=None) -> np.ndarray:
    """Mean of the data at each integer distance from ``center`` ((x, y),
    default the image centre)."""
    d, (r, _, _) = _prepare(data, center)
    return np.bincount(r, d) / np.bincount(r)


def std(data, center=None):
    """(mean, standard deviation) per radius."""
    d, (r, _, _) = _prepare(data, center)
    n = np.bincount(r)
    a = np.bincount(r, d) / n
    return a, np.sqrt(np.bincount(r, (d - a[r]) ** 2) / n)


def med(data, center=None) -> np.ndarray:
    """Median per radius; NaN where no pixel falls."""
    d, (_, order, edges) = _prepare(data, center)
    ds = d[order]
    out = np.full(len(edges) - 1, np.nan)
    for k in range(len(out)):
        if edges[k] < edges[k + 1]:
            out[k] = np.median(ds[edges[k]:edges[k + 1]])
    return out


def mad(data, center=None):
    """(median, median absolute deviation) per radius."""
    d, (_, order, edges) = _prepare(data, center)
    ds = d[order]
    m = np.full(len(edges) - 1, np.nan)
    s = np.full(len(edges) - 1, np.nan)
    for k in range(len(m)):
        seg = ds[edges[k]:edges[k + 1]]
        if seg.size:
            m[k] = np.median(seg)
            s[k] = np.median(np.abs(seg - m[k]))
    return m, s
