"""Numpy restatement of pylorenzmie's frame pre-processing.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): the checker for the HIP
frame kernels, never called by the product path.

The reference's background normalization is a direct call of np.median or
scipy.ndimage.median_filter followed by a guarded division
(utilities/normalizer.py:64-76); ``normalize`` restates it, and
``median_filter_brute`` recomputes the filter from its definition (explicit
reflected windows and an exact rank) so that the library call is itself
pinned.  ``VMedian`` restates the running median of a video
(utilities/vmedian.py:6-121).
"""
from __future__ import annotations

import logging

import numpy as np
from scipy.ndimage import median_filter as _scipy_median_filter

__all__ = ['median', 'median_filter', 'median_filter_brute', 'normalize',
           'reflect_index', 'VMedian']

logger = logging.getLogger(__name__)


def median(image):
    """np.median of the whole image (utilities/normalizer.py:74)."""
    return np.median(image)


def median_filter(image, size=51):
    """scipy.ndimage.median_filter(image, size) (utilities/normalizer.py:72).

    scipy 1.15's rank filter reads outside its extended line buffer when
    the window exceeds the image by several image sizes (a 51-wide window
    on a 3 x 2 image); there ``median_filter_brute`` is the reference."""
    return _scipy_median_filter(image, size=size)


def reflect_index(i, n):
    """scipy.ndimage mode 'reflect' (d c b a | a b c d | d c b a)."""
    p = 2 * n
    m = np.mod(i, p)
    return np.where(m < n, m, p - 1 - m)


def median_filter_brute(image, size, points=None):
    """The rank-(size^2 // 2) element of each size x size window with
    offsets -(size // 2) .. size - size // 2 - 1 under 'reflect' boundaries,
    at every pixel or at the (row, col) ``points``.  Definition of the
    filter, independent of scipy's implementation."""
    ny, nx = image.shape
    lo = size // 2
    off = np.arange(size) - lo
    if points is None:
        rows, cols = np.divmod(np.arange(ny * nx), nx)
    else:
        rows, cols = (np.asarray(v) for v in points)
    out = np.empty(rows.shape, dtype=image.dtype)
    rank = size * size // 2
    for i, (r, c) in enumerate(zip(rows.ravel(), cols.ravel())):
        win = image[np.ix_(reflect_index(r + off, ny), reflect_index(c + off, nx))]
        out.flat[i] = np.partition(win.ravel(), rank)[rank]
    return out if points is not None else out.reshape(ny, nx)


def normalize(image, method='median', size=51, reference=None):
    """Normalizer.__call__ (utilities/normalizer.py:51-76): the image divided
    by its background, where(background > 0, background, 1)."""
    if method == 'reference' and reference is not None:
        background = np.asarray(reference)
        if background.ndim > 0 and background.shape != image.shape:
            logger.warning('reference shape %s != image shape %s; falling back '
                           'to median', background.shape, image.shape)
            background = np.median(image)
    elif method == 'filter':
        background = median_filter(image, size=size)
    else:
        background = np.median(image)
    safe = np.where(background > 0, background, 1.)
    return image / safe


class VMedian:
    """Running median of a video (utilities/vmedian.py:6-121): a tree of
    3-frame median buffers of depth ``order``; a level's output updates once
    per three updates of the level below."""

    def __init__(self, order=0, shape=None):
        self.child = None
        self._shape = None
        self.index = 0
        self.order = int(np.clip(order, 0, 10))
        if self.order > 0:
            self.child = VMedian(self.order - 1, shape)
        self.shape = shape
        self.initialized = False
        self.cycled = False

    @property
    def shape(self):
        return self._shape

    @shape.setter
    def shape(self, shape):
        """vmedian.py:100-110: reallocate and restart."""
        self._shape = shape
        if shape is None:
            return
        if self.child is not None:
            self.child.shape = shape
        self.buffer = np.zeros((3, int(np.prod(shape))), dtype=np.uint8)
        self.index = 0
        self.initialized = False

    def get(self, reshape=True):
        return self._data.reshape(self.shape) if reshape else self._data

    def add(self, data):
        """vmedian.py:50-77."""
        if data.shape != self.shape:
            self._data = data.astype(np.uint8).ravel()
            self.shape = data.shape
        if self.order == 0:
            self.buffer[self.index, :] = data.astype(np.uint8).ravel()
            self.index += 1
        else:
            self.child.add(data)
            if self.child.initialized:
                self.buffer[self.index, :] = self.child.get(reshape=False)
            if self.child.cycled:
                self.index += 1
        if self.index == 3:
            self.index = 0
            self._data = np.median(self.buffer, axis=0).astype(np.uint8)
            self.initialized = True
            self.cycled = True
        else:
            self.cycled = False

    def filter(self, data):
        self.add(data)
        return self.get()

    def reset(self):
        """vmedian.py:79-84."""
        self.initialized = False
        self.cycled = False
        if self.order > 0:
            self.child.reset()
