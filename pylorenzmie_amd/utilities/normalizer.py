"""Background normalization of hologram frames on the GPU, mirroring
utilities/normalizer.py:11-76.

``Normalizer(method, size, reference)(image)`` divides the image by its
background, ``where(background > 0, background, 1.)``:

* ``'median'`` -- the scalar median of the image (lmz_median, an exact radix
  select in device memory, then lmz_normalize);
* ``'filter'`` -- scipy.ndimage.median_filter(image, size) with its 'reflect'
  boundary (lmz_median_filter: a sliding 256-bin histogram for 8-bit frames,
  a two-level one for 16-bit frames, a per-pixel radix select for float
  frames), fused with the division;
* ``'reference'`` -- a background image or scalar (lmz_normalize); an image
  of another shape falls back to the median with a warning, as the
  reference does.

A numpy image returns a numpy array of numpy's result type with the values
numpy computes (the quotient of the exact operands rounded once).  A torch
tensor on the GPU stays there: the result is a device tensor, float32 by
default when ``dtype`` is not given, ready for ``Frame`` and ``LMBatch``.
"""
from __future__ import annotations

import logging

import numpy as np

from .. import _lmz

logger = logging.getLogger(__name__)

_SUPPORTED = (np.uint8, np.uint16, np.float32, np.float64)


def result_dtype(dtype, method: str = 'median', reference=None) -> np.dtype:
    """numpy's result type of image / where(bg > 0, bg, 1.) for an image of
    ``dtype`` (utilities/normalizer.py:64-76), evaluated on one pixel."""
    img = np.ones(1, dtype)
    if method == 'reference' and reference is not None:
        bg = np.asarray(reference)
        if bg.ndim > 0:
            bg = np.ones(1, bg.dtype)
    elif method == 'filter':
        bg = np.ones(1, dtype)          # median_filter keeps the image's type
    else:
        bg = np.median(img)
    safe = np.where(bg > 0, bg, 1.)
    return (img / safe).dtype


def _torch_dtype(t, dt: np.dtype):
    return {np.dtype(np.float32): t.float32, np.dtype(np.float64): t.float64,
            np.dtype(np.uint8): t.uint8, np.dtype(np.uint16): t.uint16}[np.dtype(dt)]


def _np_dtype_of(tensor) -> np.dtype:
    t = _lmz.torch()
    m = {t.uint8: np.uint8, t.uint16: np.uint16, t.float32: np.float32,
         t.float64: np.float64}
    if tensor.dtype not in m:
        raise TypeError(f'unsupported frame dtype {tensor.dtype}: use uint8, '
                        'uint16, float32 or float64')
    return np.dtype(m[tensor.dtype])


def is_device_tensor(x) -> bool:
    try:
        import torch
    except ImportError:                                   # pragma: no cover
        return False
    return isinstance(x, torch.Tensor) and x.is_cuda


class Normalizer:
    """Normalize a hologram image by its estimated background
    (utilities/normalizer.py:11-76)."""

    methods = ('filter', 'median', 'reference')

    def __init__(self, method: str = 'median', size: int = 51,
                 reference=None) -> None:
        self.method = method
        self.size = size
        self.reference = reference

    def __call__(self, image, dtype=None):
        """Image divided by its estimated background.  ``dtype`` (an
        extension) chooses float32 or float64 output instead of numpy's
        result type."""
        t = _lmz.torch()
        on_device = is_device_tensor(image)
        if on_device:
            img = image.contiguous()
            idt = _np_dtype_of(img)
        else:
            arr = np.asarray(image)
            if arr.dtype not in [np.dtype(d) for d in _SUPPORTED]:
                if arr.dtype.kind in 'iub' or arr.dtype == np.float16:
                    arr = arr.astype(np.float64 if arr.dtype.kind != 'f'
                                     else np.float32)
                else:
                    raise TypeError(f'unsupported image dtype {arr.dtype}')
            idt = arr.dtype
            img = t.from_numpy(np.ascontiguousarray(arr)).to(_lmz.device())
        if img.dim() != 2:
            raise ValueError(f'image must be 2-D, got shape {tuple(img.shape)}')
        ny, nx = img.shape
        n = ny * nx
        method, reference = self._resolve(img.shape)
        cast = None
        if dtype is not None:
            odt = np.dtype(dtype)
            if odt not in (np.dtype(np.float32), np.dtype(np.float64)):
                raise TypeError(f'output dtype must be float32 or float64, not {odt}')
        elif on_device:
            odt = np.dtype(np.float32)
        else:
            odt = result_dtype(image.dtype if hasattr(image, 'dtype') else idt,
                               method, reference)
            if odt not in (np.dtype(np.float32), np.dtype(np.float64)):
                # float16 frames: computed in float64, rounded to numpy's type
                cast, odt = odt, np.dtype(np.float64)
        out = t.empty((ny, nx), dtype=_torch_dtype(t, odt), device=img.device)
        lib = _lmz.load_library()
        code, ocode = _lmz.NP_DTYPE_CODE[np.dtype(idt)], _lmz.NP_DTYPE_CODE[odt]
        stream = _lmz.stream_handle()
        if method == 'filter':
            size = int(self.size)
            if not 1 <= size <= _lmz.MAX_FILTER_SIZE:
                raise ValueError(f'size must be in [1, {_lmz.MAX_FILTER_SIZE}]')
            _lmz.check(lib.lmz_median_filter(_lmz.ptr(img), code, ny, nx, size,
                                             None, _lmz.ptr(out), ocode, stream))
        elif method == 'reference':
            bg = reference
            if is_device_tensor(bg) or np.ndim(bg) > 0:
                if is_device_tensor(bg):
                    ref = bg.contiguous()
                    rdt = _np_dtype_of(ref)
                else:
                    rarr = np.asarray(bg)
                    if rarr.dtype not in [np.dtype(d) for d in _SUPPORTED]:
                        rarr = rarr.astype(np.float64)
                    rdt = rarr.dtype
                    ref = t.from_numpy(np.ascontiguousarray(rarr)).to(img.device)
                _lmz.check(lib.lmz_normalize(
                    _lmz.ptr(img), code, n, None, _lmz.ptr(ref),
                    _lmz.NP_DTYPE_CODE[np.dtype(rdt)], 0., _lmz.ptr(out), ocode,
                    stream))
            else:
                _lmz.check(lib.lmz_normalize(
                    _lmz.ptr(img), code, n, None, None, 0, float(np.asarray(bg)),
                    _lmz.ptr(out), ocode, stream))
        else:
            med = self.median(img)
            _lmz.check(lib.lmz_normalize(_lmz.ptr(img), code, n, _lmz.ptr(med),
                                         None, 0, 0., _lmz.ptr(out), ocode,
                                         stream))
        if on_device:
            return out
        host = out.cpu().numpy()
        return host.astype(cast) if cast is not None else host

    def _resolve(self, shape):
        """Method and background after the reference's shape check
        (utilities/normalizer.py:64-70)."""
        if self.method == 'reference' and self.reference is not None:
            bg = self.reference
            bshape = tuple(bg.shape) if hasattr(bg, 'shape') else ()
            if len(bshape) > 0 and bshape != tuple(shape):
                logger.warning(f'Reference shape {bshape} != image shape '
                               f'{tuple(shape)}; falling back to median')
                return 'median', None
            return 'reference', bg
        if self.method == 'filter':
            return 'filter', None
        return 'median', None

    @staticmethod
    def median(image):
        """np.median of a device image as a one-element float64 device
        tensor (lmz_median)."""
        t = _lmz.torch()
        img = image.contiguous()
        lib = _lmz.load_library()
        scratch = t.empty(int(lib.lmz_median_scratch_bytes()), dtype=t.uint8,
                          device=img.device)
        med = t.empty(1, dtype=t.float64, device=img.device)
        _lmz.check(lib.lmz_median(_lmz.ptr(img),
                                  _lmz.NP_DTYPE_CODE[_np_dtype_of(img)],
                                  img.numel(), _lmz.ptr(med), _lmz.ptr(scratch),
                                  _lmz.stream_handle()))
        return med
