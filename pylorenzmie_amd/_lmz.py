"""ctypes binding of liblmz.so (include/lmz.h) and device-memory plumbing.

PyTorch provides device allocations and the current stream; every number
on the path is computed by the HIP kernels in liblmz.so.  There is no CPU
fallback: importing this module raises if the library is missing.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, 'liblmz.so')

NPARAM = 8
MAX_LAYERS = 8
NJAC = 7
NRED = 36
NRES = 40
TILE = 1024
VJP_STATIC_MAX_TILES = 65536
TUNE_FORCE = 1
CARTESIAN = 1
BOHREN = 2
ACCUMULATE = 4
FP32 = 8
MAX_FILTER_SIZE = 255

# pixel types of frames (LMZ_U8 ... LMZ_F64)
U8, U16, F32, F64 = 0, 1, 2, 3
NP_DTYPE_CODE = {np.dtype(np.uint8): U8, np.dtype(np.uint16): U16,
                 np.dtype(np.float32): F32, np.dtype(np.float64): F64}

LM_RUNNING, LM_FTOL, LM_XTOL, LM_GTOL = 0, 1, 2, 4
LM_MAXITER, LM_STALLED, LM_NONFINITE = 8, 16, 32
DAMPING_MINPACK, DAMPING_NIELSEN = 0, 1


class LMZError(RuntimeError):
    """Failure reported by liblmz (lmz_last_error)."""


class Feature(ctypes.Structure):
    """lmz_feature, include/lmz.h."""
    _fields_ = [('pix_offset', ctypes.c_int64),
                ('npix', ctypes.c_int32),
                ('width', ctypes.c_int32),
                ('origin', ctypes.c_double * 3),
                ('n_m', ctypes.c_double),
                ('wavelength', ctypes.c_double),
                ('magnification', ctypes.c_double),
                ('noise', ctypes.c_double),
                ('nmax_cap', ctypes.c_int32),
                ('coef_offset', ctypes.c_int32),
                ('tile_offset', ctypes.c_int32),
                ('ntiles', ctypes.c_int32)]


class BatchSizes(ctypes.Structure):
    """lmz_batch_sizes, include/lmz.h."""
    _fields_ = [('nfeat', ctypes.c_int32),
                ('ntiles', ctypes.c_int32),
                ('npix_total', ctypes.c_int64),
                ('coef_buffers', ctypes.c_int32),
                ('coef_slots', ctypes.c_int32),
                ('cap_max', ctypes.c_int32),
                ('work_doubles', ctypes.c_int64),
                ('coef_stride', ctypes.c_int32),
                ('dcoef_stride', ctypes.c_int32)]


class Schedule(ctypes.Structure):
    """lmz_schedule, include/lmz.h."""
    _fields_ = [('jac_threads', ctypes.c_int32),
                ('fwd_threads', ctypes.c_int32),
                ('jac_blocks_per_cu', ctypes.c_int32),
                ('fwd_blocks_per_cu', ctypes.c_int32)]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class LMSettings(ctypes.Structure):
    """lmz_lm_settings, include/lmz.h."""
    _fields_ = [('max_iter', ctypes.c_int32),
                ('check_every', ctypes.c_int32),
                ('ftol', ctypes.c_double),
                ('xtol', ctypes.c_double),
                ('gtol', ctypes.c_double),
                ('lambda0', ctypes.c_double),
                ('vary', ctypes.c_uint8 * NJAC),
                ('use_graph', ctypes.c_uint8),
                ('fd_step', ctypes.c_double),
                ('damping', ctypes.c_int32),
                ('reserved', ctypes.c_int32)]


class DESettings(ctypes.Structure):
    """lmz_de_settings, include/lmz.h."""
    _fields_ = [('ndim', ctypes.c_int32),
                ('cols', ctypes.c_int32 * NJAC),
                ('npop', ctypes.c_int32),
                ('maxiter', ctypes.c_int32),
                ('check_every', ctypes.c_int32),
                ('reserved', ctypes.c_int32),
                ('tol', ctypes.c_double),
                ('atol', ctypes.c_double),
                ('mutation_lo', ctypes.c_double),
                ('mutation_hi', ctypes.c_double),
                ('recombination', ctypes.c_double),
                ('seed', ctypes.c_uint64)]


_P = ctypes.c_void_p
_I = ctypes.c_int32

# name -> (restype, argtypes): every symbol include/lmz.h declares
SIGNATURES = {
    'lmz_batch_create': (_I, [_I, _P, _I, ctypes.POINTER(_P)]),
    'lmz_batch_destroy': (_I, [_P]),
    'lmz_batch_query': (_I, [_P, ctypes.POINTER(BatchSizes)]),
    'lmz_mie_ab': (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
    'lmz_mie_ab_layered': (_I, [_P, _P, _P, _P, _P, _P, _P, _P]),
    'lmz_mie_ab_layered_fit': (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
    'lmz_coef_from_ab': (_I, [_P, _P, _P, _P, _P, _P]),
    'lmz_field': (_I, [_P, _P, _P, _P, _P, _P, _I, _P, _P]),
    'lmz_hologram': (_I, [_P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P]),
    'lmz_intensity': (_I, [_P, _P, _P, _I, _P, _P]),
    'lmz_hologram_superposed': (_I, [_P, _I, _P, _P, _P, _P, _I, _P, _I,
                                     ctypes.c_double, _P, _P, _P, _P]),
    'lmz_jacobian': (_I, [_P, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P,
                          _P, _P]),
    'lmz_hologram_vjp': (_I, [_P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P]),
    'lmz_hologram_vjp_static': (_I, [_P, _P, _P, _P, _P, _P, _I, _P, _P, _P]),
    'lmz_batch_set_schedule': (_I, [_P, ctypes.POINTER(Schedule)]),
    'lmz_batch_get_schedule': (_I, [_P, ctypes.POINTER(Schedule)]),
    'lmz_autotune_scratch_bytes': (_I, [_P, ctypes.POINTER(ctypes.c_size_t)]),
    'lmz_autotune': (_I, [_P, _P, _P, _P, _P, _P, _I, _P, _P, _I,
                          ctypes.POINTER(Schedule), _P]),
    'lmz_normal_eq': (_I, [_P, _P, _P, _P, _P]),
    'lmz_lm_workspace_bytes': (_I, [_P, ctypes.POINTER(ctypes.c_size_t)]),
    'lmz_jacobian_tabulated_bytes': (_I, [_P, ctypes.POINTER(ctypes.c_size_t)]),
    'lmz_jacobian_tabulated': (_I, [_P, _P, _P, _P, _P, _I, _P, _P, _P, _P]),
    'lmz_residual_tabulated_bytes': (_I, [_P, ctypes.POINTER(ctypes.c_size_t)]),
    'lmz_residual_tabulated': (_I, [_P, _P, _P, _P, _I, _P, _P, _P, _P, _P]),
    'lmz_lm_plan_create': (_I, [_P, ctypes.POINTER(_P)]),
    'lmz_lm_plan_destroy': (_I, [_P]),
    'lmz_lm_run': (_I, [_P, _P, _I, _P, ctypes.POINTER(LMSettings), _P, _P,
                        _P, _P]),
    'lmz_lm_run_fixed_ab': (_I, [_P, _P, _I, _P, ctypes.POINTER(LMSettings),
                                 _P, _P, _P, _P, _P, _P]),
    'lmz_lm_run_layered': (_I, [_P, _P, _I, _P, ctypes.POINTER(LMSettings),
                                _P, _P, _P, _P, _P, _P]),
    'lmz_de_workspace_bytes': (_I, [_P, ctypes.POINTER(DESettings),
                                    ctypes.POINTER(ctypes.c_size_t)]),
    'lmz_de_run': (_I, [_P, ctypes.POINTER(DESettings), _P, _P, _P, _I, _P,
                        _P, _P, _P, _P, _P]),
    'lmz_median_scratch_bytes': (ctypes.c_size_t, []),
    'lmz_median': (_I, [_P, _I, ctypes.c_int64, _P, _P, _P]),
    'lmz_median_filter': (_I, [_P, _I, _I, _I, _I, _P, _P, _I, _P]),
    'lmz_normalize': (_I, [_P, _I, ctypes.c_int64, _P, _P, _I, ctypes.c_double,
                           _P, _I, _P]),
    'lmz_median3_u8': (_I, [_P, ctypes.c_int64, _P, _P]),
    'lmz_gather_frame': (_I, [_P, _P, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
    'lmz_last_error': (ctypes.c_char_p, []),
    'lmz_version': (_I, []),
}

_lib = None
_lock = threading.Lock()


def load_library(path: str = LIBPATH) -> ctypes.CDLL:
    """Load liblmz.so and declare its C ABI.  Raises if it is missing."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise ImportError(
                f'{path} is missing: build it with __graft_entry__.build() '
                '(hipcc --offload-arch=gfx950); there is no CPU fallback')
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load_library().lmz_last_error().decode(errors='replace')
        raise LMZError(f'liblmz error {rc}: {msg}')


# ------------------------------------------------------------------ torch
_torch = None


def torch():
    """Import torch lazily and require a GPU."""
    global _torch
    if _torch is None:
        import torch as t
        _torch = t
    return _torch


def device():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError('pylorenzmie_amd needs a ROCm GPU (gfx950); '
                           'torch.cuda.is_available() is False')
    return t.device('cuda', t.cuda.current_device())


def stream_handle(stream=None) -> int:
    t = torch()
    s = stream if stream is not None else t.cuda.current_stream()
    return s.cuda_stream


def ptr(tensor) -> int | None:
    """Device pointer of a tensor, or None.  The C ABI reads every buffer
    as densely packed row-major memory, so a strided tensor (a transposed
    view, or one copied from a Fortran-ordered array) is refused."""
    if tensor is None:
        return None
    if not tensor.is_contiguous():
        raise LMZError(f'liblmz needs contiguous buffers; got a tensor of '
                       f'shape {tuple(tensor.shape)} with strides '
                       f'{tuple(tensor.stride())}')
    return tensor.data_ptr()


class Batch:
    """Owner of an lmz_batch: the validated feature table on the device."""

    def __init__(self, features: list[Feature] | ctypes.Array,
                 coef_buffers: int = 1):
        lib = load_library()
        n = len(features)
        arr = (Feature * n)(*features) if not isinstance(
            features, ctypes.Array) else features
        handle = _P()
        check(lib.lmz_batch_create(n, ctypes.cast(arr, _P), coef_buffers,
                                   ctypes.byref(handle)))
        self._lib = lib
        self.handle = handle
        self.features = arr
        sizes = BatchSizes()
        check(lib.lmz_batch_query(handle, ctypes.byref(sizes)))
        self.nfeat = sizes.nfeat
        self.ntiles = sizes.ntiles
        self.npix_total = sizes.npix_total
        self.coef_buffers = sizes.coef_buffers
        self.coef_slots = sizes.coef_slots
        self.cap_max = sizes.cap_max
        self.work_doubles = sizes.work_doubles
        # floats per row of the coefficient tables, as the library lays them
        self.coef_stride = sizes.coef_stride
        self.dcoef_stride = sizes.dcoef_stride
        if self.coef_stride <= 0 or self.dcoef_stride <= 0:
            raise LMZError('liblmz reports no coefficient row stride: the '
                           'library predates include/lmz.h; rebuild it')

    def coef_floats(self) -> int:
        return self.coef_slots * self.coef_stride

    def dcoef_floats(self) -> int:
        return self.coef_slots * self.dcoef_stride

    def lm_workspace_bytes(self) -> int:
        n = ctypes.c_size_t()
        check(self._lib.lmz_lm_workspace_bytes(self.handle, ctypes.byref(n)))
        return n.value

    @property
    def schedule(self) -> Schedule:
        sc = Schedule()
        check(self._lib.lmz_batch_get_schedule(self.handle, ctypes.byref(sc)))
        return sc

    @schedule.setter
    def schedule(self, sc) -> None:
        if isinstance(sc, dict):
            sc = Schedule(**sc)
        check(self._lib.lmz_batch_set_schedule(self.handle, ctypes.byref(sc)))

    def autotune(self, params, nmax, coef, dcoef, data, coords=None,
                 flags: int = CARTESIAN | BOHREN, force: bool = False,
                 stream=None) -> Schedule:
        """Pick the pixel-kernel launch shapes by timing them on this
        batch (lmz_autotune); device tensors as for lmz_jacobian."""
        t = torch()
        n = ctypes.c_size_t()
        check(self._lib.lmz_autotune_scratch_bytes(self.handle, ctypes.byref(n)))
        buf = t.empty(n.value + 256, dtype=t.uint8, device=params.device)
        base = (buf.data_ptr() + 255) & ~255
        sc = Schedule()
        check(self._lib.lmz_autotune(self.handle, ptr(params), ptr(nmax),
                                     ptr(coef), ptr(dcoef), ptr(coords), flags,
                                     ptr(data), base, TUNE_FORCE if force else 0,
                                     ctypes.byref(sc), stream_handle(stream)))
        return sc

    def close(self):
        if getattr(self, 'handle', None):
            self._lib.lmz_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_feature(npix, width, origin, instrument, nmax_cap, pix_offset=0,
                 noise=None) -> Feature:
    f = Feature()
    f.pix_offset = int(pix_offset)
    f.npix = int(npix)
    f.width = int(width)
    f.origin[0], f.origin[1], f.origin[2] = (float(v) for v in origin)
    f.n_m = float(np.real(instrument.n_m))
    f.wavelength = float(instrument.wavelength)
    f.magnification = float(instrument.magnification)
    f.noise = float(instrument.noise if noise is None else noise)
    f.nmax_cap = int(nmax_cap)
    return f


def wiscombe_yang(x: float, m: complex) -> int:
    """n_max of theory/Sphere.py:91-123, used to size coefficient tables."""
    xl = abs(x)
    if xl <= 8.:
        ns = np.floor(xl + 4. * np.cbrt(xl) + 1.)
    elif xl <= 4200.:
        ns = np.floor(xl + 4.05 * np.cbrt(xl) + 2.)
    else:
        ns = np.floor(xl + 4. * np.cbrt(xl) + 2.)
    return int(max(ns, abs(x * m)))


def nmax_cap(a_p: float, n_p: float, k_p: float, n_m: float,
             wavelength: float, headroom: float = 0.25) -> int:
    """Orders to reserve for a sphere whose a_p, n_p may grow during a fit."""
    k = 2. * np.pi / wavelength * np.real(n_m)
    x = k * a_p * (1. + headroom)
    m = (n_p * (1. + 0.25 * headroom) + 1j * k_p) / np.real(n_m)
    return wiscombe_yang(x, m) + 4
