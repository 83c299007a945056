"""The C-ABI library loads, exports every symbol include/lmz.h declares, and
rejects bad arguments on the host before any launch (no GPU needed)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, 'include', 'lmz.h')


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(lmz_\w+)\s*\(', text)))


@pytest.fixture(scope='module')
def lib():
    from pylorenzmie_amd import _lmz
    return _lmz.load_library()


def test_exports_every_declared_symbol(lib):
    from pylorenzmie_amd import _lmz
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lmz.SIGNATURES, name


def test_version(lib):
    assert lib.lmz_version() == 104


def test_struct_layouts_match_header():
    from pylorenzmie_amd import _lmz
    assert ctypes.sizeof(_lmz.Feature) == 88
    assert ctypes.sizeof(_lmz.BatchSizes) == 48
    assert ctypes.sizeof(_lmz.LMSettings) == 64
    assert ctypes.sizeof(_lmz.Schedule) == 16
    assert ctypes.sizeof(_lmz.DESettings) == 96


def _feature(**kw):
    from pylorenzmie_amd import _lmz
    f = _lmz.Feature()
    f.npix, f.width, f.nmax_cap = 100, 10, 30
    f.n_m, f.wavelength, f.magnification, f.noise = 1.34, 0.447, 0.048, 0.05
    for k, v in kw.items():
        setattr(f, k, v)
    return f


@pytest.mark.parametrize('bad', [dict(npix=0), dict(width=-1),
                                 dict(width=100000), dict(nmax_cap=0),
                                 dict(noise=0.), dict(wavelength=-1.)])
def test_batch_create_rejects_bad_features(lib, bad):
    from pylorenzmie_amd import _lmz
    arr = (_lmz.Feature * 1)(_feature(**bad))
    h = ctypes.c_void_p()
    rc = lib.lmz_batch_create(1, ctypes.cast(arr, ctypes.c_void_p), 1, ctypes.byref(h))
    assert rc == 1
    assert b'invalid feature' in lib.lmz_last_error()
    assert not h.value


def test_batch_create_rejects_overlapping_pixels(lib):
    from pylorenzmie_amd import _lmz
    arr = (_lmz.Feature * 2)(_feature(), _feature(pix_offset=50))
    h = ctypes.c_void_p()
    assert lib.lmz_batch_create(2, ctypes.cast(arr, ctypes.c_void_p), 1,
                                ctypes.byref(h)) == 1


def test_calls_reject_null_batch(lib):
    assert lib.lmz_mie_ab(None, None, None, None, None, None, None, None, None) == 1
    assert b'batch is NULL' in lib.lmz_last_error()
    assert lib.lmz_hologram(None, *([None] * 5), 3, *([None] * 5)) == 1
    assert lib.lmz_lm_run(None, None, 3, None, None, None, None, None, None) == 1
    assert lib.lmz_batch_set_schedule(None, None) == 1
    assert lib.lmz_autotune(None, *([None] * 5), 3, None, None, 0, None, None) == 1
    assert lib.lmz_mie_ab_layered(None, *([None] * 7)) == 1
    assert lib.lmz_hologram_vjp(None, *([None] * 5), 3, *([None] * 4)) == 1
    assert lib.lmz_de_run(None, None, None, None, None, 3, *([None] * 6)) == 1
    assert lib.lmz_de_workspace_bytes(None, None, None) == 1


def test_frame_calls_reject_bad_arguments(lib):
    """Frame entry points check their arguments on the host, before any
    launch."""
    P = ctypes.c_void_p
    fake = P(256)                   # never dereferenced: rejected first
    assert lib.lmz_median_scratch_bytes() >= 256 * 4 + 6 * 8
    assert lib.lmz_median(fake, 7, 10, fake, fake, None) == 1
    assert lib.lmz_median(fake, 2, -1, fake, fake, None) == 1
    assert lib.lmz_median(fake, 2, 10, None, fake, None) == 1
    assert lib.lmz_median(fake, 2, 10, fake, P(257), None) == 1
    for ny, nx, size, odt in ((0, 5, 3, 2), (5, 5, 0, 2), (5, 5, 256, 2),
                              (5, 5, 3, 0)):
        assert lib.lmz_median_filter(fake, 0, ny, nx, size, None, fake, odt,
                                     None) == 1
    assert lib.lmz_median_filter(fake, 0, 5, 5, 3, None, None, 2, None) == 1
    assert lib.lmz_normalize(fake, 0, 10, fake, fake, 0, 0., fake, 2, None) == 1
    assert lib.lmz_normalize(fake, 0, 10, None, None, 0, 1., fake, 1, None) == 1
    assert lib.lmz_normalize(fake, 0, 10, None, fake, 9, 1., fake, 2, None) == 1
    assert lib.lmz_normalize(None, 0, 0, None, None, 0, 1., None, 2, None) == 0
    assert lib.lmz_median3_u8(None, 5, None, None) == 1
    assert lib.lmz_median3_u8(None, 0, None, None) == 0
    assert lib.lmz_gather_frame(None, fake, 5, 5, None, None, fake, None, fake,
                                None, None) == 1
    assert b'batch' in lib.lmz_last_error()


def test_last_error_is_per_thread(lib):
    """lmz_last_error holds the calling thread's last failure: a failure on
    another thread does not replace it (the ABI promises thread-safe calls
    on distinct streams)."""
    import threading
    assert lib.lmz_mie_ab(None, None, None, None, None, None, None, None, None) == 1
    mine = lib.lmz_last_error()
    assert b'batch is NULL' in mine
    seen = []

    def other():
        seen.append(lib.lmz_last_error())
        assert lib.lmz_median3_u8(None, 5, None, None) == 1
        seen.append(lib.lmz_last_error())
    t = threading.Thread(target=other)
    t.start()
    t.join()
    assert seen[0] in (b'', None)
    assert seen[1] != mine
    assert lib.lmz_last_error() == mine


def test_ptr_refuses_strided_tensors():
    """Every buffer crosses the C ABI as packed row-major memory: a
    transposed view (or a tensor made from a Fortran-ordered array) is
    refused instead of being read in the wrong order."""
    import numpy as np
    import torch
    from pylorenzmie_amd import _lmz
    a = torch.zeros((3, 5), dtype=torch.float32)
    assert _lmz.ptr(a) == a.data_ptr()
    assert _lmz.ptr(None) is None
    with pytest.raises(_lmz.LMZError):
        _lmz.ptr(a.t())
    f = torch.from_numpy(np.asfortranarray(np.zeros((3, 5), np.float32)))
    with pytest.raises(_lmz.LMZError):
        _lmz.ptr(f)
