"""CPU checks of frame normalization: the oracle against the definition of
the median filter and numpy's promotion rules, and the device algorithms
(pylorenzmie_amd/csrc/lmz_frame.h, compiled for the host into
tests/native/liblmz_hostcheck.so) against scipy.ndimage.median_filter and
np.median, bit for bit.  The GPU tests repeat these through liblmz.so."""
import ctypes
import os

import numpy as np
import pytest
from scipy.ndimage import median_filter

from conftest import ROOT
from oracle import normalize as ref

LIB = os.path.join(ROOT, 'tests', 'native', 'liblmz_hostcheck.so')
DTYPES = {0: np.uint8, 1: np.uint16, 2: np.float32, 3: np.float64}


@pytest.fixture(scope='module')
def chk():
    if not os.path.exists(LIB):
        pytest.skip('host checker not built (run __graft_entry__.build())')
    lib = ctypes.CDLL(LIB)
    P = ctypes.c_void_p
    lib.chk_median_filter.argtypes = [P, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, P, ctypes.c_int]
    lib.chk_key.restype = ctypes.c_ulonglong
    lib.chk_key.argtypes = [ctypes.c_int, ctypes.c_double]
    lib.chk_mean_of_two.restype = ctypes.c_double
    lib.chk_mean_of_two.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
    lib.chk_reflect.argtypes = [ctypes.c_int, ctypes.c_int]
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _image(rng, dt, shape, smooth=False):
    if dt == 0:
        return rng.integers(0, 256, shape, dtype=np.uint8)
    if dt == 1:
        return rng.integers(0, 4096, shape, dtype=np.uint16)
    if smooth:      # a background near 1 with fringes: keys share high bits
        y, x = np.mgrid[:shape[0], :shape[1]]
        v = 1. + 0.3 * np.cos(0.4 * np.hypot(x - 5.3, y - 3.1)) + \
            0.01 * rng.standard_normal(shape)
    else:
        v = rng.standard_normal(shape) * 50.
    return v.astype(DTYPES[dt])


def _host_filter(chk, img, dt, size, path):
    out = np.empty_like(img)
    assert chk.chk_median_filter(_p(img), dt, img.shape[0], img.shape[1], size,
                                 _p(out), path) == 0
    return out


# ---------------------------------------------------------------- the oracle
@pytest.mark.parametrize('dt', [0, 1, 2, 3])
@pytest.mark.parametrize('shape,size', [((7, 9), 3), ((7, 9), 4), ((6, 5), 1),
                                        ((5, 6), 2), ((9, 8), 5), ((5, 6), 15)])
def test_filter_definition(dt, shape, size):
    """scipy's filter is the rank-(size^2 // 2) element of the reflected
    window with offsets -(size//2) .. size - size//2 - 1, also for even sizes
    and windows wider than the image."""
    img = _image(np.random.default_rng(dt * 100 + size), dt, shape)
    np.testing.assert_array_equal(ref.median_filter(img, size),
                                  ref.median_filter_brute(img, size))


def test_filter_definition_rejects_a_wrong_window():
    """A window off by one (odd/even centring) or an edge mode other than
    'reflect' fails the definition check."""
    img = _image(np.random.default_rng(5), 2, (9, 10))
    good = ref.median_filter_brute(img, 4)
    assert not np.array_equal(median_filter(img, size=4, origin=-1), good)
    assert not np.array_equal(median_filter(img, size=4, mode='mirror'), good)


def test_golden_backgrounds_from_the_definition(frame_golden):
    g = frame_golden
    rng = np.random.default_rng(0)
    pts = (rng.integers(0, 256, 150), rng.integers(0, 320, 150))
    for size in (51, 8):
        np.testing.assert_array_equal(
            ref.median_filter_brute(g['image'], size, pts),
            g[f'bg{size}'][pts])
    assert g['median'] == np.median(g['image'])


def test_normalize_result_types():
    """The result types of image / where(bg > 0, bg, 1.) that the GPU path
    reproduces (utilities/normalizer.py:75-76)."""
    from pylorenzmie_amd.utilities.normalizer import result_dtype
    for dt in DTYPES.values():
        img = np.ones((3, 4), dt)
        for method in ('median', 'filter'):
            want = ref.normalize(img, method, size=3).dtype
            assert result_dtype(img.dtype, method) == want, (dt, method)
        for rdt in DTYPES.values():
            r = np.full((3, 4), 2, rdt)
            want = ref.normalize(img, 'reference', reference=r).dtype
            assert result_dtype(img.dtype, 'reference', r) == want, (dt, rdt)
        for r in (2.5, np.float32(2.5), np.asarray(2.5, np.float32), 3):
            want = ref.normalize(img, 'reference', reference=r).dtype
            assert result_dtype(img.dtype, 'reference', r) == want, (dt, r)


def test_vmedian_oracle():
    rng = np.random.default_rng(1)
    frames = rng.integers(0, 256, (9, 4, 5), dtype=np.uint8)
    v0 = ref.VMedian(order=0)
    for i in range(3):
        v0.add(frames[i])
    assert v0.initialized
    np.testing.assert_array_equal(v0.get(), np.median(frames[:3], axis=0))
    v1 = ref.VMedian(order=1, shape=(4, 5))
    for i in range(9):
        v1.add(frames[i])
    meds = np.median(frames.reshape(3, 3, 4, 5), axis=1)
    np.testing.assert_array_equal(v1.get(), np.median(meds, axis=0))


# ---------------------------------------------------- the device algorithms
def test_keys_preserve_order(chk):
    vals = np.array([-np.inf, -1e30, -2.5, -1e-40, -0.0, 0.0, 1e-45, 1e-40,
                     1.0, 1.5, 3e38, np.inf])
    for dt in (2, 3):
        v = vals.astype(DTYPES[dt])
        keys = [chk.chk_key(dt, float(x)) for x in v]
        assert keys == sorted(keys) and len(set(keys)) == len(keys)
        assert chk.chk_key(dt, float('nan')) == max(
            chk.chk_key(dt, float('nan')), keys[-1] + 1)
    for dt in (0, 1):
        assert [chk.chk_key(dt, x) for x in range(5)] == list(range(5))


def test_mean_of_two_is_numpys(chk):
    rng = np.random.default_rng(2)
    for dt, t in DTYPES.items():
        for _ in range(200):
            a = _image(rng, dt, (2,))
            if dt >= 2:
                a = (a * 10. ** rng.integers(-3, 4)).astype(t)
            assert chk.chk_mean_of_two(dt, float(a[0]), float(a[1])) == \
                float(np.median(a))


def test_reflect_index(chk):
    for n in (1, 2, 5, 8):
        for i in range(-3 * n - 2, 4 * n + 3):
            assert chk.chk_reflect(i, n) == int(ref.reflect_index(i, n))


@pytest.mark.parametrize('dt', [0, 1, 2, 3])
@pytest.mark.parametrize('shape,size', [((19, 23), 5), ((40, 33), 8),
                                        ((9, 70), 1), ((6, 5), 15),
                                        ((33, 41), 2)])
def test_filter_algorithms_match_scipy(chk, dt, shape, size):
    rng = np.random.default_rng(dt * 7 + size)
    for smooth in ((False, True) if dt >= 2 else (False,)):
        img = _image(rng, dt, shape, smooth)
        want = median_filter(img, size=size)
        paths = {0: (0, 1), 1: (1, 2)}.get(dt, (1,))
        for path in paths:
            got = _host_filter(chk, img, dt, size, path)
            np.testing.assert_array_equal(got, want, err_msg=f'path {path}')


def test_filter_algorithms_on_the_camera_frame(chk, frame_golden):
    g = frame_golden
    img = g['image']
    for size in (51, 8):
        for path in (0, 1):
            np.testing.assert_array_equal(_host_filter(chk, img, 0, size, path),
                                          g[f'bg{size}'])
    f32 = (img.astype(np.float32) / np.float32(100.))
    np.testing.assert_array_equal(_host_filter(chk, f32, 2, 51, 1),
                                  median_filter(f32, size=51))
    # 16-bit: the frame scaled to 12 bits, so the median's high byte moves
    u16 = (img.astype(np.uint16) * 16 + np.arange(img.size, dtype=np.uint16).reshape(
        img.shape) % 16)
    np.testing.assert_array_equal(_host_filter(chk, u16, 1, 51, 2),
                                  median_filter(u16, size=51))


def test_filter_planted_mistake_is_caught(chk):
    """A median off by one rank differs from scipy on ordinary data."""
    img = _image(np.random.default_rng(9), 0, (20, 20))
    want = median_filter(img, size=5)
    ranked = np.empty_like(img)
    pad = np.pad(img, 2, mode='symmetric')
    for y in range(20):
        for x in range(20):
            ranked[y, x] = np.sort(pad[y:y + 5, x:x + 5].ravel())[13]
    assert not np.array_equal(ranked, want)
    np.testing.assert_array_equal(_host_filter(chk, img, 0, 5, 0), want)
