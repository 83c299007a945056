"""Frame normalization and batch feeding on the GPU against numpy, scipy
and the oracle, bit for bit: lmz_median (np.median), lmz_median_filter
(scipy.ndimage.median_filter, both the 8-bit sliding kernel and the radix
kernel, staged in LDS and not), lmz_normalize (utilities/normalizer.py),
lmz_median3_u8 (utilities/vmedian.py) and lmz_gather_frame (a Frame fitted
from a device frame equals the same Frame fitted from host memory)."""
import numpy as np
import pytest
from scipy.ndimage import median_filter

from oracle import normalize as ref

pytestmark = pytest.mark.gpu

DT = {np.uint8: 0, np.uint16: 1, np.float32: 2, np.float64: 3}


@pytest.fixture(scope='module')
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


@pytest.fixture(scope='module')
def lib(torch):
    from pylorenzmie_amd import _lmz
    return _lmz.load_library()


def _img(rng, dt, shape, smooth=False):
    if dt is np.uint8:
        return rng.integers(0, 256, shape, dtype=np.uint8)
    if dt is np.uint16:
        return rng.integers(0, 4096, shape, dtype=np.uint16)
    if smooth:
        y, x = np.mgrid[:shape[0], :shape[1]]
        v = 1. + 0.3 * np.cos(0.3 * np.hypot(x - shape[1] / 2.3, y - shape[0] / 1.7))
        v += 0.01 * rng.standard_normal(shape)
    else:
        v = 40. * rng.standard_normal(shape)
    return v.astype(dt)


def _hologram_frame(rng, ny=1024, nx=1280):
    """A camera-like 8-bit frame: fringes of two features on a sloped
    background, with shot noise."""
    y, x = np.mgrid[:ny, :nx].astype(float)
    bg = 90. + 20. * x / nx + 10. * y / ny
    v = bg.copy()
    for (cx, cy, k) in ((400., 500., 0.35), (900., 300., 0.25)):
        r = np.hypot(x - cx, y - cy)
        v += 40. * np.cos(k * r) * np.exp(-r / 150.)
    v += rng.normal(0., 3., v.shape)
    return np.clip(v, 0, 255).astype(np.uint8)


def _filter_gpu(torch, lib, img, size):
    from pylorenzmie_amd import _lmz
    d = torch.from_numpy(np.ascontiguousarray(img)).cuda()
    bg = torch.empty_like(d)
    _lmz.check(lib.lmz_median_filter(d.data_ptr(), DT[img.dtype.type],
                                     img.shape[0], img.shape[1], size,
                                     bg.data_ptr(), None, 2,
                                     _lmz.stream_handle()))
    return bg.cpu().numpy()


# ------------------------------------------------------------------ median
@pytest.mark.parametrize('dt', [np.uint8, np.uint16, np.float32, np.float64])
def test_median_matches_numpy(torch, dt):
    from pylorenzmie_amd.utilities import Normalizer
    rng = np.random.default_rng(1)
    cases = [_img(rng, dt, (1, 1)), _img(rng, dt, (1, 2)), _img(rng, dt, (7, 9)),
             _img(rng, dt, (8, 9)), _img(rng, dt, (256, 320)),
             _img(rng, dt, (1024, 1280)), np.full((6, 5), 7, dt),
             _img(rng, dt, (3, 4)).repeat(5, axis=0)]
    if dt in (np.float32, np.float64):
        cases.append(np.array([[-0.0, 0.0, -1.5, np.inf, -np.inf, 3.]], dt))
        cases.append(_img(rng, dt, (100, 101), smooth=True))
    for c in cases:
        med = Normalizer.median(torch.from_numpy(np.ascontiguousarray(c)).cuda())
        got = float(med.cpu().numpy()[0])
        want = np.median(c)
        assert got == float(want), (c.shape, got, want)


def test_median_nan_and_empty(torch):
    from pylorenzmie_amd.utilities import Normalizer
    a = np.arange(12, dtype=np.float32).reshape(3, 4)
    a[1, 2] = np.nan
    med = Normalizer.median(torch.from_numpy(a).cuda())
    assert np.isnan(float(med.cpu()[0])) and np.isnan(np.median(a))
    med = Normalizer.median(torch.empty((0, 5), dtype=torch.float32).cuda())
    assert np.isnan(float(med.cpu()[0]))


# ------------------------------------------------------------------ filter
@pytest.mark.parametrize('dt', [np.uint8, np.uint16, np.float32, np.float64])
@pytest.mark.parametrize('shape,size', [((37, 45), 5), ((40, 33), 8),
                                        ((9, 70), 1), ((6, 5), 15),
                                        ((33, 41), 2), ((64, 80), 51)])
def test_filter_matches_scipy(torch, lib, dt, shape, size):
    rng = np.random.default_rng(DT[dt] * 31 + size)
    for smooth in ((False, True) if dt in (np.float32, np.float64) else (False,)):
        img = _img(rng, dt, shape, smooth)
        np.testing.assert_array_equal(_filter_gpu(torch, lib, img, size),
                                      median_filter(img, size=size))


@pytest.mark.parametrize('dt', [np.float32, np.float64])
def test_filter_wide_windows_unstaged(torch, lib, dt):
    """Windows whose keys do not fit LDS are read from the frame."""
    rng = np.random.default_rng(4)
    img = _img(rng, dt, (40, 150), smooth=True)
    for size in (127, 255):
        np.testing.assert_array_equal(_filter_gpu(torch, lib, img, size),
                                      median_filter(img, size=size))


def test_filter_camera_frame_golden(torch, lib, frame_golden):
    g = frame_golden
    for size in (51, 8):
        np.testing.assert_array_equal(_filter_gpu(torch, lib, g['image'], size),
                                      g[f'bg{size}'])


def test_filter_full_frame(torch, lib):
    """1280 x 1024 at size 51: the 8-bit sliding kernel, the radix kernel
    on the same values as uint16 and as float32 / 100 (an increasing map
    commutes with the median) agree exactly, and both match the definition
    at sampled pixels, borders included."""
    rng = np.random.default_rng(7)
    img = _hologram_frame(rng)
    b8 = _filter_gpu(torch, lib, img, 51)
    b16 = _filter_gpu(torch, lib, img.astype(np.uint16), 51)
    np.testing.assert_array_equal(b16, b8)
    f = img.astype(np.float32) / np.float32(100.)
    bf = _filter_gpu(torch, lib, f, 51)
    np.testing.assert_array_equal(bf, b8.astype(np.float32) / np.float32(100.))
    ny, nx = img.shape
    rows = np.concatenate([rng.integers(0, ny, 200), [0, ny - 1, 0, ny - 1, 25, 1000]])
    cols = np.concatenate([rng.integers(0, nx, 200), [0, 0, nx - 1, nx - 1, 1279, 24]])
    np.testing.assert_array_equal(ref.median_filter_brute(img, 51, (rows, cols)),
                                  b8[rows, cols])


# -------------------------------------------------------------- normalizer
@pytest.mark.parametrize('dt', [np.uint8, np.uint16, np.float32, np.float64])
def test_normalizer_matches_reference_numpy(torch, dt):
    from pylorenzmie_amd.utilities import Normalizer
    rng = np.random.default_rng(11)
    img = _img(rng, dt, (48, 61), smooth=True)
    if dt in (np.float32, np.float64):
        img = np.abs(img)
    cases = [dict(method='median'), dict(method='filter', size=9),
             dict(method='filter', size=4),
             dict(method='reference', reference=_img(rng, np.uint8, (48, 61))),
             dict(method='reference', reference=np.abs(_img(rng, np.float32, (48, 61)))),
             dict(method='reference', reference=2.5),
             dict(method='reference', reference=np.float32(3.25)),
             dict(method='reference', reference=0.)]
    for kw in cases:
        got = Normalizer(**kw)(img)
        want = ref.normalize(img, **kw)
        assert got.dtype == want.dtype, kw
        np.testing.assert_array_equal(got, want, err_msg=str(kw))


def test_normalizer_zero_background_and_fallback(torch, caplog):
    from pylorenzmie_amd.utilities import Normalizer
    img = np.zeros((20, 30), np.uint8)
    img[5:9, 4:20] = 200
    np.testing.assert_array_equal(Normalizer('median')(img),
                                  ref.normalize(img, 'median'))
    np.testing.assert_array_equal(Normalizer('filter', size=3)(img),
                                  ref.normalize(img, 'filter', size=3))
    bad = np.ones((5, 5), np.uint8)
    with caplog.at_level('WARNING'):
        got = Normalizer('reference', reference=bad)(img)
    assert 'falling back to median' in caplog.text
    np.testing.assert_array_equal(got, ref.normalize(img, 'reference',
                                                     reference=bad))


def test_normalizer_device_frame(torch, frame_golden):
    """A device frame stays on the device, float32 by default."""
    from pylorenzmie_amd.utilities import Normalizer
    g = frame_golden
    d = torch.from_numpy(g['image']).cuda()
    out = Normalizer('filter', size=51)(d)
    assert out.is_cuda and out.dtype == torch.float32
    bg = g['bg51']
    want = (g['image'] / np.where(bg > 0, bg, 1.)).astype(np.float32)
    np.testing.assert_array_equal(out.cpu().numpy(), want)
    out = Normalizer('reference', reference=torch.from_numpy(g['background']).cuda())(d)
    want = ref.normalize(g['image'], 'reference', reference=g['background'])
    np.testing.assert_array_equal(out.cpu().numpy(), want.astype(np.float32))
    out = Normalizer()(d, dtype=np.float64)
    np.testing.assert_array_equal(out.cpu().numpy(), ref.normalize(g['image']))


# ---------------------------------------------------------------- vmedian
@pytest.mark.parametrize('order', [0, 1, 2])
def test_vmedian_matches_oracle(torch, order):
    from pylorenzmie_amd.utilities import VMedian
    rng = np.random.default_rng(order)
    frames = rng.integers(0, 256, (3 ** (order + 1) + 5, 12, 17), dtype=np.uint8)
    mine, theirs = VMedian(order=order), ref.VMedian(order=order)
    dev = VMedian(order=order)
    for f in frames:
        mine.add(f)
        theirs.add(f)
        dev.add(torch.from_numpy(f).cuda())
        assert mine.initialized == theirs.initialized
        assert mine.cycled == theirs.cycled
        np.testing.assert_array_equal(mine.get(), theirs.get())
        np.testing.assert_array_equal(dev.get().cpu().numpy(), theirs.get())


def test_vmedian_results_are_not_changed_by_later_frames(torch):
    """Each get() is a new array, as the reference's np.median result is
    (utilities/vmedian.py:73): the in-place running median does not reach a
    result returned earlier, for host or device frames."""
    from pylorenzmie_amd.utilities import VMedian
    rng = np.random.default_rng(9)
    frames = rng.integers(0, 256, (12, 8, 9), dtype=np.uint8)
    for device in (False, True):
        vm = VMedian(order=1)
        kept = []
        for f in frames:
            vm.add(torch.from_numpy(f).cuda() if device else f)
            got = vm.get()
            kept.append((got, (got.cpu().numpy() if device else got).copy()))
        for got, snap in kept:
            now = got.cpu().numpy() if device else got
            np.testing.assert_array_equal(now, snap)


# ----------------------------------------------------------- gather / Frame
def _frame_pair(torch, corner, mask_fraction):
    from pylorenzmie_amd.analysis import Frame
    from pylorenzmie_amd.theory import Instrument, LorenzMie, Sphere
    inst = Instrument(n_m=1.34, wavelength=0.447, magnification=0.048, noise=0.05)
    ny, nx = 300, 420
    m = LorenzMie(instrument=inst,
                  particle=[Sphere(x_p=corner[0] + 120.3, y_p=corner[1] + 140.6,
                                   z_p=110., a_p=0.75, n_p=1.44),
                            Sphere(x_p=corner[0] + 300.2, y_p=corner[1] + 150.1,
                                   z_p=140., a_p=0.6, n_p=1.5)])
    from pylorenzmie_amd.lib import meshgrid
    m.coordinates = meshgrid((ny, nx), corner=corner)
    rng = np.random.default_rng(3)
    holo = (m.hologram() + 0.05 * rng.standard_normal(ny * nx)).reshape(ny, nx)
    frames = []
    for data in (holo.astype(np.float32),
                 torch.from_numpy(holo.astype(np.float32)).cuda()):
        f = Frame(data, corner=corner, instrument=inst)
        f.bboxes = [((60, 80), 121, 121), ((240, 90), 121, 121)]
        for feat, (x, y) in zip(f.features, ((120.3, 140.6), (300.2, 150.1))):
            feat.particle.x_p, feat.particle.y_p = corner[0] + x + 0.4, corner[1] + y - 0.3
            feat.particle.z_p, feat.particle.a_p, feat.particle.n_p = 115., 0.72, 1.45
            if mask_fraction is None:
                feat.optimizer.mask = None
            else:
                feat.mask.fraction = mask_fraction
        frames.append(f)
    return frames


@pytest.mark.parametrize('corner,fraction', [((0., 0.), 0.3),
                                             ((100., 250.), 0.3),
                                             ((0., 0.), None)])
def test_frame_from_device_equals_host(torch, corner, fraction):
    from pylorenzmie_amd.analysis import lm as lmmod
    host, dev = _frame_pair(torch, corner, fraction)
    seen = []
    orig = lmmod.LMBatch.fetch

    def spy(self):
        seen.append((self.data.cpu().numpy().copy(),
                     None if self.coords is None else self.coords.cpu().numpy().copy()))
        return orig(self)
    lmmod.LMBatch.fetch = spy
    try:
        np.random.seed(21)
        rh = host.optimize()
        np.random.seed(21)
        rd = dev.optimize()
    finally:
        lmmod.LMBatch.fetch = orig
    (dh, ch), (dd, cd) = seen
    np.testing.assert_array_equal(dd, dh)
    if ch is None:
        assert cd is None
    else:
        np.testing.assert_array_equal(cd, ch)
    for k in ('x_p', 'y_p', 'z_p', 'a_p', 'n_p'):
        np.testing.assert_allclose(rd[k].to_numpy(), rh[k].to_numpy(),
                                   rtol=0, atol=0)
    assert list(rd['success']) == list(rh['success'])


def test_gather_out_of_box_reads_nan(torch):
    """Indices past the box and boxes past the frame give NaN, not faults."""
    from pylorenzmie_amd import _lmz
    from pylorenzmie_amd.analysis.lm import LMBatch
    from pylorenzmie_amd.theory import Instrument
    inst = Instrument()
    frame = torch.arange(20 * 30, dtype=torch.float32).reshape(20, 30).cuda()
    xs, ys = np.arange(30.), np.arange(20.)
    crops = [dict(box=(2, 3, 4, 5), npix=20, width=4, origin=(2., 3., 0.),
                  nmax_cap=20),
             dict(box=(10, 10, 5, 5), npix=4, width=0, origin=(12., 12., 0.),
                  index=np.array([0, 6, 24, 3]), nmax_cap=20)]
    lm = LMBatch(crops, inst, frame=(frame, xs, ys))
    d = lm.data.cpu().numpy()
    want0 = frame.cpu().numpy()[3:8, 2:6].ravel()
    np.testing.assert_array_equal(d[:20], want0)
    np.testing.assert_array_equal(d[20:], frame.cpu().numpy()[[10, 11, 14, 10],
                                                              [10, 11, 14, 13]])
    c = lm.coords.cpu().numpy()[:, 20:]
    np.testing.assert_array_equal(c[0], np.array([10, 11, 14, 13]) - 12.)
    np.testing.assert_array_equal(c[1], np.array([10, 11, 14, 10]) - 12.)
    lm.close()
    # a device-side box past the frame edge reads NaN
    lm = LMBatch(crops[:1], inst, frame=(frame, xs, ys))
    lm._box.copy_(torch.tensor([[28, 3, 4, 5]], dtype=torch.int32))
    _lmz.check(_lmz.load_library().lmz_gather_frame(
        lm.batch.handle, frame.data_ptr(), 20, 30, lm._xs.data_ptr(),
        lm._ys.data_ptr(), lm._box.data_ptr(), None, lm.data.data_ptr(), None,
        _lmz.stream_handle()))
    d = lm.data.cpu().numpy().reshape(5, 4)
    assert np.isnan(d[:, 2:]).all() and np.isfinite(d[:, :2]).all()
    lm.close()


# ------------------------------------------------------------ edge cases
def test_filter_one_row_one_column_and_tiny(torch, lib):
    """Frames narrower than a tile and windows far wider than the frame,
    against the filter's definition (periodic 'reflect' windows).  scipy
    1.15's rank filter reads uninitialized memory once the window exceeds
    the frame by several frame sizes (a 51-wide window on a 3 x 2 image
    holds values like 3e-322), so it is no reference there."""
    rng = np.random.default_rng(13)
    for shape in ((1, 1), (1, 37), (29, 1), (3, 2), (65, 1), (1, 70)):
        for dt in (np.uint8, np.float32):
            img = _img(rng, dt, shape)
            for size in (1, 2, 5, 51):
                np.testing.assert_array_equal(_filter_gpu(torch, lib, img, size),
                                              ref.median_filter_brute(img, size),
                                              err_msg=f'{shape} {dt} {size}')


def test_normalizer_rejects_what_it_cannot_do(torch):
    from pylorenzmie_amd import _lmz
    from pylorenzmie_amd.utilities import Normalizer
    with pytest.raises(ValueError):
        Normalizer('filter', size=0)(np.ones((4, 4), np.uint8))
    with pytest.raises(ValueError):
        Normalizer('filter', size=_lmz.MAX_FILTER_SIZE + 1)(np.ones((4, 4), np.uint8))
    with pytest.raises(ValueError):
        Normalizer()(np.ones(5, np.uint8))
    with pytest.raises(TypeError):
        Normalizer()(np.ones((4, 4), np.uint8), dtype=np.int32)
    # int16 frames are converted exactly and give numpy's float64 result;
    # float16 frames come back as float16, computed in float64
    img = (np.arange(30).reshape(5, 6) - 7).astype(np.int16)
    got = Normalizer('median')(img)
    want = ref.normalize(img, 'median')
    assert got.dtype == want.dtype
    np.testing.assert_array_equal(got, want)
    img = (np.arange(30).reshape(5, 6) + 3).astype(np.float16)
    got = Normalizer('median')(img)
    want = ref.normalize(img, 'median')
    assert got.dtype == want.dtype == np.float16
    np.testing.assert_allclose(got.astype(float), want.astype(float), rtol=1e-3)


def test_frame_device_boxes_at_edges(torch):
    """Boxes clipped by the frame's edge (numpy slicing) on a device frame
    give the same batch as on a host frame."""
    from pylorenzmie_amd.analysis import Frame
    from pylorenzmie_amd.analysis import lm as lmmod
    from pylorenzmie_amd.theory import Instrument
    rng = np.random.default_rng(5)
    data = (1. + 0.05 * rng.standard_normal((120, 150))).astype(np.float32)
    inst = Instrument(noise=0.05)
    seen = []
    orig = lmmod.LMBatch.fetch

    def spy(self):
        seen.append(self.data.cpu().numpy().copy())
        return orig(self)
    lmmod.LMBatch.fetch = spy
    try:
        for d in (data, torch.from_numpy(data).cuda()):
            f = Frame(d, instrument=inst)
            f.bboxes = [((100, 80), 80, 80), ((0, 0), 50, 50)]
            for ft in f.features:
                ft.optimizer.mask = None
                ft.particle.z_p = 100.
                ft.optimizer.settings['max_nfev'] = 3
            f.optimize()
    finally:
        lmmod.LMBatch.fetch = orig
    assert seen[0].size == 40 * 50 + 50 * 50
    np.testing.assert_array_equal(seen[1], seen[0])
