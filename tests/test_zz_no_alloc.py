"""lmz_hologram_vjp_static (the allocation-free VJP: per-tile partials in a
fixed device array) against lmz_hologram_vjp (caller-supplied tile_work),
called through the C ABI on the same inputs: identical gradients on two
calls in a row, the second with a smaller batch, and the size limit
rejected on the host."""
import ctypes

import numpy as np
import pytest

from conftest import INSTRUMENT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def _draw(rng, n):
    shapes, params = [], np.zeros((n, 7))
    for i in range(n):
        s = int(rng.integers(40, 161))
        shapes.append((s, s))
        c = (s - 1) / 2
        params[i] = (c + rng.uniform(-5, 5), c + rng.uniform(-5, 5),
                     rng.uniform(60., 300.), rng.uniform(0.4, 2.0),
                     rng.uniform(1.38, 1.65), rng.uniform(0., 0.01),
                     rng.uniform(0.8, 1.2))
    return shapes, params


def _both(torch, shapes, params, seed):
    """Gradients of sum(g I) from both entry points on one batch."""
    from pylorenzmie_amd import _lmz
    from pylorenzmie_amd.theory import Instrument
    from pylorenzmie_amd.theory.batch import HologramBatch
    inst = Instrument(**INSTRUMENT)
    caps = [_lmz.nmax_cap(p[3], p[4], p[5], inst.n_m, inst.wavelength)
            for p in params]
    hb = HologramBatch(shapes, caps, instrument=inst)
    hb.load(params)
    b = hb.batch
    dev = hb.params.device
    lib = _lmz.load_library()
    s = _lmz.stream_handle()
    nmax = torch.zeros(hb.nfeat, dtype=torch.int32, device=dev)
    coef = torch.zeros(b.coef_floats(), dtype=torch.float32, device=dev)
    dcoef = torch.zeros(b.dcoef_floats(), dtype=torch.float32, device=dev)
    _lmz.check(lib.lmz_mie_ab(b.handle, _lmz.ptr(hb.params), None,
                              _lmz.ptr(nmax), _lmz.ptr(coef), _lmz.ptr(dcoef),
                              None, _lmz.ptr(hb.work), s))
    gen = torch.Generator(device='cuda')
    gen.manual_seed(seed)
    g = torch.randn(hb.npix_total, generator=gen, device=dev,
                    dtype=torch.float32)
    tiles = torch.zeros(b.ntiles * _lmz.NJAC, dtype=torch.float64, device=dev)
    want = torch.full((hb.nfeat, _lmz.NJAC), np.nan, dtype=torch.float64,
                      device=dev)
    got = torch.full_like(want, np.nan)
    _lmz.check(lib.lmz_hologram_vjp(
        b.handle, _lmz.ptr(hb.params), _lmz.ptr(nmax), _lmz.ptr(coef),
        _lmz.ptr(dcoef), None, hb.flags, _lmz.ptr(g), _lmz.ptr(tiles),
        _lmz.ptr(want), s))
    _lmz.check(lib.lmz_hologram_vjp_static(
        b.handle, _lmz.ptr(hb.params), _lmz.ptr(nmax), _lmz.ptr(coef),
        _lmz.ptr(dcoef), None, hb.flags, _lmz.ptr(g), _lmz.ptr(got), s))
    torch.cuda.synchronize()
    return b.ntiles, want.cpu().numpy(), got.cpu().numpy()


def test_static_vjp_equals_vjp_on_two_calls(torch):
    rng = np.random.default_rng(7)
    # first call: 48 features (~ 0.6 M pixels); second: 5 smaller ones, so
    # rows left in the fixed array by the first call lie past the second's
    # tiles and must not reach its sums
    for n, seed in ((48, 1), (5, 2)):
        shapes, params = _draw(rng, n)
        if n == 5:
            shapes = [(s // 2, s // 2) for s, _ in shapes]
            params[:, :2] = [((s - 1) / 2, (s - 1) / 2) for s, _ in shapes]
        ntiles, want, got = _both(torch, shapes, params, seed)
        print(f'{n} features, {ntiles} tiles: max |static - vjp| '
              f'{np.abs(got - want).max():.3e}')
        assert np.all(np.isfinite(want))
        np.testing.assert_array_equal(got, want)


def test_static_vjp_rejects_batches_beyond_its_scratch(torch):
    """More tiles than the fixed array holds is LMZ_EINVAL, before any
    launch."""
    from pylorenzmie_amd import _lmz
    lib = _lmz.load_library()
    assert _lmz.VJP_STATIC_MAX_TILES == 65536
    side, rows = 8192, 2000          # 16,000 tiles per dense crop
    n = side * rows
    feats = [_lmz.make_feature(n, side, (0., 0., 0.), _instrument(), 30, i * n)
             for i in range(5)]
    b = _lmz.Batch(feats, 1)
    assert b.ntiles > _lmz.VJP_STATIC_MAX_TILES
    fake = ctypes.c_void_p(256)       # never dereferenced: rejected first
    rc = lib.lmz_hologram_vjp_static(b.handle, fake, fake, fake, fake, None,
                                     3, fake, fake, None)
    assert rc == 1
    assert b'LMZ_VJP_STATIC_MAX_TILES' in lib.lmz_last_error()


def _instrument():
    from pylorenzmie_amd.theory import Instrument
    return Instrument(**INSTRUMENT)
