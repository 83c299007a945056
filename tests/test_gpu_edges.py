"""Ragged and tiny inputs at tile boundaries, under every threads-per-tile
schedule: features of 1, 7, 1023, 1024, 1025 and 2049 pixels (explicit
coordinate lists and dense strips) give the FP64 oracle's hologram and
the Jacobian of single-feature evaluation; an LM over such a batch fits
each feature as it would alone."""
import numpy as np
import pytest

from conftest import INSTRUMENT
from oracle import lorenzmie as ref

pytestmark = pytest.mark.gpu

SIZES = (1, 7, 1023, 1024, 1025, 2049)


@pytest.fixture(scope='module')
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def ragged_batch(torch, rng):
    from pylorenzmie_amd import _lmz
    from pylorenzmie_amd.theory import Instrument
    inst = Instrument(noise=0.05, **INSTRUMENT)
    feats, coords, params, off = [], [], [], 0
    for i, n in enumerate(SIZES):
        q = (rng.uniform(-20, 20), rng.uniform(-20, 20), rng.uniform(80, 200),
             rng.uniform(0.5, 1.5), rng.uniform(1.4, 1.6), 0.002, 1.0)
        if i % 2:        # dense strip of width 37 (last row ragged)
            w = 37
            c = ref.meshgrid(((n + w - 1) // w, w))[:, :n]
            origin = (0., 0., 0.)
            feats.append(_lmz.make_feature(n, w, origin, inst, 60, off))
            coords.append(np.zeros((3, n)))
            q = (q[0] + 18, q[1] + 10) + q[2:]
        else:            # explicit coordinates scattered around the particle
            c = np.vstack([rng.uniform(-30, 30, n), rng.uniform(-30, 30, n)])
            origin = (0., 0., 0.)
            feats.append(_lmz.make_feature(n, 0, origin, inst, 60, off))
            coords.append(np.vstack([c, np.zeros((1, n))]))
        params.append((q, c))
        off += n
    return _lmz.Batch(feats, 1), np.concatenate(coords, 1), params, off, inst


@pytest.mark.parametrize('nt', [64, 128, 256])
def test_ragged_hologram_and_jacobian(torch, nt):
    from pylorenzmie_amd import _lmz
    rng = np.random.default_rng(nt)
    b, coords, params, npix, inst = ragged_batch(torch, rng)
    b.schedule = dict(jac_threads=nt, fwd_threads=nt, jac_blocks_per_cu=0,
                      fwd_blocks_per_cu=0)
    lib = _lmz.load_library()
    dev = 'cuda'
    p = np.zeros((len(SIZES), _lmz.NPARAM))
    p[:, :7] = [q for q, _ in params]
    P = torch.from_numpy(p).to(dev)
    C = torch.from_numpy(coords.astype(np.float32)).to(dev)
    nmax = torch.zeros(len(SIZES), dtype=torch.int32, device=dev)
    coef = torch.zeros(b.coef_floats(), dtype=torch.float32, device=dev)
    dcoef = torch.zeros(b.dcoef_floats(), dtype=torch.float32, device=dev)
    work = torch.zeros(b.work_doubles, dtype=torch.float64, device=dev)
    holo = torch.full((npix,), -1., dtype=torch.float32, device=dev)
    jac = torch.zeros((7, npix), dtype=torch.float32, device=dev)
    s = _lmz.stream_handle()
    ptr = _lmz.ptr
    _lmz.check(lib.lmz_mie_ab(b.handle, ptr(P), None, ptr(nmax), ptr(coef),
                              ptr(dcoef), None, ptr(work), s))
    _lmz.check(lib.lmz_hologram(b.handle, ptr(P), None, ptr(nmax), ptr(coef),
                                ptr(C), 3, None, ptr(holo), None, None, s))
    _lmz.check(lib.lmz_jacobian(b.handle, ptr(P), None, ptr(nmax), ptr(coef),
                                ptr(dcoef), ptr(C), 3, None, ptr(jac), None,
                                None, None, s))
    h = holo.cpu().numpy()
    J = jac.cpu().numpy()
    off = 0
    for n, (q, c) in zip(SIZES, params):
        want = ref.hologram(c, dict(zip(ref.PARAMS, q)), **INSTRUMENT)
        assert np.abs(h[off:off + n] - want).max() < 1e-5, n
        Jref = ref.jacobian(c, dict(zip(ref.PARAMS, q)), INSTRUMENT, order=6)
        for k in range(7):
            scale = max(np.abs(Jref[k]).max(), 1e-30)
            assert np.abs(J[k, off:off + n] - Jref[k]).max() <= 1e-3 * scale, (n, k)
        off += n
    assert off == npix


def test_ragged_lm_equals_single_fits(torch):
    from pylorenzmie_amd.analysis.lm import LMBatch, LMSettings
    from pylorenzmie_amd.theory import Instrument
    rng = np.random.default_rng(1)
    inst = Instrument(noise=0.05, **INSTRUMENT)
    crops, starts = [], []
    for n in (40, 1023, 1025, 2049, 3000):
        q = (0.3, -0.2, rng.uniform(100, 180), rng.uniform(0.6, 1.2), 1.5, 0., 1.)
        c = np.vstack([rng.uniform(-35, 35, n), rng.uniform(-35, 35, n)])
        d = ref.hologram(c, dict(zip(ref.PARAMS, q)), **INSTRUMENT) + \
            0.05 * rng.standard_normal(n)
        crops.append(dict(data=d.astype(np.float32), width=0, origin=(0., 0., 0.),
                          coords=np.vstack([c, np.zeros((1, n))]), nmax_cap=60))
        s = np.array(q)
        s[:3] += (0.5, -0.5, 2.)
        s[3] *= 1.02
        starts.append(s)
    settings = LMSettings(ftol=1e-10, xtol=1e-10, gtol=1e-12)
    lm = LMBatch(crops, inst)
    lm.load(np.array(starts))
    lm.run(settings, ref.PARAMS[:5])
    pb, rb = lm.fetch()
    for i, (c, s) in enumerate(zip(crops, starts)):
        one = LMBatch([c], inst)
        one.load(s[None])
        one.run(settings, ref.PARAMS[:5])
        p1, r1 = one.fetch()
        assert rb[i, 35] == pytest.approx(r1[0, 35], rel=1e-9), i
        np.testing.assert_allclose(pb[i, :5], p1[0, :5], rtol=1e-7)
