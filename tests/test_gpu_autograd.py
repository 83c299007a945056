"""Reverse-mode use of the hand-written Jacobian: HologramFunction (torch
autograd) with the lmz_hologram_vjp kernel in backward.

Checked against FP64 central differences of the oracle's loss (h = 1e-6
relative) and against the FP64 reference gradient 2 sum (I - d) dI/dp with
the oracle's 6th-order Jacobian, on uneven crops; and, at the full size of
256 features on 256 x 256 crops (alpha in [0.8, 1.2]), against the
contraction of the Jacobian written to HBM by lmz_jacobian."""
import numpy as np
import pytest

from conftest import INSTRUMENT
from oracle import lorenzmie as ref

pytestmark = pytest.mark.gpu

GRAD_TOL = 2e-3          # relative, per component, vs the FP64 gradients


@pytest.fixture(scope='module')
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def draw(rng, n, sides):
    """Config-2 parameter draws, positions at the crop centre +- 5 px."""
    p = np.zeros((n, 7))
    for i, (ny, nx) in enumerate(sides):
        p[i] = ((nx - 1) / 2 + rng.uniform(-5, 5), (ny - 1) / 2 + rng.uniform(-5, 5),
                rng.uniform(50., 400.), rng.uniform(0.4, 3.0),
                rng.uniform(1.38, 1.65), rng.uniform(0., 0.01),
                rng.uniform(0.8, 1.2))
    return p


def perturb(p):
    """The cfg0-style perturbed point: +-2 px, x(1 +- 0.05) a_p, +-0.02 n_p."""
    q = p.copy()
    q[:, 0] += 2.
    q[:, 1] -= 2.
    q[:, 2] += 2.
    q[:, 3] *= 1.05
    q[:, 4] -= 0.02
    return q


def batch(torch, sides, corners, params):
    from pylorenzmie_amd import _lmz
    from pylorenzmie_amd.theory import Instrument
    from pylorenzmie_amd.theory.batch import HologramBatch
    inst = Instrument(**INSTRUMENT)
    caps = [_lmz.nmax_cap(p[3], p[4], p[5], inst.n_m, inst.wavelength)
            for p in params]
    return HologramBatch(sides, caps, corners, inst)


def targets(torch, hb, truth, sigma, seed):
    hb.load(truth)
    clean = hb.compute().clone()
    g = torch.Generator(device='cuda')
    g.manual_seed(seed)
    return clean + sigma * torch.randn(clean.shape, generator=g,
                                       device='cuda', dtype=torch.float32)


def oracle_gradients(q, coords, d):
    """FP64 central differences of L = sum (I - d)^2 (h = 1e-6 relative,
    n_max held at the point's) and the FP64 reference gradient."""
    nmax = ref.wiscombe_yang(ref.size_parameter(q[3]),
                             (q[4] + 1j * q[5]) / INSTRUMENT['n_m'])

    def holo(v):
        return ref.hologram(coords, dict(x_p=v[0], y_p=v[1], z_p=v[2], a_p=v[3],
                                         n_p=v[4], k_p=v[5], nmax=nmax),
                            alpha=v[6], **INSTRUMENT)
    fd = np.zeros(7)
    for i in range(7):
        h = 1e-6 * max(abs(q[i]), 1e-2)
        vp, vm = q.copy(), q.copy()
        vp[i] += h
        vm[i] -= h
        fd[i] = (np.sum((holo(vp) - d) ** 2) - np.sum((holo(vm) - d) ** 2)) / (2 * h)
    J = ref.jacobian(coords, dict(zip(ref.PARAMS, q)), INSTRUMENT, order=6)
    g64 = 2 * J @ (holo(q) - d)
    return fd, g64


def rel_err(got, want):
    return np.abs(got - want) / np.maximum(np.abs(want), 1e-3 * np.abs(want).max())


def test_autograd_matches_fp64_gradients_uneven_crops(torch):
    from pylorenzmie_amd.theory.autograd import hologram
    sides = [(96, 96), (80, 112), (128, 72)]
    corners = [(0., 0.), (300., 40.), (17., 950.)]
    rng = np.random.default_rng(11)
    truth = draw(rng, 3, sides)
    truth[:, 0] += [c[0] for c in corners]
    truth[:, 1] += [c[1] for c in corners]
    truth[1, 3] = 2.9          # a large sphere: bright forward lobe
    hb = batch(torch, sides, corners, truth)
    d = targets(torch, hb, truth, 0.05, 5)
    q0 = perturb(truth)
    p = torch.tensor(q0, device='cuda', requires_grad=True)
    I = hologram(p, hb)
    L = ((I.double() - d.double()) ** 2).sum()
    L.backward()
    grad = p.grad.cpu().numpy()
    dh = d.cpu().double().numpy()
    for f, ((ny, nx), c) in enumerate(zip(sides, corners)):
        coords = ref.meshgrid((ny, nx), corner=c)
        o = hb.offsets[f]
        fd, g64 = oracle_gradients(q0[f], coords, dh[o:o + ny * nx])
        e_fd, e_64 = rel_err(grad[f], fd), rel_err(grad[f], g64)
        print(f'feature {f}: rel err vs FD {e_fd.max():.2e}, vs FP64 J {e_64.max():.2e}')
        assert np.all(e_fd <= GRAD_TOL), (f, e_fd)
        assert np.all(e_64 <= GRAD_TOL), (f, e_64)


def test_autograd_linear_in_upstream_gradient(torch):
    from pylorenzmie_amd.theory.autograd import hologram
    sides = [(64, 64)] * 4
    rng = np.random.default_rng(3)
    truth = draw(rng, 4, sides)
    hb = batch(torch, sides, None, truth)
    q0 = perturb(truth)
    w = torch.randn(hb.npix_total, device='cuda', dtype=torch.float32)
    grads = []
    for scale in (1., -4.):     # a power of two: every rounding scales too
        p = torch.tensor(q0, device='cuda', requires_grad=True)
        (hologram(p, hb) * (scale * w)).sum().backward()
        grads.append(p.grad.cpu().numpy())
    np.testing.assert_array_equal(grads[1], -4. * grads[0])
    # the forward value is the FP64-kernel hologram
    hb.load(q0)
    np.testing.assert_array_equal(hologram(torch.tensor(q0, device='cuda'), hb).cpu(),
                                  hb.compute().cpu())


def test_config2_full_size_vjp_vs_materialized_jacobian(torch):
    """256 features, 256 x 256 crops, alpha in [0.8, 1.2], sigma = 0.05."""
    from pylorenzmie_amd import _lmz
    from pylorenzmie_amd.theory.autograd import hologram
    n, side = 256, 256
    sides = [(side, side)] * n
    rng = np.random.default_rng(2)
    truth = draw(rng, n, sides)
    hb = batch(torch, sides, None, truth)
    d = targets(torch, hb, truth, 0.05, 2)
    q0 = perturb(truth)
    p = torch.tensor(q0, device='cuda', requires_grad=True)
    I = hologram(p, hb)
    r2 = 2 * (I.detach().double() - d.double())
    ((I.double() - d.double()) ** 2).sum().backward()
    grad = p.grad.cpu().numpy()
    assert np.all(np.isfinite(grad))
    # J written to HBM by lmz_jacobian, contracted in FP64
    lib = _lmz.load_library()
    b = hb.batch
    pp = torch.zeros((n, _lmz.NPARAM), dtype=torch.float64, device='cuda')
    pp[:, :7] = p.detach()
    nmax = torch.zeros(n, dtype=torch.int32, device='cuda')
    coef = torch.zeros(b.coef_floats(), dtype=torch.float32, device='cuda')
    dcoef = torch.zeros(b.dcoef_floats(), dtype=torch.float32, device='cuda')
    jac = torch.empty((7, hb.npix_total), dtype=torch.float32, device='cuda')
    s = _lmz.stream_handle()
    _lmz.check(lib.lmz_mie_ab(b.handle, _lmz.ptr(pp), None, _lmz.ptr(nmax),
                              _lmz.ptr(coef), _lmz.ptr(dcoef), None,
                              _lmz.ptr(hb.work), s))
    _lmz.check(lib.lmz_jacobian(b.handle, _lmz.ptr(pp), None, _lmz.ptr(nmax),
                                _lmz.ptr(coef), _lmz.ptr(dcoef), None, 3, None,
                                _lmz.ptr(jac), None, None, None, s))
    Jg = jac.double().view(7, n, side * side) * r2.view(1, n, side * side)
    want = Jg.sum(2).T.cpu().numpy()
    scale = Jg.abs().sum(2).T.cpu().numpy()
    err = np.abs(grad - want) / scale
    print('max |vjp - J^T g| / sum |J g|:', err.max())
    assert err.max() <= 1e-5
    # two features against the FP64 oracle at full size
    dh = d.cpu().double().numpy()
    coords = ref.meshgrid((side, side))
    for f in (0, int(np.argmax(truth[:, 3]))):
        o = hb.offsets[f]
        fd, g64 = oracle_gradients(q0[f], coords, dh[o:o + side * side])
        e = np.maximum(rel_err(grad[f], fd), rel_err(grad[f], g64))
        print(f'feature {f} (a_p {truth[f, 3]:.2f}): rel err {e.max():.2e}')
        assert np.all(e <= GRAD_TOL), (f, e)
