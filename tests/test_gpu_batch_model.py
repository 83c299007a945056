"""LorenzMieBatch: the reference's TorchLorenzMieBatch tests
(tests/test_torchlorenzmie.py:63-115) case by case, and every row against
the FP64 oracle (normalized hologram within 1e-5) for dense and explicit
coordinates, several particles per list, a Cluster, empty lists and the
aberration mask."""
import numpy as np
import pytest

from conftest import INSTRUMENT
from oracle import lorenzmie as ref

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def theory():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    from pylorenzmie_amd import theory
    return theory


def sphere(theory, x, y, z, a_p, n_p, k_p=0.):
    p = theory.Sphere()
    p.r_p = [x, y, z]
    p.a_p, p.n_p, p.k_p = a_p, n_p, k_p
    return p


@pytest.fixture
def case(theory):
    """tests/test_torchlorenzmie.py:65-83 of the reference."""
    shape = (32, 32)
    model = theory.LorenzMieBatch()
    model.coordinates = model.meshgrid(shape)
    lists = [[sphere(theory, 16., 16., 150., 0.75, 1.45)],
             [sphere(theory, 10., 20., 200., 1.0, 1.40)],
             [sphere(theory, 20., 10., 100., 0.5, 1.50),
              sphere(theory, 24., 24., 180., 0.8, 1.42)]]
    return shape, model, lists


def _oracle(coords, plist, **kw):
    ps = [dict(x_p=p.x_p + p.x_0, y_p=p.y_p + p.y_0, z_p=p.z_p + p.z_0,
               a_p=p.a_p, n_p=p.n_p, k_p=p.k_p) for q in plist for p in q]
    return ref.hologram(coords, ps, **INSTRUMENT, **kw)


def test_reference_batch_cases(case):
    """test_batch_hologram_shape, _dtype, _positive, _is_numpy_array,
    test_batch_matches_sequential."""
    shape, model, lists = case
    h = model.batch_hologram(lists)
    assert isinstance(h, np.ndarray)
    assert h.shape == (len(lists), shape[0] * shape[1])
    assert h.dtype == np.float64
    assert np.all(h > 0.)
    for i, plist in enumerate(lists):
        model.particle = plist
        np.testing.assert_allclose(h[i], model.hologram(), rtol=1e-5, atol=1e-7)


def test_rows_match_oracle(case):
    shape, model, lists = case
    h = model.batch_hologram(lists)
    c = ref.meshgrid(shape)
    for i, plist in enumerate(lists):
        assert np.abs(h[i] - _oracle(c, plist)).max() < 1e-5, i


def test_batch_field_matches_sequential(case, theory):
    import torch
    shape, model, lists = case
    f = model.batch_field(lists)
    assert isinstance(f, torch.Tensor) and f.dtype == torch.complex64
    assert tuple(f.shape) == (len(lists), 3, shape[0] * shape[1])
    fh = f.cpu().numpy()
    for i, plist in enumerate(lists):
        model.particle = plist
        want = model.field().copy()
        assert np.abs(fh[i] - want).max() < 1e-6 * max(1., np.abs(want).max())
    for cart, boh in ((False, True), (True, False)):
        f = model.batch_field(lists, cartesian=cart, bohren=boh).cpu().numpy()
        model.particle = lists[2]
        want = model.field(cartesian=cart, bohren=boh)
        assert np.abs(f[2] - want).max() < 1e-6 * max(1., np.abs(want).max())


def test_explicit_coordinates_cluster_and_empty_lists(theory):
    c = ref.meshgrid((40, 36), corner=(300., 200.))[:, ::3]
    model = theory.LorenzMieBatch(coordinates=c)
    cl = theory.Cluster(x_p=318., y_p=220., z_p=120.,
                        particles=[sphere(theory, -4., 1., 0., 0.6, 1.5),
                                   sphere(theory, 5., -2., 10., 0.7, 1.45, 0.002)])
    lists = [[sphere(theory, 317.3, 221.1, 140., 0.9, 1.55)], [], [cl]]
    h = model.batch_hologram(lists)
    assert h.shape == (3, c.shape[1])
    for i, plist in enumerate(lists):
        want = _oracle(c, plist) if plist else np.ones(c.shape[1])
        assert np.abs(h[i] - want).max() < 1e-5, i
    assert np.array_equal(model.batch_hologram([[], []]), np.ones((2, c.shape[1])))
    assert model.batch_hologram([]).shape == (0, c.shape[1])
    assert tuple(model.batch_field([[]]).shape) == (1, 3, c.shape[1])


def test_aberrated_batch(theory):
    AB = theory.Aberrated(theory.LorenzMieBatch)
    model = AB(spherical=0.6)
    model.coordinates = model.meshgrid((48, 48))
    lists = [[sphere(theory, 24.2, 23.7, 130., 0.8, 1.5)],
             [sphere(theory, 20., 26., 90., 1.1, 1.42),
              sphere(theory, 30., 18., 160., 0.6, 1.6)]]
    h = model.batch_hologram(lists)
    c = ref.meshgrid((48, 48))
    for i, plist in enumerate(lists):
        assert np.abs(h[i] - _oracle(c, plist, spherical=0.6)).max() < 1e-5, i


def test_batch_hologram_on_device(case):
    """device=True: the same holograms as a float32 tensor on the GPU."""
    import torch
    _, model, lists = case
    h = model.batch_hologram(lists)
    hd = model.batch_hologram(lists, device=True)
    assert isinstance(hd, torch.Tensor) and hd.dtype == torch.float32 and hd.is_cuda
    np.testing.assert_array_equal(hd.cpu().numpy().astype(np.float64), h)
    assert tuple(model.batch_hologram([[]], device=True).shape) == (1, h.shape[1])
