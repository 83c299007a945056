"""The reference's own unit tests of the classes that compute on the GPU,
case by case: tests/test_sphere.py (ab and mie_coefficients),
test_particle.py, test_lorenzmie.py (field and hologram with every flag)
and test_feature.py (estimate, predicted, residuals on the reference's
test hologram tests/data/crop.png, carried as tests/golden/crop.npz).
The values behind these shapes are checked against the oracle and the
reference's golden vectors in tests/test_gpu_forward.py."""
import numpy as np
import pandas as pd
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def api():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    from pylorenzmie_amd import analysis, theory
    return theory, analysis


def test_sphere_ab_cases(api):
    """test_ab_shape, test_ab_dtype, test_ab_zeroth_order,
    test_mie_coefficients_shape, _dtype, _zeroth_order."""
    theory, _ = api
    s = theory.Sphere(a_p=1., n_p=1.4, k_p=0.)
    for ab in (s.ab(1.34, 0.447),
               theory.Sphere.mie_coefficients(s.a_p, s.n_p, s.k_p, 1.34, 0.447)):
        assert ab.ndim == 2 and ab.shape[1] == 2
        assert ab.dtype == np.complex128
        np.testing.assert_array_equal(ab[0, :], 0j)
    np.testing.assert_array_equal(s.ab(1.34, 0.447),
                                  theory.Sphere.mie_coefficients(1., 1.4, 0., 1.34, 0.447))


def test_lorenzmie_field_cases(api):
    """test_field_nocoordinates, test_field, test_field_bohren,
    test_field_cartesian, test_field_both, test_hologram."""
    theory, _ = api
    m = theory.LorenzMie()
    m.coordinates = None
    assert isinstance(m.field(), np.ndarray)
    p = m.particle
    p.a_p, p.n_p = 1., 1.4
    p.r_p = [64, 64, 100]
    c = m.meshgrid([128, 128])
    m.coordinates = c
    fields = {}
    for bohren in (False, True):
        for cartesian in (False, True):
            f = m.field(bohren=bohren, cartesian=cartesian)
            assert f.shape[1] == c.shape[1]
            fields[bohren, cartesian] = f.copy()
    # every call refills the model's one field buffer, as the reference's
    # field() returns self._field (theory/LorenzMie.py:178-181)
    assert m.field() is m.field()
    # the flags change the field (each is checked against the reference's
    # golden fields elsewhere)
    assert not np.allclose(fields[False, False], fields[True, True])
    holo = m.hologram()
    assert holo.size == c.shape[1]


@pytest.fixture
def feature(api, crop_golden):
    """tests/test_feature.py:16-32 of the reference."""
    theory, analysis = api
    data = crop_golden['image'].astype(float) / 100.
    model = theory.LorenzMie()
    model.instrument.wavelength = 0.447
    model.instrument.magnification = 0.048
    model.instrument.n_m = 1.34
    h, w = data.shape
    model.particle.r_p = [w / 2., h / 2., 330.]
    model.particle.a_p = 1.1
    model.particle.n_p = 1.4
    f = analysis.Feature(analysis.Hologram(data), model=model)
    f.mask.fraction = 0.1
    return f


def test_feature_estimate_cases(feature):
    """test_estimate_returns_series, test_estimate_sets_z_p,
    test_estimate_sets_a_p, test_estimate_uses_coordinates_for_position."""
    result = feature.estimate()
    assert isinstance(result, pd.Series)
    assert feature.particle.z_p > 0
    assert feature.particle.a_p > 0
    assert feature.particle.x_p == pytest.approx(
        float(feature.coordinates[0].mean()), abs=1e-5)
    assert feature.particle.y_p == pytest.approx(
        float(feature.coordinates[1].mean()), abs=1e-5)


def test_feature_predicted_and_residuals(feature):
    """test_predicted_shape, test_residuals_shape."""
    assert feature.predicted().shape == feature.shape
    r = feature.residuals()
    assert r.shape == feature.shape
    np.testing.assert_allclose(r, feature.predicted() - feature.data)


@pytest.fixture
def estimator_case(api):
    """tests/test_deestimator.py:13-24 and test_radialestimator.py:13-24."""
    theory, analysis = api
    shape = (101, 101)
    model = theory.LorenzMie()
    model.coordinates = theory.LorenzMie.meshgrid(shape)
    model.particle.r_p = [50., 50., 200.]
    model.particle.a_p = 0.75
    model.particle.n_p = 1.45
    hologram = analysis.Hologram(model.hologram().reshape(shape))
    return model, hologram


def _quick(est):
    est.settings['maxiter'] = 1
    est.settings['workers'] = 1
    return est


@pytest.mark.parametrize('name', ['DEEstimator', 'RadialEstimator'])
def test_estimator_cases(api, estimator_case, name):
    """test_returns_series, test_result_has_particle_keys,
    test_xy_pinned_to_coordinate_means, test_model_coordinates_restored,
    test_model_coordinates_restored_on_exception, test_custom_bounds,
    test_corner_aware (radial), test_exclude_filters_pixels (DE)."""
    _, analysis = api
    cls = getattr(analysis, name)
    model, hologram = estimator_case
    est = _quick(cls(model=model, seed=0))
    before = model.coordinates.copy()
    result = est.estimate(hologram)
    assert isinstance(result, pd.Series)
    for key in ('x_p', 'y_p', 'z_p', 'a_p', 'n_p'):
        assert key in result
    assert result['x_p'] == pytest.approx(float(hologram.coordinates[0].mean()),
                                          abs=5e-4)
    assert result['y_p'] == pytest.approx(float(hologram.coordinates[1].mean()),
                                          abs=5e-4)
    np.testing.assert_array_equal(model.coordinates, before)
    bad = cls(model=model, bounds={'z_p': (500., 100.)}, seed=0)
    try:
        bad.estimate(hologram)
    except Exception:
        pass
    np.testing.assert_array_equal(model.coordinates, before)
    custom = _quick(cls(model=model, bounds={'z_p': (150., 250.)}, seed=0))
    assert 150. <= custom.estimate(hologram)['z_p'] <= 250.
    h = analysis.Hologram(hologram.data, corner=(10., 20.))
    result = est.estimate(h)
    assert result['x_p'] == pytest.approx(float(h.coordinates[0].mean()), abs=5e-4)
    assert result['y_p'] == pytest.approx(float(h.coordinates[1].mean()), abs=5e-4)
    if name == 'DEEstimator':
        exclude = np.zeros(hologram.shape, dtype=bool)
        exclude[:, hologram.shape[1] // 2:] = True
        est = _quick(cls(model=model, seed=0, exclude=exclude))
        assert isinstance(est.estimate(hologram), pd.Series)


def test_single_precision_hologram(api):
    """double_precision = False (theory/cupyLorenzMie.py:46-65): the
    hologram in FP32 arithmetic, within 1e-4 + 2e-5 I of the FP64 one and
    of the oracle; a non-Cartesian hologram stays FP64."""
    from oracle import lorenzmie as ref
    theory, _ = api
    p = theory.Sphere(x_p=64.3, y_p=63.6, z_p=120., a_p=1.2, n_p=1.5, k_p=0.001)
    m = theory.LorenzMie(coordinates=theory.LorenzMie.meshgrid((128, 128)),
                         particle=p)
    h64 = m.hologram()
    m.double_precision = False
    h32 = m.hologram()
    assert not np.array_equal(h32, h64)
    assert np.all(np.abs(h32 - h64) <= 1e-4 + 2e-5 * h64)
    want = ref.hologram(ref.meshgrid((128, 128)),
                        dict(x_p=64.3, y_p=63.6, z_p=120., a_p=1.2, n_p=1.5,
                             k_p=0.001), n_m=1.34, wavelength=0.447,
                        magnification=0.048)
    assert np.all(np.abs(h32 - want) <= 1e-4 + 2e-5 * want)
    np.testing.assert_array_equal(m.hologram(cartesian=False),
                                  theory.LorenzMie(coordinates=m.coordinates,
                                                   particle=p).hologram(cartesian=False))


def test_scattered_field_and_lorenzmie_hooks(api):
    """LorenzMie.scattered_field (theory/LorenzMie.py:135-158: one particle,
    r_0 included, the phase exp(-i k z_p), no alpha) and the per-order sum
    lorenzmie(ab, kdr) (:183-342) against the oracle's, for every flag.
    lorenzmie's displacements reach the kernel as float32 offsets from a
    rounded origin, whence its 1e-5 bound."""
    from oracle import lorenzmie as ref
    theory, _ = api
    c = theory.LorenzMie.meshgrid((48, 40))
    m = theory.LorenzMie(coordinates=c, alpha=0.7)
    own = m.particle
    p = theory.Sphere(x_p=20.3, y_p=23.8, z_p=110., a_p=0.9, n_p=1.5, k_p=0.002)
    p.r_0 = [1.5, -2., 10.]
    k = ref.wavenumber()
    ab = ref.mie_coefficients(0.9, 1.5, 0.002, 1.34, 0.447)
    for cart in (True, False):
        for boh in (True, False):
            got = m.scattered_field(p, cartesian=cart, bohren=boh)
            want = ref.scattered_field(c, 20.3, 23.8, 110., ab, k, (1.5, -2., 10.),
                                       cart, boh)
            assert np.abs(got - want).max() < 1e-6 * np.abs(want).max(), (cart, boh)
    assert m.alpha == 0.7 and m.particle is own
    rng = np.random.default_rng(8)
    n = 600
    kdr = k * np.vstack([rng.uniform(-30., 30., n), rng.uniform(-30., 30., n),
                         np.concatenate([rng.uniform(-200., -20., n // 2),
                                         rng.uniform(20., 150., n - n // 2)])])
    for cart in (True, False):
        for boh in (True, False):
            got = m.lorenzmie(ab, kdr, cartesian=cart, bohren=boh)
            want = ref.lorenzmie(ab, kdr, cart, boh)
            assert got.shape == (3, n)
            assert np.abs(got - want).max() < 1e-5 * np.abs(want).max(), (cart, boh)


def test_aberrated_scattered_field(api):
    """scattered_field of an aberrated model carries the particle's mask
    (theory/Aberrated.py:78-84)."""
    from oracle import lorenzmie as ref
    theory, _ = api
    c = theory.LorenzMie.meshgrid((40, 44))
    m = theory.AberratedLorenzMie(coordinates=c, spherical=0.7, alpha=0.8)
    p = theory.Sphere(x_p=21.3, y_p=19.6, z_p=95., a_p=0.7, n_p=1.55)
    k = ref.wavenumber()
    ab = ref.mie_coefficients(0.7, 1.55, 0., 1.34, 0.447)
    want = ref.scattered_field(c, 21.3, 19.6, 95., ab, k) * \
        ref.aberration(c, np.array([21.3, 19.6, 95.]), k, 1.34, 0.7)
    got = m.scattered_field(p)
    assert np.abs(got - want).max() < 1e-6 * np.abs(want).max()
    assert m.alpha == 0.8 and m.particle is not p
