"""The reference's own unit tests of the host-side classes, case by case:
tests/test_hologram.py, test_mask.py, test_instrument.py,
test_particle.py, test_sphere.py, test_lorenzmie.py, test_optimizer.py,
test_feature.py, test_frame.py, test_cluster.py, test_dimer.py,
test_deestimator.py, test_radialestimator.py, test_radialmask.py,
test_baseestimator.py and test_lib.py (LMObject).  Each test names the
reference cases it covers.  Cases that compute on the GPU (Sphere.ab,
field, hologram, optimize, estimate, predicted, residuals) are in
tests/test_gpu_contracts.py and tests/test_gpu_crop.py; nothing here
touches the GPU."""
import numpy as np
import pandas as pd
import pytest

from pylorenzmie_amd.analysis import Feature, Frame, Hologram, Mask, Optimizer
from pylorenzmie_amd.lib import LMObject
from pylorenzmie_amd.theory import Instrument, LorenzMie, Particle, Sphere


# ------------------------------------------------------------ Hologram
@pytest.fixture
def holo():
    data = np.random.default_rng(0).random((50, 60))
    return data, Hologram(data)


def test_hologram_construction(holo):
    """test_shape, test_default_corner, test_custom_corner."""
    data, h = holo
    assert h.shape == (50, 60)
    assert h.corner == (0., 0.)
    h2 = Hologram(data, corner=(10., 20.))
    assert h2.coordinates[0, 0, 0] == pytest.approx(10.)
    assert h2.coordinates[1, 0, 0] == pytest.approx(20.)


def test_hologram_coordinates_and_flat_views(holo):
    """test_coordinates_shape, _origin, test_flat_data_shape, _is_view,
    test_flat_coordinates_shape, _is_view."""
    _, h = holo
    assert h.coordinates.shape == (2, 50, 60)
    assert h.coordinates[0, 0, 0] == 0. and h.coordinates[1, 0, 0] == 0.
    assert h.flat_data.shape == (3000,)
    assert np.shares_memory(h.flat_data, h.data)
    assert h.flat_coordinates.shape == (2, 3000)
    assert np.shares_memory(h.flat_coordinates, h.coordinates)


def test_hologram_getitem(holo):
    """test_getitem_shape, _corner, _coordinates_consistent,
    _corner_accumulates, _none_start."""
    data, h = holo
    crop = h[10:30, 15:45]
    assert crop.shape == (20, 30)
    assert crop.corner == (15., 10.)
    np.testing.assert_array_equal(crop.flat_coordinates,
                                  h.coordinates[:, 10:30, 15:45].reshape(2, -1))
    assert Hologram(data, corner=(5., 3.))[10:30, 15:45].corner == (20., 13.)
    crop = h[:30, :45]
    assert crop.corner == (0., 0.) and crop.shape == (30, 45)


def test_hologram_equality(holo):
    """test_equality, test_inequality_data, test_inequality_corner."""
    data, h = holo
    assert h == Hologram(data.copy())
    assert h != Hologram(data + 1.)
    assert h != Hologram(data, corner=(1., 0.))


# ------------------------------------------------------------ Mask
@pytest.fixture
def mask():
    np.random.seed(1)
    return Mask(shape=(128, 128))


@pytest.mark.parametrize('value', [0.2, 1.])
def test_mask_fraction(mask, value):
    """test_fraction, test_fraction_unity."""
    mask.fraction = value
    assert mask.fraction == value
    assert round(abs(value - mask().sum() / mask().size), 1) == 0


def test_mask_exclude_shape_update(mask):
    """test_exclude, test_shape_sets_mask, test_update_regenerates_mask,
    test_exclude_suppresses_pixels."""
    exclude = np.zeros((128, 128), bool)
    mask.exclude = exclude
    assert exclude.size == mask().size
    assert mask().shape == (128, 128)
    m1 = mask().copy()
    mask.update()
    assert not np.array_equal(m1, mask())
    mask.fraction = 1.
    mask.exclude = np.ones((128, 128), bool)
    assert mask().sum() == 0


def test_mask_properties_and_json(mask):
    """test_properties, test_json_roundtrip."""
    assert mask.properties['fraction'] == mask.fraction
    mask.fraction = 0.3
    s = mask.to_json()
    mask.fraction = 0.1
    mask.from_json(s)
    assert mask.fraction == pytest.approx(0.3)


def test_mask_apply(mask):
    """test_apply_returns_tuple, _data_shape, _coords_shape, _sets_shape,
    _selects_correct_pixels, _respects_exclude."""
    rng = np.random.default_rng(2)
    h = Hologram(rng.random((128, 128)))
    result = mask.apply(h)
    assert isinstance(result, tuple) and len(result) == 2
    data, coords = result
    assert data.ndim == 1 and data.shape[0] == coords.shape[1]
    assert coords.shape[0] == 2
    m = mask()
    np.testing.assert_array_equal(data, h.data[m])
    np.testing.assert_array_equal(coords, h.coordinates[:, m])
    mask.apply(Hologram(rng.random((64, 80))))
    assert mask.shape == (64, 80)
    exclude = np.zeros((128, 128), bool)
    exclude[0, :] = True
    mask.fraction = 1.
    mask.exclude = exclude
    _, coords = mask.apply(Hologram(rng.random((128, 128))))
    assert not np.any(coords[1] == 0.)


def test_mask_exclude_before_shape_change(mask):
    """test_exclude_before_shape_change_does_not_raise."""
    exclude = np.zeros((129, 128), bool)
    exclude[0, :] = True
    mask.exclude = exclude
    mask.fraction = 1.
    _, coords = mask.apply(Hologram(np.random.default_rng(3).random((129, 128))))
    assert not np.any(coords[1] == 0.)


# ------------------------------------------------------------ Instrument
def test_instrument_defaults_and_setters():
    """test_default_* and test_set_*."""
    inst = Instrument()
    assert inst.wavelength == pytest.approx(0.447)
    assert inst.magnification == pytest.approx(0.048)
    assert inst.numerical_aperture == pytest.approx(1.45)
    assert inst.noise == pytest.approx(0.05)
    assert inst.n_m == pytest.approx(1.340)
    for name, value in (('wavelength', 0.532), ('magnification', 0.120),
                        ('numerical_aperture', 1.20), ('noise', 0.1),
                        ('n_m', 1.34)):
        setattr(inst, name, value)
        assert getattr(inst, name) == pytest.approx(value)


def test_instrument_properties_serialization_repr():
    """test_properties_keys, _roundtrip, test_json_roundtrip,
    test_pandas_roundtrip, test_repr."""
    inst = Instrument()
    for key in ('wavelength', 'magnification', 'numerical_aperture', 'noise',
                'n_m'):
        assert key in inst.properties
    inst.n_m = 1.339
    b = Instrument()
    b.properties = inst.properties
    assert b.n_m == pytest.approx(1.339)
    inst.n_m = 1.341
    b = Instrument()
    b.from_json(inst.to_json())
    assert b.n_m == pytest.approx(1.341)
    b = Instrument()
    b.from_pandas(inst.to_pandas())
    assert b.n_m == pytest.approx(1.341)
    assert isinstance(repr(inst), str)


def test_instrument_wavenumber():
    """test_wavenumber_default, _in_vacuum, _unscaled, _vacuum_unscaled."""
    inst = Instrument()
    lam, nm, mag = inst.wavelength, inst.n_m, inst.magnification
    assert inst.wavenumber() == pytest.approx(2 * np.pi / lam * nm * mag)
    assert inst.wavenumber(in_medium=False) == pytest.approx(2 * np.pi / lam * mag)
    assert inst.wavenumber(scaled=False) == pytest.approx(2 * np.pi / lam * nm)
    assert inst.wavenumber(in_medium=False, scaled=False) == \
        pytest.approx(2 * np.pi / lam)


# ------------------------------------------------------------ Particle
def test_particle_positions():
    """test_default_values, test_r_p_setter/_getter, test_r_0_setter/_getter."""
    p = Particle()
    assert (p.x_p, p.y_p, p.z_p, p.x_0, p.y_0, p.z_0) == (0., 0., 100., 0., 0., 0.)
    p.r_p = [10., 20., 30.]
    assert (p.x_p, p.y_p, p.z_p) == (10., 20., 30.)
    p.x_p, p.y_p, p.z_p = 1., 2., 3.
    np.testing.assert_array_equal(p.r_p, [1., 2., 3.])
    p.r_0 = [1., 2., 3.]
    assert (p.x_0, p.y_0, p.z_0) == (1., 2., 3.)
    p.x_0, p.y_0, p.z_0 = 4., 5., 6.
    np.testing.assert_array_equal(p.r_0, [4., 5., 6.])


def test_particle_properties_and_serialization():
    """test_properties_getter/_setter, test_to_json_roundtrip,
    test_to_pandas_roundtrip."""
    p = Particle()
    p.x_p, p.y_p, p.z_p = 5., 6., 7.
    assert p.properties == {'x_p': 5., 'y_p': 6., 'z_p': 7.}
    p.properties = {'x_p': 9., 'y_p': 8., 'z_p': 7.}
    assert (p.x_p, p.y_p, p.z_p) == (9., 8., 7.)
    p.r_p = [1., 2., 3.]
    b = Particle()
    b.from_json(p.to_json())
    assert (b.x_p, b.y_p, b.z_p) == (1., 2., 3.)
    b = Particle()
    b.from_pandas(p.to_pandas())
    assert (b.x_p, b.y_p, b.z_p) == (1., 2., 3.)


def test_particle_ab_and_sequence():
    """test_ab_shape, test_ab_values, test_ab_accepts_parameters, test_len,
    test_iter, test_iter_reentrant, test_getitem, _out_of_range, test_repr."""
    p = Particle()
    ab = p.ab()
    assert ab.shape == (1, 2) and ab.dtype == complex
    np.testing.assert_array_equal(ab, np.ones((1, 2), complex))
    assert p.ab(n_m=1.33 + 0.01j, wavelength=0.447).shape == (1, 2)
    assert len(p) == 1
    assert list(p) == [p] and list(p)[0] is p
    assert list(p) == list(p)
    assert p[0] is p
    with pytest.raises(IndexError):
        p[1]
    assert isinstance(repr(p), str)


# ------------------------------------------------------------ Sphere (host)
def test_sphere_host_cases():
    """test_default_values, test_d_p_getter/_setter, test_repr,
    test_properties_keys/_roundtrip, test_wiscombe_yang_*,
    test_neves_pisignano_*."""
    s = Sphere()
    assert (s.a_p, s.n_p, s.k_p) == (1., 1.5, 0.)
    s = Sphere(a_p=1., n_p=1.4, k_p=0.)
    assert s.d_p == pytest.approx(2. * s.a_p)
    s.d_p = 3.
    assert s.a_p == pytest.approx(1.5)
    assert isinstance(repr(s), str)
    props = s.properties
    for key in ('x_p', 'y_p', 'z_p', 'a_p', 'n_p', 'k_p'):
        assert key in props
    props['a_p'] += 0.25
    s.properties = props
    assert s.a_p == props['a_p']
    assert Sphere.wiscombe_yang(1., 10. + 0.j) == 10
    assert Sphere.wiscombe_yang(10., 10. + 0.j) == 100
    assert Sphere.wiscombe_yang(4210., 10. + 0.j) == 42100
    assert isinstance(Sphere.neves_pisignano(10.), int)
    assert Sphere.neves_pisignano(10.) == 11


# ------------------------------------------------------------ LorenzMie (host)
def test_lorenzmie_host_cases():
    """test_repr, test_meshgrid, test_coordinates_none, _point, _1d, _2d,
    _3dlist, test_properties, test_serialize."""
    m = LorenzMie()
    assert isinstance(repr(m), str)
    xy = m.meshgrid((32, 24), flatten=False)
    assert xy.shape == (2, 32, 24)
    assert m.meshgrid((32, 24), flatten=True).shape == (2, 32 * 24)
    m.coordinates = None
    assert isinstance(m.coordinates, np.ndarray)
    point = np.array([1, 2, 3]).reshape((3, 1))
    m.coordinates = point
    assert np.allclose(m.coordinates, point)
    m.coordinates = np.arange(256)
    assert m.coordinates.shape[0] == 3
    c = m.meshgrid((256, 256))
    m.coordinates = c
    assert m.coordinates.shape[0] == 3 and np.allclose(m.coordinates[0:2], c)
    c = [[1, 1, 1, 1], [2, 2, 2, 2], [3, 3, 3, 3]]
    m.coordinates = c
    assert np.allclose(m.coordinates, np.array(c))
    m.properties = dict(x_p=-42)
    assert m.particle.x_p == -42
    m.particle.n_p = 1.5
    s = m.to_json()
    m.particle.n_p = 1.4
    m.from_json(s)
    assert m.particle.n_p == 1.5


# ------------------------------------------------------------ Optimizer
@pytest.fixture
def optimizer():
    shape = (390, 390)
    model = LorenzMie(coordinates=LMObject.meshgrid(shape))
    model.instrument.wavelength = 0.447
    model.instrument.magnification = 0.048
    model.instrument.n_m = 1.34
    model.particle.r_p = [shape[0] // 2, shape[1] // 2, 330]
    model.particle.a_p = 1.1
    model.particle.n_p = 1.4
    return Optimizer(model=model)


def test_optimizer_configuration_cases(optimizer):
    """test_result_none, test_mask_default_none, test_mask_set,
    test_metadata, test_properties, test_default_fixed_includes_noise,
    _numerical_aperture, test_fixed_list_is_copied,
    test_robust_false_uses_lm, test_robust_true_uses_trf,
    test_report_before_optimize_raises."""
    opt = optimizer
    assert opt.result is None
    assert opt.mask is None
    m = Mask(fraction=0.1)
    assert Optimizer(model=opt.model, mask=m).mask is m
    assert isinstance(opt.metadata, pd.Series)
    props = opt.properties
    opt.properties = props
    assert 'settings' in props
    assert 'noise' in opt.fixed and 'numerical_aperture' in opt.fixed
    fixed = ['wavelength']
    o2 = Optimizer(model=opt.model, fixed=fixed)
    fixed.append('n_m')
    assert 'n_m' not in o2.fixed
    opt.robust = False
    assert not opt.robust and opt.settings['method'] == 'lm'
    opt.robust = True
    assert opt.robust and opt.settings['method'] == 'trf'
    with pytest.raises(RuntimeError):
        opt.report()


def test_optimizer_variables_and_fixed(optimizer):
    """test_variables_setter_updates_fixed,
    test_fixed_setter_updates_variables."""
    opt = optimizer
    variables = ['x_p', 'y_p', 'z_p', 'a_p']
    opt.variables = variables
    for v in variables:
        assert v in opt.variables and v not in opt.fixed
    for f in opt.fixed:
        assert f not in variables
    fixed = list(opt.model.properties)[:2]
    opt.fixed = fixed
    for f in fixed:
        assert f in opt.fixed and f not in opt.variables


# ------------------------------------------------------------ Feature (host)
@pytest.fixture
def feature():
    data = np.random.default_rng(4).random((64, 80)) + 0.5
    model = LorenzMie()
    model.instrument.wavelength = 0.447
    model.instrument.magnification = 0.048
    model.instrument.n_m = 1.34
    h, w = data.shape
    model.particle.r_p = [w / 2., h / 2., 330.]
    model.particle.a_p = 1.1
    model.particle.n_p = 1.4
    f = Feature(Hologram(data), model=model)
    f.mask.fraction = 0.1
    return f


def test_feature_host_cases(feature):
    """test_exclude_shared, test_is_hologram, test_model,
    test_getsetparticle."""
    f = feature
    assert f.estimator.exclude is f.optimizer.mask.exclude
    assert isinstance(f, Hologram)
    model = LorenzMie()
    f.model = model
    assert f.model is model
    p = f.particle
    p.z_p = 200.
    f.particle = p
    assert f.particle.z_p == 200.


# ------------------------------------------------------------ Frame (host)
def test_frame_host_cases(frame_golden):
    """test_is_hologram, test_data, test_shape_from_data,
    test_shape_no_data, test_bboxes, test_bboxes_propagate_instrument,
    test_detect_no_data, test_getitem_returns_feature, test_getitem_corner
    (on a crop of the reference's tests/data/image0010.png)."""
    data = frame_golden['image'].astype(float) / 100.
    frame = Frame()
    assert isinstance(frame, Hologram)
    assert frame.shape == (0, 0)
    frame.data = data
    assert data.size == frame.coordinates.size // 2
    assert frame.shape == data.shape
    frame.bboxes = ((100, 200), 30, 40)
    assert len(frame.features) == 1
    frame = Frame()
    frame.instrument.wavelength = 0.447
    frame.data = data
    frame.bboxes = ((100, 200), 100, 100)
    assert frame.features[0].model.instrument.wavelength == pytest.approx(0.447)
    crop = frame[50:150, 50:150]
    assert isinstance(crop, Feature)
    assert crop.corner == (50., 50.)
    frame.data = None
    assert frame.detect() == 0


# ------------------------------------------------------------ Cluster, Dimer
def test_cluster_cases():
    """tests/test_cluster.py: test_len, test_getitem, test_iteration,
    test_update_positions, test_setattr_particles."""
    from pylorenzmie_amd.theory import Cluster
    p1 = Particle(x_p=10., y_p=20., z_p=30.)
    p2 = Particle(x_p=40., y_p=50., z_p=60.)
    c = Cluster(particles=[p1, p2])
    assert len(c) == 2
    assert c[0] == p1 and c[1] == p2
    with pytest.raises(IndexError):
        c[2]
    assert list(c) == [p1, p2]
    c.x_p, c.y_p, c.z_p = 100., 200., 300.
    for p in c:
        assert tuple(p.r_0) == (100., 200., 300.)
    new = [Particle(x_p=1., y_p=2., z_p=3.), Particle(x_p=4., y_p=5., z_p=6.)]
    c.particles = new
    assert len(c) == 2 and c[0] == new[0] and c[1] == new[1]


def test_dimer_cases():
    """tests/test_dimer.py: test_initialization, test_update_positions,
    test_setattr_particles, test_invalid_particle_count."""
    from pylorenzmie_amd.theory import Dimer
    d = Dimer()
    assert len(d) == 2
    for p in d:
        assert isinstance(p, Sphere)
        assert (p.a_p, p.n_p, p.k_p) == (d.a_p, d.n_p, d.k_p)
    a_p, theta, phi, mag = 1.0, np.pi / 3, np.pi / 4, 0.05
    d.a_p, d.theta, d.phi, d.magnification = a_p, theta, phi, mag
    r_p = (a_p / mag) * np.array([np.cos(theta) * np.cos(phi),
                                  np.cos(theta) * np.sin(phi), np.sin(theta)])
    assert np.allclose(d.particles[0].r_p, r_p)
    assert np.allclose(d.particles[1].r_p, -r_p)
    d.a_p, d.n_p, d.k_p = 1.5, 1.55, 0.1
    for p in d:
        assert (p.a_p, p.n_p, p.k_p) == (1.5, 1.55, 0.1)
    d.particles = [Sphere(a_p=0.75, n_p=1.45, k_p=0.)]
    d.update_positions()
    assert len(d) == 1


# ------------------------------------------------------------ estimators (host)
@pytest.mark.parametrize('name', ['DEEstimator', 'RadialEstimator'])
def test_estimator_host_cases(name):
    """tests/test_deestimator.py and test_radialestimator.py:
    test_is_base_estimator, test_default_bounds_keys,
    test_default_bounds_valid_ranges, test_default_bounds_match_de,
    test_properties_keys, test_exclude_none_by_default."""
    from pylorenzmie_amd import analysis
    from pylorenzmie_amd.analysis.DEEstimator import DEFAULT_BOUNDS
    model = LorenzMie(coordinates=LMObject.meshgrid((101, 101)))
    est = getattr(analysis, name)(model=model, seed=0)
    assert isinstance(est, analysis.BaseEstimator)
    for key in ('z_p', 'a_p', 'n_p'):
        assert key in DEFAULT_BOUNDS and key in est.bounds
    for key, (lo, hi) in DEFAULT_BOUNDS.items():
        assert lo < hi, key
    assert est.bounds == DEFAULT_BOUNDS
    keys = ('popsize', 'bounds', 'settings')
    if name == 'DEEstimator':
        keys += ('fraction',)
        assert est.exclude is None
    for key in keys:
        assert key in est.properties


# ------------------------------------------------------------ RadialMask
def test_radial_mask_cases():
    """tests/test_radialmask.py: test_output_shape, _dtype,
    test_non_square_shape, test_center_excluded_at_low_fraction,
    test_center_selected_at_high_fraction, test_update_regenerates,
    test_inherits_exclude, test_any_pixels_selected."""
    from pylorenzmie_amd.analysis import RadialMask
    np.random.seed(2)
    shape = (101, 121)
    h, w = shape
    m = RadialMask(shape=shape, fraction=0.5)
    assert m().shape == shape and m().dtype == bool
    assert RadialMask(shape=(61, 101), fraction=0.3)().shape == (61, 101)
    m = RadialMask(shape=shape, fraction=0.2)
    hits = 0
    for _ in range(20):
        m.update()
        hits += int(m()[h // 2, w // 2])
    assert hits == 0
    m = RadialMask(shape=shape, fraction=0.8)
    hits = 0
    for _ in range(20):
        m.update()
        hits += int(m()[h // 2, w // 2])
    assert hits > 0
    m = RadialMask(shape=shape, fraction=0.3)
    m.update()
    m1 = m().copy()
    m.update()
    assert not np.array_equal(m1, m())
    exclude = np.zeros(shape, dtype=bool)
    exclude[h // 2, w // 2] = True
    assert not RadialMask(shape=shape, fraction=0.9, exclude=exclude)()[h // 2, w // 2]
    assert RadialMask(shape=shape, fraction=0.3)().any()


@pytest.mark.parametrize('fraction', [0.15, 0.35, 0.6, 0.85])
def test_radial_mask_probability_law(fraction):
    """Each pixel's selection rate over many draws is min(1, hypot(a + b u,
    a + b v)) with u, v in [-1, 1] (analysis/RadialMask.py:21-33): the
    centre-to-edge law for either branch of (a, b), which a branch taken
    for the wrong fraction or unscaled positions would miss."""
    from pylorenzmie_amd.analysis import RadialMask
    np.random.seed(3)
    shape = (9, 13)
    m = RadialMask(shape=shape, fraction=fraction)
    rate = np.zeros(shape)
    n = 4000
    for _ in range(n):
        m.update()
        rate += m()
    rate /= n
    f = fraction
    a, b = (0., 2. * f) if f < 0.5 else (2. * f - 1., 2. * (1. - f))
    y, x = np.mgrid[:shape[0], :shape[1]]
    u = 2. * x / (shape[1] - 1.) - 1.
    v = 2. * y / (shape[0] - 1.) - 1.
    want = np.minimum(1., np.hypot(a + b * v, a + b * u))
    assert np.abs(rate - want).max() < 0.035
    wrong = np.minimum(1., np.hypot(2. * f * v, 2. * f * u)) if f >= 0.5 else \
        np.minimum(1., np.hypot(2. * f - 1. + 2. * (1. - f) * v,
                                2. * f - 1. + 2. * (1. - f) * u))
    assert np.abs(rate - wrong).max() > 0.1


def test_optimizer_takes_a_radial_mask():
    """A RadialMask plugs into the Optimizer as the reference's Mask does
    (analysis/Optimizer.py:211-215): the pixels it keeps are the fit's."""
    from pylorenzmie_amd.analysis import RadialMask
    np.random.seed(4)
    model = LorenzMie(coordinates=LMObject.meshgrid((40, 50)))
    opt = Optimizer(model=model, mask=RadialMask(fraction=0.3))
    h = Hologram(np.random.default_rng(5).random((40, 50)))
    opt._select(h)
    keep = opt.mask()
    assert keep.shape == (40, 50) and opt._count == keep.sum()
    np.testing.assert_array_equal(opt._data, h.data[keep])


def test_double_precision_switch():
    """tests/test_cupylorenzmie.py: test_doubleprecision (the switch of
    theory/cupyLorenzMie.py:46-65)."""
    m = LorenzMie()
    assert m.double_precision
    m.double_precision = False
    assert not m.double_precision
    m.double_precision = True
    assert m.double_precision
    assert not LorenzMie(double_precision=False).double_precision


def test_to_json_encodes_numpy_scalars():
    """lib/LMObject.py:110-127: numpy scalars in the properties serialize
    as numbers."""
    s = Sphere(a_p=np.float32(0.5), n_p=np.float64(1.4))
    s.x_p = np.int64(3)
    b = Sphere()
    b.from_json(s.to_json())
    assert (b.a_p, b.n_p, b.x_p) == (0.5, 1.4, 3)
    with pytest.raises(TypeError):
        Sphere(a_p=object()).to_json()


def test_lmobject_equality():
    """tests/test_lib.py:51-64: test_equal_same_class,
    test_not_equal_same_class, test_not_equal_different_class."""
    assert Sphere(a_p=1.0, n_p=1.5) == Sphere(a_p=1.0, n_p=1.5)
    assert Sphere(a_p=1.0, n_p=1.5) != Sphere(a_p=2.0, n_p=1.5)
    assert Sphere() != Instrument()


def test_base_estimator_contract():
    """tests/test_baseestimator.py: test_cannot_instantiate_abstract,
    test_concrete_subclass_instantiates, test_estimate_returns_series,
    test_missing_estimate_raises, test_missing_properties_raises."""
    from pylorenzmie_amd.analysis import BaseEstimator

    class Concrete(BaseEstimator):
        @BaseEstimator.properties.getter
        def properties(self):
            return dict()

        def estimate(self, hologram):
            return pd.Series({'x_p': 0., 'y_p': 0.})

    with pytest.raises(TypeError):
        BaseEstimator()
    est = Concrete()
    assert isinstance(est, BaseEstimator)
    assert isinstance(est.estimate(Hologram(np.random.rand(50, 50))), pd.Series)

    class NoEstimate(BaseEstimator):
        @BaseEstimator.properties.getter
        def properties(self):
            return dict()

    class NoProperties(BaseEstimator):
        def estimate(self, hologram):
            return pd.Series(dtype=float)

    with pytest.raises(TypeError):
        NoEstimate()
    with pytest.raises(TypeError):
        NoProperties()
