"""Host-side interface, mirroring the reference's unit tests
(tests/test_lorenzmie.py, test_sphere.py, test_instrument.py,
test_optimizer.py:40-109, test_hologram.py, test_mask.py); nothing here
touches the GPU."""
import numpy as np
import pandas as pd
import pytest

from pylorenzmie_amd.theory import Instrument, LorenzMie, Particle, Sphere
from pylorenzmie_amd.theory.LorenzMie import dense_layout
from pylorenzmie_amd.analysis import Hologram, Mask, Optimizer, hipOptimizer


# ---------------------------------------------------------------- Instrument
def test_instrument_wavenumber():
    inst = Instrument()
    assert inst.wavenumber() == pytest.approx(2 * np.pi / 0.447 * 1.34 * 0.048)
    assert inst.wavenumber(in_medium=False) == pytest.approx(2 * np.pi / 0.447 * 0.048)
    assert inst.wavenumber(scaled=False) == pytest.approx(2 * np.pi / 0.447 * 1.34)
    assert list(inst.properties) == ['n_m', 'wavelength', 'magnification',
                                     'numerical_aperture', 'noise']


def test_instrument_pandas_roundtrip():
    a = Instrument(n_m=1.33)
    b = Instrument()
    b.from_pandas(a.to_pandas())
    assert b.n_m == 1.33 and a == b


# ---------------------------------------------------------------- Sphere
def test_sphere_defaults_and_properties():
    s = Sphere()
    assert (s.a_p, s.n_p, s.k_p) == (1., 1.5, 0.)
    assert list(s.properties) == ['x_p', 'y_p', 'z_p', 'a_p', 'n_p', 'k_p']
    s.d_p = 3.
    assert s.a_p == 1.5
    props = s.properties
    props['a_p'] += 0.25
    s.properties = props
    assert s.a_p == 1.75


def test_sphere_order_counts():
    assert Sphere.wiscombe_yang(1., 10. + 0.j) == 10
    assert Sphere.wiscombe_yang(10., 10. + 0.j) == 100
    assert Sphere.wiscombe_yang(4210., 10. + 0.j) == 42100
    assert Sphere.neves_pisignano(10.) == 11


def test_particle_sequence_protocol():
    p = Particle()
    assert len(p) == 1 and list(p) == [p] and p[0] is p
    with pytest.raises(IndexError):
        p[1]
    assert np.array_equal(Particle().ab(), np.ones((1, 2)))


# ---------------------------------------------------------------- LorenzMie
def test_lorenzmie_method_and_jac_params():
    m = LorenzMie()
    assert 'hip' in m.method and 'numpy' in m.method
    assert m.jac_params == frozenset({'x_p', 'y_p', 'z_p', 'a_p', 'n_p', 'k_p',
                                      'alpha'})


def test_coordinates_handling():
    m = LorenzMie()
    assert m.coordinates.shape == (3, 201 * 201)
    assert m._width == 201
    m.coordinates = np.array([1, 2, 3]).reshape(3, 1)
    assert np.allclose(m.coordinates, [[1], [2], [3]])
    m.coordinates = np.arange(256)
    assert m.coordinates.shape[0] == 3
    m.coordinates = [[1, 1, 1, 1], [2, 2, 2, 2], [3, 3, 3, 3]]
    assert np.allclose(m.coordinates, [[1] * 4, [2] * 4, [3] * 4])
    assert m._width == 0
    with pytest.raises(ValueError):
        m.coordinates = np.zeros((4, 10))


def test_properties_setter_routes():
    m = LorenzMie()
    m.properties = dict(x_p=-42, n_m=1.33, alpha=0.8, bogus=1)
    assert m.particle.x_p == -42 and m.instrument.n_m == 1.33 and m.alpha == 0.8
    s = m.to_json()
    m.particle.n_p = 1.4
    m.from_json(s)
    assert m.particle.n_p == 1.5


def test_dense_layout_detection():
    c = LorenzMie.meshgrid((30, 40), corner=(12., 7.))
    c3 = np.vstack([c, np.zeros((1, c.shape[1]))])
    assert dense_layout(c3) == (40, (12., 7., 0.))
    assert dense_layout(c3[:, ::3]) is None
    c3[2, 5] = 1.
    assert dense_layout(c3) is None


# ---------------------------------------------------------------- Hologram, Mask
def test_hologram_views_and_crop():
    data = np.random.rand(60, 80)
    h = Hologram(data, corner=(5., 9.))
    assert h.shape == (60, 80)
    assert h.flat_data.base is not None
    assert h.flat_coordinates.shape == (2, 4800)
    crop = h[10:20, 30:45]
    assert crop.corner == (35., 19.)
    assert np.array_equal(crop.data, data[10:20, 30:45])


def test_mask_fraction_and_exclude():
    np.random.seed(0)
    ex = np.zeros((50, 50), bool)
    ex[:10] = True
    m = Mask(shape=(50, 50), fraction=0.3, exclude=ex)
    assert not m()[:10].any()
    assert 0.15 < m().mean() < 0.3
    d, c = m.apply(Hologram(np.ones((50, 50))))
    assert d.size == c.shape[1]


# ---------------------------------------------------------------- Optimizer
def test_optimizer_configuration():
    opt = Optimizer(model=LorenzMie())
    assert opt.result is None and opt.mask is None
    assert opt.variables == ['x_p', 'y_p', 'z_p', 'a_p', 'n_p']
    assert 'noise' in opt.fixed and 'numerical_aperture' in opt.fixed
    assert 'alpha' in opt.fixed
    assert opt.settings['method'] == 'lm' and opt.settings['ftol'] == 1e-4
    opt.robust = True
    assert opt.robust and opt.settings['method'] == 'trf'
    opt.robust = False
    assert opt.settings['loss'] == 'linear'
    opt.variables = ['x_p', 'y_p', 'z_p', 'a_p']
    assert 'n_p' in opt.fixed
    assert isinstance(opt.metadata, pd.Series)
    with pytest.raises(RuntimeError):
        opt.report()
    fixed = ['wavelength']
    o2 = Optimizer(model=opt.model, fixed=fixed)
    fixed.append('n_m')
    assert 'n_m' not in o2.fixed
    opt.fraction = 0.2
    assert isinstance(opt.mask, Mask) and opt.fraction == 0.2


def test_optimizer_rejects_incompatible_model():
    class Other:
        method = 'numpy'
        properties = {}
    with pytest.raises(TypeError):
        Optimizer(model=Other())
    assert hipOptimizer.method in LorenzMie.method


# ---------------------------------------------------------------- Cluster, Dimer
def test_cluster_moves_constituents():
    """tests/test_cluster.py of the reference: sequence protocol and r_0."""
    from pylorenzmie_amd.theory import Cluster
    p1, p2 = Sphere(x_p=-1.), Sphere(x_p=2.)
    c = Cluster(x_p=10., y_p=20., z_p=30., particles=[p1, p2])
    assert len(c) == 2 and c[0] is p1 and list(c) == [p1, p2]
    with pytest.raises(IndexError):
        c[2]
    c.x_p, c.z_p = 11., 31.
    assert all(tuple(p.r_0) == (11., 20., 31.) for p in c)
    q = [Sphere(), Sphere(), Sphere()]
    c.particles = q
    assert len(c) == 3 and tuple(q[2].r_0) == (11., 20., 31.)
    assert set(c.properties) == {'x_p', 'y_p', 'z_p'}


def test_dimer_geometry_matches_reference(superposed_golden):
    """tests/test_dimer.py of the reference, and its member positions."""
    from pylorenzmie_amd.theory import Dimer
    g = superposed_golden
    x, y, z, a, n, k, theta, phi = g['dimer_params']
    d = Dimer(a_p=a, n_p=n, k_p=k, theta=theta, phi=phi, magnification=0.048)
    assert len(d) == 2 and all(isinstance(p, Sphere) for p in d)
    np.testing.assert_allclose([p.r_p for p in d.particles], g['dimer_member_rp'])
    d.a_p = 0.7
    assert all(p.a_p == 0.7 for p in d)
    assert np.allclose(d.particles[0].r_p, -d.particles[1].r_p)
    assert set(d.properties) == {'x_p', 'y_p', 'z_p', 'a_p', 'n_p', 'k_p',
                                 'theta', 'phi'}


def test_aberrated_class():
    from pylorenzmie_amd.theory import Aberrated, AberratedLorenzMie
    m = AberratedLorenzMie(spherical=0.5)
    assert isinstance(m, LorenzMie) and m.properties['spherical'] == 0.5
    m.properties = dict(spherical=-0.2, z_p=77.)
    assert m.spherical == -0.2 and m.particle.z_p == 77.
    assert Aberrated(LorenzMie).__qualname__ == 'AberratedLorenzMie'


def test_estimators_share_the_base_class():
    """tests/test_radialestimator.py:25-33 of the reference."""
    from pylorenzmie_amd.analysis import (BaseEstimator, DEEstimator,
                                          RadialEstimator)
    from pylorenzmie_amd.analysis.DEEstimator import DEFAULT_BOUNDS
    for cls in (DEEstimator, RadialEstimator):
        est = cls(model=LorenzMie())
        assert isinstance(est, BaseEstimator)
        assert est.bounds == DEFAULT_BOUNDS and est.bounds is not DEFAULT_BOUNDS


def test_lorenzmie_batch_without_particles():
    """theory/torchLorenzMieBatch.py:313-316 of the reference: nothing to
    scatter gives unit holograms; no device work is needed for that."""
    from pylorenzmie_amd.theory import LorenzMieBatch, hipLorenzMieBatch
    assert issubclass(LorenzMieBatch, LorenzMie) and hipLorenzMieBatch is LorenzMieBatch
    m = LorenzMieBatch(coordinates=LorenzMie.meshgrid((8, 6)))
    assert m.batch_hologram([]).shape == (0, 48)
    assert np.array_equal(m.batch_hologram([[], []]), np.ones((2, 48)))
