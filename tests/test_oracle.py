"""Pins of the FP64 oracle (CPU): the reference's known-answer tests and
golden vectors, textbook Mie coefficients in mpmath, an independent
literal B&H multipole sum, invariants and closed forms.  A dropped term,
a wrong sign or a transposed index in oracle/lorenzmie.py fails one of
these."""
import numpy as np
import pytest

from conftest import INSTRUMENT
from oracle import independent, lorenzmie as ref


# ---- known answers from the reference's own tests ------------------------
def test_wiscombe_yang_known_answers():
    # tests/test_sphere.py:169-176
    assert ref.wiscombe_yang(1., 10. + 0.j) == 10
    assert ref.wiscombe_yang(10., 10. + 0.j) == 100
    assert ref.wiscombe_yang(4210., 10. + 0.j) == 42100


def test_neves_pisignano_known_answer():
    # tests/test_sphere.py:180-185
    assert ref.neves_pisignano(10.) == 11
    assert isinstance(ref.neves_pisignano(10.), int)


def test_wavenumber_closed_forms():
    # tests/test_instrument.py:81-103
    lam, nm, mag = 0.447, 1.34, 0.048
    assert ref.wavenumber(nm, lam, mag) == pytest.approx(2 * np.pi / lam * nm * mag)
    assert ref.wavenumber(nm, lam, mag, in_medium=False) == \
        pytest.approx(2 * np.pi / lam * mag)
    assert ref.wavenumber(nm, lam, mag, scaled=False) == \
        pytest.approx(2 * np.pi / lam * nm)
    assert ref.wavenumber(nm, lam, mag, False, False) == pytest.approx(2 * np.pi / lam)
    assert ref.wavenumber() == pytest.approx(0.904104, abs=1e-6)


def test_ab_zeroth_order_and_dtype():
    # tests/test_sphere.py:159-165
    ab = ref.mie_coefficients(1., 1.4, 0., 1.34, 0.447)
    assert ab.dtype == np.complex128
    np.testing.assert_array_equal(ab[0], 0j)


def test_ab_docs_spot_values():
    """Values printed in docs/tutorials/Theory.ipynb for Sphere(a_p=1,
    n_p=1.5) at n_m = 1.34, wavelength 0.447 (8 printed digits)."""
    ab = ref.mie_coefficients(1., 1.5, 0., 1.34, 0.447)
    spot = np.array([[0.5568827 + 0.49675394j, 0.6651625 + 0.47193378j],
                     [0.67591265 + 0.46803299j, 0.56705786 + 0.49548296j],
                     [0.59026502 + 0.49178486j, 0.69229663 + 0.46154323j]])
    assert np.abs(ab[1:4] - spot).max() < 1e-7


# ---- golden vectors from the reference numpy path -------------------------
def test_golden_ab(forward_golden):
    g = forward_golden
    offs = g['ab_offsets']
    for i, (a, n, k) in enumerate(g['ab_cases']):
        ab = ref.mie_coefficients(a, n, k, 1.34, 0.447)
        gold = g['ab_values'][offs[i]:offs[i + 1]]
        assert ab.shape == gold.shape
        assert np.abs(ab - gold).max() < 1e-11


def test_golden_fields(forward_golden):
    g = forward_golden
    p = dict(zip(ref.PARAMS, g['field64_params']))
    c = ref.meshgrid((64, 64))
    for cart in (1, 0):
        for boh in (1, 0):
            f = ref.field(c, p, cartesian=bool(cart), bohren=bool(boh), **INSTRUMENT)
            assert np.abs(f - g[f'field64_c{cart}_b{boh}']).max() < 1e-12


def test_golden_holograms(forward_golden):
    g = forward_golden
    h = ref.hologram(ref.meshgrid((201, 201)),
                     dict(zip(ref.PARAMS, g['cfg0_params'])), **INSTRUMENT)
    assert np.abs(h - g['cfg0_holo']).max() < 1e-12
    offs = g['holo_offsets']
    for i, p in enumerate(g['holo_cases']):
        c = g['holo_coords'][:, offs[i]:offs[i + 1]]
        h = ref.hologram(c, dict(zip(ref.PARAMS, p)), **INSTRUMENT)
        assert np.abs(h - g['holo_values'][offs[i]:offs[i + 1]]).max() < 1e-12
    pair = [dict(zip(ref.PARAMS, q)) for q in g['pair_params']]
    h = ref.hologram(ref.meshgrid((48, 48)), pair, **INSTRUMENT)
    assert np.abs(h - g['pair_holo']).max() < 1e-12


# ---- textbook Mie coefficients (mpmath) ------------------------------------
@pytest.mark.parametrize('a_p,n_p,k_p', [(0.75, 1.45, 0.), (1.5, 1.59, 0.001),
                                         (0.3, 2.0, 0.05)])
def test_ab_match_bohren_huffman_mpmath(a_p, n_p, k_p):
    """The recursion is exact B&H 4.53 once D1 starts deep enough; the
    reference's own start at D1[n_max] = 0 (theory/Sphere.py:190) perturbs
    every order (1e-7 to 1e-4 here), which the golden vectors pin."""
    ab = ref.mie_coefficients(a_p, n_p, k_p, 1.34, 0.447)
    nmax = ab.shape[0] - 1
    deep = ref.mie_coefficients(a_p, n_p, k_p, 1.34, 0.447, nmax=nmax + 40)
    bh = independent.mie_ab_mpmath(a_p, n_p, k_p, 1.34, 0.447, nmax)
    scale = np.abs(bh).max()
    assert np.abs(deep[:nmax + 1] - bh).max() < 1e-11 * scale
    assert np.abs(ab - bh).max() < 1e-3 * scale


def test_lossless_invariant():
    """|a_n - 1/2| = |b_n - 1/2| = 1/2 for a real index (deep D1 start)."""
    n = ref.mie_coefficients(2.1, 1.55, 0., 1.34, 0.447).shape[0] - 1
    ab = ref.mie_coefficients(2.1, 1.55, 0., 1.34, 0.447, nmax=n + 40)[:n + 1]
    assert np.abs(np.abs(ab[1:] - 0.5) - 0.5).max() < 1e-12


def test_layered_reduces_to_homogeneous():
    ab = ref.mie_coefficients(1.2, 1.5, 0.01, 1.34, 0.447)
    for radii in ([1.2], [0.5, 1.2], [0.3, 0.8, 1.2]):
        n = len(radii)
        abl = ref.mie_coefficients_layered(radii, [1.5] * n, [0.01] * n, 1.34, 0.447)
        assert np.abs(abl - ab).max() < 1e-12


@pytest.mark.parametrize('core,shell', [
    ((0.4, 1.60, 0.0), (0.9, 1.40, 0.0)),       # dense core, thin shell
    ((0.7, 1.38, 0.002), (1.1, 1.95, 0.01)),    # high-index absorbing shell
    ((0.2, 2.00, 0.0), (0.35, 1.33, 0.0))])     # small, index-matched shell
def test_layered_matches_coated_sphere_closed_form(core, shell):
    """Yang's recursion for two layers against B&H Sec. 8.1 in mpmath; D1
    started 40 orders deep so only the recursion itself is compared."""
    radii = [core[0], shell[0]]
    nl, kl = [core[1], shell[1]], [core[2], shell[2]]
    ab = ref.mie_coefficients_layered(radii, nl, kl, 1.34, 0.447)
    nmax = ab.shape[0] - 1
    deep = ref.mie_coefficients_layered(radii, nl, kl, 1.34, 0.447,
                                        nmax=nmax + 40)[:nmax + 1]
    bh = independent.coated_ab_mpmath(core[0], shell[0], core[1], core[2],
                                      shell[1], shell[2], 1.34, 0.447, nmax)
    scale = np.abs(bh).max()
    assert np.abs(deep - bh).max() < 1e-10 * scale
    # a swapped layer order or a dropped layer fails by far
    swap = ref.mie_coefficients_layered(radii, nl[::-1], kl[::-1], 1.34, 0.447,
                                        nmax=nmax + 40)[:nmax + 1]
    assert np.abs(swap - bh).max() > 1e-3 * scale


# ---- independent B&H multipole sum -----------------------------------------
@pytest.mark.parametrize('p', [(12.3, 11.6, 90., 0.8, 1.5, 0.),
                               (10.0, 10.0, 60., 1.3, 1.42, 0.002)])
def test_field_matches_literal_bh_sum(p):
    c = ref.meshgrid((24, 24))
    k = ref.wavenumber()
    ab = ref.mie_coefficients(p[3], p[4], p[5], 1.34, 0.447)
    f = ref.field(c, dict(zip(ref.PARAMS, p)), **INSTRUMENT)
    g = independent.field_bh(c, p[:3], ab, k)
    assert np.abs(f - g).max() < 1e-10 * np.abs(g).max()


def test_null_coefficients_and_alpha_zero():
    c = ref.meshgrid((16, 16))
    p = dict(x_p=8., y_p=8., z_p=50., a_p=1., n_p=1.5, k_p=0.)
    zero = dict(p, ab=np.zeros((20, 2), complex))
    assert np.all(ref.hologram(c, zero, **INSTRUMENT) == 1.)
    assert np.all(ref.hologram(c, p, alpha=0., **INSTRUMENT) == 1.)


def test_alpha_closed_form():
    c = ref.meshgrid((20, 20))
    p = dict(x_p=10.2, y_p=9.7, z_p=70., a_p=0.9, n_p=1.5, k_p=0.)
    e = ref.field(c, p, **INSTRUMENT)
    for a in (0.5, 1.3):
        expect = 1 + 2 * a * e[0].real + a * a * np.sum(np.abs(e) ** 2, 0)
        assert np.abs(ref.hologram(c, p, alpha=a, **INSTRUMENT) - expect).max() < 1e-12


def test_mirror_symmetries():
    c = ref.meshgrid((31, 31))
    h = ref.hologram(c, dict(x_p=15., y_p=15., z_p=80., a_p=1., n_p=1.5, k_p=0.),
                     **INSTRUMENT).reshape(31, 31)
    assert np.abs(h - h[:, ::-1]).max() < 1e-12
    assert np.abs(h - h[::-1, :]).max() < 1e-12


def test_small_sphere_dipole_limit():
    """x << 1: the field is the dipole a_1 term; a_1 ~ -(2i/3)(m^2-1)/(m^2+2) x^3."""
    a_p, n_p = 0.002, 1.5
    ab = ref.mie_coefficients(a_p, n_p, 0., 1.34, 0.447, nmax=20)
    x = ref.size_parameter(a_p)
    m = n_p / 1.34
    a1 = -2j / 3 * (m * m - 1) / (m * m + 2) * x ** 3
    assert abs(ab[1, 0] - a1) < 2e-3 * abs(a1)
    assert abs(ab[1, 1]) < 1e-2 * abs(ab[1, 0])


def test_jacobian_stencils_agree():
    """5- and 7-point stencils agree; a dropped or wrong-signed term would
    not survive both orders and the loss-gradient FD in the GPU tests."""
    c = ref.meshgrid((16, 16))
    p = dict(x_p=8.3, y_p=7.6, z_p=60., a_p=0.8, n_p=1.45, k_p=0.002, alpha=0.95)
    j4 = ref.jacobian(c, p, INSTRUMENT, order=4)
    j6 = ref.jacobian(c, p, INSTRUMENT, order=6)
    for i in range(7):
        assert np.abs(j4[i] - j6[i]).max() < 1e-6 * np.abs(j6[i]).max()
    # alpha column in closed form: dI/dalpha = 2 Re(E_x) + 2 alpha |E|^2
    e = ref.field(c, p, **INSTRUMENT)
    da = 2 * e[0].real + 2 * 0.95 * np.sum(np.abs(e) ** 2, 0)
    assert np.abs(j6[6] - da).max() < 1e-8 * np.abs(da).max()



# ---- superposition and aberration (reference golden) ------------------------
def _members(g):
    centre = g['cluster_centre']
    return [dict(zip(ref.PARAMS[:6], m), r_0=centre) for m in g['cluster_members']]


def test_aberrated_and_cluster_golden(superposed_golden):
    g = superposed_golden
    c = g['coords']
    p = dict(zip(ref.PARAMS[:6], g['aberrated_particle']))
    for tag, s in (('s09', 0.9), ('sm04', -0.4)):
        h = ref.hologram(c, p, spherical=s, **INSTRUMENT)
        assert np.abs(h - g['aberrated_' + tag]).max() < 1e-12
    # the aberration changes the hologram, and s = 0 is the plain one
    assert np.abs(g['aberrated_s09'] - ref.hologram(c, p, **INSTRUMENT)).max() > 1e-3
    members = _members(g)
    h = ref.hologram(c, members, spherical=0.5, **INSTRUMENT)
    assert np.abs(h - g['cluster_holo']).max() < 1e-12
    h = ref.hologram(c, members, **INSTRUMENT)
    assert np.abs(h - g['cluster_holo_plain']).max() < 1e-12


def test_dimer_golden(superposed_golden):
    g = superposed_golden
    x, y, z, a, n, k, theta, phi = g['dimer_params']
    axis = np.array([np.cos(theta) * np.cos(phi), np.cos(theta) * np.sin(phi),
                     np.sin(theta)])
    off = (a / INSTRUMENT['magnification']) * axis
    assert np.allclose(g['dimer_member_rp'], [off, -off])
    members = [dict(x_p=o[0], y_p=o[1], z_p=o[2], a_p=a, n_p=n, k_p=k,
                    r_0=(x, y, z)) for o in (off, -off)]
    h = ref.hologram(g['coords'], members, **INSTRUMENT)
    assert np.abs(h - g['dimer_holo']).max() < 1e-12


# ---- the fitter (oracle/fit.py) against the reference's fits ---------------
@pytest.mark.parametrize('name', ['crop00', 'crop10', 'crop07'])
def test_fit_reproduces_reference_fits(fits_golden, name):
    """MINPACK LM on the oracle's hologram at the reference's settings lands
    where the reference's Optimizer did: values, SVD uncertainties, reduced
    chi2, success and npix, at default and tight tolerances."""
    from oracle import fit
    z, meta = fits_golden
    data = z[name + '_data']
    c = ref.meshgrid(data.shape)
    start = dict(zip(fit.VARS6, z[name + '_start']))
    for mode, extra in (('default', {}), ('tight', meta['tight_settings'])):
        r = fit.fit(data, c, start, instrument=INSTRUMENT,
                    noise=float(z[name + '_sigma']), settings=extra)
        g = meta['fits'][f'{name}_{mode}']
        for k in ('redchi',) + tuple(fit.VARS6) + tuple('d' + v for v in fit.VARS6):
            assert r[k] == pytest.approx(g[k], rel=1e-9, abs=1e-12), (mode, k)
        assert r['success'] == g['success'] and r['npix'] == g['npix']


def test_fit_planted_mistakes_are_caught(fits_golden, monkeypatch):
    """A finite-difference step other than the reference's 1e-5, or the
    reduced chi2 taken over npix instead of npix - nvar, misses the
    reference's numbers at the tolerance used above."""
    from oracle import fit
    z, meta = fits_golden
    data = z['crop00_data']
    c = ref.meshgrid(data.shape)
    start = dict(zip(fit.VARS6, z['crop00_start']))
    g = meta['fits']['crop00_default']
    kw = dict(instrument=INSTRUMENT, noise=float(z['crop00_sigma']))
    r = fit.fit(data, c, start, settings={'diff_step': 1e-3}, **kw)
    assert max(abs(r[v] - g[v]) / max(abs(g[v]), 1e-3) for v in fit.VARS6) > 1e-9
    good = fit.statistics
    monkeypatch.setattr(fit, 'statistics',
                        lambda jac, cost, npix, nvar: good(jac, cost, npix, 0))
    r = fit.fit(data, c, start, **kw)
    assert r['x_p'] == pytest.approx(g['x_p'], rel=1e-9)
    assert r['redchi'] != pytest.approx(g['redchi'], rel=1e-9)


def test_layered_fit_model_invariants(fits_golden):
    """The oracle's layered fitting model (a_p scales every radius, n_p + i
    k_p is the outer index): one layer reduces to the homogeneous sphere,
    fit included; an inner layer of the outer index is invisible; and the
    coefficients depend on radii only through k r, so scaling a_p and the
    wavelength together leaves them unchanged (which fails if the inner
    radii do not follow a_p)."""
    from oracle import fit
    inst = dict(INSTRUMENT)
    p = dict(x_p=0., y_p=0., z_p=100., a_p=1.2, n_p=1.5, k_p=0.002)
    hom = ref.mie_coefficients(1.2, 1.5, 0.002, inst['n_m'], inst['wavelength'])
    one = fit.layered_properties(p, [[0.7, 1.9, 0.]], inst)['ab']
    np.testing.assert_allclose(one, hom, rtol=0, atol=1e-13)
    two = fit.layered_properties(p, [[0.3, 1.5, 0.002], [0.9, 1.7, 0.]], inst)['ab']
    np.testing.assert_allclose(two, hom, rtol=0, atol=1e-10)
    lay = np.array([[0.4, 1.8, 0.], [0.7, 1.35, 0.], [1.0, 1.6, 0.001]])
    base = ref.mie_coefficients_layered(lay[:, 0], lay[:, 1], lay[:, 2],
                                        inst['n_m'], inst['wavelength'])
    q = dict(p, a_p=2.5, n_p=1.6, k_p=0.001)
    scaled = fit.layered_properties(q, lay, dict(inst, wavelength=2.5 * inst['wavelength']))
    np.testing.assert_allclose(scaled['ab'], base, rtol=0, atol=1e-12)
    # a whole fit: the one-layer model lands where the homogeneous one does
    z, _ = fits_golden
    data = z['crop00_data']
    c = ref.meshgrid(data.shape)
    start = dict(zip(fit.VARS6, z['crop00_start']))
    kw = dict(instrument=INSTRUMENT, noise=float(z['crop00_sigma']))
    r0 = fit.fit(data, c, start, **kw)
    r1 = fit.fit(data, c, start, layers=[[1., 1.5, 0.]], **kw)
    for v in fit.VARS6:
        assert r1[v] == pytest.approx(r0[v], rel=1e-9, abs=1e-12), v


def _gauss_newton_step(data, c, values, noise=0.05):
    """One Gauss-Newton step of the 5-variable fit from ``values``, with a
    central-difference Jacobian of the oracle's hologram."""
    names = ['x_p', 'y_p', 'z_p', 'a_p', 'n_p']
    base = dict(zip(names, values), k_p=0.)
    r = (ref.hologram(c, base, **INSTRUMENT) - data) / noise
    steps = [1e-3, 1e-3, 1e-3, 1e-5, 1e-5]
    J = np.empty((r.size, 5))
    for j, (v, h) in enumerate(zip(names, steps)):
        hp = ref.hologram(c, dict(base, **{v: base[v] + h}), **INSTRUMENT)
        hm = ref.hologram(c, dict(base, **{v: base[v] - h}), **INSTRUMENT)
        J[:, j] = (hp - hm) / (2. * h * noise)
    return -np.linalg.solve(J.T @ J, J.T @ r)


def test_crop_fixture_is_the_minimum(crop_golden):
    """The stored tight fit of the reference's test hologram
    (tests/data/crop.png) is the minimum of the oracle's chi2: a
    Gauss-Newton step from it moves no parameter by 2% of its uncertainty
    (MINPACK stops where its forward-difference Jacobian, step 1e-5 |p|,
    has a zero gradient, which on this real hologram, reduced chi2 4.2,
    lies 0.2-0.9% of an uncertainty from the exact minimum), while from the
    default-tolerance fit it still moves one by more than 0.1 of its
    uncertainty."""
    g = crop_golden
    data = g['image'].astype(float)
    data = (data / data.mean()).ravel()
    c = ref.meshgrid(g['image'].shape)
    unc = g['tight_uncertainties']
    step = _gauss_newton_step(data, c, g['tight_values'])
    assert np.all(np.abs(step) < 2e-2 * unc), step / unc
    step = _gauss_newton_step(data, c, g['default_values'])
    assert np.abs(step / unc).max() > 0.1, step / unc
