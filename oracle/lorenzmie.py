"""Float64 numpy restatement of pylorenzmie's forward model.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): the checker for the HIP
kernels, never called by the product path.

Each function cites the reference file:line whose algorithm it follows.
The scattered-field amplitude ``alpha`` extends the reference as
I = |x_hat + alpha * E_s|^2 (theory/obsolete/pycudaLMHologram.py:48-49);
alpha = 1 is exactly the live reference hologram (theory/LorenzMie.py:131-133).
"""
from __future__ import annotations

import numpy as np

__all__ = ['wavenumber', 'wiscombe_yang', 'neves_pisignano', 'size_parameter',
           'mie_coefficients', 'mie_coefficients_layered', 'meshgrid',
           'lorenzmie', 'scattered_field', 'field', 'hologram', 'jacobian',
           'PARAMS']

PARAMS = ('x_p', 'y_p', 'z_p', 'a_p', 'n_p', 'k_p', 'alpha')


def wavenumber(n_m=1.340, wavelength=0.447, magnification=0.048,
               in_medium=True, scaled=True):
    """k = 2 pi / lambda [* n_m] [* magnification].

    Follows theory/Instrument.py:46-70.
    """
    k = 2. * np.pi / wavelength
    if in_medium:
        k *= n_m
    if scaled:
        k *= magnification
    return k


def wiscombe_yang(x, m) -> int:
    """Number of partial waves: Wiscombe (1980) with Yang (2003) Eq. 30.

    Follows theory/Sphere.py:91-123.  ``x`` and ``m`` may be arrays over
    layers (innermost first) for the layered extension; for a scalar the
    rolled term equals |x m| as in the reference.
    """
    x = np.atleast_1d(np.asarray(x, dtype=float))
    m = np.atleast_1d(np.asarray(m, dtype=complex))
    xl = abs(x[-1])
    if xl <= 8.:
        ns = np.floor(xl + 4. * np.cbrt(xl) + 1.)
    elif xl <= 4200.:
        ns = np.floor(xl + 4.05 * np.cbrt(xl) + 2.)
    else:
        ns = np.floor(xl + 4. * np.cbrt(xl) + 2.)
    xm = np.abs(x * m)
    xm_1 = np.abs(x[:-1] * m[1:]) if x.size > 1 else xm
    return int(max(ns, np.max(xm), np.max(xm_1)))


def neves_pisignano(x, precision=6.) -> int:
    """Neves & Pisignano (2012) truncation; theory/Sphere.py:125-145."""
    return int(x + 0.76 * np.cbrt(precision * precision * x) - 4.1)


def size_parameter(a_p, n_m=1.340, wavelength=0.447):
    """x = 2 pi Re(n_m) a_p / lambda; theory/Sphere.py:174-177."""
    return 2. * np.pi / wavelength * np.real(n_m) * a_p


def _d1_down(z: complex, nmax: int) -> np.ndarray:
    """Logarithmic derivative D1_n(z), downward from D1[nmax] = 0.

    theory/Sphere.py:190, 195-196 (Eq. 16b).
    """
    d1 = np.empty(nmax + 1, np.complex128)
    d1[nmax] = 0.
    for n in range(nmax, 0, -1):
        d1[n - 1] = n / z - 1. / (d1[n] + n / z)
    return d1


def _psi_zeta_up(x: float, d1x: np.ndarray, nmax: int):
    """Upward Psi_n(x), zeta_n(x) with D3 via Psi*zeta.

    theory/Sphere.py:206-214 (Eqs. 18a-d, 20a-b, 21a-b).
    """
    psi = np.empty(nmax + 1, np.complex128)
    zeta = np.empty(nmax + 1, np.complex128)
    d3 = np.empty(nmax + 1, np.complex128)
    d3[0] = 1.j
    psi[0] = np.sin(x)
    zeta[0] = -1.j * np.exp(1.j * x)
    psizeta = 0.5 * (1. - np.exp(2.j * x))
    for n in range(1, nmax + 1):
        psi[n] = psi[n - 1] * (n / x - d1x[n - 1])
        zeta[n] = zeta[n - 1] * (n / x - d3[n - 1])
        psizeta *= (n / x - d1x[n - 1]) * (n / x - d3[n - 1])
        d3[n] = d1x[n] + 1.j / psizeta
    return psi, zeta


def _ab_from_h(ha, hb, m, x, psi, zeta, nmax):
    """Eqs. 5 and 6 of Pena & Pal (2009); theory/Sphere.py:217-224."""
    n = np.arange(nmax + 1)
    psir = np.roll(psi, 1)
    zetar = np.roll(zeta, 1)
    ab = np.empty((nmax + 1, 2), np.complex128)
    fac = ha / m + n / x
    ab[:, 0] = (fac * psi - psir) / (fac * zeta - zetar)
    fac = hb * m + n / x
    ab[:, 1] = (fac * psi - psir) / (fac * zeta - zetar)
    ab[0, :] = 0.j
    return ab


def mie_coefficients(a_p, n_p, k_p, n_m, wavelength, nmax=None) -> np.ndarray:
    """Homogeneous-sphere Mie coefficients, (nmax+1, 2) complex128.

    Follows theory/Sphere.py:147-226 step by step (one-layer Yang
    recursion with H_a = H_b = D1(m x)).  ``nmax`` overrides the
    Wiscombe-Yang count, which a Jacobian needs to freeze the order.
    """
    k = 2. * np.pi / wavelength * np.real(n_m)
    x = k * a_p
    m = (n_p + 1.j * k_p) / n_m
    if nmax is None:
        nmax = wiscombe_yang(x, m)
    ha = _d1_down(x * m, nmax)
    d1x = _d1_down(complex(x), nmax)
    psi, zeta = _psi_zeta_up(x, d1x, nmax)
    return _ab_from_h(ha, ha, m, x, psi, zeta, nmax)


def _d3_up(z: complex, d1: np.ndarray, nmax: int) -> np.ndarray:
    """D3_n(z) upward through Psi*zeta (Yang 2003 Eqs. 18a-d)."""
    d3 = np.empty(nmax + 1, np.complex128)
    d3[0] = 1.j
    psizeta = 0.5 * (1. - np.exp(2.j * z))
    for n in range(1, nmax + 1):
        psizeta *= (n / z - d1[n - 1]) * (n / z - d3[n - 1])
        d3[n] = d1[n] + 1.j / psizeta
    return d3


def mie_coefficients_layered(radii, n_layers, k_layers, n_m, wavelength,
                             nmax=None) -> np.ndarray:
    """Layered-sphere Mie coefficients by Yang's (2003) recursion.

    ``radii`` are outer radii in um, innermost layer first and increasing;
    ``n_layers``/``k_layers`` the real/imaginary indices of each layer.
    Written from Yang, Appl. Opt. 42, 1710 (2003) Eqs. 7-8, 19 and
    Pena & Pal, Comput. Phys. Commun. 180, 2348 (2009) Eqs. 5-6, in the
    structure of lib/obsolete/sphere_coefficients.py:126-215 (which is
    broken for layered input and is not followed line by line).  The D1
    recurrences start at D1[nmax] = 0 as theory/Sphere.py:190 does, so
    one layer reproduces :func:`mie_coefficients` exactly.
    """
    radii = np.atleast_1d(np.asarray(radii, float))
    m = ((np.atleast_1d(np.asarray(n_layers, float)) +
          1.j * np.atleast_1d(np.asarray(k_layers, float))) / n_m)
    k = 2. * np.pi / wavelength * np.real(n_m)
    x = k * radii
    nl = x.size
    if nmax is None:
        nmax = wiscombe_yang(x, m)
    n = np.arange(nmax + 1)
    # layer 1: H_a = H_b = D1(m_1 x_1)
    ha = _d1_down(m[0] * x[0], nmax)
    hb = ha.copy()
    for ll in range(1, nl):
        za = m[ll] * x[ll - 1]
        zb = m[ll] * x[ll]
        d1a = _d1_down(za, nmax)
        d1b = _d1_down(zb, nmax)
        d3a = _d3_up(za, d1a, nmax)
        d3b = _d3_up(zb, d1b, nmax)
        # Q_n = [Psi_n(za)/zeta_n(za)] / [Psi_n(zb)/zeta_n(zb)], Yang Eq. 19
        q = np.empty(nmax + 1, np.complex128)
        q[0] = (np.exp(2.j * (zb - za)) * (np.exp(2.j * za) - 1.) /
                (np.exp(2.j * zb) - 1.))
        r2 = (x[ll - 1] / x[ll]) ** 2
        for j in range(1, nmax + 1):
            num = (zb * d1b[j] + j) * (j - zb * d3b[j - 1])
            den = (za * d1a[j] + j) * (j - za * d3a[j - 1])
            q[j] = q[j - 1] * r2 * num / den
        g1 = m[ll] * ha - m[ll - 1] * d1a
        g2 = m[ll] * ha - m[ll - 1] * d3a
        ha = (g2 * d1b - q * g1 * d3b) / (g2 - q * g1)
        g1 = m[ll - 1] * hb - m[ll] * d1a
        g2 = m[ll - 1] * hb - m[ll] * d3a
        hb = (g2 * d1b - q * g1 * d3b) / (g2 - q * g1)
    xl = x[-1]
    d1x = _d1_down(complex(xl), nmax)
    psi, zeta = _psi_zeta_up(xl, d1x, nmax)
    del n
    return _ab_from_h(ha, hb, m[-1], xl, psi, zeta, nmax)


def meshgrid(shape, corner=(0., 0.), flatten=True):
    """(2, ny*nx) pixel grid, x fastest; lib/meshgrid.py:10-40."""
    ny, nx = shape
    left, top = corner
    x = np.arange(left, left + nx, dtype=float)
    y = np.arange(top, top + ny, dtype=float)
    xy = np.array(np.meshgrid(x, y))
    return xy.reshape((2, -1)) if flatten else xy


def lorenzmie(ab, kdr, cartesian=True, bohren=True) -> np.ndarray:
    """Partial-wave sum for the scattered field, (3, npts) complex128.

    Follows theory/LorenzMie.py:183-342: geometry (:221-237), starting
    values of xi (:249-251), Wiscombe's angular recurrence (:279-281,
    :311-312), upward xi (:284), Deirmendjian derivative (:287),
    vector spherical harmonics (:290-294), E_n prefactor (:297-298),
    accumulation (:302-306), geometric factors (:322-325) and the
    Cartesian projection (:331-340).
    """
    kdr = np.asarray(kdr, float)
    norders = ab.shape[0]
    kx = kdr[0]
    ky = kdr[1]
    kz = -kdr[2]
    krho = np.hypot(kx, ky)
    kr = np.hypot(krho, kz)
    sinkr = np.sin(kr)
    coskr = np.cos(kr)
    phi = np.arctan2(ky, kx)
    cosphi = np.cos(phi)
    sinphi = np.sin(phi)
    theta = np.arctan2(krho, kz)
    costheta = np.cos(theta)
    sintheta = np.sin(theta)
    f = (1.j if bohren else -1.j) * np.sign(kz)
    xi_nm2 = coskr + f * sinkr
    xi_nm1 = sinkr - f * coskr
    pi_nm1 = np.zeros_like(kr)
    pi_n = np.ones_like(kr)
    es = np.zeros((3,) + kr.shape, np.complex128)
    ifac = 1.
    for n in range(1, norders):
        s = pi_n * costheta
        t = s - pi_nm1
        tau = pi_nm1 - n * t          # -tau_n(cos theta)
        xi_n = (2. * n - 1.) * (xi_nm1 / kr) - xi_nm2
        dn = n * (xi_n / kr) - xi_nm1
        m_t = pi_n * xi_n
        m_p = tau * xi_n
        n_r = (n * n + n) * m_t
        n_t = tau * dn
        n_p = pi_n * dn
        ifac = ifac * 1.j
        en = ifac * (2. * n + 1.) / (n * n + n)
        an = 1.j * en * ab[n, 0]
        bn = en * ab[n, 1]
        es[0] += an * n_r
        es[1] += an * n_t - bn * m_t
        es[2] += an * n_p - bn * m_p
        pi_nm1 = pi_n
        pi_n = s + (1. + 1. / n) * t
        xi_nm2 = xi_nm1
        xi_nm1 = xi_n
    es[0] *= cosphi * sintheta / kr
    es[1] *= cosphi
    es[2] *= sinphi
    es /= kr
    if not cartesian:
        return es
    ec = np.empty_like(es)
    ec[0] = es[0] * sintheta * cosphi + es[1] * costheta * cosphi - es[2] * sinphi
    ec[1] = es[0] * sintheta * sinphi + es[1] * costheta * sinphi + es[2] * cosphi
    ec[2] = es[0] * costheta - es[1] * sintheta
    return ec


def _coords3(coordinates):
    c = np.atleast_2d(np.asarray(coordinates, float))
    ndim, npts = c.shape
    if ndim > 3:
        raise ValueError(f'Incompatible shape: {c.shape=}')
    return np.vstack([c, np.zeros((3 - ndim, npts))])


def scattered_field(coordinates, x_p, y_p, z_p, ab, k,
                    r_0=(0., 0., 0.), cartesian=True, bohren=True):
    """Field of one particle times exp(-i k z_p); theory/LorenzMie.py:135-158."""
    c = _coords3(coordinates)
    r_p = np.array([x_p, y_p, z_p], float) + np.asarray(r_0, float)
    kdr = k * (c - r_p[:, None])
    return lorenzmie(ab, kdr, cartesian, bohren) * np.exp(-1.j * k * r_p[2])


def aberration(coordinates, r_p, k, n_m, spherical):
    """Spherical-aberration phase mask of a particle at r_p (absolute),
    theory/Aberrated.py:61-76: exp(i (4 n_m^3 / k) s r^4 / (r^2 + z_p^2)^2)
    with r the in-plane pixel distance and k in rad/px."""
    if spherical == 0.:
        return 1.
    c = _coords3(coordinates)
    x = c[0] - r_p[0]
    y = c[1] - r_p[1]
    rsq = x * x + y * y
    zpsq = r_p[2] ** 2
    phase = (4. * n_m ** 3 / k) * spherical * rsq ** 2 / (rsq + zpsq) ** 2
    return np.exp(1j * phase)


def field(coordinates, particles, n_m=1.340, wavelength=0.447,
          magnification=0.048, cartesian=True, bohren=True, spherical=0.):
    """Superposed scattered field, theory/LorenzMie.py:160-181, each
    particle's field times its aberration mask when ``spherical`` is not 0
    (theory/Aberrated.py:78-84).

    ``particles`` is a dict or a list of dicts with keys x_p, y_p, z_p,
    a_p, n_p, k_p and optionally r_0, nmax, or ``ab`` (precomputed).
    """
    if isinstance(particles, dict):
        particles = [particles]
    k = wavenumber(n_m, wavelength, magnification)
    total = 0.
    for p in particles:
        ab = p.get('ab')
        if ab is None:
            ab = mie_coefficients(p['a_p'], p['n_p'], p.get('k_p', 0.),
                                  n_m, wavelength, p.get('nmax'))
        r_0 = np.asarray(p.get('r_0', (0., 0., 0.)), float)
        e = scattered_field(coordinates, p['x_p'], p['y_p'], p['z_p'], ab, k,
                            r_0, cartesian, bohren)
        r_p = np.array([p['x_p'], p['y_p'], p['z_p']]) + r_0
        total = total + e * aberration(coordinates, r_p, k, n_m, spherical)
    return total


def hologram(coordinates, particles, n_m=1.340, wavelength=0.447,
             magnification=0.048, alpha=1.0, cartesian=True, bohren=True,
             spherical=0.):
    """I = |x_hat + alpha E_s|^2 summed over components.

    theory/LorenzMie.py:113-133 (alpha = 1); alpha as in
    theory/obsolete/pycudaLMHologram.py:48-49.
    """
    e = alpha * field(coordinates, particles, n_m, wavelength, magnification,
                      cartesian, bohren, spherical)
    e[0] += 1.
    return np.sum(e.real ** 2 + e.imag ** 2, axis=0)


def _holo7(coordinates, q, instrument, nmax):
    p = dict(x_p=q[0], y_p=q[1], z_p=q[2], a_p=q[3], n_p=q[4], k_p=q[5],
             nmax=nmax)
    return hologram(coordinates, p, alpha=q[6], **instrument)


def jacobian(coordinates, params, instrument=None, order=4, rel_step=None):
    """dI/d(x_p, y_p, z_p, a_p, n_p, k_p, alpha), (7, npts) float64.

    Central finite differences of :func:`hologram` with n_max frozen at
    the central point, as theory/jaxLorenzMie.py:359-361 freezes it for
    jax.jacfwd.  ``order`` 4 uses the 5-point stencil, 6 the 7-point one.
    """
    instrument = instrument or {}
    n_m = instrument.get('n_m', 1.340)
    wavelength = instrument.get('wavelength', 0.447)
    q0 = np.array([params[k] if k in params else (1. if k == 'alpha' else 0.)
                   for k in PARAMS], float)
    nmax = wiscombe_yang(size_parameter(q0[3], n_m, wavelength),
                         (q0[4] + 1.j * q0[5]) / n_m)
    # absolute steps: pixels for positions, um for a_p, index units
    steps = np.array([1e-3, 1e-3, 1e-3, 1e-4, 1e-4, 1e-4, 1e-4])
    if rel_step is not None:
        steps = np.maximum(np.abs(q0), 1e-2) * rel_step
    if order == 4:
        w = {1: 8. / 12., 2: -1. / 12.}
    else:
        w = {1: 45. / 60., 2: -9. / 60., 3: 1. / 60.}
    c = _coords3(coordinates)
    jac = np.zeros((7, c.shape[1]))
    for j in range(7):
        for s, wt in w.items():
            qp = q0.copy()
            qm = q0.copy()
            qp[j] += s * steps[j]
            qm[j] -= s * steps[j]
            jac[j] += wt * (_holo7(c, qp, instrument, nmax) -
                            _holo7(c, qm, instrument, nmax))
        jac[j] /= steps[j]
    return jac
