"""Textbook restatements used to pin the oracle (TEST INFRASTRUCTURE ONLY).

Nothing here shares a code path with oracle/lorenzmie.py:
  * mie_ab_mpmath: Bohren & Huffman Eq. 4.53 with Riccati-Bessel functions
    from mpmath spherical Bessel functions at 30 digits;
  * coated_ab_mpmath: the coated-sphere closed form of B&H Sec. 8.1 in
    mpmath (two layers);
  * field_bh: the scattered field of B&H Eqs. 4.45-4.50 summed literally,
    with spherical Hankel functions from scipy.special.spherical_jn/yn and
    pi_n, tau_n from associated Legendre functions (scipy.special.lpmv).
The scattered field is outgoing h_n^(1) (particle above the focal plane,
theory/LorenzMie.py:240-251) in the frame theory/LorenzMie.py:221-223 uses.
"""
from __future__ import annotations

import mpmath as mp
import numpy as np
from scipy import special


def mie_ab_mpmath(a_p, n_p, k_p, n_m, wavelength, nmax, dps=30):
    """a_n, b_n (n = 1..nmax) from B&H Eq. 4.53 in mpmath."""
    mp.mp.dps = dps
    x = mp.mpf(2) * mp.pi / mp.mpf(wavelength) * mp.mpf(n_m) * mp.mpf(a_p)
    m = mp.mpc(n_p, k_p) / mp.mpf(n_m)
    mx = m * x

    def psi(n, z):           # Riccati-Bessel z j_n(z)
        return z * mp.sqrt(mp.pi / (2 * z)) * mp.besselj(n + mp.mpf(1) / 2, z)

    def dpsi(n, z):
        return psi(n - 1, z) - n * psi(n, z) / z

    def xi(n, z):            # z h_n^(1)(z)
        return z * mp.sqrt(mp.pi / (2 * z)) * (mp.besselj(n + mp.mpf(1) / 2, z) +
                                               1j * mp.bessely(n + mp.mpf(1) / 2, z))

    def dxi(n, z):
        return xi(n - 1, z) - n * xi(n, z) / z

    out = np.zeros((nmax + 1, 2), complex)
    for n in range(1, nmax + 1):
        a = ((m * psi(n, mx) * dpsi(n, x) - psi(n, x) * dpsi(n, mx)) /
             (m * psi(n, mx) * dxi(n, x) - xi(n, x) * dpsi(n, mx)))
        b = ((psi(n, mx) * dpsi(n, x) - m * psi(n, x) * dpsi(n, mx)) /
             (psi(n, mx) * dxi(n, x) - m * xi(n, x) * dpsi(n, mx)))
        out[n] = complex(a), complex(b)
    return out


def coated_ab_mpmath(a_core, a_shell, n_core, k_core, n_shell, k_shell, n_m,
                     wavelength, nmax, dps=40):
    """a_n, b_n (n = 1..nmax) of a coated sphere, Bohren & Huffman Sec. 8.1
    (the closed form of Aden & Kerker), in mpmath.

    x = k a_core, y = k a_shell, m1, m2 the core and shell indices relative
    to the medium; psi_n(z) = z j_n(z), chi_n(z) = -z y_n(z),
    xi_n = psi_n - i chi_n.
    """
    mp.mp.dps = dps
    k = mp.mpf(2) * mp.pi / mp.mpf(wavelength) * mp.mpf(n_m)
    x, y = k * mp.mpf(a_core), k * mp.mpf(a_shell)
    m1 = mp.mpc(n_core, k_core) / mp.mpf(n_m)
    m2 = mp.mpc(n_shell, k_shell) / mp.mpf(n_m)
    half = mp.mpf(1) / 2

    def psi(n, z):
        return mp.sqrt(mp.pi * z / 2) * mp.besselj(n + half, z)

    def chi(n, z):
        return -mp.sqrt(mp.pi * z / 2) * mp.bessely(n + half, z)

    def d(f, n, z):              # f_n'(z) = f_{n-1}(z) - n f_n(z) / z
        return f(n - 1, z) - n * f(n, z) / z

    def xi(n, z):
        return psi(n, z) - 1j * chi(n, z)

    out = np.zeros((nmax + 1, 2), complex)
    for n in range(1, nmax + 1):
        p1, dp1 = psi(n, m1 * x), d(psi, n, m1 * x)
        p2, dp2 = psi(n, m2 * x), d(psi, n, m2 * x)
        c2, dc2 = chi(n, m2 * x), d(chi, n, m2 * x)
        A = (m2 * p2 * dp1 - m1 * dp2 * p1) / (m2 * c2 * dp1 - m1 * dc2 * p1)
        B = (m2 * p1 * dp2 - m1 * p2 * dp1) / (m2 * dc2 * p1 - m1 * dp1 * c2)
        py, dpy = psi(n, y), d(psi, n, y)
        xy, dxy = xi(n, y), d(xi, n, y)
        p2y, dp2y = psi(n, m2 * y), d(psi, n, m2 * y)
        c2y, dc2y = chi(n, m2 * y), d(chi, n, m2 * y)
        ua = dp2y - A * dc2y
        va = p2y - A * c2y
        a = (py * ua - m2 * dpy * va) / (xy * ua - m2 * dxy * va)
        ub = dp2y - B * dc2y
        vb = p2y - B * c2y
        b = (m2 * py * ub - dpy * vb) / (m2 * xy * ub - dxy * vb)
        out[n] = complex(a), complex(b)
    return out


def field_bh(coords, r_p, ab, k):
    """Cartesian scattered field by the literal B&H multipole sum.

    E_s = sum_n E_n (i a_n N_e1n - b_n M_o1n) with E_n = i^n (2n+1)/(n(n+1)),
    outgoing spherical Hankel functions h_n^(1)(kr), and the frame of
    theory/LorenzMie.py:221-223 (z measured from the particle toward the
    focal plane).  Returns (3, npts) complex, times exp(-i k z_p).
    """
    c = np.asarray(coords, float)
    if c.shape[0] == 2:
        c = np.vstack([c, np.zeros((1, c.shape[1]))])
    X = k * (c[0] - r_p[0])
    Y = k * (c[1] - r_p[1])
    Z = -k * (c[2] - r_p[2])
    rho = np.hypot(X, Y)
    R = np.hypot(rho, Z)
    theta = np.arctan2(rho, Z)
    phi = np.arctan2(Y, X)
    ct, st = np.cos(theta), np.sin(theta)
    cp, sp = np.cos(phi), np.sin(phi)
    Er = np.zeros_like(R, complex)
    Et = np.zeros_like(R, complex)
    Ep = np.zeros_like(R, complex)
    nmax = ab.shape[0] - 1
    for n in range(1, nmax + 1):
        En = (1j ** n) * (2 * n + 1) / (n * (n + 1))
        h = special.spherical_jn(n, R) + 1j * special.spherical_yn(n, R)
        hm = special.spherical_jn(n - 1, R) + 1j * special.spherical_yn(n - 1, R)
        xi = R * h
        dxi = R * hm - n * h          # d(R h_n)/dR
        # pi_n = P_n^1 / sin, tau_n = dP_n^1 / dtheta (B&H 4.46, sign of P_n^1)
        P1 = -special.lpmv(1, n, ct)
        with np.errstate(invalid='ignore', divide='ignore'):
            pi_n = np.where(st > 1e-12, P1 / np.where(st > 1e-12, st, 1.),
                            0.5 * n * (n + 1) * np.sign(ct) ** (n + 1))
        pn_m1 = -special.lpmv(1, n - 1, ct) if n > 1 else np.zeros_like(ct)
        with np.errstate(invalid='ignore', divide='ignore'):
            pim1 = np.where(st > 1e-12, pn_m1 / np.where(st > 1e-12, st, 1.),
                            0.5 * (n - 1) * n * np.sign(ct) ** n)
        tau_n = n * ct * pi_n - (n + 1) * pim1
        # B&H 4.50 with z h_n and the geometric factors
        Mt = cp * pi_n * h
        Mp = -sp * tau_n * h
        Nr = n * (n + 1) * cp * st * pi_n * h / R
        Nt = cp * tau_n * dxi / R
        Np = -sp * pi_n * dxi / R
        Er += En * (1j * ab[n, 0] * Nr)
        Et += En * (1j * ab[n, 0] * Nt - ab[n, 1] * Mt)
        Ep += En * (1j * ab[n, 0] * Np - ab[n, 1] * Mp)
    Ex = Er * st * cp + Et * ct * cp - Ep * sp
    Ey = Er * st * sp + Et * ct * sp + Ep * cp
    Ez = Er * ct - Et * st
    return np.array([Ex, Ey, Ez]) * np.exp(-1j * k * r_p[2])
