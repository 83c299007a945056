"""CPU checks of the kernel arithmetic (pylorenzmie_amd/csrc/lmz_pixel.h)
compiled for the host into tests/native/liblmz_hostcheck.so, against the
FP64 oracle and the golden vectors.  The GPU tests repeat these through
liblmz.so on the device."""
import ctypes
import os

import numpy as np
import pytest

from conftest import INSTRUMENT, ROOT
from oracle import lorenzmie as ref

LIB = os.path.join(ROOT, 'tests', 'native', 'liblmz_hostcheck.so')
COEF_STRIDE = 40
DCOEF_STRIDE = 24


def order_sign(nn):
    """sigma_n = +1, +1, -1, -1, ... for n = -1, 0, 1, 2, ...: the sign the
    a_p and n_p tangent rows of dcoef carry (lmz_pixel.h order_sign)."""
    return np.where(((np.asarray(nn) + 1) & 2) != 0, -1., 1.)


@pytest.fixture(scope='module')
def chk():
    if not os.path.exists(LIB):
        pytest.skip('host checker not built (run __graft_entry__.build())')
    from pylorenzmie_amd._lmz import Feature
    lib = ctypes.CDLL(LIB)
    P = ctypes.c_void_p
    lib.chk_mie.restype = ctypes.c_int
    lib.chk_mie.argtypes = [ctypes.c_double] * 5 + [ctypes.c_int, P, P, P]
    lib.chk_mie_layered.restype = ctypes.c_int
    lib.chk_mie_layered.argtypes = [P, P, P, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int, P]
    lib.chk_mie_layered_tangent.restype = ctypes.c_int
    lib.chk_mie_layered_tangent.argtypes = [P, P, P, ctypes.c_int,
                                            ctypes.c_double, ctypes.c_double,
                                            ctypes.c_int, P]
    lib.chk_pixels.argtypes = [P, P, P, P, ctypes.c_int, P, ctypes.c_int,
                               ctypes.c_int, P, P, P]
    lib.chk_dense_coords.argtypes = [ctypes.c_int, ctypes.c_int, P, P]
    lib.chk_holo32.argtypes = [P, P, P, ctypes.c_int, P, ctypes.c_int,
                               ctypes.c_int, P]
    return lib, Feature


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def mie(chk, a, n, k, cap=400):
    lib, _ = chk
    coef = np.zeros((cap, COEF_STRIDE), np.float32)
    dcoef = np.zeros((cap, DCOEF_STRIDE), np.float32)
    ab = np.zeros((cap, 4))
    N = lib.chk_mie(a, n, k, 1.34, 0.447, cap, _p(coef), _p(dcoef), _p(ab))
    return N, coef, dcoef, ab


def pixels(chk, q, coords, flags=3, origin=(0., 0., 0.)):
    lib, Feature = chk
    N, coef, dcoef, _ = mie(chk, q[3], q[4], q[5])
    F = Feature()
    F.npix = coords.shape[1]
    F.origin[:] = origin
    F.n_m, F.wavelength, F.magnification, F.noise = 1.34, 0.447, 0.048, 0.05
    F.nmax_cap = 400
    p = np.zeros(8)
    p[:len(q)] = q
    if len(q) < 7:
        p[6] = 1.
    c = np.ascontiguousarray((coords - np.asarray(origin)[:, None]).astype(np.float32))
    n = c.shape[1]
    field = np.zeros((3, n, 2), np.float32)
    holo = np.zeros(n, np.float32)
    jac = np.zeros((7, n), np.float32)
    lib.chk_pixels(ctypes.byref(F), _p(p), _p(coef), _p(dcoef), N, _p(c), n,
                   flags, _p(field), _p(holo), _p(jac))
    return field[..., 0] + 1j * field[..., 1], holo, jac


def grid3(shape):
    c = ref.meshgrid(shape)
    return np.vstack([c, np.zeros((1, c.shape[1]))])


def test_mie_matches_golden_and_tangents(chk, forward_golden):
    g = forward_golden
    offs = g['ab_offsets']
    for i, (a, n, k) in enumerate(g['ab_cases']):
        N, _, _, ab = mie(chk, a, n, k)
        gold = g['ab_values'][offs[i] + 1:offs[i + 1]]
        assert N == gold.shape[0]
        got = np.stack([ab[:N, 0] + 1j * ab[:N, 1], ab[:N, 2] + 1j * ab[:N, 3]], 1)
        assert np.abs(got - gold).max() < 1e-11
    # tangents d a_n / d(a_p, n_p, k_p) against 5-point FD of the oracle
    a, n, k = 1.3, 1.55, 0.004
    N, _, dcoef, _ = mie(chk, a, n, k)
    nn = np.arange(1, N + 1)
    en = (2 * nn + 1) / (nn * (nn + 1))
    for j in range(3):
        h = 1e-4
        q = np.array([a, n, k])
        fd = 0
        for s, w in ((1, 8 / 12), (2, -1 / 12)):
            qp, qm = q.copy(), q.copy()
            qp[j] += s * h
            qm[j] -= s * h
            fd = fd + w * (ref.mie_coefficients(*qp, 1.34, 0.447, nmax=N) -
                           ref.mie_coefficients(*qm, 1.34, 0.447, nmax=N))
        fd = fd[1:] / h
        if j < 2:   # (Re, Im, -Im) triples of dPR, dP, -dQ at 9 j
            b = 9 * j
            sg = order_sign(nn)
            dP = sg * (dcoef[:N, b + 3] + 1j * dcoef[:N, b + 4])
            dQ = -sg * (dcoef[:N, b + 6] + 1j * dcoef[:N, b + 7])
            assert np.array_equal(dcoef[:N, b + 5], -dcoef[:N, b + 4])
            assert np.array_equal(dcoef[:N, b + 8], -dcoef[:N, b + 7])
        else:       # k_p: (Re, Im) of dPR, dP, dQ at 18
            dP = dcoef[:N, 20] + 1j * dcoef[:N, 21]
            dQ = dcoef[:N, 22] + 1j * dcoef[:N, 23]
        da = dP / (1j ** (nn + 1) * en)
        db = dQ / (1j ** nn * en)
        scale = np.abs(fd).max()
        assert np.abs(da - fd[:, 0]).max() < 1e-5 * scale
        assert np.abs(db - fd[:, 1]).max() < 1e-5 * scale


@pytest.mark.parametrize('a,n,k', [(5.0, 1.9, 0.002), (9.0, 1.45, 0.002),
                                   (3.0, 2.0, 0.0)])
def test_large_sphere_tangents_match_fine_differences(chk, a, n, k):
    """Large and high-index spheres (N to ~250), where the coefficients
    swing fast with n_p and a_p: the FP64 tangents (stored as FP32) against
    5-point differences of the FP64 oracle with steps of 1e-7 relative,
    fine enough that the stencil resolves the swings (1e-4 does not at
    a_p = 5 um, n_p = 1.9)."""
    N, _, dcoef, _ = mie(chk, a, n, k)
    nn = np.arange(1, N + 1)
    en = (2 * nn + 1) / (nn * (nn + 1))
    q = np.array([a, n, k])
    for j in range(3):
        h = 1e-7 * max(abs(q[j]), 1e-2)
        fd = 0
        for s, w in ((1, 8 / 12), (2, -1 / 12)):
            qp, qm = q.copy(), q.copy()
            qp[j] += s * h
            qm[j] -= s * h
            fd = fd + w * (ref.mie_coefficients(*qp, 1.34, 0.447, nmax=N) -
                           ref.mie_coefficients(*qm, 1.34, 0.447, nmax=N))
        fd = fd[1:] / h
        if j < 2:
            b = 9 * j
            sg = order_sign(nn)
            dP = sg * (dcoef[:N, b + 3] + 1j * dcoef[:N, b + 4])
            dQ = -sg * (dcoef[:N, b + 6] + 1j * dcoef[:N, b + 7])
        else:
            dP = dcoef[:N, 20] + 1j * dcoef[:N, 21]
            dQ = dcoef[:N, 22] + 1j * dcoef[:N, 23]
        da = dP / (1j ** (nn + 1) * en)
        db = dQ / (1j ** nn * en)
        scale = np.abs(fd).max()
        ea = np.abs(da - fd[:, 0]).max() / scale
        eb = np.abs(db - fd[:, 1]).max() / scale
        print(f'a_p {a} n_p {n} param {j}: N {N}, tangent vs FD {ea:.1e} {eb:.1e}')
        assert ea < 2e-6 and eb < 2e-6, j


def test_forward_matches_golden(chk, forward_golden):
    g = forward_golden
    _, h, _ = pixels(chk, g['cfg0_params'], grid3((201, 201)))
    assert np.abs(h - g['cfg0_holo']).max() < 1e-6
    for cart in (1, 0):
        for boh in (1, 0):
            f, _, _ = pixels(chk, g['field64_params'], grid3((64, 64)),
                             flags=cart + 2 * boh)
            assert np.abs(f - g[f'field64_c{cart}_b{boh}']).max() < 2e-7
    offs = g['holo_offsets']
    for i, p in enumerate(g['holo_cases']):
        c = g['holo_coords'][:, offs[i]:offs[i + 1]]
        org = (np.round(c[0].mean()), np.round(c[1].mean()), 0.)
        _, h, _ = pixels(chk, p, c, origin=org)
        assert np.abs(h - g['holo_values'][offs[i]:offs[i + 1]]).max() < 1e-6, i


def test_bright_forward_lobe(chk):
    """A 1.6 um sphere 3 um above the focal plane (I up to ~40)."""
    p = (17.41, 11.047, 61.918, 1.595, 1.608, 0.006)
    c = grid3((32, 36))
    truth = ref.hologram(c, dict(zip(ref.PARAMS, p)), **INSTRUMENT)
    _, h, _ = pixels(chk, p, c)
    assert truth.max() > 35
    assert np.abs(h - truth).max() < 5e-6


@pytest.mark.parametrize('q', [(24.3, 23.6, 140., 0.9, 1.5, 0.003, 0.9),
                               (24.0, 24.0, 100., 2.0, 1.6, 0.0, 1.1),
                               (23.7, 24.4, -80., 0.8, 1.45, 0.0, 1.0)])
def test_jacobian_matches_oracle(chk, q):
    c = grid3((40, 40))
    _, _, jac = pixels(chk, q, c)
    Jref = ref.jacobian(c, dict(zip(ref.PARAMS, q)), INSTRUMENT, order=6)
    for i in range(7):
        assert np.abs(jac[i] - Jref[i]).max() < 2e-4 * np.abs(Jref[i]).max()


@pytest.mark.parametrize('q', [(20.3, 19.6, 140., 0.9, 1.5, 0.003, 0.9),
                               (20.0, 21.0, 100., 2.0, 1.6, 0.0, 1.1),
                               (19.7, 20.4, -80., 0.8, 1.45, 0.0, 1.0),
                               (55.0, -8.0, 120., 1.2, 1.45, 0.0, 1.0)])
def test_radial_table_jacobian_matches_oracle(chk, q):
    """Dense crop (41 x 41): the Jacobian from the radial tables of the sums
    and from the order loop at every pixel, each within 2e-4 of the column
    scale of the oracle's 6th-order differences (the last case's axis lies
    outside the crop).  Read through windows of half the table by groups of
    four crop rows and of one (as the kernel reads a table larger than its
    LDS), the table gives the same bits: every row a pixel reads lies in its
    group's window."""
    lib, Feature = chk
    P = ctypes.c_void_p
    I = ctypes.c_int
    lib.chk_jac_rtab.restype = ctypes.c_int
    lib.chk_jac_rtab.argtypes = [P, P, P, P, I, I, P, P, P, I, I, P]
    side = 41
    N, coef, dcoef, _ = mie(chk, q[3], q[4], q[5])
    F = Feature()
    F.npix, F.width = side * side, side
    F.n_m, F.wavelength, F.magnification, F.noise = 1.34, 0.447, 0.048, 0.05
    F.nmax_cap = 400
    p = np.zeros(8)
    p[:7] = q
    jl = np.zeros((7, side * side), np.float32)
    jt = np.zeros_like(jl)
    h = ctypes.c_float()
    nwin = ctypes.c_int()
    m = lib.chk_jac_rtab(ctypes.byref(F), _p(p), _p(coef), _p(dcoef), N, 3,
                         _p(jl), _p(jt), ctypes.byref(h), 0, 1024,
                         ctypes.byref(nwin))
    assert 12 <= m < 1024 and 0. < h.value <= 1.
    for rows in (4, 1):
        jw = np.zeros_like(jl)
        lib.chk_jac_rtab(ctypes.byref(F), _p(p), _p(coef), _p(dcoef), N, 3,
                         _p(jl), _p(jw), ctypes.byref(h), m // 2, rows * side,
                         ctypes.byref(nwin))
        print('m', m, 'rows', rows, 'windowed groups', nwin.value)
        assert nwin.value > 0
        assert np.array_equal(jw, jt)
    Jref = ref.jacobian(grid3((side, side)), dict(zip(ref.PARAMS, q)),
                        INSTRUMENT, order=6)
    for i in range(7):
        s = np.abs(Jref[i]).max()
        assert np.abs(jl[i] - Jref[i]).max() < 2e-4 * s
        assert np.abs(jt[i] - Jref[i]).max() < 2e-4 * s


@pytest.mark.parametrize('q', [(100.3, 99.6, 150., 1.5, 1.5, 0.002, 1.0),
                               (100.2, 100.7, 100., 3.0, 1.6, 0.01, 0.9),
                               (99.8, 100.4, 70., 3.0, 1.6, 0.0, 1.0),
                               (100.6, 99.1, -90., 1.5, 1.5, 0.0, 1.1),
                               (240.0, -30.0, 120., 1.2, 1.45, 0.0, 1.0)])
def test_forward_radial_table_matches_order_loop(chk, q):
    """Dense 201 x 201 crop: the FP64 hologram from the radial tables of the
    sums equals the order loop's at every pixel to 2e-13 of the brightest
    pixel, and is no farther from the oracle's (numpy, no FMA) than the
    loop's is (FP64 rounding of the two sums, up to 4e-12 on the bright
    lobes of 3 um spheres).  The last case's axis lies outside the crop.
    Read through windows of half the table by 256-pixel tiles (as the
    kernel reads a table larger than its LDS), the table gives the same
    bits."""
    lib, Feature = chk
    P = ctypes.c_void_p
    I = ctypes.c_int
    lib.chk_fwd_ftab.restype = ctypes.c_int
    lib.chk_fwd_ftab.argtypes = [P, P, P, I, I, P, P, P, I, I, P]
    side = 201
    N, coef, _, _ = mie(chk, q[3], q[4], q[5])
    F = Feature()
    F.npix, F.width = side * side, side
    F.n_m, F.wavelength, F.magnification, F.noise = 1.34, 0.447, 0.048, 0.05
    F.nmax_cap = 400
    p = np.zeros(8)
    p[:7] = q
    hl = np.zeros(side * side)
    ht = np.zeros(side * side)
    h = ctypes.c_double()
    nwin = ctypes.c_int()
    m = lib.chk_fwd_ftab(ctypes.byref(F), _p(p), _p(coef), N, 3, _p(hl), _p(ht),
                         ctypes.byref(h), 0, 256, ctypes.byref(nwin))
    assert 2 * 12 <= m < 2048 and 0. < h.value <= 1.
    hw = np.zeros_like(ht)
    lib.chk_fwd_ftab(ctypes.byref(F), _p(p), _p(coef), N, 3, _p(hl), _p(hw),
                     ctypes.byref(h), m // 2, 256, ctypes.byref(nwin))
    print('windowed tiles', nwin.value)
    assert nwin.value > 0
    assert np.array_equal(hw, ht)
    print('m', m, 'h', h.value, 'max |tab - loop|', np.abs(ht - hl).max())
    assert np.abs(ht - hl).max() < 2e-13 * hl.max()
    c = grid3((side, side))
    truth = ref.hologram(c, dict(zip(ref.PARAMS, q[:6])), alpha=q[6], **INSTRUMENT)
    assert np.abs(ht - truth).max() <= 1.5 * np.abs(hl - truth).max() + 1e-14


def test_dense_pixel_mapping(chk):
    lib, _ = chk
    for w, n in ((201, 201 * 201), (7, 7 * 13), (4096, 4096 * 3), (1, 17)):
        ax = np.zeros(n, np.float32)
        ay = np.zeros(n, np.float32)
        lib.chk_dense_coords(w, n, _p(ax), _p(ay))
        j = np.arange(n)
        assert np.array_equal(ax, (j % w).astype(np.float32))
        assert np.array_equal(ay, (j // w).astype(np.float32))


def mie_layered(chk, radii, nl, kl, cap=600):
    lib, _ = chk
    r, n, k = (np.ascontiguousarray(v, float) for v in (radii, nl, kl))
    ab = np.zeros((cap, 4))
    N = lib.chk_mie_layered(_p(r), _p(n), _p(k), len(r), 1.34, 0.447, cap, _p(ab))
    return N, ab[:N, 0::2] + 1j * ab[:N, 1::2]


@pytest.mark.parametrize('seed', range(6))
def test_layered_mie_matches_oracle(chk, seed):
    """2-5 layers, radii sorted uniform within a_p <= 10 um, n in
    [1.33, 2.0] (the hard-regime draws), against the oracle's Yang
    recursion with the same D1 start."""
    rng = np.random.default_rng(100 + seed)
    L = int(rng.integers(2, 6))
    a_p = rng.uniform(0.5, 10.)
    radii = np.sort(rng.uniform(0.05, 1., L)) * a_p
    radii[-1] = a_p
    nl = rng.uniform(1.33, 2.0, L)
    kl = rng.uniform(0., 0.01, L) * (rng.random(L) < 0.5)
    N, ab = mie_layered(chk, radii, nl, kl)
    want = ref.mie_coefficients_layered(radii, nl, kl, 1.34, 0.447)
    assert N == want.shape[0] - 1
    scale = np.abs(want).max()
    err = np.abs(ab - want[1:]).max() / scale
    print(f'L={L} a_p={a_p:.2f} N={N}: max err / max |ab| {err:.2e}')
    assert err < 1e-9


def test_layered_one_layer_equals_homogeneous(chk):
    for a, n, k in ((0.75, 1.45, 0.), (2.5, 1.6, 0.003), (9.5, 1.9, 0.)):
        N1, ab1 = mie_layered(chk, [a], [n], [k])
        N, _, _, ab = mie(chk, a, n, k, cap=600)
        assert N1 == N
        got = np.stack([ab[:N, 0] + 1j * ab[:N, 1], ab[:N, 2] + 1j * ab[:N, 3]], 1)
        assert np.abs(ab1 - got).max() < 1e-12 * np.abs(got).max()


@pytest.mark.parametrize('q,tol', [((100.3, 99.6, 150., 0.75, 1.45, 0.), 3e-6),
                                   ((17.41, 11.047, 61.918, 1.595, 1.608, 0.006), 6e-5),
                                   ((24.0, 24.0, 100., 2.0, 1.6, 0.0), 8e-5)])
def test_fp32_residual_arithmetic(chk, q, tol):
    """The FP32 forward (LMZ_FP32 residuals) against the FP64 oracle: about
    1e-6 of I, up to ~4e-5 on bright forward lobes (I ~ 40-55)."""
    lib, Feature = chk
    c = grid3((201, 201)) if q[0] > 50 else grid3((40, 40))
    N, coef, _, _ = mie(chk, q[3], q[4], q[5])
    F = Feature()
    F.npix = c.shape[1]
    F.n_m, F.wavelength, F.magnification, F.noise = 1.34, 0.447, 0.048, 0.05
    F.nmax_cap = 400
    p = np.zeros(8)
    p[:6] = q
    p[6] = 1.
    cc = np.ascontiguousarray(c.astype(np.float32))
    h = np.zeros(cc.shape[1], np.float32)
    lib.chk_holo32(ctypes.byref(F), _p(p), _p(coef), N, _p(cc), cc.shape[1], 3,
                   _p(h))
    truth = ref.hologram(c, dict(zip(ref.PARAMS, q)), **INSTRUMENT)
    assert np.abs(h - truth).max() < tol


@pytest.mark.parametrize('seed', [0, 1, 2, 3])
def test_layered_tangents_match_differences(chk, seed):
    """Tangents of layered coefficients in the fitting parametrization
    (a_p scales every radius, n_p + i k_p is the outer layer's index)
    against central differences of the oracle's layered recursion with
    n_max frozen; the k_p tangent against differences in k_p (it is i
    times the n_p tangent by holomorphy in the outer index)."""
    lib, _ = chk
    rng = np.random.default_rng(seed)
    L = 2 + seed
    a_p = rng.uniform(0.5, 3.)
    radii = np.sort(rng.uniform(0.1, 1., L)) * a_p
    radii[-1] = a_p
    nl = rng.uniform(1.33, 2.0, L)
    kl = np.zeros(L)
    kl[-1] = 0.003 * seed
    cap = 400
    dcoef = np.zeros((cap, DCOEF_STRIDE), np.float32)
    N = lib.chk_mie_layered_tangent(_p(radii), _p(nl), _p(kl), L, 1.34, 0.447,
                                    cap, _p(dcoef))
    nn = np.arange(1, N + 1)
    en = (2 * nn + 1) / (nn * (nn + 1))

    def coeffs(r, n, k):
        return ref.mie_coefficients_layered(r, n, k, 1.34, 0.447, nmax=N)[1:]
    for name, base in (('a_p', 0), ('n_p', 9), ('k_p', 18)):
        h = 1e-5
        fd = 0
        for sgn, w in ((1, 8 / 12), (2, -1 / 12)):
            r_p, r_m = radii.copy(), radii.copy()
            n_p, n_m = nl.copy(), nl.copy()
            k_p, k_m = kl.copy(), kl.copy()
            if name == 'a_p':
                r_p *= (a_p + sgn * h) / a_p
                r_m *= (a_p - sgn * h) / a_p
            elif name == 'n_p':
                n_p[-1] += sgn * h
                n_m[-1] -= sgn * h
            else:
                k_p[-1] += sgn * h
                k_m[-1] -= sgn * h
            fd = fd + w * (coeffs(r_p, n_p, k_p) - coeffs(r_m, n_m, k_m))
        fd = fd / h
        if base < 18:
            sg = order_sign(np.arange(1, N + 1))
            dP = sg * (dcoef[:N, base + 3] + 1j * dcoef[:N, base + 4])
            dQ = -sg * (dcoef[:N, base + 6] + 1j * dcoef[:N, base + 7])
        else:
            dP = dcoef[:N, 20] + 1j * dcoef[:N, 21]
            dQ = dcoef[:N, 22] + 1j * dcoef[:N, 23]
        da = dP / (1j ** (nn + 1) * en)
        db = dQ / (1j ** nn * en)
        scale = np.abs(fd).max()
        err = max(np.abs(da - fd[:, 0]).max(), np.abs(db - fd[:, 1]).max())
        print(name, 'layers', L, 'N', N, 'max err / scale', err / scale)
        assert err <= 2e-6 * scale, name
