"""The LM's per-feature step logic (pylorenzmie_amd/csrc/lmz_lmstep.h, run
by k_lm_solve / k_lm_accept on the device) compiled for the host and driven
here as the kernels drive it, against scipy.optimize.least_squares
(method='lm', x_scale='jac': MINPACK lmder), the library the reference's
Optimizer.optimize calls (analysis/Optimizer.py:221).

With LMZ_DAMPING_MINPACK and the same analytic Jacobian, every trial point
the device rule evaluates is the point MINPACK evaluates (to rounding: QR
there, Cholesky of the normal equations here), and the two stop at the same
point for the same reason.  Test problems: Rosenbrock (rejected steps and
lmpar iterations from the standard start), the Kowalik-Osborne and Box 3-D
problems of the MINPACK test set (More, Garbow, Hillstrom, ACM TOMS 7
(1981)), and a damped-sine fit with five parameters."""
import ctypes
import os

import numpy as np
import pytest
from scipy.optimize import least_squares

from conftest import ROOT

LIB = os.path.join(ROOT, 'tests', 'native', 'liblmz_hostcheck.so')
FTOL, XTOL, GTOL = 1, 2, 4


@pytest.fixture(scope='module')
def chk():
    if not os.path.exists(LIB):
        pytest.skip('host checker not built (run __graft_entry__.build())')
    lib = ctypes.CDLL(LIB)
    P, D = ctypes.c_void_p, ctypes.c_double
    lib.chk_lm_state_bytes.restype = ctypes.c_int
    lib.chk_lm_init.argtypes = [P, D]
    lib.chk_lm_set_chi2.argtypes = [P, D]
    lib.chk_lm_fresh.argtypes = [P, P, P, P]
    lib.chk_lm_solve.argtypes = [P, P, P, P]
    lib.chk_lm_accept.argtypes = [P, P, D, P]
    lib.chk_lm_read.argtypes = [P, P]
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def device_rule(chk, fun, jac, x0, ftol, xtol, gtol, damping=0,
                max_iter=500, lambda0=1e-3):
    """The device LM's loop for one feature: Jacobian reduction when the
    last step was accepted, solve, trial residual, accept/reject.
    Returns (x, status bits, trial points, number of Jacobians)."""
    from pylorenzmie_amd import _lmz
    n = len(x0)
    assert n <= 7
    s = _lmz.LMSettings()
    s.max_iter, s.check_every = max_iter, 1
    s.ftol, s.xtol, s.gtol, s.lambda0 = ftol, xtol, gtol, lambda0
    for i in range(7):
        s.vary[i] = 1 if i < n else 0
    s.damping = damping
    st = ctypes.create_string_buffer(chk.chk_lm_state_bytes())
    chk.chk_lm_init(st, lambda0)
    p = np.zeros(8)
    p[:n] = x0
    pt = np.zeros(8)
    r = fun(p[:n])
    chk.chk_lm_set_chi2(st, float(r @ r))
    trials, njev = [p[:n].copy()], 0
    out = np.zeros(7)
    need_j = True
    iu = np.triu_indices(7)
    while True:
        if need_j:
            J = np.zeros((r.size, 7))
            J[:, :n] = jac(p[:n])
            red = np.concatenate([(J.T @ J)[iu], J.T @ r, [r @ r]])
            njev += 1
            if chk.chk_lm_fresh(st, _ptr(red), ctypes.byref(s), _ptr(p)):
                break
        assert chk.chk_lm_solve(st, ctypes.byref(s), _ptr(p), _ptr(pt))
        rt = fun(pt[:n])
        trials.append(pt[:n].copy())
        if chk.chk_lm_accept(st, ctypes.byref(s), float(rt @ rt), _ptr(pt)):
            p, r = pt.copy(), rt
        chk.chk_lm_read(st, _ptr(out))
        if int(out[3]) != 0:
            break
        need_j = bool(out[6])
    chk.chk_lm_read(st, _ptr(out))
    return p[:n], int(out[3]), np.array(trials), njev


def minpack(fun, jac, x0, ftol, xtol, gtol):
    """scipy's MINPACK lmder, recording every point fun is evaluated at."""
    pts = []

    def f(x):
        pts.append(np.array(x, float))
        return fun(x)
    res = least_squares(f, x0, jac=jac, method='lm', x_scale='jac',
                        ftol=ftol, xtol=xtol, gtol=gtol, max_nfev=5000)
    # least_squares evaluates x0 once itself before MINPACK does
    while len(pts) > 1 and np.array_equal(pts[0], pts[1]):
        pts.pop(0)
    return res, np.array(pts)


# scipy's status -> the device's status bits (MINPACK info 4, 1, 2, 3)
STATUS = {1: GTOL, 2: FTOL, 3: XTOL, 4: FTOL | XTOL}


def rosenbrock():
    fun = lambda x: np.array([10. * (x[1] - x[0] ** 2), 1. - x[0]])
    jac = lambda x: np.array([[-20. * x[0], 10.], [-1., 0.]])
    return fun, jac, np.array([-1.2, 1.])


def kowalik_osborne():
    y = np.array([0.1957, 0.1947, 0.1735, 0.1600, 0.0844, 0.0627, 0.0456,
                  0.0342, 0.0323, 0.0235, 0.0246])
    u = np.array([4., 2., 1., 0.5, 0.25, 0.167, 0.125, 0.1, 0.0833, 0.0714,
                  0.0625])

    def fun(x):
        return y - x[0] * (u * u + u * x[1]) / (u * u + u * x[2] + x[3])

    def jac(x):
        num = u * u + u * x[1]
        den = u * u + u * x[2] + x[3]
        return np.column_stack([-num / den, -x[0] * u / den,
                                x[0] * num * u / den ** 2,
                                x[0] * num / den ** 2])
    return fun, jac, np.array([0.25, 0.39, 0.415, 0.39])


def box3d():
    t = 0.1 * np.arange(1, 11)

    def fun(x):
        return (np.exp(-t * x[0]) - np.exp(-t * x[1])
                - x[2] * (np.exp(-t) - np.exp(-10. * t)))

    def jac(x):
        return np.column_stack([-t * np.exp(-t * x[0]), t * np.exp(-t * x[1]),
                                -(np.exp(-t) - np.exp(-10. * t))])
    return fun, jac, np.array([0., 10., 20.])


def damped_sine():
    t = np.linspace(0., 10., 80)
    truth = np.array([2.0, 0.3, 1.5, 1.7, 0.4])
    rng = np.random.default_rng(5)

    def model(x):
        return x[0] * np.exp(-x[1] * t) * np.sin(x[3] * t + x[4]) + x[2]
    y = model(truth) + 0.02 * rng.standard_normal(t.size)

    def jac(x):
        e = np.exp(-x[1] * t)
        sn, cs = np.sin(x[3] * t + x[4]), np.cos(x[3] * t + x[4])
        return np.column_stack([e * sn, -x[0] * t * e * sn, np.ones_like(t),
                                x[0] * e * cs * t, x[0] * e * cs])
    return (lambda x: model(x) - y), jac, np.array([1.5, 0.2, 1.0, 1.6, 0.2])


PROBLEMS = {'rosenbrock': rosenbrock, 'kowalik_osborne': kowalik_osborne,
            'box3d': box3d, 'damped_sine': damped_sine}


@pytest.mark.parametrize('name', sorted(PROBLEMS))
@pytest.mark.parametrize('tol', [1e-8, 1e-12])
def test_minpack_rule_follows_minpack(chk, name, tol):
    fun, jac, x0 = PROBLEMS[name]()
    res, pts = minpack(fun, jac, x0, tol, tol, tol)
    x, status, trials, njev = device_rule(chk, fun, jac, x0, tol, tol, tol)
    print(f'{name}: scipy nfev {res.nfev} njev {res.njev} status '
          f'{res.status}; device trials {len(trials)} njev {njev} status '
          f'{status}')
    # the same sequence of evaluated points; on a zero-residual problem
    # the last step can differ by rounding (QR there, Cholesky here), which
    # decides whether chi2 lands on exactly 0 (gtol) or one step later
    m = min(len(trials), len(pts))
    scale = np.maximum(np.abs(pts[:m]), 1.)
    dev = np.abs(trials[:m] - pts[:m]) / scale
    assert dev.max() <= 1e-10, dev.max()
    r = fun(x)
    if len(trials) == len(pts):
        assert njev == res.njev
        assert status == STATUS[res.status]
    else:
        assert abs(len(trials) - len(pts)) <= 1 and r @ r <= 1e-28
        assert abs(njev - res.njev) <= 1 and status != 0
    np.testing.assert_allclose(x, res.x, rtol=1e-9, atol=1e-12)


def test_minpack_rule_radius_and_parameter(chk):
    """lmpar's contract at every solve: either par = 0 and ||D dp|| <= 1.1
    delta, or | ||D dp|| - delta | <= 0.1 delta (MINPACK lmpar), checked on
    the Rosenbrock run through the state after each solve."""
    from pylorenzmie_amd import _lmz
    fun, jac, x0 = rosenbrock()
    s = _lmz.LMSettings()
    s.max_iter, s.check_every = 100, 1
    s.ftol = s.xtol = s.gtol = 1e-12
    s.vary[0] = s.vary[1] = 1
    st = ctypes.create_string_buffer(chk.chk_lm_state_bytes())
    chk.chk_lm_init(st, 1e-3)
    p = np.zeros(8)
    p[:2] = x0
    pt = np.zeros(8)
    r = fun(p[:2])
    chk.chk_lm_set_chi2(st, float(r @ r))
    out = np.zeros(7)
    iu = np.triu_indices(7)
    need_j, checked, D = True, 0, None
    for _ in range(100):
        if need_j:
            J = np.zeros((2, 7))
            J[:, :2] = jac(p[:2])
            red = np.concatenate([(J.T @ J)[iu], J.T @ r, [r @ r]])
            cn = np.linalg.norm(J[:, :2], axis=0)
            D = cn if D is None else np.maximum(D, cn)
            if chk.chk_lm_fresh(st, _ptr(red), ctypes.byref(s), _ptr(p)):
                break
        chk.chk_lm_read(st, _ptr(out))
        delta_before = out[2]
        assert chk.chk_lm_solve(st, ctypes.byref(s), _ptr(p), _ptr(pt))
        chk.chk_lm_read(st, _ptr(out))
        par, delta = out[1], out[2]
        dn = np.linalg.norm(D * (pt[:2] - p[:2]))
        first = out[5] == 1            # the first outer iteration clips delta
        ref = delta_before if first else delta
        if par == 0.:
            assert dn <= 1.1 * ref * (1 + 1e-12)
        else:
            assert abs(dn - ref) <= 0.1 * ref * (1 + 1e-12)
        checked += 1
        rt = fun(pt[:2])
        if chk.chk_lm_accept(st, ctypes.byref(s), float(rt @ rt), _ptr(pt)):
            p, r = pt.copy(), rt
        chk.chk_lm_read(st, _ptr(out))
        if int(out[3]):
            break
        need_j = bool(out[6])
    assert checked > 5
    np.testing.assert_allclose(p[:2], [1., 1.], atol=1e-6)


def test_nielsen_rule_still_converges(chk):
    for name in sorted(PROBLEMS):
        fun, jac, x0 = PROBLEMS[name]()
        res, _ = minpack(fun, jac, x0, 1e-12, 1e-12, 1e-12)
        x, status, trials, _ = device_rule(chk, fun, jac, x0, 1e-12, 1e-12,
                                           1e-12, damping=1)
        r = fun(x)
        print(name, 'nielsen trials', len(trials), 'cost', r @ r,
              'minpack', res.cost * 2)
        assert status & (FTOL | XTOL | GTOL)
        assert r @ r <= 2 * res.cost * (1 + 1e-6) + 1e-20
