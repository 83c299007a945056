"""The differential-evolution step of the device search
(pylorenzmie_amd/csrc/lmz_de.h, run by lmz_de.hip's kernels) compiled for
the host and driven in the kernels' order, on the CPU.

The algorithm is scipy's differential_evolution with best1bin, deferred
updating and a Latin-hypercube start, the search the reference's
DEEstimator runs per feature (analysis/DEEstimator.py:119-173).  Checks:
the start is a Latin hypercube; the difference members are distinct and
never the member itself, and uniform over the others; the forced
crossover dimension always comes from the mutant; out-of-range
coordinates are redrawn inside [0, 1]; the stopping test is numpy's
std <= atol + tol |mean| (and waits while an energy is infinite); the
scale stays in the mutation range; and whole searches on bowls with
different minima land at each minimum, reproducibly for a seed, while a
flat energy stops at the first test."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, 'tests', 'native', 'liblmz_hostcheck.so')
D = ctypes.c_double
I = ctypes.c_int
P = ctypes.c_void_p
U64 = ctypes.c_uint64


@pytest.fixture(scope='module')
def chk():
    if not os.path.exists(LIB):
        pytest.skip('host checker not built (run __graft_entry__.build())')
    lib = ctypes.CDLL(LIB)
    lib.chk_de_lhs.argtypes = [U64, I, I, I, P]
    lib.chk_de_trial.argtypes = [U64, I, I, I, I, I, P, I, D, D, D, P]
    lib.chk_de_trial.restype = D
    lib.chk_de_converged.argtypes = [P, I, D, D]
    lib.chk_de_converged.restype = I
    lib.chk_de_bowl.argtypes = [U64, I, I, I, I, D, D, D, D, D, P, P, P, I,
                                P, P, P]
    lib.chk_de_settings_layout.argtypes = [P]
    return lib


def _p(a):
    return a.ctypes.data_as(P)


def test_settings_layout_matches_binding(chk):
    from pylorenzmie_amd import _lmz
    out = np.zeros(4, np.int32)
    chk.chk_de_settings_layout(_p(out))
    S = _lmz.DESettings
    assert out.tolist() == [ctypes.sizeof(S), S.npop.offset, S.tol.offset,
                            S.seed.offset]


def test_latin_hypercube_start(chk):
    for npop, ndim in ((5, 1), (30, 3), (70, 7)):
        u = np.zeros((npop, ndim))
        chk.chk_de_lhs(11, 2, npop, ndim, _p(u))
        seg = np.floor(u * npop).astype(int)
        for j in range(ndim):
            assert sorted(seg[:, j]) == list(range(npop))     # one per segment
        # another search or seed: another hypercube
        v = np.zeros_like(u)
        chk.chk_de_lhs(11, 3, npop, ndim, _p(v))
        assert not np.array_equal(u, v)


def test_trial_draws(chk):
    """Member i's trial: for the forced dimension (and every crossed one)
    best + F (u_r0 - u_r1) with r0 != r1 != i, else u_i; values outside
    [0, 1] redrawn inside.  The difference pairs come out uniform over the
    ordered pairs of other members."""
    npop, ndim = 6, 2
    best = 4
    # members on a line make r0 - r1 recoverable; F = 1 and CR = 1 cross
    # every dimension: trial = u_best + u_r0 - u_r1, inside [0, 1] here
    k = np.arange(npop, dtype=float)
    base = np.stack([0.3 + 0.01 * k, 0.3 + 0.05 * k], axis=1)
    count = {}
    for gen in range(1, 4001):
        i = gen % npop
        t = np.zeros(ndim)
        F = chk.chk_de_trial(7, gen, 0, i, npop, ndim, _p(base), best, 1.0, 1.0,
                             1.0, _p(t))
        assert F == 1.0
        d = np.round((t - base[best]) * 100., 6)
        assert d[1] == pytest.approx(5. * d[0])
        diff = int(round(d[0]))
        assert diff != 0
        count[(i, diff)] = count.get((i, diff), 0) + 1
    # for each i the differences r0 - r1 over ordered pairs of the 5 others
    for i in range(npop):
        others = [k for k in range(npop) if k != i]
        want = {}
        for a in others:
            for b in others:
                if a != b:
                    want[a - b] = want.get(a - b, 0) + 1
        got = {dd: c for (ii, dd), c in count.items() if ii == i}
        assert set(got) == set(want), i
        n = sum(got.values())
        for dd, w in want.items():
            assert abs(got[dd] / n - w / 20.) < 0.08, (i, dd)


def test_crossover_and_redraw(chk):
    npop, ndim = 8, 7
    rng = np.random.default_rng(0)
    u = rng.random((npop, ndim))
    forced_from_mutant = 0
    for gen in range(1, 2001):
        t = np.zeros(ndim)
        # CR = 0: only the forced dimension crosses
        F = chk.chk_de_trial(3, gen, 1, 2, npop, ndim, _p(u), 5, 0.5, 1.0, 0.0,
                             _p(t))
        assert 0.5 <= F < 1.0
        same = np.isclose(t, u[2])
        assert same.sum() >= ndim - 1
        forced_from_mutant += int(not same.all())
        assert np.all((t >= 0.) & (t <= 1.))
    assert forced_from_mutant > 1900
    # a scale of 2 on spread members pushes coordinates out: all redrawn in range
    for gen in range(1, 200):
        t = np.zeros(ndim)
        chk.chk_de_trial(3, gen, 1, 2, npop, ndim, _p(u), 5, 2.0, 2.0, 1.0, _p(t))
        assert np.all((t >= 0.) & (t <= 1.))


def test_convergence_test_is_numpys(chk):
    rng = np.random.default_rng(1)
    for _ in range(200):
        n = int(rng.integers(5, 40))
        E = rng.normal(10., rng.choice([1e-4, 1e-2, 1.]), n)
        tol, atol = rng.choice([1e-3, 1e-2, 0.1]), rng.choice([0., 1e-3])
        want = np.std(E) <= atol + tol * abs(np.mean(E))
        assert chk.chk_de_converged(_p(E), n, tol, atol) == int(want)
    E = np.ones(10)
    E[3] = np.inf
    assert chk.chk_de_converged(_p(E), 10, 1.0, 1.0) == 0


def run_bowl(chk, targets, scales, bounds, seed, npop=30, maxiter=1000,
             tol=1e-6, atol=0., flat=0):
    ns, nd = targets.shape
    best = np.zeros((ns, nd))
    energy = np.zeros(ns)
    gens = np.zeros(ns, np.int32)
    chk.chk_de_bowl(seed, ns, npop, nd, maxiter, tol, atol, 0.5, 1.0, 0.7,
                    _p(np.ascontiguousarray(bounds)), _p(targets), _p(scales),
                    flat, _p(best), _p(energy), _p(gens))
    return best, energy, gens


def test_bowls_find_each_minimum(chk):
    targets = np.array([[120., 0.8, 1.45], [300., 1.6, 1.60], [75., 0.5, 1.38],
                        [390., 1.95, 1.79]])
    scales = np.tile([50., 0.2, 0.05], (4, 1))
    bounds = np.tile(np.array([[50., 400.], [0.3, 2.0], [1.35, 1.8]]), (4, 1, 1))
    best, energy, gens = run_bowl(chk, targets, scales, bounds, seed=5)
    err = np.abs(best - targets) / scales
    print(err.max(axis=1), gens)
    assert np.all(err < 0.01)
    assert np.all(energy < 1. + 1e-4) and np.all(gens < 1000)
    again = run_bowl(chk, targets, scales, bounds, seed=5)
    assert np.array_equal(again[0], best) and np.array_equal(again[2], gens)
    other = run_bowl(chk, targets, scales, bounds, seed=6)
    assert not np.array_equal(other[2], gens)
    assert np.all(np.abs(other[0] - targets) / scales < 0.01)
    # the default tol 0.01 stops earlier, near each minimum; seeds differ
    loose = [run_bowl(chk, targets, scales, bounds, seed=sd, tol=0.01)
             for sd in (5, 6)]
    assert not np.array_equal(loose[0][0], loose[1][0])
    for b, _, g in loose:
        print(np.abs(b - targets).max(axis=0), g)
        assert np.all(np.abs(b - targets) / scales < 0.3)
        assert np.all(g < gens)


def test_flat_energy_stops_at_first_test_and_maxiter_caps(chk):
    targets = np.zeros((2, 2))
    scales = np.ones((2, 2))
    bounds = np.tile(np.array([[0., 1.], [0., 1.]]), (2, 1, 1))
    _, _, gens = run_bowl(chk, targets, scales, bounds, seed=1, flat=1)
    assert gens.tolist() == [1, 1]
    _, _, gens = run_bowl(chk, targets, scales, bounds, seed=1, maxiter=3,
                          tol=0.)
    assert gens.tolist() == [3, 3]
    _, _, gens = run_bowl(chk, targets, scales, bounds, seed=1, maxiter=0)
    assert gens.tolist() == [0, 0]
