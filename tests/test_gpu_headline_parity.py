"""Fit parity on the headline workload: the batched device LM against the
FP64 oracle's MINPACK fits and the reference's own Optimizer fits from the
same starts (tests/golden/make_golden_headline.py writes the fixtures).

  * 32 features of the headline batch (201 x 201, sigma = 0.05, 7
    parameters): the oracle's fits (MINPACK lmdif: forward-difference
    Jacobian, tolerances 1e-12), and the same MINPACK with an exact
    Jacobian (the oracle's 5-point differences);
  * the same features with alpha held at 1 (6 parameters): the reference's
    analysis/Optimizer.py fits, identical to the oracle's;
  * 25 hard small crops of the uneven-crop workload (side <= 80 px,
    a_p > 2.4 um), where chi2 has neighbouring minima.

The accuracy gate: every fitted parameter within 1e-4 relative of the
reference fit (k_p 1e-5 absolute, its truth being near 0).  The device LM
runs the benchmark's tolerances (ftol = xtol = 1e-9, gtol = 1e-10) with the
reference's evaluation budget (max_nfev 2000: 500 steps here) and MINPACK's
step rule (lmz_lmstep.h).

With lmdif's own forward-difference Jacobian (LMSettings.fd_step = 1e-5,
the reference's diff_step) the device runs the reference's algorithm and
passes the gate on every headline feature.  With the analytic Jacobian it
passes on 31 of 32, and on all 32 against MINPACK given an exact Jacobian:
where it leaves the lmdif oracle's minimum, MINPACK with the exact
Jacobian goes to the same other minimum, so the difference is the
Jacobian rule's, not the device's arithmetic."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, INSTRUMENT

pytestmark = pytest.mark.gpu

PARAMS = ['x_p', 'y_p', 'z_p', 'a_p', 'n_p', 'k_p', 'alpha']


@pytest.fixture(scope='module')
def fx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return np.load(os.path.join(GOLDEN, 'headline.npz'))


def settings(**kw):
    """The headline benchmark's tolerances (bench.py) with the reference's
    evaluation budget in place of the benchmark's 50-step cap."""
    from pylorenzmie_amd.analysis.lm import LMSettings
    return LMSettings(**dict(dict(max_iter=500, ftol=1e-9, xtol=1e-9,
                                  gtol=1e-10), **kw))


def run_lm(datas, sides, starts, sigmas, vary, st=None):
    from pylorenzmie_amd import _lmz
    from pylorenzmie_amd.analysis.lm import LMBatch
    from pylorenzmie_amd.theory import Instrument
    inst = Instrument(**INSTRUMENT)
    crops = []
    for d, s, p, sg in zip(datas, sides, starts, sigmas):
        crops.append(dict(data=np.asarray(d, np.float32).reshape(s, s),
                          width=s, origin=(0., 0., 0.), noise=float(sg),
                          nmax_cap=_lmz.nmax_cap(p[3], p[4], p[5], inst.n_m,
                                                 inst.wavelength)))
    lm = LMBatch(crops, inst)
    lm.load(np.asarray(starts))
    lm.run(st or settings(), vary)
    p, res = lm.fetch()
    lm.close()
    status = res[:, 38].astype(int)
    assert not np.any(status & 32)
    return p[:, :7], res


def gate(p, ref_p, nvar):
    """Per feature: every parameter within 1e-4 relative (k_p 1e-5
    absolute) of ref_p."""
    d = np.abs(p[:, :nvar] - ref_p[:, :nvar])
    rel = d / np.abs(ref_p[:, :nvar])
    rel[:, 5] = d[:, 5] * 10.                 # k_p: 1e-5 absolute
    return rel.max(axis=1) <= 1e-4, rel


def report(name, chi2, ref_chi2, ok):
    rel = (chi2 - ref_chi2) / ref_chi2
    for i in range(len(chi2)):
        print(f'{name} {i:2d}: chi2 {chi2[i]:.6f} ref {ref_chi2[i]:.6f} '
              f'rel {rel[i]:+.2e} gate {"ok" if ok[i] else "--"}')
    print(f'{name}: {ok.sum()}/{len(ok)} within the gate')
    return rel


def headline(fx):
    n = len(fx['head_truth'])
    return fx['head_data'], [201] * n, fx['head_start'], [float(fx['head_sigma'])] * n


def small(fx):
    sides = fx['small_side']
    off = fx['small_offsets']
    datas = [fx['small_data'][off[k]:off[k + 1]] for k in range(len(sides))]
    return datas, sides, fx['small_start'], fx['small_sigma']


def test_headline_matches_oracle(fx):
    p, res = run_lm(*headline(fx), PARAMS)
    ok, _ = gate(p, fx['oracle7_values'], 7)
    rel = report('head7', res[:, 35], fx['oracle7_chi2'], ok)
    okj, _ = gate(p, fx['oracle7j_values'], 7)
    relj = report('head7 exact-J', res[:, 35], fx['oracle7j_chi2'], okj)
    assert ok.sum() >= 31
    # MINPACK given an exact Jacobian (the oracle's 5-point differences)
    # lands where the device's analytic-Jacobian LM does on every feature
    assert okj.all()


def test_headline_fd_jacobian_matches_oracle(fx):
    """lmdif's forward differences inside the device LM: the reference's
    algorithm end to end."""
    p, res = run_lm(*headline(fx), PARAMS, settings(fd_step=1e-5))
    ok, _ = gate(p, fx['oracle7_values'], 7)
    report('head7 fd', res[:, 35], fx['oracle7_chi2'], ok)
    assert ok.all()


def test_headline_alpha_held_matches_reference(fx):
    """alpha held at 1: the reference's own Optimizer fits (its model has
    no alpha) and the oracle's, which are identical."""
    np.testing.assert_array_equal(fx['oracle6_values'], fx['reference6_values'])
    data, sides, start, sig = headline(fx)
    start = start.copy()
    start[:, 6] = 1.
    p, res = run_lm(data, sides, start, sig, PARAMS[:6])
    assert np.all(p[:, 6] == 1.)
    ok, _ = gate(p, fx['reference6_values'], 6)
    report('ref6', res[:, 35], fx['reference6_chi2'], ok)
    assert ok.all()
    # uncertainties from J^T J at the solution, the reference's SVD recipe
    from pylorenzmie_amd.analysis.lm import stats_from_jtj
    for i in range(len(p)):
        _, dp = stats_from_jtj(res[i, :28], res[i, 35], 201 * 201,
                               list(range(6)))
        np.testing.assert_allclose(dp, fx['reference6_uncertainties'][i],
                                   rtol=0.02)


def test_small_crops_against_oracle(fx):
    """Hard small crops with neighbouring minima: the analytic-Jacobian fit
    lands in the lmdif oracle's minimum or in exact-Jacobian MINPACK's on
    21-24 of 25 and in a minimum higher than both of theirs (to 1e-6
    relative) on one to three.  These crops are chaotic: Jacobians that
    agree to FP32 rounding move single fits between neighbouring minima.
    Measured with four of them (the order loop at every pixel, and radial
    tables of the sums at spacings h, 3h/4 and h/2): 22, 21, 22 and 24 of
    25 in one of the two minima (exact-Jacobian MINPACK's alone), higher
    than both on 1, 3, 2 and 1.  The bounds are the worst of those: at
    least 21 and at most three.  With lmdif's differences it lands in the
    oracle's minimum on 24 and within 1e-6 of its chi2 on all."""
    p, res = run_lm(*small(fx), PARAMS)
    ok, _ = gate(p, fx['small7_values'], 7)
    rel = report('small', res[:, 35], fx['small7_chi2'], ok)
    okj, _ = gate(p, fx['small7j_values'], 7)
    relj = report('small exact-J', res[:, 35], fx['small7j_chi2'], okj)
    worse = ~ok & ~okj & (rel > 1e-6) & (relj > 1e-6)
    print('small: higher than both MINPACK fits on', np.flatnonzero(worse))
    assert (ok | okj).sum() >= 21
    assert worse.sum() <= 3
    p, res = run_lm(*small(fx), PARAMS, settings(fd_step=1e-5))
    ok, _ = gate(p, fx['small7_values'], 7)
    rel = report('small fd', res[:, 35], fx['small7_chi2'], ok)
    assert ok.sum() >= 24
    assert np.all(ok | (rel <= 1e-6))
