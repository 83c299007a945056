"""Calls from worker threads, as the reference's lmtool makes them
(Optimizer.optimize inside a QThread, lmtool/FitWidget.py:97-103, 148):
a fit run in a threading.Thread, and two models fitted and evaluated on
two threads at once, give exactly the results of the same calls made one
after another on the main thread."""
import threading

import numpy as np
import pytest

from conftest import INSTRUMENT

pytestmark = pytest.mark.gpu

VARS6 = ['x_p', 'y_p', 'z_p', 'a_p', 'n_p', 'k_p']
TIGHT = {'ftol': 1e-10, 'xtol': 1e-10, 'gtol': 1e-10}


@pytest.fixture(scope='module')
def api():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    from pylorenzmie_amd import theory, analysis
    return theory, analysis


def _job(api, fx, name):
    """Fit one golden feature from its start, then evaluate the hologram
    and Jacobian at the solution; everything the caller gets back."""
    theory, analysis = api
    data = fx[f'{name}_data']
    ny, nx = data.shape
    inst = theory.Instrument(noise=float(fx[f'{name}_sigma']), **INSTRUMENT)
    s = theory.Sphere(**dict(zip(VARS6, fx[f'{name}_start'])))
    m = theory.LorenzMie(coordinates=theory.LorenzMie.meshgrid((ny, nx)),
                         particle=s, instrument=inst)
    opt = analysis.Optimizer(model=m)
    opt.variables = VARS6
    opt.settings.update(TIGHT)
    res = opt.optimize(analysis.Hologram(data.astype(float)))
    holo = m.hologram()
    jac = m.jac()
    return (res.to_numpy(dtype=float), holo.copy(),
            np.stack([jac[p] for p in VARS6]))


def _same(a, b):
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def _in_threads(fns):
    out = [None] * len(fns)
    err = []
    start = threading.Barrier(len(fns))

    def run(i):
        try:
            start.wait()
            out[i] = fns[i]()
        except BaseException as e:      # reported on the main thread
            err.append(e)
    ts = [threading.Thread(target=run, args=(i,)) for i in range(len(fns))]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=300)
    assert not any(t.is_alive() for t in ts), 'worker thread did not finish'
    if err:
        raise err[0]
    return out


def test_fit_in_worker_thread_equals_main_thread(api, fits_golden):
    fx, _ = fits_golden
    want = _job(api, fx, 'cfg0')
    (got,) = _in_threads([lambda: _job(api, fx, 'cfg0')])
    _same(got, want)


def test_two_threads_at_once_equal_sequential(api, fits_golden):
    fx, _ = fits_golden
    names = ['cfg0', 'crop03']
    want = [_job(api, fx, n) for n in names]
    for _ in range(2):
        got = _in_threads([lambda n=n: _job(api, fx, n) for n in names])
        for g, w in zip(got, want):
            _same(g, w)


def test_batched_fits_on_two_threads(api, fits_golden):
    """Two LMBatch fits (each with its own plan, stream and graph) run
    concurrently from two threads."""
    from pylorenzmie_amd import _lmz
    from pylorenzmie_amd.analysis.lm import LMBatch, LMSettings
    theory, _ = api
    fx, _ = fits_golden
    names = [f'crop{i:02d}' for i in range(12)]
    groups = [names[:6], names[6:]]

    def fit(group):
        crops, p0 = [], []
        inst = theory.Instrument(**INSTRUMENT)
        for n in group:
            d = fx[f'{n}_data']
            st = fx[f'{n}_start']
            crops.append(dict(data=d, width=d.shape[1], origin=(0., 0., 0.),
                              noise=float(fx[f'{n}_sigma']),
                              nmax_cap=_lmz.nmax_cap(st[3], st[4], st[5],
                                                     inst.n_m, inst.wavelength)))
            p0.append(list(st) + [1.])
        lm = LMBatch(crops, inst)
        lm.load(np.array(p0))
        lm.run(LMSettings(), VARS6 + ['alpha'])
        out = lm.fetch()
        lm.close()
        return out

    want = [fit(g) for g in groups]
    got = _in_threads([lambda g=g: fit(g) for g in groups])
    for (gp, gr), (wp, wr) in zip(got, want):
        np.testing.assert_array_equal(gp, wp)
        np.testing.assert_array_equal(gr, wr)
