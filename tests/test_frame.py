"""Frame / Feature drivers on the host (analysis/Frame.py, Feature.py):
boxes become features with the right crops and starting positions, the
mask drops bad pixels, and detection/estimation report that they are
outside the library.  No GPU."""
import numpy as np
import pytest

from pylorenzmie_amd.analysis import Feature, Frame, Hologram
from pylorenzmie_amd.theory import Instrument


def test_bboxes_make_features():
    data = np.random.rand(300, 400) + 0.5
    fr = Frame(data, instrument=Instrument(n_m=1.33))
    fr.bboxes = [((10, 20), 64, 64), ((200, 100), 80, 70)]
    assert len(fr.features) == 2
    f0, f1 = fr.features
    assert f0.shape == (64, 64) and f1.shape == (70, 70)
    assert f1.corner == (200., 100.)
    assert np.array_equal(f1.data, data[100:170, 200:270])
    assert (f1.particle.x_p, f1.particle.y_p) == (235., 135.)
    assert f1.model.instrument is fr.instrument
    fr.data = data
    assert fr.features == [] and fr.bboxes == []


def test_feature_mask_excludes_bad_pixels():
    """Saturated (= max) pixels and non-finite ones, as
    analysis/Feature.py:39-43 (np.max of data holding a NaN is NaN, so
    with a NaN present only the non-finite pixels are caught)."""
    d = np.ones((50, 50))
    d[10, 10] = d[20, 5] = 7.
    f = Feature(Hologram(d))
    assert f.mask.exclude[10, 10] and f.mask.exclude[20, 5]
    assert f.mask.exclude.sum() == 2
    assert not f.mask()[10, 10] and 0.05 < f.mask().mean() < 0.15
    d[3, 4] = np.nan
    d[4, 4] = np.inf
    f = Feature(Hologram(d))
    assert f.mask.exclude[3, 4] and f.mask.exclude[4, 4]
    assert not f.mask()[3, 4] and not f.mask()[4, 4]


def test_detection_is_outside_estimation_is_de():
    fr = Frame(np.ones((64, 64)))
    with pytest.raises(NotImplementedError):
        fr.detect()
    with pytest.raises(NotImplementedError):
        Frame(localizer=object())
    fr.bboxes = [((0, 0), 32, 32)]
    est = fr.features[0].estimator
    assert est.exclude is fr.features[0].mask.exclude
    assert set(est.bounds) == {'z_p', 'a_p', 'n_p'} and est.fraction == 0.02


def test_batched_de_settings_reach_the_device_search(monkeypatch):
    """DEEstimator.estimate_batch on the CPU, with the device search
    (_FrameSearch / lmz_de_run) replaced by a stub: scipy's population
    floor max(5, popsize * D), the settings it passes to lmz_de_settings,
    each estimator's model receiving its own result, and the settings and
    estimators it refuses.  The search itself is tested on the host build
    of lmz_de.h (tests/test_de_step.py) and on the GPU
    (tests/test_gpu_estimator.py)."""
    import importlib
    from pylorenzmie_amd import _lmz
    from pylorenzmie_amd.analysis import DEEstimator
    from pylorenzmie_amd.theory import LorenzMie
    de_mod = importlib.import_module('pylorenzmie_amd.analysis.DEEstimator')
    seen = {}

    class Stub(de_mod._FrameSearch):
        def __init__(self, probs, npop, lo, hi):
            seen.update(npop=npop, lo=lo, hi=hi, cols=probs[0]['cols'])
            self.nsearch, self.npop, self.ndim = len(probs), npop, lo.shape[1]
            self.cols = probs[0]['cols']

        def run(self, st):
            seen['st'] = st
            x = 0.5 * (seen['lo'] + seen['hi']) + np.arange(self.nsearch)[:, None]
            return x, np.zeros(self.nsearch), np.zeros(self.nsearch, np.int32)

    monkeypatch.setattr(de_mod, '_FrameSearch', Stub)
    bounds = {'z_p': (50., 400.), 'a_p': (0.3, 2.0), 'n_p': (1.35, 1.8)}
    ests = [DEEstimator(model=LorenzMie(), seed=5, bounds=dict(bounds))
            for _ in range(3)]
    for e in ests:
        e.settings.update(tol=1e-6, atol=0.5, mutation=(0.4, 0.9),
                          recombination=0.8, maxiter=77)
    holos = [Hologram(np.random.default_rng(i).random((20, 20))) for i in range(3)]
    out = de_mod.estimate_batch(ests, holos)
    st = seen['st']
    assert seen['npop'] == st.npop == 30 and st.ndim == 3
    assert list(st.cols)[:3] == seen['cols'] == [2, 3, 4]
    assert (st.tol, st.atol, st.mutation_lo, st.mutation_hi, st.recombination,
            st.maxiter, st.seed) == (1e-6, 0.5, 0.4, 0.9, 0.8, 77, 5)
    for i, (res, est) in enumerate(zip(out, ests)):
        assert res['z_p'] == pytest.approx(225. + i)
        assert est.model.particle.z_p == res['z_p']
    with pytest.raises(ValueError):
        de_mod.estimate_batch([ests[0], DEEstimator(model=LorenzMie(), popsize=5)],
                              holos[:2])
    # scipy's population floor max(5, popsize * D): one bound, popsize 2
    small = [DEEstimator(model=LorenzMie(), seed=1, popsize=2,
                         bounds={'z_p': (50., 400.)}) for _ in range(3)]
    de_mod.estimate_batch(small, holos)
    assert seen['npop'] == 5 and seen['st'].ndim == 1
    # settings the batched step cannot honour are refused, others differ
    odd = DEEstimator(model=LorenzMie(), seed=1, popsize=2,
                      bounds={'z_p': (50., 400.)})
    odd.settings['strategy'] = 'rand1bin'
    with pytest.raises(ValueError):
        de_mod.estimate_batch([odd], holos[:1])
    other = DEEstimator(model=LorenzMie(), seed=1, popsize=2,
                        bounds={'z_p': (50., 400.)})
    other.settings['tol'] = 0.5
    with pytest.raises(ValueError):
        de_mod.estimate_batch([small[0], other], holos[:2])
    from pylorenzmie_amd.analysis import RadialEstimator
    with pytest.raises(TypeError):
        de_mod.estimate_batch([RadialEstimator(model=LorenzMie())], holos[:1])
    assert _lmz.DESettings is not None
