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


def test_batched_de_finds_each_minimum(monkeypatch):
    """DEEstimator.estimate_batch, the search itself, on the CPU: with the
    device objective replaced by a bowl whose minimum differs per feature,
    every search lands at its own minimum (the best1bin step, the index
    draws and the per-search stopping rule), and each estimator's model
    receives it."""
    import importlib
    from pylorenzmie_amd.analysis import DEEstimator
    from pylorenzmie_amd.theory import LorenzMie
    de_mod = importlib.import_module('pylorenzmie_amd.analysis.DEEstimator')
    targets = np.array([[120., 0.8, 1.45], [300., 1.6, 1.60], [75., 0.5, 1.38]])
    scale = np.array([50., 0.2, 0.05])

    class Bowl:
        def __init__(self, probs, npop):
            self.npop = npop

        def __call__(self, rows):
            x = rows[:, [2, 3, 4]].reshape(len(targets), self.npop, 3)
            return (((x - targets[:, None, :]) / scale) ** 2).sum(-1).ravel()

    monkeypatch.setattr(de_mod, '_FramePopulation', Bowl)
    ests, holos = [], []
    for i in range(3):
        m = LorenzMie()
        ests.append(DEEstimator(model=m, seed=5,
                                bounds={'z_p': (50., 400.), 'a_p': (0.3, 2.0),
                                        'n_p': (1.35, 1.8)}))
        holos.append(Hologram(np.random.default_rng(i).random((20, 20))))
    for e in ests:
        e.settings['tol'] = 1e-6
    out = de_mod.estimate_batch(ests, holos)
    for i, (res, est) in enumerate(zip(out, ests)):
        got = np.array([res['z_p'], res['a_p'], res['n_p']])
        err = np.abs(got - targets[i]) / scale
        print(i, got, err)
        assert np.all(err < 0.05), i
        assert est.model.particle.z_p == res['z_p']
    with pytest.raises(ValueError):
        de_mod.estimate_batch([ests[0], DEEstimator(model=LorenzMie(), popsize=5)],
                              holos[:2])
    # scipy's population floor max(5, popsize * D): one bound, popsize 2
    small = [DEEstimator(model=LorenzMie(), seed=1, popsize=2,
                         bounds={'z_p': (50., 400.)}) for _ in range(3)]
    seen = []

    class Line:
        def __init__(self, probs, npop):
            seen.append(npop)
            self.npop = npop

        def __call__(self, rows):
            return ((rows[:, 2] - 200.) / 50.) ** 2
    monkeypatch.setattr(de_mod, '_FramePopulation', Line)
    out = de_mod.estimate_batch(small, holos)
    assert seen == [5]
    assert all(abs(r['z_p'] - 200.) < 5. for r in out)
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
