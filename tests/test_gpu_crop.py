"""The reference's optimizer tests (tests/test_optimizer.py:25-144) on its
own test hologram, tests/data/crop.png (a 390 x 390 camera crop of one
sphere, carried as tests/golden/crop.npz), fitted on the GPU: the contract
checks at max_nfev = 1, and full fits against the FP64 oracle's fits of
the same hologram (oracle/fit.py, the reference's MINPACK fitter).

Tolerances: at tight tolerances both fits sit at the minimum, so the
values agree to 1e-4 relative and 0.05 of an uncertainty (MINPACK's
forward-difference Jacobian places its stop up to 0.01 of an uncertainty
from the exact minimum on this hologram), the uncertainties to 2% and the
reduced chi2 to 1e-6 relative.  At the default tolerances each fitter
stops where its own ftol says; the GPU fit must lie within half an
uncertainty of the minimum with a reduced chi2 no worse than the
oracle's default fit."""
import numpy as np
import pandas as pd
import pytest

pytestmark = pytest.mark.gpu

VARS5 = ['x_p', 'y_p', 'z_p', 'a_p', 'n_p']


@pytest.fixture(scope='module')
def api():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    from pylorenzmie_amd import analysis, lib, theory
    return theory, analysis, lib


def setup(api, g, mask=None):
    """tests/test_optimizer.py:25-38 of the reference."""
    theory, analysis, lib = api
    img = g['image'].astype(float)
    img /= np.mean(img)
    shape = img.shape
    model = theory.LorenzMie(coordinates=lib.LMObject.meshgrid(shape))
    model.instrument.wavelength = 0.447
    model.instrument.magnification = 0.048
    model.instrument.n_m = 1.34
    model.particle.r_p = [shape[0] // 2, shape[1] // 2, 330]
    model.particle.a_p = 1.1
    model.particle.n_p = 1.4
    return analysis.Hologram(img), analysis.Optimizer(model=model, mask=mask)


def test_optimize_contract_at_one_evaluation(api, crop_golden):
    """test_optimize_returns_series, _with_mask_returns_series,
    _npix_all_pixels and _npix_masked."""
    _, analysis, _ = api
    hologram, opt = setup(api, crop_golden)
    assert opt.result is None and opt.mask is None
    opt.settings['max_nfev'] = 1
    res = opt.optimize(hologram)
    assert isinstance(res, pd.Series)
    ny, nx = hologram.shape
    assert res['npix'] == ny * nx
    assert list(res.index) == sum(([v, 'd' + v] for v in VARS5), []) + \
        ['success', 'npix', 'redchi']
    hologram, opt = setup(api, crop_golden, mask=analysis.Mask(fraction=0.1))
    opt.settings['max_nfev'] = 1
    res = opt.optimize(hologram)
    assert isinstance(res, pd.Series)
    assert res['npix'] < ny * nx
    assert abs(res['npix'] / (ny * nx) - 0.1) < 0.01


def test_fit_matches_oracle_at_tight_tolerances(api, crop_golden):
    g = crop_golden
    hologram, opt = setup(api, g)
    opt.settings.update(ftol=1e-10, xtol=1e-10, gtol=1e-10)
    res = opt.optimize(hologram)
    assert bool(res['success'])
    want, unc = g['tight_values'], g['tight_uncertainties']
    got = np.array([res[v] for v in VARS5])
    print('tight: gpu', got, 'oracle', want, 'in sigma', (got - want) / unc)
    assert np.all(np.abs(got - want) <= 1e-4 * np.abs(want))
    assert np.all(np.abs(got - want) <= 0.05 * unc)
    dgot = np.array([res['d' + v] for v in VARS5])
    assert np.all(np.abs(dgot - unc) <= 0.02 * unc)
    assert res['redchi'] == pytest.approx(float(g['tight_redchi']), rel=1e-6)
    assert 'npixels = 152100' in opt.report()


def test_fit_matches_oracle_at_default_tolerances(api, crop_golden):
    g = crop_golden
    hologram, opt = setup(api, g)
    res = opt.optimize(hologram)
    assert bool(res['success'])
    want, unc = g['tight_values'], g['tight_uncertainties']
    got = np.array([res[v] for v in VARS5])
    print('default: gpu', got, 'in sigma', (got - want) / unc,
          'redchi', res['redchi'], 'oracle default', float(g['default_redchi']))
    assert np.all(np.abs(got - want) <= 0.5 * unc)
    assert res['redchi'] <= float(g['default_redchi']) * (1 + 1e-6)
    # the model carries the fitted values, as the reference's does
    assert opt.model.particle.z_p == pytest.approx(res['z_p'])
