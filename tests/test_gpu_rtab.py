"""The LM's Jacobian on dense crops interpolates radial tables of the
multipole sums (k_jac_rtab, k_jacobian_rtab); on crops given by explicit
coordinates it runs the order loop at every pixel (k_jacobian).  The same
data fitted both ways in one batch reaches the same minimum."""
import os

import numpy as np
import pytest

from conftest import INSTRUMENT, ROOT
from oracle import lorenzmie as ref

pytestmark = pytest.mark.gpu

PARAMS = ['x_p', 'y_p', 'z_p', 'a_p', 'n_p', 'k_p', 'alpha']


@pytest.fixture(scope='module')
def fx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return np.load(os.path.join(ROOT, 'tests', 'golden', 'headline.npz'))


def test_tabulated_and_loop_jacobians_reach_the_same_minimum(fx):
    from pylorenzmie_amd import _lmz
    from pylorenzmie_amd.analysis.lm import LMBatch, LMSettings
    from pylorenzmie_amd.theory import Instrument
    inst = Instrument(**INSTRUMENT)
    n, side = 8, 201
    sigma = float(fx['head_sigma'])
    starts = np.asarray(fx['head_start'][:n])
    grid = ref.meshgrid((side, side))
    coords = np.vstack([grid, np.zeros((1, grid.shape[1]))])
    crops = []
    for k in range(2 * n):
        i = k % n
        cap = _lmz.nmax_cap(starts[i, 3], starts[i, 4], starts[i, 5],
                            inst.n_m, inst.wavelength)
        d = np.asarray(fx['head_data'][i], np.float32)
        if k < n:     # dense: tabulated sums
            crops.append(dict(data=d.reshape(side, side), width=side,
                              origin=(0., 0., 0.), noise=sigma, nmax_cap=cap))
        else:         # explicit coordinates: the order loop
            crops.append(dict(data=d, width=0, origin=(0., 0., 0.),
                              coords=coords, noise=sigma, nmax_cap=cap))
    lm = LMBatch(crops, inst)
    lm.load(np.vstack([starts, starts]))
    lm.run(LMSettings(max_iter=500, ftol=1e-9, xtol=1e-9, gtol=1e-10), PARAMS)
    p, res = lm.fetch()
    lm.close()
    assert not np.any(res[:, 38].astype(int) & 32)
    pt, pl = p[:n, :7], p[n:, :7]
    d = np.abs(pt - pl)
    rel = d / np.abs(pl)
    rel[:, 5] = d[:, 5] * 10.                 # k_p: 1e-5 absolute
    chi_t, chi_l = res[:n, 35], res[n:, 35]
    print('max param rel', rel.max(axis=1))
    print('chi2 rel', (chi_t - chi_l) / chi_l)
    assert np.all(rel.max(axis=1) <= 1e-4)
    assert np.all(np.abs(chi_t - chi_l) <= 1e-8 * chi_l)
