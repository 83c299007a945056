"""The batched LM with a forward-difference Jacobian of the FP64 hologram
(lmz_lm_settings.fd_step > 0, MINPACK lmdif's steps h_j = eps |p_j|): the
finite-difference form the analytic Jacobian replaces, inside the same
device LM.  Its fits reach the reference's minima (golden fits, reference
Optimizer with diff_step 1e-5 at tight tolerances) and the analytic-
Jacobian fits, to 1e-4 relative (k_p 1e-5 absolute)."""
import numpy as np
import pytest

from conftest import INSTRUMENT

pytestmark = pytest.mark.gpu

VARS6 = ['x_p', 'y_p', 'z_p', 'a_p', 'n_p', 'k_p']
NAMES = [f'crop{i:02d}' for i in range(12)] + ['cfg0']


@pytest.fixture(scope='module')
def fits(fits_golden):
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    from pylorenzmie_amd import analysis, theory
    fx, meta = fits_golden
    holos, parts, sig = [], [], []
    for n in NAMES:
        holos.append(analysis.Hologram(fx[f'{n}_data'].astype(float)))
        parts.append(theory.Sphere(**dict(zip(VARS6, fx[f'{n}_start']))))
        sig.append(float(fx[f'{n}_sigma']))
    out = {}
    for step in (0.0, 1e-5):
        bo = analysis.BatchOptimizer(
            theory.Instrument(**INSTRUMENT), variables=VARS6,
            settings=analysis.LMSettings(ftol=1e-10, xtol=1e-10, gtol=1e-10,
                                         max_iter=100, fd_step=step))
        parts = [theory.Sphere(**dict(zip(VARS6, fx[f'{n}_start']))) for n in NAMES]
        out[step] = bo.optimize(holos, parts, noise=sig)
    return out, meta


def _close(a, b, v):
    return abs(a - b) <= (1e-5 if v == 'k_p' else 1e-4 * abs(b))


def test_fd_fits_match_reference(fits):
    out, meta = fits
    df = out[1e-5]
    same = 0
    for i, n in enumerate(NAMES):
        ref = meta['fits'][f'{n}_tight']
        rc = df.loc[i, 'redchi']
        print(n, 'redchi fd %.8f ref %.8f' % (rc, ref['redchi']))
        if abs(rc - ref['redchi']) <= 1e-6 * ref['redchi']:
            same += 1
            for v in VARS6:
                assert _close(df.loc[i, v], ref[v], v), (n, v, df.loc[i, v], ref[v])
        else:
            assert rc <= ref['redchi'] * (1 + 1e-6), n
    assert same >= 8


def test_fd_fits_match_analytic_fits(fits):
    out, _ = fits
    fd, an = out[1e-5], out[0.0]
    agree = 0
    for i, n in enumerate(NAMES):
        if abs(fd.loc[i, 'redchi'] - an.loc[i, 'redchi']) <= 1e-6 * an.loc[i, 'redchi']:
            agree += 1
            for v in VARS6:
                assert _close(fd.loc[i, v], an.loc[i, v], v), (n, v)
            # uncertainties from the FD and the analytic J^T J
            for v in VARS6:
                assert fd.loc[i, 'd' + v] == pytest.approx(an.loc[i, 'd' + v], rel=0.02)
    assert agree >= 11


def test_fd_step_is_validated():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    from pylorenzmie_amd import _lmz
    from pylorenzmie_amd.analysis.lm import LMBatch, LMSettings
    from pylorenzmie_amd.theory import Instrument
    lm = LMBatch([dict(data=np.ones((32, 32), np.float32), width=32,
                       origin=(0., 0., 0.), nmax_cap=30)], Instrument())
    lm.load(np.array([[16., 16., 100., 0.5, 1.4, 0., 1.]]))
    with pytest.raises(_lmz.LMZError, match='fd_step'):
        lm.run(LMSettings(fd_step=1.5), VARS6)
    lm.close()
