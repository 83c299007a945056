"""Write tests/golden/large.npz: fit-parity fixtures for large homogeneous
spheres of the hard-regime workload, started as the uneven-crop workload
starts its fits.

Inputs: 16 spheres with a_p U[3, 10] um (n_max to ~250) on 512 x 512 crops,
n_p U[1.38, 1.65], k_p U[0, 0.01], alpha U[0.8, 1.2], x_p, y_p at the crop
centre +-5 px, z_p = a_p / magnification + 30 + U[0, 250] px (the focal
plane clear of the sphere); a 10% pixel mask per crop (explicit coordinate
lists, as Mask draws them); the FP64 oracle's hologram plus sigma = 0.05
numpy noise (seeded), stored as float32.  Starts: the truth +-2 px in x, y,
z, x(1 +- 0.05) a_p, +-0.02 n_p with random signs.

Fits: oracle.fit (the reference's MINPACK fitter restated in FP64,
analysis/Optimizer.py:196-278) with all 7 parameters, tolerances 1e-10,
from those starts.  Runs no reference code.  About 4 minutes on 8 CPUs.

Usage:  python tests/golden/make_golden_large.py
"""
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

INSTRUMENT = dict(wavelength=0.447, magnification=0.048, n_m=1.340)
PARAMS = ('x_p', 'y_p', 'z_p', 'a_p', 'n_p', 'k_p', 'alpha')
TIGHT = dict(ftol=1e-10, xtol=1e-10, gtol=1e-10, max_nfev=4000)
SIDE, N, FRACTION, SIGMA = 512, 16, 0.10, 0.05


def draw():
    rng = np.random.default_rng(40)
    truth, start, index = np.zeros((N, 7)), np.zeros((N, 7)), []
    c = (SIDE - 1) / 2
    for i in range(N):
        a = rng.uniform(3., 10.)
        truth[i] = (c + rng.uniform(-5, 5), c + rng.uniform(-5, 5),
                    a / INSTRUMENT['magnification'] + 30. + rng.uniform(0., 250.),
                    a, rng.uniform(1.38, 1.65), rng.uniform(0., 0.01),
                    rng.uniform(0.8, 1.2))
        sg = rng.choice([-1., 1.], size=5)
        start[i] = truth[i] + (2 * sg[0], 2 * sg[1], 2 * sg[2], 0., 0.02 * sg[4],
                               0., 0.)
        start[i, 3] = truth[i, 3] * (1 + 0.05 * sg[3])
        index.append(np.sort(rng.choice(SIDE * SIDE, int(FRACTION * SIDE * SIDE),
                                        replace=False)))
    return truth, start, np.array(index)


def _fit(job):
    from oracle import fit
    data, coords, start = job
    t0 = time.perf_counter()
    r = fit.fit(data, coords, dict(zip(PARAMS, start)), variables=PARAMS,
                instrument=INSTRUMENT, noise=SIGMA, settings=TIGHT)
    r['seconds'] = time.perf_counter() - t0
    r['chi2'] = r['redchi'] * (r['npix'] - 7)
    return r


def main():
    from oracle import lorenzmie as lm
    truth, start, index = draw()
    grid = lm.meshgrid((SIDE, SIDE))
    rng = np.random.default_rng(41)
    data = np.zeros(index.shape, np.float32)
    for i in range(N):
        c = grid[:, index[i]]
        h = lm.hologram(c, dict(zip(PARAMS[:6], truth[i, :6])), alpha=truth[i, 6],
                        **INSTRUMENT)
        data[i] = h + SIGMA * rng.standard_normal(h.size)
    os.environ.update(OMP_NUM_THREADS='1', MKL_NUM_THREADS='1',
                      OPENBLAS_NUM_THREADS='1')
    jobs = [(data[i].astype(float), grid[:, index[i]], start[i]) for i in range(N)]
    t0 = time.perf_counter()
    with mp.get_context('spawn').Pool(7) as pool:
        fits = pool.map(_fit, jobs)
    print(f'fits done in {time.perf_counter() - t0:.0f} s')
    np.savez_compressed(
        os.path.join(HERE, 'large.npz'), truth=truth, start=start,
        index=index.astype(np.int32), data=data, side=np.array(SIDE),
        sigma=np.array(SIGMA),
        oracle_values=np.array([[r[v] for v in PARAMS] for r in fits]),
        oracle_uncertainties=np.array([[r['d' + v] for v in PARAMS] for r in fits]),
        oracle_chi2=np.array([r['chi2'] for r in fits]),
        oracle_nfev=np.array([r['nfev'] for r in fits]),
        oracle_success=np.array([r['success'] for r in fits]),
        oracle_seconds=np.array([r['seconds'] for r in fits]))


if __name__ == '__main__':
    main()
