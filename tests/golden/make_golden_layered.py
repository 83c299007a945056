"""Write tests/golden/layered.npz: fit-parity fixtures for layered spheres
fitted in size and index, started as the uneven-crop workload starts its
fits.

Inputs: 16 spheres of 2-5 layers on 128 x 128 crops; a_p U[0.5, 3] um,
layer radii sorted U[0.05, 1] a_p with the outermost at a_p, layer n
U[1.33, 2.0], k 0; alpha U[0.8, 1.2]; x_p, y_p at the crop centre +-5 px,
z_p = a_p / magnification + 30 + U[0, 250] px.  Data: the FP64 oracle's
hologram plus sigma = 0.05 numpy noise (seeded), stored as float32.
Starts: the truth +-2 px in x, y, z, x(1 +- 0.05) a_p, +-0.02 n_p with
random signs.

Fits: oracle.fit with ``layers`` (a_p scales every radius, n_p + i k_p is
the outer layer's index; the reference's MINPACK fitter,
analysis/Optimizer.py:196-278, restated in FP64) over all 7 parameters,
tolerances 1e-10, from those starts.  Runs no reference code.  About 2
minutes on 8 CPUs.

Usage:  python tests/golden/make_golden_layered.py
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
SIDE, N, SIGMA, MAXL = 128, 16, 0.05, 5


def draw():
    rng = np.random.default_rng(50)
    truth, start = np.zeros((N, 7)), np.zeros((N, 7))
    layers = np.zeros((N, MAXL, 3))
    counts = np.zeros(N, np.int32)
    c = (SIDE - 1) / 2
    for i in range(N):
        L = int(rng.integers(2, MAXL + 1))
        a = rng.uniform(0.5, 3.)
        radii = np.sort(rng.uniform(0.05, 1., L)) * a
        radii[-1] = a
        nl = rng.uniform(1.33, 2.0, L)
        layers[i, :L] = np.stack([radii, nl, np.zeros(L)], 1)
        counts[i] = L
        truth[i] = (c + rng.uniform(-5, 5), c + rng.uniform(-5, 5),
                    a / INSTRUMENT['magnification'] + 30. + rng.uniform(0., 250.),
                    a, nl[-1], 0., rng.uniform(0.8, 1.2))
        sg = rng.choice([-1., 1.], size=5)
        start[i] = truth[i] + (2 * sg[0], 2 * sg[1], 2 * sg[2], 0., 0.02 * sg[4],
                               0., 0.)
        start[i, 3] = truth[i, 3] * (1 + 0.05 * sg[3])
    return truth, start, layers, counts


def _fit(job):
    from oracle import fit
    data, coords, start, layers = job
    t0 = time.perf_counter()
    r = fit.fit(data, coords, dict(zip(PARAMS, start)), variables=PARAMS,
                instrument=INSTRUMENT, noise=SIGMA, settings=TIGHT,
                layers=layers)
    r['seconds'] = time.perf_counter() - t0
    r['chi2'] = r['redchi'] * (r['npix'] - 7)
    return r


def main():
    from oracle import fit, lorenzmie as lm
    truth, start, layers, counts = draw()
    grid = lm.meshgrid((SIDE, SIDE))
    rng = np.random.default_rng(51)
    data = np.zeros((N, SIDE * SIDE), np.float32)
    for i in range(N):
        lay = layers[i, :counts[i]]
        p = fit.layered_properties(dict(zip(PARAMS, truth[i])), lay, INSTRUMENT)
        h = lm.hologram(grid, p, alpha=truth[i, 6], **INSTRUMENT)
        data[i] = h + SIGMA * rng.standard_normal(h.size)
    os.environ.update(OMP_NUM_THREADS='1', MKL_NUM_THREADS='1',
                      OPENBLAS_NUM_THREADS='1')
    jobs = [(data[i].astype(float), grid, start[i], layers[i, :counts[i]])
            for i in range(N)]
    t0 = time.perf_counter()
    with mp.get_context('spawn').Pool(7) as pool:
        fits = pool.map(_fit, jobs)
    print(f'fits done in {time.perf_counter() - t0:.0f} s')
    np.savez_compressed(
        os.path.join(HERE, 'layered.npz'), truth=truth, start=start,
        layers=layers, counts=counts, data=data, side=np.array(SIDE),
        sigma=np.array(SIGMA),
        oracle_values=np.array([[r[v] for v in PARAMS] for r in fits]),
        oracle_uncertainties=np.array([[r['d' + v] for v in PARAMS] for r in fits]),
        oracle_chi2=np.array([r['chi2'] for r in fits]),
        oracle_nfev=np.array([r['nfev'] for r in fits]),
        oracle_success=np.array([r['success'] for r in fits]),
        oracle_seconds=np.array([r['seconds'] for r in fits]))


if __name__ == '__main__':
    main()
