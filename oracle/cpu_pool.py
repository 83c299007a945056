"""The FP64 oracle's fitter on every host core: one process per core, each
fitting its share of a sample of features one at a time, as a multi-core
CPU baseline.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): bench.py's CPU-baseline
leg starts it as a child process (python -m oracle.cpu_pool SAMPLE.npz
WORKERS), so that the processes that fork here never touched the GPU; it
imports numpy and scipy only.  SAMPLE.npz holds data (n, side, side),
start (n, 7) in the order x_p y_p z_p a_p n_p k_p alpha, and sigma.
Prints one JSON object: fits, seconds (wall), workers, value (fits/s).
"""
from __future__ import annotations

import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

PARAMS = ('x_p', 'y_p', 'z_p', 'a_p', 'n_p', 'k_p', 'alpha')
INSTRUMENT = dict(n_m=1.340, wavelength=0.447, magnification=0.048)


def _fit_share(args):
    path, idx = args
    from oracle import fit, lorenzmie as lm
    z = np.load(path)
    data, start, sigma = z['data'], z['start'], float(z['sigma'])
    c = lm.meshgrid(data.shape[1:])
    for i in idx:
        fit.fit(data[i], c, dict(zip(PARAMS, start[i])), variables=PARAMS,
                instrument=INSTRUMENT, noise=sigma)
    return len(idx)


def _warm(_):
    from oracle import fit  # noqa: F401
    return os.getpid()


def main(path: str, workers: int) -> dict:
    n = int(np.load(path)['start'].shape[0])
    shares = [(path, list(range(w, n, workers))) for w in range(workers)]
    ctx = mp.get_context('fork')
    with ctx.Pool(workers) as pool:
        pool.map(_warm, range(workers))
        t0 = time.perf_counter()
        done = sum(pool.map(_fit_share, shares, chunksize=1))
        dt = time.perf_counter() - t0
    return dict(fits=done, seconds=dt, workers=workers, value=done / dt,
                omp_threads=os.environ.get('OMP_NUM_THREADS', 'unset'))


if __name__ == '__main__':
    # the caller sets OMP_NUM_THREADS=1 in the environment (numpy reads it
    # when it is imported, before this line)
    print(json.dumps(main(sys.argv[1], int(sys.argv[2]))))
