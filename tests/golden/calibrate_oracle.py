"""Write tests/golden/cpu_calibration.json: the FP64 oracle's fitter
(oracle/fit.py) timed on the inputs of the reference fits in fits.npz,
against the reference's own wall times for the same fits, recorded in
fits.json when make_golden.py ran them on this container's CPU.  The ratio
converts the oracle's CPU baseline (bench.py, timed on the GPU host) into
reference time.  Runs no reference code.  One thread (OMP/MKL = 1).

Usage:  OMP_NUM_THREADS=1 python tests/golden/calibrate_oracle.py
"""
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import fit, lorenzmie as lm  # noqa: E402

INSTRUMENT = dict(n_m=1.34, wavelength=0.447, magnification=0.048)


def main():
    z = np.load(os.path.join(HERE, 'fits.npz'))
    with open(os.path.join(HERE, 'fits.json')) as f:
        meta = json.load(f)
    rows = {}
    names = sorted({k.rsplit('_', 1)[0] for k in z.files})
    for name in names:
        data = z[name + '_data']
        c = lm.meshgrid(data.shape)
        start = dict(zip(fit.VARS6, z[name + '_start']))
        best = None
        for _ in range(3):     # best of three against scheduler noise
            t0 = time.perf_counter()
            r = fit.fit(data, c, start, instrument=INSTRUMENT,
                        noise=float(z[name + '_sigma']))
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        ref_s = meta['wall_seconds'][f'{name}_default']
        rows[name] = dict(pixels=int(data.size), oracle_s=best,
                          reference_s=ref_s, ratio=best / ref_s,
                          nfev=r['nfev'])
        print(name, rows[name], flush=True)
    tot_o = sum(r['oracle_s'] for r in rows.values())
    tot_r = sum(r['reference_s'] for r in rows.values())
    lscpu = subprocess.run(['lscpu'], capture_output=True, text=True).stdout
    model = next((ln.split(':', 1)[1].strip() for ln in lscpu.splitlines()
                  if ln.startswith('Model name')), 'unknown')
    out = dict(what='oracle.fit / reference Optimizer.optimize wall time, '
                    'default settings, 6 parameters, one thread',
               fits=rows, total_oracle_s=tot_o, total_reference_s=tot_r,
               ratio=tot_o / tot_r, cpu=model,
               reference_host=meta.get('host'),
               omp_threads=os.environ.get('OMP_NUM_THREADS', 'unset'))
    with open(os.path.join(HERE, 'cpu_calibration.json'), 'w') as f:
        json.dump(out, f, indent=1)
    print('ratio', out['ratio'])


if __name__ == '__main__':
    main()
