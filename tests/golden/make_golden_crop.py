"""Write tests/golden/crop.npz: the reference's own test hologram
(tests/data/crop.png, 390 x 390, 8-bit, a real camera crop of one sphere)
and the FP64 oracle's fits of it (oracle/fit.py, the reference's MINPACK
fitter restated and pinned to its golden fits) from the start that the
reference's tests/test_optimizer.py:25-38 uses.

Reads the PNG only; runs no reference code.  About one minute of CPU.

Usage:  python tests/golden/make_golden_crop.py
"""
import os
import sys
import time

import numpy as np
from PIL import Image

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import fit, lorenzmie as lm  # noqa: E402

PNG = '/root/reference/tests/data/crop.png'
INSTRUMENT = dict(n_m=1.34, wavelength=0.447, magnification=0.048)
VARS5 = ['x_p', 'y_p', 'z_p', 'a_p', 'n_p']
TIGHT = {'ftol': 1e-10, 'xtol': 1e-10, 'gtol': 1e-10}


def main():
    image = np.asarray(Image.open(PNG))
    assert image.dtype == np.uint8 and image.ndim == 2
    data = image.astype(float)
    data /= np.mean(data)
    ny, nx = data.shape
    start = dict(x_p=ny // 2, y_p=nx // 2, z_p=330., a_p=1.1, n_p=1.4, k_p=0.)
    c = lm.meshgrid(data.shape)
    out = dict(image=image,
               start=np.array([start[k] for k in lm.PARAMS[:6]], float),
               noise=np.array(0.05))
    for mode, settings in (('default', {}), ('tight', TIGHT)):
        t0 = time.perf_counter()
        r = fit.fit(data, c, start, variables=VARS5, instrument=INSTRUMENT,
                    noise=0.05, settings=settings)
        print(mode, f'{time.perf_counter() - t0:.1f} s', r)
        out[f'{mode}_values'] = np.array([r[v] for v in VARS5])
        out[f'{mode}_uncertainties'] = np.array([r['d' + v] for v in VARS5])
        out[f'{mode}_redchi'] = np.array(r['redchi'])
        out[f'{mode}_success'] = np.array(r['success'])
    np.savez_compressed(os.path.join(HERE, 'crop.npz'), **out)


if __name__ == '__main__':
    main()
