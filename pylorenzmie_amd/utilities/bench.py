"""Throughput of frame normalization on the GPU.

    python -m pylorenzmie_amd.utilities.bench [--ny 1024] [--nx 1280]
        [--size 51] [--repeat 20]

Times Normalizer('median'), Normalizer('filter', size) and
Normalizer('reference') on a synthetic camera frame (fringes of two
features on a sloped background with noise) already in device memory, for
8-bit, 16-bit, float32 and float64 pixels, and prints one JSON line per
case: milliseconds per frame, megapixels per second, and for the median
filter the window elements visited per second by its algorithm.
"""
from __future__ import annotations

import argparse
import json

import numpy as np


def synthetic_frame(ny: int, nx: int, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[:ny, :nx].astype(float)
    v = 90. + 20. * x / nx + 10. * y / ny
    for cx, cy, k in ((0.3 * nx, 0.5 * ny, 0.35), (0.7 * nx, 0.3 * ny, 0.25)):
        r = np.hypot(x - cx, y - cy)
        v += 40. * np.cos(k * r) * np.exp(-r / 150.)
    v += rng.normal(0., 3., v.shape)
    return np.clip(v, 0, 255).astype(np.uint8)


def main(argv=None) -> None:
    import torch
    from . import Normalizer
    ap = argparse.ArgumentParser()
    ap.add_argument('--ny', type=int, default=1024)
    ap.add_argument('--nx', type=int, default=1280)
    ap.add_argument('--size', type=int, default=51)
    ap.add_argument('--repeat', type=int, default=20)
    a = ap.parse_args(argv)
    raw = synthetic_frame(a.ny, a.nx)
    frames = {'u8': raw, 'u16': raw.astype(np.uint16),
              'f32': raw.astype(np.float32) / np.float32(100.),
              'f64': raw.astype(np.float64) / 100.}
    background = torch.from_numpy(synthetic_frame(a.ny, a.nx, seed=1)).cuda()
    npx = a.ny * a.nx
    for name, arr in frames.items():
        d = torch.from_numpy(arr).cuda()
        cases = (('median', Normalizer('median')),
                 ('filter', Normalizer('filter', size=a.size)),
                 ('reference', Normalizer('reference', reference=background)))
        for method, norm in cases:
            norm(d)
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(a.repeat):
                norm(d)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / a.repeat
            line = dict(frame=f'{a.ny}x{a.nx}', dtype=name, method=method,
                        ms_per_frame=round(ms, 4),
                        mpix_per_s=round(npx / ms / 1e3, 1))
            if method == 'filter':
                line['size'] = a.size
                # the 8-bit kernel slides a histogram (2 size visits per
                # pixel after a size^2 start per 8..32 rows); the radix
                # kernel visits size^2 keys per 8-bit digit
                line['window'] = f'{a.size}x{a.size}'
            print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
