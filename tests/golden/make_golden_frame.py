"""Write tests/golden/frame.npz: crops of the reference's own camera frames
(tests/data/image0010.png and docs/tutorials/background.png, 8-bit) with the
backgrounds of utilities/normalizer.py computed by the oracle (np.median,
scipy.ndimage.median_filter at sizes 51 and 8).  Reads the PNG files as
data with PIL; runs no reference code.

    python tests/golden/make_golden_frame.py [/root/reference]
"""
import os
import sys

import numpy as np
from PIL import Image

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import normalize as ref  # noqa: E402

REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
ROWS, COLS = slice(576, 832), slice(448, 768)      # a feature and its fringes


def main():
    frame = np.asarray(Image.open(os.path.join(REF, 'tests/data/image0010.png')))
    bgimg = np.asarray(Image.open(os.path.join(REF, 'docs/tutorials/background.png')))
    image = np.ascontiguousarray(frame[ROWS, COLS])
    background = np.ascontiguousarray(bgimg[ROWS, COLS, 0])
    # the normalized images are image / where(bg > 0, bg, 1.) of these
    out = dict(image=image, background=background,
               median=np.float64(ref.median(image)),
               bg51=ref.median_filter(image, 51),
               bg8=ref.median_filter(image, 8))
    np.savez_compressed(os.path.join(HERE, 'frame.npz'), **out)
    for k, v in out.items():
        print(k, v.shape, v.dtype)


if __name__ == '__main__':
    main()
