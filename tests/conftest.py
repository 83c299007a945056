import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, 'tests', 'golden')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (ROCm) GPU')


@pytest.fixture(scope='session')
def forward_golden():
    return np.load(os.path.join(GOLDEN, 'forward.npz'))


@pytest.fixture(scope='session')
def fits_golden():
    import json
    with open(os.path.join(GOLDEN, 'fits.json')) as f:
        meta = json.load(f)
    return np.load(os.path.join(GOLDEN, 'fits.npz')), meta


INSTRUMENT = dict(n_m=1.34, wavelength=0.447, magnification=0.048)


@pytest.fixture(scope='session')
def superposed_golden():
    return np.load(os.path.join(GOLDEN, 'superposed.npz'))


@pytest.fixture(scope='session')
def frame_golden():
    return np.load(os.path.join(GOLDEN, 'frame.npz'))


@pytest.fixture(scope='session')
def crop_golden():
    """The reference's test hologram tests/data/crop.png and the oracle's
    fits of it (tests/golden/make_golden_crop.py)."""
    return np.load(os.path.join(GOLDEN, 'crop.npz'))
