"""Azimuthal statistics (lib/Azimuthal.py): the reference's length checks
(tests/test_lib.py:9-33) and every statistic against its definition,
computed pixel by pixel (bin int(hypot(i - y_c, j - x_c)))."""
import numpy as np
import pytest

from pylorenzmie_amd.lib import azimuthal


def _brute(data, center):
    ny, nx = data.shape
    xc, yc = (nx / 2., ny / 2.) if center is None else center
    rings = {}
    for i in range(ny):
        for j in range(nx):
            rings.setdefault(int(np.hypot(i - yc, j - xc)), []).append(data[i, j])
    n = max(rings) + 1
    out = {k: np.full(n, np.nan) for k in ('avg', 'std', 'med', 'mad')}
    for r, v in rings.items():
        v = np.array(v)
        out['avg'][r] = v.mean()
        out['std'][r] = v.std()
        out['med'][r] = np.median(v)
        out['mad'][r] = np.median(np.abs(v - np.median(v)))
    return out


@pytest.mark.parametrize('fn', ['avg', 'med'])
def test_lengths_as_reference(fn):
    data = np.ones((90, 100))
    a = getattr(azimuthal, fn)(data, (50, 45))
    assert len(a) == int(np.hypot(50, 45)) + 1


@pytest.mark.parametrize('fn', ['std', 'mad'])
def test_pair_lengths_as_reference(fn):
    a, b = getattr(azimuthal, fn)(np.ones((90, 100)), (50, 45))
    assert len(a) == len(b) == int(np.hypot(50, 45)) + 1


@pytest.mark.parametrize('center', [None, (7.3, 4.1), (0., 0.), (12.5, 9.0)])
def test_statistics_match_definition(center):
    rng = np.random.default_rng(1)
    data = rng.standard_normal((13, 17))
    want = _brute(data, center)
    np.testing.assert_allclose(azimuthal.avg(data, center), want['avg'], rtol=1e-12)
    a, s = azimuthal.std(data, center)
    np.testing.assert_allclose(a, want['avg'], rtol=1e-12)
    np.testing.assert_allclose(s, want['std'], rtol=1e-10, atol=1e-15)
    np.testing.assert_array_equal(azimuthal.med(data, center), want['med'])
    m, d = azimuthal.mad(data, center)
    np.testing.assert_array_equal(m, want['med'])
    np.testing.assert_array_equal(d, want['mad'])


def test_radial_profile_of_a_ring_pattern():
    """A function of the integer radius alone comes back (to the rounding
    of the bin sums)."""
    y, x = np.mgrid[:40, :50]
    r = np.hypot(y - 20., x - 25.).astype(int)
    np.testing.assert_allclose(azimuthal.avg(np.cos(r), (25., 20.)),
                               np.cos(np.arange(r.max() + 1)), rtol=0, atol=1e-13)
