"""The oracle's restatement of the grid preparation (oracle/gridops.py) against scipy, the
library the reference calls (fteikpy/_base.py:81-123, :165-212), bit for bit, and against
the reference's own tests of smooth / resample."""
import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator
from scipy.ndimage import gaussian_filter

from oracle import gridops

CASES = [((7, 5), (0.7, 2.5)), ((3, 4, 5), (3.0, 1.0, 5.5)), ((16, 16), (1.0, 1.0)),
         ((20, 9, 13), (0.3, 0.0, 2.2)), ((2, 3), (9.0, 9.0)), ((31, 17, 6), (4.0, 12.5, 0.9))]


@pytest.mark.parametrize("shape,sig", CASES)
def test_gaussian_filter_matches_scipy_bitwise(shape, sig):
    a = np.random.default_rng(len(shape) * 7 + shape[0]).random(shape)
    assert np.array_equal(gridops.gaussian_filter(a, sig), gaussian_filter(a, sig))


def test_smooth_meters_to_cells():
    a = np.random.default_rng(3).random((12, 15, 9))
    d = (10.0, 12.5, 7.0)
    got = gridops.smooth(a, d, 25.0)
    assert np.array_equal(got, gaussian_filter(a, 25.0 / np.asarray(d)))


RS = [((16, 16), (32, 32), (0.0, 0.0), (1.0, 1.0)), ((7, 9), (13, 4), (1.5, -2.0), (0.7, 1.3)),
      ((50, 70), (33, 101), (3.0, 1.0), (0.31, 0.77)), ((5, 6, 7), (9, 11, 3), (0.1, 0.2, 0.3),
                                                         (1.1, 0.9, 2.0)),
      ((16, 16, 16), (32, 32, 32), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)),
      ((9, 4, 12), (2, 7, 12), (-5.0, 2.0, 0.0), (10.0, 12.5, 3.0))]


@pytest.mark.parametrize("shape,new,origin,d", RS)
@pytest.mark.parametrize("method", ["linear", "nearest"])
def test_resample_matches_scipy_bitwise(shape, new, origin, d, method):
    v = np.random.default_rng(sum(shape)).random(shape)
    axes = [o + dd * np.arange(n) for o, dd, n in zip(origin, d, shape)]
    fn = RegularGridInterpolator(points=tuple(axes), values=v, method=method, bounds_error=False)
    qs = np.meshgrid(*[np.linspace(a[0], a[-1], n) for a, n in zip(axes, new)], indexing="ij")
    ref = fn(np.stack([q.ravel() for q in qs], -1)).reshape(new)
    assert np.array_equal(gridops.resample(v, d, origin, new, method), ref)


def test_reference_pins():
    # tests/test_eikonal2d.py:36-48, tests/test_eikonal3d.py:48-60 (16^2 / 16^3 ones, d = 1)
    for shape in ((16, 16), (16, 16, 16)):
        ones = np.ones(shape)
        new = tuple(2 * n for n in shape)
        assert gridops.resample(ones, (1.0,) * len(shape), (0.0,) * len(shape), new).sum() == \
            np.prod(shape) * 2 ** len(shape)
        assert gridops.smooth(ones, (1.0,) * len(shape), 1.0).sum() == ones.size


def test_planted_boundary_mistake_is_caught():
    """A 'mirror' boundary (d c b | a b c d | c b a) in place of 'reflect' fails the scipy
    comparison: the pin is sensitive to the boundary rule."""
    a = np.random.default_rng(5).random((6, 8))
    w, r = gridops.gaussian_kernel(2.0)
    orig = gridops.reflect_index
    try:
        gridops.reflect_index = lambda i, n: np.abs(((i + (n - 1)) % (2 * n - 2)) - (n - 1))
        bad = gridops.gaussian_filter(a, (2.0, 2.0))
    finally:
        gridops.reflect_index = orig
    assert not np.allclose(bad, gaussian_filter(a, (2.0, 2.0)), rtol=1e-12, atol=0)
