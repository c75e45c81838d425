"""GPU: device smoothing and resampling of grids (fteik_correlate1d, fteik_resample)
against the oracle (pinned to scipy in test_oracle_gridops.py), bit for bit, and the
solver picks up the prepared model."""
import numpy as np
import pytest

import oracle
from oracle import gridops

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fa(device):
    import fteikpy_amd

    return fteikpy_amd


@pytest.mark.parametrize("shape,d,sigma", [((16, 16), (1.0, 1.0), 1.0),
                                           ((21, 34), (10.0, 12.5), (15.0, 40.0)),
                                           ((16, 16, 16), (1.0, 1.0, 1.0), 1.0),
                                           ((9, 23, 14), (5.0, 7.5, 10.0), 30.0),
                                           ((3, 40, 5), (1.0, 2.0, 1.0), (0.0, 9.0, 6.0))])
def test_smooth_bitwise(fa, shape, d, sigma):
    v = 1500.0 + 3000.0 * np.random.default_rng(sum(shape)).random(shape)
    G = fa.Grid3D if len(shape) == 3 else fa.Grid2D
    g = G(v, d, np.zeros(len(shape)))
    g.smooth(sigma)
    assert np.array_equal(g.grid, gridops.smooth(v, d, sigma))


@pytest.mark.parametrize("shape,new,d,o", [((16, 16), (32, 32), (1.0, 1.0), (0.0, 0.0)),
                                           ((17, 25), (9, 60), (10.0, 12.5), (3.0, -1.0)),
                                           ((16, 16, 16), (32, 32, 32), (1.0, 1.0, 1.0),
                                            (0.0, 0.0, 0.0)),
                                           ((7, 12, 9), (13, 5, 9), (5.0, 7.5, 10.0),
                                            (1.0, 2.0, 3.0))])
@pytest.mark.parametrize("method", ["linear", "nearest"])
def test_resample_bitwise(fa, shape, new, d, o, method):
    v = np.random.default_rng(sum(new)).random(shape)
    G = fa.Grid3D if len(shape) == 3 else fa.Grid2D
    g = G(v, d, o)
    g.resample(new, method)
    assert g.shape == new
    assert np.array_equal(g.grid, gridops.resample(v, d, o, new, method))
    # the reference's gridsize update, computed after the grid is replaced
    assert g.gridsize == tuple(a * b / c for a, b, c in zip(d, new, new))


def test_reference_pins(fa):
    # tests/test_eikonal2d.py:36-48, tests/test_eikonal3d.py:48-60
    for E, shape in ((fa.Eikonal2D, (16, 16)), (fa.Eikonal3D, (16, 16, 16))):
        e = E(np.ones(shape), (1.0,) * len(shape))
        e.resample(tuple(2 * n for n in shape))
        assert e.grid.sum() == np.prod(shape) * 2 ** len(shape)
        e = E(np.ones(shape), (1.0,) * len(shape))
        e.smooth(1.0)
        assert e.grid.sum() == e.grid.size
    with pytest.raises(ValueError):
        fa.Grid2D(np.ones((4, 4)), (1.0, 1.0), (0.0, 0.0)).resample((8, 8), method="cubic")


def test_solve_after_smoothing(fa):
    """Eikonal3D.solve uses the smoothed model (the slowness is recomputed)."""
    rng = np.random.default_rng(8)
    shape, d = (12, 15, 11), (10.0, 10.0, 10.0)
    v = 2000.0 + 2000.0 * rng.random(shape)
    eik = fa.Eikonal3D(v, d)
    src = np.array([33.0, 71.0, 52.0])
    eik.solve(src, nsweep=1)
    eik.smooth(20.0)
    vs = gridops.smooth(v, d, 20.0)
    tt = eik.solve(src, nsweep=None)
    ref, _, _, _ = oracle.solve3d(1.0 / vs, *d, src, nsweep=0)
    assert np.array_equal(tt.grid, ref)
