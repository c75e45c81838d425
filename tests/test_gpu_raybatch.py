"""GPU: rays of many traveltime grids of one model in one launch (raytrace_batch) equal,
bit for bit, the rays each grid traces on its own (TraveltimeGrid3D.raytrace, checked
against the reference's golden rays in test_gpu_parity.py), with and without honor_grid,
for single end points and arrays of them; the oracle's ray tracer agrees too."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fa(device):
    import fteikpy_amd

    return fteikpy_amd


def _grids(fa, nsrc=5, seed=41):
    rng = np.random.default_rng(seed)
    shape, d = (18, 24, 21), (10.0, 12.0, 11.0)
    v = 2000.0 + 1500.0 * rng.random(shape)
    ext = np.array(shape) * np.array(d)
    src = rng.random((nsrc, 3)) * ext
    grids = fa.Eikonal3D(v, d).solve(src, nsweep=None, return_gradient=True)
    return v, d, src, ext, grids, rng


@pytest.mark.parametrize("honor_grid", [False, True])
def test_batch_rays_equal_single_grid_rays(fa, honor_grid):
    v, d, src, ext, grids, rng = _grids(fa)
    pts = [rng.random((k, 3)) * ext for k in (7, 1, 12, 3, 9)]
    pts[1] = pts[1][0]  # one grid's end point given as [z, x, y]
    got = fa.raytrace_batch(grids, pts, honor_grid=honor_grid)
    assert len(got) == len(grids)
    for g, p, r in zip(grids, pts, got):
        want = g.raytrace(p, honor_grid=honor_grid)
        if np.ndim(p) == 1:
            assert isinstance(r, np.ndarray) and np.array_equal(r, want)
        else:
            assert len(r) == len(want)
            for a, b in zip(r, want):
                assert np.array_equal(a, b)


def test_batch_rays_match_oracle(fa):
    v, d, src, ext, grids, rng = _grids(fa, nsrc=3, seed=43)
    pts = [rng.random((4, 3)) * ext for _ in grids]
    got = fa.raytrace_batch(grids, pts)
    nodes = [n + 1 for n in v.shape]
    axes = [d[a] * np.arange(nodes[a]) for a in range(3)]
    step = min(d)
    max_step = int(2.0 * sum((n * h) ** 2 for n, h in zip(nodes, d)) ** 0.5 / step)
    for s, p, rays in zip(src, pts, got):
        _, grad, _, _ = oracle.solve3d(1.0 / v, *d, s, nsweep=0, grad=True)
        want = oracle.ray3d(*axes, *[grad[..., a].copy() for a in range(3)], p, s, step,
                            max_step)
        for a, b in zip(rays, want):
            np.testing.assert_allclose(a, b, rtol=1e-8, atol=1e-8)


def test_batch_rays_arguments(fa):
    v, d, src, ext, grids, rng = _grids(fa, nsrc=2, seed=47)
    with pytest.raises(ValueError):
        fa.raytrace_batch(grids, [ext / 2])  # one entry per grid
    other = fa.Eikonal3D(v[:-1], d).solve(src[0] * 0.5, nsweep=1, return_gradient=True)
    with pytest.raises(ValueError):
        fa.raytrace_batch([grids[0], other], [ext / 3, ext / 4])
    assert fa.raytrace_batch([], []) == []
    with pytest.raises(ValueError):  # an end point outside the grid, as raytrace raises
        fa.raytrace_batch(grids, [ext / 2, ext * 2])
