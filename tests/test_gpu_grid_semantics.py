"""GPU: grid-object semantics of the reference (fteikpy/_base.py:10-50, :230-241).

- ``.grid`` is the live array: writes to it reach the next solve, as in the reference;
- in-place writes to ``.tensor`` or to a device tensor passed to the constructor are seen
  too (the cached slowness follows the tensor's version counter);
- a traveltime grid's gradient is the solved one: resample(), smooth() and writes to the
  times leave ``.gradient`` as it was after the solve;
- ``solve_tensors(out=...)`` refuses a buffer of the wrong shape, type or layout.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fa(device):
    import fteikpy_amd

    return fteikpy_amd


def _model(shape, seed):
    rng = np.random.default_rng(seed)
    return 1500.0 + 2000.0 * rng.random(shape)


@pytest.mark.parametrize("nd", [2, 3])
def test_writes_to_grid_reach_the_solve(fa, nd):
    shape = (9, 11) if nd == 2 else (7, 9, 8)
    d = (10.0,) * nd
    v = _model(shape, 4)
    src = np.array([23.0, 41.0, 17.0][:nd])
    Eik = fa.Eikonal2D if nd == 2 else fa.Eikonal3D
    solve = oracle.solve2d if nd == 2 else oracle.solve3d
    eik = Eik(v, d)
    first = eik.solve(src, nsweep=2).grid.copy()
    assert np.array_equal(first, solve(1.0 / v, *d, src, nsweep=2)[0])
    eik.grid[2:5] *= 0.5  # the reference's array is live: this edits the model
    v2 = v.copy()
    v2[2:5] *= 0.5
    assert np.array_equal(eik.grid, v2)
    second = eik.solve(src, nsweep=2).grid
    assert np.array_equal(second, solve(1.0 / v2, *d, src, nsweep=2)[0])
    assert not np.array_equal(first, second)
    # slices read through __getitem__ are views of the same live array
    eik[0][...] = 1000.0
    v2[0] = 1000.0
    assert np.array_equal(eik.solve(src, nsweep=2).grid, solve(1.0 / v2, *d, src, nsweep=2)[0])


def test_in_place_tensor_writes_reach_the_solve(fa):
    import torch

    shape, d = (7, 9, 8), (10.0, 10.0, 10.0)
    v = _model(shape, 5)
    src = np.array([23.0, 41.0, 17.0])
    vt = torch.tensor(v, device="cuda")
    eik = fa.Eikonal3D(vt, d)
    eik.solve(src, nsweep=2)
    vt.mul_(1.25)  # the caller's tensor, kept by reference
    ref = oracle.solve3d(1.0 / (v * 1.25), *d, src, nsweep=2)[0]
    assert np.array_equal(eik.solve(src, nsweep=2).grid, ref)
    eik.tensor[0].fill_(2500.0)  # through the accessor
    v3 = v * 1.25
    v3[0] = 2500.0
    assert np.array_equal(eik.solve(src, nsweep=2).grid,
                          oracle.solve3d(1.0 / v3, *d, src, nsweep=2)[0])


@pytest.mark.parametrize("nd", [2, 3])
def test_gradient_belongs_to_the_solved_times(fa, nd):
    shape = (9, 11) if nd == 2 else (7, 9, 8)
    d = (10.0,) * nd
    Eik = fa.Eikonal2D if nd == 2 else fa.Eikonal3D
    src = np.array([23.0, 41.0, 17.0][:nd])
    eik = Eik(_model(shape, 6), d)

    tt = eik.solve(src, nsweep=3, return_gradient=True)
    tt.resample(tuple(2 * n for n in tt.shape))
    assert tt.shape == tuple(2 * (n + 1) for n in shape)
    ref = eik.solve(src, nsweep=3, return_gradient=True)
    for g, r in zip(tt.gradient, ref.gradient):
        assert np.array_equal(g.grid, r.grid)  # the solve-time gradient, its shape kept
    with pytest.raises(ValueError, match="resampled"):
        tt.raytrace(src + 5.0)

    tt = eik.solve(src, nsweep=3, return_gradient=True)
    tt.smooth(15.0)
    for g, r in zip(tt.gradient, ref.gradient):
        assert np.array_equal(g.grid, r.grid)
    tt.raytrace(src + 5.0)  # same shape: allowed, as in the reference

    tt = eik.solve(src, nsweep=3, return_gradient=True)
    tt.grid[...] = 7.0
    for g, r in zip(tt.gradient, ref.gradient):
        assert np.array_equal(g.grid, r.grid)


def test_out_buffer_is_validated(fa):
    import torch

    shape, d = (6, 7, 8), (10.0, 10.0, 10.0)
    eik = fa.Eikonal3D(_model(shape, 7), d)
    src = np.array([[11.0, 12.0, 13.0], [31.0, 32.0, 33.0]])
    good = torch.empty((2, 7, 8, 9), dtype=torch.float64, device="cuda")
    tt = eik.solve_tensors(src, nsweep=2, out=good)[0]
    assert tt.data_ptr() == good.data_ptr()
    for bad in (torch.empty((1, 7, 8, 9), dtype=torch.float64, device="cuda"),
                torch.empty((2, 7, 8, 9), dtype=torch.float32, device="cuda"),
                torch.empty((2, 7, 9, 8), dtype=torch.float64, device="cuda").transpose(2, 3),
                torch.empty((2, 7, 8, 9), dtype=torch.float64)):
        with pytest.raises(ValueError, match="out must be"):
            eik.solve_tensors(src, nsweep=2, out=bad)
