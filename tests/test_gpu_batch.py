"""GPU parity of the multi-model batch (fteik_solve_batch): every problem's traveltimes,
signs, vzero and set counts equal its own single-model solve bit for bit, and the oracle's
(pinned to the reference in test_oracle.py), at fixed nsweep and converged."""
import numpy as np
import pytest

import oracle
from fteikpy_amd import synthetic

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fa(device):
    import fteikpy_amd

    return fteikpy_amd


def _small_batch(seed, n2=3, n3=4):
    return synthetic.uneven_batch(seed=seed, n2=n2, n3=n3, side2=(9, 40), side3=((6, 16), (6, 20)),
                                  nsrc=(1, 6))


def _solvers(fa, models):
    out = []
    for v, d, src in models:
        Eik = fa.Eikonal3D if v.ndim == 3 else fa.Eikonal2D
        out.append((Eik(v, d), src))
    return out


@pytest.mark.parametrize("nsweep", [1, 2, None])
def test_batch_equals_single_solves_and_oracle(fa, nsweep):
    models = _small_batch(11)
    probs = _solvers(fa, models)
    print(f"batch nsweep={nsweep}: {[v.shape for v, _, _ in models]}", flush=True)
    res = fa.solve_batch_tensors(probs, nsweep=nsweep, return_gradient=True)
    for (eik, src), (v, d, _), (tt, sgn, vz, sets) in zip(probs, models, res):
        tt1, sgn1, vz1, sets1 = (eik.solve_tensors(src, nsweep=nsweep, return_gradient=True)
                                 if v.ndim == 3 else _solve2d_tensors(eik, src, nsweep))
        assert np.array_equal(tt.cpu().numpy(), tt1.cpu().numpy()), v.shape
        assert np.array_equal(sgn.cpu().numpy(), sgn1.cpu().numpy()), v.shape
        assert np.array_equal(vz, vz1) and np.array_equal(sets, sets1), v.shape
        solve = oracle.solve3d if v.ndim == 3 else oracle.solve2d
        for i, s in enumerate(src):
            ref, _, rvz, k = solve(1.0 / v, *d, s, nsweep=nsweep or 0, grad=True)
            assert np.array_equal(tt[i].cpu().numpy(), ref), (v.shape, i)
            assert vz[i] == rvz
            if nsweep is None:
                assert sets[i] == k


@pytest.mark.parametrize("nsweep", [2, None])
def test_batch_2d_several_blocks(fa, nsweep):
    """2D models 60 to 150 nodes deep (two and three 64-node speculative blocks per z pass)
    beside a small 3D model: times and signs equal the single solves, times the oracle."""
    models = synthetic.uneven_batch(seed=12, n2=4, n3=1, side2=(60, 150),
                                    side3=((6, 10), (6, 10)), nsrc=(1, 4))
    probs = _solvers(fa, models)
    res = fa.solve_batch_tensors(probs, nsweep=nsweep, return_gradient=True)
    for (eik, src), (v, d, _), (tt, sgn, vz, sets) in zip(probs, models, res):
        if v.ndim == 3:
            continue
        tt1, sgn1, vz1, sets1 = _solve2d_tensors(eik, src, nsweep)
        assert np.array_equal(tt.cpu().numpy(), tt1.cpu().numpy()), v.shape
        assert np.array_equal(sgn.cpu().numpy(), sgn1.cpu().numpy()), v.shape
        assert np.array_equal(sets, sets1), v.shape
        ref = oracle.solve2d(1.0 / v, *d, src, nsweep=nsweep or 0)[0]
        assert np.array_equal(tt.cpu().numpy(), ref), v.shape


def _solve2d_tensors(eik, src, nsweep):
    grids = eik.solve(src, nsweep=nsweep, return_gradient=True)
    grids = grids if isinstance(grids, list) else [grids]
    tt = np.stack([g.grid for g in grids])
    sgn = np.stack([g._sign.cpu().numpy() for g in grids])
    import torch

    return (torch.as_tensor(tt), torch.as_tensor(sgn), np.array([g.vzero for g in grids]),
            np.array([g.sets_done for g in grids]))


def test_batch_grids_api_and_tiles(fa):
    """solve_batch returns what solve returns (a grid for a 1D source, else a list), and per
    model brick heights do not change the result."""
    models = _small_batch(12, n2=1, n3=3)
    probs = _solvers(fa, models)
    probs[1] = (probs[1][0], probs[1][1][0])  # one source given as a 1D array
    out = fa.solve_batch(probs, nsweep=None, tiles=[0, 3, 5, 64])
    for (eik, src), got in zip(probs, out):
        want = eik.solve(src, nsweep=None)
        if np.ndim(src) == 1:
            assert not isinstance(got, list)
            got, want = [got], [want]
        for g, w in zip(got, want):
            assert np.array_equal(g.grid, w.grid)
            assert g.sets_done == w.sets_done


@pytest.mark.parametrize("dims", [(3,), (2,)])
def test_batch_single_dimension(fa, dims):
    models = _small_batch(13, n2=3 if 2 in dims else 0, n3=3 if 3 in dims else 0)
    probs = _solvers(fa, models)
    res = fa.solve_batch_tensors(probs, nsweep=None, use_graph=False)
    for (eik, src), (tt, _, _, sets) in zip(probs, res):
        want = eik.solve(src, nsweep=None)
        want = want if isinstance(want, list) else [want]
        for i, w in enumerate(want):
            assert np.array_equal(tt[i].cpu().numpy(), w.grid)
            assert sets[i] == w.sets_done


def test_batch_source_out_of_bound(fa):
    models = _small_batch(14, n2=1, n3=2)
    probs = _solvers(fa, models)
    v = models[2][0]
    probs[2] = (probs[2][0], np.array([[0.0, 0.0, v.shape[2] * 10.0 + 1.0]]))
    with pytest.raises(ValueError, match="source out of bound"):
        fa.solve_batch(probs, nsweep=1)
