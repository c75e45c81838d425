"""Pins of the oracle's adjoint gradients: oracle_adjoint3d_exact (reverse mode through the
sweep history) and oracle_adjoint3d (converged upwind stencil).  Central finite
differences of the oracle solver (h = 1e-6 relative) at cells where the loss is
differentiable across +-h, Euler homogeneity, linearity and closed forms."""
import numpy as np
import pytest

import oracle
from adjoint_cases import (fd_gradient, forward, layered_start_case, pick_cells, rough_case,
                           smooth_cells)


def _residual_case(case_fn, seed, **kw):
    true, d, origin, src, rec = case_fn(seed, **kw)
    T0 = forward(true, d, origin, src, rec, nthreads=8)[0]
    Tobs = T0 * (1.0 + 0.01 * np.random.default_rng(seed + 100).standard_normal(T0.shape))
    return true, d, origin, src, rec, Tobs, T0


def _check_fd(g, slow, d, origin, src, rec, Tobs, ncells, seed, tol=1e-5):
    cells = pick_cells(g, np.random.default_rng(seed), ncells)
    fd, fwd, bwd = fd_gradient(slow, d, origin, src, rec, Tobs, cells, nthreads=8)
    ok = smooth_cells(fd, fwd, bwd)
    assert ok.sum() >= 0.8 * len(cells)  # kinks inside +-h are the exception
    adj = np.array([g[c] for c in cells])
    scale = np.abs(fd).max()
    assert scale > 0
    err = np.abs(adj - fd)[ok]
    assert err.max() <= tol * scale
    return err.max() / scale


def _true_case(seed, shape=(32, 32, 32)):
    true, start, d, origin, src, rec = layered_start_case(seed=seed, shape=shape)
    return true, d, origin, src, rec


@pytest.mark.parametrize("seed", [1, 2])
def test_exact_fd_smooth_random_32(seed):
    """The gradient workload's 32^3 sub-case on its smooth random (+-8%) model: 50 cells."""
    slow, d, origin, src, rec, Tobs, T = _residual_case(_true_case, seed)
    g, sets = oracle.adjoint3d_exact(slow, *d, origin, src, rec, T - Tobs, max_sets=300)
    _check_fd(g, slow, d, origin, src, rec, Tobs, 50, seed)


def test_exact_fd_layered_start_tie_broken():
    """At the workload's 1D starting model every lateral min() is an exact tie (the loss is
    not differentiable there); broken by 1e-5 relative noise, the sweeps are smooth."""
    true, start, d, origin, src, rec = layered_start_case(seed=2)
    Tobs = forward(true, d, origin, src, rec, nthreads=8)[0]
    m = start * (1.0 + 1e-5 * np.random.default_rng(7).standard_normal(start.shape))
    T = forward(m, d, origin, src, rec, nthreads=8)[0]
    g, _ = oracle.adjoint3d_exact(m, *d, origin, src, rec, T - Tobs, max_sets=300)
    _check_fd(g, m, d, origin, src, rec, Tobs, 50, 5)


@pytest.mark.parametrize("seed", [0, 2, 5])
def test_exact_fd_rough_models(seed):
    """Rough models, 3 interior sources, receivers anywhere (stale nodes are common)."""
    slow, d, origin, src, rec, Tobs, T = _residual_case(rough_case, seed, nsrc=3)
    g, _ = oracle.adjoint3d_exact(slow, *d, origin, src, rec, T - Tobs, max_sets=300)
    _check_fd(g, slow, d, origin, src, rec, Tobs, 30, seed + 10)


@pytest.mark.parametrize("seed", [1, 3])
def test_stencil_equals_exact_without_stale_nodes(seed):
    """When every node's time is reproduced on the final grid, the converged-stencil
    adjoint is the exact one."""
    slow, d, origin, src, rec = rough_case(seed)
    T, tt, _ = forward(slow, d, origin, src, rec)
    w = np.random.default_rng(seed).standard_normal(T.shape)
    g1, stale, _ = oracle.adjoint3d(slow, *d, origin, src, tt, rec, w)
    assert stale == 0
    g2, _ = oracle.adjoint3d_exact(slow, *d, origin, src, rec, w, max_sets=300)
    assert np.allclose(g1, g2, rtol=0, atol=1e-12 * np.abs(g2).max())


def test_stencil_fd_without_stale_nodes():
    slow, d, origin, src, rec, Tobs, T = _residual_case(rough_case, 4)
    T, tt, _ = forward(slow, d, origin, src, rec)
    g, stale, _ = oracle.adjoint3d(slow, *d, origin, src, tt, rec, T - Tobs)
    assert stale == 0
    _check_fd(g, slow, d, origin, src, rec, Tobs, 30, 3)


def test_stale_nodes_are_reported():
    """Strong contrasts leave held values no current candidate reproduces: the stencil
    adjoint counts them, and there it departs from the exact one."""
    slow, d, origin, src, rec = rough_case(0, nsrc=3)
    T, tt, _ = forward(slow, d, origin, src, rec)
    w = np.ones_like(T)
    g1, stale, _ = oracle.adjoint3d(slow, *d, origin, src, tt, rec, w)
    assert stale > 0
    g2, _ = oracle.adjoint3d_exact(slow, *d, origin, src, rec, w, max_sets=300)
    assert not np.allclose(g1, g2, rtol=0, atol=1e-9 * np.abs(g2).max())


@pytest.mark.parametrize("exact", [True, False])
def test_euler_homogeneity(exact):
    """T is homogeneous of degree 1 in the slowness (every comparison of the sweeps is
    scale invariant), so sum_c s_c dT/ds_c = T: with dL/dT = w, sum_c s_c dL/ds_c =
    sum w T."""
    slow, d, origin, src, rec = rough_case(0 if exact else 3, nsrc=3 if exact else 1, nrec=25)
    T, tt, _ = forward(slow, d, origin, src, rec)
    w = np.random.default_rng(0).standard_normal(T.shape)
    if exact:
        g, _ = oracle.adjoint3d_exact(slow, *d, origin, src, rec, w, max_sets=300)
    else:
        g, stale, _ = oracle.adjoint3d(slow, *d, origin, src, tt, rec, w)
        assert stale == 0
    lhs = np.sum(slow * g)
    rhs = np.nansum(w * T)
    assert abs(lhs - rhs) <= 1e-12 * np.nansum(np.abs(w * T))


def test_linearity_in_receiver_weights():
    slow, d, origin, src, rec = rough_case(4, nsrc=2, nrec=20)
    T, tt, _ = forward(slow, d, origin, src, rec)
    rng = np.random.default_rng(2)
    w1, w2 = rng.standard_normal(T.shape), rng.standard_normal(T.shape)
    g = [oracle.adjoint3d_exact(slow, *d, origin, src, rec, w, max_sets=300)[0]
         for w in (w1, w2, 2.0 * w1 - 3.0 * w2)]
    assert np.allclose(g[2], 2.0 * g[0] - 3.0 * g[1], rtol=0, atol=1e-12 * np.abs(g[2]).max())


@pytest.mark.parametrize("exact", [True, False])
def test_source_cell_receiver_closed_form(exact):
    """A receiver in the source cell reads vzero * distance: dT/ds is the distance at the
    source cell and zero elsewhere."""
    slow, d, origin, src, rec = rough_case(0, nsrc=1, nrec=1)
    cell = np.floor((src[0] - origin) / np.array(d))
    r = (origin + (cell + 0.3) * np.array(d))[None]
    T, tt, _ = forward(slow, d, origin, src, r)
    if exact:
        g, _ = oracle.adjoint3d_exact(slow, *d, origin, src, r, np.ones((1, 1)), max_sets=300)
    else:
        g, _, _ = oracle.adjoint3d(slow, *d, origin, src, tt, r, np.ones((1, 1)))
    ci = tuple(int(c) for c in cell)
    assert g[ci] == pytest.approx(np.linalg.norm(src[0] - r[0]), rel=1e-15)
    g[ci] = 0.0
    assert not g.any()


def test_receivers_outside_contribute_nothing():
    slow, d, origin, src, rec = rough_case(1, nsrc=1, nrec=3)
    out = np.array([[origin[0] - 1.0, origin[1], origin[2]]])
    T, tt, _ = forward(slow, d, origin, src, out)
    assert np.isnan(T).all()
    g, _ = oracle.adjoint3d_exact(slow, *d, origin, src, out, np.ones((1, 1)), max_sets=300)
    assert not g.any()
    g, _, _ = oracle.adjoint3d(slow, *d, origin, src, tt, out, np.ones((1, 1)))
    assert not g.any()
