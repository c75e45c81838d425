"""GPU adjoint gradients against the oracle (whose adjoints are pinned by central finite
differences of the oracle solver in test_oracle_adjoint.py) and against finite
differences directly.

Bar: dL/dslow within 1e-10 of the largest component of the oracle's on the same grids
(FP64 atomics make the summation order differ); within 1e-5 of central differences
(h = 1e-6 relative) at cells where the loss is differentiable across +-h.
"""
import ctypes

import numpy as np
import pytest

import oracle
from adjoint_cases import (fd_gradient, forward, layered_start_case, pick_cells, rough_case,
                           smooth_cells)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fa(device):
    import fteikpy_amd

    return fteikpy_amd


def _grad(fa, slow, d, origin, src, rec, w, mode, stats=None):
    import torch

    s = torch.tensor(slow, device="cuda", requires_grad=True)
    T = fa.receiver_traveltimes(s, d, src, rec, origin, mode=mode, stats=stats)
    T.backward(torch.as_tensor(w, device="cuda"))
    return T.detach().cpu().numpy(), s.grad.cpu().numpy()


def _close(a, b, tol=1e-10):
    scale = np.abs(b).max()
    assert scale > 0
    assert np.max(np.abs(a - b)) <= tol * scale


@pytest.mark.parametrize("seed", [0, 2, 5])
def test_exact_matches_oracle_rough(fa, seed):
    slow, d, origin, src, rec = rough_case(seed, nsrc=3)
    T, tt, _ = forward(slow, d, origin, src, rec)
    w = np.random.default_rng(seed).standard_normal(T.shape)
    stats = {}
    Tg, g = _grad(fa, slow, d, origin, src, rec, w, "exact", stats)
    assert np.array_equal(np.isnan(Tg), np.isnan(T))
    assert np.allclose(Tg, T, rtol=1e-12, atol=0, equal_nan=True)
    go, _ = oracle.adjoint3d_exact(slow, *d, origin, src, rec, w, max_sets=300)
    _close(g, go)
    assert stats["events"] > 0


def test_exact_matches_oracle_layered(fa):
    true, start, d, origin, src, rec = layered_start_case(seed=3)
    T, tt, _ = forward(true, d, origin, src, rec, nthreads=8)
    w = np.random.default_rng(3).standard_normal(T.shape)
    _, g = _grad(fa, true, d, origin, src, rec, w, "exact")
    go, _ = oracle.adjoint3d_exact(true, *d, origin, src, rec, w, max_sets=300)
    _close(g, go)


@pytest.mark.parametrize("seed", [0, 1, 3])
def test_stencil_matches_oracle(fa, seed):
    slow, d, origin, src, rec = rough_case(seed, nsrc=2)
    T, tt, _ = forward(slow, d, origin, src, rec)
    w = np.random.default_rng(seed).standard_normal(T.shape)
    stats = {}
    _, g = _grad(fa, slow, d, origin, src, rec, w, "stencil", stats)
    go, stale, _ = oracle.adjoint3d(slow, *d, origin, src, tt, rec, w)
    assert stats["stale"] == stale
    _close(g, go)


def test_fd_smooth_random_32(fa):
    """The gradient workload's 32^3 sub-case through the autograd function: 50 cells."""
    import torch

    true, start, d, origin, src, rec = layered_start_case(seed=1)
    T0 = forward(true, d, origin, src, rec, nthreads=8)[0]
    Tobs = T0 * (1.0 + 0.01 * np.random.default_rng(101).standard_normal(T0.shape))
    s = torch.tensor(true, device="cuda", requires_grad=True)
    T = fa.receiver_traveltimes(s, d, src, rec, origin)
    loss = 0.5 * ((T - torch.as_tensor(Tobs, device="cuda")) ** 2).sum()
    loss.backward()
    g = s.grad.cpu().numpy()
    cells = pick_cells(g, np.random.default_rng(1), 50)
    fd, fwd, bwd = fd_gradient(true, d, origin, src, rec, Tobs, cells, nthreads=8)
    ok = smooth_cells(fd, fwd, bwd)
    assert ok.sum() >= 40
    adj = np.array([g[c] for c in cells])
    assert np.max(np.abs(adj - fd)[ok]) <= 1e-5 * np.abs(fd).max()


def test_euler_homogeneity(fa):
    slow, d, origin, src, rec = rough_case(5, nsrc=3, nrec=30)
    T, _, _ = forward(slow, d, origin, src, rec)
    w = np.random.default_rng(1).standard_normal(T.shape)
    _, g = _grad(fa, slow, d, origin, src, rec, w, "exact")
    assert abs(np.sum(slow * g) - np.nansum(w * T)) <= 1e-12 * np.nansum(np.abs(w * T))


def test_history_rollback_restores_initial_grid(fa):
    """Reversing the whole history with zero adjoint returns every grid to its initial
    state: Big everywhere but the 8 seeded corners (every logged old value is right)."""
    import torch

    from fteikpy_amd import _adjoint

    slow, d, origin, src, rec = rough_case(2, nsrc=3)
    sd = torch.tensor(slow, device="cuda")
    tt, vz, sets, hist, events = _adjoint._solve(sd, d, src - origin, 300, True)
    tto, _, vzo, _ = oracle.solve3d(slow, *d, src - origin, nsweep=0, max_sets=300)
    assert np.array_equal(tt.cpu().numpy(), tto)
    assert events > 0
    recd = torch.as_tensor(rec, device="cuda")
    dldt = torch.zeros((len(src), len(rec)), dtype=torch.float64, device="cuda")
    g = _adjoint._adjoint_history(sd, d, origin, src, tt, hist, events, recd, dldt)
    assert not g.cpu().numpy().any()
    back = tt.cpu().numpy()
    big = back == 1.0e5
    assert big.sum() == back.size - 8 * len(src)
    for s in range(len(src)):
        corner = ~big[s]
        zsi, xsi, ysi = (np.floor((src[s] - origin) / np.array(d))).astype(int)
        assert corner[zsi:zsi + 2, xsi:xsi + 2, ysi:ysi + 2].all()


def test_small_history_is_reported(fa):
    import torch

    from fteikpy_amd import _lib
    from fteikpy_amd._grid import _ptr, stream_ptr

    slow, d, origin, src, rec = rough_case(1, nsrc=1)
    L = _lib.lib()
    sd = torch.tensor(slow, device="cuda")
    nz, nx, ny = slow.shape
    tt = torch.empty((1, nz + 1, nx + 1, ny + 1), dtype=torch.float64, device="cuda")
    vz = torch.empty(1, dtype=torch.float64, device="cuda")
    sets = torch.empty(1, dtype=torch.int32, device="cuda")
    wsb = L.fteik3d_workspace_size(nz, nx, ny, 1)
    ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
    hb = L.fteik3d_history_size(16, 8 * 50)
    hist = torch.empty(hb, dtype=torch.uint8, device="cuda")
    rel, rp = _lib.host_f8(src - origin)
    ev = ctypes.c_longlong(0)
    st = L.fteik3d_solve_logged(_ptr(sd), nz, nx, ny, *d, rp, 1, 0, 50, 0, 0, _ptr(tt), _ptr(vz),
                                _ptr(sets), _ptr(ws), wsb, _ptr(hist), hb, ctypes.byref(ev), 0,
                                stream_ptr())
    assert st == _lib.ERR_WORKSPACE
    assert ev.value > 16
    tto, _, _, _ = oracle.solve3d(slow, *d, src[0] - origin, nsweep=0, max_sets=50)
    assert np.array_equal(tt[0].cpu().numpy(), tto)


def test_exact_backward_runs_once(fa):
    import torch

    slow, d, origin, src, rec = rough_case(0, nsrc=1)
    s = torch.tensor(slow, device="cuda", requires_grad=True)
    T = fa.receiver_traveltimes(s, d, src, rec, origin)
    T.sum().backward(retain_graph=True)
    with pytest.raises(RuntimeError, match="once per forward"):
        T.sum().backward()
