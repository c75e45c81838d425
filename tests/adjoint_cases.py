"""Shared adjoint test cases: models, sources, receivers and the oracle's forward map.

The adjoint has no counterpart in the reference; it is pinned by central finite
differences of the oracle solver (the restatement of the reference's fteik3d + vinterp3d).
"""
import numpy as np

import oracle
from fteikpy_amd import synthetic


def forward(slow, d, origin, src, rec, nthreads=0):
    """Receiver traveltimes [nsrc, nrec] of converged oracle solves, and the grids."""
    nz, nx, ny = slow.shape
    tt, _, vz, sets = oracle.solve3d(slow, *d, src - origin, nsweep=0, max_sets=300,
                                     nthreads=nthreads)
    z = origin[0] + d[0] * np.arange(nz + 1)
    x = origin[1] + d[1] * np.arange(nx + 1)
    y = origin[2] + d[2] * np.arange(ny + 1)
    T = np.stack([oracle.vinterp3d(z, x, y, tt[s], rec, src[s], vz[s]) for s in range(len(src))])
    return T, tt, sets


def layered_start_case(seed=0, shape=(32, 32, 32), nsrc=4, nrec=64):
    """The gradient workload at 32^3: true model smooth random (+-8% around 1/3000 s/m),
    evaluated at its 1D (depth) average; surface sources and receivers."""
    rng = np.random.default_rng(seed)
    d = (10.0, 10.0, 10.0)
    origin = np.zeros(3)
    noise = synthetic.gaussian_filter_reflect(rng.standard_normal(shape), 4.0)
    noise /= np.abs(noise).max()
    true = (1.0 / 3000.0) * (1.0 + 0.08 * noise)
    start = np.broadcast_to(true.mean(axis=(1, 2))[:, None, None], shape).copy()
    ext = np.array(shape) * np.array(d)
    src = np.column_stack([np.full(nsrc, 5.0), rng.uniform(0, ext[1], nsrc),
                           rng.uniform(0, ext[2], nsrc)])
    rec = np.column_stack([np.zeros(nrec), rng.uniform(0, ext[1], nrec),
                           rng.uniform(0, ext[2], nrec)])
    return true, start, d, origin, src, rec


def rough_case(seed, shape=(16, 18, 20), nsrc=1, nrec=40, contrast=0.2):
    """Smooth random slowness (+-contrast), interior sources, receivers anywhere."""
    rng = np.random.default_rng(seed)
    d = (10.0, 11.0, 12.0)
    origin = np.array([5.0, -3.0, 2.0])
    noise = synthetic.gaussian_filter_reflect(rng.standard_normal(shape), 2.0)
    noise /= np.abs(noise).max()
    slow = (1.0 / 3000.0) * (1.0 + contrast * noise)
    ext = np.array(shape) * np.array(d)
    src = origin + rng.uniform(0.05, 0.95, (3, 3))[:nsrc] * ext
    rec = origin + rng.uniform(0.0, 1.0, (nrec, 3)) * ext
    return slow, d, origin, src, rec


def fd_gradient(slow, d, origin, src, rec, Tobs, cells, rel_h=1e-6, nthreads=0):
    """Central differences of 0.5 * sum((T - Tobs)^2) at the given cells, with the two
    one-sided slopes: (central, forward, backward)."""
    L0 = 0.5 * np.sum((forward(slow, d, origin, src, rec, nthreads)[0] - Tobs) ** 2)
    cen, fwd, bwd = [], [], []
    for c in cells:
        h = rel_h * slow[c]
        sp = slow.copy()
        sp[c] += h
        sm = slow.copy()
        sm[c] -= h
        Lp = 0.5 * np.sum((forward(sp, d, origin, src, rec, nthreads)[0] - Tobs) ** 2)
        Lm = 0.5 * np.sum((forward(sm, d, origin, src, rec, nthreads)[0] - Tobs) ** 2)
        cen.append((Lp - Lm) / (2.0 * h))
        fwd.append((Lp - L0) / h)
        bwd.append((L0 - Lm) / h)
    return np.array(cen), np.array(fwd), np.array(bwd)


def smooth_cells(fd, fwd, bwd, tol=3e-5):
    """Cells where the loss is differentiable across +-h.  On smooth cells the one-sided
    slopes differ by the curvature term only (h f'', below 1e-5 of the largest slope
    here); a kink inside +-h separates them by 1e-4 and more."""
    scale = np.abs(fd).max()
    return np.abs(fwd - bwd) <= tol * scale


def pick_cells(g, rng, n=50):
    """Half of the cells where the gradient is non-zero, half anywhere."""
    nz = np.argwhere(g != 0.0)
    k = min(n // 2, len(nz))
    chosen = [tuple(c) for c in nz[rng.choice(len(nz), k, replace=False)]] if k else []
    while len(chosen) < n:
        chosen.append(tuple(int(rng.integers(0, s)) for s in g.shape))
    return chosen
