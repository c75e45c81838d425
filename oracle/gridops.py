"""Oracle (test infrastructure only): numpy restatement of the grid preparation the
reference delegates to scipy -- BaseGrid2D/3D.smooth (fteikpy/_base.py:114-123, :203-212,
scipy.ndimage.gaussian_filter) and BaseGrid2D/3D.resample (fteikpy/_base.py:81-112,
:165-201, scipy.interpolate.RegularGridInterpolator).  scipy is an optional dependency of
the reference (setup.cfg extras, unpinned); this restates the published algorithms of the
scipy installed here (1.15) in their operation order, and tests/test_oracle_gridops.py
pins it bit for bit against scipy itself and against the reference's own tests.
"""
import itertools

import numpy as np


def reflect_index(i, n):
    """scipy.ndimage 'reflect' boundary (d c b a | a b c d | d c b a), period 2n."""
    p = 2 * n
    i = np.mod(i, p)
    return np.where(i < n, i, p - 1 - i)


def gaussian_kernel(sd, truncate=4.0):
    """Normalised sampled Gaussian of radius int(truncate * sd + 0.5)."""
    radius = int(truncate * sd + 0.5)
    x = np.arange(-radius, radius + 1)
    phi = np.exp(-0.5 / (sd * sd) * x ** 2)
    return phi / phi.sum(), radius


def correlate1d_symmetric(a, w, radius, axis):
    """a[i] w[r] + sum_{j=r..1} (a[i-j] + a[i+j]) w[r-j] along `axis`, reflect boundaries."""
    a = np.moveaxis(np.asarray(a, dtype=np.float64), axis, -1)
    n = a.shape[-1]
    i = np.arange(n)
    out = a * w[radius]
    for j in range(radius, 0, -1):
        out = out + (a[..., reflect_index(i - j, n)] + a[..., reflect_index(i + j, n)]) * w[radius - j]
    return np.moveaxis(out, -1, axis)


def gaussian_filter(a, sigmas):
    """Axis 0 first; axes with sigma <= 1e-15 are skipped."""
    out = np.asarray(a, dtype=np.float64)
    for axis, sd in enumerate(sigmas):
        sd = float(sd)
        if sd > 1e-15:
            w, r = gaussian_kernel(sd)
            out = correlate1d_symmetric(out, w[::-1], r, axis)
    return out


def smooth(grid, gridsize, sigma):
    """BaseGrid.smooth: sigma in meters, per axis or scalar (fteikpy/_base.py:114-123)."""
    grid = np.asarray(grid, dtype=np.float64)
    sigma = np.full(grid.ndim, sigma) if np.ndim(sigma) == 0 else np.asarray(sigma)
    return gaussian_filter(grid, sigma / np.asarray(gridsize, dtype=np.float64))


def _cells(axis, q):
    i = np.clip(np.searchsorted(axis, q, side="right") - 1, 0, len(axis) - 2)
    return i, (q - axis[i]) / (axis[i + 1] - axis[i])


def resample(grid, gridsize, origin, new_shape, method="linear"):
    """BaseGrid.resample's values (fteikpy/_base.py:81-112, :165-201)."""
    grid = np.asarray(grid, dtype=np.float64)
    axes = [origin[a] + gridsize[a] * np.arange(n) for a, n in enumerate(grid.shape)]
    qs = [np.linspace(ax[0], ax[-1], m) for ax, m in zip(axes, new_shape)]
    cells = [_cells(ax, q) for ax, q in zip(axes, qs)]
    ii = np.meshgrid(*[c[0] for c in cells], indexing="ij")
    yy = np.meshgrid(*[c[1] for c in cells], indexing="ij")
    if method == "nearest":
        return grid[tuple(np.where(y <= 0.5, i, i + 1) for i, y in zip(ii, yy))]
    if grid.ndim == 2:
        (i0, i1), (y0, y1) = ii, yy
        return (grid[i0, i1] * (1 - y0) * (1 - y1) + grid[i0, i1 + 1] * (1 - y0) * y1
                + grid[i0 + 1, i1] * y0 * (1 - y1) + grid[i0 + 1, i1 + 1] * y0 * y1)
    value = np.zeros(tuple(new_shape))
    for corner in itertools.product((0, 1), repeat=grid.ndim):
        w = np.ones(tuple(new_shape))
        for c, y in zip(corner, yy):
            w = w * (y if c else 1 - y)
        value = value + grid[tuple(i + c for i, c in zip(ii, corner))] * w
    return value
