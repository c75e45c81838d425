"""Deterministic synthetic velocity models and acquisition geometries for the benchmark
workloads (numpy only, so every machine regenerates them bit for bit).

- ``model_2d_gaussians``: 2D gradient medium with 4 Gaussian anomalies (101 x 301, 10 m)
- ``model_3d_layers``: 3D layered medium with a smooth random perturbation (200^3, 5 m)
- ``model_3d_inclusions``: 3D linear-gradient medium with 20 spherical inclusions
  (160 x 320 x 320, 12.5 m) with a 16 x 16 surface source grid and 64 x 64 receivers
- ``model_3d_smooth_slowness``: gradient workload (128 x 256 x 256, 10 m)
- ``uneven_batch``: 16 2D and 32 3D independent models of uneven size (1-32 sources each)
"""
import numpy as np


def gaussian_filter_reflect(p, sigma):
    """Separable Gaussian filter (truncate at 4 sigma, reflected edges), FP64."""
    r = int(4.0 * sigma + 0.5)
    x = np.arange(-r, r + 1, dtype=np.float64)
    w = np.exp(-0.5 * (x / sigma) ** 2)
    w /= w.sum()
    out = p
    for ax in range(p.ndim):
        pad = [(0, 0)] * p.ndim
        pad[ax] = (r, r)
        q = np.pad(out, pad, mode="symmetric")
        acc = np.zeros_like(out)
        n = out.shape[ax]
        for t, wt in enumerate(w):
            sl = [slice(None)] * p.ndim
            sl[ax] = slice(t, t + n)
            acc += wt * q[tuple(sl)]
        out = acc
    return out


def model_2d_gaussians(seed=1001):
    """v(z,x) = 1500 + 0.6 z plus 4 Gaussian anomalies of +-300 m/s, sigma 150 m."""
    rng = np.random.default_rng(seed)
    nz, nx, d = 101, 301, 10.0
    z = np.arange(nz)[:, None] * d
    x = np.arange(nx)[None, :] * d
    v = 1500.0 + 0.6 * z + 0.0 * x
    cen = rng.uniform([0.0, 0.0], [1000.0, 3000.0], size=(4, 2))
    for amp, (cz, cx) in zip((300.0, -300.0, 300.0, -300.0), cen):
        v = v + amp * np.exp(-((z - cz) ** 2 + (x - cx) ** 2) / (2.0 * 150.0 ** 2))
    src = np.array([12.3, 1013.7])
    rec = rng.uniform([0.0, 0.0], [1010.0, 3010.0], size=(200, 2))
    return v, (d, d), src, rec


def model_3d_layers(seed=1002, n=200, d=5.0, perturb_sigma=8.0):
    """5 equal horizontal layers of 1800..4500 m/s times (1 + p), max |p| = 5%."""
    rng = np.random.default_rng(seed)
    layers = np.array([1800.0, 2475.0, 3150.0, 3825.0, 4500.0])
    v = np.repeat(layers, int(np.ceil(n / 5)))[:n][:, None, None] * np.ones((n, n, n))
    p = gaussian_filter_reflect(rng.standard_normal((n, n, n)), perturb_sigma)
    v = v * (1.0 + 0.05 * p / np.abs(p).max())
    src = np.array([502.3, 497.1, 503.9]) * (n * d / 1000.0)
    rec = rng.uniform(0.0, n * d, size=(10000, 3))
    return v, (d, d, d), src, rec


def model_3d_inclusions(seed=1003, shape=(160, 320, 320), d=12.5, nsrc_side=16, nrec_side=64):
    """v = 2000 + 0.8 z times (1 +- 0.15) inside 20 spheres (radius 4..16 cells, sign
    alternating); sources on a 16 x 16 grid at z = d/2, receivers on a 64 x 64 grid at z = 0."""
    rng = np.random.default_rng(seed)
    nz, nx, ny = shape
    z = np.arange(nz)[:, None, None] * d
    v = (2000.0 + 0.8 * z) * np.ones(shape)
    iz = np.arange(nz)[:, None, None]
    ix = np.arange(nx)[None, :, None]
    iy = np.arange(ny)[None, None, :]
    for s in range(20):
        c = rng.uniform([0, 0, 0], [nz, nx, ny])
        r = rng.uniform(4.0, 16.0)
        inside = (iz - c[0]) ** 2 + (ix - c[1]) ** 2 + (iy - c[2]) ** 2 <= r * r
        v[inside] *= 1.15 if s % 2 == 0 else 0.85
    ex, ey = nx * d, ny * d
    sx = (np.arange(nsrc_side) + 0.5) * ex / nsrc_side
    sy = (np.arange(nsrc_side) + 0.5) * ey / nsrc_side
    src = np.stack(np.meshgrid(np.full(1, d / 2), sx, sy, indexing="ij"), -1).reshape(-1, 3)
    rx = (np.arange(nrec_side) + 0.5) * ex / nrec_side
    ry = (np.arange(nrec_side) + 0.5) * ey / nrec_side
    rec = np.stack(np.meshgrid(np.zeros(1), rx, ry, indexing="ij"), -1).reshape(-1, 3)
    return v, (d, d, d), src, rec


def model_3d_smooth_slowness(seed=1005, shape=(128, 256, 256), d=10.0, sigma=4.0, amp=0.08,
                             src_grid=(8, 16), rec_grid=(32, 32)):
    """Gradient workload: true slowness smooth random (+-amp around 1/3000 s/m; N(0,1)
    filtered with `sigma` cells, scaled to max |p| = 1) and the starting model, its 1D
    (depth) average.  Sources on a src_grid surface grid at z = d/2, receivers on a
    rec_grid grid at z = 0.  Returns (true, start, (d, d, d), src, rec)."""
    rng = np.random.default_rng(seed)
    p = gaussian_filter_reflect(rng.standard_normal(shape), sigma)
    p /= np.abs(p).max()
    true = (1.0 / 3000.0) * (1.0 + amp * p)
    start = np.ascontiguousarray(
        np.broadcast_to(true.mean(axis=(1, 2))[:, None, None], shape))
    nz, nx, ny = shape
    ex, ey = nx * d, ny * d
    sx = (np.arange(src_grid[0]) + 0.5) * ex / src_grid[0]
    sy = (np.arange(src_grid[1]) + 0.5) * ey / src_grid[1]
    src = np.stack(np.meshgrid(np.full(1, d / 2), sx, sy, indexing="ij"), -1).reshape(-1, 3)
    rx = (np.arange(rec_grid[0]) + 0.5) * ex / rec_grid[0]
    ry = (np.arange(rec_grid[1]) + 0.5) * ey / rec_grid[1]
    rec = np.stack(np.meshgrid(np.zeros(1), rx, ry, indexing="ij"), -1).reshape(-1, 3)
    return true, start, (d, d, d), src, rec


def smooth_random_field(rng, shape, sigma):
    """N(0,1) noise smoothed by a Gaussian of `sigma` cells (periodic, applied in Fourier
    space: fast enough for 10^7-cell models), rescaled to [0, 1]."""
    p = rng.standard_normal(shape)
    f = np.fft.rfftn(p)
    k2 = np.zeros(f.shape)
    for ax, n in enumerate(shape):
        k = np.fft.rfftfreq(n) if ax == len(shape) - 1 else np.fft.fftfreq(n)
        sh = [1] * len(shape)
        sh[ax] = len(k)
        k2 = k2 + (2.0 * np.pi * k.reshape(sh)) ** 2
    p = np.fft.irfftn(f * np.exp(-0.5 * sigma * sigma * k2), s=shape,
                      axes=tuple(range(len(shape))))
    return (p - p.min()) / (p.max() - p.min())


def uneven_batch(seed=1004, n2=16, n3=32, side2=(64, 1024), side3=((48, 192), (48, 256)),
                 d=10.0, nsrc=(1, 32), vel=(1500.0, 6000.0), sigma=(4.0, 12.0)):
    """Uneven batch of independent models: n2 2D models with sides log-uniform in side2,
    n3 3D models with nz log-uniform in side3[0] and nx, ny in side3[1]; velocities in
    `vel` from smooth random fields (sigma uniform in `sigma` cells); nsrc[0]..nsrc[1]
    sources per model, uniform inside it.  Returns a list of (velocity, gridsize, sources),
    2D models first."""
    rng = np.random.default_rng(seed)

    def logu(lo, hi):
        return int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))

    out = []
    for m in range(n2 + n3):
        if m < n2:
            shape = (logu(*side2), logu(*side2))
        else:
            shape = (logu(*side3[0]), logu(*side3[1]), logu(*side3[1]))
        p = smooth_random_field(rng, shape, rng.uniform(*sigma))
        v = vel[0] + (vel[1] - vel[0]) * p
        ns = int(rng.integers(nsrc[0], nsrc[1] + 1))
        ext = np.array(shape, dtype=np.float64) * d
        src = rng.uniform(0.0, 1.0, size=(ns, len(shape))) * ext
        out.append((v, (d,) * len(shape), src))
    return out


def marmousi_like(nz=300, nx=1000, d=0.01):
    """A layered, faulted 2D model of the Marmousi model's size and units (300 x 1000 cells
    of 10 m given in km and km/s, 1.5 to 4.7 km/s): 12 interfaces dipping and undulating
    across the model, offset by two normal faults, velocity increasing downwards."""
    z = (np.arange(nz) + 0.5)[:, None] * d
    x = (np.arange(nx) + 0.5)[None, :] * d
    v = np.full((nz, nx), 1.5)
    vel = np.linspace(1.7, 4.7, 12)
    for k in range(12):
        depth = 0.15 + 0.23 * k + 0.04 * x + 0.06 * np.sin(2.0 * np.pi * x / (3.0 + 0.4 * k))
        depth = depth + 0.12 * (x > 3.7) + 0.09 * (x > 6.4) * (k % 2)
        v = np.where(z >= depth, vel[k], v)
    return v, (d, d)
