"""Device-resident grid objects with the reference's interface.

Mirrors fteikpy/_base.py:10-241 (BaseGrid, BaseGrid2D/3D, BaseTraveltime) and
fteikpy/_grid.py:8-278 (Grid2D/3D, TraveltimeGrid2D/3D).  Arrays live in HIP device memory
as torch float64 tensors.  ``.grid`` returns a host numpy array that is live as the
reference's (fteikpy/_base.py:22-25): in-place writes to it are seen by the next device
operation on the grid (solve, interpolation, resample, smooth, gradient, rays), which first
uploads it.  ``.tensor`` returns the device tensor and makes it the live copy again: a
later ``.grid`` copies it back, and writes to a numpy array handed out before are then no
longer seen.  Interpolation, apparent-velocity traveltime interpolation, gradients and ray
tracing run as HIP kernels through the C ABI.
"""
import ctypes

import numpy as np

from . import _lib


def _torch():
    import torch

    return torch


def device():
    """The current HIP device; raises when there is none (no CPU fallback)."""
    torch = _torch()
    if not torch.cuda.is_available():
        raise RuntimeError("fteikpy_amd needs a HIP device (MI355X); none is available")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def to_device(a, dtype=None):
    """float64 (or `dtype`) tensor on the current device; views of device tensors are kept."""
    torch = _torch()
    dtype = dtype or torch.float64
    if isinstance(a, torch.Tensor):
        if a.device.type == "cuda" and a.dtype == dtype:
            return a
        return a.to(device=device(), dtype=dtype)
    return torch.as_tensor(np.asarray(a, dtype=np.float64), device=device()).to(dtype)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _element_stride(t):
    """Element stride k such that t[i0,...] sits at ((i0*n1 + i1)*n2 + ...)*k, or None."""
    k = t.stride(-1)
    expect = k
    for n, s in zip(reversed(t.shape), reversed(t.stride())):
        if s != expect:
            return None
        expect *= n
    return k


def _keep_gradient(grid):
    """Before the times of a traveltime grid can change (resample(), smooth(), writes to a
    handed-out .grid), materialise the gradient that belongs to the solved times: the
    reference computes it at solve time and keeps it."""
    keep = getattr(grid, "_keep_solve_gradient", None)
    if keep is not None:
        keep()


class BaseGrid:
    """Grid storage and properties, fteikpy/_base.py:10-50."""

    _ndim = None

    def __init__(self, grid, gridsize, origin, **kwargs):
        super().__init__(**kwargs)
        self._dev = to_device(grid)
        self._host = None  # host array handed out by .grid, the live copy while not None
        self._gridsize = tuple(float(x) for x in gridsize)
        self._origin = np.asarray(origin, dtype=np.float64)

    def __getitem__(self, islice):
        return self.grid[islice]

    @property
    def grid(self):
        """Grid values as a live host numpy array (see the module docstring)."""
        if self._host is None:
            _keep_gradient(self)  # writes to the times must not reach the solve's gradient
            self._host = self._dev.detach().cpu().numpy()
        return self._host

    @property
    def tensor(self):
        """Grid values as a device tensor (from now on the live copy)."""
        dev = self._device()
        self._host = None
        return dev

    def _device(self):
        """The device tensor, after uploading the handed-out host array if there is one."""
        if self._host is not None:
            if tuple(self._host.shape) != tuple(self._dev.shape):
                raise ValueError("the grid array changed shape; assign a new grid instead")
            self._dev.copy_(_torch().from_numpy(np.ascontiguousarray(self._host, np.float64)))
        return self._dev

    @property
    def gridsize(self):
        return self._gridsize

    @property
    def origin(self):
        return self._origin

    @property
    def shape(self):
        return tuple(self._dev.shape)

    @property
    def size(self):
        return int(self._dev.numel())

    @property
    def ndim(self):
        return self._dev.dim()

    def _axis(self, a):
        return self._origin[a] + self._gridsize[a] * np.arange(self.shape[a])

    def _strided(self):
        """(tensor, element stride) usable by the interpolation kernels."""
        dev = self._device()
        k = _element_stride(dev)
        if k is None:
            return dev.contiguous(), 1
        return dev, k

    def resample(self, new_shape, method="linear"):
        """Resample the grid on the device (fteikpy/_base.py:81-112, :165-201): values at
        linspace(axis[0], axis[-1], new_shape[a]) per axis by 'linear' or 'nearest'
        interpolation, as scipy's RegularGridInterpolator computes them (bit for bit)."""
        torch = _torch()
        if method not in ("linear", "nearest"):
            raise ValueError(f"Method '{method}' is not defined")
        nd = self._ndim
        new_shape = tuple(int(n) for n in new_shape)
        if len(new_shape) != nd:
            raise ValueError(f"new_shape must have {nd} entries")
        if min(self.shape) < 2:
            raise ValueError("resampling needs at least 2 nodes per axis")
        _keep_gradient(self)
        src = self._device().contiguous()
        out = torch.empty(new_shape, dtype=torch.float64, device=device())
        shp = (ctypes.c_int * nd)(*self.shape)
        nshp = (ctypes.c_int * nd)(*new_shape)
        o, op = _lib.host_f8(self._origin)
        d, dp = _lib.host_f8(self._gridsize)
        L = _lib.lib()
        wsb = L.fteik_resample_workspace_size(nd, nshp)
        ws = torch.empty(wsb, dtype=torch.uint8, device=device())
        st = L.fteik_resample(_ptr(src), nd, shp, op, dp, nshp, 0 if method == "linear" else 1,
                              _ptr(out), _ptr(ws), wsb, stream_ptr())
        _lib.check(st)
        self._dev = out
        self._host = None
        # as the reference computes it: after the grid is replaced, so self.shape is
        # already new_shape (fteikpy/_base.py:110-112)
        self._gridsize = tuple(a * b / c for a, b, c in zip(self.gridsize, self.shape, new_shape))

    def smooth(self, sigma):
        """Gaussian smoothing on the device (fteikpy/_base.py:114-123, :203-212): sigma in
        meters (scalar or one per axis), scipy.ndimage.gaussian_filter's kernel (truncated at
        4 sigma, normalised) and 'reflect' boundaries, bit for bit."""
        torch = _torch()
        nd = self._ndim
        sigma = np.full(nd, sigma) if np.ndim(sigma) == 0 else np.asarray(sigma)
        sig = sigma / np.asarray(self._gridsize)
        _keep_gradient(self)
        cur = self._device().contiguous()
        shp = (ctypes.c_int * nd)(*self.shape)
        L = _lib.lib()
        for axis in range(nd):
            sd = float(sig[axis])
            if not sd > 1e-15:
                continue
            radius = int(4.0 * sd + 0.5)
            x = np.arange(-radius, radius + 1)
            phi = np.exp(-0.5 / (sd * sd) * x ** 2)
            phi = phi / phi.sum()
            w = torch.as_tensor(np.ascontiguousarray(phi[::-1]), device=device())
            nxt = torch.empty_like(cur)
            st = L.fteik_correlate1d(_ptr(cur), _ptr(nxt), nd, shp, axis, _ptr(w), radius,
                                     stream_ptr())
            _lib.check(st)
            cur = nxt
        self._dev = cur
        self._host = None

    def _interp(self, points, fill_value):
        torch = _torch()
        nd = self._ndim
        q = np.asarray(points, dtype=np.float64)
        single = q.ndim == 1
        q2 = q.reshape(-1, nd)
        qd = torch.as_tensor(np.ascontiguousarray(q2), device=device())
        out = torch.empty(len(q2), dtype=torch.float64, device=device())
        v, k = self._strided()
        o, op = _lib.host_f8(self._origin)
        d, dp = _lib.host_f8(self._gridsize)
        L = _lib.lib()
        if nd == 3:
            st = L.fteik3d_interp(_ptr(v), *self.shape, k, op, dp, _ptr(qd), len(q2),
                                  float(fill_value), _ptr(out), stream_ptr())
        else:
            st = L.fteik2d_interp(_ptr(v), *self.shape, k, op, dp, _ptr(qd), len(q2),
                                  float(fill_value), _ptr(out), stream_ptr())
        _lib.check(st)
        res = out.cpu().numpy()
        return float(res[0]) if single else res


class BaseGrid2D(BaseGrid):
    """fteikpy/_base.py:53-133"""

    _ndim = 2

    def __call__(self, points, fill_value=np.nan):
        """Bilinear interpolation (fteikpy/_base.py:56-79)."""
        return self._interp(points, fill_value)

    @property
    def zaxis(self):
        return self._axis(0)

    @property
    def xaxis(self):
        return self._axis(1)


class BaseGrid3D(BaseGrid):
    """fteikpy/_base.py:136-227"""

    _ndim = 3

    def __call__(self, points, fill_value=np.nan):
        """Trilinear interpolation (fteikpy/_base.py:139-163)."""
        return self._interp(points, fill_value)

    @property
    def zaxis(self):
        return self._axis(0)

    @property
    def xaxis(self):
        return self._axis(1)

    @property
    def yaxis(self):
        return self._axis(2)


class BaseTraveltime:
    """fteikpy/_base.py:230-241, plus the device sign grid the gradient is built from."""

    def __init__(self, source, gradient, vzero, sign=None, sets_done=None, **kwargs):
        super().__init__(**kwargs)
        self._source = np.asarray(source, dtype=np.float64)
        self._gradient_dev = to_device(gradient) if gradient is not None else None
        self._sign = sign
        self._vzero = float(vzero)
        self._sets_done = sets_done

    @property
    def source(self):
        return self._source

    @property
    def vzero(self):
        """Slowness of the source cell."""
        return self._vzero

    @property
    def sets_done(self):
        """Sweep sets executed by the solver (the last one changed no node if converged)."""
        return self._sets_done

    def _rel_source(self):
        return self._source - self._origin

    def _keep_solve_gradient(self):
        """The gradient of the solved times (fteikpy/_base.py:230-236 stores it at
        construction): computed from the sign grid now, before resample() or smooth()
        replace the times it belongs to."""
        if self._gradient_dev is None and self._sign is not None:
            self._gradient_tensor()

    def _gradient_tensor(self):
        if self._gradient_dev is None:
            if self._sign is None:
                raise ValueError(
                    "no gradient grid, use option `return_gradient` to return gradient grids")
            torch = _torch()
            nd = self._ndim
            # the sign grid belongs to the solved times: resample() and smooth() keep the
            # gradient materialised first, as the reference keeps its solve-time gradient
            if tuple(self._sign.shape) != self.shape:
                raise RuntimeError("sign grid does not match the traveltime grid")
            g = torch.empty(self.shape + (nd,), dtype=torch.float64, device=device())
            cells = [n - 1 for n in self.shape]
            src, sp = _lib.host_f8(self._rel_source())
            tt = self._device().contiguous()
            L = _lib.lib()
            if nd == 3:
                st = L.fteik3d_gradient(_ptr(tt), _ptr(self._sign), *cells, *self._gridsize, sp,
                                        self._vzero, _ptr(g), stream_ptr())
            else:
                st = L.fteik2d_gradient(_ptr(tt), _ptr(self._sign), *cells, *self._gridsize, sp,
                                        self._vzero, _ptr(g), stream_ptr())
            _lib.check(st)
            self._gradient_dev = g
        return self._gradient_dev

    def _vinterp(self, points, fill_value):
        torch = _torch()
        nd = self._ndim
        q = np.asarray(points, dtype=np.float64)
        single = q.ndim == 1
        q2 = np.ascontiguousarray(q.reshape(-1, nd))
        qd = torch.as_tensor(q2, device=device())
        out = torch.empty(len(q2), dtype=torch.float64, device=device())
        tt = self._device().contiguous()
        o, op = _lib.host_f8(self._origin)
        d, dp = _lib.host_f8(self._gridsize)
        s, sp = _lib.host_f8(self._source)
        L = _lib.lib()
        fn = L.fteik3d_vinterp if nd == 3 else L.fteik2d_vinterp
        st = fn(_ptr(tt), *self.shape, op, dp, _ptr(qd), len(q2), sp, self._vzero,
                float(fill_value), _ptr(out), stream_ptr())
        _lib.check(st)
        res = out.cpu().numpy()
        return float(res[0]) if single else res

    def _raytrace(self, points, stepsize, max_step, honor_grid):
        torch = _torch()
        nd = self._ndim
        grad = self._gradient_tensor()
        stepsize, max_step = _ray_steps(self, stepsize, max_step, honor_grid)
        if tuple(grad.shape[:-1]) != self.shape:
            raise ValueError("the gradient grid belongs to the solved traveltime grid, which "
                             "was resampled since: rays need the solved shape")
        p = np.asarray(points, dtype=np.float64)
        single = p.ndim == 1
        p2 = np.ascontiguousarray(p.reshape(-1, nd))
        n = len(p2)
        pd = torch.as_tensor(p2, device=device())
        rays = torch.empty((n, int(max_step), nd), dtype=torch.float64, device=device())
        counts = torch.empty(n, dtype=torch.int32, device=device())
        status = torch.empty(n, dtype=torch.int32, device=device())
        o, op = _lib.host_f8(self._origin)
        d, dp = _lib.host_f8(self._gridsize)
        s, sp = _lib.host_f8(self._source)
        L = _lib.lib()
        fn = L.fteik3d_raytrace if nd == 3 else L.fteik2d_raytrace
        st = fn(_ptr(grad.contiguous()), *self.shape, op, dp, _ptr(pd), n, sp, float(stepsize),
                int(max_step), int(bool(honor_grid)), _ptr(rays), _ptr(counts), _ptr(status),
                stream_ptr())
        _lib.check(st)
        status = status.cpu().numpy()
        for code in (_lib.ERR_ENDPOINT_OOB, _lib.ERR_MAX_STEP):
            if (status == code).any():
                _lib.check(code)
        # only the used prefix of each ray comes back; returned source first
        out = _ray_lists(rays, counts.cpu().numpy(), n)
        return out[0] if single else out


def _ray_steps(grid, stepsize, max_step, honor_grid):
    """The step length and step budget TraveltimeGrid.raytrace uses (fteikpy/_grid.py:
    103-145, :221-265)."""
    if honor_grid or not stepsize:
        stepsize = np.min(grid._gridsize)
    if not max_step:
        max_dist = 2.0 * sum((n * d) ** 2 for n, d in zip(grid.shape, grid._gridsize)) ** 0.5
        max_step = int(max_dist / stepsize)
    return float(stepsize), int(max_step)


def _ray_lists(rays, counts, n):
    """The used prefix of each ray in the work buffer, returned source first."""
    lmax = int(counts.max()) + 1 if n else 0
    host = rays[:, :lmax].cpu().numpy()
    return [host[r, counts[r]::-1].copy() for r in range(n)]


def raytrace_batch(grids, points, stepsize=None, max_step=None, honor_grid=False):
    """Rays of several 3D traveltime grids of one model in one launch.

    No counterpart in the reference, which traces one grid's rays per
    ``TraveltimeGrid3D.raytrace`` call (fteikpy/_grid.py:221-265).  `grids`: traveltime
    grids of the same shape, spacing and origin (one per source, as a batched solve returns
    them); `points`: one entry per grid, its end points ([z, x, y] or an array of them).
    Returns, per grid, what ``grid.raytrace(points_i, stepsize, max_step, honor_grid)``
    returns, bit for bit, with every grid's rays traced together (C entry point
    ``fteik3d_raytrace_batch``).  Each grid's gradient grids are materialised on the device
    (as ``raytrace`` does); for many sources of a large model, pass the grids in chunks."""
    torch = _torch()
    grids = list(grids)
    if not grids:
        return []
    if not all(isinstance(g, TraveltimeGrid3D) for g in grids):
        raise TypeError("raytrace_batch traces TraveltimeGrid3D grids")
    g0 = grids[0]
    for g in grids[1:]:
        if (g.shape != g0.shape or not np.array_equal(g._gridsize, g0._gridsize)
                or not np.array_equal(g._origin, g0._origin)):
            raise ValueError("raytrace_batch needs grids of one model (same shape, gridsize "
                             "and origin)")
    if len(points) != len(grids):
        raise ValueError("raytrace_batch needs one entry of end points per grid")
    pts = [np.asarray(p, dtype=np.float64) for p in points]
    stepsize, max_step = _ray_steps(g0, stepsize, max_step, honor_grid)
    grads = [g._gradient_tensor().contiguous() for g in grids]
    for g, gr in zip(grids, grads):
        if tuple(gr.shape[:-1]) != g.shape:
            raise ValueError("the gradient grid belongs to the solved traveltime grid, which "
                             "was resampled since: rays need the solved shape")
    flat = [np.ascontiguousarray(p.reshape(-1, 3)) for p in pts]
    sizes = [len(f) for f in flat]
    n = int(sum(sizes))
    dev = device()
    if n == 0:
        return [[] for _ in pts]
    pd = torch.as_tensor(np.concatenate(flat), device=dev)
    which = torch.as_tensor(np.repeat(np.arange(len(grids), dtype=np.int32), sizes), device=dev)
    gptr = torch.as_tensor(np.array([gr.data_ptr() for gr in grads], dtype=np.uint64)
                           .view(np.int64), device=dev)
    srcs = torch.as_tensor(np.stack([np.asarray(g._source, dtype=np.float64) for g in grids]),
                           device=dev)
    rays = torch.empty((n, max_step, 3), dtype=torch.float64, device=dev)
    counts = torch.empty(n, dtype=torch.int32, device=dev)
    status = torch.empty(n, dtype=torch.int32, device=dev)
    o, op = _lib.host_f8(g0._origin)
    d, dp = _lib.host_f8(g0._gridsize)
    L = _lib.lib()
    st = L.fteik3d_raytrace_batch(_ptr(gptr), _ptr(which), _ptr(srcs), len(grids), *g0.shape,
                                  op, dp, _ptr(pd), n, stepsize, max_step,
                                  int(bool(honor_grid)), _ptr(rays), _ptr(counts), _ptr(status),
                                  stream_ptr())
    _lib.check(st)
    status = status.cpu().numpy()
    for code in (_lib.ERR_ARG, _lib.ERR_ENDPOINT_OOB, _lib.ERR_MAX_STEP):
        if (status == code).any():
            _lib.check(code)
    out_all = _ray_lists(rays, counts.cpu().numpy(), n)
    out, k = [], 0
    for p, m in zip(pts, sizes):
        part = out_all[k:k + m]
        k += m
        out.append(part[0] if p.ndim == 1 else part)
    del grads
    return out


class Grid2D(BaseGrid2D):
    """2D grid, fteikpy/_grid.py:8-23."""

    def __init__(self, grid, gridsize, origin):
        super().__init__(grid=grid, gridsize=gridsize, origin=origin)


class Grid3D(BaseGrid3D):
    """3D grid, fteikpy/_grid.py:26-41."""

    def __init__(self, grid, gridsize, origin):
        super().__init__(grid=grid, gridsize=gridsize, origin=origin)


class TraveltimeGrid2D(BaseGrid2D, BaseTraveltime):
    """2D traveltime grid, fteikpy/_grid.py:44-158."""

    def __init__(self, grid, gridsize, origin, source, gradient, vzero, sign=None, sets_done=None):
        super().__init__(grid=grid, gridsize=gridsize, origin=np.asarray(origin, dtype=np.float64),
                         source=source, gradient=gradient, vzero=vzero, sign=sign,
                         sets_done=sets_done)

    def __call__(self, points, fill_value=np.nan):
        """Bilinear apparent velocity interpolation (fteikpy/_grid.py:76-101)."""
        return self._vinterp(points, fill_value)

    def raytrace(self, points, stepsize=None, max_step=None, honor_grid=False):
        """2D a posteriori ray-tracing (fteikpy/_grid.py:103-145)."""
        return self._raytrace(points, stepsize, max_step, honor_grid)

    @property
    def gradient(self):
        """Z and X gradient grids as a list of Grid2D (fteikpy/_grid.py:147-158)."""
        g = self._gradient_tensor()
        return [Grid2D(g[:, :, i], self._gridsize, self._origin) for i in range(2)]


class TraveltimeGrid3D(BaseGrid3D, BaseTraveltime):
    """3D traveltime grid, fteikpy/_grid.py:161-278."""

    def __init__(self, grid, gridsize, origin, source, gradient, vzero, sign=None, sets_done=None):
        super().__init__(grid=grid, gridsize=gridsize, origin=np.asarray(origin, dtype=np.float64),
                         source=source, gradient=gradient, vzero=vzero, sign=sign,
                         sets_done=sets_done)

    def __call__(self, points, fill_value=np.nan):
        """Trilinear apparent velocity interpolation (fteikpy/_grid.py:193-219)."""
        return self._vinterp(points, fill_value)

    def raytrace(self, points, stepsize=None, max_step=None, honor_grid=False):
        """3D a posteriori ray-tracing (fteikpy/_grid.py:221-265)."""
        return self._raytrace(points, stepsize, max_step, honor_grid)

    @property
    def gradient(self):
        """Z, X and Y gradient grids as a list of Grid3D (fteikpy/_grid.py:267-278)."""
        g = self._gradient_tensor()
        return [Grid3D(g[:, :, :, i], self._gridsize, self._origin) for i in range(3)]
