"""Eikonal2D / Eikonal3D front-ends, mirroring fteikpy/_solver.py:8-149.

``solve`` keeps the reference's signature and return types (one TraveltimeGrid, or a list
for a batch of sources).  Keyword-only extensions:

- ``nsweep=None`` (or <= 0): sweep until a set changes no node, at most ``max_sets`` sets;
- ``tile_a``: z-extent of the 3D sweep bricks (0 = library default);
- ``group``: sources per workgroup (0 = default; 1, 2, 4, 8); the default is 1 below 8
  sources, else 2, except from 32 sources, where the first 2 sweep sets run 4 per
  workgroup (every source sweeps nearly every brick there) and the later ones 2;
  with both left at 0 (and ``schedule`` at None), a choice recorded by
  ``fteikpy_amd.autotune.tune3d`` for this grid shape and number of sources is used;
- ``schedule``: "overlapped" (each launch runs the bricks, of any pass, whose inputs are
  final, so a pass starts while the previous one finishes its far side; bricks under 8
  planes run pass by pass), "diagonal" (one launch per brick anti-diagonal and pass) or
  "persistent" (one cooperative launch per sweep set, a grid-wide barrier between the
  overlapped list's entries); the results are identical;
  None (the default): the autotuner's recorded choice when ``tile_a`` and ``group`` are
  left at 0 too, else "overlapped";
- ``use_graph``: capture one 3D sweep set as a hipGraph and replay it.
"""
import ctypes

import numpy as np

from . import _lib, autotune
from ._grid import (BaseGrid2D, BaseGrid3D, TraveltimeGrid2D, TraveltimeGrid3D, _ptr, _torch,
                    device, stream_ptr)


_SCHEDULES = {"persistent": 0, "diagonal": 1, "overlapped": 2}


def _nsweep(nsweep):
    return 0 if nsweep is None or nsweep <= 0 else int(nsweep)


class _SlownessCache:
    """slowness = 1.0 / velocity on the device (fteikpy/_solver.py:49), recomputed when the
    velocity tensor is replaced or written in place (torch's version counter of the tensor,
    shared with its views, counts in-place writes; uploading a handed-out .grid is one)."""

    def _slowness(self):
        vel = self._device()
        key = (vel, vel._version)
        cached = getattr(self, "_slow_key", None)
        if cached is None or cached[0] is not vel or cached[1] != key[1]:
            self._slow_dev = (1.0 / vel).contiguous()
            self._slow_key = key
        return self._slow_dev


class Eikonal2D(BaseGrid2D, _SlownessCache):
    """2D Eikonal solver, fteikpy/_solver.py:8-77."""

    def __init__(self, grid, gridsize, origin=None):
        super().__init__(grid=grid, gridsize=gridsize,
                         origin=origin if origin is not None else np.zeros(2, dtype=np.float64))

    def solve(self, sources, nsweep=2, return_gradient=False, *, max_sets=100):
        """Solve the eikonal equation for one source [z, x] or a list of sources."""
        torch = _torch()
        src = np.asarray(sources, dtype=np.float64)
        single = src.ndim == 1
        rel, rp = _lib.host_f8((src - self._origin).reshape(-1, 2))
        n = len(rel)
        nz, nx = self.shape
        dev = device()
        tt = torch.empty((n, nz + 1, nx + 1), dtype=torch.float64, device=dev)
        sgn = torch.empty((n, nz + 1, nx + 1), dtype=torch.uint8, device=dev) if return_gradient else None
        vzero = torch.empty(n, dtype=torch.float64, device=dev)
        sets = torch.empty(n, dtype=torch.int32, device=dev)
        L = _lib.lib()
        wsb = L.fteik2d_workspace_size(n)
        ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        st = L.fteik2d_solve(_ptr(self._slowness()), nz, nx, *self._gridsize, rp, n,
                             _nsweep(nsweep), int(max_sets), int(bool(return_gradient)), _ptr(tt),
                             _ptr(sgn) if sgn is not None else None, _ptr(vzero), _ptr(sets),
                             _ptr(ws), wsb, stream_ptr())
        _lib.check(st)
        vz = vzero.cpu().numpy()
        nsets = sets.cpu().numpy()
        grids = [
            TraveltimeGrid2D(tt[i], self._gridsize, self._origin, src.reshape(-1, 2)[i], None,
                             vz[i], sign=sgn[i] if sgn is not None else None,
                             sets_done=int(nsets[i]))
            for i in range(n)
        ]
        return grids[0] if single else grids


class Eikonal3D(BaseGrid3D, _SlownessCache):
    """3D Eikonal solver, fteikpy/_solver.py:80-149."""

    def __init__(self, grid, gridsize, origin=None):
        super().__init__(grid=grid, gridsize=gridsize,
                         origin=origin if origin is not None else np.zeros(3, dtype=np.float64))

    def solve(self, sources, nsweep=2, return_gradient=False, *, max_sets=100, tile_a=0,
              group=0, schedule=None, use_graph=False):
        """Solve the eikonal equation for one source [z, x, y] or a list of sources."""
        tt, sgn, vz, nsets = self.solve_tensors(sources, nsweep, return_gradient,
                                                max_sets=max_sets, tile_a=tile_a, group=group,
                                                schedule=schedule, use_graph=use_graph)
        src = np.asarray(sources, dtype=np.float64)
        single = src.ndim == 1
        srcs = src.reshape(-1, 3)
        grids = [
            TraveltimeGrid3D(tt[i], self._gridsize, self._origin, srcs[i], None, vz[i],
                             sign=sgn[i] if sgn is not None else None, sets_done=int(nsets[i]))
            for i in range(len(srcs))
        ]
        return grids[0] if single else grids

    def solve_tensors(self, sources, nsweep=2, return_gradient=False, *, max_sets=100, tile_a=0,
                      group=0, schedule=None, use_graph=False, out=None):
        """Batched solve returning device tensors: (tt [n, nz+1, nx+1, ny+1], signs or None,
        vzero (numpy [n]), sets executed (numpy [n])).  `out` may pass a preallocated tt."""
        torch = _torch()
        src = np.asarray(sources, dtype=np.float64)
        rel, rp = _lib.host_f8((src - self._origin).reshape(-1, 3))
        n = len(rel)
        nz, nx, ny = self.shape
        dev = device()
        shape = (n, nz + 1, nx + 1, ny + 1)
        if out is not None:
            if (tuple(out.shape) != shape or out.dtype != torch.float64
                    or out.device.type != "cuda" or not out.is_contiguous()):
                raise ValueError(f"out must be a contiguous float64 device tensor of shape "
                                 f"{shape}")
            tt = out
        else:
            tt = torch.empty(shape, dtype=torch.float64, device=dev)
        sgn = torch.empty((n, nz + 1, nx + 1, ny + 1), dtype=torch.uint8,
                          device=dev) if return_gradient else None
        vzero = torch.empty(n, dtype=torch.float64, device=dev)
        sets = torch.empty(n, dtype=torch.int32, device=dev)
        L = _lib.lib()
        wsb = L.fteik3d_workspace_size(nz, nx, ny, n)
        ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        slow = self._slowness()
        if not tile_a and not group and schedule is None:
            tile_a, group, schedule = autotune.choice3d(self.shape, n)
        if schedule is None:
            schedule = "overlapped"
        args = (_ptr(slow), nz, nx, ny, *self._gridsize, rp, n, _nsweep(nsweep), int(max_sets),
                int(bool(return_gradient)), int(tile_a), int(group),
                _SCHEDULES[schedule], _ptr(tt),
                _ptr(sgn) if sgn is not None else None, _ptr(vzero), _ptr(sets), _ptr(ws), wsb)
        cur = torch.cuda.current_stream()
        if use_graph and cur.cuda_stream != 0:
            st = L.fteik3d_solve(*args, 1, ctypes.c_void_p(cur.cuda_stream))
        elif use_graph:
            # capture needs a non-default stream; order it after the current one
            side = _side_stream()
            side.wait_stream(cur)
            with torch.cuda.stream(side):
                st = L.fteik3d_solve(*args, 1, ctypes.c_void_p(side.cuda_stream))
            cur.wait_stream(side)
            for t in (slow, tt, vzero, sets, ws) + ((sgn,) if sgn is not None else ()):
                t.record_stream(side)
        else:
            st = L.fteik3d_solve(*args, 0, stream_ptr())
        _lib.check(st)
        return tt, sgn, vzero.cpu().numpy(), sets.cpu().numpy()


_SIDE = {}


def _side_stream():
    torch = _torch()
    d = torch.cuda.current_device()
    if d not in _SIDE:
        _SIDE[d] = torch.cuda.Stream(device=d)
    return _SIDE[d]
