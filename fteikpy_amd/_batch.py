"""Many independent models in one call: ``solve_batch``.

No single counterpart in the reference, which solves one model per
``Eikonal2D/3D.solve`` call (fteikpy/_solver.py:29-77, :101-149).  ``solve_batch`` takes
(solver, sources) pairs and returns, for each, exactly what ``solver.solve(sources, ...)``
returns -- bit for bit -- but all 3D models share every sweep launch (each model sweeps on
its own overlapped schedule, and launch l of a set covers launch l's bricks and active
sources of every model) and the 2D models run meanwhile on a stream of their own, one wave
per source.  C entry point: ``fteik_solve_batch``.
"""
import ctypes

import numpy as np

from . import _lib, autotune
from ._grid import TraveltimeGrid2D, TraveltimeGrid3D, _torch, device, stream_ptr
from ._solver import Eikonal2D, Eikonal3D, _nsweep, _side_stream


def solve_batch_tensors(problems, nsweep=2, return_gradient=False, *, max_sets=100, group=0,
                        use_graph=True, tiles=None):
    """Solve every (solver, sources) pair of `problems` (solver: Eikonal2D or Eikonal3D).

    Returns one (tt, signs or None, vzero (numpy), sets executed (numpy)) tuple per problem,
    tt of shape [nsrc, nz+1, nx+1(, ny+1)] on the device.  `tiles` optionally gives the 3D
    brick height per problem (0 = the library's rule for a batch); left unset, the heights
    ``autotune.tune_batch`` recorded for this batch apply, else the library's rule."""
    torch = _torch()
    dev = device()
    L = _lib.lib()
    n = len(problems)
    if n == 0:
        return []
    if tiles is None:
        tiles = autotune.lookup_batch(problems)
    arr = (_lib.Problem * n)()
    keep, outs = [], []
    for i, (eik, sources) in enumerate(problems):
        if not isinstance(eik, (Eikonal2D, Eikonal3D)):
            raise TypeError("problems must hold (Eikonal2D | Eikonal3D, sources) pairs")
        nd = 3 if isinstance(eik, Eikonal3D) else 2
        src = np.asarray(sources, dtype=np.float64)
        rel, rp = _lib.host_f8((src - eik._origin).reshape(-1, nd))
        ns = len(rel)
        shape = tuple(eik.shape)
        nodes = tuple(s + 1 for s in shape)
        tt = torch.empty((ns,) + nodes, dtype=torch.float64, device=dev)
        sgn = torch.empty((ns,) + nodes, dtype=torch.uint8, device=dev) if return_gradient else None
        vz = torch.empty(ns, dtype=torch.float64, device=dev)
        sets = torch.empty(ns, dtype=torch.int32, device=dev)
        slow = eik._slowness()
        tile = int(tiles[i]) if tiles is not None and nd == 3 else 0
        p = arr[i]
        p.ndim = nd
        p.nz, p.nx = shape[0], shape[1]
        p.ny = shape[2] if nd == 3 else 0
        p.dz, p.dx = eik._gridsize[0], eik._gridsize[1]
        p.dy = eik._gridsize[2] if nd == 3 else 0.0
        p.slow = slow.data_ptr()
        p.src = rp
        p.nsrc = ns
        p.tile_a = tile
        p.tt = tt.data_ptr()
        p.sgn = sgn.data_ptr() if sgn is not None else None
        p.vzero = vz.data_ptr()
        p.sets_done = sets.data_ptr()
        keep += [rel, slow]
        outs.append((tt, sgn, vz, sets))
    wsb = L.fteik_batch_workspace_size(arr, n)
    ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=dev)
    failed = ctypes.c_int(-1)
    cur = torch.cuda.current_stream()
    if use_graph and cur.cuda_stream == 0:
        # graph capture needs a non-default stream; order it after the current one
        side = _side_stream()
        side.wait_stream(cur)
        with torch.cuda.stream(side):
            st = L.fteik_solve_batch(arr, n, _nsweep(nsweep), int(max_sets),
                                     int(bool(return_gradient)), int(group), ws.data_ptr(), wsb, 1,
                                     ctypes.byref(failed), ctypes.c_void_p(side.cuda_stream))
        cur.wait_stream(side)
        for t in [ws] + [x for o in outs for x in o if x is not None] + keep[1::2]:
            t.record_stream(side)
    else:
        st = L.fteik_solve_batch(arr, n, _nsweep(nsweep), int(max_sets),
                                 int(bool(return_gradient)), int(group), ws.data_ptr(), wsb,
                                 int(bool(use_graph)), ctypes.byref(failed), stream_ptr())
    _lib.check(st)
    return [(tt, sgn, vz.cpu().numpy(), sets.cpu().numpy()) for tt, sgn, vz, sets in outs]


def solve_batch(problems, nsweep=2, return_gradient=False, *, max_sets=100, group=0,
                use_graph=True, tiles=None):
    """Solve every (solver, sources) pair of `problems`; entry i is what
    ``problems[i][0].solve(problems[i][1], nsweep, return_gradient)`` returns (a
    TraveltimeGrid for one source [z, x(, y)], else a list of them)."""
    res = solve_batch_tensors(problems, nsweep, return_gradient, max_sets=max_sets,
                              group=group, use_graph=use_graph, tiles=tiles)
    out = []
    for (eik, sources), (tt, sgn, vz, sets) in zip(problems, res):
        nd = 3 if isinstance(eik, Eikonal3D) else 2
        cls = TraveltimeGrid3D if nd == 3 else TraveltimeGrid2D
        src = np.asarray(sources, dtype=np.float64)
        srcs = src.reshape(-1, nd)
        grids = [cls(tt[i], eik._gridsize, eik._origin, srcs[i], None, vz[i],
                     sign=sgn[i] if sgn is not None else None, sets_done=int(sets[i]))
                 for i in range(len(srcs))]
        out.append(grids[0] if src.ndim == 1 else grids)
    return out
