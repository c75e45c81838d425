"""Differentiable receiver traveltimes (no counterpart in the reference).

``receiver_traveltimes(slowness, gridsize, sources, receivers)`` returns the traveltimes
T[s, r] that ``Eikonal3D(1 / slowness, gridsize, origin).solve(sources, nsweep=None)[s]
(receivers[r])`` gives (fteikpy/_solver.py:120-126 then fteikpy/_grid.py:193-219: the
converged 3D sweep and the apparent-velocity interpolation), as a torch tensor whose
gradient with respect to ``slowness`` comes from hand-written adjoint kernels:

- ``mode="exact"`` (default): the solve records its sweep history (every node change and
  the time it replaced); the backward reverses the passes, recomputing the candidate that
  produced each change (``fteik3d_solve_logged`` + ``fteik3d_adjoint_history``).  This is
  the derivative of the sweeps themselves, what finite differences of the solver measure.
- ``mode="stencil"``: the backward recomputes one upwind stencil per node from the
  converged times and solves the adjoint system in reverse dependency order
  (``fteik3d_adjoint``).  No history is kept; it matches the exact gradient when every
  node's time is reproduced by a candidate on the final grid, and reports the nodes where
  it is not ("stale").
"""
import ctypes

import numpy as np

from . import _lib, autotune
from ._grid import _ptr, _torch, stream_ptr

# changes per node and source a converged solve makes (measured: 2.1 - 2.8); the history
# is sized for this and the solve is repeated with the exact count if it was not enough
_EVENTS_PER_NODE = 3.0


def _nodes(shape):
    nz, nx, ny = shape
    return (nz + 1) * (nx + 1) * (ny + 1)


def _chunks(n, nodes):
    """Source ranges whose (source, node) indices fit in 32 bits."""
    step = max(1, min(n, (2 ** 32 - 1) // nodes))
    return [(a, min(n, a + step)) for a in range(0, n, step)]


def _solve(slow, gridsize, rel_src, max_sets, logged):
    """Converged solve of all sources; with `logged`, also the sweep history."""
    torch = _torch()
    nz, nx, ny = slow.shape
    n = len(rel_src)
    rs, rp = _lib.host_f8(rel_src)
    dev = slow.device
    tt = torch.empty((n, nz + 1, nx + 1, ny + 1), dtype=torch.float64, device=dev)
    vzero = torch.empty(n, dtype=torch.float64, device=dev)
    sets = torch.empty(n, dtype=torch.int32, device=dev)
    L = _lib.lib()
    wsb = L.fteik3d_workspace_size(nz, nx, ny, n)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    tile_a, group, sched = autotune.choice3d((nz, nx, ny), n)
    if not logged:
        # the library's fastest plain solve: the recorded (or the overlapped) schedule, sets
        # replayed from a captured graph (launched directly on the legacy default stream)
        code = {"overlapped": 2, "diagonal": 1}[sched or "overlapped"]
        st = L.fteik3d_solve(_ptr(slow), nz, nx, ny, *gridsize, rp, n, 0, int(max_sets), 0,
                             tile_a, group, code, _ptr(tt), None, _ptr(vzero), _ptr(sets),
                             _ptr(ws), wsb, 1, stream_ptr())
        _lib.check(st)
        return tt, vzero.cpu().numpy(), sets.cpu().numpy(), None, 0
    max_pass = 8 * int(max_sets)
    cap = int(_EVENTS_PER_NODE * _nodes(slow.shape) * n) + 1024
    for _ in range(2):
        hb = L.fteik3d_history_size(cap, max_pass)
        hist = torch.empty(hb, dtype=torch.uint8, device=dev)
        ev = ctypes.c_longlong(0)
        st = L.fteik3d_solve_logged(_ptr(slow), nz, nx, ny, *gridsize, rp, n, 0, int(max_sets),
                                    tile_a, group, _ptr(tt), _ptr(vzero), _ptr(sets), _ptr(ws), wsb,
                                    _ptr(hist), hb, ctypes.byref(ev), 0, stream_ptr())
        if st == _lib.ERR_WORKSPACE and ev.value > cap:
            cap = ev.value + 1024
            del hist
            continue
        _lib.check(st)
        return tt, vzero.cpu().numpy(), sets.cpu().numpy(), hist, ev.value
    raise RuntimeError("sweep history did not fit")


def _interp(tt, gridsize, origin, src, vzero, rec_dev):
    torch = _torch()
    n = tt.shape[0]
    nrec = rec_dev.shape[0]
    out = torch.empty((n, nrec), dtype=torch.float64, device=tt.device)
    o, op = _lib.host_f8(origin)
    d, dp = _lib.host_f8(gridsize)
    L = _lib.lib()
    for s in range(n):
        sv, sp = _lib.host_f8(src[s])
        st = L.fteik3d_vinterp(_ptr(tt[s]), *tt.shape[1:], op, dp, _ptr(rec_dev), nrec, sp,
                               float(vzero[s]), float("nan"), _ptr(out[s]), stream_ptr())
        _lib.check(st)
    return out


def _adjoint_stencil(slow, gridsize, origin, src, tt, rec_dev, dldt):
    """Converged-stencil adjoint; returns (dlds, stale nodes, rounds)."""
    torch = _torch()
    nz, nx, ny = slow.shape
    L = _lib.lib()
    o, op = _lib.host_f8(origin)
    total = torch.zeros_like(slow)
    part = torch.empty_like(slow)
    stale_all, rounds_all = 0, 0
    chunks = _chunks(tt.shape[0], _nodes(slow.shape))
    wsb = L.fteik3d_adjoint_workspace_size(nz, nx, ny, chunks[0][1] - chunks[0][0])
    ws = torch.empty(wsb, dtype=torch.uint8, device=slow.device)
    for a, b in chunks:
        s, sp = _lib.host_f8(src[a:b])
        stale, rounds = ctypes.c_int(0), ctypes.c_int(0)
        g = dldt[a:b].contiguous()
        st = L.fteik3d_adjoint(_ptr(slow), nz, nx, ny, *gridsize, op, sp, b - a,
                               _ptr(tt[a:b]), _ptr(rec_dev), rec_dev.shape[0], _ptr(g),
                               _ptr(part), ctypes.byref(stale), ctypes.byref(rounds), _ptr(ws),
                               wsb, stream_ptr())
        _lib.check(st)
        total += part
        stale_all += stale.value
        rounds_all = max(rounds_all, rounds.value)
    return total, stale_all, rounds_all


def _adjoint_history(slow, gridsize, origin, src, tt, hist, events, rec_dev, dldt):
    """Exact adjoint through the logged history; rolls tt back in place."""
    torch = _torch()
    nz, nx, ny = slow.shape
    L = _lib.lib()
    o, op = _lib.host_f8(origin)
    s, sp = _lib.host_f8(src)
    dlds = torch.empty_like(slow)
    wsb = L.fteik3d_adjoint_history_workspace_size(nz, nx, ny, len(src), int(events))
    ws = torch.empty(wsb, dtype=torch.uint8, device=slow.device)
    g = dldt.contiguous()
    st = L.fteik3d_adjoint_history(_ptr(slow), nz, nx, ny, *gridsize, op, sp, len(src), _ptr(tt),
                                   _ptr(hist), _ptr(rec_dev), rec_dev.shape[0], _ptr(g),
                                   _ptr(dlds), _ptr(ws), wsb, stream_ptr())
    _lib.check(st)
    return dlds


def _function():
    torch = _torch()

    class ReceiverTraveltimes(torch.autograd.Function):
        """Forward: converged solve + receiver interpolation; backward: adjoint kernels."""

        @staticmethod
        def forward(ctx, slow, gridsize, origin, sources, receivers, max_sets, mode, stats):
            slow_c = slow.detach().contiguous()
            rec_dev = torch.as_tensor(receivers, device=slow.device)
            exact = mode == "exact"
            tts, hists, outs, sets_all = [], [], [], []
            for a, b in (_chunks(len(sources), _nodes(slow.shape)) if exact else [(0, len(sources))]):
                tt, vzero, sets, hist, events = _solve(slow_c, gridsize, sources[a:b] - origin,
                                                       max_sets, exact)
                outs.append(_interp(tt, gridsize, origin, sources[a:b], vzero, rec_dev))
                tts.append(tt)
                hists.append((a, b, hist, events))
                sets_all.append(sets)
            ctx.save_for_backward(slow_c, rec_dev, *tts)
            ctx.geom = (gridsize, origin, sources, mode, stats, hists)
            ctx.consumed = False
            if stats is not None:
                stats["sets"] = np.concatenate(sets_all)
                stats["events"] = sum(h[3] for h in hists)
            return torch.cat(outs) if len(outs) > 1 else outs[0]

        @staticmethod
        def backward(ctx, grad_out):
            slow_c, rec_dev, *tts = ctx.saved_tensors
            gridsize, origin, sources, mode, stats, hists = ctx.geom
            dldt = grad_out.to(torch.float64).contiguous()
            if mode == "stencil":
                dlds, stale, rounds = _adjoint_stencil(slow_c, gridsize, origin, sources, tts[0],
                                                       rec_dev, dldt)
                if stats is not None:
                    stats["stale"] = stale
                    stats["rounds"] = rounds
                return dlds, None, None, None, None, None, None, None
            if ctx.consumed:
                raise RuntimeError("the exact adjoint consumes its sweep history: backward runs "
                                   "once per forward")
            ctx.consumed = True
            dlds = None
            for tt, (a, b, hist, events) in zip(tts, hists):
                part = _adjoint_history(slow_c, gridsize, origin, sources[a:b], tt, hist, events,
                                        rec_dev, dldt[a:b])
                dlds = part if dlds is None else dlds + part
            return dlds, None, None, None, None, None, None, None

    return ReceiverTraveltimes


_FN = None


def receiver_traveltimes(slowness, gridsize, sources, receivers, origin=None, *, max_sets=200,
                         mode="exact", stats=None):
    """Traveltimes [nsrc, nrec] at `receivers` of the converged 3D solve, differentiable
    with respect to `slowness`.

    slowness  torch float64 tensor [nz, nx, ny] on the HIP device (may require grad)
    gridsize  (dz, dx, dy); origin: (z0, x0, y0), default zeros
    sources   [nsrc, 3] and receivers [nrec, 3]: absolute coordinates (host arrays)
    max_sets  cap on the sweep sets of the converged solve
    mode      "exact" (reverse mode through the sweep history) or "stencil" (converged
              upwind stencil, no history kept)
    stats     optional dict, filled with "sets" (per source) and "events" (logged changes)
              by the forward, and in "stencil" mode with "stale" and "rounds" by the backward
    Receivers outside the grid give NaN and contribute no gradient.
    """
    global _FN
    torch = _torch()
    if _FN is None:
        _FN = _function()
    if mode not in ("exact", "stencil"):
        raise ValueError("mode must be 'exact' or 'stencil'")
    if not isinstance(slowness, torch.Tensor) or slowness.dtype != torch.float64:
        raise TypeError("slowness must be a float64 torch tensor")
    if slowness.device.type != "cuda":
        raise ValueError("slowness must live on the HIP device (no CPU path)")
    if slowness.ndim != 3:
        raise ValueError("slowness must be 3D [nz, nx, ny]")
    gridsize = tuple(float(g) for g in gridsize)
    origin = np.zeros(3) if origin is None else np.asarray(origin, dtype=np.float64)
    src = np.ascontiguousarray(np.asarray(sources, dtype=np.float64).reshape(-1, 3))
    rec = np.ascontiguousarray(np.asarray(receivers, dtype=np.float64).reshape(-1, 3))
    return _FN.apply(slowness, gridsize, origin, src, rec, int(max_sets), mode, stats)
