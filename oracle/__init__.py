"""Oracle: CPU restatement of fteikpy 2.4.0's numeric kernels (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import this
package, and only as the checker.  The product package ``fteikpy_amd`` never imports it.

The numerics live in ``fteik_oracle.c`` (built into ``libfteik_oracle.so`` by the Makefile
beside it); this module is a ctypes front-end whose functions mirror the reference's
numba entry points:

- ``solve3d``  -> fteikpy/_fteik/_fteik3d.py:516-525 (fteik3d / fteik3d_vectorized)
- ``solve2d``  -> fteikpy/_fteik/_fteik2d.py:667-674
- ``interp3d`` / ``interp2d``   -> fteikpy/_interp/_interp3d.py:188-195, _interp2d.py:88-95
- ``vinterp3d`` / ``vinterp2d`` -> fteikpy/_interp/_vinterp3d.py:280-291, _vinterp2d.py:122-131
- ``ray3d`` / ``ray2d``         -> fteikpy/_fteik/_ray3d.py:167-208, _ray2d.py:120-153
- ``adjoint3d``  -> no reference counterpart: adjoint-state gradient of receiver traveltimes
  (vinterp3d of converged fteik3d grids) with respect to slowness, pinned by central finite
  differences of ``solve3d`` + ``vinterp3d``

Extra keyword arguments (not in the reference): ``nsweep <= 0`` runs sets until one
changes no node (``max_sets`` caps it) and the number of sets executed is returned;
``order`` selects the loop nesting (0 = reference k,j,i; 1 = i,j,k for 3D, or the
non-interleaved 4-pass set for 2D).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "libfteik_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile libfteik_oracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            build()
        _lib = ctypes.CDLL(_LIBPATH)
        L = _lib
        i, d = ctypes.c_int, ctypes.c_double
        L.oracle_fteik3d_batch.argtypes = [_dp, i, i, i, d, d, d, _dp, i, i, i, i, i, i, _dp, _dp, _dp, _ip]
        L.oracle_fteik2d_batch.argtypes = [_dp, i, i, d, d, _dp, i, i, i, i, i, i, _dp, _dp, _dp, _ip]
        L.oracle_interp3d.argtypes = [_dp, i, _dp, i, _dp, i, _dp, _dp, i, d, _dp]
        L.oracle_vinterp3d.argtypes = [_dp, i, _dp, i, _dp, i, _dp, _dp, i, _dp, d, d, _dp]
        L.oracle_interp2d.argtypes = [_dp, i, _dp, i, _dp, _dp, i, d, _dp]
        L.oracle_vinterp2d.argtypes = [_dp, i, _dp, i, _dp, _dp, i, _dp, d, d, _dp]
        L.oracle_ray3d.argtypes = [_dp, i, _dp, i, _dp, i, _dp, _dp, _dp, _dp, i, _dp, d, i, i, _dp, _ip, _ip]
        L.oracle_ray2d.argtypes = [_dp, i, _dp, i, _dp, _dp, _dp, i, _dp, d, i, i, _dp, _ip, _ip]
        L.oracle_adjoint3d.argtypes = [_dp, i, i, i, d, d, d, _dp, _dp, i, _dp, _dp, i, _dp, _dp, _ip, _ip]
        L.oracle_adjoint3d_exact.argtypes = [_dp, i, i, i, d, d, d, _dp, _dp, i, i, _dp, i, _dp, _dp, _ip]
        for name in ("oracle_adjoint3d_exact", "oracle_adjoint3d", "oracle_fteik3d_batch", "oracle_fteik2d_batch", "oracle_ray3d", "oracle_ray2d"):
            getattr(L, name).restype = ctypes.c_int
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _pi(a):
    return a.ctypes.data_as(_ip)


def _f8(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def solve3d(slow, dz, dx, dy, src, nsweep=2, grad=False, max_sets=100, order=0, nthreads=0):
    """Restates solve3d; returns (tt, ttgrad, vzero, sets_done)."""
    slow = _f8(slow)
    src = _f8(src)
    single = src.ndim == 1
    src2 = src.reshape(-1, 3)
    n = src2.shape[0]
    nz, nx, ny = slow.shape
    tt = np.empty((n, nz + 1, nx + 1, ny + 1))
    ttgrad = np.empty((n, nz + 1, nx + 1, ny + 1, 3)) if grad else np.empty((n, 0, 0, 0, 0))
    vzero = np.empty(n)
    sets = np.zeros(n, dtype=np.int32)
    st = lib().oracle_fteik3d_batch(
        _p(slow), nz, nx, ny, dz, dx, dy, _p(src2), n, int(nsweep), int(max_sets), int(grad),
        int(order), int(nthreads), _p(tt), _p(ttgrad), _p(vzero), _pi(sets))
    if st == 1:
        raise ValueError("source out of bound")
    if single:
        return tt[0], ttgrad[0], float(vzero[0]), int(sets[0])
    return tt, ttgrad, vzero, sets


def solve2d(slow, dz, dx, src, nsweep=2, grad=False, max_sets=100, order=0, nthreads=0):
    """Restates solve2d; returns (tt, ttgrad, vzero, sets_done)."""
    slow = _f8(slow)
    src = _f8(src)
    single = src.ndim == 1
    src2 = src.reshape(-1, 2)
    n = src2.shape[0]
    nz, nx = slow.shape
    tt = np.empty((n, nz + 1, nx + 1))
    ttgrad = np.empty((n, nz + 1, nx + 1, 2)) if grad else np.empty((n, 0, 0, 0))
    vzero = np.empty(n)
    sets = np.zeros(n, dtype=np.int32)
    st = lib().oracle_fteik2d_batch(
        _p(slow), nz, nx, dz, dx, _p(src2), n, int(nsweep), int(max_sets), int(grad),
        int(order), int(nthreads), _p(tt), _p(ttgrad), _p(vzero), _pi(sets))
    if st == 1:
        raise ValueError("source out of bound")
    if single:
        return tt[0], ttgrad[0], float(vzero[0]), int(sets[0])
    return tt, ttgrad, vzero, sets


def interp3d(z, x, y, v, q, fval=np.nan):
    z, x, y, v, q = map(_f8, (z, x, y, v, q))
    single = q.ndim == 1
    q2 = q.reshape(-1, 3)
    out = np.empty(q2.shape[0])
    lib().oracle_interp3d(_p(z), len(z), _p(x), len(x), _p(y), len(y), _p(v), _p(q2), len(q2), fval, _p(out))
    return float(out[0]) if single else out


def vinterp3d(z, x, y, v, q, src, vzero, fval=np.nan):
    z, x, y, v, q, src = map(_f8, (z, x, y, v, q, src))
    single = q.ndim == 1
    q2 = q.reshape(-1, 3)
    out = np.empty(q2.shape[0])
    lib().oracle_vinterp3d(_p(z), len(z), _p(x), len(x), _p(y), len(y), _p(v), _p(q2), len(q2),
                           _p(src), float(vzero), fval, _p(out))
    return float(out[0]) if single else out


def interp2d(z, x, v, q, fval=np.nan):
    z, x, v, q = map(_f8, (z, x, v, q))
    single = q.ndim == 1
    q2 = q.reshape(-1, 2)
    out = np.empty(q2.shape[0])
    lib().oracle_interp2d(_p(z), len(z), _p(x), len(x), _p(v), _p(q2), len(q2), fval, _p(out))
    return float(out[0]) if single else out


def vinterp2d(z, x, v, q, src, vzero, fval=np.nan):
    z, x, v, q, src = map(_f8, (z, x, v, q, src))
    single = q.ndim == 1
    q2 = q.reshape(-1, 2)
    out = np.empty(q2.shape[0])
    lib().oracle_vinterp2d(_p(z), len(z), _p(x), len(x), _p(v), _p(q2), len(q2), _p(src),
                           float(vzero), fval, _p(out))
    return float(out[0]) if single else out


def _ray_errors(status):
    if (status == 2).any():
        raise ValueError("end point out of bound")
    if (status == 3).any():
        raise RuntimeError("maximum number of steps reached")


def ray3d(z, x, y, gz, gx, gy, p, src, stepsize, max_step, honor_grid=False):
    """Restates ray3d: list of rays (source first) or one ray."""
    z, x, y, gz, gx, gy, p, src = map(_f8, (z, x, y, gz, gx, gy, p, src))
    single = p.ndim == 1
    p2 = p.reshape(-1, 3)
    n = p2.shape[0]
    rays = np.empty((n, max_step, 3))
    counts = np.zeros(n, dtype=np.int32)
    status = np.zeros(n, dtype=np.int32)
    lib().oracle_ray3d(_p(z), len(z), _p(x), len(x), _p(y), len(y), _p(gz), _p(gx), _p(gy),
                       _p(p2), n, _p(src), float(stepsize), int(max_step), int(bool(honor_grid)),
                       _p(rays), _pi(counts), _pi(status))
    _ray_errors(status)
    out = [rays[r, counts[r]::-1].copy() for r in range(n)]
    return out[0] if single else out


def ray2d(z, x, gz, gx, p, src, stepsize, max_step, honor_grid=False):
    z, x, gz, gx, p, src = map(_f8, (z, x, gz, gx, p, src))
    single = p.ndim == 1
    p2 = p.reshape(-1, 2)
    n = p2.shape[0]
    rays = np.empty((n, max_step, 2))
    counts = np.zeros(n, dtype=np.int32)
    status = np.zeros(n, dtype=np.int32)
    lib().oracle_ray2d(_p(z), len(z), _p(x), len(x), _p(gz), _p(gx), _p(p2), n, _p(src),
                       float(stepsize), int(max_step), int(bool(honor_grid)), _p(rays),
                       _pi(counts), _pi(status))
    _ray_errors(status)
    out = [rays[r, counts[r]::-1].copy() for r in range(n)]
    return out[0] if single else out


def adjoint3d(slow, dz, dx, dy, origin, src, tt, rec, dldt):
    """dL/d(slow) given dL/dT_rec for T_rec = vinterp3d of the converged grids ``tt``.

    ``src`` (nsrc, 3) and ``rec`` (nrec, 3) are absolute, ``tt`` is (nsrc, nz+1, nx+1, ny+1),
    ``dldt`` is (nsrc, nrec).  Returns (dlds, orphans, rounds): the number of stale nodes
    (held values no candidate reproduces) that carried adjoint, and the push rounds needed.
    """
    slow = _f8(slow)
    src = _f8(src).reshape(-1, 3)
    rec = _f8(rec).reshape(-1, 3)
    tt = _f8(tt).reshape((len(src),) + tuple(s + 1 for s in slow.shape))
    dldt = _f8(dldt).reshape(len(src), len(rec))
    nz, nx, ny = slow.shape
    dlds = np.empty_like(slow)
    orph = np.zeros(1, dtype=np.int32)
    rounds = np.zeros(1, dtype=np.int32)
    st = lib().oracle_adjoint3d(
        _p(slow), nz, nx, ny, dz, dx, dy, _p(_f8(origin)), _p(src), len(src), _p(tt), _p(rec),
        len(rec), _p(dldt), _p(dlds), _pi(orph), _pi(rounds))
    if st == 1:
        raise ValueError("source out of bound")
    if st == 2:
        raise RuntimeError("adjoint pushes did not settle")
    return dlds, int(orph[0]), int(rounds[0])


def adjoint3d_exact(slow, dz, dx, dy, origin, src, rec, dldt, max_sets=100):
    """Exact discrete adjoint: reverse mode through the converged solve's sweep history.

    Solves internally (converge mode, at most ``max_sets`` sets).  Returns (dlds, sets) with
    sets the longest history in sets.
    """
    slow = _f8(slow)
    src = _f8(src).reshape(-1, 3)
    rec = _f8(rec).reshape(-1, 3)
    dldt = _f8(dldt).reshape(len(src), len(rec))
    nz, nx, ny = slow.shape
    dlds = np.empty_like(slow)
    sets = np.zeros(1, dtype=np.int32)
    st = lib().oracle_adjoint3d_exact(
        _p(slow), nz, nx, ny, dz, dx, dy, _p(_f8(origin)), _p(src), len(src), int(max_sets),
        _p(rec), len(rec), _p(dldt), _p(dlds), _pi(sets))
    if st == 1:
        raise ValueError("source out of bound")
    if st == 2:
        raise RuntimeError("history replay mismatch")
    return dlds, int(sets[0])
