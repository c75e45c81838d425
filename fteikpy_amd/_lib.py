"""ctypes binding of libfteik_amd.so (the C ABI declared in include/fteik_amd.h).

The library is built in-tree by ``fteikpy_amd.build()`` (or ``make -C fteikpy_amd/csrc``).
There is no CPU fallback: if the library or a HIP device is missing, every solver call
raises.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libfteik_amd.so")

_lib = None

_c_int = ctypes.c_int
_c_dbl = ctypes.c_double
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)

# C return codes, include/fteik_amd.h
OK = 0
ERR_SOURCE_OOB = 1
ERR_ENDPOINT_OOB = 2
ERR_MAX_STEP = 3
ERR_ARG = 4
ERR_WORKSPACE = 5
ERR_TIMEOUT = 6
ERR_UNSUPPORTED = 7
ERR_UNSETTLED = 8


class Problem(ctypes.Structure):
    """ctypes mirror of ``fteik_problem`` (include/fteik_amd.h): one model of a batch."""

    _fields_ = [("ndim", ctypes.c_int), ("nz", ctypes.c_int), ("nx", ctypes.c_int),
                ("ny", ctypes.c_int), ("dz", ctypes.c_double), ("dx", ctypes.c_double),
                ("dy", ctypes.c_double), ("slow", ctypes.c_void_p),
                ("src", ctypes.POINTER(ctypes.c_double)), ("nsrc", ctypes.c_int),
                ("tile_a", ctypes.c_int), ("tt", ctypes.c_void_p), ("sgn", ctypes.c_void_p),
                ("vzero", ctypes.c_void_p), ("sets_done", ctypes.c_void_p)]


_pp = ctypes.POINTER(Problem)

# name -> (restype, argtypes); every entry point of include/fteik_amd.h
SIGNATURES = {
    "fteik_version": (ctypes.c_char_p, []),
    "fteik3d_workspace_size": (ctypes.c_size_t, [_c_int, _c_int, _c_int, _c_int]),
    "fteik3d_solve": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_dbl, _c_dbl, _c_dbl, _dp, _c_int,
                               _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp,
                               _vp, ctypes.c_size_t, _c_int, _vp]),
    "fteik3d_scratch_capacity": (ctypes.c_size_t, []),
    "fteik3d_sweep_launches": (_c_int, [_c_int] * 6),
    "fteik3d_graph_cache": (_c_int, [_c_int, ctypes.POINTER(ctypes.c_longlong),
                                     ctypes.POINTER(ctypes.c_longlong)]),
    "fteik3d_solve_noalloc": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_dbl, _c_dbl, _c_dbl, _dp,
                                       _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp,
                                       _vp, _vp, _vp]),
    "fteik3d_gradient": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_dbl, _c_dbl, _c_dbl, _dp,
                                  _c_dbl, _vp, _vp]),
    "fteik3d_history_size": (ctypes.c_size_t, [ctypes.c_longlong, _c_int]),
    "fteik3d_solve_logged": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_dbl, _c_dbl, _c_dbl, _dp,
                                      _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp,
                                      ctypes.c_size_t, _vp, ctypes.c_size_t,
                                      ctypes.POINTER(ctypes.c_longlong), _c_int, _vp]),
    "fteik3d_adjoint_history_workspace_size": (ctypes.c_size_t, [_c_int, _c_int, _c_int, _c_int,
                                                                 ctypes.c_longlong]),
    "fteik3d_adjoint_history": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_dbl, _c_dbl, _c_dbl, _dp,
                                         _dp, _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp,
                                         ctypes.c_size_t, _vp]),
    "fteik3d_adjoint_workspace_size": (ctypes.c_size_t, [_c_int, _c_int, _c_int, _c_int]),
    "fteik3d_adjoint": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_dbl, _c_dbl, _c_dbl, _dp, _dp,
                                 _c_int, _vp, _vp, _c_int, _vp, _vp, ctypes.POINTER(_c_int),
                                 ctypes.POINTER(_c_int), _vp, ctypes.c_size_t, _vp]),
    "fteik_batch_workspace_size": (ctypes.c_size_t, [_pp, _c_int]),
    "fteik_solve_batch": (_c_int, [_pp, _c_int, _c_int, _c_int, _c_int, _c_int, _vp,
                                   ctypes.c_size_t, _c_int, ctypes.POINTER(_c_int), _vp]),
    "fteik_correlate1d": (_c_int, [_vp, _vp, _c_int, ctypes.POINTER(_c_int), _c_int, _vp, _c_int,
                                   _vp]),
    "fteik_resample_workspace_size": (ctypes.c_size_t, [_c_int, ctypes.POINTER(_c_int)]),
    "fteik_resample": (_c_int, [_vp, _c_int, ctypes.POINTER(_c_int), _dp, _dp,
                                ctypes.POINTER(_c_int), _c_int, _vp, _vp, ctypes.c_size_t, _vp]),
    "fteik2d_workspace_size": (ctypes.c_size_t, [_c_int]),
    "fteik2d_solve": (_c_int, [_vp, _c_int, _c_int, _c_dbl, _c_dbl, _dp, _c_int, _c_int, _c_int,
                               _c_int, _vp, _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp]),
    "fteik2d_gradient": (_c_int, [_vp, _vp, _c_int, _c_int, _c_dbl, _c_dbl, _dp, _c_dbl, _vp,
                                  _vp]),
    "fteik3d_interp": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _dp, _dp, _vp, _c_int,
                                _c_dbl, _vp, _vp]),
    "fteik3d_vinterp": (_c_int, [_vp, _c_int, _c_int, _c_int, _dp, _dp, _vp, _c_int, _dp, _c_dbl,
                                 _c_dbl, _vp, _vp]),
    "fteik2d_interp": (_c_int, [_vp, _c_int, _c_int, _c_int, _dp, _dp, _vp, _c_int, _c_dbl, _vp,
                                _vp]),
    "fteik2d_vinterp": (_c_int, [_vp, _c_int, _c_int, _dp, _dp, _vp, _c_int, _dp, _c_dbl, _c_dbl,
                                 _vp, _vp]),
    "fteik3d_raytrace": (_c_int, [_vp, _c_int, _c_int, _c_int, _dp, _dp, _vp, _c_int, _dp, _c_dbl,
                                  _c_int, _c_int, _vp, _vp, _vp, _vp]),
    "fteik3d_raytrace_batch": (_c_int, [_vp, _vp, _vp, _c_int, _c_int, _c_int, _c_int, _dp, _dp,
                                        _vp, _c_int, _c_dbl, _c_int, _c_int, _vp, _vp, _vp, _vp]),
    "fteik2d_raytrace": (_c_int, [_vp, _c_int, _c_int, _dp, _dp, _vp, _c_int, _dp, _c_dbl, _c_int,
                                  _c_int, _vp, _vp, _vp, _vp]),
}


def build():
    """Compile libfteik_amd.so for gfx950 in-tree."""
    subprocess.run(["make", "-s", "-j8", "-C", os.path.join(_HERE, "csrc")], check=True)


def lib():
    """The loaded library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(
                f"{LIBPATH} is missing: build it with fteikpy_amd.build() "
                "(hipcc, gfx950); there is no CPU fallback")
        L = ctypes.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    """Map a C return code to the reference's exceptions."""
    if status == OK:
        return
    if status == ERR_SOURCE_OOB:
        raise ValueError("source out of bound")
    if status == ERR_ENDPOINT_OOB:
        raise ValueError("end point out of bound")
    if status == ERR_MAX_STEP:
        raise RuntimeError("maximum number of steps reached")
    if status == ERR_ARG:
        raise ValueError("invalid argument")
    if status == ERR_WORKSPACE:
        raise RuntimeError("workspace too small")
    if status == ERR_UNSUPPORTED:
        raise NotImplementedError("option not supported by this build")
    if status == ERR_TIMEOUT:
        raise RuntimeError("persistent sweep: a grid-barrier wait gave up (result invalid)")
    if status == ERR_UNSETTLED:
        raise RuntimeError("adjoint: the upwind stencils do not reproduce the solved times "
                           "(cycle or inconsistent history)")
    raise RuntimeError(f"HIP error {-status}")


def host_f8(a):
    """Contiguous float64 numpy array and a ctypes double* to it (keep the array alive)."""
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    return arr, arr.ctypes.data_as(_dp)
