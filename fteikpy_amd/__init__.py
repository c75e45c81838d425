"""fteikpy_amd: MI355X-native fast-sweeping eikonal solver with fteikpy's interface.

Drop-in names of fteikpy/__init__.py:1-19.  Solves, interpolation, gradients and ray
tracing run as hand-written HIP kernels (gfx950) through the C ABI of
include/fteik_amd.h; arrays live in device memory as torch tensors.
"""
from . import autotune
from ._adjoint import receiver_traveltimes
from ._batch import solve_batch, solve_batch_tensors
from ._grid import Grid2D, Grid3D, TraveltimeGrid2D, TraveltimeGrid3D, raytrace_batch
from ._helpers import get_num_threads, set_num_threads
from ._io import grid_to_meshio, ray_to_meshio
from ._lib import build as _build_lib
from ._solver import Eikonal2D, Eikonal3D

__version__ = "0.1.0"


def build():
    """Compile the HIP library (libfteik_amd.so, gfx950) in-tree."""
    _build_lib()


__all__ = [
    "Eikonal2D",
    "Eikonal3D",
    "Grid2D",
    "Grid3D",
    "TraveltimeGrid2D",
    "TraveltimeGrid3D",
    "get_num_threads",
    "set_num_threads",
    "grid_to_meshio",
    "ray_to_meshio",
    "receiver_traveltimes",
    "solve_batch",
    "solve_batch_tensors",
    "raytrace_batch",
    "autotune",
    "build",
    "__version__",
]
