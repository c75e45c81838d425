"""meshio export of grids and raypaths, fteikpy/_io.py:7-153.

Same functions, arguments, mesh layout and errors as the reference: grids become a
structured quad (2D) or hexahedron (3D) mesh with x, y, elevation (= -z) axes, velocity
models as cell data, traveltimes and their gradient direction as point data; rays become
polylines.  meshio is imported when called, as in the reference (an ImportError without
it).  The arrays are built with vectorized index arithmetic instead of the reference's
per-cell loops; they are the same arrays (tests/test_io.py checks the reference's pins).
"""
import numpy as np

from ._grid import TraveltimeGrid2D, TraveltimeGrid3D
from ._solver import Eikonal2D, Eikonal3D


def grid_to_meshio(*args):
    """Return a meshio.Mesh from Eikonal2D/3D and TraveltimeGrid2D/3D objects
    (fteikpy/_io.py:7-106)."""
    import meshio

    for i, arg in enumerate(args):
        if not isinstance(arg, (Eikonal2D, Eikonal3D, TraveltimeGrid2D, TraveltimeGrid3D)):
            raise ValueError(f"argument {i + 1} is not a supported grid")
        if i == 0:
            ndim = arg._ndim
            shape = arg.shape
            gridsize = arg.gridsize
            origin = arg.origin

    node_grid = isinstance(args[0], (TraveltimeGrid2D, TraveltimeGrid3D))
    if ndim == 2:
        nz, nx = shape
        dz, dx = gridsize
        z0, x0 = origin
        if node_grid:
            nz, nx = nz - 1, nx - 1
        points, cells = _mesh_2d(nx, nz, dx, dz, x0, z0)
        points = np.column_stack((points, np.zeros(len(points))))[:, [0, 2, 1]]
    else:
        nz, nx, ny = shape
        dz, dx, dy = gridsize
        z0, x0, y0 = origin
        if node_grid:
            nz, nx, ny = nz - 1, nx - 1, ny - 1
        points, cells = _mesh_3d(nx, ny, nz, dx, dy, dz, x0, y0, z0)
    points[:, 2] *= -1.0  # depth -> elevation

    point_data, cell_data = {}, {}
    vel_count = tt_count = 0
    for arg in args:
        if isinstance(arg, (Eikonal2D, Eikonal3D)):
            vel_count += 1
            name = f"Velocity {vel_count}" if vel_count > 1 else "Velocity"
            cell_data[name] = [_ravel_grid(arg.grid, ndim)]
        else:
            tt_count += 1
            name = f"Traveltime {tt_count}" if tt_count > 1 else "Traveltime"
            point_data[name] = _ravel_grid(arg.grid, ndim)
            if arg._sign is not None or arg._gradient_dev is not None:
                name = f"Gradient {tt_count}" if tt_count > 1 else "Gradient"
                g = np.column_stack([_ravel_grid(c.grid, ndim) for c in arg.gradient])
                if ndim == 2:
                    g = np.column_stack((g, np.zeros(len(points))))
                g = g[:, [1, 2, 0]]
                g[:, 2] *= -1.0
                point_data[name] = g
    return meshio.Mesh(points, cells, point_data, cell_data)


def ray_to_meshio(*args):
    """Return a meshio.Mesh of polylines from raypaths (fteikpy/_io.py:109-153)."""
    import meshio

    for i, arg in enumerate(args):
        if np.ndim(arg) != 2:
            raise ValueError(f"argument {i + 1} does not seem to be a ray")
        if i == 0:
            ndim = np.shape(arg)[1]
    cells, parts, n = [], [], 0
    for ray in args:
        ray = np.asarray(ray, dtype=np.float64)
        k = np.arange(len(ray)) + n
        cells.append(("line", np.column_stack((k[:-1], k[1:]))))
        parts.append(ray)
        n += len(ray)
    points = np.concatenate(parts)
    if ndim == 2:
        points = np.column_stack((points, np.zeros(len(points))))
    points = points[:, [1, 2, 0]]
    points[:, 2] *= -1.0
    return meshio.Mesh(points, cells)


def _mesh_2d(nx, ny, dx, dy, x0, y0):
    """Structured quads (fteikpy/_io.py:156-195): points in Fortran order over (x, y),
    vertices (i, j) (i+1, j) (i+1, j+1) (i, j+1) of each cell, cells in Fortran order."""
    xs = np.arange(nx + 1) * dx + x0
    ys = np.arange(ny + 1) * dy + y0
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    points = np.column_stack((X.ravel("F"), Y.ravel("F"))).astype(float)
    I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    i, j = I.ravel("F"), J.ravel("F")

    def idx(a, b):
        return a + b * (nx + 1)

    cells = np.column_stack((idx(i, j), idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)))
    return points, [("quad", cells)]


def _mesh_3d(nx, ny, nz, dx, dy, dz, x0, y0, z0):
    """Structured hexahedra (fteikpy/_io.py:198-237): points in C order over (x, y, z),
    vertices of the lower face then the upper face of each cell, cells in C order."""
    xs = np.arange(nx + 1) * dx + x0
    ys = np.arange(ny + 1) * dy + y0
    zs = np.arange(nz + 1) * dz + z0
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    points = np.column_stack((X.ravel(), Y.ravel(), Z.ravel())).astype(float)
    I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = I.ravel(), J.ravel(), K.ravel()

    def idx(a, b, c):
        return (a * (ny + 1) + b) * (nz + 1) + c

    cells = np.column_stack((idx(i, j, k), idx(i + 1, j, k), idx(i + 1, j + 1, k),
                             idx(i, j + 1, k), idx(i, j, k + 1), idx(i + 1, j, k + 1),
                             idx(i + 1, j + 1, k + 1), idx(i, j + 1, k + 1)))
    return points, [("hexahedron", cells)]


def _ravel_grid(grid, ndim):
    """fteikpy/_io.py:240-242: (z, x) C order in 2D, (x, y, z) C order in 3D."""
    return grid.ravel() if ndim == 2 else np.transpose(grid, axes=[1, 2, 0]).ravel()
