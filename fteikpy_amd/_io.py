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


_GRIDS = (Eikonal2D, Eikonal3D, TraveltimeGrid2D, TraveltimeGrid3D)
_NODE_GRIDS = (TraveltimeGrid2D, TraveltimeGrid3D)


def _numbered(label, k):
    """Data array names: the first of a kind keeps the bare label, later ones are numbered
    from 2 ("Velocity", "Velocity 2", ...)."""
    return label if k == 1 else f"{label} {k}"


def _elevation_frame(zxy):
    """Columns in the grids' (z, x[, y]) axis order -> the mesh frame (x, y, -z): depth
    becomes elevation and a 2D grid lies in the y = 0 plane."""
    zxy = np.asarray(zxy, dtype=np.float64)
    y = zxy[:, 2] if zxy.shape[1] == 3 else np.zeros(len(zxy))
    return np.column_stack((zxy[:, 1], y, -zxy[:, 0]))


def grid_to_meshio(*args):
    """Return a meshio.Mesh from Eikonal2D/3D and TraveltimeGrid2D/3D objects
    (fteikpy/_io.py:7-106).  The first argument sets the mesh: its cells are the velocity
    cells (a traveltime grid's nodes are their corners).  Velocity models become cell data,
    traveltimes and their gradient direction point data."""
    import meshio

    for i, arg in enumerate(args):
        if not isinstance(arg, _GRIDS):
            raise ValueError(f"argument {i + 1} is not a supported grid")
    first = args[0]
    ndim = first._ndim
    ncell = [n - 1 for n in first.shape] if isinstance(first, _NODE_GRIDS) else list(first.shape)
    zxy, cells = _structured_mesh(ncell, first.gridsize, first.origin)
    points = _elevation_frame(zxy)

    counts = {}
    point_data, cell_data = {}, {}
    for arg in args:
        kind = "Traveltime" if isinstance(arg, _NODE_GRIDS) else "Velocity"
        counts[kind] = counts.get(kind, 0) + 1
        k = counts[kind]
        if kind == "Velocity":
            cell_data[_numbered(kind, k)] = [_ravel_grid(arg.grid, ndim)]
            continue
        point_data[_numbered(kind, k)] = _ravel_grid(arg.grid, ndim)
        if arg._sign is not None or arg._gradient_dev is not None:
            comps = np.column_stack([_ravel_grid(c.grid, ndim) for c in arg.gradient])
            point_data[_numbered("Gradient", k)] = _elevation_frame(comps)
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


def _structured_mesh(ncell, gridsize, origin):
    """Nodes and cells of a structured grid of `ncell` cells per (z, x[, y]) axis
    (fteikpy/_io.py:156-237).

    Returns the node coordinates as (z, x[, y]) columns and the cell blocks.  Nodes and
    cells are numbered in the export order of the data arrays (`_ravel_grid`): in 2D C
    order over (z, x), in 3D C order over (x, y, z).  A quad lists its corners
    (x, z), (x+1, z), (x+1, z+1), (x, z+1); a hexahedron its lower face
    (x, y, z), (x+1, y, z), (x+1, y+1, z), (x, y+1, z), then the same four at z+1."""
    nd = len(ncell)
    nodes = [n + 1 for n in ncell]
    axes = [o + d * np.arange(n) for o, d, n in zip(origin, gridsize, nodes)]
    mesh = np.meshgrid(*axes, indexing="ij")
    zxy = np.column_stack([_ravel_grid(m, nd) for m in mesh])
    # node number of (z, x[, y]) index arrays, and the corner offsets of a cell
    if nd == 2:
        number = lambda i: i[0] * nodes[1] + i[1]  # noqa: E731
        corners = [(0, 0), (0, 1), (1, 1), (1, 0)]
        ctype = "quad"
    else:
        number = lambda i: (i[1] * nodes[2] + i[2]) * nodes[0] + i[0]  # noqa: E731
        corners = [(0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1),
                   (1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1)]
        ctype = "hexahedron"
    low = [_ravel_grid(m, nd) for m in np.meshgrid(*[np.arange(n) for n in ncell], indexing="ij")]
    conn = np.column_stack([number([a + o for a, o in zip(low, off)]) for off in corners])
    return zxy, [(ctype, conn)]


def _ravel_grid(grid, ndim):
    """fteikpy/_io.py:240-242: (z, x) C order in 2D, (x, y, z) C order in 3D."""
    return grid.ravel() if ndim == 2 else np.transpose(grid, axes=[1, 2, 0]).ravel()
