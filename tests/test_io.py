"""CPU checks of the meshio export (fteikpy/_io.py): mesh layout against a per-cell
restatement of the reference's construction on small grids, ray meshes, and the errors."""
import sys

import numpy as np
import pytest

import fake_meshio
from fteikpy_amd import _io


def _loop_mesh_2d(nx, ny, dx, dy, x0, y0):
    # the reference's construction (fteikpy/_io.py:156-195): per cell, per vertex
    xs = np.arange(nx + 1) * dx + x0
    ys = np.arange(ny + 1) * dy + y0
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    pts = [[x, y] for x, y in zip(X.ravel("F"), Y.ravel("F"))]
    I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    cells = [[np.ravel_multi_index(v, [nx + 1, ny + 1], order="F")
              for v in ([i, j], [i + 1, j], [i + 1, j + 1], [i, j + 1])]
             for i, j in zip(I.ravel("F"), J.ravel("F"))]
    return np.array(pts, dtype=float), np.array(cells)


def _loop_mesh_3d(nx, ny, nz, dx, dy, dz, x0, y0, z0):
    xs = np.arange(nx + 1) * dx + x0
    ys = np.arange(ny + 1) * dy + y0
    zs = np.arange(nz + 1) * dz + z0
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    pts = [[x, y, z] for x, y, z in zip(X.ravel(), Y.ravel(), Z.ravel())]
    I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    verts = lambda i, j, k: ([i, j, k], [i + 1, j, k], [i + 1, j + 1, k], [i, j + 1, k],
                             [i, j, k + 1], [i + 1, j, k + 1], [i + 1, j + 1, k + 1],
                             [i, j + 1, k + 1])
    cells = [[np.ravel_multi_index(v, [nx + 1, ny + 1, nz + 1]) for v in verts(i, j, k)]
             for i, j, k in zip(I.ravel(), J.ravel(), K.ravel())]
    return np.array(pts, dtype=float), np.array(cells)


def _reference_frame_2d(pts):
    # the reference's 2D steps (fteikpy/_io.py:50-57, :71): append y = 0, swap the last two
    # columns, negate the vertical one
    p = np.column_stack((pts, np.zeros(len(pts))))[:, [0, 2, 1]]
    p[:, 2] *= -1.0
    return p


@pytest.mark.parametrize("n", [(2, 3), (4, 1), (5, 7)])
def test_mesh_2d_matches_per_cell_construction(n):
    nx, nz = n
    zxy, cells = _io._structured_mesh((nz, nx), (2.0, 0.5), (3.0, -1.0))
    rp, rc = _loop_mesh_2d(nx, nz, 0.5, 2.0, -1.0, 3.0)
    assert np.array_equal(_io._elevation_frame(zxy), _reference_frame_2d(rp))
    assert cells[0][0] == "quad" and np.array_equal(cells[0][1], rc)


@pytest.mark.parametrize("n", [(2, 3, 4), (1, 1, 1), (3, 5, 2)])
def test_mesh_3d_matches_per_cell_construction(n):
    nx, ny, nz = n
    zxy, cells = _io._structured_mesh((nz, nx, ny), (1.5, 0.5, 2.0), (7.0, -1.0, 3.0))
    rp, rc = _loop_mesh_3d(nx, ny, nz, 0.5, 2.0, 1.5, -1.0, 3.0, 7.0)
    rp = rp.copy()
    rp[:, 2] *= -1.0
    assert np.array_equal(_io._elevation_frame(zxy), rp) and cells[0][0] == "hexahedron"
    assert np.array_equal(cells[0][1], rc)


def test_ray_mesh(monkeypatch):
    monkeypatch.setitem(sys.modules, "meshio", fake_meshio.module())
    r1 = np.array([[0.0, 1.0, 2.0], [1.0, 1.5, 2.5], [2.0, 2.0, 3.0]])
    r2 = np.array([[5.0, 0.0, 0.0], [4.0, 1.0, 0.5]])
    m = _io.ray_to_meshio(r1, r2)
    assert np.array_equal(m.points[:, 0], np.r_[r1[:, 1], r2[:, 1]])   # x
    assert np.array_equal(m.points[:, 1], np.r_[r1[:, 2], r2[:, 2]])   # y
    assert np.array_equal(m.points[:, 2], -np.r_[r1[:, 0], r2[:, 0]])  # elevation
    assert [c.type for c in m.cells] == ["line", "line"]
    assert np.array_equal(m.cells[0].data, [[0, 1], [1, 2]])
    assert np.array_equal(m.cells[1].data, [[3, 4]])
    m2 = _io.ray_to_meshio(np.array([[1.0, 2.0], [3.0, 4.0]]))
    assert np.array_equal(m2.points, [[2.0, 0.0, -1.0], [4.0, 0.0, -3.0]])
    with pytest.raises(ValueError, match="argument 2 does not seem to be a ray"):
        _io.ray_to_meshio(r1, np.zeros(3))


def test_grid_to_meshio_rejects_non_grids(monkeypatch):
    monkeypatch.setitem(sys.modules, "meshio", fake_meshio.module())
    with pytest.raises(ValueError, match="argument 1 is not a supported grid"):
        _io.grid_to_meshio(np.ones((3, 3)))
