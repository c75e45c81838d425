"""GPU: meshio export (fteikpy/_io.py) of device-solved grids and rays, with a stand-in
for meshio.Mesh (meshio is not installed here).

Unit-velocity models with the source on the centre node.  The exported arrays are checked
against the solved grids themselves (node order, elevation frame, cell connectivity), and
against the values the reference pins for the same models (tests/test_meshio.py), kept
here as data: the traveltime sum, and the sum of the line-cell indices of the honor-grid
ray from the origin corner, which is (m - 1)^2 for a ray of m points."""
import sys

import numpy as np
import pytest

import fake_meshio

pytestmark = pytest.mark.gpu

# cells per axis (z, x[, y]) -> (traveltime sum, points of the corner ray)
PINNED = {
    (8, 10): (378.23469225, 9),
    (8, 10, 12): (6909.90160991, 14),
}


@pytest.fixture
def fa(device, monkeypatch):
    import fteikpy_amd

    monkeypatch.setitem(sys.modules, "meshio", fake_meshio.module())
    return fteikpy_amd


def _export_order(a):
    """(z, x[, y]) grid -> the export's node/cell order."""
    return a.ravel() if a.ndim == 2 else np.transpose(a, (1, 2, 0)).ravel()


@pytest.mark.parametrize("cells", sorted(PINNED, key=len))
def test_unit_model_export(fa, cells):
    nd = len(cells)
    tt_sum, ray_points = PINNED[cells]
    Eik = fa.Eikonal2D if nd == 2 else fa.Eikonal3D
    model = Eik(np.ones(cells), (1.0,) * nd)
    centre = tuple(float(n // 2) for n in cells)
    tt = model.solve(centre, return_gradient=True)
    mesh = fa.grid_to_meshio(model, tt)

    nodes = [n + 1 for n in cells]
    # geometry: every node once, in the export order, in the (x, y, -z) frame
    axes = [np.arange(n, dtype=float) for n in nodes]
    zxy = np.column_stack([_export_order(m) for m in np.meshgrid(*axes, indexing="ij")])
    expect = np.column_stack((zxy[:, 1], zxy[:, 2] if nd == 3 else 0.0 * zxy[:, 0], -zxy[:, 0]))
    assert np.array_equal(mesh.points, expect)
    (block,) = mesh.cells
    assert block.type == ("quad" if nd == 2 else "hexahedron")
    assert block.data.shape == (int(np.prod(cells)), 4 if nd == 2 else 8)
    # each cell's corners span exactly one unit in every axis of the frame
    span = mesh.points[block.data].max(axis=1) - mesh.points[block.data].min(axis=1)
    assert np.array_equal(span[:, [0, 2] if nd == 2 else [0, 1, 2]],
                          np.ones((len(span), nd)))

    # data: the solved arrays in the same order
    times = mesh.point_data["Traveltime"]
    assert np.array_equal(times, _export_order(tt.grid))
    src_node = int(np.argmin(times))
    assert times[src_node] == 0.0 and np.array_equal(mesh.points[src_node][[0, -1]],
                                                      [centre[1], -centre[0]])
    np.testing.assert_allclose(times.sum(), tt_sum, rtol=1e-9)
    grad = mesh.point_data["Gradient"]
    g = [_export_order(c.grid) for c in tt.gradient]
    assert np.array_equal(grad[:, 0], g[1]) and np.array_equal(grad[:, 2], -g[0])
    assert np.all(grad[src_node] == 0.0)
    np.testing.assert_allclose(grad.sum(axis=0), 0.0, atol=1e-8)  # symmetric about the source
    assert np.array_equal(mesh.cell_data["Velocity"][0], np.ones(int(np.prod(cells))))

    ray = tt.raytrace(np.zeros(nd), honor_grid=True)
    assert len(ray) == ray_points
    line = fa.ray_to_meshio(ray)
    assert len(line.points) == ray_points
    segs = line.cells[0].data
    assert np.array_equal(segs, np.column_stack((np.arange(ray_points - 1),
                                                 np.arange(1, ray_points))))
    assert segs.sum() == (ray_points - 1) ** 2


def test_several_grids_are_numbered(fa):
    e = fa.Eikonal3D(np.ones((4, 5, 6)), (1.0, 1.0, 1.0))
    t1 = e.solve((1.0, 1.0, 1.0))
    t2 = e.solve((2.0, 2.0, 2.0), return_gradient=True)
    mesh = fa.grid_to_meshio(t1, t2, e, e)
    assert set(mesh.point_data) == {"Traveltime", "Traveltime 2", "Gradient 2"}
    assert set(mesh.cell_data) == {"Velocity", "Velocity 2"}
    assert len(mesh.points) == 5 * 6 * 7
