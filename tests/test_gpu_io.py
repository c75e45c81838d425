"""GPU: the reference's meshio tests (tests/test_meshio.py) on device-solved grids and rays,
with a stand-in for meshio.Mesh (meshio is not installed here)."""
import sys

import numpy as np
import pytest

import fake_meshio

pytestmark = pytest.mark.gpu


@pytest.fixture
def fa(device, monkeypatch):
    import fteikpy_amd

    monkeypatch.setitem(sys.modules, "meshio", fake_meshio.module())
    return fteikpy_amd


def test_meshio_2d(fa):
    nz, nx = 8, 10
    eik = fa.Eikonal2D(np.ones((nz, nx)), (1.0, 1.0))
    tt = eik.solve((float(nz // 2), float(nx // 2)), return_gradient=True)
    ray = tt.raytrace((0.0, 0.0), honor_grid=True)
    mesh = fa.grid_to_meshio(eik, tt)
    mesh_ray = fa.ray_to_meshio(ray)
    npts = (nz + 1) * (nx + 1)
    assert len(mesh.points) == npts
    assert sum(len(cell) for cell in mesh.cells) == nz * nx
    assert mesh.point_data["Traveltime"][npts // 2] == 0.0
    assert np.allclose(mesh.point_data["Traveltime"].sum(), 378.23469225)
    for grad in mesh.point_data["Gradient"].T:
        assert grad[npts // 2] == 0.0
        assert np.allclose(grad.sum(), 0.0)
    assert mesh.cell_data["Velocity"][0].sum() == nz * nx
    assert len(mesh_ray.points) == len(ray)
    assert len(mesh_ray.cells[0].data) == len(ray) - 1
    assert mesh_ray.cells[0].data.sum() == 64.0


def test_meshio_3d(fa):
    nz, nx, ny = 8, 10, 12
    eik = fa.Eikonal3D(np.ones((nz, nx, ny)), (1.0, 1.0, 1.0))
    tt = eik.solve((float(nz // 2), float(nx // 2), float(ny // 2)), return_gradient=True)
    ray = tt.raytrace((0.0, 0.0, 0.0), honor_grid=True)
    mesh = fa.grid_to_meshio(eik, tt)
    mesh_ray = fa.ray_to_meshio(ray)
    npts = (nz + 1) * (nx + 1) * (ny + 1)
    assert len(mesh.points) == npts
    assert sum(len(cell) for cell in mesh.cells) == nz * nx * ny
    assert mesh.point_data["Traveltime"][npts // 2] == 0.0
    assert np.allclose(mesh.point_data["Traveltime"].sum(), 6909.90160991)
    for grad in mesh.point_data["Gradient"].T:
        assert grad[npts // 2] == 0.0
        assert np.allclose(grad.sum(), 0.0)
    assert mesh.cell_data["Velocity"][0].sum() == nz * nx * ny
    assert len(mesh_ray.points) == len(ray)
    assert len(mesh_ray.cells[0].data) == len(ray) - 1
    assert mesh_ray.cells[0].data.sum() == 169.0


def test_several_grids_are_numbered(fa):
    e = fa.Eikonal3D(np.ones((4, 5, 6)), (1.0, 1.0, 1.0))
    t1 = e.solve((1.0, 1.0, 1.0))
    t2 = e.solve((2.0, 2.0, 2.0), return_gradient=True)
    mesh = fa.grid_to_meshio(t1, t2, e, e)
    assert set(mesh.point_data) == {"Traveltime", "Traveltime 2", "Gradient 2"}
    assert set(mesh.cell_data) == {"Velocity", "Velocity 2"}
    assert len(mesh.points) == 5 * 6 * 7
