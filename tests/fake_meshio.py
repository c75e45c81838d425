"""A stand-in for meshio.Mesh (meshio is not installed here): keeps the arrays the export
functions pass, with meshio's CellBlock view of the cells (type, data, len)."""
import types


class CellBlock:
    def __init__(self, ctype, data):
        self.type = ctype
        self.data = data

    def __len__(self):
        return len(self.data)


class Mesh:
    def __init__(self, points, cells, point_data=None, cell_data=None):
        self.points = points
        self.cells = [CellBlock(t, d) for t, d in cells]
        self.point_data = point_data or {}
        self.cell_data = cell_data or {}


def module():
    m = types.ModuleType("meshio")
    m.Mesh = Mesh
    return m
