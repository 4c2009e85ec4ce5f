"""VTK output (the reference's src/Export/VTK.jl) read back with an XML parser.

The heat-equation solution from the oracle's direct solve is written with `write_solution`:
the point data must be u at every grid node, so its norm is the reference's KAT
norm(u) = 3.307743912641305 (heat_equation.jl), and the connectivity must be the grid's cells.
A Q1² field checks the component interleave and the padding of 2-vectors to three components.
"""
import xml.etree.ElementTree as ET

import numpy as np

import ferrite_amd as fa
from oracle.problems import heat_equation


def _read(path):
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    arrays = {}
    for sec in ("PointData", "CellData", "Points", "Cells"):
        for da in piece.find(sec):
            name = da.get("Name") or sec
            vals = np.array(da.text.split(), dtype=np.float64)
            arrays[(sec, name)] = vals.reshape(-1, int(da.get("NumberOfComponents")))
    return piece, arrays


def test_heat_equation_solution_kat(tmp_path):
    u, _ = heat_equation(20)
    g = fa.generate_grid("Quadrilateral", (20, 20))
    ip = fa.Lagrange("RefQuadrilateral", 1)
    dh = fa.DofHandler(g).add("u", ip).close()
    colors = fa.create_coloring(g)
    with fa.VTKGridFile(str(tmp_path / "heat"), dh) as vtk:
        fa.write_solution(vtk, dh, u)
        fa.write_cell_colors(vtk, g, colors)
    piece, a = _read(tmp_path / "heat.vtu")
    assert int(piece.get("NumberOfPoints")) == g.nnodes and int(piece.get("NumberOfCells")) == g.ncells
    un = a[("PointData", "u")][:, 0]
    assert abs(np.linalg.norm(un) - 3.307743912641305) < 1e-12
    cells = np.asarray(g.cells) - 1
    for c in (0, 57, g.ncells - 1):   # the node value is u at that node's dof
        for k in range(4):
            assert un[cells[c, k]] == u[dh.cell_dofs[c, k] - 1]
    assert np.array_equal(a[("Cells", "connectivity")][:, 0].astype(np.int64), cells.ravel())
    assert np.all(a[("Cells", "types")] == 9)
    assert np.array_equal(a[("Cells", "offsets")][:, 0], 4 * np.arange(1, g.ncells + 1))
    pts = a[("Points", "Points")]
    assert np.array_equal(pts[:, :2], g.nodes) and np.all(pts[:, 2] == 0.0)
    col = a[("CellData", "coloring")][:, 0]
    for i, cs in enumerate(colors):
        assert np.all(col[np.asarray(cs) - 1] == i + 1)


def test_vector_field_components_and_padding(tmp_path):
    g = fa.generate_grid("Quadrilateral", (5, 3))
    ip = fa.Lagrange("RefQuadrilateral", 1) ** 2
    dh = fa.DofHandler(g).add("d", ip).close()
    u = np.random.default_rng(3).normal(size=dh.ndofs)
    with fa.VTKGridFile(str(tmp_path / "vec.vtu"), dh) as vtk:
        fa.write_solution(vtk, dh, u, "_x")
        fa.write_cell_data(vtk, np.arange(g.ncells), "id")
    _, a = _read(tmp_path / "vec.vtu")
    d = a[("PointData", "d_x")]
    assert d.shape == (g.nnodes, 3) and np.all(d[:, 2] == 0.0)
    cells = np.asarray(g.cells) - 1
    for c in range(g.ncells):
        for k in range(4):
            for comp in range(2):
                assert d[cells[c, k], comp] == u[dh.cell_dofs[c, 2 * k + comp] - 1]
    assert np.array_equal(a[("CellData", "id")][:, 0], np.arange(g.ncells))


def test_hexahedron_grid(tmp_path):
    g = fa.generate_grid("Hexahedron", (3, 2, 2))
    ip = fa.Lagrange("RefHexahedron", 1)
    dh = fa.DofHandler(g).add("u", ip).close()
    u = np.arange(dh.ndofs, dtype=np.float64)
    with fa.VTKGridFile(str(tmp_path / "hex"), g) as vtk:
        fa.write_node_data(vtk, u[np.argsort(_node_of_dof(g, dh))], "by_node")
    _, a = _read(tmp_path / "hex.vtu")
    assert np.all(a[("Cells", "types")] == 12)
    assert np.array_equal(a[("Cells", "connectivity")][:, 0].astype(np.int64), (np.asarray(g.cells) - 1).ravel())


def _node_of_dof(g, dh):
    nod = np.zeros(dh.ndofs, dtype=np.int64)
    nod[dh.cell_dofs.ravel() - 1] = (np.asarray(g.cells) - 1).ravel()
    return nod
