"""VTK output of a grid and of nodal and cell values, for looking at a solution in ParaView.

The counterpart of the reference's `VTKGridFile` / `write_solution` / `write_node_data` /
`write_cell_data` / `write_nodeset` / `write_cell_colors` (src/Export/VTK.jl:29-78, 129-141,
209-225, 244-275, 276-288, 372-383).  The file is a VTU (XML UnstructuredGrid) in ASCII with
17 significant digits, written by this module instead of WriteVTK, so values round-trip
exactly.  Q1 geometry only: Line, Quadrilateral and Hexahedron, whose node order is VTK's
(VTK.jl:96).  Coordinates and 2-component vectors are padded to three components, as
WriteVTK does (VTK.jl:163-175).
"""
import numpy as np

from .refel import base_interpolation, n_components

VTK_CELL_TYPE = {"Line": 3, "Quadrilateral": 9, "Hexahedron": 12}   # VTK_LINE, VTK_QUAD, VTK_HEXAHEDRON


class VTKGridFile:
    """`VTKGridFile(filename, grid_or_dh)`: collects data, writes `filename` (+ ".vtu") on
    close; usable as a context manager like the reference's do-block form."""

    def __init__(self, filename, grid_or_dh):
        self.dh = grid_or_dh if hasattr(grid_or_dh, "cell_dofs") else None
        self.grid = self.dh.grid if self.dh is not None else grid_or_dh
        if self.grid.celltype not in VTK_CELL_TYPE:
            raise ValueError(f"no VTK cell type for {self.grid.celltype}")
        self.filename = filename if filename.endswith(".vtu") else filename + ".vtu"
        self.point_data = []   # (name, array (nnodes, ncomp))
        self.cell_data = []    # (name, array (ncells, ncomp))
        self.closed = False

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def close(self):
        if self.closed:
            return
        g = self.grid
        pts = np.zeros((g.nnodes, 3))
        pts[:, : g.nodes.shape[1]] = g.nodes
        cells = np.asarray(g.cells, dtype=np.int64) - 1
        npc = cells.shape[1]
        with open(self.filename, "w") as fh:
            w = fh.write
            w('<?xml version="1.0"?>\n')
            w('<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n')
            w(f'<UnstructuredGrid>\n<Piece NumberOfPoints="{g.nnodes}" NumberOfCells="{g.ncells}">\n')
            for tag, data in (("PointData", self.point_data), ("CellData", self.cell_data)):
                w(f"<{tag}>\n")
                for name, a in data:
                    _array(w, a, "Float64", name)
                w(f"</{tag}>\n")
            w("<Points>\n")
            _array(w, pts, "Float64", None)
            w("</Points>\n<Cells>\n")
            _array(w, cells.reshape(-1, 1), "Int64", "connectivity")
            _array(w, (np.arange(1, g.ncells + 1, dtype=np.int64) * npc).reshape(-1, 1), "Int64", "offsets")
            _array(w, np.full((g.ncells, 1), VTK_CELL_TYPE[g.celltype], dtype=np.int64), "UInt8", "types")
            w("</Cells>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")
        self.closed = True


def _array(w, a, vtype, name):
    a = np.asarray(a)
    ncomp = a.shape[1] if a.ndim == 2 else 1
    nm = f' Name="{name}"' if name is not None else ""
    w(f'<DataArray type="{vtype}"{nm} NumberOfComponents="{ncomp}" format="ascii">\n')
    fmt = "%.17g" if vtype == "Float64" else "%d"
    w(" ".join(fmt % v for v in a.ravel()))
    w("\n</DataArray>\n")


def _pad(a):
    """Scalars stay one component; 2-component vectors are padded to 3 (VTK.jl:163-175)."""
    if a.shape[1] == 2:
        return np.hstack([a, np.zeros((a.shape[0], 1))])
    return a


def _at_grid_nodes(dh, u, field):
    """Nodal values of one field: the vertex dofs of every cell (VTK.jl:209-225; Lagrange
    vertex dofs come first, components interleaved per base function)."""
    k = dh.field_names.index(field)
    ip, off = dh.field_ips[k], dh.field_offsets[k]
    vdim = n_components(ip)
    g = dh.grid
    cells = np.asarray(g.cells, dtype=np.int64) - 1
    nv = cells.shape[1]
    if base_interpolation(ip).nbasefunctions < nv:
        raise ValueError("the field has fewer base functions than the cell has nodes")
    data = np.zeros((g.nnodes, vdim))
    u = np.asarray(u, dtype=np.float64)
    for a in range(nv):
        for d in range(vdim):
            data[cells[:, a], d] = u[dh.cell_dofs[:, off + a * vdim + d] - 1]
    return data


def write_solution(vtk, dh, u, suffix=""):
    """Every field of `dh` at the grid nodes, named field + suffix (VTK.jl:209-225)."""
    for name in dh.field_names:
        vtk.point_data.append((name + suffix, _pad(_at_grid_nodes(dh, u, name))))
    return vtk


def write_node_data(vtk, nodedata, name):
    """Values already ordered by the grid's nodes (VTK.jl:261-275)."""
    a = np.asarray(nodedata, dtype=np.float64)
    a = a.reshape(len(a), -1)
    if a.shape[0] != vtk.grid.nnodes:
        raise ValueError("node data length must equal the number of nodes")
    vtk.point_data.append((name, _pad(a)))
    return vtk


def write_cell_data(vtk, celldata, name):
    """One value (or vector) per cell (VTK.jl:244-247)."""
    a = np.asarray(celldata, dtype=np.float64)
    a = a.reshape(len(a), -1)
    if a.shape[0] != vtk.grid.ncells:
        raise ValueError("cell data length must equal the number of cells")
    vtk.cell_data.append((name, a))
    return vtk


def write_nodeset(vtk, grid, nodeset):
    """1 at the nodes of the named node set, 0 elsewhere (VTK.jl:276-288)."""
    z = np.zeros(grid.nnodes)
    z[np.asarray(sorted(grid.nodesets[nodeset]), dtype=np.int64) - 1] = 1.0
    return write_node_data(vtk, z, nodeset)


def write_cell_colors(vtk, grid, cell_colors, name="coloring"):
    """The colour (1-based) of every cell, 0 for cells in no colour (VTK.jl:372-383)."""
    c = np.zeros(grid.ncells)
    for i, cells in enumerate(cell_colors):
        c[np.asarray(cells, dtype=np.int64) - 1] = i + 1
    return write_cell_data(vtk, c, name)
