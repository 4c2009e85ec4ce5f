"""Oracle: Grid and generate_grid (TEST INFRASTRUCTURE ONLY, see oracle/__init__.py).

Restates src/Grid/grid_generators.jl:78-113 (Quadrilateral), :159-204
(Hexahedron), :542-578 (_calculate_coordinate, _generate_nodes,
_extrema_to_corners) as literal loops.  Node ids follow Iterators.product
(first index fastest); cell ids run i fastest; facet sets are sorted by cell.
All indices here are 1-based like the reference.
"""
import numpy as np

from .refel import Lagrange, _product_first_fastest


class Grid:
    def __init__(self, celltype, cells, nodes, facetsets=None, nodesets=None):
        self.celltype = celltype          # "Quadrilateral" | "Hexahedron" | "Line"
        self.cells = cells                # list of tuples of 1-based node ids
        self.nodes = nodes                # (nnodes, sdim) float64
        self.facetsets = facetsets or {}  # name -> list of (cell, facet), 1-based
        self.nodesets = nodesets or {}

    @property
    def refshape(self):
        return {"Line": "RefLine", "Quadrilateral": "RefQuadrilateral",
                "Hexahedron": "RefHexahedron"}[self.celltype]

    @property
    def ncells(self):
        return len(self.cells)

    @property
    def nnodes(self):
        return self.nodes.shape[0]

    def getcoordinates(self, cell):
        """getcoordinates! (grid.jl:612-624): node coords in cell-local order; 1-based cell."""
        return self.nodes[[n - 1 for n in self.cells[cell - 1]]]


def _extrema_to_corners(rdim, left, right):
    """grid_generators.jl:565-578 for rdim == sdim."""
    ip = Lagrange({2: "RefQuadrilateral", 3: "RefHexahedron"}[rdim], 1)
    corners = []
    for xi in ip.reference_coordinates():
        dxi = [xi[d] - (-1.0) for d in range(rdim)]
        dx = []
        for i in range(rdim):
            # A ⋅ dξ with A = diag(Δx/Δξ): Σ_j A_ij dξ_j, off-diagonal terms are exact zeros
            a_ii = (right[i] - left[i]) / (1.0 - (-1.0))
            acc = 0.0
            for j in range(rdim):
                acc = acc + (a_ii if i == j else 0.0) * dxi[j]
            dx.append(acc)
        corners.append([left[i] + dx[i] for i in range(rdim)])
    return corners


def _generate_nodes(rdim, nnodes, corners):
    """grid_generators.jl:550-557."""
    ip = Lagrange({2: "RefQuadrilateral", 3: "RefHexahedron"}[rdim], 1)
    out = []
    for idx0 in _product_first_fastest(list(nnodes)):
        xi = [2 * float(idx0[d]) / (nnodes[d] - 1) - 1 for d in range(rdim)]
        x = [0.0] * rdim
        for i in range(len(corners)):
            m = ip.shape_value(xi, i + 1)
            for d in range(rdim):
                x[d] += m * corners[i][d]
        out.append(x)
    return np.array(out, dtype=np.float64)


def generate_grid(celltype, nel, left=None, right=None):
    if celltype == "Quadrilateral":
        left = left if left is not None else (-1.0, -1.0)
        right = right if right is not None else (1.0, 1.0)
        return _quad_grid(nel, left, right)
    if celltype == "Hexahedron":
        left = left if left is not None else (-1.0, -1.0, -1.0)
        right = right if right is not None else (1.0, 1.0, 1.0)
        return _hex_grid(nel, left, right)
    raise ValueError(celltype)


def _quad_grid(nel, left, right):
    """grid_generators.jl:78-113."""
    nx, ny = nel
    nnx, nny = nx + 1, ny + 1
    nodes = _generate_nodes(2, (nnx, nny), _extrema_to_corners(2, left, right))

    def nid(i, j):  # 1-based node_array[i, j]
        return (j - 1) * nnx + i
    cells = []
    for j in range(1, ny + 1):
        for i in range(1, nx + 1):
            cells.append((nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)))

    def cid(i, j):
        return (j - 1) * nx + i
    fs = {
        "bottom": [(cid(i, 1), 1) for i in range(1, nx + 1)],
        "right": [(cid(nx, j), 2) for j in range(1, ny + 1)],
        "top": [(cid(i, ny), 3) for i in range(1, nx + 1)],
        "left": [(cid(1, j), 4) for j in range(1, ny + 1)],
    }
    for k in fs:
        fs[k] = sorted(fs[k], key=lambda t: t[0])
    return Grid("Quadrilateral", cells, nodes, fs)


def _hex_grid(nel, left, right):
    """grid_generators.jl:159-204."""
    nx, ny, nz = nel
    nnx, nny, nnz = nx + 1, ny + 1, nz + 1
    nodes = _generate_nodes(3, (nnx, nny, nnz), _extrema_to_corners(3, left, right))

    def nid(i, j, k):
        return ((k - 1) * nny + (j - 1)) * nnx + i
    cells = []
    for k in range(1, nz + 1):
        for j in range(1, ny + 1):
            for i in range(1, nx + 1):
                cells.append((nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                              nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1),
                              nid(i, j + 1, k + 1)))

    def cid(i, j, k):
        return ((k - 1) * ny + (j - 1)) * nx + i
    fs = {
        "bottom": [(cid(i, j, 1), 1) for j in range(1, ny + 1) for i in range(1, nx + 1)],
        "front": [(cid(i, 1, k), 2) for k in range(1, nz + 1) for i in range(1, nx + 1)],
        "right": [(cid(nx, j, k), 3) for k in range(1, nz + 1) for j in range(1, ny + 1)],
        "back": [(cid(i, ny, k), 4) for k in range(1, nz + 1) for i in range(1, nx + 1)],
        "left": [(cid(1, j, k), 5) for k in range(1, nz + 1) for j in range(1, ny + 1)],
        "top": [(cid(i, j, nz), 6) for j in range(1, ny + 1) for i in range(1, nx + 1)],
    }
    for k in fs:
        fs[k] = sorted(fs[k], key=lambda t: t[0])
    return Grid("Hexahedron", cells, nodes, fs)
