"""Grid and generate_grid of the host interface.

Mirrors src/Grid/grid.jl (Grid, Node, Quadrilateral, Hexahedron, getcoordinates)
and src/Grid/grid_generators.jl:78-113, :159-204, :542-578.  Node coordinates
are produced by the same blend x = Σ_i M_i(ξ) corner_i (Q1 shape functions,
summed in corner order, ξ = 2(i-1)/(n-1) - 1), evaluated elementwise with
numpy in the reference's operation order, so they are bitwise identical to
the reference's.  Ids are 1-based like Ferrite's.
"""
import numpy as np

Quadrilateral, Hexahedron = "Quadrilateral", "Hexahedron"
CELL_REFSHAPE = {Quadrilateral: "RefQuadrilateral", Hexahedron: "RefHexahedron"}
NODES_PER_CELL = {Quadrilateral: 4, Hexahedron: 8}


class Grid:
    """Grid(cells, nodes; facetsets): cells (ncells, npc) int64 1-based, nodes (nnodes, sdim)."""

    def __init__(self, celltype, cells, nodes, facetsets=None, nodesets=None):
        self.celltype = celltype
        self.cells = np.ascontiguousarray(cells, dtype=np.int64)
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.facetsets = facetsets or {}
        self.nodesets = nodesets or {}
        if self.cells.ndim != 2 or self.cells.shape[1] != NODES_PER_CELL[celltype]:
            raise ValueError("cells must be (ncells, nodes_per_cell)")

    @property
    def refshape(self):
        return CELL_REFSHAPE[self.celltype]

    @property
    def ncells(self):
        return self.cells.shape[0]

    @property
    def nnodes(self):
        return self.nodes.shape[0]

    @property
    def sdim(self):
        return self.nodes.shape[1]

    def getcoordinates(self, cell):
        """Coordinates of a cell's nodes in local order (grid.jl:612-624); 1-based cell."""
        return self.nodes[self.cells[cell - 1] - 1]

    def getfacetset(self, name):
        return self.facetsets[name]

    def addfacetset(self, name, predicate, all_nodes=True):
        """addfacetset!(grid, name, f) (Grid/utils.jl:40-98): facets whose nodes satisfy f."""
        from .refel import REF_FACETS
        facets = REF_FACETS[self.refshape]
        out = []
        for f, loc in enumerate(facets, start=1):
            ok = np.ones(self.ncells, dtype=bool) if all_nodes else np.zeros(self.ncells, dtype=bool)
            for v in loc:
                x = self.nodes[self.cells[:, v - 1] - 1]
                r = np.array([bool(predicate(xi)) for xi in x])
                ok = (ok & r) if all_nodes else (ok | r)
            for c in np.nonzero(ok)[0]:
                out.append((int(c) + 1, f))
        out.sort(key=lambda t: t[0])
        self.facetsets[name] = out
        return out


class NonConformingGrid(Grid):
    """NonConformingGrid(cells, nodes; facetsets, conformity_info) (Adaptivity/ncgrid.jl:24-50):
    a grid with hanging vertices, each lying on an edge (or, in 3D, a face) of a neighbouring
    cell that does not have it as a vertex.  conformity_info maps every hanging node to its 2
    (edge midpoint) or 4 (face centre) master nodes (1-based ids)."""

    def __init__(self, celltype, cells, nodes, conformity_info, facetsets=None, nodesets=None):
        super().__init__(celltype, cells, nodes, facetsets, nodesets)
        info = {}
        for h, masters in dict(conformity_info).items():
            ms = [int(m) for m in masters]
            if not 1 <= int(h) <= self.nnodes or len(ms) not in (2, 4) or not all(1 <= m <= self.nnodes for m in ms):
                raise ValueError(f"conformity_info entry {h} => {masters}: a hanging node and 2 or 4 master nodes")
            info[int(h)] = ms
        self.conformity_info = info


def _extrema_to_corners(rdim, left, right):
    """grid_generators.jl:565-578: corner = left + (Δx/Δξ)(ξ - ξ_left) per axis."""
    from .refel import REF_COORDS
    shape = {2: "RefQuadrilateral", 3: "RefHexahedron"}[rdim]
    corners = []
    for xi in REF_COORDS[(shape, 1)]:
        c = []
        for i in range(rdim):
            a_ii = (right[i] - left[i]) / (1.0 - (-1.0))
            dx = 0.0
            for j in range(rdim):
                dx = dx + (a_ii if i == j else 0.0) * (float(xi[j]) - (-1.0))
            c.append(left[i] + dx)
        corners.append(c)
    return corners


def _q1_blend(rdim, nnodes, corners):
    """_generate_nodes (grid_generators.jl:550-557), vectorised over all nodes."""
    q = np.arange(int(np.prod(nnodes)))
    xi = []
    stride = 1
    for d in range(rdim):
        idx = (q // stride) % nnodes[d]
        stride *= nnodes[d]
        xi.append(2 * idx.astype(np.float64) / (nnodes[d] - 1) - 1)
    if rdim == 2:
        x, y = xi
        M = [(1 - x) * (1 - y) / 4, (1 + x) * (1 - y) / 4, (1 + x) * (1 + y) / 4, (1 - x) * (1 + y) / 4]
    else:
        x, y, z = xi
        M = [(1 - x) * (1 - y) * (1 - z) / 8, (1 + x) * (1 - y) * (1 - z) / 8,
             (1 + x) * (1 + y) * (1 - z) / 8, (1 - x) * (1 + y) * (1 - z) / 8,
             (1 - x) * (1 - y) * (1 + z) / 8, (1 + x) * (1 - y) * (1 + z) / 8,
             (1 + x) * (1 + y) * (1 + z) / 8, (1 - x) * (1 + y) * (1 + z) / 8]
    out = np.zeros((q.size, rdim))
    for d in range(rdim):
        acc = np.zeros(q.size)
        for a in range(len(M)):
            acc = acc + M[a] * corners[a][d]
        out[:, d] = acc
    return out


def generate_grid(celltype, nel, left=None, right=None):
    """generate_grid(Quadrilateral|Hexahedron, nel, left, right); default domain [-1, 1]^dim."""
    dim = len(nel)
    if celltype == Quadrilateral and dim != 2 or celltype == Hexahedron and dim != 3:
        raise ValueError("nel does not match the cell dimension")
    left = tuple(float(v) for v in (left if left is not None else (-1.0,) * dim))
    right = tuple(float(v) for v in (right if right is not None else (1.0,) * dim))
    nn = tuple(n + 1 for n in nel)
    nodes = _q1_blend(dim, nn, _extrema_to_corners(dim, left, right))
    if dim == 2:
        nx, ny = nel
        j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
        i, j = i.ravel(), j.ravel()
        n0 = j * nn[0] + i + 1
        cells = np.stack([n0, n0 + 1, n0 + 1 + nn[0], n0 + nn[0]], axis=1)
        cid = np.arange(nx * ny).reshape(ny, nx) + 1  # [j, i]
        fs = {
            "bottom": [(int(c), 1) for c in cid[0, :]],
            "right": [(int(c), 2) for c in cid[:, -1]],
            "top": [(int(c), 3) for c in cid[-1, :]],
            "left": [(int(c), 4) for c in cid[:, 0]],
        }
    else:
        nx, ny, nz = nel
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        i, j, k = i.ravel(), j.ravel(), k.ravel()
        n0 = (k * nn[1] + j) * nn[0] + i + 1
        px, pz = nn[0], nn[0] * nn[1]
        cells = np.stack([n0, n0 + 1, n0 + 1 + px, n0 + px,
                          n0 + pz, n0 + 1 + pz, n0 + 1 + px + pz, n0 + px + pz], axis=1)
        cid = np.arange(nx * ny * nz).reshape(nz, ny, nx) + 1  # [k, j, i]
        fs = {
            "bottom": [(int(c), 1) for c in cid[0, :, :].ravel()],
            "front": [(int(c), 2) for c in cid[:, 0, :].ravel()],
            "right": [(int(c), 3) for c in cid[:, :, -1].ravel()],
            "back": [(int(c), 4) for c in cid[:, -1, :].ravel()],
            "left": [(int(c), 5) for c in cid[:, :, 0].ravel()],
            "top": [(int(c), 6) for c in cid[-1, :, :].ravel()],
        }
    for key in fs:
        fs[key].sort(key=lambda t: t[0])
    return Grid(celltype, cells, nodes, fs)
