"""DofHandler, sparsity pattern and coloring of the host interface.

`DofHandler(grid)`, `add!(dh, name, ip)`, `close!(dh)` (src/Dofs/DofHandler.jl:169,
:373-451, :474-574), `allocate_matrix(dh)`'s pattern (sparsity_pattern.jl:628-645)
and `create_coloring(grid; alg)` (Grid/coloring.jl:308-317), computed by the
library's native setup routines (fem_setup_*) with the reference's algorithms.
"""
import numpy as np

from . import _lib as L
from .refel import base_interpolation, n_components

_REF = {"RefLine": L.FEM_REF_LINE, "RefQuadrilateral": L.FEM_REF_QUADRILATERAL,
        "RefHexahedron": L.FEM_REF_HEXAHEDRON}


class DofHandler:
    def __init__(self, grid):
        self.grid = grid
        self.field_names = []
        self.field_ips = []
        self.closed = False

    def add(self, name, ip):
        if self.closed:
            raise RuntimeError("DofHandler is closed")
        if base_interpolation(ip).shape != self.grid.refshape:
            raise ValueError("interpolation does not match the cell's reference shape")
        self.field_names.append(name)
        self.field_ips.append(ip)
        return self

    def close(self, device=None):
        """close!(dh).  With `device` (a GPU index) the numbering is computed on that GPU
        (fem_dofs_distribute, one field of quadrilaterals or hexahedra); the context it ran
        in is kept for allocate_matrix, which then builds the pattern there as well."""
        if device is not None:
            return self._close_device(device)
        g = self.grid
        nf = len(self.field_ips)
        order = np.array([base_interpolation(ip).order for ip in self.field_ips], dtype=np.int32)
        vdim = np.array([n_components(ip) for ip in self.field_ips], dtype=np.int32)
        cells = L.as_i64(g.cells)
        ndpc = L.C.c_int32(0)
        nd = L.C.c_int64(0)
        lib = L.lib()
        rc = lib.fem_setup_distribute_dofs(_REF[g.refshape], g.ncells, cells.shape[1], L.ptr(cells, L.C.c_int64),
                                           g.nnodes, nf, L.ptr(order, L.C.c_int32), L.ptr(vdim, L.C.c_int32),
                                           None, L.C.byref(ndpc), L.C.byref(nd))
        L.raise_for(rc, what="dof distribution failed")
        cd = np.zeros(g.ncells * ndpc.value, dtype=np.int64)
        rc = lib.fem_setup_distribute_dofs(_REF[g.refshape], g.ncells, cells.shape[1], L.ptr(cells, L.C.c_int64),
                                           g.nnodes, nf, L.ptr(order, L.C.c_int32), L.ptr(vdim, L.C.c_int32),
                                           L.ptr(cd, L.C.c_int64), L.C.byref(ndpc), L.C.byref(nd))
        L.raise_for(rc, what="dof distribution failed")
        self.ndofs_per_cell = ndpc.value
        self.cell_dofs = cd.reshape(g.ncells, ndpc.value)
        self.ndofs = nd.value
        offs, o = [], 0
        for ip in self.field_ips:
            offs.append(o)
            o += ip.nbasefunctions
        self.field_offsets = offs
        self.closed = True
        return self

    def _close_device(self, device):
        from .device import Context
        if len(self.field_ips) != 1:
            raise ValueError("close(device=...) distributes one field")
        g = self.grid
        ip = self.field_ips[0]
        ctx = Context(device)
        xyz = L.as_f64(g.nodes)
        cells = L.as_i64(g.cells)
        ctx.call("fem_mesh_upload", g.sdim, g.nnodes, L.ptr(xyz, L.C.c_double), g.ncells, cells.shape[1],
                 L.ptr(cells, L.C.c_int64))
        nd, ndpc = L.C.c_int64(0), L.C.c_int32(0)
        ctx.call("fem_dofs_distribute", base_interpolation(ip).order, n_components(ip), L.C.byref(nd),
                 L.C.byref(ndpc))
        cd = np.zeros(g.ncells * ndpc.value, dtype=np.int64)
        ctx.call("fem_dofs_download", L.ptr(cd, L.C.c_int64))
        self.ndofs_per_cell = ndpc.value
        self.cell_dofs = cd.reshape(g.ncells, ndpc.value)
        self.ndofs = nd.value
        self.field_offsets = [0]
        self.closed = True
        self._ctx = ctx
        return self

    def celldofs(self, cell):
        return self.cell_dofs[cell - 1]

    def dof_range(self, name):
        i = self.field_names.index(name)
        return range(self.field_offsets[i], self.field_offsets[i] + self.field_ips[i].nbasefunctions)


def add(dh, name, ip):
    """add!(dh, name, ip)."""
    return dh.add(name, ip)


def close(obj):
    """close!(dh) / close!(ch)."""
    return obj.close()


def create_sparsity_pattern(dh):
    """The CSC structure of allocate_matrix(dh): (colptr, rowval), 1-based int64."""
    if not dh.closed:
        raise RuntimeError("close! the DofHandler first")
    cd = L.as_i64(dh.cell_dofs)
    n = dh.ndofs
    colptr = np.zeros(n + 1, dtype=np.int64)
    nnz = L.C.c_int64(0)
    lib = L.lib()
    rc = lib.fem_setup_pattern(dh.grid.ncells, dh.ndofs_per_cell, L.ptr(cd, L.C.c_int64), n,
                               L.ptr(colptr, L.C.c_int64), None, 0, L.C.byref(nnz))
    L.raise_for(rc, what="pattern construction failed")
    rowval = np.zeros(nnz.value, dtype=np.int64)
    rc = lib.fem_setup_pattern(dh.grid.ncells, dh.ndofs_per_cell, L.ptr(cd, L.C.c_int64), n,
                               L.ptr(colptr, L.C.c_int64), L.ptr(rowval, L.C.c_int64), nnz.value, L.C.byref(nnz))
    L.raise_for(rc, what="pattern construction failed")
    return colptr, rowval


def add_constraint_entries(colptr, rowval, ch):
    """_add_constraint_entries!(sp, ch; keep_constrained = true) (sparsity_pattern.jl:782-843):
    every entry (r, c) of the pattern adds (r', c') for r' in E(r), c' in E(c), where E maps an
    affinely constrained dof to its masters and any other dof to itself; with the boolean
    expansion B (B[d, e] = 1 iff e in E(d)) the result is P ∪ BᵀPB."""
    import scipy.sparse as sp
    if not getattr(ch, "has_affine", False):
        return colptr, rowval
    n = len(colptr) - 1
    dofs, cptr, cdofs, _, _ = ch.affine_arrays()
    keep = np.ones(n, dtype=bool)
    keep[dofs - 1] = False
    ki = np.nonzero(keep)[0]
    ri = np.concatenate([ki, np.repeat(dofs - 1, np.diff(cptr))])
    ci = np.concatenate([ki, cdofs - 1])
    B = sp.csr_matrix((np.ones(ri.size, dtype=np.int8), (ri, ci)), shape=(n, n))
    B.data[:] = 1
    P = sp.csc_matrix((np.ones(rowval.size, dtype=np.int8), rowval - 1,
                       colptr - 1), shape=(n, n))
    Q = ((B.T @ P @ B) + P).tocsc()
    Q.sum_duplicates()
    Q.sort_indices()
    return Q.indptr.astype(np.int64) + 1, Q.indices.astype(np.int64) + 1


def create_coloring(grid, alg="workstream"):
    """create_coloring(grid; alg = WorkStream | Greedy): list of ascending 1-based cell-id arrays."""
    a = {"workstream": L.FEM_COLORING_WORKSTREAM, "greedy": L.FEM_COLORING_GREEDY}[alg.lower()]
    cells = L.as_i64(grid.cells)
    col = np.zeros(grid.ncells, dtype=np.int32)
    nc = L.C.c_int32(0)
    rc = L.lib().fem_setup_coloring(grid.ncells, cells.shape[1], L.ptr(cells, L.C.c_int64), grid.nnodes, a,
                                    L.ptr(col, L.C.c_int32), L.C.byref(nc))
    L.raise_for(rc, what="coloring failed")
    order = np.argsort(col, kind="stable")
    bounds = np.searchsorted(col[order], np.arange(1, nc.value + 2))
    return [order[bounds[k]:bounds[k + 1]].astype(np.int64) + 1 for k in range(nc.value)]
