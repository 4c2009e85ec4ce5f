"""Oracle: sparsity pattern and CSC allocation (TEST INFRASTRUCTURE ONLY, see oracle/__init__.py).

Restates the semantics of src/Dofs/sparsity_pattern.jl:223-253
(_build_pattern!: row r holds the diagonal plus the dofs of every cell that
contains r), :1136-1249 (_visit_row_candidates!) and :951-1016
(_allocate_matrix: counting transpose, rowval ascending within a column).
Written as the brute-force set union, which is what pins the reference's
counting construction.  Returns 1-based colptr/rowval like SparseMatrixCSC.
"""
import numpy as np


def create_pattern(dh):
    n = dh.ndofs
    rows = [set() for _ in range(n + 1)]
    for ci in range(1, dh.grid.ncells + 1):
        dofs = dh.celldofs(ci)
        for r in dofs:
            rows[r].update(dofs)
    for r in range(1, n + 1):
        rows[r].add(r)
    # The pattern is structurally symmetric, so column j holds the rows of row j.
    colptr = np.zeros(n + 1, dtype=np.int64)
    colptr[0] = 1
    rowval = []
    for j in range(1, n + 1):
        col = sorted(rows[j])
        rowval.extend(col)
        colptr[j] = colptr[j - 1] + len(col)
    return colptr, np.array(rowval, dtype=np.int64)


def create_pattern_fast(cell_dofs_2d, ndofs):
    """Vectorised brute force for larger oracle cases (same set semantics)."""
    cd = np.asarray(cell_dofs_2d, dtype=np.int64) - 1
    nb = cd.shape[1]
    r = np.repeat(cd, nb, axis=1).ravel()
    c = np.tile(cd, (1, nb)).ravel()
    key = np.concatenate([c * ndofs + r, np.arange(ndofs, dtype=np.int64) * (ndofs + 1)])
    key = np.unique(key)
    cols = key // ndofs
    rows = key % ndofs
    colptr = np.zeros(ndofs + 1, dtype=np.int64)
    np.add.at(colptr, cols + 1, 1)
    colptr = np.cumsum(colptr) + 1
    return colptr, rows + 1
