"""Oracle: the serial Poisson loop restated over arrays of cells (TEST INFRASTRUCTURE ONLY).

Same arithmetic as oracle/fem.py's per-cell loop, batched with numpy so the oracle reaches
meshes of 10^4-10^5 cells in seconds:
  * reinit!(cv, x) per cell and quadrature point: J = Σ_j x_j ⊗ dMdξ_j
    (src/FEValues/GeometryMapping.jl:124-130), det J > 0 (CellValues.jl:110-116),
    dNdx = dNdξ ⋅ J⁻¹ (FunctionValues.jl:163, :186-192), dΩ = det J · w;
  * Poisson element: Ke[i, j] = k Σ_q (∇N_i · ∇N_j) dΩ, fe[i] = s Σ_q N_i dΩ
    (docs/src/literate-tutorials/heat_equation.jl:143-164, with a per-cell conductivity);
  * assemble!(K, dofs, Ke, fe) for every cell in ascending cell id
    (src/assembler.jl:322-457): np.add.at over cell-major arrays adds each entry's
    contributions in ascending cell order, the order of the serial loop.
det and inv use the same cofactor formulas as oracle/fem.py.  Pinned against that serial
loop (tests/test_oracle.py: ≤1e-14 on quad and hex grids, distorted ones included).
"""
import numpy as np


def _det_inv(J):
    """Batched cofactor det and inverse of (..., d, d) matrices, d = 2 or 3."""
    if J.shape[-1] == 2:
        d = J[..., 0, 0] * J[..., 1, 1] - J[..., 0, 1] * J[..., 1, 0]
        a = np.empty_like(J)
        a[..., 0, 0] = J[..., 1, 1]
        a[..., 0, 1] = -J[..., 0, 1]
        a[..., 1, 0] = -J[..., 1, 0]
        a[..., 1, 1] = J[..., 0, 0]
        return d, a / d[..., None, None]
    d = (J[..., 0, 0] * (J[..., 1, 1] * J[..., 2, 2] - J[..., 1, 2] * J[..., 2, 1])
         - J[..., 0, 1] * (J[..., 1, 0] * J[..., 2, 2] - J[..., 1, 2] * J[..., 2, 0])
         + J[..., 0, 2] * (J[..., 1, 0] * J[..., 2, 1] - J[..., 1, 1] * J[..., 2, 0]))
    a = np.empty_like(J)
    a[..., 0, 0] = J[..., 1, 1] * J[..., 2, 2] - J[..., 1, 2] * J[..., 2, 1]
    a[..., 0, 1] = J[..., 0, 2] * J[..., 2, 1] - J[..., 0, 1] * J[..., 2, 2]
    a[..., 0, 2] = J[..., 0, 1] * J[..., 1, 2] - J[..., 0, 2] * J[..., 1, 1]
    a[..., 1, 0] = J[..., 1, 2] * J[..., 2, 0] - J[..., 1, 0] * J[..., 2, 2]
    a[..., 1, 1] = J[..., 0, 0] * J[..., 2, 2] - J[..., 0, 2] * J[..., 2, 0]
    a[..., 1, 2] = J[..., 0, 2] * J[..., 1, 0] - J[..., 0, 0] * J[..., 1, 2]
    a[..., 2, 0] = J[..., 1, 0] * J[..., 2, 1] - J[..., 1, 1] * J[..., 2, 0]
    a[..., 2, 1] = J[..., 0, 1] * J[..., 2, 0] - J[..., 0, 0] * J[..., 2, 1]
    a[..., 2, 2] = J[..., 0, 0] * J[..., 1, 1] - J[..., 0, 1] * J[..., 1, 0]
    return d, a / d[..., None, None]


def poisson_elements(cv, X, k_cell=None, fsrc=1.0):
    """Ke (nc, nb, nb) and fe (nc, nb) of every cell; X = (nc, ngeo, sdim) node coordinates."""
    X = np.asarray(X, dtype=np.float64)
    w = np.asarray(cv.qr.weights, dtype=np.float64)
    J = np.einsum("cja,jqb->cqab", X, cv.dMdxi)              # J_ab = Σ_j x_j[a] dMdξ_j[b]
    d, Jinv = _det_inv(J)
    bad = np.nonzero(~(d > 0.0))[0]
    if bad.size:
        raise ValueError(f"det(J) is not positive in cell {bad[0] // J.shape[1] + 1}")
    dNdx = np.einsum("iqa,cqab->cqib", cv.dNdxi, Jinv)       # dNdx_b = Σ_a dNdξ_a Jinv_ab
    dO = d * w[None, :]
    k = np.ones(X.shape[0]) if k_cell is None else np.asarray(k_cell, dtype=np.float64)
    Ke = k[:, None, None] * np.einsum("cq,cqid,cqjd->cij", dO, dNdx, dNdx)
    fe = fsrc * np.einsum("cq,iq->ci", dO, cv.N)
    return Ke, fe


def nz_positions(colptr, rowval, cell_dofs):
    """pos[c, i, j] = index in nzval of K[dofs_i, dofs_j] (0-based), −1 when missing."""
    cp = np.asarray(colptr, dtype=np.int64) - 1
    rv = np.asarray(rowval, dtype=np.int64) - 1
    n = len(cp) - 1
    cols = np.repeat(np.arange(n, dtype=np.int64), np.diff(cp))
    keys = cols * n + rv                                     # ascending: CSC order
    cd = np.asarray(cell_dofs, dtype=np.int64) - 1
    want = cd[:, None, :] * n + cd[:, :, None]                # [c, i (row), j (col)]
    pos = np.searchsorted(keys, want)
    pos = np.minimum(pos, len(keys) - 1)
    return np.where(keys[pos] == want, pos, -1)


def assemble_poisson_vec(grid, cv, cell_dofs, colptr, rowval, k_cell=None, fsrc=1.0):
    """nzval and f of the serial loop start_assemble + assemble! over all cells (FP64)."""
    cells = np.asarray(grid.cells, dtype=np.int64) - 1
    X = grid.nodes[cells]
    Ke, fe = poisson_elements(cv, X, k_cell, fsrc)
    pos = nz_positions(colptr, rowval, cell_dofs)
    nz = Ke != 0.0                                            # zero entries are skipped
    if np.any(nz & (pos < 0)):
        raise KeyError("entry missing in the sparsity pattern")
    nzval = np.zeros(int(colptr[-1]) - 1)
    np.add.at(nzval, pos[nz], Ke[nz])
    f = np.zeros(len(colptr) - 1)
    np.add.at(f, np.asarray(cell_dofs, dtype=np.int64).ravel() - 1, fe.ravel())
    return nzval, f
