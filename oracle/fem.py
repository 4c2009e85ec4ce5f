"""Oracle: CellValues/FacetValues, element routines, CSC assembler (TEST INFRASTRUCTURE ONLY).

Restates, as literal per-cell loops:
  * CellValues construction (tabulated N, dNdξ, M, dMdξ at the quadrature
    points): src/FEValues/CellValues.jl:66-90, FunctionValues.jl:62-98,
    GeometryMapping.jl:68-98
  * reinit!(cv, x): CellValues.jl:122-140; calculate_mapping
    GeometryMapping.jl:124-130 (J = Σ_j x_j ⊗ dMdξ_j); detJ > 0 check
    CellValues.jl:110-116 / common_values.jl:5; dNdx = dNdξ ⋅ J⁻¹
    FunctionValues.jl:163, :186-192
  * FacetValues reinit: FacetValues.jl:95-156 (dΓ = |weighted_normal| w)
  * element routines: heat/Poisson docs/src/literate-tutorials/heat_equation.jl:143-164
    (with a per-cell conductivity k), linear elasticity
    docs/src/literate-howto/threaded_assembly.jl:105-132 (εᵢ : C : εⱼ with C
    isotropic from (E, ν)), traction fe_i += t·N_i dΓ
    docs/src/literate-tutorials/topology_optimization.jl:347-383
  * start_assemble / assemble!: src/assembler.jl:287-291, :322-457 (zero Ke
    entries skipped; a missing entry with a non-zero value raises)

det and inv use the cofactor formulas (Tensors.jl is not vendored in the
reference: parity at the ulp level is unpinned there).
"""
import numpy as np

from .refel import (Lagrange, VectorizedInterpolation, base_interpolation, weighted_normal)


def det(J):
    if J.shape == (2, 2):
        return J[0, 0] * J[1, 1] - J[0, 1] * J[1, 0]
    return (J[0, 0] * (J[1, 1] * J[2, 2] - J[1, 2] * J[2, 1])
            - J[0, 1] * (J[1, 0] * J[2, 2] - J[1, 2] * J[2, 0])
            + J[0, 2] * (J[1, 0] * J[2, 1] - J[1, 1] * J[2, 0]))


def inv(J):
    d = det(J)
    if J.shape == (2, 2):
        return np.array([[J[1, 1], -J[0, 1]], [-J[1, 0], J[0, 0]]]) / d
    a = np.empty((3, 3))
    a[0, 0] = J[1, 1] * J[2, 2] - J[1, 2] * J[2, 1]
    a[0, 1] = J[0, 2] * J[2, 1] - J[0, 1] * J[2, 2]
    a[0, 2] = J[0, 1] * J[1, 2] - J[0, 2] * J[1, 1]
    a[1, 0] = J[1, 2] * J[2, 0] - J[1, 0] * J[2, 2]
    a[1, 1] = J[0, 0] * J[2, 2] - J[0, 2] * J[2, 0]
    a[1, 2] = J[0, 2] * J[1, 0] - J[0, 0] * J[1, 2]
    a[2, 0] = J[1, 0] * J[2, 1] - J[1, 1] * J[2, 0]
    a[2, 1] = J[0, 1] * J[2, 0] - J[0, 0] * J[2, 1]
    a[2, 2] = J[0, 0] * J[1, 1] - J[0, 1] * J[1, 0]
    return a / d


class DetJError(ValueError):
    pass


def _tabulate(ip, points):
    """N[i, q] and dNdξ[i, q, :] of the scalar base interpolation."""
    nb = ip.nbasefunctions
    nq = len(points)
    N = np.zeros((nb, nq))
    dN = np.zeros((nb, nq, ip.rdim))
    for q, xi in enumerate(points):
        for i in range(nb):
            g, v = ip.shape_gradient_and_value(list(xi), i + 1)
            N[i, q] = v
            dN[i, q] = g
    return N, dN


class CellValues:
    """CellValues(qr, ip[, ip_geo]) with identity mapping and first-order derivatives."""

    def __init__(self, qr, ip, ip_geo=None):
        self.qr = qr
        self.ip = ip
        self.base = base_interpolation(ip)
        self.vdim = ip.vdim if isinstance(ip, VectorizedInterpolation) else 1
        self.ip_geo = ip_geo or Lagrange(self.base.shape, 1)
        self.N, self.dNdxi = _tabulate(self.base, qr.points)
        self.M, self.dMdxi = _tabulate(self.ip_geo, qr.points)
        self.nqp = qr.nqp
        self.nbf = self.base.nbasefunctions * self.vdim

    def reinit(self, x):
        """Returns (dNdx[i, q, :] for the scalar base, detJdV[q]); raises DetJError."""
        x = np.asarray(x, dtype=np.float64)
        if x.shape[0] != self.ip_geo.nbasefunctions:
            raise ValueError("the number of (geometric) base functions does not match the number of coordinates")
        nb = self.base.nbasefunctions
        sdim = x.shape[1]
        dNdx = np.zeros((nb, self.nqp, sdim))
        detJdV = np.zeros(self.nqp)
        for q, w in enumerate(self.qr.weights):
            J = np.zeros((sdim, self.ip_geo.rdim))
            for j in range(x.shape[0]):
                J = J + np.outer(x[j], self.dMdxi[j, q])
            dJ = det(J)
            if not dJ > 0.0:
                raise DetJError(f"det(J) is not positive: det(J) = {dJ}")
            detJdV[q] = dJ * w
            Jinv = inv(J)
            for i in range(nb):
                g = self.dNdxi[i, q]
                dNdx[i, q] = [sum(g[a] * Jinv[a, b] for a in range(len(g))) for b in range(sdim)]
        return dNdx, detJdV

    def spatial_coordinate(self, q, x):
        return sum(self.M[j, q] * np.asarray(x[j]) for j in range(len(x)))


class FacetValues:
    """FacetValues(fqr, ip): per facet, tabulated values and the weighted-normal facet measure."""

    def __init__(self, fqr, ip):
        self.fqr = fqr
        self.ip = ip
        self.base = base_interpolation(ip)
        self.vdim = ip.vdim if isinstance(ip, VectorizedInterpolation) else 1
        self.ip_geo = Lagrange(self.base.shape, 1)
        self.tables = []
        for pts, w in fqr.rules:
            N, dN = _tabulate(self.base, pts)
            M, dM = _tabulate(self.ip_geo, pts)
            self.tables.append((N, dN, M, dM, w))

    def reinit(self, x, facet):
        """Returns (N[i, q] of the scalar base, dΓ[q])."""
        N, _, _, dM, w = self.tables[facet - 1]
        x = np.asarray(x, dtype=np.float64)
        dG = np.zeros(len(w))
        for q in range(len(w)):
            J = np.zeros((x.shape[1], self.ip_geo.rdim))
            for j in range(x.shape[0]):
                J = J + np.outer(x[j], dM[j, q])
            n = weighted_normal(J, self.base.shape, facet)
            dG[q] = np.sqrt(np.dot(n, n)) * w[q]
        return N, dG


# ----------------------------------------------------------------------------
# Element routines
# ----------------------------------------------------------------------------
def poisson_element(cv, x, k=1.0, fsrc=1.0):
    """Ke[i,j] += k (∇Nᵢ·∇Nⱼ) dΩ ; fe[i] += f Nᵢ dΩ (heat_equation.jl:143-164)."""
    dNdx, dV = cv.reinit(x)
    nb = cv.nbf
    Ke = np.zeros((nb, nb))
    fe = np.zeros(nb)
    for q in range(cv.nqp):
        dO = dV[q]
        for i in range(nb):
            fe[i] += fsrc * cv.N[i, q] * dO
            for j in range(nb):
                Ke[i, j] += k * np.dot(dNdx[i, q], dNdx[j, q]) * dO
    return Ke, fe


def isotropic_C(E, nu, dim):
    """Voigt-free 4th-order isotropic elasticity tensor; plane strain for dim == 2."""
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    C = np.zeros((dim, dim, dim, dim))
    for i in range(dim):
        for j in range(dim):
            for k in range(dim):
                for l in range(dim):
                    C[i, j, k, l] = (lam * (i == j) * (k == l)
                                     + mu * ((i == k) * (j == l) + (i == l) * (j == k)))
    return C


def _sym_grad(dNdx_iq, comp, dim):
    """shape_symmetric_gradient of the vector basis function (node i, component comp)."""
    G = np.zeros((dim, dim))
    G[comp, :] = dNdx_iq
    return 0.5 * (G + G.T)


def elasticity_element(cv, x, E, nu):
    """Ke[I,J] += (ε_I : C : ε_J) dΩ (threaded_assembly.jl:105-132) for ip^dim vector Lagrange.

    Written with numpy contractions over the basis index (same formula, evaluated for all
    I at once) so that Q2 hex cells stay fast enough for tests."""
    dim = cv.vdim
    dNdx, dV = cv.reinit(x)
    C = E * isotropic_C(1.0, nu, dim)
    nb = cv.nbf
    Ke = np.zeros((nb, nb))
    for q in range(cv.nqp):
        G = np.zeros((nb, dim, dim))
        for I in range(nb):
            G[I, I % dim, :] = dNdx[I // dim, q]
        eps = 0.5 * (G + G.transpose(0, 2, 1))
        Ceps = np.einsum("ijkl,Ikl->Iij", C, eps)
        Ke += np.einsum("Iij,Jij->IJ", eps, Ceps) * dV[q]
    return Ke


def traction_element(fv, x, facet, t):
    """fe[I] += t[comp]·N_node dΓ for vector Lagrange (topology_optimization.jl:369-381)."""
    N, dG = fv.reinit(x, facet)
    dim = fv.vdim
    nb = fv.base.nbasefunctions * dim
    fe = np.zeros(nb)
    for q in range(len(dG)):
        for I in range(nb):
            fe[I] += t[I % dim] * N[I // dim, q] * dG[q]
    return fe


# ----------------------------------------------------------------------------
# CSC assembler (assembler.jl:287-457)
# ----------------------------------------------------------------------------
class MissingPatternError(RuntimeError):
    pass


class CSC:
    """1-based SparseMatrixCSC restatement: colptr (n+1), rowval (nnz), nzval (nnz)."""

    def __init__(self, n, colptr, rowval, nzval=None):
        self.n = n
        self.colptr = np.asarray(colptr, dtype=np.int64)
        self.rowval = np.asarray(rowval, dtype=np.int64)
        self.nzval = np.zeros(len(self.rowval)) if nzval is None else nzval

    def find(self, r, c):
        lo, hi = self.colptr[c - 1] - 1, self.colptr[c] - 1
        k = lo + np.searchsorted(self.rowval[lo:hi], r)
        if k < hi and self.rowval[k] == r:
            return k
        return -1

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csc_matrix((self.nzval, self.rowval - 1, self.colptr - 1), shape=(self.n, self.n))


def start_assemble(K, f):
    K.nzval[:] = 0.0
    f[:] = 0.0
    return (K, f)


def assemble(A, dofs, Ke, fe=None):
    K, f = A
    if fe is not None:
        for i, d in enumerate(dofs):
            f[d - 1] += fe[i]
    order = sorted(range(len(dofs)), key=lambda i: dofs[i])
    for cj in order:
        col = dofs[cj]
        for ri in order:
            val = Ke[ri, cj]
            p = K.find(dofs[ri], col)
            if p < 0:
                if val != 0:
                    raise MissingPatternError(
                        f"You are trying to assemble values in to K[{dofs[ri]}, {col}], "
                        f"but K[{dofs[ri]}, {col}] is missing in the sparsity pattern.")
                continue
            if val != 0:
                K.nzval[p] += val
