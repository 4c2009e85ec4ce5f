"""Oracle: the serial reference loop, end to end (TEST INFRASTRUCTURE ONLY).

Restates the canonical assembly loop of docs/src/literate-tutorials/heat_equation.jl:186-205
(start_assemble; for cell in CellIterator(dh): reinit!, fill!(Ke,0),
assemble_element!, assemble!) followed by apply! and K \\ f (:212-218).  The
per-cell conductivity k_e and source f are the 3D Poisson configuration's
extension of the same loop.
"""
import numpy as np

from . import constraints as C
from .coloring import create_coloring
from .dofs import DofHandler
from .fem import CSC, CellValues, assemble, poisson_element, start_assemble
from .grid import generate_grid
from .refel import Lagrange, QuadratureRule
from .sparsity import create_pattern


def setup_poisson(celltype, nel, left=None, right=None, bc_sets=None, qorder=2):
    grid = generate_grid(celltype, nel, left, right)
    shape = grid.refshape
    ip = Lagrange(shape, 1)
    qr = QuadratureRule(shape, qorder)
    cv = CellValues(qr, ip)
    dh = DofHandler(grid).add("u", ip).close()
    colptr, rowval = create_pattern(dh)
    K = CSC(dh.ndofs, colptr, rowval)
    f = np.zeros(dh.ndofs)
    ch = C.ConstraintHandler(dh)
    if bc_sets is None:
        bc_sets = list(grid.facetsets)
    facets = []
    for s in bc_sets:
        facets.extend(grid.facetsets[s])
    ch.add(C.Dirichlet("u", facets, lambda x, t: 0.0)).close()
    return grid, cv, dh, K, f, ch


def assemble_poisson(grid, cv, dh, K, f, k_cell=None, fsrc=1.0):
    A = start_assemble(K, f)
    for ci in range(1, grid.ncells + 1):
        x = grid.getcoordinates(ci)
        k = 1.0 if k_cell is None else float(k_cell[ci - 1])
        Ke, fe = poisson_element(cv, x, k, fsrc)
        assemble(A, dh.celldofs(ci), Ke, fe)
    return K, f


def heat_equation(n=20):
    """heat_equation.jl: 20x20 Q1 quad on [-1,1]², f=1, u=0 on ∂Ω -> norm(u) KAT."""
    grid, cv, dh, K, f, ch = setup_poisson("Quadrilateral", (n, n))
    assemble_poisson(grid, cv, dh, K, f)
    C.apply(K, f, ch)
    u = C.solve(K, f)
    return u, dict(grid=grid, cv=cv, dh=dh, K=K, f=f, ch=ch)


def poisson_conductivity(nel, seed, sigma=0.5):
    """Per-cell k_e = exp(g_e), g_e ~ N(0, sigma²) i.i.d. in cell-id order."""
    ncells = int(np.prod(nel))
    rng = np.random.default_rng(seed)
    return np.exp(rng.normal(0.0, sigma, size=ncells))


__all__ = ["setup_poisson", "assemble_poisson", "heat_equation", "poisson_conductivity",
           "create_coloring"]


def assemble_elasticity(grid, cv, dh, K, f, E_cell, nu):
    """Serial loop with the linear-elastic element (threaded_assembly.jl:105-132)."""
    from .fem import elasticity_element
    A = start_assemble(K, f)
    for ci in range(1, grid.ncells + 1):
        Ke = elasticity_element(cv, grid.getcoordinates(ci), float(E_cell[ci - 1]), nu)
        assemble(A, dh.celldofs(ci), Ke)
    return K, f


def assemble_traction(grid, fv, dh, f, facets, t):
    """f[celldofs] += ∫ t·N dΓ over the facet set (FacetIterator loop, topology_optimization.jl:369-381)."""
    from .fem import traction_element
    for (ci, fc) in facets:
        fe = traction_element(fv, grid.getcoordinates(ci), fc, np.asarray(t, dtype=np.float64))
        for i, d in enumerate(dh.celldofs(ci)):
            f[d - 1] += fe[i]
    return f


def setup_vector(celltype, nel, left, right, order, qorder):
    from .fem import CSC
    grid = generate_grid(celltype, nel, left, right)
    shape = grid.refshape
    dim = len(nel)
    ip = Lagrange(shape, order) ** dim
    cv = CellValues(QuadratureRule(shape, qorder), ip)
    dh = DofHandler(grid).add("u", ip).close()
    colptr, rowval = create_pattern(dh)
    K = CSC(dh.ndofs, colptr, rowval)
    f = np.zeros(dh.ndofs)
    return grid, cv, dh, K, f
