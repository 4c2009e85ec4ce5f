"""The element-assembly solver operator (K x as per-cell products with the element matrices the
assembly stored; docs/src/literate-howto/gpu_assembly.jl:265-304) against the oracle.

Bars: u within 1e-8 of the oracle's direct solve for the mixed solver (FP32 operator, FP64
residual) and the FP64 solver on the operator; the operator is used exactly when K is the
last fillzero assembly followed by apply!, and the solver falls back to its copy of K
otherwise.
"""
import ctypes as C

import numpy as np
import pytest

import ferrite_amd as fa
from oracle import constraints as OC
from oracle.problems import (assemble_elasticity, assemble_poisson, poisson_conductivity, setup_poisson,
                             setup_vector)

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _distort(X, amount, seed, lo=0.0, hi=1.0):
    inner = np.all((X > lo + 1e-12) & (X < hi - 1e-12), axis=1)
    X[inner] += np.random.default_rng(seed).uniform(-amount, amount, (int(inner.sum()), X.shape[1]))


def _solves(K, f, u_ref, modes=("mixed", "cg")):
    """mixed on the element operator, FP64 on it, and the FP64 solver on the sliced copy."""
    out = {}
    fa.set_tuning(K, lattice=-1, elem_operator=1)
    for mode in modes:
        r = fa.solve(K, f, mode=mode, rtol=1e-12, maxit=20000)
        assert r.operator == "elements", (mode, r.operator)
        assert r.relres <= 1e-12 and _rel(r.u, u_ref) < 1e-8, (mode, _rel(r.u, u_ref))
        out[mode] = r
    fa.set_tuning(K, lattice=-1, elem_operator=-1)
    s = fa.solve(K, f, mode="cg", rtol=1e-12, maxit=20000)
    assert s.operator == "sell" and _rel(s.u, out[modes[0]].u) < 1e-8
    fa.set_tuning(K, lattice=0, elem_operator=0)
    again = fa.solve(K, f, mode=modes[0], rtol=1e-12, maxit=20000)
    return out, again


def test_distorted_q1_hex_poisson():
    print("case elemop distorted q1 hex", flush=True)
    nel = (10, 8, 7)
    k = poisson_conductivity(nel, seed=21)
    g = fa.generate_grid("Hexahedron", nel, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    _distort(g.nodes, 0.3 / 10, 5)
    ip = fa.Lagrange("RefHexahedron", 1)
    cv = fa.CellValues(fa.QuadratureRule("RefHexahedron", 2), ip)
    dh = fa.DofHandler(g).add("u", ip).close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh)
    f = fa.allocate_vector(K)
    el = fa.PoissonElement(k=k)
    colors = fa.create_coloring(g)
    fa.assemble_global(fa.start_assemble(K, f), cv, colors, el)
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.Dirichlet("u", g.facetsets["left"] + g.facetsets["top"], lambda x, t: 0.5 * x[1])).close()
    fa.apply(K, f, ch)
    grid, ocv, odh, OK, of, _ = setup_poisson("Hexahedron", nel, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    grid.nodes[:] = g.nodes
    assemble_poisson(grid, ocv, odh, OK, of, k_cell=k)
    och = OC.ConstraintHandler(odh)
    och.add(OC.Dirichlet("u", grid.facetsets["left"] + grid.facetsets["top"], lambda x, t: 0.5 * x[1])).close()
    OC.apply(OK, of, och)
    u_ref = OC.solve(OK, of)
    out, again = _solves(K, f, u_ref)
    assert again.operator == "lattice"      # the lattice planes come first where they apply
    # the operator is dropped once K is no longer that assembly (+ apply!) ...
    nz = K.nzval()
    K.set_nzval(nz)
    fa.set_tuning(K, lattice=-1, elem_operator=1)
    assert fa.solve(K, f, mode="mixed", rtol=1e-12).operator == "sell"
    # ... or after an accumulating assembly
    fa.assemble_global(fa.start_assemble(K, f, fillzero=False), cv, colors, el)
    assert fa.solve(K, f, mode="mixed", rtol=1e-10, maxit=20000).operator == "sell"
    # a fresh assembly (+ apply!) makes it usable again, with the same result
    fa.assemble_global(fa.start_assemble(K, f), cv, colors, el)
    fa.apply(K, f, ch)
    r = fa.solve(K, f, mode="mixed", rtol=1e-12)
    assert r.operator == "elements" and np.array_equal(r.u, out["mixed"].u)


@pytest.mark.parametrize("stiffness", ["simp", "mild"])
def test_cantilever_q1_quad_vector(stiffness):
    """The plane-strain cantilever, with the SIMP stiffness of the differentiation workload
    (E = 1e-3 + θ³(1 − 1e-3), θ ~ U[0.2, 1]) and with E ~ U[0.5, 1].  cond(K) is large enough
    here that the mixed solver's FP32 recurrence claims its drop for several hand-overs in a
    row without moving the true residual much before it resumes (the stall rule allows six)."""
    print("case elemop cantilever q1^2", stiffness, flush=True)
    nel = (24, 12)
    rng = np.random.default_rng(3)
    if stiffness == "simp":
        E = 1e-3 + rng.uniform(0.2, 1.0, nel[0] * nel[1]) ** 3 * (1 - 1e-3)
    else:
        E = rng.uniform(0.5, 1.0, nel[0] * nel[1])
    g = fa.generate_grid("Quadrilateral", nel, (0.0, 0.0), (2.0, 1.0))
    ip = fa.Lagrange("RefQuadrilateral", 1) ** 2
    cv = fa.CellValues(fa.QuadratureRule("RefQuadrilateral", 2), ip)
    dh = fa.DofHandler(g).add("u", ip).close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh)
    f = fa.allocate_vector(K)
    for precision in ("fp64", "mixed"):
        fa.assemble_global(fa.start_assemble(K, f), cv, fa.create_coloring(g), fa.ElasticityElement(E=E, nu=0.3),
                           precision=precision)
    # FP64 element matrices for the comparison below
    fa.assemble_global(fa.start_assemble(K, f), cv, fa.create_coloring(g), fa.ElasticityElement(E=E, nu=0.3))
    tip = int(np.nonzero(np.all(np.isclose(g.nodes, [2.0, 0.5]), axis=1))[0][0]) + 1
    c = int(np.nonzero(np.any(g.cells == tip, axis=1))[0][0])
    ydof = int(dh.cell_dofs[c, 2 * int(np.nonzero(g.cells[c] == tip)[0][0]) + 1])
    fa.add_point_load(K, f, [ydof], [-1.0])
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.Dirichlet("u", g.facetsets["left"], lambda x, t: [0.0, 0.0])).close()
    fa.apply(K, f, ch)
    grid, ocv, odh, OK, of = setup_vector("Quadrilateral", nel, (0.0, 0.0), (2.0, 1.0), 1, 2)
    assemble_elasticity(grid, ocv, odh, OK, of, E, 0.3)
    of[ydof - 1] += -1.0
    och = OC.ConstraintHandler(odh)
    och.add(OC.Dirichlet("u", grid.facetsets["left"], lambda x, t: [0.0, 0.0])).close()
    OC.apply(OK, of, och)
    u_ref = OC.solve(OK, of)
    out, again = _solves(K, f, u_ref)
    assert again.operator == "sell"         # default: the planes only when asked for


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_q2_hex_elasticity(precision):
    print("case elemop q2 hex", precision, flush=True)
    nel = (4, 3, 3)
    E = np.random.default_rng(7).uniform(0.5, 2.0, int(np.prod(nel)))
    g = fa.generate_grid("Hexahedron", nel, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    _distort(g.nodes, 0.05, 8)
    ip = fa.Lagrange("RefHexahedron", 2) ** 3
    cv = fa.CellValues(fa.QuadratureRule("RefHexahedron", 3), ip)
    dh = fa.DofHandler(g).add("u", ip).close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh)
    f = fa.allocate_vector(K)
    fa.assemble_global(fa.start_assemble(K, f), cv, fa.create_coloring(g), fa.ElasticityElement(E=E, nu=0.3),
                       precision=precision)
    fv = fa.FacetValues(fa.FacetQuadratureRule("RefHexahedron", 3), ip)
    fa.apply_traction(K, f, fv, g.facetsets["right"], (0.0, 0.0, -1.0))
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.Dirichlet("u", g.facetsets["left"], lambda x, t: [0.0, 0.0, 0.0])).close()
    fa.apply(K, f, ch)
    # reference: the direct solve of the device's own constrained K (the FP32 element path's
    # K differs from the oracle's by its rounding; the operator must reproduce that K)
    import scipy.sparse.linalg as spla
    u_ref = spla.spsolve(K.to_scipy().tocsc(), f.to_numpy())
    out, again = _solves(K, f, u_ref)
    assert again.operator == "elements"     # default for the Q2 node blocks, both solvers
    assert fa.solve(K, f, mode="cg", rtol=1e-12).operator == "elements"
    if precision == "fp64":
        grid, ocv, odh, OK, of = setup_vector("Hexahedron", nel, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 2, 3)
        grid.nodes[:] = g.nodes
        assemble_elasticity(grid, ocv, odh, OK, of, E, 0.3)
        from oracle.fem import FacetValues as OFacetValues
        from oracle.problems import assemble_traction
        from oracle.refel import FacetQuadratureRule as OFQR, Lagrange as OLagrange
        ofv = OFacetValues(OFQR("RefHexahedron", 3), OLagrange("RefHexahedron", 2) ** 3)
        assemble_traction(grid, ofv, odh, of, grid.facetsets["right"], (0.0, 0.0, -1.0))
        och = OC.ConstraintHandler(odh)
        och.add(OC.Dirichlet("u", grid.facetsets["left"], lambda x, t: [0.0, 0.0, 0.0])).close()
        OC.apply(OK, of, och)
        assert _rel(out["cg"].u, OC.solve(OK, of)) < 1e-8
