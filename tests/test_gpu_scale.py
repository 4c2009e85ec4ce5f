"""Parity at scale: scalar Q1 hex Poisson on 40³ cells (64,000 cells, 68,921 dofs) through the
C ABI against the oracle.

The oracle side is the vectorised restatement of the serial loop (oracle/fem_vec.py, pinned
against the per-cell loop in tests/test_oracle.py), the brute-force pattern and the oracle's
ConstraintHandler / apply!.  u is compared with an oracle Jacobi-CG solve to a relative
residual of 1e-14 in the scaled system, ≤1e-12 unscaled (SuperLU needs minutes at this size;
cond(K) ≈ 1e3 here, so the iterate is within ~1e-9 of the direct solution, below the 1e-8
bar).

Bars: pattern and prescribed dofs bit-exact; K.nzval and f within 1e-12 (normwise, and
entrywise against 1e-12·max|K|) before and after apply!; u within 1e-8 of the oracle for the
FP64 and the mixed solver.  The distorted case moves interior nodes by up to 30% of a cell
(trilinear geometry form, lattice-plane solver operator); the straight one is the box of the
benchmark (affine geometry form, cell-form solver operator).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import ferrite_amd as fa
from oracle import constraints as OC
from oracle.dofs import DofHandler as ODofHandler
from oracle.fem import CSC, CellValues as OCellValues
from oracle.fem_vec import assemble_poisson_vec
from oracle.grid import generate_grid as o_generate_grid
from oracle.problems import poisson_conductivity
from oracle.refel import Lagrange as OLagrange, QuadratureRule as OQuadratureRule
from oracle.sparsity import create_pattern_fast

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _oracle_cg(K, f, rtol=1e-14):
    A = K.to_scipy().tocsr()
    d = 1.0 / np.sqrt(A.diagonal())
    S = sp.diags(d)
    As = (S @ A @ S).tocsr()
    it = {"n": 0}
    y, info = spla.cg(As, d * f, rtol=rtol, atol=0.0, maxiter=20000,
                      callback=lambda _: it.__setitem__("n", it["n"] + 1))
    u = d * y
    assert info == 0 and np.linalg.norm(f - A @ u) <= 1e-12 * np.linalg.norm(f)
    return u


@pytest.mark.parametrize("distort", [0.0, 0.3])
def test_q1_hex_40_poisson_against_oracle(distort):
    n = 40
    nel = (n, n, n)
    print("case q1 hex 40^3 poisson, distort", distort, flush=True)
    k = poisson_conductivity(nel, seed=3)
    g = fa.generate_grid("Hexahedron", nel, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    og = o_generate_grid("Hexahedron", nel, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    if distort:
        X = g.nodes.copy()
        inner = np.all((X > 1e-12) & (X < 1 - 1e-12), axis=1)
        X[inner] += np.random.default_rng(9).uniform(-distort / n, distort / n, (int(inner.sum()), 3))
        g.nodes[:] = X
        og.nodes[:] = X
    ip = fa.Lagrange("RefHexahedron", 1)
    cv = fa.CellValues(fa.QuadratureRule("RefHexahedron", 2), ip)
    dh = fa.DofHandler(g).add("u", ip).close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh)
    f = fa.allocate_vector(K)
    fa.assemble_global(fa.start_assemble(K, f), cv, fa.create_coloring(g), fa.PoissonElement(k=k))
    bc = lambda x, t: 0.0  # noqa: E731
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.Dirichlet("u", g.facetsets["left"] + g.facetsets["right"], bc)).close()

    oip = OLagrange("RefHexahedron", 1)
    ocv = OCellValues(OQuadratureRule("RefHexahedron", 2), oip)
    odh = ODofHandler(og).add("u", oip).close()
    cd = np.array(odh.cell_dofs).reshape(og.ncells, -1)
    cp, rv = create_pattern_fast(cd, odh.ndofs)
    assert np.array_equal(K.colptr, cp) and np.array_equal(K.rowval, rv)
    onz, of = assemble_poisson_vec(og, ocv, cd, cp, rv, k_cell=k)
    nz, fv = K.nzval(), f.to_numpy()
    assert _rel(nz, onz) <= 1e-12 and np.abs(nz - onz).max() <= 1e-12 * np.abs(onz).max()
    assert _rel(fv, of) <= 1e-12

    och = OC.ConstraintHandler(odh)
    och.add(OC.Dirichlet("u", og.facetsets["left"] + og.facetsets["right"], bc)).close()
    assert np.array_equal(ch.prescribed_dofs, och.prescribed_dofs)
    OK = CSC(odh.ndofs, cp, rv, onz.copy())
    OC.apply(OK, of, och)
    fa.apply(K, f, ch)
    nz, fv = K.nzval(), f.to_numpy()
    assert _rel(nz, OK.nzval) <= 1e-12 and np.abs(nz - OK.nzval).max() <= 1e-12 * np.abs(OK.nzval).max()
    assert _rel(fv, of) <= 1e-12

    u = _oracle_cg(OK, of)
    for mode in ("cg", "mixed"):
        res = fa.solve(K, f, mode=mode, rtol=1e-12)
        print(f"  {mode}: operator {res.operator}, {res.iterations} its, u err {_rel(res.u, u):.2e}", flush=True)
        assert _rel(res.u, u) <= 1e-8
        if mode == "mixed":
            assert res.operator == ("lattice" if distort else "cells")
