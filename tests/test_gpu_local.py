"""apply_local! during the device assembly (assemble_global(..., ch=ch)) against the oracle's
restatement (oracle.constraints.apply_local, pinned by the reference's local-application KATs
in tests/test_oracle_local.py) and against the properties the reference's own test checks
(test/test_constraints.jl:1566-1594): the free block of K and f equal the standard path's
(assemble, then apply!), the prescribed block is diagonal with f / diag = g, and the solution
is the same.

Bars: K and f within 1e-12 of the oracle (FP64 element math; 2e-6 with FP32 element math),
the free blocks within 1e-12 of the standard path, f / diag = g to 1e-14, u within 1e-10 of
the standard path's and 1e-8 of the oracle's direct solve.
"""
import numpy as np
import pytest

import ferrite_amd as fa
from oracle import constraints as OC
from oracle.fem import CSC, CellValues, assemble, poisson_element, start_assemble
from oracle.dofs import DofHandler as ODofHandler
from oracle.grid import generate_grid as ogenerate_grid
from oracle.refel import Lagrange as OLagrange
from oracle.refel import QuadratureRule as OQuadratureRule
from oracle.sparsity import create_pattern

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _g(x, t):
    return 1.0 + x[1]


def _device(ct, nel, k, ch_sets, local, apply_zero=False, precision="fp64"):
    dim = len(nel)
    g = fa.generate_grid(ct, nel, (0.0,) * dim, (1.0,) * dim)
    ip = fa.Lagrange(g.refshape, 1)
    cv = fa.CellValues(fa.QuadratureRule(g.refshape, 2), ip)
    dh = fa.DofHandler(g).add("u", ip).close()
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.Dirichlet("u", g.facetsets[ch_sets[0]], _g))
    ch.add(fa.Dirichlet("u", g.facetsets[ch_sets[1]], lambda x, t: 0.0))
    ch.close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh)
    f = fa.allocate_vector(K)
    A = fa.start_assemble(K, f)
    fa.assemble_global(A, cv, fa.create_coloring(g), fa.PoissonElement(k=k, source=1.0), precision=precision,
                       ch=ch if local else None, apply_zero=apply_zero)
    if not local:
        fa.apply(K, f, ch, applyzero=apply_zero)
    return g, dh, ch, K, f


def _oracle(ct, nel, k, ch_sets, apply_zero):
    dim = len(nel)
    g = ogenerate_grid(ct, nel, (0.0,) * dim, (1.0,) * dim)
    ip = OLagrange(g.refshape, 1)
    cv = CellValues(OQuadratureRule(g.refshape, 2), ip)
    dh = ODofHandler(g).add("u", ip).close()
    ch = OC.ConstraintHandler(dh)
    ch.add(OC.Dirichlet("u", g.facetsets[ch_sets[0]], _g))
    ch.add(OC.Dirichlet("u", g.facetsets[ch_sets[1]], lambda x, t: 0.0))
    ch.close()
    cp, rv = create_pattern(dh)
    K, f = CSC(dh.ndofs, cp, rv), np.zeros(dh.ndofs)
    A = start_assemble(K, f)
    for c in range(1, g.ncells + 1):
        Ke, fe = poisson_element(cv, g.getcoordinates(c), float(k[c - 1]), 1.0)
        OC.apply_local(Ke, fe, dh.celldofs(c), ch, applyzero=apply_zero)
        assemble(A, dh.celldofs(c), Ke, fe)
    return K, f, ch


CASES = [("Quadrilateral", (12, 10), ("left", "right")), ("Hexahedron", (6, 5, 4), ("left", "right"))]


@pytest.mark.parametrize("ct,nel,sets", CASES)
@pytest.mark.parametrize("apply_zero", [False, True])
def test_apply_local_against_oracle_and_standard(ct, nel, sets, apply_zero):
    print("case apply_local", ct, nel, apply_zero, flush=True)
    k = np.random.default_rng(7).uniform(0.5, 2.0, int(np.prod(nel)))
    g, dh, ch, K, f = _device(ct, nel, k, sets, True, apply_zero)
    nz, fv = K.nzval(), f.to_numpy()
    OK, of, och = _oracle(ct, nel, k, sets, apply_zero)
    assert np.array_equal(K.colptr, OK.colptr) and np.array_equal(K.rowval, OK.rowval)
    assert np.array_equal(ch.prescribed_dofs, och.prescribed_dofs)
    assert np.max(np.abs(nz - OK.nzval)) <= 1e-12 * np.max(np.abs(OK.nzval))
    assert np.max(np.abs(fv - of)) <= 1e-12 * max(np.max(np.abs(of)), 1.0)
    # the reference test's properties, against the standard path on the device
    _, _, _, Ks, fs = _device(ct, nel, k, sets, False, apply_zero)
    A, S = K.to_scipy().toarray(), Ks.to_scipy().toarray()
    pd = ch.prescribed_dofs - 1
    free = np.setdiff1d(np.arange(dh.ndofs), pd)
    assert np.max(np.abs(A[np.ix_(free, free)] - S[np.ix_(free, free)])) <= 1e-12 * np.max(np.abs(S))
    fsv = fs.to_numpy()
    assert np.max(np.abs(fv[free] - fsv[free])) <= 1e-12 * max(np.max(np.abs(fsv)), 1.0)
    blk = A[np.ix_(pd, pd)]
    assert np.array_equal(blk, np.diag(np.diag(blk)))
    assert np.all(A[np.ix_(pd, free)] == 0.0) and np.all(A[np.ix_(free, pd)] == 0.0)
    expect = np.zeros(len(pd)) if apply_zero else ch.inhomogeneities
    assert np.max(np.abs(fv[pd] / np.diag(blk) - expect)) <= 1e-14 * max(np.max(np.abs(expect)), 1.0)
    ul = fa.solve(K, f, mode="cg", rtol=1e-12).u
    us = fa.solve(Ks, fs, mode="cg", rtol=1e-12).u
    assert _rel(ul, us) < 1e-10
    uo = OC.solve(OK, of)
    assert _rel(ul, uo) < 1e-8
    mixed = fa.solve(K, f, mode="mixed", rtol=1e-12)
    assert _rel(mixed.u, ul) < 1e-8


def test_apply_local_fp32_elements():
    print("case apply_local fp32 elements", flush=True)
    nel = (6, 5, 4)
    k = np.random.default_rng(8).uniform(0.5, 2.0, int(np.prod(nel)))
    _, _, _, K, f = _device("Hexahedron", nel, k, ("left", "right"), True, precision="mixed")
    OK, of, _ = _oracle("Hexahedron", nel, k, ("left", "right"), False)
    assert _rel(K.nzval(), OK.nzval) < 2e-6 and _rel(f.to_numpy(), of) < 2e-6


def test_apply_local_rejects_affine():
    g = fa.generate_grid("Quadrilateral", (3, 3), (0.0, 0.0), (1.0, 1.0))
    ip = fa.Lagrange(g.refshape, 1)
    cv = fa.CellValues(fa.QuadratureRule(g.refshape, 2), ip)
    dh = fa.DofHandler(g).add("u", ip).close()
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.AffineConstraint(1, [(5, 1.0)], 0.0)).close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh, ch)
    f = fa.allocate_vector(K)
    with pytest.raises(fa.FemArgumentError, match="global matrix"):
        fa.assemble_global(fa.start_assemble(K, f), cv, fa.create_coloring(g), fa.PoissonElement(), ch=ch)
