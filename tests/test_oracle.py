"""Pins the CPU oracle against the reference's own known answers and closed forms."""
import json
import math
import os

import numpy as np
import pytest

from oracle import constraints as C
from oracle.coloring import create_coloring
from oracle.dofs import DofHandler
from oracle.fem import CSC, CellValues, DetJError, poisson_element
from oracle.grid import Grid, generate_grid
from oracle.problems import heat_equation
from oracle.refel import Lagrange, QuadratureRule, gauss_legendre
from oracle.sparsity import create_pattern, create_pattern_fast

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")


def kats():
    with open(GOLD) as fh:
        return json.load(fh)


def test_heat_equation_norm_kat():
    """docs/src/literate-tutorials/heat_equation.jl:234: norm(u) ≈ 3.307743912641305."""
    u, _ = heat_equation(20)
    ref = kats()["heat_equation_norm_u"]["value"]
    # the reference's ≈ is rtol = sqrt(eps) ~ 1.5e-8; the restatement is far tighter
    assert abs(np.linalg.norm(u) - ref) / ref < 1e-13


def test_gauss_closed_forms():
    p, w = gauss_legendre(2)
    assert np.allclose(p, [-1 / math.sqrt(3), 1 / math.sqrt(3)], rtol=0, atol=4e-16)
    assert np.allclose(w, [1.0, 1.0], rtol=0, atol=4e-16)
    p, w = gauss_legendre(3)
    assert np.allclose(p, [-math.sqrt(0.6), 0.0, math.sqrt(0.6)], rtol=0, atol=4e-16)
    assert np.allclose(w, [5 / 9, 8 / 9, 5 / 9], rtol=0, atol=4e-16)


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_gauss_exactness(n):
    """test_quadrules.jl:22-39: exact for x^a y^b z^c with a, b, c <= 2n-1 on the reference cube."""
    qr = QuadratureRule("RefHexahedron", n)

    def i1(d):
        return 2.0 / (d + 1) if d % 2 == 0 else 0.0
    for a in range(2 * n):
        for b in range(2 * n):
            for c in (0, 2 * n - 1):
                val = sum(w * p[0] ** a * p[1] ** b * p[2] ** c for p, w in zip(qr.points, qr.weights))
                assert abs(val - i1(a) * i1(b) * i1(c)) < 1e-13


@pytest.mark.parametrize("shape,order", [("RefQuadrilateral", 1), ("RefQuadrilateral", 2),
                                         ("RefHexahedron", 1), ("RefHexahedron", 2),
                                         ("RefLine", 1), ("RefLine", 2)])
def test_shape_functions(shape, order):
    """Kronecker delta at reference nodes (test_interpolations.jl:93), partition of unity, FD gradients."""
    ip = Lagrange(shape, order)
    coords = ip.reference_coordinates()
    nb = ip.nbasefunctions
    for a in range(nb):
        for b in range(nb):
            v = ip.shape_value(list(coords[b]), a + 1)
            assert abs(v - (1.0 if a == b else 0.0)) < 1e-15
    rng = np.random.default_rng(0)
    for _ in range(5):
        xi = rng.uniform(-1, 1, size=ip.rdim)
        assert abs(sum(ip.shape_value(list(xi), i + 1) for i in range(nb)) - 1) < 1e-14
        for i in range(nb):
            g, _ = ip.shape_gradient_and_value(list(xi), i + 1)
            h = 1e-6
            for d in range(ip.rdim):
                xp, xm = xi.copy(), xi.copy()
                xp[d] += h
                xm[d] -= h
                fd = (ip.shape_value(list(xp), i + 1) - ip.shape_value(list(xm), i + 1)) / (2 * h)
                assert abs(fd - g[d]) < 1e-8


def test_generate_grid_closed_form():
    """test_grid_generators.jl:22-40 style: corners, spacing, facetsets."""
    g = generate_grid("Hexahedron", (4, 3, 2), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    assert g.nnodes == 5 * 4 * 3 and g.ncells == 24
    assert np.allclose(g.nodes[0], [0, 0, 0]) and np.allclose(g.nodes[-1], [1, 1, 1])
    ii, jj, kk = np.meshgrid(np.arange(5), np.arange(4), np.arange(3), indexing="ij")
    exp = np.stack([ii.T.ravel() / 4, jj.T.ravel() / 3, kk.T.ravel() / 2], axis=1)
    assert np.allclose(g.nodes, exp, atol=1e-15)
    assert g.cells[0] == (1, 2, 7, 6, 21, 22, 27, 26)
    assert [c for c, _ in g.facetsets["left"]] == [1, 5, 9, 13, 17, 21]
    q = generate_grid("Quadrilateral", (2, 2))
    assert np.allclose(q.nodes[4], [0.0, 0.0])
    assert q.facetsets["right"] == [(2, 2), (4, 2)]


def test_celldofs_kats():
    """test_dofs.jl:59-89 (Line, vector Q1/Q2) and :95-127 (quad shell, 2 fields)."""
    K = kats()["celldofs"]
    nodes = np.array([[0.0, 0.0], [1.0, 1.0], [2.0, 0.0]])
    g = Grid("Line", [(1, 2), (2, 3)], nodes)
    dh = DofHandler(g).add("x", Lagrange("RefLine", 1) ** 2).close()
    assert [dh.celldofs(1), dh.celldofs(2)] == K["line_q1_vec2"]
    dh = DofHandler(g).add("x", Lagrange("RefLine", 2) ** 2).close()
    assert [dh.celldofs(1), dh.celldofs(2)] == K["line_q2_vec2"]
    dh = DofHandler(g).add("u", Lagrange("RefLine", 2) ** 3).add("t", Lagrange("RefLine", 2) ** 3).close()
    assert [dh.celldofs(1), dh.celldofs(2)] == K["line_q2_vec3_two_fields"]
    nodes = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [2, 0, 0], [2, 2, 0.0]])
    g = Grid("Quadrilateral", [(1, 2, 3, 4), (2, 5, 6, 3)], nodes)
    dh = DofHandler(g).add("u", Lagrange("RefQuadrilateral", 1) ** 3).add("t", Lagrange("RefQuadrilateral", 1) ** 3).close()
    assert [dh.celldofs(1), dh.celldofs(2)] == K["quad_shell_q1_vec3_two_fields"]
    dh = DofHandler(g).add("u", Lagrange("RefQuadrilateral", 2)).add("t", Lagrange("RefQuadrilateral", 2)).close()
    assert [dh.celldofs(1), dh.celldofs(2)] == K["quad_shell_q2_two_fields"]


def test_testdhch_kat():
    """test_dofs.jl:235-258: celldofs and prescribed dofs (the affine dof 13 is out of scope)."""
    K = kats()["testdhch"]
    grid = generate_grid("Quadrilateral", (2, 1))
    dh = DofHandler(grid).add("v", Lagrange("RefQuadrilateral", 1) ** 2).add("s", Lagrange("RefQuadrilateral", 1)).close()
    assert [dh.celldofs(1), dh.celldofs(2)] == K["celldofs"]
    ch = C.ConstraintHandler(dh)
    ch.add(C.Dirichlet("v", grid.facetsets["left"], lambda x, t: 0.0, [2]))
    ch.add(C.Dirichlet("s", grid.facetsets["left"], lambda x, t: 0.0))
    ch.close()
    assert list(ch.prescribed_dofs) == K["prescribed_dofs_without_affine"]


def test_q1_dofs_first_visit():
    """Q1 dof numbering is first-visit order of nodes, not node order (DofHandler.jl:715-744)."""
    g = generate_grid("Hexahedron", (2, 2, 2))
    dh = DofHandler(g).add("u", Lagrange("RefHexahedron", 1)).close()
    assert dh.celldofs(1) == [1, 2, 3, 4, 5, 6, 7, 8]
    assert dh.celldofs(2) == [2, 9, 10, 3, 6, 11, 12, 7]
    assert dh.ndofs == 27
    seen = sorted(set(dh.cell_dofs))
    assert seen == list(range(1, 28))


def test_q2_vector_dofs_counts():
    g = generate_grid("Hexahedron", (2, 2, 2))
    dh = DofHandler(g).add("u", Lagrange("RefHexahedron", 2) ** 3).close()
    assert dh.ndofs_per_cell == 81 and dh.ndofs == 5 ** 3 * 3
    assert sorted(set(dh.cell_dofs)) == list(range(1, 376))
    # vector components of one node are consecutive
    cd = dh.celldofs(1)
    for a in range(27):
        assert cd[3 * a + 1] == cd[3 * a] + 1 and cd[3 * a + 2] == cd[3 * a] + 2


def test_pattern_bruteforce_vs_vectorised():
    g = generate_grid("Hexahedron", (3, 2, 2))
    dh = DofHandler(g).add("u", Lagrange("RefHexahedron", 1)).close()
    cp, rv = create_pattern(dh)
    cd = np.array(dh.cell_dofs).reshape(g.ncells, -1)
    cp2, rv2 = create_pattern_fast(cd, dh.ndofs)
    assert np.array_equal(cp, cp2) and np.array_equal(rv, rv2)


@pytest.mark.parametrize("alg", ["greedy", "workstream"])
def test_coloring_validity(alg):
    """test_grid_dofhandler_vtk.jl:719-760: no two cells of one color share a node; all cells covered."""
    for g in (generate_grid("Quadrilateral", (7, 5)), generate_grid("Hexahedron", (4, 3, 3))):
        colors = create_coloring(g, alg)
        allc = sorted(c for col in colors for c in col)
        assert allc == list(range(1, g.ncells + 1))
        for col in colors:
            assert col == sorted(col)
            nodes = [n for c in col for n in g.cells[c - 1]]
            assert len(nodes) == len(set(nodes))
    if alg == "greedy":
        assert len(create_coloring(generate_grid("Hexahedron", (4, 4, 4)), alg)) == 8
        assert len(create_coloring(generate_grid("Quadrilateral", (6, 6)), alg)) == 4


def test_cellvalues_properties():
    """test_cellvalues.jl:60-175: affine field reproduced, Σ detJdV = volume."""
    rng = np.random.default_rng(3)
    for shape, n in (("RefQuadrilateral", 4), ("RefHexahedron", 8)):
        ip = Lagrange(shape, 1)
        cv = CellValues(QuadratureRule(shape, 2), ip)
        ref = np.array(ip.reference_coordinates())
        dim = ref.shape[1]
        A = np.eye(dim) * 0.7 + rng.uniform(-0.1, 0.1, (dim, dim))
        x = ref @ A.T + 0.3
        dNdx, dV = cv.reinit(x)
        assert abs(dV.sum() - 2 ** dim * abs(np.linalg.det(A))) < 1e-13
        grad = rng.normal(size=dim)
        u = x @ grad
        for q in range(cv.nqp):
            g = sum(u[i] * dNdx[i, q] for i in range(n))
            assert np.allclose(g, grad, atol=1e-12)
        dNdx, dV = cv.reinit(ref)
        assert abs(dV.sum() - 2 ** dim) < 1e-14


def test_detj_error():
    ip = Lagrange("RefQuadrilateral", 1)
    cv = CellValues(QuadratureRule("RefQuadrilateral", 2), ip)
    x = np.array(ip.reference_coordinates())[[1, 0, 3, 2]]  # mirrored -> det < 0
    with pytest.raises(DetJError):
        cv.reinit(x)


def test_poisson_element_closed_form():
    """Axis-aligned box: Ke = k (K1x⊗M1y⊗M1z + ...), 2-point Gauss integrates Q1 products exactly."""
    hx, hy, hz, k = 0.5, 0.25, 2.0, 1.7
    ip = Lagrange("RefHexahedron", 1)
    cv = CellValues(QuadratureRule("RefHexahedron", 2), ip)
    ref = (np.array(ip.reference_coordinates()) + 1) / 2 * [hx, hy, hz]
    Ke, fe = poisson_element(cv, ref, k, 1.0)

    def K1(h):
        return np.array([[1, -1], [-1, 1]]) / h

    def M1(h):
        return np.array([[2, 1], [1, 2]]) * h / 6
    # local node a -> (ia, ja, ka) in Ferrite's hex order
    idx = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    ex = np.zeros((8, 8))
    for a, (ia, ja, ka) in enumerate(idx):
        for b, (ib, jb, kb) in enumerate(idx):
            ex[a, b] = k * (K1(hx)[ia, ib] * M1(hy)[ja, jb] * M1(hz)[ka, kb]
                            + M1(hx)[ia, ib] * K1(hy)[ja, jb] * M1(hz)[ka, kb]
                            + M1(hx)[ia, ib] * M1(hy)[ja, jb] * K1(hz)[ka, kb])
    assert np.allclose(Ke, ex, rtol=1e-13, atol=1e-14)
    assert np.allclose(Ke.sum(axis=1), 0, atol=1e-13)
    assert np.allclose(fe, hx * hy * hz / 8, rtol=1e-13)


def test_apply_inhomogeneous_matches_elimination():
    """apply! with g != 0: the solve reproduces the block-eliminated system (ConstraintHandler.jl:710-747)."""
    _, d = heat_equation(6)
    grid, dh = d["grid"], d["dh"]
    from oracle.problems import assemble_poisson
    cp, rv = create_pattern(dh)
    K = CSC(dh.ndofs, cp, rv)
    f = np.zeros(dh.ndofs)
    cv = d["cv"]
    assemble_poisson(grid, cv, dh, K, f)
    A = K.to_scipy().toarray()
    b = f.copy()
    ch = C.ConstraintHandler(dh)
    ch.add(C.Dirichlet("u", grid.facetsets["left"] + grid.facetsets["top"], lambda x, t: 1.0 + x[0] * x[1])).close()
    C.apply(K, f, ch)
    u = C.solve(K, f)
    pres = ch.prescribed_dofs - 1
    free = np.setdiff1d(np.arange(dh.ndofs), pres)
    g = ch.inhomogeneities
    uf = np.linalg.solve(A[np.ix_(free, free)], b[free] - A[np.ix_(free, pres)] @ g)
    assert np.allclose(u[pres], g, atol=1e-14)
    assert np.allclose(u[free], uf, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("ct,nel,distort", [("Quadrilateral", (5, 4), 0.0), ("Hexahedron", (4, 3, 3), 0.0),
                                            ("Hexahedron", (3, 3, 2), 0.06)])
def test_vectorised_poisson_matches_serial_loop(ct, nel, distort):
    """oracle.fem_vec (batched over cells) against the per-cell serial loop of oracle.problems:
    nzval and f to ≤1e-14 (different order of the batched products only), same pattern."""
    from oracle.fem_vec import assemble_poisson_vec
    from oracle.problems import assemble_poisson, poisson_conductivity, setup_poisson
    grid, cv, dh, K, f, _ = setup_poisson(ct, nel)
    if distort:
        X = grid.nodes
        inner = np.all((X > -1 + 1e-12) & (X < 1 - 1e-12), axis=1)
        X[inner] += np.random.default_rng(2).uniform(-distort, distort, (int(inner.sum()), X.shape[1]))
    k = poisson_conductivity(nel, seed=4)
    assemble_poisson(grid, cv, dh, K, f, k_cell=k)
    cd = np.array(dh.cell_dofs).reshape(grid.ncells, -1)
    nzv, fv = assemble_poisson_vec(grid, cv, cd, K.colptr, K.rowval, k_cell=k)
    assert np.linalg.norm(nzv - K.nzval) <= 1e-14 * np.linalg.norm(K.nzval)
    assert np.abs(nzv - K.nzval).max() <= 1e-14 * np.abs(K.nzval).max()
    assert np.linalg.norm(fv - f) <= 1e-14 * np.linalg.norm(f)
