"""Host-side setup of the product vs the oracle (CPU only): bit-exact grid, dofs, pattern, colors, constraints."""
import numpy as np
import pytest

import ferrite_amd as fa
from oracle import constraints as OC
from oracle.coloring import create_coloring as o_coloring
from oracle.dofs import DofHandler as ODofHandler
from oracle.grid import generate_grid as o_generate_grid
from oracle.refel import Lagrange as OLagrange
from oracle.refel import QuadratureRule as OQuadratureRule
from oracle.fem import CellValues as OCellValues
from oracle.sparsity import create_pattern

CASES = [
    ("Quadrilateral", (7, 5), 1, 1, (-1.0, -1.0), (1.0, 1.0)),
    ("Quadrilateral", (6, 4), 1, 2, (0.0, 0.0), (2.0, 1.0)),
    ("Quadrilateral", (3, 3), 2, 1, (0.1, 0.2), (0.7, 0.9)),
    ("Hexahedron", (3, 4, 2), 1, 1, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)),
    ("Hexahedron", (3, 2, 2), 2, 3, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)),
    ("Hexahedron", (2, 2, 3), 1, 3, (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)),
]


def _pair(ct, nel, order, vdim, left, right):
    g = fa.generate_grid(ct, nel, left, right)
    og = o_generate_grid(ct, nel, left, right)
    ip = fa.Lagrange(g.refshape, order)
    oip = OLagrange(og.refshape, order)
    if vdim > 1:
        ip, oip = ip ** vdim, oip ** vdim
    return g, og, ip, oip


@pytest.mark.parametrize("case", CASES)
def test_grid_dofs_pattern_bitexact(case):
    g, og, ip, oip = _pair(*case)
    assert np.array_equal(g.nodes, og.nodes)          # bitwise: same blend, same operation order
    assert np.array_equal(g.cells, np.array(og.cells))
    assert g.facetsets == og.facetsets
    dh = fa.DofHandler(g).add("u", ip).close()
    odh = ODofHandler(og).add("u", oip).close()
    assert dh.ndofs == odh.ndofs
    assert np.array_equal(dh.cell_dofs.ravel(), np.array(odh.cell_dofs))
    cp, rv = fa.create_sparsity_pattern(dh)
    ocp, orv = create_pattern(odh)
    assert np.array_equal(cp, ocp) and np.array_equal(rv, orv)


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("alg", ["greedy", "workstream"])
def test_coloring_matches_reference_algorithm(case, alg):
    g, og, _, _ = _pair(*case)
    c1 = fa.create_coloring(g, alg)
    c2 = o_coloring(og, alg)
    assert len(c1) == len(c2)
    for a, b in zip(c1, c2):
        assert np.array_equal(a, np.array(b))


def test_two_field_dofs_kat():
    """test_dofs.jl:235-258 via the product's native distribution."""
    g = fa.generate_grid("Quadrilateral", (2, 1))
    dh = fa.DofHandler(g).add("v", fa.Lagrange("RefQuadrilateral", 1) ** 2).add(
        "s", fa.Lagrange("RefQuadrilateral", 1)).close()
    assert dh.celldofs(1).tolist() == list(range(1, 13))
    assert dh.celldofs(2).tolist() == [3, 4, 13, 14, 15, 16, 5, 6, 10, 17, 18, 11]
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.Dirichlet("v", g.facetsets["left"], lambda x, t: 0.0, [2]))
    ch.add(fa.Dirichlet("s", g.facetsets["left"], lambda x, t: 0.0))
    ch.close()
    assert ch.prescribed_dofs.tolist() == [2, 8, 9, 12]


@pytest.mark.parametrize("case", CASES)
def test_prescribed_dofs_and_inhomogeneities(case):
    g, og, ip, oip = _pair(*case)
    dh = fa.DofHandler(g).add("u", ip).close()
    odh = ODofHandler(og).add("u", oip).close()
    names = sorted(g.facetsets)[:2]
    gf = lambda x, t: 1.0 + 2.0 * x[0] - x[1]  # noqa: E731
    ch = fa.ConstraintHandler(dh)
    och = OC.ConstraintHandler(odh)
    for nm in names:
        ch.add(fa.Dirichlet("u", g.facetsets[nm], gf))
        och.add(OC.Dirichlet("u", og.facetsets[nm], gf))
    ch.close()
    och.close()
    assert np.array_equal(ch.prescribed_dofs, och.prescribed_dofs)
    assert np.allclose(ch.inhomogeneities, och.inhomogeneities, rtol=0, atol=1e-15)


@pytest.mark.parametrize("shape,order,q", [("RefQuadrilateral", 1, 2), ("RefHexahedron", 1, 2),
                                           ("RefHexahedron", 2, 3), ("RefQuadrilateral", 2, 3)])
def test_cellvalues_tables_match(shape, order, q):
    cv = fa.CellValues(fa.QuadratureRule(shape, q), fa.Lagrange(shape, order))
    ocv = OCellValues(OQuadratureRule(shape, q), OLagrange(shape, order))
    assert np.allclose(cv.N, ocv.N.T, rtol=0, atol=4e-16)
    assert np.allclose(cv.dNdxi, ocv.dNdxi.transpose(1, 0, 2), rtol=0, atol=8e-16)
    assert np.allclose(cv.dMdxi, ocv.dMdxi.transpose(1, 0, 2), rtol=0, atol=8e-16)
    assert np.allclose(cv.weights, ocv.qr.weights, rtol=0, atol=2e-15)


def test_c2_scale_setup_properties():
    """Hex 128^3 Q1: sizes from the 3D Poisson configuration; size-independent pattern invariants."""
    g = fa.generate_grid("Hexahedron", (128, 128, 128), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    dh = fa.DofHandler(g).add("u", fa.Lagrange("RefHexahedron", 1)).close()
    assert g.ncells == 2097152 and dh.ndofs == 2146689
    cp, rv = fa.create_sparsity_pattern(dh)
    assert len(rv) == 57066625 == 385 ** 3
    counts = np.diff(cp)
    assert counts.min() == 8 and counts.max() == 27
    colors = fa.create_coloring(g)
    allc = np.sort(np.concatenate(colors))
    assert np.array_equal(allc, np.arange(1, g.ncells + 1))
    # no two cells of one colour share a node (checked on a sample of colours)
    for col in colors[:3]:
        nodes = g.cells[col - 1].ravel()
        assert len(np.unique(nodes)) == nodes.size


def test_batch_problem_sizes():
    """The 64-problem batch: 2D/3D mix, cell counts log-uniform in [2e3, 2e5], near-square/cube grids."""
    from ferrite_amd.batch import batch_sizes
    sz = batch_sizes(64, 0)
    cells = np.array([np.prod(n) for _, n in sz])
    assert len(sz) == 64
    assert 0.8 * 2e3 <= cells.min() and cells.max() <= 1.2 * 2e5
    assert {ct for ct, _ in sz} == {"Quadrilateral", "Hexahedron"}
    for ct, n in sz:
        assert len(n) == (2 if ct == "Quadrilateral" else 3)
        assert max(n) <= 1.2 * min(n) + 1
    assert batch_sizes(64, 0) == sz and batch_sizes(64, 1) != sz


def test_batch_problem_setup_matches_oracle():
    """A batch problem's dofs, pattern and prescribed dofs equal the oracle's setup_poisson."""
    from ferrite_amd.batch import poisson_problem
    from ferrite_amd.dofs import create_sparsity_pattern
    from oracle.problems import setup_poisson
    for ct, nel in (("Quadrilateral", (7, 5)), ("Hexahedron", (3, 4, 2))):
        g, dh, colors, ch, _ = poisson_problem(ct, nel, np.ones(int(np.prod(nel))))
        og, ocv, odh, oK, of, och = setup_poisson(ct, nel)
        assert np.array_equal(dh.cell_dofs.ravel(), np.asarray(odh.cell_dofs).ravel())
        cp, rv = create_sparsity_pattern(dh)
        assert np.array_equal(cp, oK.colptr) and np.array_equal(rv, oK.rowval)
        assert np.array_equal(ch.prescribed_dofs, np.asarray(och.prescribed_dofs))
        assert sorted(np.concatenate(colors).tolist()) == list(range(1, g.ncells + 1))


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_q2_hex_dofs_single_valued_geometry(impl):
    """Q2 hex dof numbering pinned by geometry (no reference table of hex Q2 celldofs exists).

    On a grid with moved interior nodes, every local dof of every cell is placed at its
    physical point x(ξ_a) (Q1 geometry at the Q2 reference coordinates).  A global dof must
    be the same point from every cell that holds it, distinct dofs distinct points, and each
    cell's edge / face / volume dofs must sit at the midpoint of Ferrite's edge (grid.jl:218-
    229) / the centre of its face / the cell centre, which the trilinear map preserves.  A
    misread edge or face key (dofs shared between the wrong cells) or a local entity order
    that disagrees with the reference coordinates fails here."""
    from oracle.refel import REF_EDGES, REF_FACES
    nel = (3, 2, 2)
    og = o_generate_grid("Hexahedron", nel, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    X = og.nodes.copy()
    inner = np.all((X > 1e-12) & (X < 1 - 1e-12), axis=1)
    X[inner] += np.random.default_rng(3).uniform(-0.05, 0.05, (int(inner.sum()), 3))
    oip = OLagrange("RefHexahedron", 2) ** 3
    if impl == "oracle":
        cd = np.array(ODofHandler(og).add("u", oip).close().cell_dofs).reshape(og.ncells, 81)
        ndofs = int(cd.max())
    else:
        g = fa.generate_grid("Hexahedron", nel, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
        dh = fa.DofHandler(g).add("u", fa.Lagrange("RefHexahedron", 2) ** 3).close()
        cd, ndofs = np.asarray(dh.cell_dofs), dh.ndofs
    geo = OLagrange("RefHexahedron", 1)
    refc = OLagrange("RefHexahedron", 2).reference_coordinates()
    M = np.array([[geo.shape_value(xi, j) for j in range(1, 9)] for xi in refc])   # (27, 8)
    pos = {}
    for c in range(og.ncells):
        X8 = X[np.array(og.cells[c]) - 1]
        P = M @ X8
        for a in range(27):
            d = int(cd[c, 3 * a])
            assert (d - 1) % 3 == 0 and cd[c, 3 * a + 1] == d + 1 and cd[c, 3 * a + 2] == d + 2
            if d in pos:
                assert np.abs(pos[d] - P[a]).max() <= 1e-14
            else:
                pos[d] = P[a]
        for k, (v1, v2) in enumerate(REF_EDGES["RefHexahedron"]):
            assert np.abs(P[8 + k] - 0.5 * (X8[v1 - 1] + X8[v2 - 1])).max() <= 1e-14
        for k, face in enumerate(REF_FACES["RefHexahedron"]):
            assert np.abs(P[20 + k] - X8[np.array(face) - 1].mean(axis=0)).max() <= 1e-14
        assert np.abs(P[26] - X8.mean(axis=0)).max() <= 1e-14
    assert len(pos) * 3 == ndofs
    pts = np.array(list(pos.values()))
    assert len(np.unique(np.round(pts, 10), axis=0)) == len(pos)     # distinct dofs, distinct points
