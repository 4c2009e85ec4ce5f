"""The oracle's C restatement (the timed CPU baselines) against the Python oracle (CPU).

  * scalar Poisson on Q1 quad, Q2 quad and Q1 hex distorted meshes: K and f within 1e-13 of the
    serial loop (oracle.problems / oracle.fem), and the fixed-table Q1 hex entry point agrees
    with the generic one;
  * small-strain elasticity on distorted Q2 hex and plane-strain Q1 quad meshes: K within 1e-12
    of the oracle's element routine and serial assembly;
  * the heat-equation known answer norm(u) = 3.307743912641305 (heat_equation.jl:234) through
    the timed heat run (assembly, apply!, CG);
  * the compliance adjoint on a small cantilever: J, dJ/dθ and dJ/dX against torch float64
    autograd through a dense assembly (oracle/adjoint_torch.py);
  * the error codes: det(J) <= 0 and a missing pattern entry.
"""
import os
import subprocess
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import cbaseline as CB
from oracle import constraints as OC
from oracle.coloring import create_coloring
from oracle.dofs import DofHandler
from oracle.fem import CSC, CellValues, assemble, elasticity_element, poisson_element, start_assemble
from oracle.grid import generate_grid
from oracle.refel import Lagrange, QuadratureRule
from oracle.sparsity import create_pattern


@pytest.fixture(scope="module", autouse=True)
def _built():
    if not os.path.exists(CB.LIB):
        subprocess.run(["make", "-C", CB.HERE], check=True, capture_output=True)


def _distort(grid, rng, amount):
    X = grid.nodes
    lo, hi = X.min(axis=0), X.max(axis=0)
    inner = np.all((X > lo + 1e-12) & (X < hi - 1e-12), axis=1)
    X[inner] += rng.uniform(-amount, amount, (int(inner.sum()), X.shape[1]))


def _setup(celltype, nel, order, vdim, rng, amount=0.0, left=None, right=None):
    g = generate_grid(celltype, nel, left, right)
    if amount:
        _distort(g, rng, amount)
    base = Lagrange(g.refshape, order)
    ip = base ** vdim if vdim > 1 else base
    dh = DofHandler(g).add("u", ip).close()
    cp, rv = create_pattern(dh)
    cd = np.array([dh.celldofs(c) for c in range(1, g.ncells + 1)], dtype=np.int64)
    colors = create_coloring(g)
    s = CB.System(g.nodes, np.array(g.cells), cd, cp, rv, colors)
    return g, ip, dh, cp, rv, cd, colors, s


def _rel_max(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("celltype,nel,order,qo", [("Quadrilateral", (7, 5), 1, 2), ("Quadrilateral", (4, 3), 2, 3),
                                                   ("Hexahedron", (3, 3, 2), 1, 2)])
def test_poisson_matches_serial_loop(celltype, nel, order, qo):
    rng = np.random.default_rng(3)
    g, ip, dh, cp, rv, cd, colors, s = _setup(celltype, nel, order, 1, rng, amount=0.15 / max(nel))
    k = rng.uniform(0.5, 2.0, g.ncells)
    CB.assemble_poisson(s, CB.Tables(g.refshape, order, qo), k, source=1.3, threads=4)
    cv = CellValues(QuadratureRule(g.refshape, qo), ip)
    K, f = CSC(dh.ndofs, cp, rv), np.zeros(dh.ndofs)
    A = start_assemble(K, f)
    for c in range(1, g.ncells + 1):
        Ke, fe = poisson_element(cv, g.getcoordinates(c), float(k[c - 1]), 1.3)
        assemble(A, dh.celldofs(c), Ke, fe)
    assert _rel_max(s.nzval, K.nzval) <= 1e-13
    assert _rel_max(s.f, f) <= 1e-13
    if celltype == "Hexahedron":
        nz, fv = s.nzval.copy(), s.f.copy()
        CB.assemble(s, k, source=1.3, threads=4)
        assert _rel_max(s.nzval, nz) <= 1e-14 and _rel_max(s.f, fv) <= 1e-14


@pytest.mark.parametrize("celltype,nel,order,qo", [("Hexahedron", (2, 2, 1), 2, 3), ("Quadrilateral", (6, 4), 1, 2)])
def test_elasticity_matches_serial_loop(celltype, nel, order, qo):
    rng = np.random.default_rng(5)
    dim = len(nel)
    g, ip, dh, cp, rv, cd, colors, s = _setup(celltype, nel, order, dim, rng, amount=0.12 / max(nel))
    E = rng.uniform(0.5, 2.0, g.ncells)
    CB.assemble_elasticity(s, CB.Tables(g.refshape, order, qo), E, 0.3, threads=4)
    cv = CellValues(QuadratureRule(g.refshape, qo), ip)
    K, f = CSC(dh.ndofs, cp, rv), np.zeros(dh.ndofs)
    A = start_assemble(K, f)
    for c in range(1, g.ncells + 1):
        assemble(A, dh.celldofs(c), elasticity_element(cv, g.getcoordinates(c), float(E[c - 1]), 0.3))
    assert _rel_max(s.nzval, K.nzval) <= 1e-12


def test_heat_run_known_answer():
    rng = np.random.default_rng(0)
    g, ip, dh, cp, rv, cd, colors, s = _setup("Quadrilateral", (20, 20), 1, 1, rng)
    ch = OC.ConstraintHandler(dh)
    facets = [fc for name in g.facetsets for fc in g.facetsets[name]]
    ch.add(OC.Dirichlet("u", facets, lambda x, t: 0.0)).close()
    grid = SimpleNamespace(nodes=g.nodes, cells=np.array(g.cells), ncells=g.ncells, refshape=g.refshape)
    rec = CB.run_heat(grid, SimpleNamespace(cell_dofs=cd), colors, cp, rv, ch.prescribed_dofs, threads=2,
                      rtol=1e-13, reps=2, log=lambda *a: None)
    assert abs(rec["u_norm"] - 3.307743912641305) <= 1e-10 * 3.307743912641305
    assert rec["relres"] <= 1e-13 and rec["value"] > 0


def test_cantilever_adjoint_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    from oracle.adjoint_torch import compliance_torch
    rng = np.random.default_rng(7)
    g, ip, dh, cp, rv, cd, colors, s = _setup("Quadrilateral", (8, 4), 1, 2, rng, amount=0.03,
                                              left=(0.0, 0.0), right=(2.0, 1.0))
    theta = rng.uniform(0.2, 1.0, g.ncells)
    ch = OC.ConstraintHandler(dh)
    ch.add(OC.Dirichlet("u", g.facetsets["left"], lambda x, t: [0.0, 0.0])).close()
    tip = int(np.nonzero(np.all(np.isclose(g.nodes, [2.0, 0.5]), axis=1))[0][0]) + 1
    cells = np.array(g.cells)
    c = int(np.nonzero(np.any(cells == tip, axis=1))[0][0])
    a = int(np.nonzero(cells[c] == tip)[0][0])
    ydof = int(cd[c, 2 * a + 1])
    grid = SimpleNamespace(nodes=g.nodes, cells=cells, ncells=g.ncells, refshape=g.refshape)
    rec = CB.run_cantilever(grid, SimpleNamespace(cell_dofs=cd), theta, colors, cp, rv, ch.prescribed_dofs, ydof,
                            threads=2, rtol=1e-14, log=lambda *a: None)
    assert rec["solve_complete"]

    fvec = np.zeros(dh.ndofs)
    fvec[ydof - 1] = -1.0
    free = np.setdiff1d(np.arange(dh.ndofs), ch.prescribed_dofs - 1)
    X = torch.tensor(g.nodes, dtype=torch.float64, requires_grad=True)
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    J = compliance_torch(X, cells, cd, dh.ndofs, th, fvec, free)
    J.backward()
    assert abs(rec["J"] - J.item()) <= 1e-11 * abs(J.item())
    assert _rel_max(rec["dJdtheta"], th.grad.numpy()) <= 1e-9
    assert _rel_max(rec["dJdX"], X.grad.numpy()) <= 1e-9


def test_batch_run_completes():
    rng = np.random.default_rng(2)
    probs = []
    for celltype, nel in (("Quadrilateral", (9, 7)), ("Hexahedron", (4, 3, 3))):
        g, ip, dh, cp, rv, cd, colors, s = _setup(celltype, nel, 1, 1, rng)
        ch = OC.ConstraintHandler(dh)
        ch.add(OC.Dirichlet("u", [fc for n in g.facetsets for fc in g.facetsets[n]], lambda x, t: 0.0)).close()
        grid = SimpleNamespace(nodes=g.nodes, cells=np.array(g.cells), ncells=g.ncells, refshape=g.refshape)
        probs.append((grid, SimpleNamespace(cell_dofs=cd), colors, cp, rv, ch.prescribed_dofs,
                      rng.uniform(0.5, 2.0, g.ncells)))
    rec = CB.run_batch(probs, threads=2, log=lambda *a: None)
    assert rec["total_iterations"] > 0 and rec["p99_ms"] >= rec["p50_ms"] > 0


def test_error_codes():
    rng = np.random.default_rng(1)
    g, ip, dh, cp, rv, cd, colors, s = _setup("Quadrilateral", (3, 3), 1, 1, rng)
    tabs = CB.Tables(g.refshape, 1, 2)
    k = np.ones(g.ncells)
    # a pattern without the entry (1st dof, 2nd dof of cell 1) and its transpose
    d0, d1 = cd[0, 0] - 1, cd[0, 1] - 1
    keep = np.ones(len(rv), dtype=bool)
    cols = np.repeat(np.arange(dh.ndofs), np.diff(np.asarray(cp) - 1))
    for r, cc in ((d0, d1), (d1, d0)):
        keep &= ~((cols == cc) & (np.asarray(rv) - 1 == r))
    ncp = np.concatenate([[1], 1 + np.cumsum(np.bincount(cols[keep], minlength=dh.ndofs))])
    s2 = CB.System(g.nodes, np.array(g.cells), cd, ncp, np.asarray(rv)[keep], colors)
    with pytest.raises(RuntimeError, match="missing sparsity pattern entry"):
        CB.assemble_poisson(s2, tabs, k, threads=2)
    cells = np.array(g.cells)
    cells[4] = cells[4][[1, 0, 3, 2]]          # a mirrored cell: det(J) < 0
    s3 = CB.System(g.nodes, cells, cd, cp, rv, colors)
    with pytest.raises(RuntimeError, match="det"):
        CB.assemble_poisson(s3, tabs, k, threads=2)
