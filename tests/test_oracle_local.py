"""The oracle's apply_local! and PeriodicDirichlet bookkeeping pinned by the reference's
"local application of bc" test (test/test_constraints.jl:1426-1611): a 5×5 Quadrilateral grid
on [-1, 1]², Q1, per-cell conductivity k = cell id and source b = 1 / cell id, Dirichlet
u = 0 on "left" and u = 1 on "right":
  * assembling after apply_local! gives the standard K/f (assemble, then apply!) on the free
    dofs, a diagonal prescribed block with f / diag = g, and the same solution,
    norm(u) = 3.8249286998373586 (apply_zero: 0.06401424182205259);
  * with PeriodicDirichlet (mirror "bottom", image "top") before the two Dirichlet sets,
    u = apply!(K \\ f, ch) has norm 3.7828270430540893 (apply_zero: 0.02672553850330505).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import constraints as OC
from oracle.affine import AffineConstraint, AffineHandler, add_constraint_entries
from oracle.affine import apply as affine_apply
from oracle.affine import apply_vector, periodic_dof_map
from oracle.dofs import DofHandler
from oracle.fem import CSC, CellValues, assemble, poisson_element, start_assemble
from oracle.grid import generate_grid
from oracle.refel import Lagrange, QuadratureRule, reference_facets
from oracle.sparsity import create_pattern


def _setup():
    grid = generate_grid("Quadrilateral", (5, 5))
    ip = Lagrange("RefQuadrilateral", 1)
    cv = CellValues(QuadratureRule("RefQuadrilateral", 2), ip)
    dh = DofHandler(grid).add("u", ip).close()
    ch = OC.ConstraintHandler(dh)
    ch.add(OC.Dirichlet("u", grid.facetsets["left"], lambda x, t: 0.0))
    ch.add(OC.Dirichlet("u", grid.facetsets["right"], lambda x, t: 1.0))
    ch.close()
    return grid, cv, dh, ch


def _elements(grid, cv):
    for c in range(1, grid.ncells + 1):
        Ke, fe = poisson_element(cv, grid.getcoordinates(c), float(c), 1.0 / c)
        yield c, Ke, fe


@pytest.mark.parametrize("azero,norm_u", [(False, 3.8249286998373586), (True, 0.06401424182205259)])
def test_apply_local_matches_standard_and_kat(azero, norm_u):
    grid, cv, dh, ch = _setup()
    cp, rv = create_pattern(dh)
    Ks, fs = CSC(dh.ndofs, cp, rv), np.zeros(dh.ndofs)
    Kl, fl = CSC(dh.ndofs, cp, rv), np.zeros(dh.ndofs)
    As, Al = start_assemble(Ks, fs), start_assemble(Kl, fl)
    for c, Ke, fe in _elements(grid, cv):
        dofs = dh.celldofs(c)
        assemble(As, dofs, Ke, fe)
        Ke2, fe2 = Ke.copy(), fe.copy()
        OC.apply_local(Ke2, fe2, dofs, ch, applyzero=azero)
        assemble(Al, dofs, Ke2, fe2)
    OC.apply(Ks, fs, ch, applyzero=azero)
    free = np.setdiff1d(np.arange(1, dh.ndofs + 1), ch.prescribed_dofs) - 1
    pd = ch.prescribed_dofs - 1
    Ss, Sl = Ks.to_scipy().toarray(), Kl.to_scipy().toarray()
    assert np.allclose(Sl[np.ix_(free, free)], Ss[np.ix_(free, free)], rtol=1e-13, atol=1e-13)
    assert np.allclose(fl[free], fs[free], rtol=1e-13, atol=1e-13)
    blk = Sl[np.ix_(pd, pd)]
    assert np.array_equal(blk, np.diag(np.diag(blk)))
    assert np.allclose(fl[pd] / np.diag(blk), 0.0 if azero else ch.inhomogeneities, rtol=1e-14, atol=1e-15)
    ul = spla.spsolve(Kl.to_scipy().tocsc(), fl)
    us = spla.spsolve(Ks.to_scipy().tocsc(), fs)
    assert abs(np.linalg.norm(us) - norm_u) <= 1e-12 * norm_u
    assert abs(np.linalg.norm(ul) - norm_u) <= 1e-12 * norm_u


def _node_xy(grid, dh):
    xy = np.zeros((dh.ndofs, 2))
    for c in range(1, grid.ncells + 1):
        x = grid.getcoordinates(c)
        for a, d in enumerate(dh.celldofs(c)):
            xy[d - 1] = x[a]
    return xy


@pytest.mark.parametrize("azero,norm_u", [(False, 3.7828270430540893), (True, 0.02672553850330505)])
def test_periodic_kat(azero, norm_u):
    grid, cv, dh, ch_dbc = _setup()
    xy = _node_xy(grid, dh)
    # PeriodicDirichlet's pairs: mirror facets ("bottom") in set order, each matched with the
    # image facet ("top") at the same x; image facet dofs in local facet order, each paired
    # with the mirror dof at the same x (ConstraintHandler.jl:1134-1186)
    def fdofs(fc):
        c, f = fc
        cd = dh.celldofs(c)
        return [int(cd[a - 1]) for a in reference_facets("RefQuadrilateral")[f - 1]]
    pairs = []
    for mf in grid.facetsets["bottom"]:
        mx = sorted(xy[d - 1][0] for d in fdofs(mf))
        imf = [fc for fc in grid.facetsets["top"] if np.allclose(sorted(xy[d - 1][0] for d in fdofs(fc)), mx)]
        assert len(imf) == 1
        mdofs = fdofs(mf)
        for idof in fdofs(imf[0]):
            cdof = [d for d in mdofs if abs(xy[d - 1][0] - xy[idof - 1][0]) < 1e-12]
            pairs.append((cdof[0], idof))
    dmap = periodic_dof_map(pairs)
    # the Dirichlet sets added after the periodic ones replace their dofs' constraints
    dir_vals = dict(zip(ch_dbc.prescribed_dofs.tolist(), ch_dbc.inhomogeneities.tolist()))
    ch = AffineHandler(dh.ndofs, dir_vals)
    for k, v in dmap.items():
        if k not in dir_vals:
            ch.add(AffineConstraint(k, [(v, 1.0)], 0.0))
    ch.close()
    cp, rv = create_pattern(dh)
    cp, rv = add_constraint_entries(cp, rv, ch)
    K, f = CSC(dh.ndofs, cp, rv), np.zeros(dh.ndofs)
    A = start_assemble(K, f)
    for c, Ke, fe in _elements(grid, cv):
        assemble(A, dh.celldofs(c), Ke, fe)
    affine_apply(K, f, ch, applyzero=azero)
    u = apply_vector(spla.spsolve(K.to_scipy().tocsc(), f), ch, applyzero=azero)
    assert abs(np.linalg.norm(u) - norm_u) <= 1e-12 * norm_u
