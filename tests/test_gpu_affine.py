"""Affine and periodic constraints on the device (apply! with _condense!, apply!(u, ch)) against
the oracle (oracle/affine.py) and against properties the mathematics fixes.

  * apply!(K, f, ch): starting from the device's own assembled K and f, every entry of K and f
    equals the oracle's in-place condensation bit for bit, except the prescribed diagonal and
    right-hand side entries, which carry meandiag(K) (a reduction, within 1e-14 relative);
  * apply!(u, ch) after the solve equals the oracle's apply!(u, ch) bit for bit, and u agrees
    with the oracle's direct solve of its own condensed system within 1e-8;
  * periodic Poisson: u_mirror == u_image exactly, and shifting a random per-cell conductivity
    cyclically along a periodic axis shifts the solution the same way (within 1e-9), which
    homogeneous Neumann sides would not do;
  * a pattern without the constraint entries is rejected with the reference's message;
  * ConformityConstraint on quadrilateral and hexahedral grids with hanging nodes (2 and 4
    masters): apply! and the solve against the oracle, and the linear patch test through the
    hanging nodes.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import ferrite_amd as fa
from oracle.affine import AffineConstraint as OAC
from oracle.affine import AffineHandler
from oracle.affine import apply as oracle_apply
from oracle.affine import apply_vector as oracle_apply_vector
from oracle.fem import CSC

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _oracle_handler(ch):
    aff = set(int(d) for d in ch.affine_dofs)
    dirichlet = {int(d): float(v) for d, v in zip(ch.prescribed_dofs, ch.inhomogeneities) if int(d) not in aff}
    oh = AffineHandler(ch.dh.ndofs, dirichlet)
    dofs, ptr, cd, cv, b = ch.affine_arrays()
    for i, d in enumerate(dofs):
        oh.add(OAC(d, list(zip(cd[ptr[i]:ptr[i + 1]], cv[ptr[i]:ptr[i + 1]])), b[i]))
    return oh.close()


def _assemble(g, dh, ch, k, qorder=2, pattern_ch=True):
    ip = dh.field_ips[0]
    cv = fa.CellValues(fa.QuadratureRule(g.refshape, qorder), ip)
    K = fa.allocate_matrix(fa.MI355XCSC, dh, ch if pattern_ch else None)
    f = fa.allocate_vector(K)
    A = fa.start_assemble(K, f)
    fa.assemble_global(A, cv, fa.create_coloring(g), fa.PoissonElement(k=k, source=1.0))
    return K, f


def _check_apply(K, f, ch):
    """apply! on the device against the oracle's in-place condensation of the same K, f."""
    n = K.n
    K0, f0 = K.nzval(), f.to_numpy()
    m = fa.apply(K, f, ch)
    oh = _oracle_handler(ch)
    OK = CSC(n, K.colptr, K.rowval, K0.copy())
    of = f0.copy()
    om = oracle_apply(OK, of, oh)
    assert abs(m - om) <= 1e-14 * abs(om)
    nz, fv = K.nzval(), f.to_numpy()
    diag = np.array([OK.find(d, d) for d in oh.prescribed_dofs], dtype=np.int64)
    other = np.ones(len(nz), dtype=bool)
    other[diag] = False
    assert np.array_equal(nz[other], OK.nzval[other])
    assert np.all(np.abs(nz[diag] - OK.nzval[diag]) <= 1e-14 * abs(om))
    free = np.ones(n, dtype=bool)
    free[oh.prescribed_dofs - 1] = False
    assert np.array_equal(fv[free], of[free])
    assert np.all(np.abs(fv[~free] - of[~free]) <= 1e-14 * abs(om) * (1 + np.abs(oh.inhomogeneities)))
    return OK, of, oh


def _solve_and_check(K, f, ch, OK, of, oh):
    res = fa.solve(K, f, mode="cg", rtol=1e-13, maxit=20000)
    assert res.operator != "elements"
    raw = res.u.copy()
    u = fa.apply_solution(K, ch)
    assert np.array_equal(u, oracle_apply_vector(raw.copy(), oh))
    u_ref = oracle_apply_vector(spla.spsolve(OK.to_scipy().tocsc(), of), oh)
    assert _rel(u, u_ref) < 1e-8, _rel(u, u_ref)
    return u


def test_affine_quad_poisson():
    """Dirichlet left (u = 1 + y), and general affine constraints with Dirichlet and free
    masters, including one replacing a Dirichlet dof."""
    print("case affine quad 12x10", flush=True)
    g = fa.generate_grid("Quadrilateral", (12, 10), (0.0, 0.0), (1.0, 1.0))
    dh = fa.DofHandler(g).add("u", fa.Lagrange(g.refshape, 1)).close()
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.Dirichlet("u", g.facetsets["left"], lambda x, t: 1.0 + x[1]))
    left = sorted(ch._entries)
    right = sorted({int(dh.cell_dofs[c - 1][i]) for c, fct in g.facetsets["right"]
                    for i in ([1, 2] if fct == 2 else [])})
    rng = np.random.default_rng(3)
    slaves = set(right[::2]) | {left[3]}
    for j, d in enumerate(right[::2]):
        masters = [int(rng.integers(30, 100)), left[(2 * j) % len(left)], int(rng.integers(30, 100))]
        masters = [mm for mm in dict.fromkeys(masters) if mm not in slaves]
        ch.add(fa.AffineConstraint(d, [(mm, float(rng.uniform(-0.7, 0.7))) for mm in masters],
                                   float(rng.uniform(-0.2, 0.2))))
    assert 60 not in slaves
    ch.add(fa.AffineConstraint(left[3], [(60, 0.5)], 0.25))
    ch.close()
    k = np.random.default_rng(5).uniform(0.5, 2.0, g.ncells)
    K, f = _assemble(g, dh, ch, k)
    OK, of, oh = _check_apply(K, f, ch)
    u = _solve_and_check(K, f, ch, OK, of, oh)
    dofs, ptr, cd, cv, b = ch.affine_arrays()
    for i, d in enumerate(dofs):
        v = b[i]
        for j in range(ptr[i], ptr[i + 1]):
            v += cv[j] * u[cd[j] - 1]
        assert u[d - 1] == v


def _node_dofs(g, dh):
    nd = np.zeros(g.nnodes, dtype=np.int64)
    nd[g.cells.ravel() - 1] = dh.cell_dofs.ravel()
    return nd


@pytest.mark.parametrize("ct,nel,axes", [
    ("Quadrilateral", (16, 8), (0,)),
    ("Hexahedron", (8, 6, 4), (0, 1)),
])
def test_periodic_poisson_shift(ct, nel, axes):
    """Periodic along `axes`, Dirichlet u = 0 on the remaining last axis' two sides."""
    print("case periodic", ct, nel, axes, flush=True)
    dim = len(nel)
    g = fa.generate_grid(ct, nel, (0.0,) * dim, (1.0,) * dim)
    dh = fa.DofHandler(g).add("u", fa.Lagrange(g.refshape, 1)).close()
    sides = ({0: ("left", "right"), 1: ("bottom", "top")} if dim == 2 else
             {0: ("left", "right"), 1: ("front", "back"), 2: ("bottom", "top")})
    ch = fa.ConstraintHandler(dh)
    pairs = []
    for ax in axes:
        pairs += fa.collect_periodic_facets(g, sides[ax][1], sides[ax][0])
    ch.add(fa.PeriodicDirichlet("u", pairs))
    last = dim - 1
    ch.add(fa.Dirichlet("u", g.facetsets[sides[last][0]] + g.facetsets[sides[last][1]], lambda x, t: 0.0))
    ch.close()
    kc = np.random.default_rng(11).uniform(0.5, 2.0, tuple(reversed(nel)))   # cell index: x fastest
    nd = _node_dofs(g, dh)
    ijk = np.stack(np.unravel_index(np.arange(g.nnodes), tuple(n + 1 for n in reversed(nel))), axis=1)[:, ::-1]
    sols = []
    for s in (0, 3):
        k = np.roll(kc, s, axis=dim - 1).ravel()     # shift along x by s cells
        K, f = _assemble(g, dh, ch, k)
        OK, of, oh = _check_apply(K, f, ch)
        sols.append(_solve_and_check(K, f, ch, OK, of, oh))
    u0, u3 = sols
    # u_mirror == u_image exactly
    dofs, ptr, cd, _, _ = ch.affine_arrays()
    assert np.array_equal(u0[dofs - 1], u0[cd - 1])
    # the solution of the shifted conductivity is the shifted solution
    nx = nel[0]
    shifted = ijk.copy()
    shifted[:, 0] = (ijk[:, 0] + 3) % nx
    lin = np.ravel_multi_index(tuple(shifted[:, ::-1].T), tuple(n + 1 for n in reversed(nel)))
    a, b = u0[nd - 1], u3[nd[lin] - 1]
    assert np.max(np.abs(a - b)) <= 1e-9 * np.max(np.abs(a)), np.max(np.abs(a - b))
    assert np.max(np.abs(a)) > 1e-3


def test_pattern_without_constraint_entries_rejected():
    print("case affine missing pattern entries", flush=True)
    g = fa.generate_grid("Quadrilateral", (6, 6), (0.0, 0.0), (1.0, 1.0))
    dh = fa.DofHandler(g).add("u", fa.Lagrange(g.refshape, 1)).close()
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.Dirichlet("u", g.facetsets["left"], lambda x, t: 0.0))
    ch.add(fa.AffineConstraint(dh.ndofs, [(1 + dh.ndofs // 2, 1.0)], 0.0))
    ch.close()
    K, f = _assemble(g, dh, ch, None, pattern_ch=False)
    with pytest.raises(fa.FemPatternError, match=r"allocate_matrix\(dh, ch\)"):
        fa.apply(K, f, ch)


def test_conformity_constraints_on_a_nonconforming_grid():
    """Hanging nodes (ConformityConstraint on a 6x5 grid with three cells split in four): the
    device apply! against the oracle's condensation of the same K, f; and the patch test a
    conforming space passes exactly: with u = 1 + x + 2y on the boundary, no source and a
    constant conductivity, the discrete solution is that linear function at every node,
    hanging nodes included (each is the mean of its edge's two ends), to the solver's
    tolerance.  Without the constraints the same solve misses the linear function by 0.25 at
    the hanging nodes (the oracle's dense solve), so the test sees a dropped constraint."""
    from ncgrid_util import refine_quads
    print("case conformity constraints", flush=True)
    g = refine_quads(fa.generate_grid("Quadrilateral", (6, 5)), [8, 10, 21])
    assert len(g.conformity_info) == 12
    dh = fa.DofHandler(g).add("u", fa.Lagrange(g.refshape, 1)).close()
    bnd = g.facetsets["left"] + g.facetsets["right"] + g.facetsets["bottom"] + g.facetsets["top"]
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.Dirichlet("u", bnd, lambda x, t: 1.0 + x[0] + 2.0 * x[1]))
    ch.add(fa.ConformityConstraint("u"))
    ch.close()
    # random conductivity, unit source: device apply! and solve against the oracle
    k = np.random.default_rng(9).uniform(0.5, 2.0, g.ncells)
    K, f = _assemble(g, dh, ch, k)
    OK, of, oh = _check_apply(K, f, ch)
    _solve_and_check(K, f, ch, OK, of, oh)
    # the patch test
    cv = fa.CellValues(fa.QuadratureRule(g.refshape, 2), dh.field_ips[0])
    K2 = fa.allocate_matrix(fa.MI355XCSC, dh, ch)
    f2 = fa.allocate_vector(K2)
    fa.assemble_global(fa.start_assemble(K2, f2), cv, fa.create_coloring(g), fa.PoissonElement(source=0.0))
    fa.apply(K2, f2, ch)
    fa.solve(K2, f2, mode="cg", rtol=1e-14, maxit=20000)
    u = fa.apply_solution(K2, ch)
    nd = _node_dofs(g, dh)
    exact = 1.0 + g.nodes[:, 0] + 2.0 * g.nodes[:, 1]
    assert np.max(np.abs(u[nd - 1] - exact)) <= 1e-10 * np.max(np.abs(exact))
    for h, ms in g.conformity_info.items():
        mean = sum(u[nd[m - 1] - 1] for m in ms) / len(ms)
        assert abs(u[nd[h - 1] - 1] - mean) <= 1e-14 * abs(mean)


def test_conformity_constraints_hexahedra():
    """Hanging nodes in 3D: a 5x4x4 hex grid with three interior cells split in eight (two of
    them face neighbours, so their shared face is regular); edge midpoints take 2 masters and
    face centres 4.  apply! and the solve against the oracle, and the linear patch test
    u = 1 + x + 2y + 3z through the hanging nodes (without the constraints the oracle's dense
    solve misses it by 1.2)."""
    from ncgrid_util import refine_hexes
    print("case conformity constraints, hexahedra", flush=True)
    g = refine_hexes(fa.generate_grid("Hexahedron", (5, 4, 4)), [27, 28, 54])
    nm = [len(v) for v in g.conformity_info.values()]
    assert len(nm) == 48 and nm.count(4) == 16
    dh = fa.DofHandler(g).add("u", fa.Lagrange(g.refshape, 1)).close()
    bnd = [fc for n in ("left", "right", "front", "back", "bottom", "top") for fc in g.facetsets[n]]
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.Dirichlet("u", bnd, lambda x, t: 1.0 + x[0] + 2.0 * x[1] + 3.0 * x[2]))
    ch.add(fa.ConformityConstraint("u"))
    ch.close()
    k = np.random.default_rng(11).uniform(0.5, 2.0, g.ncells)
    K, f = _assemble(g, dh, ch, k)
    OK, of, oh = _check_apply(K, f, ch)
    _solve_and_check(K, f, ch, OK, of, oh)
    cv = fa.CellValues(fa.QuadratureRule(g.refshape, 2), dh.field_ips[0])
    K2 = fa.allocate_matrix(fa.MI355XCSC, dh, ch)
    f2 = fa.allocate_vector(K2)
    fa.assemble_global(fa.start_assemble(K2, f2), cv, fa.create_coloring(g), fa.PoissonElement(source=0.0))
    fa.apply(K2, f2, ch)
    fa.solve(K2, f2, mode="cg", rtol=1e-14, maxit=20000)
    u = fa.apply_solution(K2, ch)
    nd = _node_dofs(g, dh)
    exact = 1.0 + g.nodes[:, 0] + 2.0 * g.nodes[:, 1] + 3.0 * g.nodes[:, 2]
    assert np.max(np.abs(u[nd - 1] - exact)) <= 1e-10 * np.max(np.abs(exact))
    for h, ms in g.conformity_info.items():
        mean = sum(u[nd[m - 1] - 1] for m in ms) / len(ms)
        assert abs(u[nd[h - 1] - 1] - mean) <= 1e-14 * abs(mean)
