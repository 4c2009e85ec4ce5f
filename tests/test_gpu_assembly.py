"""Device parity: the HIP assembly path (through the C ABI) vs the CPU oracle.

Bar: pattern bit-exact (shared structure), ‖Δnzval‖₂/‖nzval‖₂ ≤ 1e-12 with an entrywise
floor |Δ| ≤ 1e-12·max|K| (FP64 element path), f likewise, prescribed dofs bit-exact,
u within 1e-8 of the oracle's direct solve.  Mixed path (FP32 element math, FP64
accumulation): K within 1e-6 relative (FP32 rounding of Ke), u after the FP64 solve
within 1e-8.
"""
import numpy as np
import pytest

import ferrite_amd as fa
from oracle import constraints as OC
from oracle.problems import assemble_poisson, heat_equation, poisson_conductivity, setup_poisson

pytestmark = pytest.mark.gpu


def _relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


ASM = {"gather": 2, "scatter": 1}


def _device_poisson(ct, nel, left, right, k, bc_sets, alg="workstream", precision="fp64", q=2, asm="gather"):
    g = fa.generate_grid(ct, nel, left, right)
    ip = fa.Lagrange(g.refshape, 1)
    cv = fa.CellValues(fa.QuadratureRule(g.refshape, q), ip)
    dh = fa.DofHandler(g).add("u", ip).close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh)
    fa.set_tuning(K, asm_mode=ASM[asm])
    f = fa.allocate_vector(K)
    colors = fa.create_coloring(g, alg)
    A = fa.start_assemble(K, f)
    fa.assemble_global(A, cv, colors, fa.PoissonElement(k=k, source=1.0), precision=precision)
    ch = fa.ConstraintHandler(dh)
    facets = [fc for s in bc_sets for fc in g.facetsets[s]]
    ch.add(fa.Dirichlet("u", facets, lambda x, t: 0.0)).close()
    return g, dh, K, f, ch


def _check_K(K, f, OK, of, tol):
    assert np.array_equal(K.colptr, OK.colptr) and np.array_equal(K.rowval, OK.rowval)
    nz = K.nzval()
    fv = f.to_numpy()
    assert _relerr(nz, OK.nzval) <= tol
    assert np.max(np.abs(nz - OK.nzval)) <= tol * np.max(np.abs(OK.nzval))
    assert _relerr(fv, of) <= tol
    return nz, fv


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 2e-6)])
def test_heat_equation_kat(precision, tol):
    """heat_equation.jl: 20x20 quads, norm(u) = 3.307743912641305."""
    print("case heat_equation", precision, flush=True)
    g, dh, K, f, ch = _device_poisson("Quadrilateral", (20, 20), None, None, None,
                                      ["left", "right", "top", "bottom"], precision=precision)
    u_ref, _ = heat_equation(20)
    OK, of = _oracle_sys(20)
    _check_K(K, f, OK, of, tol)
    m = fa.apply(K, f, ch)
    res = fa.solve(K, f, mode="cg", rtol=1e-13)
    assert abs(np.linalg.norm(res.u) - 3.307743912641305) / 3.307743912641305 < (1e-10 if precision == "fp64" else 1e-5)
    assert _relerr(res.u, u_ref) < (1e-10 if precision == "fp64" else 1e-5)
    assert m > 0
    # the mixed solver (FP32 operator, FP64 reliable updates) meets the same KAT on the cell
    # form of K (k_e times one element matrix: the FP64-assembled K passes the entry check; the
    # FP32-element one may fall back to the planes) and on the lattice planes
    mix = fa.solve(K, f, mode="mixed", rtol=1e-13)
    assert mix.operator == "cells" if precision == "fp64" else mix.operator in ("cells", "lattice")
    assert _relerr(mix.u, res.u) < 1e-10
    assert abs(np.linalg.norm(mix.u) - 3.307743912641305) / 3.307743912641305 < (1e-10 if precision == "fp64" else 1e-5)
    fa.set_tuning(K, cell_operator=-1)
    planes = fa.solve(K, f, mode="mixed", rtol=1e-13)
    fa.set_tuning(K, cell_operator=0)
    assert planes.operator == "lattice"
    assert _relerr(planes.u, res.u) < 1e-10


_cache = {}


def _oracle_sys(n):
    if n not in _cache:
        grid, cv, dh, K, f, ch = setup_poisson("Quadrilateral", (n, n))
        assemble_poisson(grid, cv, dh, K, f)
        _cache[n] = (K, f)
    return _cache[n]


def test_c1_quad_100_parity():
    """C1: 100x100 Q1 quads on [-1,1]^2, pattern bit-exact, K/f <= 1e-12, apply!, u vs direct solve."""
    print("case C1 100x100", flush=True)
    g, dh, K, f, ch = _device_poisson("Quadrilateral", (100, 100), None, None, None,
                                      ["left", "right", "top", "bottom"])
    grid, cv, odh, OK, of, och = setup_poisson("Quadrilateral", (100, 100))
    assemble_poisson(grid, cv, odh, OK, of)
    _check_K(K, f, OK, of, 1e-12)
    assert np.array_equal(ch.prescribed_dofs, och.prescribed_dofs)
    m = fa.apply(K, f, ch)
    om = OC.apply(OK, of, och)
    assert abs(m - om) <= 1e-12 * abs(om)  # meandiag: sequential vs tree summation order
    _check_K(K, f, OK, of, 1e-12)
    res = fa.solve(K, f, mode="cg", rtol=1e-12)
    u_ref = OC.solve(OK, of)
    assert _relerr(res.u, u_ref) < 1e-8
    assert res.relres <= 1e-12


@pytest.mark.parametrize("asm", ["gather", "scatter"])
@pytest.mark.parametrize("alg", ["workstream", "greedy"])
def test_hex_poisson_random_k_parity(alg, asm):
    """3D Poisson with k_e = exp(N(0, 0.25)), left/right Dirichlet, 6x5x4 hexes."""
    print("case hex 6x5x4", alg, asm, flush=True)
    nel = (6, 5, 4)
    k = poisson_conductivity(nel, seed=7)
    g, dh, K, f, ch = _device_poisson("Hexahedron", nel, (0, 0, 0), (1, 1, 1), k, ["left", "right"], alg=alg,
                                      asm=asm)
    grid, cv, odh, OK, of, och = setup_poisson("Hexahedron", nel, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0),
                                              bc_sets=["left", "right"])
    assemble_poisson(grid, cv, odh, OK, of, k_cell=k)
    _check_K(K, f, OK, of, 1e-12)
    fa.apply(K, f, ch)
    OC.apply(OK, of, och)
    _check_K(K, f, OK, of, 1e-12)
    u_ref = OC.solve(OK, of)
    for mode in ("cg", "mixed"):
        res = fa.solve(K, f, mode=mode, rtol=1e-12)
        assert _relerr(res.u, u_ref) < 1e-8, mode


@pytest.mark.parametrize("asm", ["gather", "scatter"])
def test_mixed_element_path_hex(asm):
    print("case mixed hex", asm, flush=True)
    nel = (5, 4, 3)
    k = poisson_conductivity(nel, seed=3)
    g, dh, K, f, ch = _device_poisson("Hexahedron", nel, (0, 0, 0), (1, 1, 1), k, ["left", "right"],
                                      precision="mixed", asm=asm)
    grid, cv, odh, OK, of, och = setup_poisson("Hexahedron", nel, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0),
                                              bc_sets=["left", "right"])
    assemble_poisson(grid, cv, odh, OK, of, k_cell=k)
    _check_K(K, f, OK, of, 2e-6)


@pytest.mark.parametrize("asm", ["gather", "scatter"])
def test_assembly_deterministic_and_fillzero(asm):
    """Run-twice bitwise equality (threaded_assembly.jl:398-403) and start_assemble(fillzero)."""
    print("case determinism", asm, flush=True)
    nel = (9, 7, 5)
    k = poisson_conductivity(nel, seed=11)
    g = fa.generate_grid("Hexahedron", nel, (0, 0, 0), (1, 1, 1))
    ip = fa.Lagrange("RefHexahedron", 1)
    cv = fa.CellValues(fa.QuadratureRule("RefHexahedron", 2), ip)
    dh = fa.DofHandler(g).add("u", ip).close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh)
    fa.set_tuning(K, asm_mode=ASM[asm])
    f = fa.allocate_vector(K)
    colors = fa.create_coloring(g)
    el = fa.PoissonElement(k=k)
    fa.assemble_global(fa.start_assemble(K, f), cv, colors, el)
    a1, f1 = K.nzval(), f.to_numpy()
    fa.assemble_global(fa.start_assemble(K, f), cv, colors, el)
    a2, f2 = K.nzval(), f.to_numpy()
    assert np.array_equal(a1, a2) and np.array_equal(f1, f2)
    fa.assemble_global(fa.start_assemble(K, f, fillzero=False), cv, colors, el)
    a3, f3 = K.nzval(), f.to_numpy()
    assert np.allclose(a3, 2 * a1, rtol=1e-15, atol=0) and np.allclose(f3, 2 * f1, rtol=1e-15, atol=0)
    # the other assembly strategy sums in another order: equal to rounding
    fa.set_tuning(K, asm_mode=ASM["scatter" if asm == "gather" else "gather"])
    fa.assemble_global(fa.start_assemble(K, f), cv, colors, el)
    a4 = K.nzval()
    assert np.abs(a4 - a1).max() <= 1e-14 * np.abs(a1).max()


def test_hex_geometry_paths():
    """The element kernel's three geometry forms: generic dM tables, trilinear coefficient
    form (a distorted mesh), and the constant-Jacobian form of affine meshes — each against
    the oracle's quadrature loop to 1e-12, and affine vs quadrature on one mesh to 1e-14."""
    print("case hex geometry paths", flush=True)
    nel = (5, 4, 3)
    k = poisson_conductivity(nel, seed=21)
    # distorted: interior nodes moved by up to 15% of the cell size (det J stays positive)
    g = fa.generate_grid("Hexahedron", nel, (0, 0, 0), (1, 1, 1))
    X = g.nodes
    inner = np.all((X > 1e-12) & (X < 1 - 1e-12), axis=1)
    h = np.array([1.0 / n for n in nel])
    X[inner] += np.random.default_rng(22).uniform(-0.15, 0.15, size=(int(inner.sum()), 3)) * h
    ip = fa.Lagrange("RefHexahedron", 1)
    cv = fa.CellValues(fa.QuadratureRule("RefHexahedron", 2), ip)
    dh = fa.DofHandler(g).add("u", ip).close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh)
    f = fa.allocate_vector(K)
    colors = fa.create_coloring(g)
    fa.assemble_global(fa.start_assemble(K, f), cv, colors, fa.PoissonElement(k=k))
    grid, ocv, odh, OK, of, _ = setup_poisson("Hexahedron", nel, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    grid.nodes[...] = X
    assemble_poisson(grid, ocv, odh, OK, of, k_cell=k)
    _check_K(K, f, OK, of, 1e-12)
    # affine mesh: constant-J form (default) vs the per-point form, and vs the oracle
    g2, dh2, K2, f2, _ = _device_poisson("Hexahedron", nel, (0, 0, 0), (1, 2, 3), k, ["left"])
    a_aff, f_aff = K2.nzval(), f2.to_numpy()
    fa.set_tuning(K2, affine=-1)
    fa.assemble_global(fa.start_assemble(K2, f2), cv, fa.create_coloring(g2), fa.PoissonElement(k=k))
    a_q, f_q = K2.nzval(), f2.to_numpy()
    assert np.abs(a_aff - a_q).max() <= 1e-14 * np.abs(a_q).max()
    assert np.abs(f_aff - f_q).max() <= 1e-14 * np.abs(f_q).max()
    grid2, ocv2, odh2, OK2, of2, _ = setup_poisson("Hexahedron", nel, (0.0, 0.0, 0.0), (1.0, 2.0, 3.0))
    assemble_poisson(grid2, ocv2, odh2, OK2, of2, k_cell=k)
    _check_K(K2, f2, OK2, of2, 1e-12)


@pytest.mark.parametrize("asm", ["gather", "scatter"])
def test_inverted_cell_raises_detj(asm):
    """det(J) <= 0 -> ArgumentError (common_values.jl:5), reported with the cell id."""
    print("case detJ error", asm, flush=True)
    g = fa.generate_grid("Quadrilateral", (4, 4))
    c = 6
    g.cells[c - 1] = g.cells[c - 1][[1, 0, 3, 2]]  # mirror cell 6
    ip = fa.Lagrange("RefQuadrilateral", 1)
    cv = fa.CellValues(fa.QuadratureRule("RefQuadrilateral", 2), ip)
    dh = fa.DofHandler(g).add("u", ip).close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh)
    fa.set_tuning(K, asm_mode=ASM[asm])
    with pytest.raises(fa.FemArgumentError, match="det\\(J\\) is not positive in cell 6"):
        fa.assemble_global(fa.start_assemble(K), cv, fa.create_coloring(g), fa.PoissonElement())


@pytest.mark.parametrize("asm", ["gather", "scatter"])
def test_missing_pattern_entry_raises(asm):
    """A coupling absent from the pattern with a nonzero value -> the reference's ErrorException text."""
    print("case missing pattern entry", asm, flush=True)
    import ctypes as C
    from ferrite_amd import _lib as L
    g = fa.generate_grid("Quadrilateral", (3, 3))
    ip = fa.Lagrange("RefQuadrilateral", 1)
    cv = fa.CellValues(fa.QuadratureRule("RefQuadrilateral", 2), ip)
    dh = fa.DofHandler(g).add("u", ip).close()
    cp, rv = fa.create_sparsity_pattern(dh)
    # drop one off-diagonal entry (row r, col c) from column c
    cidx = 4
    lo, hi = cp[cidx] - 1, cp[cidx + 1] - 1
    drop = next(p for p in range(lo, hi) if rv[p] != cidx + 1)
    rv2 = np.delete(rv, drop)
    cp2 = cp.copy()
    cp2[cidx + 1:] -= 1
    ctx = fa.Context(0)
    xyz = L.as_f64(g.nodes)
    cells = L.as_i64(g.cells)
    ctx.call("fem_mesh_upload", 2, g.nnodes, L.ptr(xyz, C.c_double), g.ncells, 4, L.ptr(cells, C.c_int64))
    cd = L.as_i64(dh.cell_dofs)
    ctx.call("fem_dofs_upload", dh.ndofs, 4, L.ptr(cd, C.c_int64))
    ctx.call("fem_pattern_upload", dh.ndofs, L.ptr(L.as_i64(cp2), C.c_int64), L.ptr(L.as_i64(rv2), C.c_int64))
    K = fa.device.DeviceMatrix(dh, ctx, cp2, rv2)
    fa.set_tuning(K, asm_mode=ASM[asm])
    with pytest.raises(fa.FemPatternError, match="missing in the sparsity pattern"):
        fa.assemble_global(fa.start_assemble(K), cv, fa.create_coloring(g), fa.PoissonElement())


def test_apply_inhomogeneous_bitexact_f():
    """f -= K g with g != 0: the row-gather matches the serial loop's f bit for bit (same K)."""
    print("case inhomogeneous apply", flush=True)
    nel = (6, 6)
    g = fa.generate_grid("Quadrilateral", nel)
    ip = fa.Lagrange("RefQuadrilateral", 1)
    cv = fa.CellValues(fa.QuadratureRule("RefQuadrilateral", 2), ip)
    dh = fa.DofHandler(g).add("u", ip).close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh)
    f = fa.allocate_vector(K)
    fa.assemble_global(fa.start_assemble(K, f), cv, fa.create_coloring(g), fa.PoissonElement())
    gf = lambda x, t: 1.0 + x[0] * x[1]  # noqa: E731
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.Dirichlet("u", g.facetsets["left"] + g.facetsets["top"], gf)).close()
    # oracle apply! on the device's own K and f
    from oracle.fem import CSC
    OK = CSC(dh.ndofs, K.colptr, K.rowval, K.nzval().copy())
    of = f.to_numpy()
    grid, _, odh, _, _, _ = setup_poisson("Quadrilateral", nel)
    och = OC.ConstraintHandler(odh)
    och.add(OC.Dirichlet("u", grid.facetsets["left"] + grid.facetsets["top"], gf)).close()
    assert np.array_equal(och.prescribed_dofs, ch.prescribed_dofs)
    om = OC.apply(OK, of, och)
    m = fa.apply(K, f, ch)
    nz = K.nzval()
    fv = f.to_numpy()
    free = np.ones(dh.ndofs, bool)
    free[ch.prescribed_dofs - 1] = False
    assert np.array_equal(fv[free], of[free])     # bitwise: same order, no FMA
    assert abs(m - om) <= 4e-16 * abs(om)
    assert np.allclose(nz, OK.nzval, rtol=1e-15, atol=0)
    assert np.allclose(fv, of, rtol=1e-15, atol=0)
