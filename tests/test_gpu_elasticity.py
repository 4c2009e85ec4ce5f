"""Device parity for linear elasticity (Q2 hex^3 with MFMA/FP64 element paths, Q1 quad^2 plane
strain), FacetValues traction, point loads, and the compliance adjoint.

Bars: FP64 element path K ≤ 1e-12 (normwise and entrywise floor) vs the oracle; FP32 (MFMA)
element path K ≤ 1e-5 relative (FP32 rounding of Ke); traction f ≤ 1e-12; u ≤ 1e-8 vs the
oracle's direct solve; adjoint gradients ≤ 1e-6 relative vs central differences (h = 1e-6)
and vs torch float64 autograd.
"""
import numpy as np
import pytest

import ferrite_amd as fa
from oracle import constraints as OC
from oracle.fem import FacetValues as OFacetValues
from oracle.problems import assemble_elasticity, assemble_traction, setup_vector
from oracle.refel import FacetQuadratureRule as OFQR
from oracle.refel import Lagrange as OLagrange

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _device_vector(ct, nel, left, right, order, q, colors="workstream"):
    g = fa.generate_grid(ct, nel, left, right)
    dim = len(nel)
    ip = fa.Lagrange(g.refshape, order) ** dim
    cv = fa.CellValues(fa.QuadratureRule(g.refshape, q), ip)
    dh = fa.DofHandler(g).add("u", ip).close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh)
    f = fa.allocate_vector(K)
    return g, ip, cv, dh, K, f, fa.create_coloring(g, colors)


ASM = {"gather": 0, "scatter": 1}


@pytest.mark.parametrize("asm", ["gather", "scatter"])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-5)])
def test_q2_hex_elasticity_parity(precision, tol, asm):
    print("case q2 hex elasticity", precision, asm, flush=True)
    nel = (3, 2, 2)
    rng = np.random.default_rng(5)
    E = rng.uniform(0.5, 2.0, int(np.prod(nel)))
    g, ip, cv, dh, K, f, colors = _device_vector("Hexahedron", nel, (0, 0, 0), (1, 1, 1), 2, 3)
    fa.set_tuning(K, asm_mode=ASM[asm])
    fa.assemble_global(fa.start_assemble(K, f), cv, colors, fa.ElasticityElement(E=E, nu=0.3), precision=precision)
    t = fa.timings(K)
    assert (t["element_ms"] > 0 and t["scatter_ms"] > 0) == (asm == "gather")
    grid, ocv, odh, OK, of = setup_vector("Hexahedron", nel, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 2, 3)
    assemble_elasticity(grid, ocv, odh, OK, of, E, 0.3)
    assert np.array_equal(K.colptr, OK.colptr) and np.array_equal(K.rowval, OK.rowval)
    nz = K.nzval()
    assert _rel(nz, OK.nzval) <= tol
    assert np.abs(nz - OK.nzval).max() <= tol * np.abs(OK.nzval).max()
    # traction (0, 0, -1) on 'right' with the 3x3 facet rule
    fv = fa.FacetValues(fa.FacetQuadratureRule("RefHexahedron", 3), ip)
    fa.apply_traction(K, f, fv, g.facetsets["right"], (0.0, 0.0, -1.0))
    ofv = OFacetValues(OFQR("RefHexahedron", 3), OLagrange("RefHexahedron", 2) ** 3)
    assemble_traction(grid, ofv, odh, of, grid.facetsets["right"], (0.0, 0.0, -1.0))
    fvd = f.to_numpy()
    assert _rel(fvd, of) <= 1e-12
    assert abs(fvd[2::3].sum() + 1.0) < 1e-12      # total force = t * area(right) = -1
    # clamp 'left', apply!, solve
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.Dirichlet("u", g.facetsets["left"], lambda x, t: [0.0, 0.0, 0.0])).close()
    och = OC.ConstraintHandler(odh)
    och.add(OC.Dirichlet("u", grid.facetsets["left"], lambda x, t: [0.0, 0.0, 0.0])).close()
    assert np.array_equal(ch.prescribed_dofs, och.prescribed_dofs)
    fa.apply(K, f, ch)
    OC.apply(OK, of, och)
    res = fa.solve(K, f, mode="cg", rtol=1e-13)
    u = OC.solve(OK, of)
    assert _rel(res.u, u) < (1e-8 if precision == "fp64" else 1e-4)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-5)])
def test_q2_hex_gather_parity_distorted(precision, tol):
    """Owner-computes Q2 path on a 6x5x4 mesh (vertex nodes in 8 cells) with moved interior
    nodes: K vs the oracle, start_assemble(fillzero=false) accumulation, run-twice bitwise,
    and the colored scatter within rounding."""
    print("case q2 hex gather distorted", precision, flush=True)
    nel = (6, 5, 4)
    rng = np.random.default_rng(11)
    E = rng.uniform(0.5, 2.0, int(np.prod(nel)))
    g, ip, cv, dh, K, f, colors = _device_vector("Hexahedron", nel, (0, 0, 0), (1, 1, 1), 2, 3)
    grid, ocv, odh, OK, of = setup_vector("Hexahedron", nel, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 2, 3)
    # move interior nodes (the same on both sides); boundary nodes stay on the box
    X = g.nodes.copy()
    inner = np.all((X > 1e-12) & (X < 1 - 1e-12), axis=1)
    X[inner] += rng.uniform(-0.04, 0.04, (int(inner.sum()), 3))
    g.nodes[:] = X
    grid.nodes[:] = X
    K = fa.allocate_matrix(fa.MI355XCSC, dh)
    f = fa.allocate_vector(K)
    el = fa.ElasticityElement(E=E, nu=0.3)
    fa.assemble_global(fa.start_assemble(K, f), cv, colors, el, precision=precision)
    assemble_elasticity(grid, ocv, odh, OK, of, E, 0.3)
    nz = K.nzval()
    assert _rel(nz, OK.nzval) <= tol
    assert np.abs(nz - OK.nzval).max() <= tol * np.abs(OK.nzval).max()
    S = K.to_scipy()
    assert abs(S - S.T).max() == 0.0        # symmetric bit for bit, like Ferrite's K
    fa.assemble_global(fa.start_assemble(K, f), cv, colors, el, precision=precision)
    assert np.array_equal(K.nzval(), nz)
    fa.assemble_global(fa.start_assemble(K, f, fillzero=False), cv, colors, el, precision=precision)
    assert _rel(K.nzval(), 2 * OK.nzval) <= tol
    fa.set_tuning(K, asm_mode=1)
    fa.assemble_global(fa.start_assemble(K, f), cv, colors, el, precision=precision)
    assert _rel(K.nzval(), nz) <= (1e-14 if precision == "fp64" else 1e-6)


def _cantilever(nx, ny, seed):
    g, ip, cv, dh, K, f, colors = _device_vector("Quadrilateral", (nx, ny), (0.0, 0.0), (2.0, 1.0), 1, 2)
    rng = np.random.default_rng(seed)
    theta = rng.uniform(0.2, 1.0, g.ncells)
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.Dirichlet("u", g.facetsets["left"], lambda x, t: [0.0, 0.0])).close()
    tip = int(np.nonzero(np.all(np.isclose(g.nodes, [2.0, 0.5]), axis=1))[0][0]) + 1
    # dof of the tip node's y component: find a cell holding the node
    c = int(np.nonzero(np.any(g.cells == tip, axis=1))[0][0])
    a = int(np.nonzero(g.cells[c] == tip)[0][0])
    ydof = int(dh.cell_dofs[c, 2 * a + 1])

    def loads(K_, f_):
        fa.add_point_load(K_, f_, [ydof], [-1.0])
    from ferrite_amd.autodiff import ComplianceProblem
    prob = ComplianceProblem(K, f, cv, colors, ch, loads, rtol=1e-12)
    return g, dh, K, f, ch, prob, theta, ydof


def test_plane_strain_q1_parity():
    print("case plane strain q1", flush=True)
    nel = (8, 4)
    rng = np.random.default_rng(8)
    E = rng.uniform(0.2, 1.0, 32)
    g, ip, cv, dh, K, f, colors = _device_vector("Quadrilateral", nel, (0.0, 0.0), (2.0, 1.0), 1, 2)
    fa.assemble_global(fa.start_assemble(K, f), cv, colors, fa.ElasticityElement(E=E, nu=0.3))
    grid, ocv, odh, OK, of = setup_vector("Quadrilateral", nel, (0.0, 0.0), (2.0, 1.0), 1, 2)
    assemble_elasticity(grid, ocv, odh, OK, of, E, 0.3)
    nz = K.nzval()
    assert _rel(nz, OK.nzval) <= 1e-12
    assert np.abs(nz - OK.nzval).max() <= 1e-12 * np.abs(OK.nzval).max()


def test_compliance_adjoint_vs_fd_and_torch():
    """dJ/dθ and dJ/dX from the adjoint kernel vs central differences and torch autograd."""
    print("case compliance adjoint", flush=True)
    g, dh, K, f, ch, prob, theta, ydof = _cantilever(16, 8, seed=9)
    X0 = g.nodes.copy()
    J, dth, dx, res = prob.compliance(theta)
    assert res.relres < 1e-12
    rng = np.random.default_rng(10)
    h = 1e-6
    # A central difference at h = 1e-6 cannot resolve gradient entries much below its
    # noise: each FP64 evaluation of J (here ~150) carries ~1e-14 |J| of rounding even with
    # an exact dense solve, which moves (J+ − J−)/2h by ~4e-8 |J| (measured with the torch
    # restatement: entry 31, |g| = 0.256, FD off by 2.5e-5 relative at h = 1e-6, by 1e-7 at
    # h = 1e-4).  Entries are therefore checked to 1e-6 relative above a floor of 1e-7 |J|,
    # and the sampled gradient as a whole to 1e-6 in norm; the entry-wise 1e-6 check against
    # exact derivatives is the torch autograd comparison below.
    floor = 1e-7 * abs(J)
    fds, ads = [], []
    for e in rng.choice(g.ncells, 12, replace=False):
        tp, tm = theta.copy(), theta.copy()
        tp[e] += h
        tm[e] -= h
        fd = (prob.compliance(tp)[0] - prob.compliance(tm)[0]) / (2 * h)
        assert abs(fd - dth[e]) <= 1e-6 * abs(dth[e]) + floor, (e, fd, dth[e])
        fds.append(fd)
        ads.append(dth[e])
    assert _rel(np.array(ads), np.array(fds)) <= 1e-6
    fds, ads = [], []
    for k in rng.choice(g.nnodes * 2, 12, replace=False):
        nd, d = divmod(int(k), 2)
        Xp, Xm = X0.copy(), X0.copy()
        Xp[nd, d] += h
        Xm[nd, d] -= h
        fp = prob.compliance(theta, Xp)[0]
        fm = prob.compliance(theta, Xm)[0]
        fd = (fp - fm) / (2 * h)
        assert abs(fd - dx[nd, d]) <= 1e-6 * abs(dx[nd, d]) + floor, (nd, d, fd, dx[nd, d])
        fds.append(fd)
        ads.append(dx[nd, d])
    assert _rel(np.array(ads), np.array(fds)) <= 1e-6
    prob.compliance(theta, X0)
    # independent AD: torch float64 autograd through a dense restatement
    import torch
    from oracle.adjoint_torch import compliance_torch
    free = np.setdiff1d(np.arange(dh.ndofs), ch.prescribed_dofs - 1)
    fvec = np.zeros(dh.ndofs)
    fvec[ydof - 1] = -1.0
    th_t = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    X_t = torch.tensor(X0, dtype=torch.float64, requires_grad=True)
    Jt = compliance_torch(X_t, g.cells, dh.cell_dofs, dh.ndofs, th_t, fvec, free)
    Jt.backward()
    assert abs(Jt.item() - J) <= 1e-10 * abs(J)
    gt, gx = th_t.grad.numpy(), X_t.grad.numpy()
    assert _rel(dth, gt) <= 1e-8
    assert _rel(dx, gx) <= 1e-8
    # entry by entry, to 1e-6 relative (absolute floor 1e-8 max|g| for entries that vanish)
    assert np.all(np.abs(dth - gt) <= 1e-6 * np.abs(gt) + 1e-8 * np.abs(gt).max())
    assert np.all(np.abs(dx - gx) <= 1e-6 * np.abs(gx) + 1e-8 * np.abs(gx).max())


def test_compliance_adjoint_vs_exact_central_differences():
    """dJ/dθ and dJ/dX against central differences at h = 1e-6 that carry no solve noise: each
    J(· ± h) − J is a Woodbury update of one factored K_ff, and the adjoint kernel runs on the
    same LU solution.  Every entry to 1e-6 relative; dJ/dX entries that are small through
    cancellation between the four cells around the node (here one entry of 3e-4 beside entries
    of 1-10) keep an absolute floor of 1e-9 max|g|, the rounding of the assembled K(X ± h)
    (~eps E per entry) times the patch displacement squared, divided by 2h."""
    print("case compliance adjoint exact differences 96x48", flush=True)
    from ferrite_amd import _lib as L
    from ferrite_amd.profile_configs import exact_central_differences, pick_fd_entries
    g, dh, K, f, ch, prob, theta, ydof = _cantilever(96, 48, seed=4)
    cells, coords = pick_fd_entries(g, np.random.default_rng(5), 16, 16)
    assert len(cells) == 16 and len(coords) == 16
    h = 1e-6
    u, refine_res, fd = exact_central_differences(prob, theta, cells, coords, (h, 2 * h))
    assert refine_res < 1e-10          # ‖f − K u‖/‖f‖ of the LU solution after refinement
    K.ctx.call("fem_solution_upload", L.ptr(u, L.C.c_double))
    J, dth, dx = fa.compliance_gradient(K, theta)
    assert abs(J - (-u[ydof - 1])) <= 1e-12 * abs(J)
    adt = dth[cells]
    adx = np.array([dx[nd, d] for nd, d in coords])
    (t1, x1), (t2, x2) = fd[h], fd[2 * h]
    print("dJ/dtheta rel err", np.abs(t1 - adt) / np.abs(adt), flush=True)
    print("dJ/dX rel err h", np.abs(x1 - adx) / np.abs(adx), flush=True)
    print("dJ/dX rel err 2h", np.abs(x2 - adx) / np.abs(adx), flush=True)
    assert np.all(np.abs(t1 - adt) <= 1e-6 * np.abs(adt)), (t1, adt)
    assert np.all(np.abs(x1 - adx) <= 1e-6 * np.abs(adx) + 1e-9 * np.abs(adx).max()), (x1, adx)
    assert _rel(x1, adx) <= 1e-8 and _rel(t1, adt) <= 1e-9


def test_torch_autograd_function():
    print("case torch autograd function", flush=True)
    import torch
    from ferrite_amd.autodiff import torch_compliance
    g, dh, K, f, ch, prob, theta, ydof = _cantilever(8, 4, seed=3)
    fn = torch_compliance(prob)
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    X = torch.tensor(g.nodes.copy(), dtype=torch.float64, requires_grad=True)
    J = fn(th, X)
    (2.0 * J).backward()
    J2, dth, dx, _ = prob.compliance(theta, g.nodes.copy())
    assert np.allclose(th.grad.numpy(), 2 * dth, rtol=1e-14, atol=0)
    assert np.allclose(X.grad.numpy(), 2 * dx, rtol=1e-14, atol=0)
