"""Pins the oracle's elasticity element, facet traction and torch adjoint by invariants (CPU)."""
import numpy as np
import pytest

from oracle.fem import CellValues, FacetValues, elasticity_element, traction_element
from oracle.refel import FacetQuadratureRule, Lagrange, QuadratureRule


def _cell(shape, order, q, rng):
    ip = Lagrange(shape, order)
    dim = ip.rdim
    cv = CellValues(QuadratureRule(shape, q), ip ** dim)
    ref = np.array(Lagrange(shape, 1).reference_coordinates())
    A = np.eye(dim) * 0.6 + rng.uniform(-0.1, 0.1, (dim, dim))
    x = ref @ A.T + rng.uniform(-0.05, 0.05, ref.shape) + 1.0
    return cv, x, dim


@pytest.mark.parametrize("shape,order,q", [("RefQuadrilateral", 1, 2), ("RefHexahedron", 1, 2),
                                           ("RefHexahedron", 2, 3)])
def test_elasticity_symmetry_and_rigid_modes(shape, order, q):
    """Symmetric Ke; rigid translations and rotations are in the null space."""
    rng = np.random.default_rng(1)
    cv, x, dim = _cell(shape, order, q, rng)
    Ke = elasticity_element(cv, x, 1.7, 0.3)
    assert np.allclose(Ke, Ke.T, rtol=0, atol=1e-13 * np.abs(Ke).max())
    # dof coordinates: geometric map of the interpolation's reference nodes
    ip = Lagrange(shape, order)
    geo = Lagrange(shape, 1)
    X = np.array([sum(geo.shape_value(list(xi), j + 1) * x[j] for j in range(len(x)))
                  for xi in ip.reference_coordinates()])
    modes = []
    for d in range(dim):
        m = np.zeros((len(X), dim))
        m[:, d] = 1.0
        modes.append(m.ravel())
    if dim == 2:
        modes.append(np.stack([-X[:, 1], X[:, 0]], 1).ravel())
    else:
        for a, b in ((0, 1), (1, 2), (0, 2)):
            m = np.zeros_like(X)
            m[:, a], m[:, b] = -X[:, b], X[:, a]
            modes.append(m.ravel())
    for m in modes:
        assert np.abs(Ke @ m).max() <= 1e-11 * np.abs(Ke).max() * np.abs(m).max()
    # positive definite on the complement of the rigid modes
    ev = np.linalg.eigvalsh(Ke)
    assert (ev > 1e-10 * ev.max()).sum() == Ke.shape[0] - len(modes)


def test_elasticity_patch_energy():
    """Uniform strain ε0 on a box: uᵀKu = vol · ε0:C:ε0 (closed form)."""
    rng = np.random.default_rng(2)
    ip = Lagrange("RefHexahedron", 2)
    cv = CellValues(QuadratureRule("RefHexahedron", 3), ip ** 3)
    h = np.array([0.5, 0.7, 1.3])
    x = (np.array(Lagrange("RefHexahedron", 1).reference_coordinates()) + 1) / 2 * h
    X = (np.array(ip.reference_coordinates()) + 1) / 2 * h
    eps0 = rng.normal(size=(3, 3))
    eps0 = 0.5 * (eps0 + eps0.T)
    u = X @ eps0.T
    E, nu = 2.0, 0.3
    Ke = elasticity_element(cv, x, E, nu)
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    sig = lam * np.trace(eps0) * np.eye(3) + 2 * mu * eps0
    assert abs(u.ravel() @ Ke @ u.ravel() - h.prod() * np.sum(eps0 * sig)) < 1e-12 * h.prod() * np.abs(sig).max()


def test_traction_total_force():
    """Σ fe = t · area of the facet for a unit traction (FacetValues measure)."""
    ip = Lagrange("RefHexahedron", 2)
    fv = FacetValues(FacetQuadratureRule("RefHexahedron", 3), ip ** 3)
    h = np.array([0.5, 0.7, 1.3])
    x = (np.array(Lagrange("RefHexahedron", 1).reference_coordinates()) + 1) / 2 * h
    fe = traction_element(fv, x, 3, np.array([0.0, 0.0, -1.0]))
    assert abs(fe[2::3].sum() + h[1] * h[2]) < 1e-14
    assert abs(fe[0::3]).max() == 0.0


def test_torch_adjoint_matches_finite_differences():
    torch = pytest.importorskip("torch")
    from oracle.adjoint_torch import compliance_torch
    from oracle.dofs import DofHandler
    from oracle.grid import generate_grid
    g = generate_grid("Quadrilateral", (4, 2), (0.0, 0.0), (2.0, 1.0))
    dh = DofHandler(g).add("u", Lagrange("RefQuadrilateral", 1) ** 2).close()
    cd = np.array(dh.cell_dofs).reshape(g.ncells, 8)
    # clamp x = 0, unit downward load at the tip node (2, 0.5)
    node_dof = {}
    for c in range(g.ncells):
        for a, nd in enumerate(g.cells[c]):
            node_dof[nd] = (cd[c, 2 * a], cd[c, 2 * a + 1])
    fixed = [d for nd, ds in node_dof.items() if abs(g.nodes[nd - 1][0]) < 1e-12 for d in ds]
    free = np.setdiff1d(np.arange(dh.ndofs), np.array(fixed) - 1)
    tip = next(nd for nd in node_dof if np.allclose(g.nodes[nd - 1], [2.0, 0.5]))
    f = np.zeros(dh.ndofs)
    f[node_dof[tip][1] - 1] = -1.0
    rng = np.random.default_rng(4)
    theta = torch.tensor(rng.uniform(0.2, 1.0, g.ncells), dtype=torch.float64, requires_grad=True)
    X = torch.tensor(g.nodes, dtype=torch.float64, requires_grad=True)
    J = compliance_torch(X, g.cells, cd, dh.ndofs, theta, f, free)
    J.backward()
    h = 1e-6
    for e in (0, 3, 7):
        tp, tm = theta.detach().clone(), theta.detach().clone()
        tp[e] += h
        tm[e] -= h
        fd = (compliance_torch(X.detach(), g.cells, cd, dh.ndofs, tp, f, free)
              - compliance_torch(X.detach(), g.cells, cd, dh.ndofs, tm, f, free)) / (2 * h)
        assert abs(fd.item() - theta.grad[e].item()) <= 1e-6 * abs(theta.grad[e].item())
    for nd, d in ((5, 1), (12, 0)):
        xp, xm = X.detach().clone(), X.detach().clone()
        xp[nd, d] += h
        xm[nd, d] -= h
        fd = (compliance_torch(xp, g.cells, cd, dh.ndofs, theta.detach(), f, free)
              - compliance_torch(xm, g.cells, cd, dh.ndofs, theta.detach(), f, free)) / (2 * h)
        assert abs(fd.item() - X.grad[nd, d].item()) <= 1e-6 * max(abs(X.grad[nd, d].item()), 1e-3)
