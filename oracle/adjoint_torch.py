"""Oracle: compliance and its gradients by torch float64 autograd (TEST INFRASTRUCTURE ONLY).

An independent restatement of the plane-strain Q1 cantilever of the differentiation
configuration (docs/src/literate-tutorials/topology_optimization.jl:75-127, :320-383
for the SIMP element, here with E_e = E_min + θ_e^p (1 − E_min) and plane strain):
dense K(θ, X) assembled with torch ops, clamped dofs eliminated, u = K_ff⁻¹ f_f,
J = fᵀu.  torch.autograd gives dJ/dθ and dJ/dX, the independent check of the
hand-written adjoint (ForwardDiff through Ferrite is not available here).
"""
import math

import numpy as np
import torch


def compliance_torch(nodes, cells, cell_dofs, ndofs, theta, fvec, free, nu=0.3, emin=1e-3, penal=3.0):
    """nodes (nn, 2) and theta (nc,) are torch float64 tensors (may require grad)."""
    g = 1.0 / math.sqrt(3.0)
    pts = [(-g, -g), (g, -g), (-g, g), (g, g)]
    sx = [-1.0, 1.0, 1.0, -1.0]
    sy = [-1.0, -1.0, 1.0, 1.0]
    lam = nu / ((1 + nu) * (1 - 2 * nu))
    mu = 1.0 / (2 * (1 + nu))
    K = torch.zeros((ndofs, ndofs), dtype=torch.float64)
    cells_t = torch.as_tensor(np.asarray(cells) - 1)
    X = nodes[cells_t]                     # (nc, 4, 2)
    E = emin + theta ** penal * (1.0 - emin)
    Ke = torch.zeros((len(cells), 8, 8), dtype=torch.float64)
    for (xi, eta) in pts:
        dN = torch.tensor([[sx[a] * (1 + sy[a] * eta) / 4, sy[a] * (1 + sx[a] * xi) / 4] for a in range(4)],
                          dtype=torch.float64)              # (4, 2) dN/dξ
        J = torch.einsum("cai,aj->cij", X, dN)              # J_ij = Σ x_a,i dN_a/dξ_j
        det = J[:, 0, 0] * J[:, 1, 1] - J[:, 0, 1] * J[:, 1, 0]
        Jinv = torch.stack([torch.stack([J[:, 1, 1], -J[:, 0, 1]], -1),
                            torch.stack([-J[:, 1, 0], J[:, 0, 0]], -1)], -2) / det[:, None, None]
        G = torch.einsum("al,clk->cak", dN, Jinv)           # ∇N_a (nc, 4, 2)
        B = torch.zeros((len(cells), 3, 8), dtype=torch.float64)
        B[:, 0, 0::2] = G[:, :, 0]
        B[:, 1, 1::2] = G[:, :, 1]
        B[:, 2, 0::2] = G[:, :, 1]
        B[:, 2, 1::2] = G[:, :, 0]
        D = torch.tensor([[lam + 2 * mu, lam, 0], [lam, lam + 2 * mu, 0], [0, 0, mu]], dtype=torch.float64)
        Ke = Ke + torch.einsum("cki,kl,clj->cij", B, D, B) * det[:, None, None]
    Ke = Ke * E[:, None, None]
    cd = torch.as_tensor(np.asarray(cell_dofs) - 1)
    idx_r = cd[:, :, None].expand(-1, -1, 8).reshape(-1)
    idx_c = cd[:, None, :].expand(-1, 8, -1).reshape(-1)
    K = K.index_put((idx_r, idx_c), Ke.reshape(-1), accumulate=True)
    fr = torch.as_tensor(fvec, dtype=torch.float64)
    fi = torch.as_tensor(free)
    Kff = K[fi][:, fi]
    uf = torch.linalg.solve(Kff, fr[fi])
    return torch.dot(fr[fi], uf)
