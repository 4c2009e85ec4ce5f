"""Compliance as a differentiable function (the Python counterpart of a ChainRulesCore rrule).

`Compliance(problem)` wraps a linear-elastic problem on the device: the forward pass
assembles K(θ, X), applies the Dirichlet conditions, solves K u = f and returns
J = fᵀu; the backward pass runs the hand-written adjoint kernel (no stored Ke):

    dJ/dθ_e = −p θ_e^{p−1} (1 − E_min) u_eᵀ K⁰_e u_e,   dJ/dX = −uᵀ (∂K/∂X) u.
"""
import numpy as np

from . import device as D


class ComplianceProblem:
    """Holds the device system and the loads of a compliance evaluation."""

    def __init__(self, K, f, cv, colors, ch, loads, emin=1e-3, penal=3.0, nu=0.3, rtol=1e-12):
        self.K, self.f, self.cv, self.colors, self.ch = K, f, cv, colors, ch
        self.loads = loads            # callable(K, f) adding the (θ, X independent) loads
        self.emin, self.penal, self.nu, self.rtol = emin, penal, nu, rtol

    def compliance(self, theta, X=None):
        """Forward: J(θ, X) and the solve info."""
        if X is not None:
            self.K.dh.grid.nodes[...] = X
            self.K.ctx.call("fem_mesh_upload", *_mesh_args(self.K))
            self.K._colors = None
        E = D.simp_modulus(theta, self.emin, self.penal)
        D.assemble_global(D.start_assemble(self.K, self.f), self.cv, self.colors,
                          D.ElasticityElement(E=E, nu=self.nu))
        self.loads(self.K, self.f)
        D.apply(self.K, self.f, self.ch)
        res = D.solve(self.K, self.f, mode="cg", rtol=self.rtol, download=False)
        J, dth, dx = D.compliance_gradient(self.K, theta, self.emin, self.penal, self.nu)
        return J, dth, dx, res


def _mesh_args(K):
    import ctypes as C
    from . import _lib as L
    g = K.dh.grid
    xyz = L.as_f64(g.nodes)
    cells = L.as_i64(g.cells)
    K._keep = (xyz, cells)
    return (g.sdim, g.nnodes, L.ptr(xyz, C.c_double), g.ncells, cells.shape[1], L.ptr(cells, C.c_int64))


def torch_compliance(problem):
    """A torch.autograd.Function computing J(θ, X) with the adjoint backward."""
    import torch

    class _Compliance(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, X):
            th = theta.detach().cpu().double().numpy()
            x = X.detach().cpu().double().numpy()
            J, dth, dx, _ = problem.compliance(th, x)
            ctx.save_for_backward(torch.from_numpy(dth), torch.from_numpy(dx))
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gJ):
            dth, dx = ctx.saved_tensors
            return gJ * dth, gJ * dx

    return _Compliance.apply
