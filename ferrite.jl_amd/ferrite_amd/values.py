"""CellValues / FacetValues of the host interface.

On this path the host side of `CellValues(qr, ip[, ip_geo])` (CellValues.jl:66-90)
is the tabulation of N, ∂N/∂ξ (FunctionValues.jl:62-98) and ∂M/∂ξ
(GeometryMapping.jl:68-98) at the quadrature points; `reinit!` itself
(CellValues.jl:122-140) runs inside the fused device kernel.  These objects
carry the tables to the device.
"""
import numpy as np

from .refel import FacetQuadratureRule, Lagrange, QuadratureRule, base_interpolation, n_components


class CellValues:
    def __init__(self, qr: QuadratureRule, ip, ip_geo=None):
        self.qr = qr
        self.ip = ip
        self.base = base_interpolation(ip)
        self.vdim = n_components(ip)
        self.ip_geo = ip_geo or Lagrange(self.base.shape, 1)
        if self.ip_geo.order != 1:
            raise ValueError("only the linear geometric interpolation is supported on this path")
        self.N, self.dNdxi = self.base.values_and_gradients(qr.points)
        self.M, self.dMdxi = self.ip_geo.values_and_gradients(qr.points)
        self.weights = np.asarray(qr.weights, dtype=np.float64)

    @property
    def nqp(self):
        return len(self.weights)

    def getnbasefunctions(self):
        return self.base.nbasefunctions * self.vdim

    def device_tables(self):
        """(rdim, nqp, nbf, ngeo, vdim, N[q][i], dNdξ[q][i][d], dMdξ[q][j][d], w[q]) contiguous."""
        return (self.base.rdim, self.nqp, self.base.nbasefunctions, self.ip_geo.nbasefunctions, self.vdim,
                np.ascontiguousarray(self.N), np.ascontiguousarray(self.dNdxi),
                np.ascontiguousarray(self.dMdxi), np.ascontiguousarray(self.weights))


class FacetValues:
    """FacetValues(fqr, ip): per-facet tables (FacetValues.jl:39-94)."""

    def __init__(self, fqr: FacetQuadratureRule, ip):
        self.fqr = fqr
        self.ip = ip
        self.base = base_interpolation(ip)
        self.vdim = n_components(ip)
        self.ip_geo = Lagrange(self.base.shape, 1)
        self.N, self.dMdxi, self.weights = [], [], []
        for pts, w in zip(fqr.points, fqr.weights):
            N, _ = self.base.values_and_gradients(pts)
            _, dM = self.ip_geo.values_and_gradients(pts)
            self.N.append(N)
            self.dMdxi.append(dM)
            self.weights.append(np.asarray(w, dtype=np.float64))

    @property
    def nfacets(self):
        return len(self.N)

    @property
    def nqp(self):
        return len(self.weights[0])

    def device_tables(self):
        """(nfq, nfacets, N[f][q][i], dMdξ[f][q][j][d], w[f][q]) contiguous."""
        return (self.nqp, self.nfacets, np.ascontiguousarray(np.stack(self.N)),
                np.ascontiguousarray(np.stack(self.dMdxi)), np.ascontiguousarray(np.stack(self.weights)))
