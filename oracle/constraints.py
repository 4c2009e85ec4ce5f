"""Oracle: Dirichlet constraints and apply! (TEST INFRASTRUCTURE ONLY, see oracle/__init__.py).

Restates src/Dofs/ConstraintHandler.jl:
  * _add!(ch, dbc, facets): :404-429, local facet dofs :431-444
    ((fdof-1)*field_dim + d + offset for every component d in the BC)
  * close!: :303-366 (prescribed dofs sorted; isconstrained mask)
  * update!: :504-583 (g evaluated at the dof's spatial coordinate)
  * apply!: :710-747 (m = meandiag(K) before zeroing, f -= K g, zero prescribed
    columns and rows, K[d,d] = m, f[d] = g_d m), meandiag :952-958,
    add_inhomogeneities_csc! :755-789, zero_out_columns!/rows! :931-950
  * apply_local!: :1750-1820 (on one element matrix and vector before assembly, Dirichlet
    constraints: fe -= Ke[:, a] g_a per constrained local dof a, m = meandiag(Ke) before
    any change, constrained rows and columns zeroed, Ke[a, a] = m, fe[a] = g_a m)
"""
import numpy as np

from .refel import base_interpolation


class Dirichlet:
    def __init__(self, field, facets, g, components=None):
        self.field = field
        self.facets = list(facets)
        self.g = g
        self.components = components


class ConstraintHandler:
    def __init__(self, dh):
        self.dh = dh
        self.dbcs = []
        self._dofs = {}  # dof -> (dbc index, spatial coordinate)

    def add(self, dbc):
        dh = self.dh
        fi = dh.field_names.index(dbc.field)
        ip = dh.field_ips[fi]
        base = base_interpolation(ip)
        vdim = getattr(ip, "vdim", 1)
        comps = dbc.components or list(range(1, vdim + 1))
        offset = dh.field_offsets[fi]
        facetdofs = base.dirichlet_facetdof_indices()
        refc = base.reference_coordinates()
        from .refel import Lagrange
        geo = Lagrange(base.shape, 1)
        for (cell, facet) in dbc.facets:
            cd = dh.celldofs(cell)
            x = dh.grid.getcoordinates(cell)
            for fdof in facetdofs[facet - 1]:
                # spatial coordinate of the dof: Σ M_j(ξ_dof) x_j
                xi = list(refc[fdof - 1])
                pos = sum(geo.shape_value(xi, j + 1) * x[j] for j in range(x.shape[0]))
                for d in range(1, vdim + 1):
                    if d in comps:
                        ldof = (fdof - 1) * vdim + d + offset
                        self._dofs[cd[ldof - 1]] = (len(self.dbcs), pos, d)
        self.dbcs.append(dbc)
        return self

    def close(self):
        self.prescribed_dofs = np.array(sorted(self._dofs), dtype=np.int64)
        self.isconstrained = np.zeros(self.dh.ndofs + 1, dtype=bool)
        self.isconstrained[self.prescribed_dofs] = True
        self.update(0.0)
        return self

    def update(self, t=0.0):
        vals = []
        for d in self.prescribed_dofs:
            bi, pos, comp = self._dofs[int(d)]
            v = self.dbcs[bi].g(pos, t)
            v = np.atleast_1d(np.asarray(v, dtype=np.float64))
            vals.append(v[comp - 1] if v.size > 1 else v[0])
        self.inhomogeneities = np.array(vals, dtype=np.float64)


def meandiag(K):
    z = 0.0
    for i in range(1, K.n + 1):
        p = K.find(i, i)
        z += abs(K.nzval[p]) if p >= 0 else 0.0
    return z / K.n


def apply(K, f, ch, applyzero=False):
    m = meandiag(K)
    if not applyzero:
        for i, d in enumerate(ch.prescribed_dofs):
            v = ch.inhomogeneities[i]
            if v != 0:
                for p in range(K.colptr[d - 1] - 1, K.colptr[d] - 1):
                    r = K.rowval[p]
                    f[r - 1] -= v * K.nzval[p]
    for d in ch.prescribed_dofs:
        K.nzval[K.colptr[d - 1] - 1:K.colptr[d] - 1] = 0.0
    for p in range(len(K.rowval)):
        if ch.isconstrained[K.rowval[p]]:
            K.nzval[p] = 0.0
    for i, d in enumerate(ch.prescribed_dofs):
        K.nzval[K.find(d, d)] = m
        if len(f) != 0:
            vz = 0.0 if applyzero else ch.inhomogeneities[i]
            f[d - 1] = vz * m
    return m


def apply_local(Ke, fe, dofs, ch, applyzero=False):
    """apply_local!(Ke, fe, global_dofs, ch; apply_zero) in place (dense Ke, 1-based dofs);
    Dirichlet constraints only (affine ones need the global arrays, _condense_local!)."""
    idx = {int(d): i for i, d in enumerate(ch.prescribed_dofs)}
    n = len(dofs)
    has = False
    for a in range(n):
        i = idx.get(int(dofs[a]))
        if i is None:
            continue
        has = True
        v = ch.inhomogeneities[i]
        if not applyzero and v != 0:
            for j in range(n):
                fe[j] -= v * Ke[j, a]
    if not has:
        return
    m = 0.0
    for i in range(n):
        m += abs(Ke[i, i])
    m = m / n
    for a in range(n):
        i = idx.get(int(dofs[a]))
        if i is None:
            continue
        Ke[:, a] = 0.0
        Ke[a, :] = 0.0
        Ke[a, a] = m
        fe[a] = 0.0 if applyzero else ch.inhomogeneities[i] * m


def solve(K, f):
    """u = K \\ f by SuperLU (the reference's `\\` goes to SuiteSparse, heat_equation.jl:218)."""
    import scipy.sparse.linalg as spla
    return spla.spsolve(K.to_scipy().tocsc(), f)
