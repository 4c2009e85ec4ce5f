"""Dirichlet / ConstraintHandler of the host interface.

Mirrors src/Dofs/ConstraintHandler.jl: `Dirichlet(field, facets, f, components)`
(:41-60), `add!` on facet sets (:404-444: local facet dofs (fdof-1)*dim + d +
field offset), `close!` (:303-366: prescribed dofs sorted), `update!`
(:504-583: g evaluated at the dof's spatial coordinate).  `apply!` runs on the
device (device.py).
"""
import numpy as np

from .refel import Lagrange, base_interpolation, n_components


class Dirichlet:
    def __init__(self, field, facets, f, components=None):
        self.field = field
        self.facets = list(facets)
        self.f = f
        self.components = list(components) if components is not None else None


class ConstraintHandler:
    def __init__(self, dh):
        if not dh.closed:
            raise RuntimeError("close! the DofHandler first")
        self.dh = dh
        self.dbcs = []
        self._entries = {}   # dof -> (dbc index, x, component)
        self.closed = False

    def add(self, dbc):
        if self.closed:
            raise RuntimeError("ConstraintHandler is closed")
        dh = self.dh
        fi = dh.field_names.index(dbc.field)
        ip = dh.field_ips[fi]
        base = base_interpolation(ip)
        vdim = n_components(ip)
        comps = dbc.components or list(range(1, vdim + 1))
        offset = dh.field_offsets[fi]
        fdofs = base.facet_dof_indices()
        refc = base.reference_coordinates()
        geo = Lagrange(base.shape, 1)
        if not dbc.facets:
            self.dbcs.append(dbc)
            return self
        cells = np.array([c for c, _ in dbc.facets], dtype=np.int64)
        facets = np.array([f for _, f in dbc.facets], dtype=np.int64)
        xc = dh.grid.nodes[dh.grid.cells[cells - 1] - 1]           # (nf, ngeo, sdim)
        cd = dh.cell_dofs[cells - 1]                                # (nf, ndpc)
        for f in np.unique(facets):
            sel = np.nonzero(facets == f)[0]
            for fdof in fdofs[f - 1]:
                M, _ = geo.values_and_gradients(refc[fdof - 1][None, :])
                pos = np.einsum("j,njd->nd", M[0], xc[sel])
                for d in range(1, vdim + 1):
                    if d in comps:
                        ldof = (fdof - 1) * vdim + d + offset
                        for k, dof in zip(sel, cd[sel, ldof - 1]):
                            self._entries[int(dof)] = (len(self.dbcs), k, fdof, d, f)
        self._pos_cache = None
        self.dbcs.append(dbc)
        return self

    def close(self):
        self.prescribed_dofs = np.array(sorted(self._entries), dtype=np.int64)
        self.closed = True
        self.update(0.0)
        return self

    def update(self, t=0.0):
        """update!(ch, t): inhomogeneities g(x, t) at each prescribed dof's coordinate."""
        dh = self.dh
        ent = [self._entries[int(d)] for d in self.prescribed_dofs]
        vals = np.zeros(len(ent))
        mcache = {}
        for i, (bi, k, fdof, comp, _) in enumerate(ent):
            dbc = self.dbcs[bi]
            key = (bi, fdof)
            if key not in mcache:
                base = base_interpolation(dh.field_ips[dh.field_names.index(dbc.field)])
                M, _ = Lagrange(base.shape, 1).values_and_gradients(base.reference_coordinates()[fdof - 1][None, :])
                mcache[key] = M[0]
            x = mcache[key] @ dh.grid.getcoordinates(dbc.facets[k][0])
            v = np.atleast_1d(np.asarray(dbc.f(x, t), dtype=np.float64))
            vals[i] = v[comp - 1] if v.size > 1 else v[0]
        self.inhomogeneities = vals
        return self
