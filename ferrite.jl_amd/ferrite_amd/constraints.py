"""Dirichlet / AffineConstraint / PeriodicDirichlet / ConstraintHandler of the host interface.

Mirrors src/Dofs/ConstraintHandler.jl: `Dirichlet(field, facets, f, components)`
(:41-60), `add!` on facet sets (:404-444: local facet dofs (fdof-1)*dim + d +
field offset), `AffineConstraint(dof, entries, b)` (:140-150, add! :368-402: a later
constraint on a dof replaces the earlier one), `PeriodicDirichlet(field, facet_pairs,
components)` (:1060-1294: u_mirror = u_image as affine constraints), `close!` (:303-366:
prescribed dofs sorted, nested affine constraints rejected), `update!` (:504-583: g evaluated
at the dof's spatial coordinate; affine constraints get the effective inhomogeneity
b + Σ a g over Dirichlet masters), `ConformityConstraint(field)` (Adaptivity/constraints.jl:
15-101: on a NonConformingGrid, every hanging vertex dof the mean of its masters' dofs, per
component).  `apply!` runs on the device (device.py).
"""
import warnings

import numpy as np

from .refel import Lagrange, base_interpolation, n_components


class Dirichlet:
    def __init__(self, field, facets, f, components=None):
        self.field = field
        self.facets = list(facets)
        self.f = f
        self.components = list(components) if components is not None else None


class AffineConstraint:
    """u[constrained_dof] = Σ a_j u[j] + b over entries (j, a_j) (1-based dofs)."""

    def __init__(self, constrained_dof, entries, b=0.0):
        self.constrained_dof = int(constrained_dof)
        self.entries = [(int(j), float(a)) for j, a in entries]
        self.b = float(b)


class ConformityConstraint:
    """ConformityConstraint(field) (Adaptivity/constraints.jl:15-18): the affine constraints that
    make a linear Lagrange field (scalar or vectorized) conforming on a NonConformingGrid."""

    def __init__(self, field):
        self.field = field


class PeriodicDirichlet:
    """u(x_mirror) = u(x_image) + f(x_mirror) - f(x_image) on matched facet pairs
    ((mirror cell, facet), (image cell, facet)) from collect_periodic_facets, for the given
    components (default all); f = None gives u_mirror = u_image."""

    def __init__(self, field, facet_pairs, f=None, components=None):
        self.field = field
        self.facet_pairs = list(facet_pairs)
        self.f = f
        self.components = list(components) if components is not None else None


def collect_periodic_facets(grid, mirror, image, tol=1e-12):
    """collect_periodic_facets(grid, mset, iset) (ConstraintHandler.jl:1424-1437) for facet sets
    that are translates of each other (the opposite sides of a box): pairs each mirror facet
    with the image facet whose node coordinates are the mirror's shifted by the translation
    between the two sets."""
    ms = grid.facetsets[mirror] if isinstance(mirror, str) else list(mirror)
    im = grid.facetsets[image] if isinstance(image, str) else list(image)
    if len(ms) != len(im):
        raise ValueError("mirror and image facet sets differ in size")
    from .refel import REF_FACETS
    local = REF_FACETS[grid.refshape]

    def centroid(fc):
        c, f = fc
        return grid.nodes[grid.cells[c - 1][np.asarray(local[f - 1]) - 1] - 1].mean(axis=0)
    cm = np.array([centroid(fc) for fc in ms])
    ci = np.array([centroid(fc) for fc in im])
    t = ci.mean(axis=0) - cm.mean(axis=0)
    scale = max(1.0, float(np.abs(grid.nodes).max()))
    key = lambda x: tuple(np.round(x / (scale * 1e-9)).astype(np.int64))  # noqa: E731
    lookup = {key(c): k for k, c in enumerate(ci)}
    pairs = []
    for k, c in enumerate(cm):
        j = lookup.get(key(c + t))
        if j is None or np.abs(ci[j] - c - t).max() > 1e3 * tol * scale:
            raise ValueError(f"no image facet for mirror facet {ms[k]}")
        pairs.append((tuple(ms[k]), tuple(im[j])))
    return pairs


class ConstraintHandler:
    def __init__(self, dh):
        if not dh.closed:
            raise RuntimeError("close! the DofHandler first")
        self.dh = dh
        self.dbcs = []
        self._entries = {}   # dof -> (dbc index, x, component)
        self._affine = {}    # dof -> (entries, b); the later of a Dirichlet and an affine wins
        self.closed = False

    def add(self, dbc):
        if self.closed:
            raise RuntimeError("ConstraintHandler is closed")
        if isinstance(dbc, AffineConstraint):
            self._entries.pop(dbc.constrained_dof, None)
            self._affine[dbc.constrained_dof] = (dbc.entries, dbc.b)
            return self
        if isinstance(dbc, PeriodicDirichlet):
            return self._add_periodic(dbc)
        if isinstance(dbc, ConformityConstraint):
            return self._add_conformity(dbc)
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
                            self._affine.pop(int(dof), None)
        self._pos_cache = None
        self.dbcs.append(dbc)
        return self

    def _facet_dofs(self, field, facet, components):
        """(global dof, component, coordinate) of a facet's dofs (cell, facet), in the order of
        the reference's local facet dofs: facet dof, then component."""
        dh = self.dh
        fi = dh.field_names.index(field)
        ip = dh.field_ips[fi]
        base = base_interpolation(ip)
        vdim = n_components(ip)
        geo = Lagrange(base.shape, 1)
        cell, f = facet
        x = dh.grid.getcoordinates(cell)
        out = []
        for fdof in base.facet_dof_indices()[f - 1]:
            M, _ = geo.values_and_gradients(base.reference_coordinates()[fdof - 1][None, :])
            pos = M[0] @ x
            for d in components or range(1, vdim + 1):
                out.append((int(dh.cell_dofs[cell - 1][(fdof - 1) * vdim + d - 1 + dh.field_offsets[fi]]), d, pos))
        return out

    def _add_periodic(self, pdbc):
        """PeriodicDirichlet's dof bookkeeping (ConstraintHandler.jl:1134-1294) with the facet dofs
        matched by position: mirror -> image per pair in facet order, remapped through an image
        that is already a mirror, the first mapping of a mirror kept, self pairs skipped, chains
        untangled once; every pair becomes AffineConstraint(mirror, [image => 1], b) with
        b = f(x_mirror) - f(x_image) (zero without f).  The legacy corner locking and rotated
        pairs are not provided."""
        dof_map = {}
        where = {}
        for (mf, imf) in pdbc.facet_pairs:
            mdofs = self._facet_dofs(pdbc.field, mf, pdbc.components)
            idofs = self._facet_dofs(pdbc.field, imf, pdbc.components)
            t = np.mean([p for _, _, p in idofs], axis=0) - np.mean([p for _, _, p in mdofs], axis=0)
            scale = max(1.0, float(np.abs(t).max()))
            for idof, comp, ipos in idofs:
                cands = [(md, mp) for md, mc, mp in mdofs
                         if mc == comp and np.abs(mp + t - ipos).max() <= 1e-9 * scale]
                if len(cands) != 1:
                    raise ValueError("periodic facets do not match")
                (cdof, cpos), mdof = cands[0], idof
                where[cdof] = (cpos, comp)
                where[mdof] = (ipos, comp)
                if mdof in dof_map:
                    dof_map[cdof] = dof_map[mdof]
                elif cdof in dof_map or cdof == mdof:
                    continue
                else:
                    dof_map[cdof] = mdof
        remaps = {}
        for k, v in dof_map.items():
            if v in dof_map:
                remaps[k] = remaps.get(v, dof_map[v])
        dof_map.update(remaps)
        if set(dof_map) & set(dof_map.values()):
            raise ValueError("periodic constraints chain beyond one step")
        def g(dof):
            x, comp = where[dof]
            v = np.atleast_1d(np.asarray(pdbc.f(x, 0.0), dtype=np.float64))
            return v[comp - 1] if v.size > 1 else v[0]
        for k, v in dof_map.items():
            self.add(AffineConstraint(k, [(v, 1.0)], 0.0 if pdbc.f is None else -g(v) + g(k)))
        return self

    def _add_conformity(self, cc):
        """add!(ch, ConformityConstraint(field)) (Adaptivity/constraints.jl:20-101): skipped with a
        warning on a conforming grid; on a NonConformingGrid one AffineConstraint per hanging
        node and component, u_h = Σ u_m / (number of masters), b = 0.  Only linear Lagrange
        fields (and their vectorizations) are supported, as in the reference."""
        from ._lib import FEM_ERR_ARG, FemArgumentError
        dh = self.dh
        info = getattr(dh.grid, "conformity_info", None)
        if info is None:
            warnings.warn(f"Trying to add conformity constraint to {cc.field} on a conforming grid. Skipping.")
            return self
        if cc.field not in dh.field_names:
            raise FemArgumentError(FEM_ERR_ARG, f"Field {cc.field} not found in provided dof handler. "
                                                f"Available fields are {dh.field_names}.")
        fi = dh.field_names.index(cc.field)
        ip = dh.field_ips[fi]
        base = base_interpolation(ip)
        if not (isinstance(base, Lagrange) and base.order == 1):
            raise FemArgumentError(FEM_ERR_ARG, "ConformityConstraint supports only linear Lagrange interpolations "
                                                f"(and their vectorizations), got {ip}.")
        vdim = n_components(ip)
        # first dof of every vertex of the field (its components follow)
        vdof = np.zeros(dh.grid.nnodes + 1, dtype=np.int64)
        off = dh.field_offsets[fi]
        for a in range(dh.grid.cells.shape[1]):
            vdof[dh.grid.cells[:, a]] = dh.cell_dofs[:, off + a * vdim]
        for h, masters in info.items():
            w = 1.0 / len(masters)
            for c in range(vdim):
                self.add(AffineConstraint(int(vdof[h]) + c, [(int(vdof[m]) + c, w) for m in masters], 0.0))
        return self

    def close(self):
        dofs = sorted(set(self._entries) | set(self._affine))
        self.prescribed_dofs = np.array(dofs, dtype=np.int64)
        for d, (entries, _) in self._affine.items():
            for (j, _) in entries:
                if j in self._affine:
                    raise ValueError("nested affine constraints currently not supported")
        self.affine_dofs = np.array(sorted(self._affine), dtype=np.int64)
        self.closed = True
        self.update(0.0)
        return self

    @property
    def has_affine(self):
        return len(self._affine) > 0

    def affine_arrays(self):
        """(dofs, coef_ptr, coef_dofs, coef_vals, b) of the affine constraints, dofs ascending."""
        ptr, cdofs, cvals, bs = [0], [], [], []
        for d in self.affine_dofs:
            entries, b = self._affine[int(d)]
            cdofs.extend(j for j, _ in entries)
            cvals.extend(a for _, a in entries)
            ptr.append(len(cdofs))
            bs.append(b)
        return (self.affine_dofs.astype(np.int64), np.array(ptr, dtype=np.int64), np.array(cdofs, dtype=np.int64),
                np.array(cvals, dtype=np.float64), np.array(bs, dtype=np.float64))

    def update(self, t=0.0):
        """update!(ch, t): inhomogeneities g(x, t) at each Dirichlet dof's coordinate, then the
        effective inhomogeneity b + Σ a_j g_j of each affine constraint (Dirichlet masters j)."""
        dh = self.dh
        vals = np.zeros(len(self.prescribed_dofs))
        mcache = {}
        gmap = {}
        for i, d in enumerate(self.prescribed_dofs):
            if int(d) not in self._entries:
                continue
            bi, k, fdof, comp, _ = self._entries[int(d)]
            dbc = self.dbcs[bi]
            key = (bi, fdof)
            if key not in mcache:
                base = base_interpolation(dh.field_ips[dh.field_names.index(dbc.field)])
                M, _ = Lagrange(base.shape, 1).values_and_gradients(base.reference_coordinates()[fdof - 1][None, :])
                mcache[key] = M[0]
            x = mcache[key] @ dh.grid.getcoordinates(dbc.facets[k][0])
            v = np.atleast_1d(np.asarray(dbc.f(x, t), dtype=np.float64))
            vals[i] = v[comp - 1] if v.size > 1 else v[0]
            gmap[int(d)] = vals[i]
        for i, d in enumerate(self.prescribed_dofs):
            if int(d) in self._affine:
                entries, b = self._affine[int(d)]
                h = b
                for (j, a) in entries:
                    if j in gmap:
                        h += gmap[j] * a
                vals[i] = h
        self.inhomogeneities = vals
        return self
