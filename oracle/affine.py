"""Oracle: affine (linear) constraints (TEST INFRASTRUCTURE ONLY, see oracle/__init__.py).

Restates src/Dofs/ConstraintHandler.jl for constraints u_i = Σ_j a_ij u_j + b_i:
  * AffineConstraint :140-150, add! / add_prescribed_dof! :368-402 (a later constraint on
    the same dof overrides the earlier one)
  * close! :303-366 (prescribed dofs sorted; masters that are themselves affinely
    constrained are rejected: "nested affine constraints currently not supported")
  * update! :504-540 (effective inhomogeneity h_i = b_i + Σ_j a_ij g_j over masters j that
    are prescribed by a Dirichlet condition)
  * apply! :710-747 with _condense! :809-867 (the serial column loop, K := Cᵀ K C and
    f := Cᵀ f in place), add_inhomogeneities_csc! :755-789, zero_out_columns!/rows!,
    K[d,d] = meandiag, f[d] = h_d meandiag
  * apply!(u, ch) :686-700 (u[prescribed] = g, then u_i = b_i + Σ a_ij u_j)
  * create_constraint_matrix :889-929 (C, g with a_c = C a_f + g)
and src/Dofs/sparsity_pattern.jl:782-843 (_add_constraint_entries!, keep_constrained = true)
for the pattern, and the dof bookkeeping of PeriodicDirichlet (ConstraintHandler.jl:1134-1294:
mirror -> image map, remap when the image is itself a mirror, skip self-constraints, then one
untangling pass; every remaining pair is AffineConstraint(mirror, [image => 1], 0)).

Pinned by the reference's own tests (test/test_constraints.jl:323-477: CᵀKC on the free
dofs, the condensed solve against the in-place one, the analytical spring solution) in
tests/test_oracle_affine.py.
"""
import numpy as np

from .constraints import meandiag


class AffineConstraint:
    def __init__(self, constrained_dof, entries, b):
        self.constrained_dof = int(constrained_dof)
        self.entries = [(int(d), float(a)) for d, a in entries]
        self.b = float(b)


class AffineHandler:
    """Dirichlet dofs (dof -> value, from a closed oracle ConstraintHandler or a dict) plus
    affine constraints, closed like ch.close!()."""

    def __init__(self, ndofs, dirichlet=None):
        self.ndofs = ndofs
        self._pres = {}          # dof -> (value or None, coefficients or None, b or None)
        for d, v in (dirichlet or {}).items():
            self._pres[int(d)] = (float(v), None, None)

    def add(self, ac):
        self._pres[ac.constrained_dof] = (None, ac.entries, ac.b)
        return self

    def close(self):
        self.prescribed_dofs = np.array(sorted(self._pres), dtype=np.int64)
        self.dofmapping = {int(d): i for i, d in enumerate(self.prescribed_dofs)}
        self.dofcoefficients = [self._pres[int(d)][1] for d in self.prescribed_dofs]
        self.affine_b = [self._pres[int(d)][2] for d in self.prescribed_dofs]
        for coeffs in self.dofcoefficients:
            for (d, _) in coeffs or []:
                i = self.dofmapping.get(d)
                if i is not None and self.dofcoefficients[i]:
                    raise ValueError("nested affine constraints currently not supported")
        self.isconstrained = np.zeros(self.ndofs + 1, dtype=bool)
        self.isconstrained[self.prescribed_dofs] = True
        self.free_dofs = np.array([d for d in range(1, self.ndofs + 1) if not self.isconstrained[d]], dtype=np.int64)
        self.update()
        return self

    def update(self):
        g = [self._pres[int(d)][0] if self._pres[int(d)][1] is None else None for d in self.prescribed_dofs]
        for i, coeffs in enumerate(self.dofcoefficients):
            if coeffs is None:
                continue
            h = self.affine_b[i]
            for (d, w) in coeffs:
                j = self.dofmapping.get(d)
                if j is None:
                    continue
                h += g[j] * w
            g[i] = h
        self.inhomogeneities = np.array(g, dtype=np.float64)

    def coefficients_for_dof(self, dof):
        i = self.dofmapping.get(int(dof))
        return None if i is None else self.dofcoefficients[i]


def _addindex(K, v, r, c):
    p = K.find(r, c)
    if p < 0:
        raise KeyError(f"K[{r}, {c}] is missing in the sparsity pattern")
    K.nzval[p] += v


def condense(K, f, ch):
    """_condense!(K, f, dofcoefficients, dofmapping): the serial column loop, in place."""
    if not any(c for c in ch.dofcoefficients if c):
        return
    for col in range(1, K.n + 1):
        col_coeffs = ch.coefficients_for_dof(col)
        lo, hi = K.colptr[col - 1] - 1, K.colptr[col] - 1
        if col_coeffs is None:
            for a in range(lo, hi):
                kv = K.nzval[a]
                if kv == 0.0:
                    continue
                row_coeffs = ch.coefficients_for_dof(K.rowval[a])
                if row_coeffs is None:
                    continue
                for (d, v) in row_coeffs:
                    _addindex(K, v * kv, d, col)
        else:
            for a in range(lo, hi):
                kv = K.nzval[a]
                if kv == 0.0:
                    continue
                row = K.rowval[a]
                row_coeffs = ch.coefficients_for_dof(row)
                if row_coeffs is None:
                    for (d, v) in col_coeffs:
                        _addindex(K, v * kv, row, d)
                else:
                    for (d1, v1) in row_coeffs:
                        for (d2, v2) in col_coeffs:
                            _addindex(K, v1 * v2 * kv, d1, d2)
            if len(f):
                for (d, v) in col_coeffs:
                    f[d - 1] += f[col - 1] * v
                f[col - 1] = 0.0


def apply(K, f, ch, applyzero=False):
    """apply!(K, f, ch) with affine constraints; returns meandiag(K)."""
    m = meandiag(K)
    if not applyzero and len(f):
        for i, d in enumerate(ch.prescribed_dofs):
            v = ch.inhomogeneities[i]
            if v != 0:
                for p in range(K.colptr[d - 1] - 1, K.colptr[d] - 1):
                    f[K.rowval[p] - 1] -= v * K.nzval[p]
    condense(K, f, ch)
    for d in ch.prescribed_dofs:
        K.nzval[K.colptr[d - 1] - 1:K.colptr[d] - 1] = 0.0
    for p in range(len(K.rowval)):
        if ch.isconstrained[K.rowval[p]]:
            K.nzval[p] = 0.0
    for i, d in enumerate(ch.prescribed_dofs):
        K.nzval[K.find(d, d)] = m
        if len(f):
            f[d - 1] = (0.0 if applyzero else ch.inhomogeneities[i]) * m
    return m


def apply_vector(u, ch, applyzero=False):
    """apply!(u, ch) / apply_zero!(u, ch)."""
    u[ch.prescribed_dofs - 1] = 0.0 if applyzero else ch.inhomogeneities
    for i, d in enumerate(ch.prescribed_dofs):
        coeffs = ch.dofcoefficients[i]
        if coeffs is None:
            continue
        v = 0.0 if applyzero else ch.affine_b[i]
        for (dm, s) in coeffs:
            v += s * u[dm - 1]
        u[d - 1] = v
    return u


def constraint_matrix(ch):
    """create_constraint_matrix(ch): C (ndofs x nfree) and g with a = C a_f + g."""
    import scipy.sparse as sp
    I, J, V = [], [], []
    col_of = {int(d): j for j, d in enumerate(ch.free_dofs)}
    for j, d in enumerate(ch.free_dofs):
        I.append(d - 1); J.append(j); V.append(1.0)
    for i, d in enumerate(ch.prescribed_dofs):
        for (dm, s) in ch.dofcoefficients[i] or []:
            I.append(d - 1); J.append(col_of[dm]); V.append(s)
    g = np.zeros(ch.ndofs)
    g[ch.prescribed_dofs - 1] = ch.inhomogeneities
    C = sp.csr_matrix((V, (I, J)), shape=(ch.ndofs, len(ch.free_dofs)))
    return C, g


def add_constraint_entries(colptr, rowval, ch):
    """_add_constraint_entries!(sp, ch; keep_constrained = true) on a 1-based CSC pattern
    (structurally symmetric: column c lists the rows r with (r, c) present)."""
    n = len(colptr) - 1
    rows = [set(rowval[colptr[c] - 1:colptr[c + 1] - 1]) for c in range(n)]   # column c+1's rows
    new = {}
    for col in range(1, n + 1):
        col_coeffs = ch.coefficients_for_dof(col)
        for row in rows[col - 1]:
            row_coeffs = ch.coefficients_for_dof(row)
            if row_coeffs is None and col_coeffs is not None:
                for (c2, _) in col_coeffs:
                    new.setdefault(c2, set()).add(row)
            elif row_coeffs is not None and col_coeffs is None:
                for (r2, _) in row_coeffs:
                    new.setdefault(col, set()).add(r2)
            elif row_coeffs is not None and col_coeffs is not None:
                for (r2, _) in row_coeffs:
                    for (c2, _) in col_coeffs:
                        new.setdefault(c2, set()).add(r2)
    for c, rs in new.items():
        rows[c - 1] |= rs
    cp = np.zeros(n + 1, dtype=np.int64)
    cp[0] = 1
    rv = []
    for c in range(n):
        col = sorted(rows[c])
        rv.extend(col)
        cp[c + 1] = cp[c] + len(col)
    return cp, np.array(rv, dtype=np.int64)


def periodic_dof_map(pairs):
    """Mirror -> image dof map of PeriodicDirichlet from (mirror dof, image dof) pairs in
    facet order: remap through an image that is already a mirror, keep the first mapping of
    a mirror, skip self-constraints, then untangle chains once."""
    dof_map = {}
    for cdof, mdof in pairs:
        if mdof in dof_map:
            dof_map[cdof] = dof_map[mdof]
        elif cdof in dof_map:
            pass
        elif cdof == mdof:
            pass
        else:
            dof_map[cdof] = mdof
    remaps = {}
    for k, v in dof_map.items():
        if v in dof_map:
            remaps[k] = remaps.get(v, dof_map[v])
    dof_map.update(remaps)
    assert not (set(dof_map) & set(dof_map.values()))
    return dof_map
