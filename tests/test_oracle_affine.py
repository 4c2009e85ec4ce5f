"""The oracle's affine constraints pinned by the reference's own tests
(test/test_constraints.jl:323-477), restated on a chain of 1D two-node elements (the
reference's generate_grid(Line, (n,)) with Lagrange{RefLine, 1}: dof i = node i):

  * apply!(K, ch) on a dense K leaves K[free, free] == Cᵀ K C bit for bit (small integer
    values, so every sum is exact);
  * for four sets of affine constraints (with a Dirichlet dof), the in-place route apply!,
    K \\ f, apply!(a, ch) equals the explicit condensation C (Cᵀ K C)⁻¹ Cᵀ (f − K g) + g;
  * a spring chain with u_1 = a u_3 + b and u_2 = 0 has the closed-form solution
    u_3 = (f/k − b) / (1 + a), u_1 = a u_3 + b;
and by the PeriodicDirichlet bookkeeping on hand-made pairs.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle.affine import (AffineConstraint, AffineHandler, add_constraint_entries, apply, apply_vector,
                           constraint_matrix, periodic_dof_map)
from oracle.fem import CSC
from oracle.sparsity import create_pattern_fast


def _chain(ncells):
    cd = np.array([[i + 1, i + 2] for i in range(ncells)])
    return cd, ncells + 1


def test_nonsymmetric_condensation_exact():
    cd, n = _chain(10)
    ch = AffineHandler(n).add(AffineConstraint(1, [(3, 2.0)], 0.0)).add(AffineConstraint(2, [(4, 3.0)], 0.0)).close()
    C, _ = constraint_matrix(ch)
    dense = np.arange(1.0, n * n + 1).reshape(n, n, order="F")      # reshape(1.0:n^2, n, n)
    cp = np.arange(0, n * n + 1, n, dtype=np.int64) + 1
    rv = np.tile(np.arange(1, n + 1, dtype=np.int64), n)
    K = CSC(n, cp, rv, dense.ravel(order="F").copy())
    Kc = (C.T @ dense @ C)
    apply(K, np.zeros(0), ch, applyzero=True)
    Kd = K.to_scipy().toarray()
    fd = ch.free_dofs - 1
    assert np.array_equal(Kd[np.ix_(fd, fd)], Kc)


ACS = [
    [AffineConstraint(4, [(7, 1.0)], 0.0)],
    [AffineConstraint(2, [(5, 1.0), (6, 2.0)], 1.0)],
    [AffineConstraint(2, [(9, 1.0)], 0.0), AffineConstraint(3, [(9, 1.0)], 0.0)],
    [AffineConstraint(2, [(7, 3.0), (8, 1.0)], -1.0), AffineConstraint(4, [(9, -1.0)], 2.0)],
]


@pytest.mark.parametrize("acs", range(len(ACS)))
def test_in_place_equals_explicit_condensation(acs):
    cd, n = _chain(10)
    ch = AffineHandler(n, {1: 0.0})        # Dirichlet on 'left' (node 1)
    for ac in ACS[acs]:
        ch.add(ac)
    ch.close()
    C, g = constraint_matrix(ch)
    cp, rv = create_pattern_fast(cd, n)
    cp, rv = add_constraint_entries(cp, rv, ch)       # allocate_matrix(dh, ch)
    K = CSC(n, cp, rv)
    for c in cd:
        for i in range(2):
            for j in range(2):
                K.nzval[K.find(c[i], c[j])] += 2.0 * (1 if i == j else -1)
    f = np.zeros(n)
    f[-1] = 1.0
    Ks = K.to_scipy().tocsc()
    aa = C @ spla.spsolve((C.T @ Ks @ C).tocsc(), C.T @ (f - Ks @ g)) + g
    apply(K, f, ch)
    a = apply_vector(spla.spsolve(K.to_scipy().tocsc(), f), ch)
    assert np.allclose(a, aa, rtol=1.5e-8, atol=0)


def test_spring_chain_closed_form():
    k, fval, aa, bb = 1.0, 1.0, 1.0, 0.2
    cd, n = _chain(2)
    ch = AffineHandler(n, {2: 0.0}).add(AffineConstraint(1, [(3, aa)], bb)).close()
    cp, rv = create_pattern_fast(cd, n)
    cp, rv = add_constraint_entries(cp, rv, ch)
    K = CSC(n, cp, rv)
    for c in cd:
        for i in range(2):
            for j in range(2):
                K.nzval[K.find(c[i], c[j])] += k * (1 if i == j else -1)
    f = np.zeros(n)
    f[-1] = fval
    apply(K, f, ch)
    u = apply_vector(spla.spsolve(K.to_scipy().tocsc(), f), ch)
    a3 = (fval / k - bb) / (1 + aa)
    assert np.allclose(u, [aa * a3 + bb, 0.0, a3], rtol=1e-14, atol=1e-15)


def test_nested_affine_rejected():
    ch = AffineHandler(5).add(AffineConstraint(1, [(2, 1.0)], 0.0)).add(AffineConstraint(2, [(3, 1.0)], 0.0))
    with pytest.raises(ValueError, match="nested"):
        ch.close()


def test_periodic_bookkeeping():
    # corner: 1 -> 2 and later 2 -> 3 untangle to 1 -> 3, 2 -> 3; a repeated mirror keeps its
    # first image; a self pair is skipped
    m = periodic_dof_map([(1, 2), (2, 3), (1, 5), (4, 4), (6, 1)])
    assert m == {1: 3, 2: 3, 6: 3}
