"""Allocation-free apply! (fem_apply_dirichlet_noalloc) against fem_apply_dirichlet.

Both entry points run on identical systems in separate contexts; the allocation-free one
must give the same K.nzval, f and meandiag bit for bit.  Two systems in a row, the second
smaller, so the second call runs on the device scratch the first one left behind.  The
larger system carries inhomogeneous Dirichlet values (the path whose dense g vector the
allocating entry point allocates per call), the smaller one runs apply! and apply_zero!.
"""
import ctypes as C

import numpy as np
import pytest

import ferrite_amd as fa

pytestmark = pytest.mark.gpu


def _system(nel, gfun):
    g = fa.generate_grid("Hexahedron", nel)
    ip = fa.Lagrange("RefHexahedron", 1)
    cv = fa.CellValues(fa.QuadratureRule("RefHexahedron", 2), ip)
    dh = fa.DofHandler(g).add("u", ip).close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh)
    f = fa.allocate_vector(K)
    k = np.exp(0.5 * np.random.default_rng(7).standard_normal(g.ncells))
    fa.assemble_global(fa.start_assemble(K, f), cv, fa.create_coloring(g), fa.PoissonElement(k=k))
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.Dirichlet("u", g.facetsets["left"] + g.facetsets["top"], gfun)).close()
    p = np.ascontiguousarray(ch.prescribed_dofs, dtype=np.int64)
    gv = np.ascontiguousarray(ch.inhomogeneities, dtype=np.float64)
    K.ctx.call("fem_constraints_upload", len(p), p.ctypes.data_as(C.POINTER(C.c_int64)),
               gv.ctypes.data_as(C.POINTER(C.c_double)))
    return K, f


def _compare(nel, gfun, applyzero):
    print("case noalloc apply", nel, "applyzero", applyzero, flush=True)
    K1, f1 = _system(nel, gfun)
    K2, f2 = _system(nel, gfun)
    assert np.array_equal(K1.nzval(), K2.nzval()) and np.array_equal(f1.to_numpy(), f2.to_numpy())
    m1, m2 = C.c_double(0.0), C.c_double(0.0)
    K1.ctx.call("fem_apply_dirichlet", applyzero, C.byref(m1))
    K2.ctx.call("fem_apply_dirichlet_noalloc", applyzero, C.byref(m2))
    assert m1.value == m2.value
    assert np.array_equal(K1.nzval(), K2.nzval())
    assert np.array_equal(f1.to_numpy(), f2.to_numpy())
    return K2, f2


def test_noalloc_apply_matches_two_calls_second_smaller():
    inhom = lambda x, t: 1.0 + x[0] * x[1] - 0.5 * x[2]  # noqa: E731
    K, f = _compare((14, 11, 9), inhom, 0)
    fv = f.to_numpy()
    assert np.count_nonzero(fv) > 0
    _compare((5, 4, 3), lambda x, t: 0.0, 0)
    _compare((5, 4, 3), inhom, 1)
