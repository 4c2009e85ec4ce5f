"""Batched assembly + solve (one persistent work-queue kernel, batched mixed CG) vs the oracle.

Every problem of a batch must match the oracle's serial loop for that problem alone
(heat_equation.jl:186-218 with per-cell k): FP64 element path K and f within 1e-12
(before and after apply!), FP32 element path within 2e-6, prescribed rows/cols exact,
u within the solver tolerance of the oracle's direct solve.  At the full 64-problem
size the checks are size-independent: every problem's FP64 residual ||f - K u||/||f||
meets rtol on the downloaded system, and runs are bitwise reproducible across launch
shapes.
"""
import numpy as np
import pytest

from ferrite_amd.batch import BatchedPoisson, conductivity, make_batch, poisson_problem
from oracle import constraints as OC
from oracle.problems import assemble_poisson, setup_poisson

pytestmark = pytest.mark.gpu

SMALL = [("Quadrilateral", (9, 7)), ("Hexahedron", (4, 3, 5)), ("Quadrilateral", (25, 18)),
         ("Hexahedron", (6, 6, 5)), ("Quadrilateral", (2, 2)), ("Hexahedron", (8, 7, 6))]


def _relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


_oracle = {}


def _oracle_problem(i, ct, nel, k):
    if i not in _oracle:
        grid, cv, dh, K, f, ch = setup_poisson(ct, nel)
        assemble_poisson(grid, cv, dh, K, f, k_cell=k, fsrc=1.0)
        K0, f0 = K.nzval.copy(), f.copy()
        OC.apply(K, f, ch)
        u = OC.solve(K, f)
        _oracle[i] = (K0, f0, K, f, u, ch)
    return _oracle[i]


def _small_batch():
    rng = np.random.default_rng(7)
    b = BatchedPoisson(0)
    ks = []
    for ct, nel in SMALL:
        k = conductivity(int(np.prod(nel)), rng)
        g, dh, colors, ch, k = poisson_problem(ct, nel, k)
        b.add(g, dh, colors, ch, k)
        ks.append(k)
    b.finalize()
    return b, ks


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 2e-6)])
def test_small_batch_assembly_matches_oracle(precision, tol):
    print("case batch-assembly", precision, flush=True)
    b, ks = _small_batch()
    for cpw, chunk in ((64, 64), (128, 256)):
        b.assemble(precision, cpw, chunk)
        for i, (ct, nel) in enumerate(SMALL):
            K0, f0, *_ = _oracle_problem(i, ct, nel, ks[i])
            nz, f, _ = b.download(i)
            assert _relerr(nz, K0) <= tol, (i, cpw, chunk)
            assert np.max(np.abs(nz - K0)) <= tol * np.max(np.abs(K0))
            assert _relerr(f, f0) <= tol
    b.close()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 2e-6)])
def test_small_batch_apply_and_solve(precision, tol):
    print("case batch-solve", precision, flush=True)
    b, ks = _small_batch()
    res = b.run(precision, rtol=1e-11, cells_per_wg=64, chunk=128)
    assert res.converged == len(SMALL)
    assert np.all(res.completion_ms > 0) and np.all(res.iterations > 0)
    for i, (ct, nel) in enumerate(SMALL):
        _, _, K, f, u_ref, ch = _oracle_problem(i, ct, nel, ks[i])
        nz, fd, u = b.download(i)
        pres = ch.prescribed_dofs - 1
        # apply!: prescribed rows/cols zero except the diagonal, which holds meandiag
        assert _relerr(nz, K.nzval) <= tol
        assert np.array_equal(fd[pres], f[pres])
        assert _relerr(fd, f) <= tol
        assert np.all(u[pres] == 0.0)
        assert _relerr(u, u_ref) <= (1e-8 if precision == "fp64" else 2e-5), i
        # the FP64 residual of the device system itself
        import scipy.sparse as sp
        A = sp.csc_matrix((nz, K.rowval - 1, K.colptr - 1), shape=(K.n, K.n))
        assert np.linalg.norm(fd - A @ u) <= 1e-11 * np.linalg.norm(fd) * 1.0001
    b.close()


def test_batch_with_a_distorted_problem():
    """One problem with a moved interior node (a general quadrilateral mesh) in the batch:
    every problem still meets rtol on its own downloaded system."""
    print("case batch-fallback", flush=True)
    import scipy.sparse as sp
    rng = np.random.default_rng(3)
    b = BatchedPoisson(0)
    for j, (ct, nel) in enumerate(SMALL[:3]):
        k = conductivity(int(np.prod(nel)), rng)
        g, dh, colors, ch, k = poisson_problem(ct, nel, k)
        if j == 2:
            inner = np.nonzero(np.all(np.abs(g.nodes) < 0.99, axis=1))[0][0]
            g.nodes[inner] += 0.01
        b.add(g, dh, colors, ch, k)
    b.finalize()
    res = b.run("mixed", rtol=1e-10)
    assert res.converged == 3
    for i in range(3):
        p = b.problems[i]
        nz, f, u = b.download(i)
        A = sp.csc_matrix((nz, p["rowval"] - 1, p["colptr"] - 1), shape=(p["ndofs"], p["ndofs"]))
        assert np.linalg.norm(f - A @ u) <= 1e-10 * np.linalg.norm(f) * 1.0001
    b.close()


def test_small_batch_error_paths():
    print("case batch-errors", flush=True)
    import ferrite_amd as fa
    b = BatchedPoisson(0)
    g, dh, colors, ch, k = poisson_problem("Quadrilateral", (4, 4), np.ones(16))
    # a colouring that puts neighbours in one colour is rejected (the scatter is atomic-free)
    with pytest.raises(fa.FemArgumentError):
        b.add(g, dh, [np.arange(1, 17)], ch, k)
    with pytest.raises(ValueError):
        b.add(g, dh, colors, ch, np.ones(3))
    # inverted cell -> det(J) error at assembly (reinit!'s check)
    g2, dh2, colors2, ch2, k2 = poisson_problem("Quadrilateral", (3, 3), np.ones(9))
    g2.nodes[:, 0] *= -1.0
    b.add(g2, dh2, colors2, ch2, k2)
    b.finalize()
    with pytest.raises(fa.FemArgumentError):
        b.assemble("fp64", 64, 64)
    b.close()


def test_full_batch_converges_and_is_reproducible():
    """64 problems (2D/3D, 2e3..2e5 cells), FP32 element math, batched mixed CG to 1e-8."""
    print("case batch-64", flush=True)
    b, setups = make_batch(64, seed=0)
    cpw, chunk = b.autotune("mixed")
    print("autotuned", cpw, chunk, flush=True)
    r1 = b.run("mixed", rtol=1e-8)
    assert r1.converged == 64
    print(f"batch-64 total {r1.total_ms:.2f} ms assemble {r1.assemble_ms:.3f} ms "
          f"p50 {r1.percentile(50):.2f} p99 {r1.percentile(99):.2f}", flush=True)
    import scipy.sparse as sp
    big2 = max((i for i, s in enumerate(setups) if s[0] == "Quadrilateral"), key=lambda i: np.prod(setups[i][1]))
    big3 = max((i for i, s in enumerate(setups) if s[0] == "Hexahedron"), key=lambda i: np.prod(setups[i][1]))
    small = min(range(64), key=lambda i: np.prod(setups[i][1]))
    keep = {}
    for i in (big2, big3, small, 17):
        p = b.problems[i]
        nz, f, u = b.download(i)
        A = sp.csc_matrix((nz, p["rowval"] - 1, p["colptr"] - 1), shape=(p["ndofs"], p["ndofs"]))
        assert np.linalg.norm(f - A @ u) <= 1e-8 * np.linalg.norm(f), i
        keep[i] = (nz, u)
    # another launch shape: bitwise-identical assembly (the colour order fixes every sum) and solve
    r2 = b.run("mixed", rtol=1e-8, cells_per_wg=64 if cpw != 64 else 128, chunk=1024)
    assert r2.converged == 64
    for i, (nz, u) in keep.items():
        nz2, _, u2 = b.download(i)
        assert np.array_equal(nz, nz2) and np.array_equal(u, u2), i
    # FP64 element math on the same batch: the FP32-tuned shape is replaced when it does
    # not fit the FP64 staging (256 cells per workgroup)
    r3 = b.run("fp64", rtol=1e-8)
    assert r3.converged == 64
    for i, (nz, u) in keep.items():
        nz3, _, u3 = b.download(i)
        assert np.linalg.norm(nz3 - nz) <= 1e-5 * np.linalg.norm(nz3), i
        assert np.linalg.norm(u3 - u) <= 1e-4 * np.linalg.norm(u3), i   # FP32 element rounding
    b.close()


@pytest.mark.skip(reason="the batched solver's per-problem cell-coefficient operator (k_bc_cell) is "
                         "not built: under rocprofv3's kernel trace its first solve ended with an "
                         "unspecified launch failure, twice, with no cause found in its indexing "
                         "(checked on the CPU over the 64-problem batch); it was removed")
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 2e-6)])
def test_small_batch_cell_operator(precision, tol):
    """The per-problem cell coefficients (every problem a box lattice) against the oracle."""
    b, ks = _small_batch()
    res = b.run(precision, rtol=1e-11, cells_per_wg=64, chunk=128, operator="cells")
    assert res.operator == "cells" and res.converged == len(SMALL)
    for i, (ct, nel) in enumerate(SMALL):
        *_, u_ref, _ = _oracle_problem(i, ct, nel, ks[i])
        _, _, u = b.download(i)
        assert _relerr(u, u_ref) <= (1e-8 if precision == "fp64" else 2e-5), i
    b.close()
