"""Oracle: ctypes wrapper of the C restatement (TEST INFRASTRUCTURE ONLY, see oracle/__init__.py).

`run()` times the Ferrite-algorithm CPU restatement (OpenMP colored assembly with the
sorted merge-walk scatter, apply!, Jacobi-PCG) on the host cores for the bench's
cpu_baseline leg of the 128³ Poisson workload; `run_heat()`, `run_elasticity()`,
`run_cantilever()` and `run_batch()` do the same for the quad heat problem, the Q2 hex
elasticity assembly, the cantilever compliance with its gradients and the batch of Poisson
problems.  The CPU tests check every entry point against the Python oracle (and the adjoint
against torch autograd).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "libferrite_cpu_oracle.so")
_h = None

I32P, I64P, DP = C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_double)


def lib():
    global _h
    if _h is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        h = C.CDLL(LIB)
        h.cb_assemble_q1hex.restype = C.c_int
        h.cb_assemble_q1hex.argtypes = [C.c_int64, DP, I32P, I32P, DP, C.c_double, C.c_int64, I32P, I32P, C.c_int,
                                        I64P, I32P, C.c_int, DP, DP]
        h.cb_apply_zero_dirichlet.restype = C.c_double
        h.cb_apply_zero_dirichlet.argtypes = [C.c_int64, I32P, I32P, DP, DP, C.c_int64, I32P, C.c_int]
        h.cb_pcg.restype = C.c_int
        h.cb_pcg.argtypes = [C.c_int64, I32P, I32P, DP, DP, DP, C.c_double, C.c_int, C.c_int, DP]
        h.cb_assemble_poisson.restype = C.c_int
        h.cb_assemble_poisson.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, DP, DP, DP, DP, C.c_int64, DP, I32P,
                                          I32P, DP, C.c_double, C.c_int64, I32P, I32P, C.c_int, I64P, I32P, C.c_int,
                                          DP, DP]
        h.cb_assemble_elasticity.restype = C.c_int
        h.cb_assemble_elasticity.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, DP, DP, DP, C.c_int64, DP, I32P,
                                             I32P, DP, C.c_double, C.c_int64, I32P, I32P, C.c_int, I64P, I32P,
                                             C.c_int, DP]
        h.cb_compliance_adjoint.restype = C.c_int
        h.cb_compliance_adjoint.argtypes = [C.c_int, C.c_int, C.c_int, DP, DP, C.c_int64, C.c_int64, DP, I32P, I32P,
                                            DP, C.c_double, C.c_double, C.c_double, DP, C.c_int, I64P, I32P,
                                            C.c_int, DP, DP]
        h.cb_warm.restype = C.c_double
        h.cb_warm.argtypes = [C.c_int]
        h.cb_dot.restype = C.c_double
        h.cb_dot.argtypes = [C.c_int64, DP, DP]
        h.cb_wtime.restype = C.c_double
        h.cb_max_threads.restype = C.c_int
        _h = h
    return _h


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class System:
    """0-based int32 tables of one system (grid, cell dofs, CSC pattern, colours), built from
    1-based Ferrite arrays."""

    def __init__(self, nodes, cells, cell_dofs, colptr, rowval, colors):
        self.xyz = np.ascontiguousarray(nodes, dtype=np.float64)
        self.cells = np.ascontiguousarray(np.asarray(cells) - 1, dtype=np.int32)
        self.dofs = np.ascontiguousarray(np.asarray(cell_dofs) - 1, dtype=np.int32)
        self.colptr = np.ascontiguousarray(np.asarray(colptr) - 1, dtype=np.int32)
        self.rowval = np.ascontiguousarray(np.asarray(rowval) - 1, dtype=np.int32)
        self.n = len(colptr) - 1
        sizes = np.array([len(c) for c in colors], dtype=np.int64)
        self.cptr = np.zeros(len(colors) + 1, dtype=np.int64)
        self.cptr[1:] = np.cumsum(sizes)
        self.ccells = np.ascontiguousarray(np.concatenate([np.asarray(c) - 1 for c in colors]), dtype=np.int32)
        self.ncolors = len(colors)
        # written, not only allocated: the pages are mapped before any timed call, as K and f
        # of an allocated system are
        self.nzval = np.full(len(self.rowval), 0.0)
        self.f = np.full(self.n, 0.0)


_ERR = {-1: "det(J) is not positive", -2: "missing sparsity pattern entry", -3: "tables exceed the C limits"}


class Tables:
    """Shape tables at the quadrature points in the C layout: N[q, i], dN[q, i, k] of the scalar
    function base, dM[q, a, k] of the cell's Lagrange Q1 geometry, w[q] (from oracle.fem.CellValues)."""

    def __init__(self, refshape, order, qorder):
        from .fem import CellValues
        from .refel import Lagrange, QuadratureRule
        cv = CellValues(QuadratureRule(refshape, qorder), Lagrange(refshape, order))
        self.N = np.ascontiguousarray(cv.N.T)
        self.dN = np.ascontiguousarray(cv.dNdxi.transpose(1, 0, 2))
        self.dM = np.ascontiguousarray(cv.dMdxi.transpose(1, 0, 2))
        self.w = np.ascontiguousarray(cv.qr.weights, dtype=np.float64)
        self.nq, self.nb, self.dim = self.dN.shape
        self.ng = self.dM.shape[1]


def assemble_poisson(sys_, tabs, k, source=1.0, threads=0):
    """Scalar Poisson on any Lagrange base (cb_assemble_poisson); K and f into sys_."""
    k = np.ascontiguousarray(k, dtype=np.float64)
    t = tabs
    rc = lib().cb_assemble_poisson(t.dim, t.nb, t.ng, t.nq, _p(t.N, C.c_double), _p(t.dN, C.c_double),
                                   _p(t.dM, C.c_double), _p(t.w, C.c_double), len(sys_.cells),
                                   _p(sys_.xyz, C.c_double), _p(sys_.cells, C.c_int32), _p(sys_.dofs, C.c_int32),
                                   _p(k, C.c_double), source, sys_.n, _p(sys_.colptr, C.c_int32),
                                   _p(sys_.rowval, C.c_int32), sys_.ncolors, _p(sys_.cptr, C.c_int64),
                                   _p(sys_.ccells, C.c_int32), threads, _p(sys_.nzval, C.c_double),
                                   _p(sys_.f, C.c_double))
    if rc != 0:
        raise RuntimeError(_ERR.get(rc, rc))
    return sys_.nzval, sys_.f


def assemble_elasticity(sys_, tabs, E, nu, threads=0):
    """Small-strain elasticity on the vector Lagrange base (cb_assemble_elasticity); K into sys_."""
    E = np.ascontiguousarray(E, dtype=np.float64)
    t = tabs
    rc = lib().cb_assemble_elasticity(t.dim, t.nb, t.ng, t.nq, _p(t.dN, C.c_double), _p(t.dM, C.c_double),
                                      _p(t.w, C.c_double), len(sys_.cells), _p(sys_.xyz, C.c_double),
                                      _p(sys_.cells, C.c_int32), _p(sys_.dofs, C.c_int32), _p(E, C.c_double),
                                      nu, sys_.n, _p(sys_.colptr, C.c_int32), _p(sys_.rowval, C.c_int32),
                                      sys_.ncolors, _p(sys_.cptr, C.c_int64), _p(sys_.ccells, C.c_int32),
                                      threads, _p(sys_.nzval, C.c_double))
    if rc != 0:
        raise RuntimeError(_ERR.get(rc, rc))
    return sys_.nzval


def compliance_adjoint(sys_, tabs, theta, u, nu=0.3, emin=1e-3, penal=3.0, threads=0):
    """(dJ/dθ per cell, dJ/dX per node (nnodes, dim)) of J = fᵀu for the SIMP element."""
    t = tabs
    th = np.ascontiguousarray(theta, dtype=np.float64)
    uu = np.ascontiguousarray(u, dtype=np.float64)
    nn = sys_.xyz.shape[0]
    dth = np.zeros(len(sys_.cells))
    dx = np.zeros((nn, t.dim))
    rc = lib().cb_compliance_adjoint(t.dim, t.ng, t.nq, _p(t.dM, C.c_double), _p(t.w, C.c_double),
                                     len(sys_.cells), nn, _p(sys_.xyz, C.c_double), _p(sys_.cells, C.c_int32),
                                     _p(sys_.dofs, C.c_int32), _p(th, C.c_double), nu, emin, penal,
                                     _p(uu, C.c_double), sys_.ncolors, _p(sys_.cptr, C.c_int64),
                                     _p(sys_.ccells, C.c_int32), threads, _p(dth, C.c_double), _p(dx, C.c_double))
    if rc != 0:
        raise RuntimeError(_ERR.get(rc, rc))
    return dth, dx


def dot(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return lib().cb_dot(len(a), _p(a, C.c_double), _p(b, C.c_double))


def assemble(sys_, k, source=1.0, threads=0):
    k = np.ascontiguousarray(k, dtype=np.float64)
    rc = lib().cb_assemble_q1hex(len(sys_.cells), _p(sys_.xyz, C.c_double), _p(sys_.cells, C.c_int32),
                                 _p(sys_.dofs, C.c_int32), _p(k, C.c_double), source, sys_.n,
                                 _p(sys_.colptr, C.c_int32), _p(sys_.rowval, C.c_int32), sys_.ncolors,
                                 _p(sys_.cptr, C.c_int64), _p(sys_.ccells, C.c_int32), threads,
                                 _p(sys_.nzval, C.c_double), _p(sys_.f, C.c_double))
    if rc != 0:
        raise RuntimeError({-1: "det(J) is not positive", -2: "missing sparsity pattern entry"}.get(rc, rc))
    return sys_.nzval, sys_.f


def apply_zero(sys_, prescribed, threads=0):
    pres = np.ascontiguousarray(np.asarray(prescribed) - 1, dtype=np.int32)
    return lib().cb_apply_zero_dirichlet(sys_.n, _p(sys_.colptr, C.c_int32), _p(sys_.rowval, C.c_int32),
                                         _p(sys_.nzval, C.c_double), _p(sys_.f, C.c_double), len(pres),
                                         _p(pres, C.c_int32), threads)


def pcg(sys_, rtol, maxit=100000, threads=0):
    u = np.zeros(sys_.n)
    rel = C.c_double(0.0)
    it = lib().cb_pcg(sys_.n, _p(sys_.colptr, C.c_int32), _p(sys_.rowval, C.c_int32), _p(sys_.nzval, C.c_double),
                      _p(sys_.f, C.c_double), _p(u, C.c_double), rtol, maxit, threads, C.byref(rel))
    return u, it, rel.value


def run(g, dh, k, colors, colptr, rowval, prescribed, threads=0, rtol=1e-10, max_seconds=60.0, log=print):
    """Time assemble + apply! + PCG for the full grid on the host cores (bounded by max_seconds
    of CG: if the solve would exceed it, the CG part is measured on its first iterations and
    extrapolated to the iteration count the same algorithm needs, stated in `sample`)."""
    h = lib()
    nt = _threads(threads)
    s = System(g.nodes, g.cells, dh.cell_dofs, colptr, rowval, colors)
    t0 = h.cb_wtime()
    assemble(s, k, 1.0, nt)
    t1 = h.cb_wtime()
    apply_zero(s, prescribed, nt)
    t2 = h.cb_wtime()
    # CG: time a bounded number of iterations
    probe = 50
    u, it, rel = pcg(s, rtol, maxit=probe, threads=nt)
    t3 = h.cb_wtime()
    per_it = (t3 - t2) / max(it, 1)
    est_iters = None
    if rel > rtol and per_it * 2000 < max_seconds:
        u, it2, rel = pcg(s, rtol, maxit=100000, threads=nt)
        t3 = h.cb_wtime()
        solve_s = t3 - t2 - per_it * it
        iters = it2
    else:
        solve_s = per_it * it
        iters = it
        est_iters = it
    total = (t1 - t0) + (t2 - t1) + solve_s
    log(f"cpu baseline: assembly {t1 - t0:.3f}s apply {t2 - t1:.3f}s CG {solve_s:.3f}s ({iters} its, "
        f"relres {rel:.2e}) on {nt} threads")
    return {"value": g.ncells / total, "unit": "cells/s", "cores": nt, "kind": "port",
            "sample": f"full {g.ncells}-cell assembly + apply! + Jacobi-PCG to {rtol:g} "
                      f"({iters} iterations{'' if est_iters is None else ' measured, rest extrapolated'}), "
                      f"Ferrite-algorithm CPU restatement in C/OpenMP",
            "assembly_s": t1 - t0, "apply_s": t2 - t1, "solve_s": solve_s, "cg_iterations": iters,
            "assembly_cells_per_s": g.ncells / (t1 - t0)}


def _threads(threads):
    """Thread count (16 at most by default), with the OpenMP pool started before any timing."""
    nt = threads or min(16, os.cpu_count() or 1)
    lib().cb_warm(nt)
    return nt


def _timed_pcg(h, s, rtol, nt, max_seconds):
    """Jacobi-PCG to rtol, bounded: 50 iterations are timed first; if the whole solve would take
    more than max_seconds only those are kept and the iteration count is reported as measured
    ('extrapolated' is then the caller's to state).  Returns (u, iterations, relres, seconds,
    complete)."""
    t0 = h.cb_wtime()
    u, it, rel = pcg(s, rtol, maxit=50, threads=nt)
    t1 = h.cb_wtime()
    per_it = (t1 - t0) / max(it, 1)
    if rel <= rtol:
        return u, it, rel, t1 - t0, True
    if per_it * 20000 > max_seconds:
        return u, it, rel, t1 - t0, False
    t0 = h.cb_wtime()
    u, it, rel = pcg(s, rtol, maxit=200000, threads=nt)
    return u, it, rel, h.cb_wtime() - t0, True


def run_heat(g, dh, colors, colptr, rowval, prescribed, threads=0, rtol=1e-10, reps=20, log=print):
    """Heat equation on a quad grid (k = 1, f = 1): assembly timed as the best of `reps`
    (the grid is small), then apply! and Jacobi-PCG to rtol."""
    h = lib()
    nt = _threads(threads)
    s = System(g.nodes, g.cells, dh.cell_dofs, colptr, rowval, colors)
    tabs = Tables(g.refshape, 1, 2)
    k = np.ones(g.ncells)
    best = float("inf")
    for _ in range(reps):
        t0 = h.cb_wtime()
        assemble_poisson(s, tabs, k, 1.0, nt)
        best = min(best, h.cb_wtime() - t0)
    t0 = h.cb_wtime()
    apply_zero(s, prescribed, nt)
    t_apply = h.cb_wtime() - t0
    u, it, rel, t_cg, _ = _timed_pcg(h, s, rtol, nt, 60.0)
    log(f"cpu baseline (heat): assembly {best * 1e3:.3f} ms, apply {t_apply * 1e3:.3f} ms, "
        f"CG {t_cg * 1e3:.2f} ms ({it} its) on {nt} threads")
    return {"value": g.ncells / best, "unit": "cells/s (assembly)", "cores": nt, "kind": "port",
            "sample": f"{g.ncells}-cell Q1 quad assembly, best of {reps}; apply! and Jacobi-PCG to {rtol:g} "
                      "timed beside it; Ferrite-algorithm CPU restatement in C/OpenMP",
            "assembly_s": best, "apply_s": t_apply, "solve_s": t_cg, "cg_iterations": it, "relres": rel,
            "assemble_solve_cells_per_s": g.ncells / (best + t_apply + t_cg), "u_norm": float(np.linalg.norm(u))}


def run_elasticity(g, dh, E, nu, colors, colptr, rowval, threads=0, max_cells=None, log=print):
    """Q2 hex elasticity assembly (FP64) of K, on the first colours up to `max_cells` cells
    (the whole grid when None): cells/s over the cells assembled."""
    h = lib()
    nt = _threads(threads)
    if max_cells is not None:
        kept, total = [], 0
        for c in colors:
            if total and total + len(c) > max_cells:
                break
            kept.append(c)
            total += len(c)
        colors = kept
    s = System(g.nodes, g.cells, dh.cell_dofs, colptr, rowval, colors)
    tabs = Tables(g.refshape, 2, 3)
    ncells = int(s.cptr[-1])
    t0 = h.cb_wtime()
    assemble_elasticity(s, tabs, E, nu, nt)
    dt = h.cb_wtime() - t0
    log(f"cpu baseline (elasticity): {ncells} cells in {dt:.2f} s on {nt} threads")
    return {"value": ncells / dt, "unit": "cells/s (assembly)", "cores": nt, "kind": "port",
            "sample": f"{ncells} of {g.ncells} cells ({len(colors)} whole colours) of the FP64 Q2 hex elasticity "
                      "assembly into K (zeroing all of nzval included); Ferrite-algorithm CPU restatement in C/OpenMP",
            "assembly_s": dt, "cells_assembled": ncells}


def run_cantilever(g, dh, theta, colors, colptr, rowval, prescribed, ydof, threads=0, rtol=2e-9, nu=0.3, emin=1e-3,
                   penal=3.0, max_seconds=60.0, log=print):
    """Cantilever compliance: assembly with E = emin + θ^p (1 − emin), apply!, the point load,
    Jacobi-PCG to rtol, J = fᵀu and the adjoint.  The CG is bounded by max_seconds: when the
    whole solve would take longer, only its first 50 iterations are timed, the record says so
    (`solve_complete` false) and gives the time per iteration."""
    h = lib()
    nt = _threads(threads)
    s = System(g.nodes, g.cells, dh.cell_dofs, colptr, rowval, colors)
    tabs = Tables(g.refshape, 1, 2)
    E = emin + np.asarray(theta) ** penal * (1.0 - emin)
    t0 = h.cb_wtime()
    assemble_elasticity(s, tabs, E, nu, nt)
    t1 = h.cb_wtime()
    s.f[:] = 0.0
    apply_zero(s, prescribed, nt)
    s.f[ydof - 1] += -1.0
    t2 = h.cb_wtime()
    u, it, rel, t_cg, complete = _timed_pcg(h, s, rtol, nt, max_seconds)
    t3 = h.cb_wtime()
    J = dot(s.f, u)
    dth, dx = compliance_adjoint(s, tabs, theta, u, nu, emin, penal, nt)
    t4 = h.cb_wtime()
    per_it = t_cg / max(it, 1)
    log(f"cpu baseline (cantilever): assembly {t1 - t0:.3f} s, apply {t2 - t1:.3f} s, CG {t_cg:.2f} s ({it} its, "
        f"relres {rel:.2e}{'' if complete else ', stopped early'}), J + adjoint {t4 - t3:.3f} s on {nt} threads")
    return {"unit": "s", "cores": nt, "kind": "port",
            "sample": "full cantilever: assembly, apply!, point load, Jacobi-PCG to "
                      f"{rtol:g}{'' if complete else ' (first 50 iterations timed; the solve would exceed the bound)'}"
                      ", J = fᵀu, dJ/dθ and dJ/dX; Ferrite-algorithm CPU restatement in C/OpenMP",
            "assembly_s": t1 - t0, "apply_s": t2 - t1, "solve_s": t_cg, "cg_iterations": it, "relres": rel,
            "solve_complete": complete, "cg_s_per_iteration": per_it, "J_and_gradients_s": t4 - t3,
            "J": J, "u": u, "dJdtheta": dth, "dJdX": dx}


def run_batch(problems, threads=0, rtol=1e-8, log=print):
    """The batch one problem after another, each with the threaded assembly, apply! and
    Jacobi-PCG: total cells/s and the completion time of every problem from the start."""
    h = lib()
    nt = _threads(threads)
    done, ncells, iters = [], 0, 0
    tabs = {}
    systems = [(System(g.nodes, g.cells, dh.cell_dofs, colptr, rowval, colors),
                tabs.setdefault(g.refshape, Tables(g.refshape, 1, 2)), g.ncells, prescribed, k)
               for (g, dh, colors, colptr, rowval, prescribed, k) in problems]
    t0 = h.cb_wtime()
    for s, t, nc, prescribed, k in systems:
        assemble_poisson(s, t, k, 1.0, nt)
        apply_zero(s, prescribed, nt)
        _, it, rel = pcg(s, rtol, maxit=100000, threads=nt)
        iters += it
        ncells += nc
        done.append((h.cb_wtime() - t0) * 1e3)
    total = done[-1] * 1e-3
    log(f"cpu baseline (batch): {len(problems)} problems, {ncells} cells in {total:.2f} s on {nt} threads")
    return {"value": ncells / total, "unit": "cells/s", "cores": nt, "kind": "port",
            "sample": f"all {len(problems)} problems one after another (systems allocated beforehand), each "
                      f"assembled in FP64, apply!, Jacobi-PCG to {rtol:g}; Ferrite-algorithm CPU restatement in "
                      "C/OpenMP",
            "total_ms": total * 1e3, "p50_ms": float(np.percentile(done, 50)),
            "p99_ms": float(np.percentile(done, 99)), "total_iterations": iters}
