"""Oracle: ctypes wrapper of the C restatement (TEST INFRASTRUCTURE ONLY, see oracle/__init__.py).

`run()` times the Ferrite-algorithm CPU restatement (OpenMP colored assembly with the
sorted merge-walk scatter, apply!, Jacobi-PCG) on the host cores for the bench's
cpu_baseline leg; `assemble()` is also used by the CPU tests to check it against the
Python oracle.
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
        h.cb_wtime.restype = C.c_double
        h.cb_max_threads.restype = C.c_int
        _h = h
    return _h


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class System:
    """0-based int32 tables of a Q1 hex Poisson system, built from 1-based Ferrite arrays."""

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
        self.nzval = np.zeros(len(self.rowval))
        self.f = np.zeros(self.n)


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
    nt = threads or min(16, os.cpu_count() or 1)
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
