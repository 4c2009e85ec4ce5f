"""ctypes binding of libferrite_mi355x.so (the C ABI declared in include/ferrite_mi355x*.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (make -C
ferrite.jl_amd/csrc).  There is no fallback: if it is missing, every entry
point raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libferrite_mi355x.so")

FEM_OK, FEM_ERR_ARG, FEM_ERR_HIP, FEM_ERR_PATTERN, FEM_ERR_DETJ, FEM_ERR_STATE, FEM_ERR_NOCONV, FEM_ERR_OOM = range(8)
FEM_KERNEL_POISSON, FEM_KERNEL_ELASTICITY = 1, 2
FEM_FP64, FEM_MIXED = 0, 1
FEM_SOLVE_CG_FP64, FEM_SOLVE_IR_MIXED = 0, 1
FEM_REF_LINE, FEM_REF_QUADRILATERAL, FEM_REF_HEXAHEDRON = 1, 2, 3
FEM_COLORING_GREEDY, FEM_COLORING_WORKSTREAM = 0, 1


class FemError(RuntimeError):
    """A failed library call; ``code`` is the FEM_ERR_* status."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class FemArgumentError(FemError, ValueError):
    """Ferrite's ArgumentError (e.g. det(J) <= 0, length mismatch)."""


class FemPatternError(FemError):
    """Ferrite's ErrorException for an entry missing in the sparsity pattern."""


class FemConvergenceError(FemError):
    pass


class AssembleCfg(C.Structure):
    _fields_ = [("kernel", C.c_int32), ("precision", C.c_int32), ("fillzero", C.c_int32),
                ("local_constraints", C.c_int32), ("source", C.c_double), ("nu", C.c_double)]


class SolveCfg(C.Structure):
    _fields_ = [("mode", C.c_int32), ("maxit", C.c_int32), ("rtol", C.c_double),
                ("inner_rtol", C.c_double), ("precond", C.c_int32), ("reserved", C.c_int32)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("outer_iterations", C.c_int32), ("relres", C.c_double),
                ("operator_format", C.c_int32), ("operator_shape", C.c_int32)]


class Tuning(C.Structure):
    _fields_ = [("spmv_cap64", C.c_int32), ("spmv_lanes64", C.c_int32), ("sell_compress", C.c_int32),
                ("asm_block", C.c_int32), ("asm_mode", C.c_int32), ("time_kernels", C.c_int32),
                ("sell_unroll", C.c_int32), ("sell_map", C.c_int32), ("gather_group", C.c_int32),
                ("cg_elems", C.c_int32), ("sell_nt", C.c_int32),
                ("affine", C.c_int32), ("lattice", C.c_int32), ("lattice_blocks", C.c_int32),
                ("cell_operator", C.c_int32), ("cell_blocks", C.c_int32),
                ("cell_rows", C.c_int32), ("elem_operator", C.c_int32)]


class Timings(C.Structure):
    _fields_ = [("assemble_kernel_ms", C.c_double), ("fill_ms", C.c_double), ("apply_ms", C.c_double),
                ("spmv_ms", C.c_double), ("spmv_launches", C.c_int32), ("assemble_launches", C.c_int32),
                ("solve_ms", C.c_double), ("element_ms", C.c_double), ("scatter_ms", C.c_double)]


class BatchCfg(C.Structure):
    _fields_ = [("precision", C.c_int32), ("cells_per_wg", C.c_int32), ("chunk", C.c_int32),
                ("maxit", C.c_int32), ("rtol", C.c_double), ("ru_delta", C.c_double)]


class BatchInfo(C.Structure):
    _fields_ = [("assemble_ms", C.c_double), ("apply_ms", C.c_double), ("solve_ms", C.c_double),
                ("total_ms", C.c_double), ("converged", C.c_int32), ("total_iterations", C.c_int32)]


P = C.c_void_p
I32, I64, D = C.c_int32, C.c_int64, C.c_double
PI64, PI32, PD = C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_double)

# name -> (restype, argtypes); the exported C ABI
SIGNATURES = {
    "fem_ctx_create": (I32, [I32, C.POINTER(P)]),
    "fem_ctx_destroy": (None, [P]),
    "fem_last_error": (C.c_char_p, [P]),
    "fem_last_error_info": (I32, [P, PI64, PI64, PI64]),
    "fem_mesh_upload": (I32, [P, I32, I64, PD, I64, I32, PI64]),
    "fem_dofs_upload": (I32, [P, I64, I32, PI64]),
    "fem_pattern_upload": (I32, [P, I64, PI64, PI64]),
    "fem_colors_upload": (I32, [P, I32, PI64, PI64]),
    "fem_dofs_distribute": (I32, [P, I32, I32, PI64, PI32]),
    "fem_dofs_download": (I32, [P, PI64]),
    "fem_pattern_build": (I32, [P, PI64]),
    "fem_pattern_download": (I32, [P, PI64, PI64]),
    "fem_cellvalues_upload": (I32, [P, I32, I32, I32, I32, I32, PD, PD, PD, PD]),
    "fem_cellparam_upload": (I32, [P, I64, PD]),
    "fem_assemble": (I32, [P, C.POINTER(AssembleCfg)]),
    "fem_facet_traction": (I32, [P, I64, PI64, I32, I32, PD, PD, PD, PD]),
    "fem_point_load": (I32, [P, I64, PI64, PD]),
    "fem_constraints_upload": (I32, [P, I64, PI64, PD]),
    "fem_apply_dirichlet": (I32, [P, I32, PD]),
    "fem_apply_dirichlet_noalloc": (I32, [P, I32, PD]),
    "fem_affine_upload": (I32, [P, I64, PI64, PI64, PI64, PD, PD]),
    "fem_apply_solution": (I32, [P, I32]),
    "fem_solve": (I32, [P, C.POINTER(SolveCfg), C.POINTER(SolveInfo)]),
    "fem_matrix_download": (I32, [P, PD, PD]),
    "fem_matrix_upload": (I32, [P, PD, PD]),
    "fem_solution_download": (I32, [P, PD]),
    "fem_solution_upload": (I32, [P, PD]),
    "fem_device_pointer": (I32, [P, I32, C.POINTER(P), PI64]),
    "fem_get_timings": (I32, [P, C.POINTER(Timings)]),
    "fem_spmv_bench": (I32, [P, I32, I32, PD]),
    "fem_set_tuning": (I32, [P, C.POINTER(Tuning)]),
    "fem_get_tuning": (I32, [P, C.POINTER(Tuning)]),
    "fem_adjoint_compliance": (I32, [P, PD, D, D, D, PD, PD, PD]),
    "fem_setup_distribute_dofs": (I32, [I32, I64, I32, PI64, I64, I32, PI32, PI32, PI64, PI32, PI64]),
    "fem_setup_pattern": (I32, [I64, I32, PI64, I64, PI64, PI64, I64, PI64]),
    "fem_setup_coloring": (I32, [I64, I32, PI64, I64, I32, PI32, PI32]),
    "fem_batch_create": (I32, [I32, C.POINTER(P)]),
    "fem_batch_destroy": (None, [P]),
    "fem_batch_last_error": (C.c_char_p, [P]),
    "fem_batch_add_problem": (I32, [P, I32, I64, PD, I64, PI64, I64, PI64, PI64, PI64, I32, PI64, PI64, PD,
                                    I64, PI64, D]),
    "fem_batch_finalize": (I32, [P, PD, PD, PD, PD, PD, PD]),
    "fem_batch_autotune": (I32, [P, I32, PI32, PI32]),
    "fem_batch_run": (I32, [P, C.POINTER(BatchCfg), C.POINTER(BatchInfo)]),
    "fem_batch_problem_stats": (I32, [P, PD, PI32]),
    "fem_batch_download": (I32, [P, I32, PD, PD, PD]),
    "fem_batch_assemble_only": (I32, [P, I32, I32, I32, PD]),
}

_lib = None


def lib():
    """Load (once) and return the library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FemError(FEM_ERR_STATE, f"{LIB_PATH} is missing: run __graft_entry__.build() "
                                          "(make -C ferrite.jl_amd/csrc) first")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def exported_symbols():
    h = lib()
    return {name for name in SIGNATURES if getattr(h, name, None) is not None}


def raise_for(code, ctx=None, what="", batch=False):
    if code == FEM_OK:
        return
    msg = what
    if ctx is not None:
        m = (lib().fem_batch_last_error if batch else lib().fem_last_error)(ctx)
        msg = (m.decode() if m else "") or what
    if code == FEM_ERR_PATTERN:
        raise FemPatternError(code, msg)
    if code in (FEM_ERR_ARG, FEM_ERR_DETJ):
        raise FemArgumentError(code, msg)
    if code == FEM_ERR_NOCONV:
        raise FemConvergenceError(code, msg)
    raise FemError(code, msg)


def ptr(a, ctype):
    """Pointer to a contiguous numpy array of the matching dtype."""
    return a.ctypes.data_as(C.POINTER(ctype))


def as_i64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64))


def as_f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))
