"""Device-resident system and the drop-in assembly API.

The counterparts of Ferrite's GPU path (docs/src/literate-howto/gpu_assembly.jl,
ext/FerriteCudaExt.jl, ext/FerriteKAExt/*):

  allocate_matrix(MI355XCSC, dh)   allocate_matrix(CuSparseMatrixCSC, dh)  FerriteCudaExt.jl:95-101
  start_assemble(K, f)             start_assemble(K, f; fillzero)          assembler.jl:287-291
  assemble_global(A, cv, colors, element)
                                   the colored threaded cell loop           threaded_assembly.jl:232-265
                                   (reinit!, element routine, assemble! fused on the device)
  apply(K, f, ch)                  apply!(K, f, ch)                         ConstraintHandler.jl:710-747
  solve(K, f, ...)                 u = K \\ f                               heat_equation.jl:218

Every call goes through the C ABI of libferrite_mi355x.so; nothing here computes
on the host beyond argument marshalling.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .dofs import add_constraint_entries, create_sparsity_pattern


class MI355XCSC:
    """Marker type for allocate_matrix(MI355XCSC, dh)."""


class Context:
    """Owns one fem_ctx (one HIP stream on one device)."""

    def __init__(self, device=0):
        self._lib = L.lib()
        h = C.c_void_p()
        rc = self._lib.fem_ctx_create(int(device), C.byref(h))
        L.raise_for(rc, what=f"fem_ctx_create(device={device}) failed (is a GPU visible?)")
        self.h = h

    def call(self, name, *args):
        rc = getattr(self._lib, name)(self.h, *args)
        L.raise_for(rc, self.h, name)
        return rc

    def close(self):
        if getattr(self, "h", None):
            self._lib.fem_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceMatrix:
    """K on the device: Ferrite's pattern (colptr, rowval; 1-based) plus nzval in HBM."""

    def __init__(self, dh, ctx, colptr, rowval):
        self.dh = dh
        self.ctx = ctx
        self.colptr = colptr
        self.rowval = rowval
        self.n = dh.ndofs
        self.nnz = len(rowval)
        self._colors = None
        self._cv = None
        self._param = None

    @property
    def shape(self):
        return (self.n, self.n)

    def nzval(self):
        out = np.empty(self.nnz)
        self.ctx.call("fem_matrix_download", L.ptr(out, C.c_double), None)
        return out

    def set_nzval(self, values):
        """Writes K.nzval from the host (the in-place nzval write of a SparseMatrixCSC)."""
        v = L.as_f64(values)
        if v.shape != (self.nnz,):
            raise ValueError("nzval length must equal nnz")
        self.ctx.call("fem_matrix_upload", L.ptr(v, C.c_double), None)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csc_matrix((self.nzval(), self.rowval - 1, self.colptr - 1), shape=self.shape)


class DeviceVector:
    """f (or u) resident in HBM next to its matrix."""

    def __init__(self, K, which):
        self.K, self.which = K, which

    def __len__(self):
        return self.K.n

    def to_numpy(self):
        out = np.empty(self.K.n)
        if self.which == 1:
            self.K.ctx.call("fem_matrix_download", None, L.ptr(out, C.c_double))
        else:
            self.K.ctx.call("fem_solution_download", L.ptr(out, C.c_double))
        return out


def allocate_matrix(mtype, dh, ch=None, device=0, ctx=None):
    """allocate_matrix(MI355XCSC, dh[, ch]): upload mesh, dofs and the pattern; nzval zeroed in
    HBM.  With a closed ConstraintHandler holding affine constraints the pattern also gets the
    entries their condensation fills (allocate_matrix(dh, ch), sparsity_pattern.jl:782-843)."""
    if mtype is not MI355XCSC:
        raise TypeError("only MI355XCSC is provided by this package")
    if ch is not None and not ch.closed:
        raise RuntimeError("close! the ConstraintHandler first")
    affine = ch is not None and ch.has_affine
    dctx = getattr(dh, "_ctx", None)
    if dctx is not None and ctx is None:
        # dofs distributed on the device (DofHandler.close(device=...)): mesh and dofs are in
        # that context already; the pattern is built there too, and copied out for the host
        dh._ctx = None
        nnz = C.c_int64(0)
        dctx.call("fem_pattern_build", C.byref(nnz))
        colptr = np.zeros(dh.ndofs + 1, dtype=np.int64)
        rowval = np.zeros(nnz.value, dtype=np.int64)
        dctx.call("fem_pattern_download", L.ptr(colptr, C.c_int64), L.ptr(rowval, C.c_int64))
        if affine:
            colptr, rowval = add_constraint_entries(colptr, rowval, ch)
            dctx.call("fem_pattern_upload", dh.ndofs, L.ptr(colptr, C.c_int64), L.ptr(rowval, C.c_int64))
        return DeviceMatrix(dh, dctx, colptr, rowval)
    colptr, rowval = create_sparsity_pattern(dh)
    if affine:
        colptr, rowval = add_constraint_entries(colptr, rowval, ch)
    ctx = ctx or Context(device)
    g = dh.grid
    xyz = L.as_f64(g.nodes)
    cells = L.as_i64(g.cells)
    ctx.call("fem_mesh_upload", g.sdim, g.nnodes, L.ptr(xyz, C.c_double), g.ncells, cells.shape[1],
             L.ptr(cells, C.c_int64))
    cd = L.as_i64(dh.cell_dofs)
    ctx.call("fem_dofs_upload", dh.ndofs, dh.ndofs_per_cell, L.ptr(cd, C.c_int64))
    ctx.call("fem_pattern_upload", dh.ndofs, L.ptr(colptr, C.c_int64), L.ptr(rowval, C.c_int64))
    return DeviceMatrix(dh, ctx, colptr, rowval)


def allocate_vector(K):
    """The device f paired with K (zeros(ndofs(dh)) in the reference)."""
    return DeviceVector(K, 1)


class PoissonElement:
    """Ke += k_e (∇Nᵢ·∇Nⱼ) dΩ ; fe += s Nᵢ dΩ  (heat_equation.jl:143-164 with per-cell k_e)."""
    kernel = L.FEM_KERNEL_POISSON

    def __init__(self, k=None, source=1.0):
        self.k, self.source = k, float(source)

    def param(self, ncells):
        return np.ones(ncells) if self.k is None else L.as_f64(self.k)


class ElasticityElement:
    """Ke += E_e (εᵢ : C(ν) : εⱼ) dΩ with isotropic C (threaded_assembly.jl:105-132)."""
    kernel = L.FEM_KERNEL_ELASTICITY

    def __init__(self, E=None, nu=0.3):
        self.E, self.nu, self.source = E, float(nu), 0.0

    def param(self, ncells):
        return np.ones(ncells) if self.E is None else L.as_f64(self.E)


class Assembler:
    def __init__(self, K, f, fillzero):
        self.K, self.f, self.fillzero = K, f, fillzero


def start_assemble(K, f=None, fillzero=True):
    if f is not None and f.K is not K:
        raise ValueError("f must be the device vector allocated with K")
    return Assembler(K, f or allocate_vector(K), fillzero)


def _upload_colors(K, colors):
    if K._colors is colors:
        return
    sizes = np.array([len(c) for c in colors], dtype=np.int64)
    ptr = np.zeros(len(colors) + 1, dtype=np.int64)
    ptr[1:] = np.cumsum(sizes)
    cells = L.as_i64(np.concatenate([np.asarray(c, dtype=np.int64) for c in colors]))
    K.ctx.call("fem_colors_upload", len(colors), L.ptr(ptr, C.c_int64), L.ptr(cells, C.c_int64))
    K._colors = colors
    K._param = None


def _upload_cellvalues(K, cv):
    if K._cv is cv:
        return
    rdim, nqp, nbf, ngeo, vdim, N, dN, dM, w = cv.device_tables()
    K.ctx.call("fem_cellvalues_upload", rdim, nqp, nbf, ngeo, vdim, L.ptr(N, C.c_double), L.ptr(dN, C.c_double),
               L.ptr(dM, C.c_double), L.ptr(w, C.c_double))
    K._cv = cv


def set_cell_parameters(K, values):
    p = L.as_f64(values)
    K.ctx.call("fem_cellparam_upload", len(p), L.ptr(p, C.c_double))
    K._param = values


def assemble_global(assembler, cv, colors, element, precision="fp64", ch=None, apply_zero=False):
    """The fused colored cell loop: K[dofs, dofs] += Ke, f[dofs] += fe for every cell.  With a
    ConstraintHandler `ch`, every cell's Ke and fe first go through apply_local!(Ke, fe, dofs,
    ch; apply_zero) (ConstraintHandler.jl:1750-1820; Dirichlet constraints, Poisson kernel), so
    K and f come out constrained and no apply! follows."""
    K = assembler.K
    _upload_colors(K, colors)
    _upload_cellvalues(K, cv)
    vals = element.param(K.dh.grid.ncells)
    if K._param is None or K._param is not vals and not np.array_equal(np.asarray(K._param), vals):
        set_cell_parameters(K, vals)
    local = 0
    if ch is not None:
        if getattr(ch, "has_affine", False):
            raise L.FemArgumentError(L.FEM_ERR_ARG, "can not condense constraint without the global matrix and vector")
        p = L.as_i64(ch.prescribed_dofs)
        g = L.as_f64(ch.inhomogeneities)
        K.ctx.call("fem_constraints_upload", len(p), L.ptr(p, C.c_int64), L.ptr(g, C.c_double))
        local = 2 if apply_zero else 1
    cfg = L.AssembleCfg(kernel=element.kernel,
                        precision=L.FEM_MIXED if precision == "mixed" else L.FEM_FP64,
                        fillzero=1 if assembler.fillzero else 0, local_constraints=local,
                        source=float(getattr(element, "source", 0.0)), nu=float(getattr(element, "nu", 0.0)))
    K.ctx.call("fem_assemble", C.byref(cfg))
    assembler.fillzero = False
    return K, assembler.f


def _upload_constraints(K, ch):
    p = L.as_i64(ch.prescribed_dofs)
    g = L.as_f64(ch.inhomogeneities)
    K.ctx.call("fem_constraints_upload", len(p), L.ptr(p, C.c_int64), L.ptr(g, C.c_double))
    if getattr(ch, "has_affine", False):
        dofs, cptr, cdofs, cvals, b = ch.affine_arrays()
        K.ctx.call("fem_affine_upload", len(dofs), L.ptr(dofs, C.c_int64), L.ptr(cptr, C.c_int64),
                   L.ptr(cdofs, C.c_int64), L.ptr(cvals, C.c_double), L.ptr(b, C.c_double))


def apply(K, f, ch, applyzero=False):
    """apply!(K, f, ch) on the device; returns meandiag(K).  Affine constraints of ch are
    condensed into K and f (K := CᵀKC, f := Cᵀf in the serial column order of _condense!,
    ConstraintHandler.jl:809-867) before the prescribed rows and columns are cleared."""
    _upload_constraints(K, ch)
    m = C.c_double(0.0)
    K.ctx.call("fem_apply_dirichlet", 1 if applyzero else 0, C.byref(m))
    return m.value


def apply_zero(K, f, ch):
    return apply(K, f, ch, applyzero=True)


def apply_solution(K, ch, applyzero=False, download=True):
    """apply!(u, ch) / apply_zero!(u, ch) (ConstraintHandler.jl:686-700) on the device solution
    u of the last solve: u[prescribed] = g, then u_i = b_i + Σ a_ij u_j for affine dofs."""
    _upload_constraints(K, ch)
    K.ctx.call("fem_apply_solution", 1 if applyzero else 0)
    return DeviceVector(K, 2).to_numpy() if download else None


class SolveResult:
    def __init__(self, info, u):
        self.iterations = info.iterations
        self.outer_iterations = info.outer_iterations
        self.relres = info.relres
        # the solver's copy of K: "sell" (sliced ELL), "lattice" (value planes per neighbour)
        # or "cells" (per-cell coefficients times cell 0's unit-coefficient element matrix)
        self.operator = {1: "lattice", 2: "lattice", 3: "cells", 4: "elements"}.get(info.operator_format, "sell")
        self.unit_diagonal = info.operator_format == 2
        # the 3D cell form's SpMV template (MAP, ZN, KN, stages, rows per thread), or None
        sh = info.operator_shape
        self.cell_shape = None if sh <= 0 else (sh // 10000, sh // 1000 % 10, sh // 100 % 10, sh // 10 % 10, sh % 10)
        self.u = u


def solve(K, f, mode="cg", rtol=1e-10, maxit=0, precond="jacobi", inner_rtol=0.0, download=True):
    """u = K \\ f on the device: FP64 PCG (mode='cg') or FP64-refined FP32 PCG (mode='mixed')."""
    cfg = L.SolveCfg(mode=L.FEM_SOLVE_IR_MIXED if mode == "mixed" else L.FEM_SOLVE_CG_FP64,
                     maxit=int(maxit), rtol=float(rtol), inner_rtol=float(inner_rtol),
                     precond=1 if precond == "jacobi" else 0, reserved=0)
    info = L.SolveInfo()
    K.ctx.call("fem_solve", C.byref(cfg), C.byref(info))
    u = DeviceVector(K, 2).to_numpy() if download else None
    return SolveResult(info, u)


def timings(K):
    t = L.Timings()
    K.ctx.call("fem_get_timings", C.byref(t))
    return {k: getattr(t, k) for k, _ in L.Timings._fields_}


def spmv_bench(K, mixed=False, reps=20):
    ms = C.c_double(0.0)
    K.ctx.call("fem_spmv_bench", 1 if mixed else 0, int(reps), C.byref(ms))
    return ms.value


def set_tuning(K, **kw):
    """Launch-shape knobs of the context (see fem_tuning in include/ferrite_mi355x.h)."""
    t = L.Tuning()
    K.ctx.call("fem_get_tuning", C.byref(t))
    for k, v in kw.items():
        setattr(t, k, int(v))
    K.ctx.call("fem_set_tuning", C.byref(t))


def get_tuning(K):
    t = L.Tuning()
    K.ctx.call("fem_get_tuning", C.byref(t))
    return {k: getattr(t, k) for k, _ in L.Tuning._fields_ if k != "reserved"}


def apply_traction(K, f, fv, facets, traction):
    """f[dofs] += ∫ t·N dΓ over `facets` ((cell, facet) pairs, 1-based), the FacetValues
    traction term of an element routine (topology_optimization.jl:369-381)."""
    nfq, nft, N, dM, w = fv.device_tables()
    fc = L.as_i64(np.asarray(facets, dtype=np.int64).reshape(-1, 2))
    t = L.as_f64(np.resize(np.asarray(traction, dtype=np.float64), 3))
    K.ctx.call("fem_facet_traction", len(fc), L.ptr(fc, C.c_int64), nfq, nft, L.ptr(N, C.c_double),
               L.ptr(dM, C.c_double), L.ptr(w, C.c_double), L.ptr(t, C.c_double))


def add_point_load(K, f, dofs, values):
    """f[dofs] += values (nodal forces), 1-based dofs."""
    d = L.as_i64(np.atleast_1d(dofs))
    v = L.as_f64(np.atleast_1d(values))
    K.ctx.call("fem_point_load", len(d), L.ptr(d, C.c_int64), L.ptr(v, C.c_double))


def compliance_gradient(K, theta, emin=1e-3, penal=3.0, nu=0.3):
    """J = fᵀu and its adjoint gradients for E_e = E_min + θ_e^p (1 − E_min):
    returns (J, dJ/dθ (ncells,), dJ/dX (nnodes, dim)).  Needs the solved system in K."""
    g = K.dh.grid
    th = L.as_f64(theta)
    dth = np.empty(g.ncells)
    dx = np.empty((g.nnodes, g.sdim))
    J = C.c_double(0.0)
    K.ctx.call("fem_adjoint_compliance", L.ptr(th, C.c_double), float(emin), float(penal), float(nu), C.byref(J),
               L.ptr(dth, C.c_double), L.ptr(dx, C.c_double))
    return J.value, dth, dx


def simp_modulus(theta, emin=1e-3, penal=3.0):
    """E_e = E_min + θ_e^p (1 − E_min)."""
    theta = np.asarray(theta, dtype=np.float64)
    return emin + theta ** penal * (1.0 - emin)
