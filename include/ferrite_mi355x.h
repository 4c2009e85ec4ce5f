/*
 * ferrite_mi355x.h -- C ABI of libferrite_mi355x.so, the MI355X (gfx950)
 * finite-element assembly engine that sits behind Ferrite.jl's cell-loop API.
 *
 * The reference (Ferrite.jl 1.6.0) is pure Julia; a Julia package extension
 * would bind these entry points with `ccall` (INTEGRATION.md shows the stub).
 * Every entry point replaces one step of the reference's assembly path; the
 * reference interface each one stands in for is cited beside it.
 *
 * Conventions
 *   - status codes: 0 = OK, otherwise a FEM_ERR_* code; fem_last_error()
 *     returns the message (Ferrite's exception text where one exists).
 *   - integer arrays coming from Julia are Int64 and 1-based, in the memory
 *     layout of Ferrite's own containers (Vector{Node}, Vector{Hexahedron},
 *     dh.cell_dofs, SparseMatrixCSC colptr/rowval).  The library copies them
 *     to 0-based int32 device tables; it never keeps a host pointer.
 *   - the library owns all device memory behind the opaque context; K.nzval,
 *     f and u stay resident in HBM between calls and are copied out only by
 *     the *_download calls (or read in place through fem_device_pointer).
 *   - one context = one HIP stream on one device; a context is not thread
 *     safe.  Every call returns after its work on the stream has completed.
 */
#ifndef FERRITE_MI355X_H
#define FERRITE_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fem_ctx fem_ctx;

enum fem_status {
  FEM_OK = 0,
  FEM_ERR_ARG = 1,      /* ArgumentError (length/shape mismatch, common_values.jl:19-31) */
  FEM_ERR_HIP = 2,      /* HIP runtime error */
  FEM_ERR_PATTERN = 3,  /* entry missing in the sparsity pattern (assembler.jl:459-467) */
  FEM_ERR_DETJ = 4,     /* det(J) <= 0 in reinit! (common_values.jl:5, CellValues.jl:110-116) */
  FEM_ERR_STATE = 5,    /* call out of order (e.g. assemble before the pattern is uploaded) */
  FEM_ERR_NOCONV = 6,   /* CG did not reach the tolerance within maxit */
  FEM_ERR_OOM = 7       /* device allocation failed */
};

enum fem_kernel {
  FEM_KERNEL_POISSON = 1,    /* Ke += k (∇Nᵢ·∇Nⱼ) dΩ, fe += s Nᵢ dΩ  (heat_equation.jl:143-164) */
  FEM_KERNEL_ELASTICITY = 2  /* Ke += E (εᵢ : C(ν) : εⱼ) dΩ (threaded_assembly.jl:105-132) */
};

enum fem_precision {
  FEM_FP64 = 0,   /* element math and accumulation in FP64 */
  FEM_MIXED = 1   /* element math in FP32, accumulation into nzval/f in FP64 */
};

/* ---- context ------------------------------------------------------------ */
int32_t fem_ctx_create(int32_t device, fem_ctx** out);
void fem_ctx_destroy(fem_ctx* ctx);
const char* fem_last_error(const fem_ctx* ctx);
/* Error details of the last failed assemble: cell id (1-based), row, col. */
int32_t fem_last_error_info(const fem_ctx* ctx, int64_t* cell, int64_t* row, int64_t* col);

/* ---- device mesh and dof layout (replaces adapt(backend, dh), KAExt/dof_handler.jl:56-86) ---- */
/* Grid.nodes (Vector{Node{sdim,Float64}}, AoS) and Grid.cells (NTuple{npc,Int}, 1-based),
 * src/Grid/grid.jl:12-14, :385-393.  Stored on device as SoA coordinates + int32 tables. */
int32_t fem_mesh_upload(fem_ctx* ctx, int32_t sdim, int64_t nnodes, const double* xyz,
                        int64_t ncells, int32_t nodes_per_cell, const int64_t* cell_nodes);
/* dh.cell_dofs (DofHandler.jl:121-166, celldofs! :248-275), constant ndofs per cell. */
int32_t fem_dofs_upload(fem_ctx* ctx, int64_t ndofs, int32_t ndofs_per_cell, const int64_t* cell_dofs);
/* allocate_matrix(dh) pattern (sparsity_pattern.jl:628-645): SparseMatrixCSC colptr/rowval.
 * Allocates K.nzval (FP64) on the device. */
int32_t fem_pattern_upload(fem_ctx* ctx, int64_t n, const int64_t* colptr, const int64_t* rowval);
/* close!(dh) on the device for one field Lagrange{order}^vdim on the uploaded quadrilateral or
 * hexahedral mesh (DofHandler.jl:576-800: first-visit numbering, cells ascending, vertices ->
 * edges -> faces -> interior per cell; the same numbering as fem_setup_distribute_dofs).  Sets
 * the context's dofs as fem_dofs_upload would; *ndofs and *ndofs_per_cell out. */
int32_t fem_dofs_distribute(fem_ctx* ctx, int32_t order, int32_t vdim, int64_t* ndofs, int32_t* ndofs_per_cell);
/* dh.cell_dofs of the context (1-based Int64, ncells x ndofs_per_cell). */
int32_t fem_dofs_download(fem_ctx* ctx, int64_t* cell_dofs);
/* allocate_matrix(dh)'s pattern built on the device from the context's dofs
 * (sparsity_pattern.jl:223-253, :1136-1249, :951-1016); installs it as fem_pattern_upload
 * does (nzval allocated, zero).  *nnz out. */
int32_t fem_pattern_build(fem_ctx* ctx, int64_t* nnz);
/* The pattern as SparseMatrixCSC's colptr[n+1] / rowval[nnz] (1-based Int64). */
int32_t fem_pattern_download(fem_ctx* ctx, int64_t* colptr, int64_t* rowval);
/* create_coloring(grid) (coloring.jl:308-317) flattened: color c holds
 * color_cells[color_ptr[c] .. color_ptr[c+1]-1] (0-based offsets, 1-based cell ids). */
int32_t fem_colors_upload(fem_ctx* ctx, int32_t ncolors, const int64_t* color_ptr, const int64_t* color_cells);
/* CellValues(qr, ip) tables (CellValues.jl:66-90, FunctionValues.jl:62-98,
 * GeometryMapping.jl:68-98): N[q][i], dNdxi[q][i][d] of the scalar base interpolation,
 * dMdxi[q][j][d] of the geometric interpolation, weights[q].  vdim = number of
 * vectorized components (1 = scalar field, rdim = ip^rdim). */
int32_t fem_cellvalues_upload(fem_ctx* ctx, int32_t rdim, int32_t nqp, int32_t nbf, int32_t ngeo,
                              int32_t vdim, const double* N, const double* dNdxi,
                              const double* dMdxi, const double* weights);
/* Per-cell material parameter in cell-id order (k_e for Poisson, E_e for elasticity). */
int32_t fem_cellparam_upload(fem_ctx* ctx, int64_t ncells, const double* param);

/* ---- assembly (replaces the colored threaded cell loop, threaded_assembly.jl:232-265,
 *      i.e. start_assemble assembler.jl:287-291 + per cell reinit! CellValues.jl:122-140
 *      + element routine + assemble! assembler.jl:322-457) ---- */
typedef struct {
  int32_t kernel;      /* fem_kernel */
  int32_t precision;   /* fem_precision */
  int32_t fillzero;    /* start_assemble(K, f; fillzero) */
  int32_t local_constraints; /* apply_local!(Ke, fe, dofs, ch) on every cell before it is added
                                (ConstraintHandler.jl:1750-1820) with the constraints of the last
                                fem_constraints_upload: 0 none, 1 apply_local!, 2 apply_zero;
                                scalar Poisson (colored scatter), Dirichlet constraints only */
  double source;       /* Poisson: constant body source s */
  double nu;           /* elasticity: Poisson's ratio */
} fem_assemble_cfg;

int32_t fem_assemble(fem_ctx* ctx, const fem_assemble_cfg* cfg);

/* FacetValues traction load f[dofs] += t·N dΓ on a facet set (FacetValues.jl:95-156,
 * topology_optimization.jl:369-381).  facets = (cell, facet) pairs, 1-based;
 * fN[q][i], fdMdxi[q][j][d], fweights[q] are the facet tables of EVERY facet, facet-major. */
int32_t fem_facet_traction(fem_ctx* ctx, int64_t nfacets, const int64_t* facets,
                           int32_t nfq, int32_t nfacet_types, const double* fN,
                           const double* fdMdxi, const double* fweights, const double* traction);
/* Point load f[dof] += value (nodal force), 1-based dofs. */
int32_t fem_point_load(fem_ctx* ctx, int64_t n, const int64_t* dofs, const double* values);

/* ---- Dirichlet apply! (ConstraintHandler.jl:710-747; GPU counterpart KAExt/constraints.jl:115-148) ---- */
/* ch.prescribed_dofs (sorted, 1-based) and ch.inhomogeneities. */
int32_t fem_constraints_upload(fem_ctx* ctx, int64_t np, const int64_t* prescribed, const double* g);
/* apply!(K, f, ch) (applyzero = 0) / apply_zero!(K, f, ch) (applyzero = 1).  Deterministic,
 * no atomics.  Returns meandiag(K) (ConstraintHandler.jl:952-958) in *m_out if non-NULL. */
int32_t fem_apply_dirichlet(fem_ctx* ctx, int32_t applyzero, double* m_out);
/* The same apply! / apply_zero!, with the same results bit for bit, allocating nothing: no
 * counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc) once the matrix and
 * the constraints are uploaded.  Its reduction partials, meandiag and the pattern-symmetry
 * flag live in one fixed 8,208-byte device array per GPU, held until the process ends, instead
 * of buffers allocated by the context or the call (fem_apply_dirichlet allocates a dense g
 * vector per inhomogeneous call); no problem size needs more.  Every call on a GPU shares that
 * array: calls that overlap on one GPU (two contexts on two host threads) race. */
int32_t fem_apply_dirichlet_noalloc(fem_ctx* ctx, int32_t applyzero, double* m_out);

/* Affine constraints u_c = Σ_k a_ck u_mk + b_c (AffineConstraint, ConstraintHandler.jl:140-150,
 * :368-402; PeriodicDirichlet's mirror = image pairs, :1134-1294), after fem_constraints_upload
 * whose prescribed list holds their dofs (sorted, 1-based) with the effective inhomogeneities
 * of update! (:521-540).  coef_ptr[naff+1] are 0-based offsets into coef_dofs (1-based masters,
 * none of them affinely constrained) / coef_vals; b holds each constraint's own b.  apply!
 * then condenses K := Cᵀ K C, f := Cᵀ f (:809-867) before zeroing the prescribed rows and
 * columns; the pattern needs allocate_matrix(dh, ch)'s constraint entries
 * (sparsity_pattern.jl:782-843).  naff = 0 drops them. */
int32_t fem_affine_upload(fem_ctx* ctx, int64_t naff, const int64_t* dofs, const int64_t* coef_ptr,
                          const int64_t* coef_dofs, const double* coef_vals, const double* b);
/* apply!(u, ch) / apply_zero!(u, ch) on the resident u (ConstraintHandler.jl:686-700). */
int32_t fem_apply_solution(fem_ctx* ctx, int32_t applyzero);

/* ---- solve (replaces u = K \ f, heat_equation.jl:218; the reference has no CG) ---- */
enum fem_solve_mode {
  FEM_SOLVE_CG_FP64 = 0,  /* CG, FP64 SpMV and vectors */
  FEM_SOLVE_IR_MIXED = 1  /* mixed CG: FP32 SpMV and vectors, FP64 dot products, FP64 reliable
                             updates (u += x32, r = f − K u with the FP64 K) */
};
typedef struct {
  int32_t mode;        /* fem_solve_mode */
  int32_t maxit;       /* total CG iterations allowed */
  double rtol;         /* ‖f − K u‖₂ / ‖f‖₂ target (FP64 residual) */
  double inner_rtol;   /* mixed: reliable update each time the FP32 residual fell by this factor
                          (0 -> 0.1) */
  int32_t precond;     /* 0 = none, 1 = Jacobi */
  int32_t reserved;
} fem_solve_cfg;

typedef struct {
  int32_t iterations;        /* total CG iterations */
  int32_t outer_iterations;  /* mixed: reliable updates; FP64: 1 + residual replacements */
  double relres;             /* final true relative residual ‖f − K u‖/‖f‖ in FP64 */
  int32_t operator_format;   /* the CG operator copy used: 0 sliced ELL, 1 lattice planes,
                                2 lattice planes whose (scaled) diagonal is all ones, so the
                                SpMV does not read the centre plane, 3 cell coefficients
                                (tuning cell_operator), 4 the element matrices of the last
                                assembly (tuning elem_operator) */
  int32_t operator_shape;    /* the z-marching cell-form SpMV's template (format 3, 3D):
                                MAP·10⁴ + ZN·10³ + KN·10² + stages·10 + rows per thread of
                                k_cell_spmv_z<X, MAP, ZN, KN, stages, rows>; 0 otherwise */
} fem_solve_info;

int32_t fem_solve(fem_ctx* ctx, const fem_solve_cfg* cfg, fem_solve_info* info);

/* ---- results ---- */
int32_t fem_matrix_download(fem_ctx* ctx, double* nzval, double* f);
/* K.nzval (and f, either may be NULL) written from the host, as a Julia caller writes the
 * SparseMatrixCSC's nzval in place; the solver rebuilds its operator copies at the next solve. */
int32_t fem_matrix_upload(fem_ctx* ctx, const double* nzval, const double* f);
int32_t fem_solution_download(fem_ctx* ctx, double* u);
int32_t fem_solution_upload(fem_ctx* ctx, const double* u);
/* Device address of a resident array (0 = nzval, 1 = f, 2 = u) for zero-copy wrapping.  Asking
 * for nzval marks K as changed: the solver rebuilds its operator copies at the next solve, so
 * a caller may write K in place between solves. */
int32_t fem_device_pointer(fem_ctx* ctx, int32_t which, void** ptr, int64_t* nelem);

/* ---- launch-shape tuning (autotuner interface; no counterpart in the reference) ---- */
typedef struct {
  int32_t spmv_cap64, spmv_lanes64;  /* FP64 residual f − K u (CSR-stream over the CSC arrays): LDS
                                        products per chunk (2048/4096) and lanes per row (1/2/4);
                                        0 = default */
  int32_t sell_compress;             /* solver operator (sliced ELL): 16-bit slice-relative column
                                        offsets when every slice spans < 65536 columns (0, default)
                                        or int32 columns (-1) */
  int32_t asm_block;                 /* cells per workgroup of the element kernels (0 = default) */
  int32_t asm_mode;                  /* scalar Q1 Poisson assembly: 0 = owner-computes gather when
                                        supported (default), 1 = colored scatter, 2 = gather only */
  int32_t time_kernels;              /* 1: the solver launches its CG kernels directly (no graph
                                        replay) with HIP events around every SpMV
                                        (fem_timings.spmv_ms / spmv_launches) */
  int32_t sell_unroll;               /* sliced-ELL SpMV: entries loaded per lane before their x
                                        gathers, 4 / 8 / 16 (0 = default: 8 FP32, 4 FP64) */
  int32_t sell_map;                  /* solver SpMV rows to blocks: 1 = grid-stride, 2 = contiguous,
                                        one range per XCD share; 0 = default (sliced ELL: grid-stride,
                                        lattice planes: XCD ranges) */
  int32_t gather_group;              /* owner-computes assembly: cells whose element values a column
                                        lane loads before adding them, 1 / 2 / 4 / 8 (0 = default 4) */
  int32_t cg_elems;                  /* CG vector step: elements per thread per pass, 2 / 4 / 8 / 16
                                        (0 = default 8) */
  int32_t sell_nt;                   /* sliced-ELL SpMV: non-temporal loads of the operator's values
                                        and offsets (0 = default, 1) or ordinary loads (-1) */
  int32_t affine;                    /* owner-computes assembly on trilinear hexes: constant-Jacobian
                                        form for parallelepiped cells (0 = default, on) or the
                                        quadrature loop for every cell (-1) */
  int32_t lattice;                   /* solver operator when the dofs of a scalar Q1 field form a
                                        tensor-product lattice and K is symmetric: value planes
                                        per neighbour, no column indices, half the values
                                        (0 = default: mixed solver only; 1: FP64 solver too;
                                        -1: never, the sliced ELL copy); the cell form
                                        (cell_operator) is preferred where it applies */
  int32_t lattice_blocks;            /* blocks of the lattice SpMV, a multiple of 8 (0 = default) */
  int32_t cell_operator;             /* solver operator on a lattice of congruent cells whose K is
                                        the sum of k_e times cell 0's unit-coefficient element
                                        matrix (checked entry by entry against K): applied from
                                        the per-cell parameter, 4 bytes per cell instead of value
                                        planes (0 = default: mixed solver only; 1: FP64 solver too;
                                        -1: never) */
  int32_t cell_blocks;               /* blocks of that SpMV, a multiple of 8 (0 = default 768) */
  int32_t cell_rows;                 /* 3D cell form, FP32: rows per thread of the marching SpMV,
                                        1 or 2 (0 = default 2; FP64 always 1) */
  int32_t elem_operator;             /* solver operator from the element matrices the last owner-
                                        computes assembly stored (scalar Q1, Q1 quad^2, Q2 hex^3;
                                        any mesh), applied as per-cell products, when K is still
                                        that assembly followed by apply! and no lattice form
                                        applies: 0 = default: both solvers, for the Q2 node blocks
                                        (fewer bytes per SpMV than the sliced copy); 1: for every
                                        kind; -1: never */
} fem_tuning;
int32_t fem_set_tuning(fem_ctx* ctx, const fem_tuning* t);
int32_t fem_get_tuning(const fem_ctx* ctx, fem_tuning* t);

/* ---- timing: HIP-event durations of the last call's kernels on the context stream ---- */
typedef struct {
  double assemble_kernel_ms;   /* element + scatter kernels, summed over colors */
  double fill_ms;              /* start_assemble zero fill */
  double apply_ms;             /* all apply! kernels */
  double spmv_ms;              /* solver SpMV launches of the last solve, summed (with time_kernels) */
  int32_t spmv_launches;       /* number of those launches (0 without time_kernels) */
  int32_t assemble_launches;
  double solve_ms;             /* the CG loop of the last solve (HIP events) */
  double element_ms;           /* two-phase (owner-computes) assembly: the element-matrix phase */
  double scatter_ms;           /* ... and the phase that sums them into nzval/f (0 for the
                                  colored scatter, whose kernels do both) */
} fem_timings;
int32_t fem_get_timings(const fem_ctx* ctx, fem_timings* t);
/* Average duration (HIP events) of `reps` back-to-back SpMV launches y = K x of the solver,
 * FP64 (mixed = 0) or FP32 values and vectors (mixed = 1), on the operator copy the solver
 * would use (cell form, lattice planes or sliced ELL).  The solver's dominant kernel. */
int32_t fem_spmv_bench(fem_ctx* ctx, int32_t mixed, int32_t reps, double* avg_ms);

/* ---- compliance adjoint (no counterpart in the reference; the compliance of
 *      topology_optimization.jl:320-383 with the SIMP modulus E = E_min + θ^p (1 − E_min)) ----
 * J = fᵀu, dJ/dθ_e = −p θ_e^{p−1}(1−E_min) u_eᵀ K⁰_e u_e, dJ/dX by recomputing Ke per cell.
 * Requires a solved system (fem_solve) assembled with FEM_KERNEL_ELASTICITY. */
int32_t fem_adjoint_compliance(fem_ctx* ctx, const double* theta, double emin, double penal,
                               double nu, double* J_out, double* dJdtheta, double* dJdX);

#ifdef __cplusplus
}
#endif
#endif /* FERRITE_MI355X_H */
