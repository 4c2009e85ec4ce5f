/* Batched assembly + solve of many independent scalar Q1 Poisson problems.
 *
 * Reference counterpart: the serial per-problem loop of
 * docs/src/literate-tutorials/heat_equation.jl:186-218 (start_assemble; for cell in
 * CellIterator(dh): reinit!, Ke/fe, assemble!; apply!(K, f, ch); u = K \ f), run for
 * every problem of a set, each with its own Grid, DofHandler, create_sparsity_pattern
 * (src/Dofs/sparsity_pattern.jl), create_coloring (src/Grid/coloring.jl) and
 * ConstraintHandler.  Here all problems live in one device context; assembly is one
 * persistent kernel draining a (problem, colour, chunk) work queue, captured in a
 * hipGraph; the solve is one batched mixed-precision CG with per-problem scalars.
 *
 * Conventions are those of ferrite_mi355x.h: 1-based indices in the host arrays,
 * int32 status codes, fem_batch_last_error() for the message.
 */
#ifndef FERRITE_MI355X_BATCH_H
#define FERRITE_MI355X_BATCH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fem_batch fem_batch;

typedef struct {
  int32_t precision;     /* FEM_FP64 or FEM_MIXED (element math in FP32, FP64 accumulation) */
  int32_t cells_per_wg;  /* 0 = autotuned / default; else 64, 128 or 256 (256: FEM_MIXED only) */
  int32_t chunk;         /* cells per work item; 0 = autotuned / default; must be >= cells_per_wg */
  int32_t maxit;         /* CG iteration cap per problem (0 = 20000) */
  double rtol;           /* per-problem ||f - K u|| <= rtol ||f|| (FP64 residual) */
  double ru_delta;       /* reliable-update trigger: FP32 residual fell by this factor (0 = 0.1) */
} fem_batch_cfg;

typedef struct {
  double assemble_ms;        /* persistent assembly graph (zero + kernel) */
  double apply_ms;           /* batched apply! (meandiag, zero rows/cols, diagonal) */
  double solve_ms;           /* batched CG until the last problem converged */
  double total_ms;           /* whole run, start of assembly to the last convergence */
  int32_t converged;         /* problems that reached rtol */
  int32_t total_iterations;  /* CG iterations executed by the batch (max over problems) */
} fem_batch_info;

/* Create a batch on a device (owns its own stream and context). */
int32_t fem_batch_create(int32_t device, fem_batch** out);
void fem_batch_destroy(fem_batch* b);
const char* fem_batch_last_error(const fem_batch* b);

/* Add one problem.  dim 2: Quadrilateral Q1 (4 nodes/cell), dim 3: Hexahedron Q1
 * (8 nodes/cell).  xyz[nnodes*dim] (Grid.nodes), cell_nodes[ncells*npc] (Grid.cells),
 * cell_dofs[ncells*npc] (celldofs of a DofHandler with one scalar Lagrange{1} field,
 * src/Dofs/DofHandler.jl), colptr/rowval of its create_sparsity_pattern, colours as
 * color_ptr[ncolors+1] (0-based offsets) + color_cells[ncells] (1-based cell ids, the
 * concatenated vectors of create_coloring), k[ncells] the per-cell conductivity,
 * prescribed[np] its homogeneous Dirichlet dofs (ConstraintHandler.prescribed_dofs),
 * source the constant right-hand side f. */
int32_t fem_batch_add_problem(fem_batch* b, int32_t dim, int64_t nnodes, const double* xyz, int64_t ncells,
                              const int64_t* cell_nodes, int64_t ndofs, const int64_t* cell_dofs,
                              const int64_t* colptr, const int64_t* rowval, int32_t ncolors,
                              const int64_t* color_ptr, const int64_t* color_cells, const double* k, int64_t np,
                              const int64_t* prescribed, double source);

/* Upload the batch.  Reference tables of CellValues(QuadratureRule(2), Lagrange{1}):
 * quad N[4 qp][4], dNdxi[4][4][2], w[4]; hex N[8 qp][8], dNdxi[8][8][3], w[8]. */
int32_t fem_batch_finalize(fem_batch* b, const double* quad_N, const double* quad_dNdxi, const double* quad_w,
                           const double* hex_N, const double* hex_dNdxi, const double* hex_w);

/* Time every (cells_per_wg, chunk) candidate on this batch and keep the fastest. */
int32_t fem_batch_autotune(fem_batch* b, int32_t precision, int32_t* cells_per_wg, int32_t* chunk);

/* Assemble, apply! and solve every problem.  Fails with FEM_ERR_NOCONV if a problem
 * misses rtol within maxit, FEM_ERR_DETJ / FEM_ERR_PATTERN as the single-problem
 * assembler does. */
int32_t fem_batch_run(fem_batch* b, const fem_batch_cfg* cfg, fem_batch_info* info);

/* Per problem of the last run: completion time (ms from the start of assembly) and CG
 * iterations; arrays of length nproblems. */
int32_t fem_batch_problem_stats(const fem_batch* b, double* completion_ms, int32_t* iterations);

/* Assembly only (zero + persistent kernel), timed with HIP events: *ms. */
int32_t fem_batch_assemble_only(fem_batch* b, int32_t precision, int32_t cells_per_wg, int32_t chunk, double* ms);

/* Copy problem `prob` (0-based) back: K.nzval[nnz], f[ndofs], u[ndofs]; NULL skips. */
int32_t fem_batch_download(fem_batch* b, int32_t prob, double* nzval, double* f, double* u);

#ifdef __cplusplus
}
#endif

#endif
