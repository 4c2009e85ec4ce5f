// Internal declarations shared by the libferrite_mi355x translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "ferrite_mi355x.h"

namespace fem {

constexpr int kMaxQP = 27;     // 3x3x3 Gauss
constexpr int kMaxNB = 27;     // scalar base functions (Q2 hex)
constexpr int kMaxNG = 8;      // geometric nodes (Q1 geometry)
constexpr int kMaxNFQ = 9;     // facet quadrature points (3x3)

// Reference tables of CellValues as uploaded (host copy); the element kernels read a
// packed copy from device memory with lane-uniform indices (scalar loads).
struct RefTables {
  double N[kMaxQP][kMaxNB];
  double dN[kMaxQP][kMaxNB][3];
  double dM[kMaxQP][kMaxNG][3];
  double w[kMaxQP];
};
struct RefTablesF {
  float N[kMaxQP][kMaxNB];
  float dN[kMaxQP][kMaxNB][3];
  float dM[kMaxQP][kMaxNG][3];
  float w[kMaxQP];
};

// Device error word: atomicMin of the first offending cell (in color order) per kind.
struct DevError {
  int detj_cell;     // cell id (0-based) with det(J) <= 0, INT_MAX if none
  int pattern_cell;  // cell id with a nonzero value for a missing pattern entry
  int pattern_entry; // local entry (col * nb + row) of that cell (valid when pattern_cell set)
  int pad;
};

struct SolverWork;

// CG scalars of the scaled system (hat quantities): γ = r̂·r̂, δ = r̂·Â r̂.
struct Scal {
  double rz;      // γ of the current residual
  double pq;      // p·Â p at a reseed
  double rr;      // true residual ||f − K u||² after residual() (unscaled)
  double alpha;   // step length of the last step
  double beta;
  double bnorm2;
  double gstop;   // hand the recurrence over (done = 1) once γ <= gstop
  double rzp;     // γ of the previous step
  double gam;     // (r̂·p)/(r̂·r̂) at a reseed: p is made orthogonal to the new residual
  int done;       // 1: recurrence stopped (claims convergence or fell by δ), or broke down
  int iters;
  int restart;    // breakdown happened: the next reseed restarts p = r̂
  int fresh;      // p, s = Â p and α were just set by a reseed: the next step only moves x̂, r̂
  int sfresh;     // (record written by k_cg_scal) the step k_cg_move runs now is such a step
  int pad;
  // x̂ moves every other step: k_cg_move adds xa0 p_old + xa1 p_new (both 0: x̂ is not touched
  // and α p of this step is pending); pend = the α whose step's p is not yet in x̂ (k_fold
  // adds it)
  double xa0, xa1, pend;
  double rz0;     // γ at the last reseed (the host's batch-length choice)
};

}  // namespace fem

struct fem_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int64_t err_cell = -1, err_row = -1, err_col = -1;

  // mesh
  int sdim = 0, npc = 0;
  int64_t nnodes = 0, ncells = 0;
  double* d_x[3] = {nullptr, nullptr, nullptr};
  std::vector<int32_t> h_cell_nodes;   // 0-based, cell-id order

  // dofs
  int64_t ndofs = 0;
  int ndpc = 0;
  std::vector<int32_t> h_cell_dofs;    // 0-based, cell-id order

  // pattern (CSC, 0-based, int32 indices; nnz < 2^31)
  int64_t n = 0, nnz = 0;
  int max_row_nnz = 0;                 // longest column of the pattern
  int pattern_sym = -1;                // structurally symmetric pattern: -1 unknown, 0 no, 1 yes
  int32_t* d_colptr = nullptr;
  int32_t* d_rowval = nullptr;
  int32_t* d_diag = nullptr;           // position of K[j,j] in column j (-1 if absent)
  double* d_nzval = nullptr;
  uint64_t nzval_version = 1;          // bumped whenever nzval changes (solver operator copies track it)
  uint64_t aux_version = 1;            // bumped when cell parameters, cell values or constraints change

  // colors: cells permuted into color-major order on the device
  int ncolors = 0;
  std::vector<int64_t> color_ptr;      // ncolors + 1
  std::vector<int32_t> h_color_cells;  // 0-based cell id per color slot
  int32_t* d_ccell = nullptr;          // [ncells] cell id of color slot
  int32_t* d_cnodes = nullptr;         // [ncells][npc]
  int32_t* d_cdofs = nullptr;          // [ncells][ndpc]
  int32_t* d_cnz = nullptr;            // [ncells][ndpc*ndpc] nzval position, -1 = missing
  bool tables_ready = false;

  // per-cell parameter, color order (d_cparam) and cell-id order (d_param_cell)
  double* d_cparam = nullptr;
  double* d_param_cell = nullptr;
  bool have_param = false;

  // owner-computes (gather) assembly: cell-order nodes, dof -> (cell, local) lists, slot
  // table, element-matrix planes (FP64 and FP32 element math)
  int32_t* d_gcnodes = nullptr;
  int32_t* d_d2c_ptr = nullptr;
  int32_t* d_d2c = nullptr;
  uint8_t* d_gslot = nullptr;
  double* d_keb64 = nullptr;
  double* d_feb64 = nullptr;
  float* d_keb32 = nullptr;
  float* d_feb32 = nullptr;
  int64_t keb64_cells = 0, keb32_cells = 0;
  bool gather_ready = false;
  // owner-computes assembly of vector Q2 hex elasticity: the three dofs 3m..3m+2 of a dof node
  // m share one row set made of such triples; node -> (cell, local node) lists n2c (packed
  // cell*27 + b, ascending cell), the rank of each local node's row triple among column 3m's
  // triples (one byte per list entry and local node a, 255 = absent), and the element matrices
  // [cell][b][(i, j)][a] (28 rows a per block, row 27 padding) in FP32 / FP64
  int q2_ready = -1;                   // 1 tables built, 0 not applicable, -1 not checked
  int32_t* d_n2c_ptr = nullptr;
  int32_t* d_n2c = nullptr;
  uint8_t* d_nrank = nullptr;
  float* d_keq32 = nullptr;
  double* d_keq64 = nullptr;
  int64_t keq32_cells = 0, keq64_cells = 0;

  // cellvalues
  int rdim = 0, nqp = 0, nbf = 0, ngeo = 0, vdim = 1;
  fem::RefTables ref{};
  bool have_cv = false;
  bool tab_dirty = true;
  double* d_tab = nullptr;             // packed N, dN, dM, w for the element kernels (FP64)
  float* d_tabf = nullptr;             // the same in FP32
  // trilinear hex geometry (Ferrite's node order, tables checked on the host): per
  // quadrature point ξ, η, ζ, ηζ, ξζ, ξη for the coefficient form of the Jacobian
  bool trilinear = false;
  double* d_qc = nullptr;
  float* d_qcf = nullptr;
  // affine (parallelepiped) cells of that geometry: Ke = Σ_m A_m P[t][m] with the reference
  // integrals P[t][m] = Σ_q w_q (∇N_a,i ∇N_b,j + ∇N_a,j ∇N_b,i)/(1 + δ_ij) for the 6 unique
  // (i, j), t = (a, b), a < b, row-major, followed by Σ_q w_q N_a (8 values)
  double* d_aff = nullptr;
  float* d_afff = nullptr;
  int affine_mesh = -1;                // every cell a parallelepiped: 1 / 0, −1 = not checked

  // vectors
  double* d_f = nullptr;
  double* d_u = nullptr;

  // constraints
  int64_t np = 0;
  int32_t* d_pres = nullptr;
  double* d_g = nullptr;
  uint8_t* d_isc = nullptr;
  bool any_inhom = false;
  // affine constraints u_c = Σ a u_m + b among the prescribed dofs (apply! condenses them)
  int64_t naff = 0;
  int32_t* d_affidx = nullptr;         // [n] index of each dof's affine constraint, −1 if none
  int32_t* d_afd = nullptr;            // [naff] the affinely constrained dofs, ascending
  int32_t* d_acp = nullptr;            // [naff + 1] coefficient ranges
  int32_t* d_acd = nullptr;            // master dofs
  double* d_acv = nullptr;             // coefficients
  double* d_acb = nullptr;             // [naff] b of each constraint
  int32_t* d_slp = nullptr;            // [n + 1] per master dof: its slaves (ascending), with
  int32_t* d_sld = nullptr;            //   the slave dof
  double* d_slv = nullptr;             //   and the slave's coefficient for this master
  double* d_scalar = nullptr;          // small device scratch for reductions (meandiag etc.)
  double* d_partials = nullptr;
  int npartials = 0;

  fem::DevError* d_err = nullptr;
  fem::SolverWork* solver = nullptr;

  fem_timings timings{};
  fem_tuning tune{};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evm = nullptr;            // between the two phases of an owner-computes assembly

  // element matrices that K equals (the element-assembly solver operator, fem_elemop.hip):
  // ea_kind 1 scalar Q1 planes Ke[t][cell] of the strict upper triangle (diagonal from the zero
  // row sums), 2 planes of the upper triangle with the diagonal (Q1 quad^2), 3 Q2 node blocks;
  // valid while nzval_version == ea_version, i.e. K is the last fillzero assembly, followed
  // by apply! (ea_constrained) with the constraints of aux version ea_aux
  int ea_kind = 0;
  bool ea_fp64 = false;                // which planes hold them: the FP64 or the FP32 copy
  uint64_t ea_version = 0, ea_aux = 0;
  bool ea_constrained = false;
  int32_t* d_gcdofs = nullptr;         // [ncells][ndpc] cell dofs in cell-id order
  int32_t* d_q2cnode = nullptr;        // [ncells][27] dof nodes of each Q2 cell
  double* d_ep64 = nullptr;            // Q1 quad^2 planes (36 per cell), FP64 / FP32
  float* d_ep32 = nullptr;
  int64_t ep_cells = 0;
};

namespace fem {

int set_error(fem_ctx* c, int code, const std::string& msg);
int check_hip(fem_ctx* c, hipError_t e, const char* what);

#define FEM_HIP(ctx, expr)                                   \
  do {                                                       \
    hipError_t _e = (expr);                                  \
    if (_e != hipSuccess) return fem::check_hip(ctx, _e, #expr); \
  } while (0)

template <class T>
int dev_alloc(fem_ctx* c, T** p, size_t n) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(c, FEM_ERR_OOM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  }
  return FEM_OK;
}

template <class T>
void dev_free(T** p) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
}

// Elapsed time between two events recorded on the context stream (stream synchronized).
float event_ms(fem_ctx* c);

int build_tables(fem_ctx* c);
int pattern_finish(fem_ctx* c, int64_t n, int64_t nnz, const int32_t* hcolptr);
int noalloc_scratch_init(int device);   // fem_dirichlet.hip: address of its device scratch
void free_solver(fem_ctx* c);
int build_gather_tables(fem_ctx* c);
// element-assembly solver operator (fem_elemop.hip)
struct ElemOp;
bool elem_op_valid(const fem_ctx* c);
int elem_op_prepare(fem_ctx* c, ElemOp** op, const double* sc, bool fp32, int* grid);
int elem_op_launch(fem_ctx* c, const ElemOp* op, const float* x, float* y, double* part, const Scal* s);
int elem_op_launch(fem_ctx* c, const ElemOp* op, const double* x, double* y, double* part, const Scal* s);
void elem_op_free(ElemOp** op);
const void* elem_op_key(const fem_ctx* c, const ElemOp* op, bool fp32);
void elem_op_note_apply(fem_ctx* c);
void elem_op_note_assembly(fem_ctx* c, int kind, bool fp64, bool fillzero);
int build_ea_rows_tables(fem_ctx* c);
void free_q2_tables(fem_ctx* c);
// Q2 hex elasticity through the owner-computes path when its tables apply (*done = true).
int assemble_elasticity_gather(fem_ctx* c, const fem_assemble_cfg* cfg, bool* done);
bool gather_supported(const fem_ctx* c);
int gather_poisson(fem_ctx* c, bool mixed, double src, bool fillzero, const double* tab64, const float* tab32);

constexpr int kReduceBlocks = 1024;   // fixed partial count -> deterministic reductions

// Tensor-product numbering of the dofs of a scalar Q1 hex (quad) field, from the cell
// topology (fem_lattice.hip).
struct Lattice {
  int dim = 0;                 // 3 or 2 when found
  int64_t nx = 0, ny = 0, nz = 1;
  std::vector<int32_t> pos;    // [n] ix + nx (iy + ny iz) of each dof
};
bool find_lattice(const fem_ctx* c, Lattice* L);

}  // namespace fem
