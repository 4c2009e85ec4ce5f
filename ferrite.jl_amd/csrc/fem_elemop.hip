// Element-assembly operator of the CG solver: K x applied as per-cell dense products with
// the element matrices the assembly stored, instead of a copy of the assembled K
// (docs/src/literate-howto/gpu_assembly.jl:265-304, the reference's element-assembly
// strategy: keep Ke per cell and apply the sum of the cell products).
//
// K is the sum of the stored Ke only when nzval is exactly the last fillzero assembly,
// possibly followed by apply! with the current constraints; fem_assemble and apply! keep
// that record (ea_kind / ea_version / ea_aux), and any other change of nzval (an upload, a
// device-pointer hand-out, an accumulating assembly) bumps nzval_version past it.  Works for
// any mesh, straight or distorted, structured or not.
//
// The solver runs on Â = S K S (S = diag(K)^-1/2).  With t_j = S_j x_j on free dofs (0 on
// prescribed ones, whose columns apply! zeroed):
//   free row i:        (Â x)_i = S_i Σ_{cells c ∋ i} Σ_b Ke_c[a_i, b] t_{dof_c(b)}
//   prescribed row i:  (Â x)_i = S_i K_ii S_i x_i                (K_ii = meandiag)
// sm_i = S_i (0 on prescribed rows) and pd_i = S_i² K_ii (0 on free rows) are formed once per
// solve.  Ke is symmetric (as Ferrite's is), so row a_i of Ke is its stored column.
//
// Layouts of the stored element matrices:
//   kind 1  scalar Q1 (hex / quad) Poisson: planes Ke[t][cell] of the strict upper triangle,
//           diagonal Ke[a,a] = −Σ_{b≠a} Ke[a,b] (rows of a Lagrange Poisson Ke sum to zero,
//           which is how the owner-computes assembly forms K's diagonal too);
//   kind 2  Q1 quad² elasticity: planes Ke[t][cell] of the upper triangle with the diagonal;
//   kind 3  Q2 hex³ elasticity: per cell and local node b the 81 x 3 column block as
//           [(i, j)][a] with 28 rows (row 27 padding), 252 values.
// Every row sums its cells in ascending cell id and each cell's terms in a fixed order, so
// the result is bitwise reproducible.  Both run in two passes: the cell products yc = Ke t,
// then the row sums of the cell values.
#include <algorithm>
#include <vector>

#include "fem_element.h"

namespace fem {

constexpr int kEaBlock = 256;
constexpr int kEaGrid = 2048;   // fixed block count: the CG's partial count

struct ElemOp {
  int64_t n = 0;
  float *sm32 = nullptr, *pd32 = nullptr;
  double *sm64 = nullptr, *pd64 = nullptr;
  float* keb32 = nullptr;          // FP32 copy of FP64 element matrices (mixed solver)
  float* yc32 = nullptr;           // cell products yc[a][cell] of the two-pass form (kinds 1, 2)
  double* yc64 = nullptr;
  uint64_t conv_ver = 0;           // ea_version the copy was made from (+1)
  int grid = kEaGrid;
};

__device__ __forceinline__ double ea_wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ __forceinline__ float ea_wave_sum(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

__device__ __forceinline__ double ea_block_sum(double v, double* sh) {
  v = ea_wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kEaBlock / 64; ++i) r += sh[i];
  return r;
}

// Index of (lo, hi), lo < hi, in the row-major strict upper triangle of an nb x nb matrix.
constexpr __host__ __device__ int ea_off_idx(int lo, int hi, int nb) {
  return lo * (2 * nb - lo - 1) / 2 + (hi - lo - 1);
}

template <typename X>
__global__ void k_ea_prep(int64_t n, const double* __restrict__ sc, const uint8_t* __restrict__ isc,
                          const int32_t* __restrict__ diag, const double* __restrict__ nzval, X* __restrict__ sm,
                          X* __restrict__ pd) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const bool pres = isc && isc[i];
    const double s = sc[i];
    sm[i] = pres ? X(0) : (X)s;
    pd[i] = pres ? (X)(s * nzval[diag[i]] * s) : X(0);
  }
}

template <typename P>
__global__ void k_ea_to32(int64_t m, const P* __restrict__ a, float* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// Scalar / small vector elements, two passes.
//   k_ea_cells: one lane per cell: t_b = sm[dof_b] x[dof_b] of its NB dofs, yc = Ke t (NB x NB
//               product), stored as planes yc[a][cell].  ROWSUM: strict-upper planes and the
//               diagonal from the row sum (kind 1); otherwise upper planes with the diagonal
//               (kind 2).
//   k_ea_sum:   one lane per row: the row's cell values yc[a][cell] in ascending cell id (its
//               (cell, local index) list d2c), scaled; partial x·y per block.
template <typename X, typename P, int NB, bool ROWSUM>
__global__ __launch_bounds__(kEaBlock) void k_ea_cells(int64_t ncells, const int32_t* __restrict__ cdofs,
                                                       const P* __restrict__ planes, const X* __restrict__ sm,
                                                       const X* __restrict__ x, X* __restrict__ yc,
                                                       const Scal* __restrict__ s) {
  if (s->done || s->fresh) return;   // the recurrence stopped, or this step only moves x̂, r̂
  for (int64_t cell = blockIdx.x * (int64_t)kEaBlock + threadIdx.x; cell < ncells;
       cell += (int64_t)gridDim.x * kEaBlock) {
    int32_t dof[NB];
    if constexpr (NB == 8) {
      const int4* p = reinterpret_cast<const int4*>(cdofs + cell * 8);
      const int4 u0 = p[0], u1 = p[1];
      dof[0] = u0.x; dof[1] = u0.y; dof[2] = u0.z; dof[3] = u0.w;
      dof[4] = u1.x; dof[5] = u1.y; dof[6] = u1.z; dof[7] = u1.w;
    } else {
      const int4 u0 = *reinterpret_cast<const int4*>(cdofs + cell * 4);
      dof[0] = u0.x; dof[1] = u0.y; dof[2] = u0.z; dof[3] = u0.w;
    }
    X t[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) t[b] = sm[dof[b]] * x[dof[b]];
    X y[NB];
#pragma unroll
    for (int a = 0; a < NB; ++a) y[a] = X(0);
    if constexpr (ROWSUM) {
      // Ke[a,b] (a < b) from the plane; Ke[a,a] = −Σ_{b≠a} Ke[a,b] in ascending b
      X dg[NB];
#pragma unroll
      for (int a = 0; a < NB; ++a) dg[a] = X(0);
      X v[NB * (NB - 1) / 2];
#pragma unroll
      for (int q = 0; q < NB * (NB - 1) / 2; ++q) v[q] = (X)planes[(int64_t)q * ncells + cell];
#pragma unroll
      for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b)
          if (a != b) {
            const X kab = v[ea_off_idx(a < b ? a : b, a < b ? b : a, NB)];
            dg[a] -= kab;
            y[a] += kab * t[b];
          }
#pragma unroll
      for (int a = 0; a < NB; ++a) y[a] += dg[a] * t[a];
    } else {
      X v[NB * (NB + 1) / 2];
#pragma unroll
      for (int q = 0; q < NB * (NB + 1) / 2; ++q) v[q] = (X)planes[(int64_t)q * ncells + cell];
#pragma unroll
      for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) y[a] += v[tri_idx(a < b ? a : b, a < b ? b : a, NB)] * t[b];
    }
#pragma unroll
    for (int a = 0; a < NB; ++a) yc[(int64_t)a * ncells + cell] = y[a];
  }
}

template <typename X, int NB>
__global__ __launch_bounds__(kEaBlock) void k_ea_sum(int64_t n, int64_t ncells, const int32_t* __restrict__ d2c_ptr,
                                                     const int32_t* __restrict__ d2c, const X* __restrict__ yc,
                                                     const X* __restrict__ sm, const X* __restrict__ pd,
                                                     const X* __restrict__ x, X* __restrict__ y,
                                                     double* __restrict__ part, const Scal* __restrict__ s) {
  if (s->done || s->fresh) return;
  __shared__ double sh[kEaBlock / 64];
  double d = 0.0;
  for (int64_t row = blockIdx.x * (int64_t)kEaBlock + threadIdx.x; row < n; row += (int64_t)gridDim.x * kEaBlock) {
    X acc = X(0);
    const int32_t e0 = d2c_ptr[row], e1 = d2c_ptr[row + 1];
    int32_t e = e0;
    for (; e + 4 <= e1; e += 4) {   // four cells' values in flight, added in ascending order
      int32_t pk[4];
      X v[4];
#pragma unroll
      for (int g = 0; g < 4; ++g) pk[g] = d2c[e + g];
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int64_t cell = pk[g] / NB;
        v[g] = yc[(int64_t)(pk[g] - (int)cell * NB) * ncells + cell];
      }
#pragma unroll
      for (int g = 0; g < 4; ++g) acc += v[g];
    }
    for (; e < e1; ++e) {
      const int32_t pk = d2c[e];
      const int64_t cell = pk / NB;
      acc += yc[(int64_t)(pk - (int)cell * NB) * ncells + cell];
    }
    const X xr = x[row];
    const X yv = sm[row] * acc + pd[row] * xr;
    y[row] = yv;
    d += (double)xr * (double)yv;
  }
  const double bsum = ea_block_sum(d, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bsum;
}

// Q2 hex³, two passes.
//   k_ea_q2_cells: three cells per 256-thread block, one thread per row r = (a, i) of a cell's
//                  81 x 81 Ke: yc[cell][r] = Σ_b Σ_j Ke[(a, i), (b, j)] t_(b, j), reading the
//                  stored column blocks (Ke is symmetric: row (a, i) = column (a, i)); t in LDS.
//                  The 81 threads of a cell read each block as runs of 27 consecutive values.
//   k_ea_q2_sum:   one lane per row 3m + j: Σ yc[cell][3b + j] over the node's (cell, b) list.
constexpr int kQ2Cells = 3;   // cells per block of k_ea_q2_cells (243 of 256 threads busy)

template <typename X, typename P>
__global__ __launch_bounds__(kEaBlock) void k_ea_q2_cells(int64_t ncells, const int32_t* __restrict__ cnode,
                                                          const P* __restrict__ keb, const X* __restrict__ sm,
                                                          const X* __restrict__ x, X* __restrict__ yc,
                                                          const Scal* __restrict__ s) {
  if (s->done || s->fresh) return;
  constexpr int kKS = 252, kKC = 27 * kKS;
  __shared__ X st[kQ2Cells][81];
  const int lc = threadIdx.x / 81, r = threadIdx.x - 81 * lc;
  const int64_t cell = (int64_t)blockIdx.x * kQ2Cells + lc;
  const bool act = lc < kQ2Cells && cell < ncells;
  if (act) {   // t of this cell: the thread of row r loads entry r (node r / 3, component r % 3)
    const int64_t dof = 3 * (int64_t)cnode[cell * 27 + r / 3] + (r - 3 * (r / 3));
    st[lc][r] = sm[dof] * x[dof];
  }
  __syncthreads();
  if (!act) return;
  const int a = r / 3, i = r - 3 * a;
  const P* blk = keb + cell * (int64_t)kKC + (i * 3) * 28 + a;
  X acc = X(0);
#pragma unroll 3
  for (int b = 0; b < 27; ++b) {
    const P* bb = blk + b * kKS;
    const X v0 = (X)bb[0], v1 = (X)bb[28], v2 = (X)bb[56];   // (i, j = 0, 1, 2)
    acc += v0 * st[lc][3 * b] + v1 * st[lc][3 * b + 1] + v2 * st[lc][3 * b + 2];
  }
  yc[cell * 81 + r] = acc;
}

template <typename X>
__global__ __launch_bounds__(kEaBlock) void k_ea_q2_sum(int64_t n, const int32_t* __restrict__ n2c_ptr,
                                                        const int32_t* __restrict__ n2c, const X* __restrict__ yc,
                                                        const X* __restrict__ sm, const X* __restrict__ pd,
                                                        const X* __restrict__ x, X* __restrict__ y,
                                                        double* __restrict__ part, const Scal* __restrict__ s) {
  if (s->done || s->fresh) return;
  __shared__ double sh[kEaBlock / 64];
  double d = 0.0;
  for (int64_t row = blockIdx.x * (int64_t)kEaBlock + threadIdx.x; row < n; row += (int64_t)gridDim.x * kEaBlock) {
    const int64_t m = row / 3;
    const int j = (int)(row - 3 * m);
    X acc = X(0);
    const int32_t e0 = n2c_ptr[m], e1 = n2c_ptr[m + 1];
    int32_t e = e0;
    for (; e + 4 <= e1; e += 4) {   // four cells' values in flight, added in ascending order
      int32_t ent[4];
      X v[4];
#pragma unroll
      for (int g = 0; g < 4; ++g) ent[g] = n2c[e + g];
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int64_t cell = ent[g] / 27;
        v[g] = yc[cell * 81 + 3 * (ent[g] - (int)cell * 27) + j];
      }
#pragma unroll
      for (int g = 0; g < 4; ++g) acc += v[g];
    }
    for (; e < e1; ++e) {
      const int64_t cell = n2c[e] / 27;
      acc += yc[cell * 81 + 3 * (n2c[e] - (int)cell * 27) + j];
    }
    const X xr = x[row];
    const X yv = sm[row] * acc + pd[row] * xr;
    y[row] = yv;
    d += (double)xr * (double)yv;
  }
  const double bsum = ea_block_sum(d, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bsum;
}

static int gride(int64_t n, int b = 256) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + b - 1) / b, 8192)); }

// Row lists d2c and cell-order dofs for kind 2 (kind 1 has them from the gather tables).
int build_ea_rows_tables(fem_ctx* c) {
  if (c->d_d2c_ptr && c->d_d2c && c->d_gcdofs) return FEM_OK;
  const int64_t nc = c->ncells, nd = c->ndofs;
  const int nb = c->ndpc;
  std::vector<int32_t> ptr(nd + 1, 0), ent(nc * nb);
  for (int64_t i = 0; i < nc * nb; ++i) ++ptr[c->h_cell_dofs[i] + 1];
  for (int64_t j = 0; j < nd; ++j) ptr[j + 1] += ptr[j];
  std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
  for (int64_t cell = 0; cell < nc; ++cell)
    for (int b = 0; b < nb; ++b) ent[fill[c->h_cell_dofs[cell * nb + b]]++] = (int32_t)(cell * nb + b);
  int rc;
  if ((rc = dev_alloc(c, &c->d_d2c_ptr, nd + 1)) || (rc = dev_alloc(c, &c->d_d2c, nc * nb)) ||
      (rc = dev_alloc(c, &c->d_gcdofs, nc * nb)))
    return rc;
  FEM_HIP(c, hipMemcpyAsync(c->d_d2c_ptr, ptr.data(), 4 * (nd + 1), hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipMemcpyAsync(c->d_d2c, ent.data(), 4 * nc * nb, hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipMemcpyAsync(c->d_gcdofs, c->h_cell_dofs.data(), 4 * nc * nb, hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  c->gather_ready = false;   // the gather tables own d2c too; they rebuild it with their slots
  return FEM_OK;
}

bool elem_op_valid(const fem_ctx* c) {
  if (c->ea_kind == 0 || c->ea_version != c->nzval_version) return false;
  if (c->ea_constrained && (c->ea_aux != c->aux_version || !c->d_isc)) return false;
  if (c->ea_kind == 3)
    return c->d_n2c && c->d_q2cnode && (c->ea_fp64 ? (void*)c->d_keq64 : (void*)c->d_keq32) != nullptr;
  if (!c->d_d2c_ptr || !c->d_d2c || !c->d_gcdofs) return false;
  if (c->ea_kind == 1) return (c->ea_fp64 ? (void*)c->d_keb64 : (void*)c->d_keb32) != nullptr;
  return (c->ea_fp64 ? (void*)c->d_ep64 : (void*)c->d_ep32) != nullptr;
}

void elem_op_free(ElemOp** op) {
  if (!*op) return;
  ElemOp* o = *op;
  dev_free(&o->sm32); dev_free(&o->pd32); dev_free(&o->sm64); dev_free(&o->pd64); dev_free(&o->keb32);
  dev_free(&o->yc32); dev_free(&o->yc64);
  delete o;
  *op = nullptr;
}

static int64_t ea_plane_count(const fem_ctx* c) {
  const int nb = c->ndpc;
  if (c->ea_kind == 1) return c->ncells * (int64_t)(nb * (nb - 1) / 2);
  if (c->ea_kind == 2) return c->ncells * (int64_t)(nb * (nb + 1) / 2);
  return c->ncells * (int64_t)(27 * 252);
}

// sm, pd of this solve (and an FP32 copy of FP64 element matrices for the FP32 operator).
int elem_op_prepare(fem_ctx* c, ElemOp** opp, const double* sc, bool fp32, int* grid) {
  if (!*opp) *opp = new ElemOp();
  ElemOp* o = *opp;
  const int64_t n = c->n;
  int rc;
  if (o->n != n) {
    dev_free(&o->sm32); dev_free(&o->pd32); dev_free(&o->sm64); dev_free(&o->pd64);
    dev_free(&o->yc32); dev_free(&o->yc64);
    o->n = n;
  }
  const uint8_t* isc = c->ea_constrained ? c->d_isc : nullptr;
  const int64_t ycn = c->ncells * (int64_t)c->ndpc;   // one value per cell and local dof
  if (fp32) {
    if (!o->sm32 && ((rc = dev_alloc(c, &o->sm32, n)) || (rc = dev_alloc(c, &o->pd32, n)))) return rc;
    if (!o->yc32 && (rc = dev_alloc(c, &o->yc32, ycn))) return rc;
    k_ea_prep<float><<<gride(n), 256, 0, c->stream>>>(n, sc, isc, c->d_diag, c->d_nzval, o->sm32, o->pd32);
    if (c->ea_fp64 && o->conv_ver != c->ea_version + 1) {
      const int64_t m = ea_plane_count(c);
      if ((rc = dev_alloc(c, &o->keb32, m))) return rc;
      const double* src = c->ea_kind == 1 ? c->d_keb64 : c->ea_kind == 2 ? c->d_ep64 : c->d_keq64;
      k_ea_to32<double><<<gride(m), 256, 0, c->stream>>>(m, src, o->keb32);
      o->conv_ver = c->ea_version + 1;
    }
  } else {
    if (!o->sm64 && ((rc = dev_alloc(c, &o->sm64, n)) || (rc = dev_alloc(c, &o->pd64, n)))) return rc;
    if (!o->yc64 && (rc = dev_alloc(c, &o->yc64, ycn))) return rc;
    k_ea_prep<double><<<gride(n), 256, 0, c->stream>>>(n, sc, isc, c->d_diag, c->d_nzval, o->sm64, o->pd64);
  }
  FEM_HIP(c, hipGetLastError());
  *grid = o->grid;
  return FEM_OK;
}

template <typename X, typename P>
static void launch_kind(fem_ctx* c, const ElemOp* o, const P* planes, const X* sm, const X* pd, const X* x, X* y,
                        X* yc, double* part, const Scal* s) {
  const int g = o->grid;
  const int nb = c->ndpc;
  if (c->ea_kind == 3) {
    const int64_t gq = (c->ncells + kQ2Cells - 1) / kQ2Cells;
    k_ea_q2_cells<X, P><<<(unsigned)gq, kEaBlock, 0, c->stream>>>(c->ncells, c->d_q2cnode, planes, sm, x, yc, s);
    k_ea_q2_sum<X><<<g, kEaBlock, 0, c->stream>>>(c->n, c->d_n2c_ptr, c->d_n2c, yc, sm, pd, x, y, part, s);
    return;
  }
  const int gc = gride(c->ncells, kEaBlock);
  if (c->ea_kind == 1 && nb == 8)
    k_ea_cells<X, P, 8, true><<<gc, kEaBlock, 0, c->stream>>>(c->ncells, c->d_gcdofs, planes, sm, x, yc, s);
  else if (c->ea_kind == 1)
    k_ea_cells<X, P, 4, true><<<gc, kEaBlock, 0, c->stream>>>(c->ncells, c->d_gcdofs, planes, sm, x, yc, s);
  else
    k_ea_cells<X, P, 8, false><<<gc, kEaBlock, 0, c->stream>>>(c->ncells, c->d_gcdofs, planes, sm, x, yc, s);
  if (nb == 8)
    k_ea_sum<X, 8><<<g, kEaBlock, 0, c->stream>>>(c->n, c->ncells, c->d_d2c_ptr, c->d_d2c, yc, sm, pd, x, y, part, s);
  else
    k_ea_sum<X, 4><<<g, kEaBlock, 0, c->stream>>>(c->n, c->ncells, c->d_d2c_ptr, c->d_d2c, yc, sm, pd, x, y, part, s);
}

// y = Â x on the element matrices (partial x·y per block); returns the partial count.
int elem_op_launch(fem_ctx* c, const ElemOp* o, const float* x, float* y, double* part, const Scal* s) {
  if (c->ea_fp64) {
    launch_kind<float, float>(c, o, o->keb32, o->sm32, o->pd32, x, y, o->yc32, part, s);
  } else {
    const float* pl = c->ea_kind == 1 ? c->d_keb32 : c->ea_kind == 2 ? c->d_ep32 : c->d_keq32;
    launch_kind<float, float>(c, o, pl, o->sm32, o->pd32, x, y, o->yc32, part, s);
  }
  return o->grid;
}

int elem_op_launch(fem_ctx* c, const ElemOp* o, const double* x, double* y, double* part, const Scal* s) {
  if (c->ea_fp64) {
    const double* pl = c->ea_kind == 1 ? c->d_keb64 : c->ea_kind == 2 ? c->d_ep64 : c->d_keq64;
    launch_kind<double, double>(c, o, pl, o->sm64, o->pd64, x, y, o->yc64, part, s);
  } else {
    const float* pl = c->ea_kind == 1 ? c->d_keb32 : c->ea_kind == 2 ? c->d_ep32 : c->d_keq32;
    launch_kind<double, float>(c, o, pl, o->sm64, o->pd64, x, y, o->yc64, part, s);
  }
  return o->grid;
}

// The element matrices the FP32 (fp32) or FP64 operator reads (a graph key).
const void* elem_op_key(const fem_ctx* c, const ElemOp* o, bool fp32) {
  if (fp32 && c->ea_fp64) return o->keb32;
  if (c->ea_fp64) return c->ea_kind == 1 ? (const void*)c->d_keb64 : c->ea_kind == 2 ? (const void*)c->d_ep64
                                                                                      : (const void*)c->d_keq64;
  return c->ea_kind == 1 ? (const void*)c->d_keb32 : c->ea_kind == 2 ? (const void*)c->d_ep32 : (const void*)c->d_keq32;
}

// apply! on a K that equals its stored element matrices keeps them usable (prescribed rows
// and columns are masked in the operator); called before apply! bumps nzval_version.
void elem_op_note_apply(fem_ctx* c) {
  if (c->ea_kind == 0) return;
  if (c->naff > 0) {   // K := Cᵀ K C couples the masters beyond the element matrices
    c->ea_kind = 0;
    return;
  }
  if (c->ea_version != c->nzval_version || (c->ea_constrained && c->ea_aux != c->aux_version)) {
    c->ea_kind = 0;
    return;
  }
  c->ea_version = c->nzval_version + 1;
  c->ea_constrained = true;
  c->ea_aux = c->aux_version;
}

// After an assembly: K equals the stored element matrices of `kind` (0: none).
void elem_op_note_assembly(fem_ctx* c, int kind, bool fp64, bool fillzero) {
  c->ea_kind = fillzero ? kind : 0;
  c->ea_fp64 = fp64;
  c->ea_version = c->nzval_version;
  c->ea_constrained = false;
}

}  // namespace fem
