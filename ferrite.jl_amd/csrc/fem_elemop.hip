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
// the result is bitwise reproducible.
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

// Scalar / small vector elements: one lane per row, walking the row's (cell, local index)
// list d2c in ascending cell id.  ROWSUM: strict-upper planes, diagonal from the row sum
// (kind 1); otherwise upper planes with the diagonal (kind 2).
template <typename X, typename P, int NB, bool ROWSUM>
__global__ __launch_bounds__(kEaBlock) void k_ea_rows(int64_t n, int64_t ncells, const int32_t* __restrict__ d2c_ptr,
                                                      const int32_t* __restrict__ d2c, const int32_t* __restrict__ cdofs,
                                                      const P* __restrict__ planes, const X* __restrict__ sm,
                                                      const X* __restrict__ pd, const X* __restrict__ x,
                                                      X* __restrict__ y, double* __restrict__ part,
                                                      const Scal* __restrict__ s) {
  if (s->done || s->fresh) return;   // the recurrence stopped, or this step only moves x̂, r̂
  __shared__ double sh[kEaBlock / 64];
  double d = 0.0;
  for (int64_t row = blockIdx.x * (int64_t)kEaBlock + threadIdx.x; row < n; row += (int64_t)gridDim.x * kEaBlock) {
    X acc = X(0);
    const int32_t e0 = d2c_ptr[row], e1 = d2c_ptr[row + 1];
    for (int32_t e = e0; e < e1; ++e) {
      const int32_t pk = d2c[e];
      const int64_t cell = pk / NB;
      const int a = pk - (int)cell * NB;
      int32_t dof[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) dof[b] = cdofs[cell * NB + b];
      X v[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int lo = a < b ? a : b, hi = a < b ? b : a;
        if constexpr (ROWSUM) v[b] = a == b ? X(0) : (X)planes[(int64_t)ea_off_idx(lo, hi, NB) * ncells + cell];
        else v[b] = (X)planes[(int64_t)tri_idx(lo, hi, NB) * ncells + cell];
      }
      if constexpr (ROWSUM) {
        X dg = X(0);
#pragma unroll
        for (int b = 0; b < NB; ++b) dg -= v[b];   // v[a] is 0 here
#pragma unroll
        for (int b = 0; b < NB; ++b)
          if (b == a) v[b] = dg;
      }
#pragma unroll
      for (int b = 0; b < NB; ++b) acc += v[b] * (sm[dof[b]] * x[dof[b]]);
    }
    const X xr = x[row];
    const X yv = sm[row] * acc + pd[row] * xr;
    y[row] = yv;
    d += (double)xr * (double)yv;
  }
  const double bsum = ea_block_sum(d, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bsum;
}

// Q2 hex³: one wave per dof node m (rows 3m + j), walking the node's (cell, local node b)
// list n2c.  Lane slot t = lane + 64 k of the node's 252-value block holds Ke[(a, i), (b, j)]
// with (i, j) = t / 28, a = t % 28, i.e. (Ke is symmetric) the coupling of row (b, j) with
// column (a, i): it multiplies t_{3 node(a) + i} into the partial sum of component j.
template <typename X, typename P>
__global__ __launch_bounds__(kEaBlock) void k_ea_q2(int64_t nnode, const int32_t* __restrict__ n2c_ptr,
                                                    const int32_t* __restrict__ n2c, const int32_t* __restrict__ cnode,
                                                    const P* __restrict__ keb, const X* __restrict__ sm,
                                                    const X* __restrict__ pd, const X* __restrict__ x,
                                                    X* __restrict__ y, double* __restrict__ part,
                                                    const Scal* __restrict__ s) {
  if (s->done || s->fresh) return;
  constexpr int kKS = 252, kKC = 27 * kKS, S4 = 4;
  __shared__ double sh[kEaBlock / 64];
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * (kEaBlock / 64);
  int sa[S4], si[S4], sj[S4];
#pragma unroll
  for (int k = 0; k < S4; ++k) {
    const int t = lane + 64 * k;
    const int ij = t / 28, a = t - 28 * ij;
    sa[k] = (t < kKS && a < 27) ? a : -1;
    si[k] = ij / 3;
    sj[k] = ij - 3 * (ij / 3);
  }
  double d = 0.0;
  for (int64_t m = blockIdx.x * (int64_t)(kEaBlock / 64) + (threadIdx.x >> 6); m < nnode; m += waves) {
    X acc0 = X(0), acc1 = X(0), acc2 = X(0);
    const int32_t e0 = n2c_ptr[m], e1 = n2c_ptr[m + 1];
    for (int32_t e = e0; e < e1; ++e) {
      const int32_t ent = n2c[e];
      const int64_t cell = ent / 27;
      const int b = ent - (int)cell * 27;
      const P* blk = keb + cell * (int64_t)kKC + b * kKS;
#pragma unroll
      for (int k = 0; k < S4; ++k) {
        if (sa[k] < 0) continue;
        const int64_t col = 3 * (int64_t)cnode[cell * 27 + sa[k]] + si[k];
        const X c = (X)blk[lane + 64 * k] * (sm[col] * x[col]);
        acc0 += sj[k] == 0 ? c : X(0);
        acc1 += sj[k] == 1 ? c : X(0);
        acc2 += sj[k] == 2 ? c : X(0);
      }
    }
    acc0 = ea_wave_sum(acc0);
    acc1 = ea_wave_sum(acc1);
    acc2 = ea_wave_sum(acc2);
    if (lane == 0) {
      const X av[3] = {acc0, acc1, acc2};
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int64_t row = 3 * m + j;
        const X xr = x[row];
        const X yv = sm[row] * av[j] + pd[row] * xr;
        y[row] = yv;
        d += (double)xr * (double)yv;
      }
    }
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
    o->n = n;
  }
  const uint8_t* isc = c->ea_constrained ? c->d_isc : nullptr;
  if (fp32) {
    if (!o->sm32 && ((rc = dev_alloc(c, &o->sm32, n)) || (rc = dev_alloc(c, &o->pd32, n)))) return rc;
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
    k_ea_prep<double><<<gride(n), 256, 0, c->stream>>>(n, sc, isc, c->d_diag, c->d_nzval, o->sm64, o->pd64);
  }
  FEM_HIP(c, hipGetLastError());
  *grid = o->grid;
  return FEM_OK;
}

template <typename X, typename P>
static void launch_kind(fem_ctx* c, const ElemOp* o, const P* planes, const X* sm, const X* pd, const X* x, X* y,
                        double* part, const Scal* s) {
  const int g = o->grid;
  const int nb = c->ndpc;
  if (c->ea_kind == 3) {
    k_ea_q2<X, P><<<g, kEaBlock, 0, c->stream>>>(c->n / 3, c->d_n2c_ptr, c->d_n2c, c->d_q2cnode, planes, sm, pd, x, y,
                                                 part, s);
  } else if (c->ea_kind == 1) {
    if (nb == 8)
      k_ea_rows<X, P, 8, true><<<g, kEaBlock, 0, c->stream>>>(c->n, c->ncells, c->d_d2c_ptr, c->d_d2c, c->d_gcdofs,
                                                              planes, sm, pd, x, y, part, s);
    else
      k_ea_rows<X, P, 4, true><<<g, kEaBlock, 0, c->stream>>>(c->n, c->ncells, c->d_d2c_ptr, c->d_d2c, c->d_gcdofs,
                                                              planes, sm, pd, x, y, part, s);
  } else {
    k_ea_rows<X, P, 8, false><<<g, kEaBlock, 0, c->stream>>>(c->n, c->ncells, c->d_d2c_ptr, c->d_d2c, c->d_gcdofs,
                                                             planes, sm, pd, x, y, part, s);
  }
}

// y = Â x on the element matrices (partial x·y per block); returns the partial count.
int elem_op_launch(fem_ctx* c, const ElemOp* o, const float* x, float* y, double* part, const Scal* s) {
  if (c->ea_fp64) {
    launch_kind<float, float>(c, o, o->keb32, o->sm32, o->pd32, x, y, part, s);
  } else {
    const float* pl = c->ea_kind == 1 ? c->d_keb32 : c->ea_kind == 2 ? c->d_ep32 : c->d_keq32;
    launch_kind<float, float>(c, o, pl, o->sm32, o->pd32, x, y, part, s);
  }
  return o->grid;
}

int elem_op_launch(fem_ctx* c, const ElemOp* o, const double* x, double* y, double* part, const Scal* s) {
  if (c->ea_fp64) {
    const double* pl = c->ea_kind == 1 ? c->d_keb64 : c->ea_kind == 2 ? c->d_ep64 : c->d_keq64;
    launch_kind<double, double>(c, o, pl, o->sm64, o->pd64, x, y, part, s);
  } else {
    const float* pl = c->ea_kind == 1 ? c->d_keb32 : c->ea_kind == 2 ? c->d_ep32 : c->d_keq32;
    launch_kind<double, float>(c, o, pl, o->sm64, o->pd64, x, y, part, s);
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
