// Dirichlet apply!(K, f, ch) on the device, deterministic and atomic-free.
//
// Follows src/Dofs/ConstraintHandler.jl:710-747 step by step:
//   m = meandiag(K)                  :952-958  (before any zeroing)
//   f -= K * g   (only if some g≠0)  :755-789  add_inhomogeneities_csc!
//   zero_out_columns!, zero_out_rows! :931-950
//   K[d,d] = m ; f[d] = g_d * m (0 for apply_zero!)
// The reference's GPU version (ext/FerriteKAExt/constraints.jl:51-61) subtracts
// the inhomogeneities with KA.@atomic; here each row r gathers its own
// contributions K[r,d]·g_d in ascending d, which is the order the serial CPU
// loop applies them in, with the multiply and subtract rounded separately:
// f is bitwise identical to the reference loop's for the same K.
#include "fem_internal.h"

#include <mutex>

namespace fem {

template <int BLOCK>
__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < BLOCK / 64; ++w) r += sh[w];
  return r;
}

// Σ|K_jj| over fixed block partitions (fixed order -> deterministic).
__global__ __launch_bounds__(256) void k_absdiag_partial(int64_t n, const int32_t* __restrict__ diag,
                                                         const double* __restrict__ nzval,
                                                         double* __restrict__ partial) {
  __shared__ double sh[4];
  double acc = 0.0;
  for (int64_t j = blockIdx.x * 256LL + threadIdx.x; j < n; j += (int64_t)gridDim.x * 256) {
    const int32_t p = diag[j];
    if (p >= 0) acc += fabs(nzval[p]);
  }
  const double s = block_sum<256>(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// One block: fixed-order sum of the partials, then the mean.
__global__ __launch_bounds__(256) void k_finish_mean(int nparts, const double* __restrict__ partial, int64_t n,
                                                     double* __restrict__ out) {
  __shared__ double sh[4];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) acc += partial[i];
  const double s = block_sum<256>(acc, sh);
  if (threadIdx.x == 0) out[0] = s / (double)n;
}

// Structurally symmetric pattern: zero column d and row d of every prescribed d directly
// (row d's entries are found in the columns r listed in column d, by binary search).
__global__ void k_zero_prescribed(int64_t np, const int32_t* __restrict__ pres, const int32_t* __restrict__ colptr,
                                  const int32_t* __restrict__ rowval, double* __restrict__ nzval) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t d = pres[i];
    for (int32_t p = colptr[d]; p < colptr[d + 1]; ++p) {
      nzval[p] = 0.0;
      const int32_t r = rowval[p];
      int32_t lo = colptr[r], hi = colptr[r + 1];
      while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (rowval[mid] < d) lo = mid + 1; else hi = mid;
      }
      if (lo < colptr[r + 1] && rowval[lo] == d) nzval[lo] = 0.0;
    }
  }
}

// flag = 0 if some entry (r, j) has no transposed entry (j, r) in the pattern.
__global__ void k_check_symmetric(int64_t n, const int32_t* __restrict__ colptr, const int32_t* __restrict__ rowval,
                                  int* __restrict__ flag) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    for (int32_t p = colptr[j]; p < colptr[j + 1]; ++p) {
      const int32_t r = rowval[p];
      int32_t lo = colptr[r], hi = colptr[r + 1];
      while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (rowval[mid] < (int32_t)j) lo = mid + 1; else hi = mid;
      }
      if (!(lo < colptr[r + 1] && rowval[lo] == (int32_t)j)) atomicExch(flag, 0);
    }
  }
}

// f[r] -= g_d K[r,d] for prescribed d (ascending d), rows found through the
// structurally symmetric pattern: row r's prescribed neighbours d are the rows
// of column r; K[r,d] is located in column d by binary search.  g_d is read from a dense
// by-dof copy (DENSE_G) or found by binary search in the sorted prescribed list; both give
// the same value, so f is the same bit for bit.
template <bool DENSE_G>
__global__ void k_add_inhom(int64_t n, const int32_t* __restrict__ colptr, const int32_t* __restrict__ rowval,
                            const double* __restrict__ nzval, const uint8_t* __restrict__ isc,
                            const double* __restrict__ gfull, int64_t np, const int32_t* __restrict__ pres,
                            const double* __restrict__ g, double* __restrict__ f) {
#pragma clang fp contract(off)
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    double fr = f[r];
    bool touched = false;
    for (int32_t p = colptr[r]; p < colptr[r + 1]; ++p) {
      const int32_t d = rowval[p];
      if (!isc[d]) continue;
      double v;
      if constexpr (DENSE_G) {
        v = gfull[d];
      } else {
        int64_t lo = 0, hi = np;   // isc[d] set: d is in pres[0, np)
        while (lo < hi) {
          const int64_t mid = (lo + hi) >> 1;
          if (pres[mid] < d) lo = mid + 1; else hi = mid;
        }
        v = g[lo];
      }
      if (v == 0.0) continue;
      int32_t lo = colptr[d], hi = colptr[d + 1];
      while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (rowval[mid] < (int32_t)r) lo = mid + 1; else hi = mid;
      }
      if (lo < colptr[d + 1] && rowval[lo] == (int32_t)r) {
        fr = fr - v * nzval[lo];
        touched = true;
      }
    }
    if (touched) f[r] = fr;
  }
}

// Zero every stored entry in a prescribed column or a prescribed row.
__global__ void k_zero_rows_cols(int64_t n, const int32_t* __restrict__ colptr, const int32_t* __restrict__ rowval,
                                 const uint8_t* __restrict__ isc, double* __restrict__ nzval) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const bool cj = isc[j];
    for (int32_t p = colptr[j]; p < colptr[j + 1]; ++p)
      if (cj || isc[rowval[p]]) nzval[p] = 0.0;
  }
}

__global__ void k_set_diag(int64_t np, const int32_t* __restrict__ pres, const int32_t* __restrict__ diag,
                           const double* __restrict__ g, const double* __restrict__ m, int applyzero,
                           double* __restrict__ nzval, double* __restrict__ f) {
#pragma clang fp contract(off)
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t d = pres[i];
    const double mm = m[0];
    const int32_t pd = diag[d];
    if (pd >= 0) nzval[pd] = mm;
    f[d] = applyzero ? 0.0 * mm : g[i] * mm;
  }
}

__global__ void k_scatter_g(int64_t np, const int32_t* __restrict__ pres, const double* __restrict__ g,
                            double* __restrict__ gfull) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x)
    gfull[pres[i]] = g[i];
}

__global__ void k_set_int(int* __restrict__ p, int v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

// ---------------------------------------------------------------------------
// Fixed device scratch of the allocation-free apply! (fem_apply_dirichlet_noalloc): the
// 1024 per-block partial sums of |K_jj|, meandiag(K) and the pattern-symmetry flag.  One
// copy per GPU, held until the process ends; every call on a GPU shares it, so calls that
// overlap on one GPU race.  Each use overwrites what it reads (every block writes its
// partial, the mean kernel writes the mean, a fill kernel sets the flag before the check),
// so nothing has to be cleared between calls.
// ---------------------------------------------------------------------------
constexpr int kScratchMean = kReduceBlocks;       // index of meandiag(K)
constexpr int kScratchFlag = kReduceBlocks + 1;   // the flag, an int stored in a double slot
constexpr int kScratchDoubles = kReduceBlocks + 2;
constexpr size_t kScratchBytes = kScratchDoubles * sizeof(double);   // 8,208 bytes
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};   // device address of `scratch` on each GPU

static int grid_of(int64_t n, int block = 256) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > 8192) g = 8192;
  return (int)g;
}

// Device address of `scratch` for the current device (the caller has made `device` current).
// Host code passes this address, never the array's name, to kernels and copies.
int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return FEM_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return FEM_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return FEM_ERR_HIP;
  g_scratch_ptr[device] = static_cast<double*>(p);
  return FEM_OK;
}

static double* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

}  // namespace fem

using namespace fem;

extern "C" {

int32_t fem_constraints_upload(fem_ctx* c, int64_t np, const int64_t* prescribed, const double* g) {
  if (!c || np < 0 || (np > 0 && (!prescribed || !g)) || c->ndofs == 0)
    return set_error(c, FEM_ERR_ARG, "fem_constraints_upload: invalid arguments (upload dofs first)");
  std::vector<int32_t> p(np > 0 ? np : 1);
  std::vector<uint8_t> isc(c->ndofs, 0);
  bool inhom = false;
  for (int64_t i = 0; i < np; ++i) {
    const int64_t d = prescribed[i];
    if (d < 1 || d > c->ndofs) return set_error(c, FEM_ERR_ARG, "prescribed dof out of range");
    if (i > 0 && d <= prescribed[i - 1]) return set_error(c, FEM_ERR_ARG, "prescribed dofs must be sorted and unique");
    p[i] = (int32_t)(d - 1);
    isc[d - 1] = 1;
    inhom |= (g[i] != 0.0);
  }
  int rc;
  if ((rc = dev_alloc(c, &c->d_pres, np)) || (rc = dev_alloc(c, &c->d_g, np)) || (rc = dev_alloc(c, &c->d_isc, c->ndofs)))
    return rc;
  if (np > 0) {
    FEM_HIP(c, hipMemcpyAsync(c->d_pres, p.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice, c->stream));
    FEM_HIP(c, hipMemcpyAsync(c->d_g, g, sizeof(double) * np, hipMemcpyHostToDevice, c->stream));
  }
  FEM_HIP(c, hipMemcpyAsync(c->d_isc, isc.data(), c->ndofs, hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  c->np = np;
  c->any_inhom = inhom;
  ++c->aux_version;
  return FEM_OK;
}

int32_t fem_apply_dirichlet(fem_ctx* c, int32_t applyzero, double* m_out) {
  if (!c || c->nnz == 0) return set_error(c, FEM_ERR_STATE, "no matrix to apply constraints to");
  if (!c->d_isc) return set_error(c, FEM_ERR_STATE, "upload constraints first");
  FEM_HIP(c, hipSetDevice(c->device));
  const int64_t n = c->n;
  if (c->pattern_sym < 0) {   // once per pattern
    int* dflag = nullptr;
    int rc = dev_alloc(c, &dflag, 1);
    if (rc) return rc;
    const int one = 1;
    int h = 0;
    FEM_HIP(c, hipMemcpyAsync(dflag, &one, sizeof(int), hipMemcpyHostToDevice, c->stream));
    k_check_symmetric<<<grid_of(n), 256, 0, c->stream>>>(n, c->d_colptr, c->d_rowval, dflag);
    FEM_HIP(c, hipMemcpyAsync(&h, dflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    FEM_HIP(c, hipStreamSynchronize(c->stream));
    dev_free(&dflag);
    c->pattern_sym = h;
  }
  if (!applyzero && c->any_inhom && c->pattern_sym != 1)
    return set_error(c, FEM_ERR_ARG, "apply! with inhomogeneous Dirichlet values needs a structurally symmetric pattern");
  FEM_HIP(c, hipEventRecord(c->ev0, c->stream));
  const int nparts = 1024;
  k_absdiag_partial<<<nparts, 256, 0, c->stream>>>(n, c->d_diag, c->d_nzval, c->d_partials);
  k_finish_mean<<<1, 256, 0, c->stream>>>(nparts, c->d_partials, n, c->d_scalar);
  if (!applyzero && c->any_inhom) {
    double* gfull = nullptr;
    int rc = dev_alloc(c, &gfull, n);
    if (rc) return rc;
    FEM_HIP(c, hipMemsetAsync(gfull, 0, sizeof(double) * n, c->stream));
    k_scatter_g<<<grid_of(c->np), 256, 0, c->stream>>>(c->np, c->d_pres, c->d_g, gfull);
    k_add_inhom<true><<<grid_of(n), 256, 0, c->stream>>>(n, c->d_colptr, c->d_rowval, c->d_nzval, c->d_isc, gfull,
                                                          0, nullptr, nullptr, c->d_f);
    FEM_HIP(c, hipStreamSynchronize(c->stream));
    dev_free(&gfull);
  }
  if (c->pattern_sym == 1) {
    if (c->np > 0)
      k_zero_prescribed<<<grid_of(c->np), 256, 0, c->stream>>>(c->np, c->d_pres, c->d_colptr, c->d_rowval,
                                                                 c->d_nzval);
  } else {
    k_zero_rows_cols<<<grid_of(n), 256, 0, c->stream>>>(n, c->d_colptr, c->d_rowval, c->d_isc, c->d_nzval);
  }
  if (c->np > 0)
    k_set_diag<<<grid_of(c->np), 256, 0, c->stream>>>(c->np, c->d_pres, c->d_diag, c->d_g, c->d_scalar, applyzero,
                                                        c->d_nzval, c->d_f);
  FEM_HIP(c, hipGetLastError());
  FEM_HIP(c, hipEventRecord(c->ev1, c->stream));
  c->timings.apply_ms = event_ms(c);
  if (m_out) FEM_HIP(c, hipMemcpy(m_out, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost));
  elem_op_note_apply(c);
  ++c->nzval_version;
  return FEM_OK;
}

int32_t fem_apply_dirichlet_noalloc(fem_ctx* c, int32_t applyzero, double* m_out) {
  if (!c || c->nnz == 0) return set_error(c, FEM_ERR_STATE, "no matrix to apply constraints to");
  if (!c->d_isc) return set_error(c, FEM_ERR_STATE, "upload constraints first");
  double* S = noalloc_scratch(c->device);
  if (!S) return set_error(c, FEM_ERR_STATE, "fem_apply_dirichlet_noalloc: no device scratch address for this GPU");
  // The partials, the mean and the flag are all the scratch this call uses, whatever the
  // problem size (the partial count is fixed), so no size can exceed it.
  static_assert(kScratchBytes >= (kReduceBlocks + 2) * sizeof(double), "scratch too small");
  FEM_HIP(c, hipSetDevice(c->device));
  const int64_t n = c->n;
  if (c->pattern_sym < 0) {   // once per pattern
    int* flag = reinterpret_cast<int*>(S + kScratchFlag);
    int h = 0;
    k_set_int<<<1, 64, 0, c->stream>>>(flag, 1);
    k_check_symmetric<<<grid_of(n), 256, 0, c->stream>>>(n, c->d_colptr, c->d_rowval, flag);
    FEM_HIP(c, hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    FEM_HIP(c, hipStreamSynchronize(c->stream));
    c->pattern_sym = h;
  }
  if (!applyzero && c->any_inhom && c->pattern_sym != 1)
    return set_error(c, FEM_ERR_ARG, "apply! with inhomogeneous Dirichlet values needs a structurally symmetric pattern");
  FEM_HIP(c, hipEventRecord(c->ev0, c->stream));
  k_absdiag_partial<<<kReduceBlocks, 256, 0, c->stream>>>(n, c->d_diag, c->d_nzval, S);
  k_finish_mean<<<1, 256, 0, c->stream>>>(kReduceBlocks, S, n, S + kScratchMean);
  if (!applyzero && c->any_inhom)
    k_add_inhom<false><<<grid_of(n), 256, 0, c->stream>>>(n, c->d_colptr, c->d_rowval, c->d_nzval, c->d_isc,
                                                           nullptr, c->np, c->d_pres, c->d_g, c->d_f);
  if (c->pattern_sym == 1) {
    if (c->np > 0)
      k_zero_prescribed<<<grid_of(c->np), 256, 0, c->stream>>>(c->np, c->d_pres, c->d_colptr, c->d_rowval,
                                                                 c->d_nzval);
  } else {
    k_zero_rows_cols<<<grid_of(n), 256, 0, c->stream>>>(n, c->d_colptr, c->d_rowval, c->d_isc, c->d_nzval);
  }
  if (c->np > 0)
    k_set_diag<<<grid_of(c->np), 256, 0, c->stream>>>(c->np, c->d_pres, c->d_diag, c->d_g, S + kScratchMean,
                                                        applyzero, c->d_nzval, c->d_f);
  FEM_HIP(c, hipGetLastError());
  FEM_HIP(c, hipEventRecord(c->ev1, c->stream));
  c->timings.apply_ms = event_ms(c);
  if (m_out) {
    FEM_HIP(c, hipMemcpyAsync(m_out, S + kScratchMean, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FEM_HIP(c, hipStreamSynchronize(c->stream));
  }
  elem_op_note_apply(c);
  ++c->nzval_version;
  return FEM_OK;
}

}  // extern "C"
