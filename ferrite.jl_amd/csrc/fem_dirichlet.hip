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

// ---------------------------------------------------------------------------
// Affine constraints, apply!'s _condense! (ConstraintHandler.jl:809-867): K := Cᵀ K C and
// f := Cᵀ f in place, before the constrained rows and columns are zeroed.  The reference
// walks the columns in order and adds into the masters' entries; here every target entry
// (i, j) with i and j free of affine constraints gathers its own contributions, visiting
// them in the order the serial loop adds them (source column ascending, then source row
// ascending; products rounded before the add), so K comes out bit for bit the same:
//   source column j:       rows r ∈ S(i):             v_ri K[r, j]
//   source column c ∈ S(j): row i:                    v_cj K[i, c]
//                           rows r ∈ S(i):             v_ri v_cj K[r, c]
// with S(d) the dofs constrained with d as a master, v_rd their coefficient.  Entries read
// belong to affine rows or columns, which no thread writes.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int32_t find_row(const int32_t* __restrict__ colptr, const int32_t* __restrict__ rowval,
                                            int32_t row, int32_t col) {
  int32_t lo = colptr[col], hi = colptr[col + 1];
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (rowval[mid] < row) lo = mid + 1; else hi = mid;
  }
  return (lo < colptr[col + 1] && rowval[lo] == row) ? lo : -1;
}

__global__ void k_condense_K(int64_t n, const int32_t* __restrict__ colptr, const int32_t* __restrict__ rowval,
                             const int32_t* __restrict__ affidx, const int32_t* __restrict__ slp,
                             const int32_t* __restrict__ sld, const double* __restrict__ slv,
                             double* __restrict__ nzval) {
#pragma clang fp contract(off)
  for (int64_t jj = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; jj < n; jj += (int64_t)gridDim.x * blockDim.x) {
    const int32_t j = (int32_t)jj;
    if (affidx[j] >= 0) continue;
    const int32_t sj0 = slp[j], sj1 = slp[j + 1];
    for (int32_t p = colptr[j]; p < colptr[j + 1]; ++p) {
      const int32_t i = rowval[p];
      if (affidx[i] >= 0) continue;
      const int32_t si0 = slp[i], si1 = slp[i + 1];
      if (si0 == si1 && sj0 == sj1) continue;
      double acc = nzval[p];
      // source columns in ascending order: j itself among the slaves c of j
      int32_t q = sj0;
      bool did_j = false;
      while (q < sj1 || !did_j) {
        const bool take_j = !did_j && (q >= sj1 || j < sld[q]);
        if (take_j) {
          did_j = true;
          for (int32_t t = si0; t < si1; ++t) {   // rows r ∈ S(i) of column j, ascending
            const int32_t pr = find_row(colptr, rowval, sld[t], j);
            if (pr < 0) continue;
            const double kv = nzval[pr];
            if (kv != 0.0) acc = acc + slv[t] * kv;
          }
        } else {
          const int32_t cc = sld[q];
          const double vc = slv[q];
          // rows i and r ∈ S(i) of column c, ascending
          int32_t t = si0;
          bool did_i = false;
          while (t < si1 || !did_i) {
            const bool take_i = !did_i && (t >= si1 || i < sld[t]);
            if (take_i) {
              did_i = true;
              const int32_t pc = find_row(colptr, rowval, i, cc);
              if (pc >= 0) {
                const double kv = nzval[pc];
                if (kv != 0.0) acc = acc + vc * kv;
              }
            } else {
              const int32_t pc = find_row(colptr, rowval, sld[t], cc);
              if (pc >= 0) {
                const double kv = nzval[pc];
                if (kv != 0.0) acc = acc + slv[t] * vc * kv;
              }
              ++t;
            }
          }
          ++q;
        }
      }
      nzval[p] = acc;
    }
  }
}

// A contribution whose target entry is absent from the pattern (the reference's addindex!
// error; allocate_matrix(dh, ch) includes them): the first such (row, col) in *bad (atomicMin
// of row * n + col), INT64_MAX if none.  Per affine dof a: column a (rows i) and, by structural
// symmetry, row a (columns j = the same indices).
__global__ void k_condense_check(int64_t naff, int64_t n, const int32_t* __restrict__ aff,
                                 const int32_t* __restrict__ colptr, const int32_t* __restrict__ rowval,
                                 const int32_t* __restrict__ affidx, const int32_t* __restrict__ acp,
                                 const int32_t* __restrict__ acd, const double* __restrict__ nzval,
                                 unsigned long long* __restrict__ bad) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < naff; k += (int64_t)gridDim.x * blockDim.x) {
    const int32_t a = aff[k];
    for (int32_t p = colptr[a]; p < colptr[a + 1]; ++p) {
      const int32_t i = rowval[p];
      const int32_t ia = affidx[i];
      const double kia = nzval[p];                                // K[i, a]
      const int32_t pt = find_row(colptr, rowval, a, i);
      const double kai = pt >= 0 ? nzval[pt] : 0.0;               // K[a, i]
      for (int32_t u = acp[k]; u < acp[k + 1]; ++u) {
        const int32_t m = acd[u];
        if (ia < 0) {
          if (kia != 0.0 && find_row(colptr, rowval, i, m) < 0)   // target (i, m)
            atomicMin(bad, (unsigned long long)i * n + m);
          if (kai != 0.0 && find_row(colptr, rowval, m, i) < 0)   // target (m, i)
            atomicMin(bad, (unsigned long long)m * n + i);
        } else if (kia != 0.0) {
          for (int32_t w = acp[ia]; w < acp[ia + 1]; ++w)         // target (m_i, m)
            if (find_row(colptr, rowval, acd[w], m) < 0) atomicMin(bad, (unsigned long long)acd[w] * n + m);
        }
      }
    }
  }
}

// f_d += Σ_{c ∈ S(d)} v_cd f_c in ascending c (the masters), then f_c = 0 (the slaves).
__global__ void k_condense_f(int64_t n, const int32_t* __restrict__ slp, const int32_t* __restrict__ sld,
                             const double* __restrict__ slv, double* __restrict__ f) {
#pragma clang fp contract(off)
  for (int64_t d = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; d < n; d += (int64_t)gridDim.x * blockDim.x) {
    const int32_t t0 = slp[d], t1 = slp[d + 1];
    if (t0 == t1) continue;
    double v = f[d];
    for (int32_t t = t0; t < t1; ++t) v = v + f[sld[t]] * slv[t];
    f[d] = v;
  }
}

__global__ void k_zero_slaves(int64_t naff, const int32_t* __restrict__ aff, double* __restrict__ f) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < naff; k += (int64_t)gridDim.x * blockDim.x)
    f[aff[k]] = 0.0;
}

// apply!(u, ch) (ConstraintHandler.jl:686-700): u[prescribed] = g (0 for apply_zero!), then for
// every affine dof u_c = b_c + Σ a u_m (0 + Σ for apply_zero!), masters never affine.
__global__ void k_set_prescribed(int64_t np, const int32_t* __restrict__ pres, const double* __restrict__ g,
                                 int applyzero, double* __restrict__ u) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x)
    u[pres[i]] = applyzero ? 0.0 : g[i];
}

__global__ void k_set_affine(int64_t naff, const int32_t* __restrict__ aff, const int32_t* __restrict__ acp,
                             const int32_t* __restrict__ acd, const double* __restrict__ acv,
                             const double* __restrict__ acb, int applyzero, double* __restrict__ u) {
#pragma clang fp contract(off)
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < naff; k += (int64_t)gridDim.x * blockDim.x) {
    double v = applyzero ? 0.0 : acb[k];
    for (int32_t t = acp[k]; t < acp[k + 1]; ++t) v = v + acv[t] * u[acd[t]];
    u[aff[k]] = v;
  }
}

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
  c->naff = 0;   // affine constraints are uploaded after the prescribed dofs they belong to
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
  if (c->naff > 0 && c->pattern_sym != 1)
    return set_error(c, FEM_ERR_ARG, "apply! with affine constraints needs a structurally symmetric pattern");
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
  if (c->naff > 0) {   // K := Cᵀ K C, f := Cᵀ f (ConstraintHandler.jl:809-867)
    unsigned long long* dbad = nullptr;
    int rc = dev_alloc(c, &dbad, 1);
    if (rc) return rc;
    const unsigned long long none = ~0ull;
    unsigned long long hbad = none;
    FEM_HIP(c, hipMemcpyAsync(dbad, &none, sizeof(none), hipMemcpyHostToDevice, c->stream));
    k_condense_check<<<grid_of(c->naff), 256, 0, c->stream>>>(c->naff, n, c->d_afd, c->d_colptr, c->d_rowval,
                                                               c->d_affidx, c->d_acp, c->d_acd, c->d_nzval, dbad);
    FEM_HIP(c, hipMemcpyAsync(&hbad, dbad, sizeof(hbad), hipMemcpyDeviceToHost, c->stream));
    FEM_HIP(c, hipStreamSynchronize(c->stream));
    dev_free(&dbad);
    if (hbad != none) {
      c->err_row = (int64_t)(hbad / (unsigned long long)n) + 1;
      c->err_col = (int64_t)(hbad % (unsigned long long)n) + 1;
      char buf[400];
      snprintf(buf, sizeof buf,
               "You are trying to assemble values in to K[%lld, %lld], but K[%lld, %lld] is missing in the "
               "sparsity pattern. Make sure you have called `K = allocate_matrix(dh, ch)` if you have affine "
               "constraints.",
               (long long)c->err_row, (long long)c->err_col, (long long)c->err_row, (long long)c->err_col);
      return set_error(c, FEM_ERR_PATTERN, buf);
    }
    k_condense_K<<<grid_of(n), 256, 0, c->stream>>>(n, c->d_colptr, c->d_rowval, c->d_affidx, c->d_slp, c->d_sld,
                                                     c->d_slv, c->d_nzval);
    k_condense_f<<<grid_of(n), 256, 0, c->stream>>>(n, c->d_slp, c->d_sld, c->d_slv, c->d_f);
    k_zero_slaves<<<grid_of(c->naff), 256, 0, c->stream>>>(c->naff, c->d_afd, c->d_f);
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
  if (c->naff > 0)
    return set_error(c, FEM_ERR_ARG, "fem_apply_dirichlet_noalloc: affine constraints are condensed by fem_apply_dirichlet");
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

int32_t fem_affine_upload(fem_ctx* c, int64_t naff, const int64_t* dofs, const int64_t* coef_ptr,
                          const int64_t* coef_dofs, const double* coef_vals, const double* b) {
  if (!c || naff < 0 || (naff > 0 && (!dofs || !coef_ptr || !coef_vals || !b)))
    return set_error(c, FEM_ERR_ARG, "fem_affine_upload: invalid arguments");
  if (!c->d_isc || c->ndofs == 0) return set_error(c, FEM_ERR_STATE, "upload the prescribed dofs first");
  const int64_t n = c->ndofs;
  std::vector<uint8_t> isc(n);
  FEM_HIP(c, hipMemcpy(isc.data(), c->d_isc, n, hipMemcpyDeviceToHost));
  std::vector<int32_t> idx(n, -1), aff(naff > 0 ? naff : 1), acp(naff + 1), acd, sl_cnt(n + 1, 0);
  std::vector<double> acv;
  for (int64_t k = 0; k < naff; ++k) {
    const int64_t d = dofs[k];
    if (d < 1 || d > n || (k > 0 && d <= dofs[k - 1]))
      return set_error(c, FEM_ERR_ARG, "affine dofs must be in range, sorted and unique");
    if (!isc[d - 1]) return set_error(c, FEM_ERR_ARG, "affine dofs must be among the prescribed dofs");
    idx[d - 1] = (int32_t)k;
    aff[k] = (int32_t)(d - 1);
  }
  acp[0] = 0;
  for (int64_t k = 0; k < naff; ++k) {
    if (coef_ptr[k + 1] < coef_ptr[k]) return set_error(c, FEM_ERR_ARG, "coef_ptr not monotone");
    for (int64_t t = coef_ptr[k]; t < coef_ptr[k + 1]; ++t) {
      const int64_t m = coef_dofs[t];
      if (m < 1 || m > n) return set_error(c, FEM_ERR_ARG, "master dof out of range");
      if (idx[m - 1] >= 0) return set_error(c, FEM_ERR_ARG, "nested affine constraints currently not supported");
      acd.push_back((int32_t)(m - 1));
      acv.push_back(coef_vals[t]);
      ++sl_cnt[m];
    }
    acp[k + 1] = (int32_t)acd.size();
  }
  // per master: slaves in ascending slave dof (the affine dofs are ascending)
  for (int64_t m = 0; m < n; ++m) sl_cnt[m + 1] += sl_cnt[m];
  std::vector<int32_t> slp(sl_cnt), fill(sl_cnt.begin(), sl_cnt.end() - 1), sld(acd.size());
  std::vector<double> slv(acd.size());
  for (int64_t k = 0; k < naff; ++k)
    for (int32_t t = acp[k]; t < acp[k + 1]; ++t) {
      const int32_t q = fill[acd[t]]++;
      sld[q] = aff[k];
      slv[q] = acv[t];
    }
  const size_t nc = std::max<size_t>(acd.size(), 1);
  int rc;
  if ((rc = dev_alloc(c, &c->d_affidx, n)) || (rc = dev_alloc(c, &c->d_afd, aff.size())) ||
      (rc = dev_alloc(c, &c->d_acp, naff + 1)) || (rc = dev_alloc(c, &c->d_acd, nc)) ||
      (rc = dev_alloc(c, &c->d_acv, nc)) || (rc = dev_alloc(c, &c->d_acb, aff.size())) ||
      (rc = dev_alloc(c, &c->d_slp, n + 1)) || (rc = dev_alloc(c, &c->d_sld, nc)) || (rc = dev_alloc(c, &c->d_slv, nc)))
    return rc;
  FEM_HIP(c, hipMemcpyAsync(c->d_affidx, idx.data(), 4 * n, hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipMemcpyAsync(c->d_afd, aff.data(), 4 * aff.size(), hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipMemcpyAsync(c->d_acp, acp.data(), 4 * (naff + 1), hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipMemcpyAsync(c->d_slp, slp.data(), 4 * (n + 1), hipMemcpyHostToDevice, c->stream));
  if (!acd.empty()) {
    FEM_HIP(c, hipMemcpyAsync(c->d_acd, acd.data(), 4 * acd.size(), hipMemcpyHostToDevice, c->stream));
    FEM_HIP(c, hipMemcpyAsync(c->d_acv, acv.data(), 8 * acv.size(), hipMemcpyHostToDevice, c->stream));
    FEM_HIP(c, hipMemcpyAsync(c->d_sld, sld.data(), 4 * sld.size(), hipMemcpyHostToDevice, c->stream));
    FEM_HIP(c, hipMemcpyAsync(c->d_slv, slv.data(), 8 * slv.size(), hipMemcpyHostToDevice, c->stream));
  }
  if (naff > 0) FEM_HIP(c, hipMemcpyAsync(c->d_acb, b, 8 * naff, hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  c->naff = naff;
  ++c->aux_version;
  return FEM_OK;
}

int32_t fem_apply_solution(fem_ctx* c, int32_t applyzero) {
  if (!c || !c->d_u) return set_error(c, FEM_ERR_STATE, "no solution vector");
  if (!c->d_isc) return set_error(c, FEM_ERR_STATE, "upload constraints first");
  FEM_HIP(c, hipSetDevice(c->device));
  if (c->np > 0)
    k_set_prescribed<<<grid_of(c->np), 256, 0, c->stream>>>(c->np, c->d_pres, c->d_g, applyzero, c->d_u);
  if (c->naff > 0)
    k_set_affine<<<grid_of(c->naff), 256, 0, c->stream>>>(c->naff, c->d_afd, c->d_acp, c->d_acd, c->d_acv, c->d_acb,
                                                          applyzero, c->d_u);
  FEM_HIP(c, hipGetLastError());
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  return FEM_OK;
}

}  // extern "C"
