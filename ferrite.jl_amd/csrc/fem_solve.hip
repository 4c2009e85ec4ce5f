// Solvers on the assembled, constrained K (replaces u = K \ f, heat_equation.jl:218;
// the reference has no iterative solver on this path).
//
//  * FEM_SOLVE_CG_FP64   Jacobi-preconditioned CG, FP64 SpMV and vectors.
//  * FEM_SOLVE_IR_MIXED  FP64 iterative refinement: r = f − K u in FP64, inner
//                        Jacobi-PCG on an FP32 copy of K with FP32 vectors and
//                        FP64-accumulated dot products, u += d in FP64.
//
// K is symmetric after apply! (rows and columns of prescribed dofs are zeroed
// together), so the CSC columns double as CSR rows: y = K x is a row-gather
// y[j] = Σ_p nzval[p] x[rowval[p]] over column j, with no atomics.
// Every reduction is a fixed-shape two-level tree (per-block partials, then the
// last block to finish sums them in a fixed order), so results are bitwise
// reproducible; the CG scalars (α, β) never leave the device.  The host looks
// at the residual only every few iterations; kernels launched past
// convergence return immediately on a device flag.
#include <cmath>

#include "fem_internal.h"

namespace fem {

constexpr int kBlock = 256;
constexpr int kParts = 1024;      // fixed number of blocks of every reducing kernel

// Device-side solver state (one cache line apart to avoid false sharing).
struct Scal {
  double rz, rz_new, pq, rr, alpha, beta, bnorm2, pad0;
  unsigned int counter[4];
  int done;
  int iters;
  int pad1[2];
};

struct SolverWork {
  int64_t n = 0;
  double *r = nullptr, *p = nullptr, *q = nullptr, *invd = nullptr, *res = nullptr;
  float *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *x32 = nullptr, *invd32 = nullptr;
  double* part = nullptr;     // [4][kParts]
  Scal* s = nullptr;
};

void free_solver(fem_ctx* c) {
  if (!c->solver) return;
  SolverWork* w = c->solver;
  dev_free(&w->r); dev_free(&w->p); dev_free(&w->q); dev_free(&w->invd); dev_free(&w->res);
  dev_free(&w->r32); dev_free(&w->p32); dev_free(&w->q32); dev_free(&w->x32); dev_free(&w->invd32);
  dev_free(&w->part); dev_free(&w->s);
  delete w;
  c->solver = nullptr;
}

// --------------------------------------------------------------------------
// reductions
// --------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Block sum; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kBlock / 64; ++i) r += sh[i];
  return r;
}

// Sum of part[0..kParts) in a fixed order by one block; result in thread 0.
__device__ __forceinline__ double sum_parts(const double* part, double* sh) {
  double v = 0.0;
  for (int i = threadIdx.x; i < kParts; i += kBlock) v += part[i];
  return block_sum(v, sh);
}

// Returns true in every thread of the block that finishes last.
__device__ __forceinline__ bool last_block(unsigned int* counter) {
  __shared__ bool last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned int t = atomicAdd(counter, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  return last;
}

// --------------------------------------------------------------------------
// SpMV  y = K x (+ optional fused dots), G lanes per row
// --------------------------------------------------------------------------
template <int G, typename V, typename X, typename A>
__device__ __forceinline__ A row_dot(int64_t j, const int32_t* __restrict__ colptr,
                                     const int32_t* __restrict__ rowval, const V* __restrict__ val,
                                     const X* __restrict__ x, int sub) {
  A acc = A(0);
  const int32_t p0 = colptr[j], p1 = colptr[j + 1];
  for (int32_t p = p0 + sub; p < p1; p += G) acc += (A)val[p] * (A)x[rowval[p]];
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_down(acc, o, G);
  return acc;
}

// q = K p ; part0 = Σ p·q.  Last block computes α = rz / pq.
template <int G, typename V, typename X>
__global__ __launch_bounds__(kBlock) void k_spmv_pq(int64_t n, const int32_t* __restrict__ colptr,
                                                    const int32_t* __restrict__ rowval,
                                                    const V* __restrict__ val, const X* __restrict__ p,
                                                    X* __restrict__ q, double* __restrict__ part,
                                                    Scal* __restrict__ s) {
  if (s->done) return;
  __shared__ double sh[kBlock / 64];
  using A = typename std::conditional<sizeof(X) == 8, double, float>::type;
  const int sub = threadIdx.x % G;
  const int64_t rows_per_pass = (int64_t)gridDim.x * (kBlock / G);
  double dot = 0.0;
  for (int64_t j = (int64_t)blockIdx.x * (kBlock / G) + threadIdx.x / G; j < n; j += rows_per_pass) {
    const A y = row_dot<G, V, X, A>(j, colptr, rowval, val, p, sub);
    if (sub == 0) {
      q[j] = (X)y;
      dot += (double)p[j] * (double)(X)y;
    }
  }
  const double b = block_sum(dot, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
  if (last_block(&s->counter[0])) {
    const double pq = sum_parts(part, sh);
    if (threadIdx.x == 0) {
      s->pq = pq;
      s->alpha = s->rz / pq;
      s->counter[0] = 0;
    }
  }
}

// x += α p ; r -= α q ; z = D⁻¹ r ; partials of r·z and r·r.  Last block: β, convergence flag.
template <typename X>
__global__ __launch_bounds__(kBlock) void k_cg_update(int64_t n, X* __restrict__ x, X* __restrict__ r,
                                                      const X* __restrict__ p, const X* q,
                                                      const X* __restrict__ invd, X* z_out,
                                                      double* __restrict__ part, Scal* __restrict__ s,
                                                      double tol2) {
  if (s->done) return;
  __shared__ double sh[kBlock / 64];
  const X alpha = (X)s->alpha;
  double rz = 0.0, rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    x[i] += alpha * p[i];
    const X ri = r[i] - alpha * q[i];
    r[i] = ri;
    const X zi = ri * invd[i];
    z_out[i] = zi;
    rz += (double)ri * (double)zi;
    rr += (double)ri * (double)ri;
  }
  double b = block_sum(rz, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
  b = block_sum(rr, sh);
  if (threadIdx.x == 0) part[kParts + blockIdx.x] = b;
  if (last_block(&s->counter[1])) {
    const double rzn = sum_parts(part, sh);
    const double rrn = sum_parts(part + kParts, sh);
    if (threadIdx.x == 0) {
      s->beta = rzn / s->rz;
      s->rz = rzn;
      s->rr = rrn;
      s->iters += 1;
      if (rrn <= tol2) s->done = 1;
      s->counter[1] = 0;
    }
  }
}

// p = z + β p   (z stored in q's slot by the update kernel's caller)
template <typename X>
__global__ __launch_bounds__(kBlock) void k_p_update(int64_t n, X* __restrict__ p, const X* __restrict__ z,
                                                     const Scal* __restrict__ s) {
  if (s->done) return;
  const X beta = (X)s->beta;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    p[i] = z[i] + beta * p[i];
}

// Start: z = D⁻¹ r, p = z, rz = r·z, rr = r·r, done = (rr <= tol2).
template <typename X>
__global__ __launch_bounds__(kBlock) void k_cg_init(int64_t n, const X* __restrict__ r, const X* __restrict__ invd,
                                                    X* __restrict__ p, double* __restrict__ part,
                                                    Scal* __restrict__ s, double tol2) {
  __shared__ double sh[kBlock / 64];
  double rz = 0.0, rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const X ri = r[i];
    const X zi = ri * invd[i];
    p[i] = zi;
    rz += (double)ri * (double)zi;
    rr += (double)ri * (double)ri;
  }
  double b = block_sum(rz, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
  b = block_sum(rr, sh);
  if (threadIdx.x == 0) part[kParts + blockIdx.x] = b;
  if (last_block(&s->counter[2])) {
    const double rzn = sum_parts(part, sh);
    const double rrn = sum_parts(part + kParts, sh);
    if (threadIdx.x == 0) {
      s->rz = rzn;
      s->rr = rrn;
      s->done = (rrn <= tol2) ? 1 : 0;
      s->counter[2] = 0;
    }
  }
}

// res = f − K u (FP64) ; part = Σ res² ; optionally scaled FP32 copy r32 = res * scale is done separately.
template <int G>
__global__ __launch_bounds__(kBlock) void k_residual(int64_t n, const int32_t* __restrict__ colptr,
                                                     const int32_t* __restrict__ rowval,
                                                     const double* __restrict__ val, const double* __restrict__ u,
                                                     const double* __restrict__ f, double* __restrict__ res,
                                                     double* __restrict__ part, Scal* __restrict__ s) {
  __shared__ double sh[kBlock / 64];
  const int sub = threadIdx.x % G;
  const int64_t rows_per_pass = (int64_t)gridDim.x * (kBlock / G);
  double acc = 0.0;
  for (int64_t j = (int64_t)blockIdx.x * (kBlock / G) + threadIdx.x / G; j < n; j += rows_per_pass) {
    const double y = row_dot<G, double, double, double>(j, colptr, rowval, val, u, sub);
    if (sub == 0) {
      const double rj = f[j] - y;
      res[j] = rj;
      acc += rj * rj;
    }
  }
  const double b = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
  if (last_block(&s->counter[3])) {
    const double rr = sum_parts(part, sh);
    if (threadIdx.x == 0) {
      s->rr = rr;
      s->counter[3] = 0;
    }
  }
}

__global__ void k_sumsq(int64_t n, const double* __restrict__ v, double* __restrict__ part, Scal* __restrict__ s) {
  __shared__ double sh[kBlock / 64];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    acc += v[i] * v[i];
  const double b = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
  if (last_block(&s->counter[3])) {
    const double t = sum_parts(part, sh);
    if (threadIdx.x == 0) {
      s->bnorm2 = t;
      s->counter[3] = 0;
    }
  }
}

template <typename X>
__global__ void k_invdiag(int64_t n, const int32_t* __restrict__ diag, const double* __restrict__ nzval,
                          int jacobi, X* __restrict__ invd) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const int32_t p = diag[j];
    const double d = (p >= 0) ? nzval[p] : 0.0;
    invd[j] = (jacobi && d != 0.0) ? (X)(1.0 / d) : (X)1.0;
  }
}

__global__ void k_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// r32 = res / ‖res‖ ; x32 = 0
__global__ void k_scale_to_f32(int64_t n, const double* __restrict__ res, const Scal* __restrict__ s,
                               float* __restrict__ r32, float* __restrict__ x32) {
  const double inv = 1.0 / sqrt(s->rr);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    r32[i] = (float)(res[i] * inv);
    x32[i] = 0.0f;
  }
}

// u += ‖res‖ d
__global__ void k_correct(int64_t n, const float* __restrict__ d, double scale, double* __restrict__ u) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    u[i] += scale * (double)d[i];
}

__global__ void k_copy(int64_t n, const double* __restrict__ a, double* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

__global__ void k_zero(int64_t n, double* __restrict__ a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0.0;
}

__global__ void k_reset_scal(Scal* s) {
  s->rz = s->rz_new = s->pq = s->rr = s->alpha = s->beta = 0.0;
  for (int i = 0; i < 4; ++i) s->counter[i] = 0;
  s->done = 0;
  s->iters = 0;
}

// --------------------------------------------------------------------------
// host driver
// --------------------------------------------------------------------------
static int lanes_per_row(const fem_ctx* c) {
  const double avg = (double)c->nnz / (double)c->n;
  if (avg <= 6) return 4;
  if (avg <= 14) return 8;
  if (avg <= 40) return 16;
  if (avg <= 90) return 32;
  return 64;
}

template <typename V, typename X>
static void launch_spmv_pq(fem_ctx* c, int G, const V* val, const X* p, X* q, SolverWork* w) {
  switch (G) {
    case 4: k_spmv_pq<4, V, X><<<kParts, kBlock, 0, c->stream>>>(c->n, c->d_colptr, c->d_rowval, val, p, q, w->part, w->s); break;
    case 8: k_spmv_pq<8, V, X><<<kParts, kBlock, 0, c->stream>>>(c->n, c->d_colptr, c->d_rowval, val, p, q, w->part, w->s); break;
    case 16: k_spmv_pq<16, V, X><<<kParts, kBlock, 0, c->stream>>>(c->n, c->d_colptr, c->d_rowval, val, p, q, w->part, w->s); break;
    case 32: k_spmv_pq<32, V, X><<<kParts, kBlock, 0, c->stream>>>(c->n, c->d_colptr, c->d_rowval, val, p, q, w->part, w->s); break;
    default: k_spmv_pq<64, V, X><<<kParts, kBlock, 0, c->stream>>>(c->n, c->d_colptr, c->d_rowval, val, p, q, w->part, w->s); break;
  }
}

static void launch_residual(fem_ctx* c, int G, SolverWork* w) {
  switch (G) {
    case 4: k_residual<4><<<kParts, kBlock, 0, c->stream>>>(c->n, c->d_colptr, c->d_rowval, c->d_nzval, c->d_u, c->d_f, w->res, w->part, w->s); break;
    case 8: k_residual<8><<<kParts, kBlock, 0, c->stream>>>(c->n, c->d_colptr, c->d_rowval, c->d_nzval, c->d_u, c->d_f, w->res, w->part, w->s); break;
    case 16: k_residual<16><<<kParts, kBlock, 0, c->stream>>>(c->n, c->d_colptr, c->d_rowval, c->d_nzval, c->d_u, c->d_f, w->res, w->part, w->s); break;
    case 32: k_residual<32><<<kParts, kBlock, 0, c->stream>>>(c->n, c->d_colptr, c->d_rowval, c->d_nzval, c->d_u, c->d_f, w->res, w->part, w->s); break;
    default: k_residual<64><<<kParts, kBlock, 0, c->stream>>>(c->n, c->d_colptr, c->d_rowval, c->d_nzval, c->d_u, c->d_f, w->res, w->part, w->s); break;
  }
}

static int ensure_work(fem_ctx* c, bool mixed) {
  if (!c->solver) c->solver = new SolverWork();
  SolverWork* w = c->solver;
  const int64_t n = c->n;
  int rc;
  if (w->n != n || !w->r) {
    if ((rc = dev_alloc(c, &w->r, n)) || (rc = dev_alloc(c, &w->p, n)) || (rc = dev_alloc(c, &w->q, n)) ||
        (rc = dev_alloc(c, &w->invd, n)) || (rc = dev_alloc(c, &w->res, n)) || (rc = dev_alloc(c, &w->part, 4 * kParts)) ||
        (rc = dev_alloc(c, &w->s, 1)))
      return rc;
    dev_free(&w->r32);
    w->n = n;
  }
  if (mixed && !w->r32) {
    if ((rc = dev_alloc(c, &w->r32, n)) || (rc = dev_alloc(c, &w->p32, n)) || (rc = dev_alloc(c, &w->q32, n)) ||
        (rc = dev_alloc(c, &w->x32, n)) || (rc = dev_alloc(c, &w->invd32, n)))
      return rc;
  }
  return FEM_OK;
}

static int read_scal(fem_ctx* c, SolverWork* w, Scal* h) {
  FEM_HIP(c, hipMemcpyAsync(h, w->s, sizeof(Scal), hipMemcpyDeviceToHost, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  return FEM_OK;
}

// Preconditioned CG on K (FP64: V = X = double; mixed inner: V = X = float).
// Solves K x = r starting from x = 0 (r overwritten with the residual).
template <typename V, typename X>
static int pcg(fem_ctx* c, SolverWork* w, const V* val, X* x, X* r, X* p, X* q, const X* invd, double tol2,
               int maxit, int* iters, double* spmv_ms, int* spmv_count) {
  const int G = lanes_per_row(c);
  const int64_t n = c->n;
  const int vgrid = kParts;
  k_reset_scal<<<1, 1, 0, c->stream>>>(w->s);
  k_cg_init<X><<<vgrid, kBlock, 0, c->stream>>>(n, r, invd, p, w->part, w->s, tol2);
  FEM_HIP(c, hipGetLastError());
  const int check_every = 16;
  int it = 0;
  Scal h{};
  float ms_total = 0.f;
  while (it < maxit) {
    const int batch = std::min(check_every, maxit - it);
    FEM_HIP(c, hipEventRecord(c->ev0, c->stream));
    for (int k = 0; k < batch; ++k) {
      launch_spmv_pq<V, X>(c, G, val, p, q, w);
      k_cg_update<X><<<vgrid, kBlock, 0, c->stream>>>(n, x, r, p, q, invd, q, w->part, w->s, tol2);
      k_p_update<X><<<vgrid, kBlock, 0, c->stream>>>(n, p, q, w->s);
    }
    FEM_HIP(c, hipGetLastError());
    FEM_HIP(c, hipEventRecord(c->ev1, c->stream));
    ms_total += event_ms(c);
    int rc = read_scal(c, w, &h);
    if (rc) return rc;
    it += batch;
    if (h.done) break;
  }
  // iterations actually performed: the device stops on the first converged update
  *iters = h.iters;
  *spmv_ms += ms_total;
  *spmv_count += h.iters;
  return h.done ? FEM_OK : FEM_ERR_NOCONV;
}

}  // namespace fem

using namespace fem;

extern "C" int32_t fem_solve(fem_ctx* c, const fem_solve_cfg* cfg, fem_solve_info* info) {
  if (!c || !cfg) return FEM_ERR_ARG;
  if (c->nnz == 0 || !c->d_f) return set_error(c, FEM_ERR_STATE, "assemble a system first");
  FEM_HIP(c, hipSetDevice(c->device));
  const bool mixed = cfg->mode == FEM_SOLVE_IR_MIXED;
  int rc = ensure_work(c, mixed);
  if (rc) return rc;
  SolverWork* w = c->solver;
  const int64_t n = c->n;
  const int grid = kParts;
  const int maxit = cfg->maxit > 0 ? cfg->maxit : 100000;
  fem_solve_info out{};
  c->timings.spmv_ms = 0.0;
  c->timings.spmv_launches = 0;
  double spmv_ms = 0.0;
  int spmv_count = 0;

  k_reset_scal<<<1, 1, 0, c->stream>>>(w->s);
  k_sumsq<<<grid, kBlock, 0, c->stream>>>(n, c->d_f, w->part, w->s);
  Scal h{};
  if ((rc = read_scal(c, w, &h))) return rc;
  const double bnorm2 = h.bnorm2;
  if (bnorm2 == 0.0) {
    k_zero<<<grid, kBlock, 0, c->stream>>>(n, c->d_u);
    FEM_HIP(c, hipStreamSynchronize(c->stream));
    out.outer_iterations = 1;
    if (info) *info = out;
    return FEM_OK;
  }
  const double tol2 = cfg->rtol * cfg->rtol * bnorm2;

  if (!mixed) {
    k_invdiag<double><<<grid, kBlock, 0, c->stream>>>(n, c->d_diag, c->d_nzval, cfg->precond, w->invd);
    k_zero<<<grid, kBlock, 0, c->stream>>>(n, c->d_u);
    k_copy<<<grid, kBlock, 0, c->stream>>>(n, c->d_f, w->r);
    int iters = 0;
    rc = pcg<double, double>(c, w, c->d_nzval, c->d_u, w->r, w->p, w->q, w->invd, tol2, maxit, &iters, &spmv_ms,
                             &spmv_count);
    out.iterations = iters;
    out.outer_iterations = 1;
    if (rc && rc != FEM_ERR_NOCONV) return rc;
  } else {
    if (!c->d_nzval32) {
      if ((rc = dev_alloc(c, &c->d_nzval32, c->nnz))) return rc;
      k_to_f32<<<4096, kBlock, 0, c->stream>>>(c->nnz, c->d_nzval, c->d_nzval32);
    }
    k_invdiag<float><<<grid, kBlock, 0, c->stream>>>(n, c->d_diag, c->d_nzval, cfg->precond, w->invd32);
    k_zero<<<grid, kBlock, 0, c->stream>>>(n, c->d_u);
    const double inner = cfg->inner_rtol > 0 ? cfg->inner_rtol : 1e-5;
    const int G = lanes_per_row(c);
    int total = 0, outer = 0;
    rc = FEM_ERR_NOCONV;
    while (total < maxit && outer < 100) {
      k_reset_scal<<<1, 1, 0, c->stream>>>(w->s);
      launch_residual(c, G, w);
      Scal hs{};
      int r2 = read_scal(c, w, &hs);
      if (r2) return r2;
      if (hs.rr <= tol2) { rc = FEM_OK; break; }
      ++outer;
      k_scale_to_f32<<<grid, kBlock, 0, c->stream>>>(n, w->res, w->s, w->r32, w->x32);
      int iters = 0;
      int r3 = pcg<float, float>(c, w, c->d_nzval32, w->x32, w->r32, w->p32, w->q32, w->invd32, inner * inner,
                                 maxit - total, &iters, &spmv_ms, &spmv_count);
      if (r3 && r3 != FEM_ERR_NOCONV) return r3;
      total += iters;
      k_correct<<<grid, kBlock, 0, c->stream>>>(n, w->x32, std::sqrt(hs.rr), c->d_u);
    }
    out.iterations = total;
    out.outer_iterations = outer;
  }
  // true FP64 residual of the returned u
  k_reset_scal<<<1, 1, 0, c->stream>>>(w->s);
  launch_residual(c, lanes_per_row(c), w);
  FEM_HIP(c, hipGetLastError());
  if ((rc = read_scal(c, w, &h))) return rc;
  out.relres = std::sqrt(h.rr / bnorm2);
  c->timings.spmv_ms = spmv_ms;
  c->timings.spmv_launches = spmv_count;
  if (info) *info = out;
  if (out.relres > cfg->rtol * 1.0000001) {
    char buf[160];
    snprintf(buf, sizeof buf, "CG did not converge: relres %.3e after %d iterations", out.relres, out.iterations);
    return set_error(c, FEM_ERR_NOCONV, buf);
  }
  return FEM_OK;
}

// Times `reps` SpMV launches y = K x (FP64 values/vectors, or FP32 with mixed != 0) on the
// context stream with HIP events; returns the average launch duration.  Used for the
// roofline figure of the solver's dominant kernel.
extern "C" int32_t fem_spmv_bench(fem_ctx* c, int32_t mixed, int32_t reps, double* avg_ms) {
  if (!c || reps <= 0 || !avg_ms) return FEM_ERR_ARG;
  if (c->nnz == 0) return set_error(c, FEM_ERR_STATE, "no matrix");
  FEM_HIP(c, hipSetDevice(c->device));
  int rc = ensure_work(c, mixed != 0);
  if (rc) return rc;
  SolverWork* w = c->solver;
  const int G = lanes_per_row(c);
  k_reset_scal<<<1, 1, 0, c->stream>>>(w->s);
  if (mixed) {
    if (!c->d_nzval32) {
      if ((rc = dev_alloc(c, &c->d_nzval32, c->nnz))) return rc;
      k_to_f32<<<4096, kBlock, 0, c->stream>>>(c->nnz, c->d_nzval, c->d_nzval32);
    }
    k_to_f32<<<kParts, kBlock, 0, c->stream>>>(c->n, c->d_f, w->p32);
  } else {
    k_copy<<<kParts, kBlock, 0, c->stream>>>(c->n, c->d_f, w->p);
  }
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  FEM_HIP(c, hipEventRecord(c->ev0, c->stream));
  for (int i = 0; i < reps; ++i) {
    if (mixed) launch_spmv_pq<float, float>(c, G, c->d_nzval32, w->p32, w->q32, w);
    else launch_spmv_pq<double, double>(c, G, c->d_nzval, w->p, w->q, w);
  }
  FEM_HIP(c, hipGetLastError());
  FEM_HIP(c, hipEventRecord(c->ev1, c->stream));
  *avg_ms = event_ms(c) / reps;
  return FEM_OK;
}
