// Batched assembly + solve of many independent scalar Q1 Poisson problems (2D quads
// and 3D hexes mixed) on one device.
//
// Layout: every problem keeps Ferrite's own numbering (its dofs, pattern and colour
// lists); on the device the problems are concatenated — nodes, dofs and the CSC
// pattern shifted by per-problem offsets — so the batch is one block-diagonal K
// and the solver kernels see a single matrix.
//
// Assembly: ONE persistent kernel drains a work queue of (problem, colour, chunk)
// items ordered colour-major.  An item may start once every chunk of the lower
// colours of its own problem has finished (a per-problem completion counter with
// release/acquire fences); problems never wait for each other.  Waits only ever
// target items claimed earlier, so the smallest unfinished item always runs and the
// kernel cannot deadlock; the wait is also bounded (a spin limit raises an error
// flag instead of hanging).  The launch (queue reset + kernel) is captured in a
// hipGraph and replayed.  Cells per workgroup and chunk size come from an
// autotuner that times the candidates on the batch itself.
//
// Solve: batched Jacobi-PCG in FP32 (operator, vectors) with FP64 dot products per
// problem and FP64 reliable updates (true residual per problem) every batch of
// iterations; a problem stops updating once its FP64 residual meets its tolerance,
// and its completion time is taken from the batch boundary at which that happened.
// The FP32 operator is a sliced ELL copy whose slices (≤ 64 rows, one wave) never cross
// a reduction segment, so one block per segment runs its slices and sums its partial.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fem_internal.h"
#include "ferrite_mi355x_batch.h"

namespace fem {

struct BItem {
  int32_t prob;    // problem index
  int32_t slot0;   // first cell slot (global)
  int32_t nslots;  // cells in this chunk
  int32_t need;    // chunks of this problem (lower colours) that must be complete
};

struct BProbDev {
  int32_t dim;       // 2 or 3
  int32_t node0;     // first node (global)
  int32_t dof0;      // first dof (global)
  int32_t ndofs;
};

struct BQueue {
  unsigned int next;
  int error;
  int pad[2];
};

struct BProblemHost {
  int dim = 0;
  int64_t nnodes = 0, ncells = 0, ndofs = 0, nnz = 0;
  std::vector<double> xyz;            // AoS
  std::vector<int32_t> cell_nodes;    // 0-based, local
  std::vector<int32_t> cell_dofs;     // 0-based, local
  std::vector<int32_t> colptr, rowval;
  std::vector<int64_t> color_ptr;
  std::vector<int32_t> color_cells;   // 0-based
  std::vector<double> k;
  std::vector<int32_t> prescribed;    // 0-based, sorted
  double source = 1.0;
};

}  // namespace fem

struct fem_batch {
  fem_ctx* c = nullptr;   // owns stream, error string and the concatenated matrix
  std::vector<fem::BProblemHost> probs;
  bool finalized = false;
  // device
  int32_t* d_cnodes8 = nullptr;     // [nslots][8] (quads use the first 4), global node ids
  int32_t* d_cdofs8 = nullptr;      // [nslots][8] global dofs
  int32_t* d_cnz = nullptr;         // [nslots][64] global nz positions (quads use 16)
  double* d_kslot = nullptr;        // [nslots]
  int32_t* d_slotprob = nullptr;    // [nslots]
  fem::BItem* d_items = nullptr;
  int nitems = 0;
  int32_t* d_done = nullptr;        // [nprob] completed chunks
  fem::BQueue* d_q = nullptr;
  fem::BProbDev* d_probs = nullptr;
  int32_t* d_rowprob = nullptr;     // [ndofs] problem of each global dof
  int64_t* d_pdof = nullptr;        // [nprob+1] dof offsets
  int32_t* d_sprob = nullptr;       // [nseg] problem of each reduction segment
  int64_t* d_srow = nullptr;        // [nseg+1] first row of each segment
  int32_t* d_pseg = nullptr;        // [nprob+1] first segment of each problem
  double* d_spart = nullptr;        // [6 nseg] segment partials: γ and ||r||² (two buffers of 2
                                    // per segment), then δ / the hand-over partials (2 per segment)
  int nseg = 0;
  int64_t nslices = 0, blen = 0;
  int32_t* d_aseg = nullptr;        // [nseg] segments of problems in the recurrence
  int32_t* d_hseg = nullptr;        // [nseg] segments of problems being handed over
  int32_t* d_nlist = nullptr;       // [2] lengths of the two lists
  int32_t* d_handing = nullptr;     // [nprob] 1 for the problems of the current hand-over
  std::vector<int32_t> h_pseg;      // [nprob+1] first segment of each problem (host copy)
  double* d_tab = nullptr;          // reference tables: quad block then hex block (float copy below)
  float* d_tabf = nullptr;
  int64_t nslots = 0;
  int cells_per_wg = 0, chunk = 0;  // tuned launch shape
  hipGraphExec_t graph = nullptr;
  int graph_cpw = 0, graph_chunk = 0, graph_prec = -1;
  // persistent solver work (allocated at finalize)
  double* d_m = nullptr;            // [nprob] meandiag per problem
  // FP32 operator as slices of ≤ 64 rows inside each segment (a segment's rows are its first
  // slices' rows; entry k of the slice's row l at boff[slice] + 64 k + l, padding value 0)
  int32_t* d_bsl = nullptr;         // [nseg+1] first slice of each segment
  int32_t* d_slrow = nullptr;       // [nslices] first row of each slice
  int32_t* d_boff = nullptr;        // [nslices+1] first entry of each slice
  int32_t* d_bbase = nullptr;       // [nslices] smallest column of the slice
  uint16_t* d_broff = nullptr;      // [blen] column − base (when every slice spans < 65536 columns)
  int32_t* d_bcol = nullptr;        // [blen] column (otherwise)
  int32_t* d_bsrc = nullptr;        // [blen] position in nzval, −1 on padding
  float* d_bval = nullptr;          // [blen] s_i K_ij s_j
  // FP32 recurrence of the scaled system Â = S K S: x̂, r̂, w = Â r̂, p, s = Â p
  float *d_x32 = nullptr, *d_r32 = nullptr, *d_w32 = nullptr, *d_p32 = nullptr, *d_s32 = nullptr;
  double* d_sc = nullptr;           // [ndofs] S = diag(K)^-1/2
  float* d_dsq = nullptr;           // [ndofs] S⁻² = |diag(K)|
  double* d_res = nullptr;          // [ndofs] FP64 true residual
  void* d_scal = nullptr;           // [2 nprob] per-problem CG scalars (two records each)
  // captured CG batches, one per launch width (a power of two ≥ the recurring segments):
  // blocks past the list's length return at once, so a narrow graph serves short lists
  std::vector<std::pair<int, hipGraphExec_t>> cg_graphs;
  hipEvent_t e_start = nullptr, e_asm = nullptr, e_apply = nullptr, e_mark = nullptr;
  std::vector<double> last_times_ms;
  std::vector<int> last_iters;
};

namespace fem {

// reference Q1 tables: quad 2x2 Gauss then hex 2x2x2 Gauss (N[q][a], dN[q][a][d], w[q])
constexpr int kTQ_N = 0, kTQ_dN = 16, kTQ_w = 16 + 32;           // quad: 4 qp x 4 nodes
constexpr int kTH_N = 64, kTH_dN = 64 + 64, kTH_w = 64 + 64 + 192;  // hex: 8 qp x 8 nodes
constexpr int kTabSize = 64 + 64 + 192 + 8;

template <typename T>
__device__ __forceinline__ void poisson_quad(const double* __restrict__ X0, const double* __restrict__ X1,
                                             const int32_t* __restrict__ nodes, T k, T src, const T* __restrict__ tab,
                                             T (&ke)[10], T (&fe)[4], bool& bad) {
  T x[4][2];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    x[a][0] = (T)X0[nodes[a]];
    x[a][1] = (T)X1[nodes[a]];
  }
#pragma unroll
  for (int i = 0; i < 10; ++i) ke[i] = T(0);
#pragma unroll
  for (int i = 0; i < 4; ++i) fe[i] = T(0);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    T J00 = 0, J01 = 0, J10 = 0, J11 = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const T d0 = tab[kTQ_dN + (q * 4 + a) * 2], d1 = tab[kTQ_dN + (q * 4 + a) * 2 + 1];
      J00 += x[a][0] * d0; J01 += x[a][0] * d1; J10 += x[a][1] * d0; J11 += x[a][1] * d1;
    }
    const T det = J00 * J11 - J01 * J10;
    bad |= !(det > T(0));
    const T id = T(1) / det;
    const T i00 = J11 * id, i01 = -J01 * id, i10 = -J10 * id, i11 = J00 * id;
    const T dV = det * tab[kTQ_w + q];
    T g[4][2];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const T d0 = tab[kTQ_dN + (q * 4 + b) * 2], d1 = tab[kTQ_dN + (q * 4 + b) * 2 + 1];
      g[b][0] = d0 * i00 + d1 * i10;
      g[b][1] = d0 * i01 + d1 * i11;
    }
    const T kdV = k * dV, sdV = src * dV;
    int t = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      fe[a] += tab[kTQ_N + q * 4 + a] * sdV;
#pragma unroll
      for (int b = a; b < 4; ++b) ke[t++] += (g[a][0] * g[b][0] + g[a][1] * g[b][1]) * kdV;
    }
  }
}

template <typename T>
__device__ __forceinline__ void poisson_hex(const double* __restrict__ X0, const double* __restrict__ X1,
                                            const double* __restrict__ X2, const int32_t* __restrict__ nodes, T k, T src,
                                            const T* __restrict__ tab, T (&ke)[36], T (&fe)[8], bool& bad) {
  T x[8][3];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    x[a][0] = (T)X0[nodes[a]];
    x[a][1] = (T)X1[nodes[a]];
    x[a][2] = (T)X2[nodes[a]];
  }
#pragma unroll
  for (int i = 0; i < 36; ++i) ke[i] = T(0);
#pragma unroll
  for (int i = 0; i < 8; ++i) fe[i] = T(0);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    T J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) J[i][j] += x[a][i] * tab[kTH_dN + (q * 8 + a) * 3 + j];
    const T a00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const T a01 = J[0][2] * J[2][1] - J[0][1] * J[2][2];
    const T a02 = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    const T det = J[0][0] * a00 + J[1][0] * a01 + J[2][0] * a02;
    bad |= !(det > T(0));
    const T id = T(1) / det;
    T Ji[3][3];
    Ji[0][0] = a00 * id; Ji[0][1] = a01 * id; Ji[0][2] = a02 * id;
    Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    const T dV = det * tab[kTH_w + q];
    T g[8][3];
#pragma unroll
    for (int b = 0; b < 8; ++b)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        g[b][j] = tab[kTH_dN + (q * 8 + b) * 3] * Ji[0][j] + tab[kTH_dN + (q * 8 + b) * 3 + 1] * Ji[1][j] +
                  tab[kTH_dN + (q * 8 + b) * 3 + 2] * Ji[2][j];
    const T kdV = k * dV, sdV = src * dV;
    int t = 0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      fe[a] += tab[kTH_N + q * 8 + a] * sdV;
#pragma unroll
      for (int b = a; b < 8; ++b) ke[t++] += (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]) * kdV;
    }
  }
}

__device__ __forceinline__ int tri8(int lo, int hi) { return lo * (17 - lo) / 2 + (hi - lo); }
__device__ __forceinline__ int tri4(int lo, int hi) { return lo * (9 - lo) / 2 + (hi - lo); }

__global__ void k_bq_reset(BQueue* q, int32_t* done, int nprob) {
  for (int i = threadIdx.x; i < nprob; i += blockDim.x) done[i] = 0;
  if (threadIdx.x == 0) {
    q->next = 0u;
    q->error = 0;
  }
}

constexpr long long kSpinLimit = 1ll << 20;   // >= 2^27 cycles of s_sleep: far above any real wait

template <typename T, int CPW>
__global__ __launch_bounds__(CPW) void k_batch_assemble(
    const BItem* __restrict__ items, int nitems, BQueue* q, int32_t* done, const BProbDev* __restrict__ probs,
    const double* __restrict__ X0, const double* __restrict__ X1, const double* __restrict__ X2,
    const int32_t* __restrict__ cnodes, const int32_t* __restrict__ cdofs, const int32_t* __restrict__ cnz,
    const double* __restrict__ kslot, const T* __restrict__ tab, T src, double* nzval, double* f) {
  __shared__ T s_ke[CPW * 36];
  __shared__ T s_fe[CPW * 8];
  __shared__ int s_item;
  const int tid = threadIdx.x;
  int prev = -1;   // problem of the item finished in the previous trip (its completion not yet published)
  // Every lane-0-only block sits between two barriers and the loop is steered by
  // readfirstlane'd (scalar, uniform) values only.  Without that, a one-wave
  // workgroup (where the barrier is only a wave barrier) may be compiled into a loop
  // in which lanes 1..63 go round again and read the previous item index before
  // lane 0 has claimed the next one.
  for (;;) {
    __syncthreads();   // previous item fully scattered by every wave
    if (tid == 0) {
      if (prev >= 0) __hip_atomic_fetch_add(&done[prev], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      s_item = (int)atomicAdd(&q->next, 1u);
    }
    __syncthreads();
    const int i = __builtin_amdgcn_readfirstlane(s_item);
    if (i >= nitems) break;
    const BItem it = items[i];
    if (tid == 0 && it.need > 0) {
      long long spins = 0;
      while (__hip_atomic_load(&done[it.prob], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < it.need) {
        __builtin_amdgcn_s_sleep(2);
        if (++spins > kSpinLimit) {
          atomicExch(&q->error, 1);
          break;
        }
      }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const int dim = probs[it.prob].dim;
    const int nb = dim == 3 ? 8 : 4;
    for (int c0 = 0; c0 < it.nslots; c0 += CPW) {
      const int ncell = min(CPW, it.nslots - c0);
      const int64_t sb = (int64_t)it.slot0 + c0;
      if (tid < ncell) {
        const int64_t s = sb + tid;
        bool bad = false;
        if (dim == 3) {
          T ke[36], fe[8];
          poisson_hex<T>(X0, X1, X2, cnodes + s * 8, (T)kslot[s], src, tab, ke, fe, bad);
#pragma unroll
          for (int e = 0; e < 36; ++e) s_ke[tid * 36 + e] = ke[e];
#pragma unroll
          for (int e = 0; e < 8; ++e) s_fe[tid * 8 + e] = fe[e];
        } else {
          T ke[10], fe[4];
          poisson_quad<T>(X0, X1, cnodes + s * 8, (T)kslot[s], src, tab, ke, fe, bad);
#pragma unroll
          for (int e = 0; e < 10; ++e) s_ke[tid * 36 + e] = ke[e];
#pragma unroll
          for (int e = 0; e < 4; ++e) s_fe[tid * 8 + e] = fe[e];
        }
        if (bad) atomicExch(&q->error, 2);
      }
      __syncthreads();
      const int E = nb * nb;
      const int tot = ncell * E;
      for (int t = tid; t < tot; t += CPW) {
        const int lc = t / E, e = t - lc * E;
        const int b = e / nb, a = e - b * nb;
        const int lo = a < b ? a : b, hi = a < b ? b : a;
        const T v = s_ke[lc * 36 + (nb == 8 ? tri8(lo, hi) : tri4(lo, hi))];
        const int32_t p = cnz[(sb + lc) * 64 + e];
        if (p >= 0) {
          if (v != T(0)) nzval[p] += (double)v;
        } else if (v != T(0)) {
          atomicExch(&q->error, 3);
        }
      }
      for (int t = tid; t < ncell * nb; t += CPW) {
        const int lc = t / nb, a = t - lc * nb;
        f[cdofs[(sb + lc) * 8 + a]] += (double)s_fe[lc * 8 + a];
      }
      __syncthreads();
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    prev = it.prob;
  }
}

// ---------------------------------------------------------------------------
// batched Dirichlet apply! (per-problem meandiag)
// ---------------------------------------------------------------------------
__global__ void k_b_meandiag(const int64_t* __restrict__ pdof, const int32_t* __restrict__ diag,
                             const double* __restrict__ nzval, double* __restrict__ m) {
  __shared__ double sh[4];
  const int p = blockIdx.x;
  const int64_t j0 = pdof[p], j1 = pdof[p + 1];
  double acc = 0.0;
  for (int64_t j = j0 + threadIdx.x; j < j1; j += 256) acc += fabs(nzval[diag[j]]);
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) m[p] = (sh[0] + sh[1] + sh[2] + sh[3]) / (double)(j1 - j0);
}

__global__ void k_b_zero_rc(int64_t n, const int32_t* __restrict__ colptr, const int32_t* __restrict__ rowval,
                            const uint8_t* __restrict__ isc, double* __restrict__ nzval) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const bool cj = isc[j];
    for (int32_t p = colptr[j]; p < colptr[j + 1]; ++p)
      if (cj || isc[rowval[p]]) nzval[p] = 0.0;
  }
}

__global__ void k_b_setdiag(int64_t np, const int32_t* __restrict__ pres, const int32_t* __restrict__ diag,
                            const int32_t* __restrict__ rowprob, const double* __restrict__ m,
                            double* __restrict__ nzval, double* __restrict__ f) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t d = pres[i];
    nzval[diag[d]] = m[rowprob[d]];
    f[d] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// batched mixed PCG: FP32 Krylov recurrence on the block-diagonal operator with
// per-problem scalars; dot products and the reliable-update residual in FP64
// ---------------------------------------------------------------------------
// Per-problem scalars of the batched Chronopoulos–Gear CG on the scaled system (two
// records per problem, read one / write the other, like the single-problem solver).
struct BScal {
  double rz;      // γ = r̂·r̂ of the current residual
  double pq;      // p·Â p at a reseed
  double rr;      // FP64 true residual ||f − K u||² of the last hand-over
  double rrec;    // recursive residual ||S⁻¹ r̂||² of the last step
  double alpha, beta;
  double tol2;    // rtol² ||f||²
  double d2;      // δ²: hand over once the recursive residual fell by this factor
  double bnorm2;  // ||f||²
  double rzp;     // γ of the previous step
  double gam;     // (r̂·p)/(r̂·r̂) at a reseed
  int done;       // converged (FP64 true residual <= rtol ||f||)
  int hold;       // recurrence stopped, waiting for a hand-over on the host
  int iters;
  int restart;    // breakdown: the next reseed restarts p = r̂
  int fresh;      // p, s, α just set by a reseed: the next step only moves x̂, r̂
  int pad;
};

// Per-problem reductions in two levels: every segment (≤ kSeg consecutive rows of one
// problem) is one block and writes its partials; the problem's blocks (or a per-problem
// kernel) sum them in segment order.  Fixed order throughout (deterministic).
// 1024-row segments: at 256 the latency-bound tail (a few large 2D problems) ran several
// rounds of small blocks per launch, each block paying the scalar record, its problem's
// partials and the block sums for 256 rows.
constexpr int kSeg = 1024;
constexpr int kSegThreads = 256;     // vector kernels: kSegR rows per thread, kSegThreads apart
constexpr int kSegR = kSeg / kSegThreads;
constexpr int kSpmvThreads = 1024;   // FP64 residual kernel: 4 lanes per row, kSpmvThreads / 4 rows a pass
static_assert(kSeg % kSegThreads == 0 && kSeg % (kSpmvThreads / 4) == 0, "segment tiling");

// Sum over the kSpmvThreads threads of a block, valid in every thread; `sh` holds 16.
__device__ __forceinline__ double spmv_block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < kSpmvThreads / 64; ++i) t += sh[i];
  return t;
}

__device__ __forceinline__ double seg_block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  return (sh[0] + sh[1]) + (sh[2] + sh[3]);   // valid in every thread
}

// w = Â r̂ on one segment per block from the sliced operator (one wave per slice of the
// segment, lane = row, 8 entries per lane loaded before their r̂ gathers; the operator is
// streamed with non-temporal loads so that r̂ stays in L2), and the segment's partial
// δ = r̂·w.  list/nlist: the segments to run; blocks past the list's length return.
// Problems held or right after a reseed (fresh) are skipped.
template <bool COMP>
__global__ __launch_bounds__(kSpmvThreads) void k_bc_sell(const int32_t* __restrict__ list,
                                                         const int32_t* __restrict__ nlist, int which,
                                                         const int32_t* __restrict__ sprob,
                                                         const int64_t* __restrict__ srow,
                                                         const int32_t* __restrict__ bsl,
                                                         const int32_t* __restrict__ slrow,
                                                         const int32_t* __restrict__ boff,
                                                         const int32_t* __restrict__ bbase,
                                                         const uint16_t* __restrict__ broff,
                                                         const int32_t* __restrict__ bcol,
                                                         const float* __restrict__ val, const float* __restrict__ r,
                                                         float* __restrict__ w, const BScal* __restrict__ sin,
                                                         double* __restrict__ part_d) {
  __shared__ double sh[kSpmvThreads / 64];
  if ((int)blockIdx.x >= nlist[which]) return;
  const int seg = list[blockIdx.x];
  const BScal& S = sin[sprob[seg]];
  if (S.done || S.hold || S.fresh) return;
  const int lane = threadIdx.x & 63;
  double d = 0.0;
  // one wave per slice (a segment has at most kSpmvThreads / 64 of them)
  for (int sl = bsl[seg] + (int)(threadIdx.x >> 6); sl < bsl[seg + 1]; sl += kSpmvThreads / 64) {
    const int64_t row = (int64_t)slrow[sl] + lane;
    const int32_t o0 = boff[sl];
    const int W = (boff[sl + 1] - o0) >> 6;
    const int32_t b = COMP ? bbase[sl] : 0;
    const int32_t e0 = o0 + lane;
    constexpr int U = 8;
    float acc = 0.f;
    int k = 0;
    for (; k + U <= W; k += U) {
      float v[U];
      int32_t ci[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t e = e0 + 64 * (k + u);
        v[u] = __builtin_nontemporal_load(val + e);
        ci[u] = COMP ? b + (int32_t)__builtin_nontemporal_load(broff + e) : __builtin_nontemporal_load(bcol + e);
      }
      float xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) xv[u] = r[ci[u]];
#pragma unroll
      for (int u = 0; u < U; ++u) acc += v[u] * xv[u];
    }
    for (; k < W; ++k) {
      const int32_t e = e0 + 64 * k;
      const int32_t ci = COMP ? b + (int32_t)__builtin_nontemporal_load(broff + e) : __builtin_nontemporal_load(bcol + e);
      acc += __builtin_nontemporal_load(val + e) * r[ci];
    }
    if (row < srow[seg + 1]) {
      w[row] = acc;
      d += (double)r[row] * (double)acc;
    }
  }
  const double t = spmv_block_sum(d, sh);
  if (threadIdx.x == 0) part_d[seg] = t;
}

// Scaled values of the sliced operator: val = s_row K[row, col] s_col (the product order of
// the single-problem copy), 0 on padding; one wave per slice.
template <bool COMP>
__global__ void k_bc_sell_vals(int64_t nsl, int64_t n, const int32_t* __restrict__ slrow,
                               const int32_t* __restrict__ boff, const int32_t* __restrict__ bbase,
                               const uint16_t* __restrict__ broff, const int32_t* __restrict__ bcol,
                               const int32_t* __restrict__ bsrc, const double* __restrict__ nzval,
                               const double* __restrict__ sc, float* __restrict__ val) {
  const int lane = threadIdx.x & 63;
  for (int64_t sl = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; sl < nsl;
       sl += ((int64_t)gridDim.x * blockDim.x) >> 6) {
    const int64_t row = (int64_t)slrow[sl] + lane;
    const double sr = row < n ? sc[row] : 0.0;
    const int32_t o0 = boff[sl];
    const int W = (boff[sl + 1] - o0) >> 6;
    const int32_t b = COMP ? bbase[sl] : 0;
    for (int k = 0; k < W; ++k) {
      const int32_t e = o0 + 64 * k + lane;
      const int32_t q = bsrc[e];
      const int32_t col = COMP ? b + (int32_t)broff[e] : bcol[e];
      val[e] = q >= 0 ? (float)(nzval[q] * sr * sc[col]) : 0.f;
    }
  }
}

// The fused vector step on one segment per block.  Every block of a problem sums that
// problem's partials (γ of the previous step, δ of the SpMV) in segment order, so all its
// blocks hold the same scalars; the problem's first segment writes its next record.
__global__ __launch_bounds__(kSegThreads) void k_bc_step(const int32_t* __restrict__ list,
                                                         const int32_t* __restrict__ nlist,
                                                         const int32_t* __restrict__ sprob,
                                                         const int64_t* __restrict__ srow,
                                                         const int32_t* __restrict__ pseg, float* __restrict__ x,
                                                         float* __restrict__ r, const float* __restrict__ w,
                                                         float* __restrict__ p, float* __restrict__ sv,
                                                         const float* __restrict__ dsq,
                                                         const double* __restrict__ part_d,
                                                         const double* __restrict__ part_in,
                                                         double* __restrict__ part_out, const BScal* __restrict__ sin,
                                                         BScal* __restrict__ sout) {
  __shared__ double sh[4];
  if ((int)blockIdx.x >= nlist[0]) return;
  const int seg = list[blockIdx.x];
  const int pr = sprob[seg];
  const bool writer = seg == pseg[pr] && threadIdx.x == 0;
  const BScal S = sin[pr];
  if (S.done || S.hold) {
    if (writer) sout[pr] = S;
    return;
  }
  const bool fresh = S.fresh != 0;
  double gam = S.rz, rrec = S.rrec, alpha = S.alpha, beta = 0.0;
  if (!fresh) {
    double g = 0.0, q = 0.0, dl = 0.0;
    for (int k = pseg[pr] + threadIdx.x; k < pseg[pr + 1]; k += kSegThreads) {
      g += part_in[2 * k];
      q += part_in[2 * k + 1];
      dl += part_d[k];
    }
    gam = seg_block_sum(g, sh);
    rrec = seg_block_sum(q, sh);
    const double del = seg_block_sum(dl, sh);
    beta = gam / S.rzp;
    const double den = del - beta * gam / S.alpha;
    // converged by the recursive residual, or fallen by δ since the last true residual
    const bool stop = rrec <= S.tol2 || rrec < S.d2 * S.rr;
    const bool brk = !(S.rzp > 0.0) || !(S.alpha != 0.0) || !(den > 0.0);
    if (stop || brk) {
      if (writer) {
        BScal o = S;
        o.hold = 1;
        o.rz = gam;
        o.rrec = rrec;
        if (!stop) o.restart = 1;
        sout[pr] = o;
      }
      return;
    }
    alpha = gam / den;
  }
  const float al = (float)alpha, be = (float)beta;
  const int64_t j0 = srow[seg], j1 = srow[seg + 1];
  double ga = 0.0, gr = 0.0;
  // the segment's kSegR rows of this thread: every load issued before any use
  float pv[kSegR], sw[kSegR], rv[kSegR], dv[kSegR], wv[kSegR], xv[kSegR];
#pragma unroll
  for (int k = 0; k < kSegR; ++k) {
    const int64_t j = j0 + threadIdx.x + k * kSegThreads;
    const bool in = j < j1;
    pv[k] = in ? p[j] : 0.f;
    sw[k] = in ? sv[j] : 0.f;
    rv[k] = in ? r[j] : 0.f;
    dv[k] = in ? dsq[j] : 0.f;
    wv[k] = in && !fresh ? w[j] : 0.f;
    xv[k] = in ? x[j] : 0.f;
  }
#pragma unroll
  for (int k = 0; k < kSegR; ++k) {
    const int64_t j = j0 + threadIdx.x + k * kSegThreads;
    if (j >= j1) continue;
    float pj = pv[k], sj = sw[k];
    const float rj0 = rv[k];
    if (!fresh) {
      pj = rj0 + be * pj;
      sj = wv[k] + be * sj;
      p[j] = pj;
      sv[j] = sj;
    }
    x[j] = xv[k] + al * pj;
    const float rj = rj0 - al * sj;
    r[j] = rj;
    const double rd = (double)rj * (double)rj;
    ga += rd;
    gr += rd * (double)dv[k];
  }
  const double t = seg_block_sum(ga, sh);
  const double t2 = seg_block_sum(gr, sh);
  if (threadIdx.x == 0) {
    part_out[2 * seg] = t;
    part_out[2 * seg + 1] = t2;
  }
  if (writer) {
    BScal o = S;
    o.fresh = 0;
    o.rrec = rrec;
    o.rzp = gam;
    o.alpha = alpha;
    o.beta = beta;
    o.rz = gam;
    o.iters = S.iters + 1;
    sout[pr] = o;
  }
}

// Hand-over, step 1 (segments of the handed-over problems): u += S x̂, x̂ = 0, then
// res = f − K u with the FP64 CSC arrays; partial ||res||² per segment.
__global__ __launch_bounds__(kSpmvThreads) void k_bc_resid(const int32_t* __restrict__ list,
                                                          const int32_t* __restrict__ nlist,
                                                          const int64_t* __restrict__ srow,
                                                          const int32_t* __restrict__ colptr,
                                                          const int32_t* __restrict__ rowval,
                                                          const double* __restrict__ val, const double* __restrict__ u,
                                                          const double* __restrict__ f, double* __restrict__ res,
                                                          double* __restrict__ part) {
  __shared__ double sh[kSpmvThreads / 64];
  if ((int)blockIdx.x >= nlist[1]) return;
  const int seg = list[blockIdx.x];
  const int sub = threadIdx.x & 3;
  double d = 0.0;
  for (int64_t j = srow[seg] + (threadIdx.x >> 2); j - (threadIdx.x >> 2) < srow[seg + 1]; j += kSpmvThreads / 4) {
    const bool in = j < srow[seg + 1];
    double acc = 0.0;
    if (in)
      for (int32_t q = colptr[j] + sub; q < colptr[j + 1]; q += 4) acc += val[q] * u[rowval[q]];
    acc += __shfl_down(acc, 2, 4);
    acc += __shfl_down(acc, 1, 4);
    if (in && sub == 0) {
      const double rj = f[j] - acc;
      res[j] = rj;
      d += rj * rj;
    }
  }
  const double t = spmv_block_sum(d, sh);
  if (threadIdx.x == 0) part[seg] = t;
}

__global__ __launch_bounds__(kSegThreads) void k_bc_fold(const int32_t* __restrict__ list,
                                                         const int32_t* __restrict__ nlist,
                                                         const int64_t* __restrict__ srow,
                                                         const double* __restrict__ sc, float* __restrict__ x,
                                                         double* __restrict__ u) {
  if ((int)blockIdx.x >= nlist[1]) return;
  const int seg = list[blockIdx.x];
  for (int64_t j = srow[seg] + threadIdx.x; j < srow[seg + 1]; j += kSegThreads) {
    u[j] += sc[j] * (double)x[j];
    x[j] = 0.f;
  }
}

// Hand-over, step 2 (one thread per problem): the true residual; converged problems are
// done, the others are set up for a reseed (hold stays 1 until then).
// Fixed-order sum of part[stride·k] over a problem's segments by one wave (lane l takes
// k ≡ l mod 64, then a fixed butterfly), valid in every lane.
__device__ __forceinline__ double wave_seg_sum(const double* __restrict__ part, int k0, int k1, int stride) {
  double t = 0.0;
  for (int k = k0 + (int)(threadIdx.x & 63); k < k1; k += 64) t += part[(int64_t)stride * k];
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  return t;
}

__global__ __launch_bounds__(64) void k_bc_true(int nprob, const int32_t* __restrict__ pseg,
                                                const double* __restrict__ part, const int32_t* __restrict__ handing,
                                                BScal* __restrict__ s) {
  const int pr = blockIdx.x;
  if (pr >= nprob || !handing[pr]) return;
  const double t = wave_seg_sum(part, pseg[pr], pseg[pr + 1], 1);
  if (threadIdx.x == 0) {
    s[pr].rr = t;
    if (t <= s[pr].tol2) s[pr].done = 1;
  }
}

// Hand-over, step 3 (segments): r̂ = S res; partials r̂·r̂ and r̂·p.
__global__ __launch_bounds__(kSegThreads) void k_bc_reseed(const int32_t* __restrict__ list,
                                                           const int32_t* __restrict__ nlist,
                                                           const int32_t* __restrict__ sprob,
                                                           const int64_t* __restrict__ srow,
                                                           const double* __restrict__ sc,
                                                           const double* __restrict__ res, float* __restrict__ r,
                                                           const float* __restrict__ p, const BScal* __restrict__ s,
                                                           double* __restrict__ part) {
  __shared__ double sh[4];
  if ((int)blockIdx.x >= nlist[1]) return;
  const int seg = list[blockIdx.x];
  if (s[sprob[seg]].done) return;
  double rr = 0.0, rp = 0.0;
  for (int64_t j = srow[seg] + threadIdx.x; j < srow[seg + 1]; j += kSegThreads) {
    const float rj = (float)(sc[j] * res[j]);
    r[j] = rj;
    rr += (double)rj * (double)rj;
    rp += (double)rj * (double)p[j];
  }
  const double a = seg_block_sum(rr, sh);
  const double b = seg_block_sum(rp, sh);
  if (threadIdx.x == 0) {
    part[2 * seg] = a;
    part[2 * seg + 1] = b;
  }
}

// Hand-over, step 4 (one wave per problem): γ, γ', β (0: restart p = r̂); the recurrence
// runs again (hold = 0).
__global__ __launch_bounds__(64) void k_bc_reseed_fin(int nprob, const int32_t* __restrict__ pseg,
                                                      const double* __restrict__ part,
                                                      const int32_t* __restrict__ handing, BScal* __restrict__ s) {
  const int pr = blockIdx.x;
  if (pr >= nprob || !handing[pr] || s[pr].done) return;
  const double ru = wave_seg_sum(part, pseg[pr], pseg[pr + 1], 2);
  const double rp = wave_seg_sum(part + 1, pseg[pr], pseg[pr + 1], 2);
  if (threadIdx.x != 0) return;
  BScal& sp = s[pr];
  const bool fresh = sp.restart || !(sp.rzp > 0.0) || !(ru > 0.0);
  sp.gam = ru > 0.0 ? rp / ru : 0.0;
  sp.beta = fresh ? 0.0 : ru / sp.rzp;
  sp.rz = ru;
  sp.rrec = sp.rr;
  sp.hold = 0;
  sp.restart = 0;
  sp.fresh = 0;
}

// Hand-over, step 5 (segments, after w = Â r̂): p = r̂ + β(p − γ' r̂), s = w + β(s − γ' w)
// (p = r̂, s = w on a restart); partial p·s.
__global__ __launch_bounds__(kSegThreads) void k_bc_pfix(const int32_t* __restrict__ list,
                                                         const int32_t* __restrict__ nlist,
                                                         const int32_t* __restrict__ sprob,
                                                         const int64_t* __restrict__ srow,
                                                         const float* __restrict__ r, const float* __restrict__ w,
                                                         float* __restrict__ p, float* __restrict__ sv,
                                                         const BScal* __restrict__ s, double* __restrict__ part) {
  __shared__ double sh[4];
  if ((int)blockIdx.x >= nlist[1]) return;
  const int seg = list[blockIdx.x];
  const BScal& sp = s[sprob[seg]];
  if (sp.done) return;
  const float be = (float)sp.beta, ga = (float)sp.gam;
  const bool keep = sp.beta != 0.0;
  double ps = 0.0;
  for (int64_t j = srow[seg] + threadIdx.x; j < srow[seg + 1]; j += kSegThreads) {
    const float rj = r[j], wj = w[j];
    const float pj = keep ? rj + be * (p[j] - ga * rj) : rj;
    const float sj = keep ? wj + be * (sv[j] - ga * wj) : wj;
    p[j] = pj;
    sv[j] = sj;
    ps += (double)pj * (double)sj;
  }
  const double t = seg_block_sum(ps, sh);
  if (threadIdx.x == 0) part[seg] = t;
}

// Hand-over, step 6 (one thread per problem): α = γ / p·s; the next step only moves x̂, r̂.
__global__ __launch_bounds__(64) void k_bc_alpha(int nprob, const int32_t* __restrict__ pseg,
                                                 const double* __restrict__ part, const int32_t* __restrict__ handing,
                                                 BScal* __restrict__ s) {
  const int pr = blockIdx.x;
  if (pr >= nprob || !handing[pr] || s[pr].done) return;
  const double ps = wave_seg_sum(part, pseg[pr], pseg[pr + 1], 1);
  if (threadIdx.x != 0) return;
  BScal& sp = s[pr];
  sp.pq = ps;
  const bool ok = ps > 0.0;
  sp.alpha = ok ? sp.rz / ps : 0.0;
  sp.fresh = ok ? 1 : 0;
  if (!ok) {   // p vanished (r̂ already zero): hand over again at once
    sp.hold = 1;
    sp.restart = 1;
  }
}

// ||f||² per problem (segment partials, then k_bc_bnorm), and the initial records.
__global__ __launch_bounds__(kSegThreads) void k_bc_fsq(int nseg, const int64_t* __restrict__ srow,
                                                        const double* __restrict__ f, double* __restrict__ part) {
  __shared__ double sh[4];
  const int seg = blockIdx.x;
  if (seg >= nseg) return;
  double d = 0.0;
  for (int64_t j = srow[seg] + threadIdx.x; j < srow[seg + 1]; j += kSegThreads) d += f[j] * f[j];
  const double t = seg_block_sum(d, sh);
  if (threadIdx.x == 0) part[seg] = t;
}

__global__ __launch_bounds__(64) void k_bc_init(int nprob, const int32_t* __restrict__ pseg,
                                                const double* __restrict__ part, double rtol2, double d2,
                                                BScal* __restrict__ s) {
  const int pr = blockIdx.x;
  if (pr >= nprob) return;
  const double t = wave_seg_sum(part, pseg[pr], pseg[pr + 1], 1);
  if (threadIdx.x != 0) return;
  BScal z{};
  z.bnorm2 = t;
  z.tol2 = rtol2 * t;
  z.d2 = d2;
  z.rr = t;
  z.hold = 1;          // every problem starts with a hand-over from u = 0
  if (t == 0.0) z.done = 1;   // f = 0: u = 0 is exact
  s[pr] = z;
  s[nprob + pr] = z;
}

// S = diag(K)^-1/2 (1 where the diagonal is zero) and the scaled FP32 operator.
// ... and dsq = S⁻² in FP32, which turns r̂·r̂ into the unscaled ||r||² row by row.
__global__ void k_bc_scale(int64_t n, const int32_t* __restrict__ diag, const double* __restrict__ nzval,
                           double* __restrict__ sc, float* __restrict__ dsq) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const double d = nzval[diag[j]];
    sc[j] = d != 0.0 ? 1.0 / sqrt(fabs(d)) : 1.0;
    dsq[j] = d != 0.0 ? (float)fabs(d) : 1.0f;
  }
}

__global__ void k_bc_zero32(int64_t n, float* __restrict__ a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0.f;
}

__global__ void k_b_zero(int64_t n, double* __restrict__ a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0.0;
}


}  // namespace fem

using namespace fem;

namespace {

int grid_n(int64_t n, int b = 256) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + b - 1) / b, 8192)); }

int build_items(fem_batch* B, int chunk) {
  fem_ctx* c = B->c;
  std::vector<BItem> items;
  // colour-major across problems; slots of problem p are contiguous and colour-ordered
  int maxc = 0;
  for (auto& p : B->probs) maxc = std::max<int>(maxc, (int)p.color_ptr.size() - 1);
  std::vector<int64_t> slot_base(B->probs.size());
  int64_t acc = 0;
  for (size_t p = 0; p < B->probs.size(); ++p) {
    slot_base[p] = acc;
    acc += B->probs[p].ncells;
  }
  std::vector<int32_t> chunks_done(B->probs.size(), 0);
  for (int col = 0; col < maxc; ++col) {
    std::vector<int32_t> added(B->probs.size(), 0);
    for (size_t p = 0; p < B->probs.size(); ++p) {
      const auto& P = B->probs[p];
      if (col + 1 >= (int)P.color_ptr.size()) continue;
      for (int64_t s = P.color_ptr[col]; s < P.color_ptr[col + 1]; s += chunk) {
        BItem it;
        it.prob = (int32_t)p;
        it.slot0 = (int32_t)(slot_base[p] + s);
        it.nslots = (int32_t)std::min<int64_t>(chunk, P.color_ptr[col + 1] - s);
        it.need = chunks_done[p];
        items.push_back(it);
        ++added[p];
      }
    }
    for (size_t p = 0; p < B->probs.size(); ++p) chunks_done[p] += added[p];
  }
  // the assembly graph holds the old queue's address: drop it before the queue is reallocated
  if (B->graph) {
    FEM_HIP(c, hipStreamSynchronize(c->stream));
    (void)hipGraphExecDestroy(B->graph);
    B->graph = nullptr;
  }
  int rc = dev_alloc(c, &B->d_items, items.size());
  if (rc) return rc;
  FEM_HIP(c, hipMemcpyAsync(B->d_items, items.data(), sizeof(BItem) * items.size(), hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  B->nitems = (int)items.size();
  B->chunk = chunk;
  return FEM_OK;
}

template <typename T, int CPW>
void launch_batch(fem_batch* B, hipStream_t st, int grid, double src) {
  fem_ctx* c = B->c;
  const T* tab = nullptr;
  if constexpr (sizeof(T) == 8) tab = B->d_tab; else tab = B->d_tabf;
  k_bq_reset<<<1, 256, 0, st>>>(B->d_q, B->d_done, (int)B->probs.size());
  k_batch_assemble<T, CPW><<<grid, CPW, 0, st>>>(B->d_items, B->nitems, B->d_q, B->d_done, B->d_probs, c->d_x[0],
                                                  c->d_x[1], c->d_x[2], B->d_cnodes8, B->d_cdofs8, B->d_cnz, B->d_kslot,
                                                  tab, (T)src, c->d_nzval, c->d_f);
}

bool shape_ok(int cpw, int chunk, bool mixed) {
  if (!(cpw == 64 || cpw == 128 || cpw == 256) || chunk < cpw || chunk > (1 << 20)) return false;
  return mixed || cpw <= 128;   // FP64 LDS staging of 256 cells exceeds 64 KiB
}

void launch_any(fem_batch* B, hipStream_t st, int cpw, bool mixed, double src) {
  // persistent grid: a few workgroups per CU, each loops over the queue
  int ncu = 0;
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, B->c->device);
  const int grid = std::max(1, ncu) * (cpw == 64 ? 8 : (cpw == 128 ? 4 : 2));
  if (mixed) {
    if (cpw == 64) launch_batch<float, 64>(B, st, grid, src);
    else if (cpw == 128) launch_batch<float, 128>(B, st, grid, src);
    else launch_batch<float, 256>(B, st, grid, src);
  } else {
    if (cpw == 64) launch_batch<double, 64>(B, st, grid, src);
    else launch_batch<double, 128>(B, st, grid, src);
  }
}

// Launch (building the graph if the shape changed) the zero + persistent assembly graph.
int assemble_graph_launch(fem_batch* B, int cpw, int chunk, bool mixed) {
  fem_ctx* c = B->c;
  int rc;
  if (chunk != B->chunk && (rc = build_items(B, chunk))) return rc;
  if (!B->graph || B->graph_cpw != cpw || B->graph_chunk != chunk || B->graph_prec != (int)mixed) {
    if (B->graph) { (void)hipGraphExecDestroy(B->graph); B->graph = nullptr; }
    hipGraph_t g = nullptr;
    FEM_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    k_b_zero<<<grid_n(c->nnz), 256, 0, c->stream>>>(c->nnz, c->d_nzval);
    k_b_zero<<<grid_n(c->ndofs), 256, 0, c->stream>>>(c->ndofs, c->d_f);
    launch_any(B, c->stream, cpw, mixed, B->probs[0].source);
    FEM_HIP(c, hipStreamEndCapture(c->stream, &g));
    hipError_t e = hipGraphInstantiate(&B->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return check_hip(c, e, "hipGraphInstantiate (batch assembly)");
    B->graph_cpw = cpw;
    B->graph_chunk = chunk;
    B->graph_prec = (int)mixed;
  }
  FEM_HIP(c, hipGraphLaunch(B->graph, c->stream));
  return FEM_OK;
}

int assemble_errors(fem_batch* B) {
  fem_ctx* c = B->c;
  BQueue q{};
  FEM_HIP(c, hipMemcpyAsync(&q, B->d_q, sizeof(BQueue), hipMemcpyDeviceToHost, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  if (q.error == 1) return set_error(c, FEM_ERR_HIP, "batch assembly: colour dependency wait exceeded its limit");
  if (q.error == 2) return set_error(c, FEM_ERR_DETJ, "det(J) is not positive in a batch cell");
  if (q.error == 3) return set_error(c, FEM_ERR_PATTERN, "batch assembly: entry missing in a sparsity pattern");
  return FEM_OK;
}

int timed_assemble(fem_batch* B, int cpw, int chunk, bool mixed, float* ms) {
  fem_ctx* c = B->c;
  FEM_HIP(c, hipEventRecord(c->ev0, c->stream));
  int rc = assemble_graph_launch(B, cpw, chunk, mixed);
  if (rc) return rc;
  FEM_HIP(c, hipEventRecord(c->ev1, c->stream));
  *ms = event_ms(c);
  return assemble_errors(B);
}

// Per-problem reduction t = Σ a·b, t2 = Σ a·c (two levels) followed by the scalar step `which`.
// One batch of kCgBatch iterations, captured once per batch object: SpMV + fused step per
// iteration over the list of recurring segments (its length is read on the device, so the
// same graph serves every active set).  kCgBatch is even: a replay starts and ends on
// scalar record 0 and partial buffer 0.  32 (measured on the 64-problem batch: total 81.7 ms
// against 84.1 with 16; a held problem waits longer for its hand-over, p50 35.7 vs 32.3 ms,
// but the host round trips halve).
constexpr int kCgBatch = 32;

// The sliced SpMV over `blocks` entries of a segment list (which = 0: recurrence list,
// 1: hand-over list).
void launch_bc_sell(fem_batch* B, int blocks, const int32_t* list, int which, const BScal* sin, double* part) {
  fem_ctx* c = B->c;
  if (B->d_broff)
    k_bc_sell<true><<<blocks, kSpmvThreads, 0, c->stream>>>(list, B->d_nlist, which, B->d_sprob, B->d_srow, B->d_bsl,
                                                           B->d_slrow, B->d_boff, B->d_bbase, B->d_broff, nullptr,
                                                           B->d_bval, B->d_r32, B->d_w32, sin, part);
  else
    k_bc_sell<false><<<blocks, kSpmvThreads, 0, c->stream>>>(list, B->d_nlist, which, B->d_sprob, B->d_srow, B->d_bsl,
                                                            B->d_slrow, B->d_boff, B->d_bbase, nullptr, B->d_bcol,
                                                            B->d_bval, B->d_r32, B->d_w32, sin, part);
}

hipGraphExec_t* cg_graph_for(fem_batch* B, int width) {
  for (auto& g : B->cg_graphs)
    if (g.first == width) return &g.second;
  return nullptr;
}

int ensure_cg_graph(fem_batch* B, int width) {
  if (cg_graph_for(B, width)) return FEM_OK;
  fem_ctx* c = B->c;
  const int np_ = (int)B->probs.size();
  BScal* s = (BScal*)B->d_scal;
  double* pg[2] = {B->d_spart, B->d_spart + 2 * B->nseg};
  double* pd = B->d_spart + 4 * B->nseg;
  hipStream_t st = c->stream;
  hipGraph_t g = nullptr;
  FEM_HIP(c, hipStreamSynchronize(st));   // captured on first use of this width, mid-solve
  FEM_HIP(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  for (int it = 0; it < kCgBatch; ++it) {
    BScal* sin = s + (it & 1) * np_;
    BScal* sout = s + ((it + 1) & 1) * np_;
    launch_bc_sell(B, width, B->d_aseg, 0, sin, pd);
    k_bc_step<<<width, kSegThreads, 0, st>>>(B->d_aseg, B->d_nlist, B->d_sprob, B->d_srow, B->d_pseg, B->d_x32,
                                               B->d_r32, B->d_w32, B->d_p32, B->d_s32, B->d_dsq, pd, pg[it & 1],
                                               pg[(it + 1) & 1], sin, sout);
  }
  FEM_HIP(c, hipStreamEndCapture(st, &g));
  hipGraphExec_t ex = nullptr;
  hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return check_hip(c, e, "hipGraphInstantiate (batch CG)");
  B->cg_graphs.emplace_back(width, ex);
  return FEM_OK;
}

}  // namespace

extern "C" {

int32_t fem_batch_create(int32_t device, fem_batch** out) {
  if (!out) return FEM_ERR_ARG;
  fem_ctx* c = nullptr;
  int rc = fem_ctx_create(device, &c);
  if (rc) return rc;
  fem_batch* B = new fem_batch();
  B->c = c;
  if (hipEventCreate(&B->e_start) != hipSuccess || hipEventCreate(&B->e_asm) != hipSuccess ||
      hipEventCreate(&B->e_apply) != hipSuccess || hipEventCreate(&B->e_mark) != hipSuccess) {
    fem_batch_destroy(B);
    return FEM_ERR_HIP;
  }
  *out = B;
  return FEM_OK;
}

void fem_batch_destroy(fem_batch* B) {
  if (!B) return;
  fem_ctx* c = B->c;
  if (c) {
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
  }
  if (B->graph) (void)hipGraphExecDestroy(B->graph);
  for (auto& g : B->cg_graphs) (void)hipGraphExecDestroy(g.second);
  for (hipEvent_t* e : {&B->e_start, &B->e_asm, &B->e_apply, &B->e_mark})
    if (*e) (void)hipEventDestroy(*e);
  dev_free(&B->d_cnodes8); dev_free(&B->d_cdofs8); dev_free(&B->d_cnz); dev_free(&B->d_kslot);
  dev_free(&B->d_slotprob); dev_free(&B->d_items); dev_free(&B->d_done); dev_free(&B->d_q);
  dev_free(&B->d_probs); dev_free(&B->d_rowprob); dev_free(&B->d_pdof); dev_free(&B->d_tab); dev_free(&B->d_tabf);
  dev_free(&B->d_sprob); dev_free(&B->d_srow); dev_free(&B->d_pseg); dev_free(&B->d_spart);
  dev_free(&B->d_m); dev_free(&B->d_bsl); dev_free(&B->d_slrow); dev_free(&B->d_boff); dev_free(&B->d_bbase);
  dev_free(&B->d_broff); dev_free(&B->d_bcol); dev_free(&B->d_bsrc); dev_free(&B->d_bval); dev_free(&B->d_x32); dev_free(&B->d_r32); dev_free(&B->d_p32);
  dev_free(&B->d_w32); dev_free(&B->d_s32); dev_free(&B->d_sc); dev_free(&B->d_dsq); dev_free(&B->d_res); dev_free(&B->d_scal);
  dev_free(&B->d_aseg); dev_free(&B->d_hseg); dev_free(&B->d_nlist); dev_free(&B->d_handing);
  if (c) fem_ctx_destroy(c);
  delete B;
}

const char* fem_batch_last_error(const fem_batch* B) { return B && B->c ? B->c->err.c_str() : "null batch"; }

int32_t fem_batch_add_problem(fem_batch* B, int32_t dim, int64_t nnodes, const double* xyz, int64_t ncells,
                              const int64_t* cell_nodes, int64_t ndofs, const int64_t* cell_dofs, const int64_t* colptr,
                              const int64_t* rowval, int32_t ncolors, const int64_t* color_ptr,
                              const int64_t* color_cells, const double* k, int64_t np, const int64_t* prescribed,
                              double source) {
  if (!B) return FEM_ERR_ARG;
  fem_ctx* c = B->c;
  if (B->finalized) return set_error(c, FEM_ERR_STATE, "fem_batch_add_problem: batch already finalized");
  if ((dim != 2 && dim != 3) || nnodes <= 0 || ncells <= 0 || ndofs <= 0 || !xyz || !cell_nodes || !cell_dofs ||
      !colptr || !rowval || ncolors <= 0 || !color_ptr || !color_cells || !k || np < 0 || (np > 0 && !prescribed))
    return set_error(c, FEM_ERR_ARG, "fem_batch_add_problem: invalid arguments");
  if (!B->probs.empty() && B->probs[0].source != source)
    return set_error(c, FEM_ERR_ARG, "fem_batch_add_problem: every problem of a batch uses the same source f");
  const int npc = dim == 3 ? 8 : 4;
  BProblemHost P;
  P.dim = dim; P.nnodes = nnodes; P.ncells = ncells; P.ndofs = ndofs; P.source = source;
  P.xyz.assign(xyz, xyz + nnodes * dim);
  P.cell_nodes.resize(ncells * npc);
  P.cell_dofs.resize(ncells * npc);
  for (int64_t i = 0; i < ncells * npc; ++i) {
    if (cell_nodes[i] < 1 || cell_nodes[i] > nnodes || cell_dofs[i] < 1 || cell_dofs[i] > ndofs)
      return set_error(c, FEM_ERR_ARG, "cell node/dof out of range");
    P.cell_nodes[i] = (int32_t)(cell_nodes[i] - 1);
    P.cell_dofs[i] = (int32_t)(cell_dofs[i] - 1);
  }
  if (colptr[0] != 1 || colptr[ndofs] - 1 < ndofs) return set_error(c, FEM_ERR_ARG, "invalid colptr");
  P.nnz = colptr[ndofs] - 1;
  P.colptr.resize(ndofs + 1);
  P.rowval.resize(P.nnz);
  for (int64_t j = 0; j <= ndofs; ++j) {
    if (j > 0 && colptr[j] < colptr[j - 1]) return set_error(c, FEM_ERR_ARG, "colptr not monotone");
    P.colptr[j] = (int32_t)(colptr[j] - 1);
  }
  for (int64_t p = 0; p < P.nnz; ++p) {
    if (rowval[p] < 1 || rowval[p] > ndofs) return set_error(c, FEM_ERR_ARG, "rowval out of range");
    P.rowval[p] = (int32_t)(rowval[p] - 1);
  }
  P.color_ptr.assign(color_ptr, color_ptr + ncolors + 1);
  if (P.color_ptr[0] != 0 || P.color_ptr[ncolors] != ncells) return set_error(c, FEM_ERR_ARG, "colors must cover all cells");
  for (int i = 0; i < ncolors; ++i)
    if (P.color_ptr[i + 1] < P.color_ptr[i]) return set_error(c, FEM_ERR_ARG, "color_ptr not monotone");
  P.color_cells.resize(ncells);
  std::vector<uint8_t> seen(ncells, 0);
  for (int64_t s = 0; s < ncells; ++s) {
    const int64_t cc = color_cells[s];
    if (cc < 1 || cc > ncells || seen[cc - 1]) return set_error(c, FEM_ERR_ARG, "colors must cover all cells once");
    seen[cc - 1] = 1;
    P.color_cells[s] = (int32_t)(cc - 1);
  }
  // a colour must not hold two cells sharing a dof (the scatter is atomic-free)
  {
    std::vector<int32_t> stamp(ndofs, -1);
    for (int col = 0; col < ncolors; ++col)
      for (int64_t s = P.color_ptr[col]; s < P.color_ptr[col + 1]; ++s) {
        const int32_t cell = P.color_cells[s];
        for (int a = 0; a < npc; ++a) {
          int32_t& st = stamp[P.cell_dofs[(int64_t)cell * npc + a]];
          if (st == col) {
            bool self = false;   // a cell listing the same dof twice is not a conflict
            for (int b = 0; b < a; ++b) self |= P.cell_dofs[(int64_t)cell * npc + b] == P.cell_dofs[(int64_t)cell * npc + a];
            if (!self) return set_error(c, FEM_ERR_ARG, "two cells of one colour share a dof");
          }
          st = col;
        }
      }
  }
  P.k.assign(k, k + ncells);
  P.prescribed.resize(np);
  for (int64_t i = 0; i < np; ++i) {
    if (prescribed[i] < 1 || prescribed[i] > ndofs) return set_error(c, FEM_ERR_ARG, "prescribed dof out of range");
    if (i > 0 && prescribed[i] <= prescribed[i - 1])
      return set_error(c, FEM_ERR_ARG, "prescribed dofs must be sorted and unique");
    P.prescribed[i] = (int32_t)(prescribed[i] - 1);
  }
  B->probs.push_back(std::move(P));
  return FEM_OK;
}

int32_t fem_batch_finalize(fem_batch* B, const double* quad_N, const double* quad_dN, const double* quad_w,
                           const double* hex_N, const double* hex_dN, const double* hex_w) {
  if (!B) return FEM_ERR_ARG;
  fem_ctx* c = B->c;
  if (B->finalized || B->probs.empty()) return set_error(c, FEM_ERR_STATE, "fem_batch_finalize: nothing to finalize");
  if (!quad_N || !quad_dN || !quad_w || !hex_N || !hex_dN || !hex_w) return set_error(c, FEM_ERR_ARG, "missing tables");
  FEM_HIP(c, hipSetDevice(c->device));
  const int np_ = (int)B->probs.size();
  int64_t nn = 0, nd = 0, nz = 0, ns = 0;
  for (auto& P : B->probs) { nn += P.nnodes; nd += P.ndofs; nz += P.nnz; ns += P.ncells; }
  if (nz >= INT32_MAX || nd >= INT32_MAX || nn >= INT32_MAX || ns * 64 >= INT32_MAX)
    return set_error(c, FEM_ERR_ARG, "batch too large for int32 tables");
  // concatenated mesh (SoA), colour-slot tables, pattern
  std::vector<double> X[3];
  for (auto& v : X) v.assign(nn, 0.0);
  std::vector<int32_t> cn8(ns * 8, 0), cd8(ns * 8, 0), slotprob(ns), rowprob(nd), pres;
  std::vector<double> kslot(ns);
  std::vector<int32_t> cp(nd + 1, 0), rv(nz);
  std::vector<int64_t> pdof(np_ + 1, 0);
  std::vector<BProbDev> pd(np_);
  int64_t n0 = 0, d0 = 0, z0 = 0, s0 = 0;
  for (int p = 0; p < np_; ++p) {
    auto& P = B->probs[p];
    const int npc = P.dim == 3 ? 8 : 4;
    pd[p] = {P.dim, (int32_t)n0, (int32_t)d0, (int32_t)P.ndofs};
    pdof[p] = d0;
    for (int64_t i = 0; i < P.nnodes; ++i)
      for (int d = 0; d < P.dim; ++d) X[d][n0 + i] = P.xyz[i * P.dim + d];
    for (int64_t s = 0; s < P.ncells; ++s) {
      const int32_t cell = P.color_cells[s];
      for (int a = 0; a < npc; ++a) {
        cn8[(s0 + s) * 8 + a] = (int32_t)(n0 + P.cell_nodes[(int64_t)cell * npc + a]);
        cd8[(s0 + s) * 8 + a] = (int32_t)(d0 + P.cell_dofs[(int64_t)cell * npc + a]);
      }
      kslot[s0 + s] = P.k[cell];
      slotprob[s0 + s] = p;
    }
    for (int64_t j = 0; j < P.ndofs; ++j) {
      cp[d0 + j] = (int32_t)(z0 + P.colptr[j]);
      rowprob[d0 + j] = p;
    }
    for (int64_t q = 0; q < P.nnz; ++q) rv[z0 + q] = (int32_t)(d0 + P.rowval[q]);
    for (int32_t d : P.prescribed) pres.push_back((int32_t)(d0 + d));
    n0 += P.nnodes; d0 += P.ndofs; z0 += P.nnz; s0 += P.ncells;
  }
  cp[nd] = (int32_t)nz;
  pdof[np_] = nd;
  int rc;
  hipStream_t st = c->stream;
  // context mesh coordinates (3 SoA arrays; 2D problems have z = 0), vectors, pattern
  for (int d = 0; d < 3; ++d) {
    if ((rc = dev_alloc(c, &c->d_x[d], nn))) return rc;
    FEM_HIP(c, hipMemcpyAsync(c->d_x[d], X[d].data(), 8 * nn, hipMemcpyHostToDevice, st));
  }
  c->sdim = 3; c->nnodes = nn; c->ncells = ns; c->npc = 8;
  c->ndofs = nd; c->ndpc = 8;
  if ((rc = dev_alloc(c, &c->d_f, nd)) || (rc = dev_alloc(c, &c->d_u, nd))) return rc;
  FEM_HIP(c, hipMemsetAsync(c->d_f, 0, 8 * nd, st));
  FEM_HIP(c, hipMemsetAsync(c->d_u, 0, 8 * nd, st));
  {
    std::vector<int64_t> cp1(nd + 1), rv1(nz);
    for (int64_t j = 0; j <= nd; ++j) cp1[j] = (int64_t)cp[j] + 1;
    for (int64_t q = 0; q < nz; ++q) rv1[q] = (int64_t)rv[q] + 1;
    if ((rc = fem_pattern_upload(c, nd, cp1.data(), rv1.data()))) return rc;
  }
  // per-slot nz table (64 per slot; quads use the first 16, entry b*nb+a)
  std::vector<int32_t> cnz(ns * 64, -1);
  for (int64_t s = 0; s < ns; ++s) {
    const int nb = B->probs[slotprob[s]].dim == 3 ? 8 : 4;
    for (int b = 0; b < nb; ++b) {
      const int32_t col = cd8[s * 8 + b];
      const int32_t* lo = rv.data() + cp[col];
      const int32_t* hi = rv.data() + cp[col + 1];
      for (int a = 0; a < nb; ++a) {
        const int32_t row = cd8[s * 8 + a];
        const int32_t* it = std::lower_bound(lo, hi, row);
        cnz[s * 64 + b * nb + a] = (it != hi && *it == row) ? (int32_t)(it - rv.data()) : -1;
      }
    }
  }
  std::vector<double> tab(kTabSize, 0.0);
  for (int i = 0; i < 16; ++i) tab[kTQ_N + i] = quad_N[i];
  for (int i = 0; i < 32; ++i) tab[kTQ_dN + i] = quad_dN[i];
  for (int i = 0; i < 4; ++i) tab[kTQ_w + i] = quad_w[i];
  for (int i = 0; i < 64; ++i) tab[kTH_N + i] = hex_N[i];
  for (int i = 0; i < 192; ++i) tab[kTH_dN + i] = hex_dN[i];
  for (int i = 0; i < 8; ++i) tab[kTH_w + i] = hex_w[i];
  std::vector<float> tabf(tab.begin(), tab.end());
  if ((rc = dev_alloc(c, &B->d_cnodes8, ns * 8)) || (rc = dev_alloc(c, &B->d_cdofs8, ns * 8)) ||
      (rc = dev_alloc(c, &B->d_cnz, ns * 64)) || (rc = dev_alloc(c, &B->d_kslot, ns)) ||
      (rc = dev_alloc(c, &B->d_slotprob, ns)) || (rc = dev_alloc(c, &B->d_done, np_)) ||
      (rc = dev_alloc(c, &B->d_q, 1)) || (rc = dev_alloc(c, &B->d_probs, np_)) ||
      (rc = dev_alloc(c, &B->d_rowprob, nd)) || (rc = dev_alloc(c, &B->d_pdof, np_ + 1)) ||
      (rc = dev_alloc(c, &B->d_tab, kTabSize)) || (rc = dev_alloc(c, &B->d_tabf, kTabSize)) ||
      (rc = dev_alloc(c, &B->d_m, np_)) ||
      (rc = dev_alloc(c, &B->d_x32, nd)) || (rc = dev_alloc(c, &B->d_r32, nd)) ||
      (rc = dev_alloc(c, &B->d_p32, nd)) || (rc = dev_alloc(c, &B->d_w32, nd)) ||
      (rc = dev_alloc(c, &B->d_s32, nd)) || (rc = dev_alloc(c, &B->d_sc, nd)) ||
      (rc = dev_alloc(c, &B->d_dsq, nd)) ||
      (rc = dev_alloc(c, &B->d_res, nd)) || (rc = dev_alloc(c, (BScal**)&B->d_scal, 2 * np_)) ||
      (rc = dev_alloc(c, &B->d_nlist, 2)) || (rc = dev_alloc(c, &B->d_handing, np_)))
    return rc;
  FEM_HIP(c, hipMemcpyAsync(B->d_cnodes8, cn8.data(), 4 * cn8.size(), hipMemcpyHostToDevice, st));
  FEM_HIP(c, hipMemcpyAsync(B->d_cdofs8, cd8.data(), 4 * cd8.size(), hipMemcpyHostToDevice, st));
  FEM_HIP(c, hipMemcpyAsync(B->d_cnz, cnz.data(), 4 * cnz.size(), hipMemcpyHostToDevice, st));
  FEM_HIP(c, hipMemcpyAsync(B->d_kslot, kslot.data(), 8 * ns, hipMemcpyHostToDevice, st));
  FEM_HIP(c, hipMemcpyAsync(B->d_slotprob, slotprob.data(), 4 * ns, hipMemcpyHostToDevice, st));
  FEM_HIP(c, hipMemcpyAsync(B->d_probs, pd.data(), sizeof(BProbDev) * np_, hipMemcpyHostToDevice, st));
  FEM_HIP(c, hipMemcpyAsync(B->d_rowprob, rowprob.data(), 4 * nd, hipMemcpyHostToDevice, st));
  FEM_HIP(c, hipMemcpyAsync(B->d_pdof, pdof.data(), 8 * (np_ + 1), hipMemcpyHostToDevice, st));
  FEM_HIP(c, hipMemcpyAsync(B->d_tab, tab.data(), 8 * kTabSize, hipMemcpyHostToDevice, st));
  FEM_HIP(c, hipMemcpyAsync(B->d_tabf, tabf.data(), 4 * kTabSize, hipMemcpyHostToDevice, st));
  FEM_HIP(c, hipStreamSynchronize(st));
  // constraints through the single-problem path (isconstrained mask, prescribed list)
  {
    std::vector<int64_t> pres1(pres.size());
    std::vector<double> g(pres.size(), 0.0);
    for (size_t i = 0; i < pres.size(); ++i) pres1[i] = (int64_t)pres[i] + 1;   // ascending by construction
    if ((rc = fem_constraints_upload(c, (int64_t)pres.size(), pres1.data(), g.data()))) return rc;
  }
  // reduction segments: ≤ kSeg consecutive rows, never crossing a problem boundary
  {
    std::vector<int32_t> sprob, pseg(np_ + 1);
    std::vector<int64_t> srow;
    for (int p = 0; p < np_; ++p) {
      pseg[p] = (int32_t)sprob.size();
      for (int64_t r = pdof[p]; r < pdof[p + 1]; r += kSeg) {
        sprob.push_back(p);
        srow.push_back(r);
      }
    }
    pseg[np_] = (int32_t)sprob.size();
    srow.push_back(nd);
    B->nseg = (int)sprob.size();
    if ((rc = dev_alloc(c, &B->d_sprob, B->nseg)) || (rc = dev_alloc(c, &B->d_srow, B->nseg + 1)) ||
        (rc = dev_alloc(c, &B->d_pseg, np_ + 1)) || (rc = dev_alloc(c, &B->d_spart, 6 * B->nseg)) ||
        (rc = dev_alloc(c, &B->d_aseg, B->nseg)) || (rc = dev_alloc(c, &B->d_hseg, B->nseg)))
      return rc;
    FEM_HIP(c, hipMemcpyAsync(B->d_sprob, sprob.data(), 4 * sprob.size(), hipMemcpyHostToDevice, st));
    FEM_HIP(c, hipMemcpyAsync(B->d_srow, srow.data(), 8 * srow.size(), hipMemcpyHostToDevice, st));
    FEM_HIP(c, hipMemcpyAsync(B->d_pseg, pseg.data(), 4 * pseg.size(), hipMemcpyHostToDevice, st));
    FEM_HIP(c, hipStreamSynchronize(st));
    B->h_pseg = pseg;
    // sliced operator: slices of ≤ 64 rows inside each segment
    std::vector<int32_t> bsl(B->nseg + 1), slrow, slend, boff(1, 0), bbase;
    int64_t blen = 0;
    int32_t span = 0;
    for (int sg = 0; sg < B->nseg; ++sg) {
      bsl[sg] = (int32_t)slrow.size();
      for (int64_t r0 = srow[sg]; r0 < srow[sg + 1]; r0 += 64) {
        const int64_t r1 = std::min<int64_t>(r0 + 64, srow[sg + 1]);
        int32_t W = 0, lo = INT32_MAX, hi = 0;
        for (int64_t r = r0; r < r1; ++r) {
          W = std::max(W, cp[r + 1] - cp[r]);
          for (int32_t q = cp[r]; q < cp[r + 1]; ++q) {
            lo = std::min(lo, rv[q]);
            hi = std::max(hi, rv[q]);
          }
        }
        if (lo > hi) lo = hi = (int32_t)r0;
        slrow.push_back((int32_t)r0);
        slend.push_back((int32_t)r1);
        bbase.push_back(lo);
        span = std::max(span, hi - lo);
        blen += 64 * (int64_t)W;
        if (blen >= INT32_MAX) return set_error(c, FEM_ERR_ARG, "batch too large for the sliced operator");
        boff.push_back((int32_t)blen);
      }
    }
    bsl[B->nseg] = (int32_t)slrow.size();
    const int64_t nsl = (int64_t)slrow.size();
    const bool comp = span < 65536;
    std::vector<int32_t> bsrc(blen, -1), bcol(comp ? 0 : blen);
    std::vector<uint16_t> broff(comp ? blen : 0);
    for (int64_t k = 0; k < nsl; ++k) {
      const int64_t r0 = slrow[k], r1 = slend[k];
      const int32_t W = (boff[k + 1] - boff[k]) / 64;
      for (int l = 0; l < 64; ++l) {
        const int64_t r = r0 + l;
        for (int32_t e = 0; e < W; ++e) {
          const int64_t at = boff[k] + 64 * (int64_t)e + l;
          const int32_t q = r < r1 ? cp[r] + e : -1;
          const bool real = r < r1 && q < cp[r + 1];
          bsrc[at] = real ? q : -1;   // rows past the slice and short rows: padding
          const int32_t col = real ? rv[q] : bbase[k];
          if (comp) broff[at] = (uint16_t)(col - bbase[k]);
          else bcol[at] = col;
        }
      }
    }
    if ((rc = dev_alloc(c, &B->d_bsl, B->nseg + 1)) || (rc = dev_alloc(c, &B->d_slrow, nsl)) ||
        (rc = dev_alloc(c, &B->d_boff, nsl + 1)) || (rc = dev_alloc(c, &B->d_bbase, nsl)) ||
        (rc = dev_alloc(c, &B->d_bsrc, blen)) || (rc = dev_alloc(c, &B->d_bval, blen)))
      return rc;
    if (comp) rc = dev_alloc(c, &B->d_broff, blen);
    else rc = dev_alloc(c, &B->d_bcol, blen);
    if (rc) return rc;
    FEM_HIP(c, hipMemcpyAsync(B->d_bsl, bsl.data(), 4 * bsl.size(), hipMemcpyHostToDevice, st));
    FEM_HIP(c, hipMemcpyAsync(B->d_slrow, slrow.data(), 4 * nsl, hipMemcpyHostToDevice, st));
    FEM_HIP(c, hipMemcpyAsync(B->d_boff, boff.data(), 4 * boff.size(), hipMemcpyHostToDevice, st));
    FEM_HIP(c, hipMemcpyAsync(B->d_bbase, bbase.data(), 4 * nsl, hipMemcpyHostToDevice, st));
    FEM_HIP(c, hipMemcpyAsync(B->d_bsrc, bsrc.data(), 4 * blen, hipMemcpyHostToDevice, st));
    if (comp) FEM_HIP(c, hipMemcpyAsync(B->d_broff, broff.data(), 2 * blen, hipMemcpyHostToDevice, st));
    else FEM_HIP(c, hipMemcpyAsync(B->d_bcol, bcol.data(), 4 * blen, hipMemcpyHostToDevice, st));
    FEM_HIP(c, hipStreamSynchronize(st));
    B->nslices = nsl;
    B->blen = blen;
  }
  B->nslots = ns;
  B->finalized = true;
  B->chunk = 0;
  return FEM_OK;
}

int32_t fem_batch_autotune(fem_batch* B, int32_t precision, int32_t* cpw_out, int32_t* chunk_out) {
  if (!B) return FEM_ERR_ARG;
  fem_ctx* c = B->c;
  if (!B->finalized) return set_error(c, FEM_ERR_STATE, "fem_batch_autotune: finalize first");
  FEM_HIP(c, hipSetDevice(c->device));
  const bool mixed = precision == FEM_MIXED;
  // 64-cell workgroups are not tried: each work item ends with an agent-scope release fence
  // (its nzval writes must reach the next colour's items on other XCDs), which writes back the
  // XCD's L2, and 64-cell items made thousands of them per launch (0.1-0.34 s per assembly,
  // against 6.6 ms for the chosen shape)
  const int cpws[2] = {128, 256};
  const int chunks[5] = {256, 512, 1024, 2048, 4096};
  float best = 1e30f;
  int bc = 128, bk = 512;
  for (int cw : cpws)
    for (int ch : chunks) {
      if (!shape_ok(cw, ch, mixed)) continue;
      float t = 0.f, ms = 0.f;
      int rc = timed_assemble(B, cw, ch, mixed, &t);  // warm-up (graph build)
      if (rc) return rc;
      for (int r = 0; r < 3; ++r) {
        if ((rc = timed_assemble(B, cw, ch, mixed, &t))) return rc;
        ms += t;
      }
      if (ms < best) { best = ms; bc = cw; bk = ch; }
    }
  B->cells_per_wg = bc;
  if (cpw_out) *cpw_out = bc;
  if (chunk_out) *chunk_out = bk;
  return build_items(B, bk);
}

int32_t fem_batch_run(fem_batch* B, const fem_batch_cfg* cfg, fem_batch_info* info) {
  if (!B || !cfg) return FEM_ERR_ARG;
  fem_ctx* c = B->c;
  if (!B->finalized) return set_error(c, FEM_ERR_STATE, "fem_batch_run: finalize first");
  FEM_HIP(c, hipSetDevice(c->device));
  const int np_ = (int)B->probs.size();
  const bool mixed = cfg->precision == FEM_MIXED;
  int cpw = cfg->cells_per_wg ? cfg->cells_per_wg : (B->cells_per_wg ? B->cells_per_wg : 128);
  const int chunk = cfg->chunk ? cfg->chunk : (B->chunk ? B->chunk : 512);
  // a shape tuned for FP32 element math may exceed the FP64 staging limit: fall back
  if (!cfg->cells_per_wg && !shape_ok(cpw, chunk, mixed)) cpw = 128;
  if (!shape_ok(cpw, chunk, mixed)) return set_error(c, FEM_ERR_ARG, "fem_batch_run: unsupported launch shape");
  if (!(cfg->rtol > 0.0)) return set_error(c, FEM_ERR_ARG, "fem_batch_run: rtol must be positive");
  const int64_t n = c->ndofs;
  hipStream_t st = c->stream;
  BScal* d_s = (BScal*)B->d_scal;
  int rc;
  // assembly: zero + persistent work-queue kernel (one graph)
  FEM_HIP(c, hipEventRecord(B->e_start, st));
  if ((rc = assemble_graph_launch(B, cpw, chunk, mixed))) return rc;
  FEM_HIP(c, hipEventRecord(B->e_asm, st));
  // apply!: per-problem meandiag of the assembled diagonal, zero rows/cols, K[d,d] = m, f[d] = 0
  k_b_meandiag<<<np_, 256, 0, st>>>(B->d_pdof, c->d_diag, c->d_nzval, B->d_m);
  k_b_zero_rc<<<grid_n(n), 256, 0, st>>>(n, c->d_colptr, c->d_rowval, c->d_isc, c->d_nzval);
  if (c->np > 0)
    k_b_setdiag<<<grid_n(c->np), 256, 0, st>>>(c->np, c->d_pres, c->d_diag, B->d_rowprob, B->d_m, c->d_nzval, c->d_f);
  FEM_HIP(c, hipEventRecord(B->e_apply, st));
  // solver set-up: S = diag(K)^-1/2, the scaled FP32 operator, records (every problem
  // starts with a hand-over from u = 0)
  double* ph = B->d_spart + 4 * B->nseg;   // hand-over partials (2 per segment)
  k_bc_scale<<<grid_n(n), 256, 0, st>>>(n, c->d_diag, c->d_nzval, B->d_sc, B->d_dsq);
  {
    const int vb = (int)std::min<int64_t>(8192, (B->nslices + 3) / 4);
    if (B->d_broff)
      k_bc_sell_vals<true><<<vb, 256, 0, st>>>(B->nslices, n, B->d_slrow, B->d_boff, B->d_bbase, B->d_broff, nullptr,
                                               B->d_bsrc, c->d_nzval, B->d_sc, B->d_bval);
    else
      k_bc_sell_vals<false><<<vb, 256, 0, st>>>(B->nslices, n, B->d_slrow, B->d_boff, B->d_bbase, nullptr, B->d_bcol,
                                                B->d_bsrc, c->d_nzval, B->d_sc, B->d_bval);
  }
  k_b_zero<<<grid_n(n), 256, 0, st>>>(n, c->d_u);
  k_bc_zero32<<<grid_n(n), 256, 0, st>>>(n, B->d_x32);
  k_bc_zero32<<<grid_n(n), 256, 0, st>>>(n, B->d_p32);
  k_bc_zero32<<<grid_n(n), 256, 0, st>>>(n, B->d_s32);
  const double delta = cfg->ru_delta > 0 ? cfg->ru_delta : 0.1;
  k_bc_fsq<<<B->nseg, kSegThreads, 0, st>>>(B->nseg, B->d_srow, c->d_f, ph);
  k_bc_init<<<np_, 64, 0, st>>>(np_, B->d_pseg, ph, cfg->rtol * cfg->rtol, delta * delta, d_s);
  FEM_HIP(c, hipGetLastError());
  std::vector<BScal> hs(np_);
  FEM_HIP(c, hipMemcpyAsync(hs.data(), d_s, sizeof(BScal) * np_, hipMemcpyDeviceToHost, st));
  FEM_HIP(c, hipStreamSynchronize(st));
  if ((rc = assemble_errors(B))) return rc;
  std::vector<double> done_ms(np_, -1.0);
  std::vector<int> iters(np_, 0);
  int ndone = 0;
  float t_asm = 0.f, t_apply = 0.f;
  (void)hipEventElapsedTime(&t_asm, B->e_start, B->e_asm);
  (void)hipEventElapsedTime(&t_apply, B->e_asm, B->e_apply);
  for (int p = 0; p < np_; ++p)
    if (hs[p].done) {   // f = 0: u = 0 is exact
      done_ms[p] = t_asm + t_apply;
      ++ndone;
    }
  std::vector<int32_t> list, handing(np_);
  // hand-over of every problem whose recurrence stopped (or of all unfinished ones at maxit):
  // fold x̂ into u, FP64 true residual, done or reseed (r̂ = S res, re-orthogonalised p, α)
  auto hand_over = [&](bool all) -> int {
    list.clear();
    bool any = false;
    for (int p = 0; p < np_; ++p) {
      handing[p] = !hs[p].done && (hs[p].hold || all);
      if (!handing[p]) continue;
      any = true;
      for (int k = B->h_pseg[p]; k < B->h_pseg[p + 1]; ++k) list.push_back(k);
    }
    if (!any) return FEM_OK;
    const int32_t len = (int32_t)list.size();
    FEM_HIP(c, hipMemcpyAsync(B->d_hseg, list.data(), 4 * list.size(), hipMemcpyHostToDevice, st));
    FEM_HIP(c, hipMemcpyAsync(B->d_nlist + 1, &len, 4, hipMemcpyHostToDevice, st));
    FEM_HIP(c, hipMemcpyAsync(B->d_handing, handing.data(), 4 * np_, hipMemcpyHostToDevice, st));
    k_bc_fold<<<len, kSegThreads, 0, st>>>(B->d_hseg, B->d_nlist, B->d_srow, B->d_sc, B->d_x32, c->d_u);
    k_bc_resid<<<len, kSpmvThreads, 0, st>>>(B->d_hseg, B->d_nlist, B->d_srow, c->d_colptr, c->d_rowval, c->d_nzval,
                                            c->d_u, c->d_f, B->d_res, ph);
    k_bc_true<<<np_, 64, 0, st>>>(np_, B->d_pseg, ph, B->d_handing, d_s);
    k_bc_reseed<<<len, kSegThreads, 0, st>>>(B->d_hseg, B->d_nlist, B->d_sprob, B->d_srow, B->d_sc, B->d_res,
                                             B->d_r32, B->d_p32, d_s, ph);
    k_bc_reseed_fin<<<np_, 64, 0, st>>>(np_, B->d_pseg, ph, B->d_handing, d_s);
    launch_bc_sell(B, len, B->d_hseg, 1, d_s, ph);
    k_bc_pfix<<<len, kSegThreads, 0, st>>>(B->d_hseg, B->d_nlist, B->d_sprob, B->d_srow, B->d_r32, B->d_w32,
                                           B->d_p32, B->d_s32, d_s, ph);
    k_bc_alpha<<<np_, 64, 0, st>>>(np_, B->d_pseg, ph, B->d_handing, d_s);
    FEM_HIP(c, hipGetLastError());
    FEM_HIP(c, hipEventRecord(B->e_mark, st));
    FEM_HIP(c, hipMemcpyAsync(hs.data(), d_s, sizeof(BScal) * np_, hipMemcpyDeviceToHost, st));
    FEM_HIP(c, hipStreamSynchronize(st));
    float el = 0.f;
    (void)hipEventElapsedTime(&el, B->e_start, B->e_mark);
    for (int p = 0; p < np_; ++p)
      if (hs[p].done && done_ms[p] < 0) {
        done_ms[p] = el;
        iters[p] = hs[p].iters;
        ++ndone;
      }
    return FEM_OK;
  };
  if ((rc = hand_over(false))) return rc;
  const int maxit = cfg->maxit > 0 ? cfg->maxit : 20000;
  int launched = 0, idle = 0;
  while (ndone < np_ && launched < maxit) {
    // the segments of every problem still in its recurrence
    list.clear();
    for (int p = 0; p < np_; ++p)
      if (!hs[p].done && !hs[p].hold)
        for (int k = B->h_pseg[p]; k < B->h_pseg[p + 1]; ++k) list.push_back(k);
    if (!list.empty()) {
      const int32_t len = (int32_t)list.size();
      FEM_HIP(c, hipMemcpyAsync(B->d_aseg, list.data(), 4 * list.size(), hipMemcpyHostToDevice, st));
      FEM_HIP(c, hipMemcpyAsync(B->d_nlist, &len, 4, hipMemcpyHostToDevice, st));
      int width = 64;
      while (width < len) width *= 2;
      width = std::min(width, B->nseg);
      if ((rc = ensure_cg_graph(B, width))) return rc;
      FEM_HIP(c, hipGraphLaunch(*cg_graph_for(B, width), st));
      launched += kCgBatch;
      FEM_HIP(c, hipMemcpyAsync(hs.data(), d_s, sizeof(BScal) * np_, hipMemcpyDeviceToHost, st));
      FEM_HIP(c, hipStreamSynchronize(st));
      idle = 0;
    } else if (++idle > 64) {
      break;   // every unfinished problem breaks down at once on each reseed: give up
    }
    if ((rc = hand_over(launched >= maxit))) return rc;
  }
  FEM_HIP(c, hipEventRecord(B->e_mark, st));
  FEM_HIP(c, hipEventSynchronize(B->e_mark));
  float total = 0.f;
  (void)hipEventElapsedTime(&total, B->e_start, B->e_mark);
  for (int p = 0; p < np_; ++p)
    if (done_ms[p] < 0) iters[p] = hs[p].iters;   // unconverged: iterations so far, completion -1
  B->last_times_ms = done_ms;
  B->last_iters = iters;
  ++c->nzval_version;
  if (info) {
    info->assemble_ms = t_asm;
    info->apply_ms = t_apply;
    info->solve_ms = total - t_asm - t_apply;
    info->total_ms = total;
    info->converged = ndone;
    info->total_iterations = launched;
  }
  if (ndone < np_) {
    std::string msg = "batched CG: problems missed rtol within maxit:";
    int shown = 0;
    for (int p = 0; p < np_ && shown < 8; ++p) {
      if (hs[p].done) continue;
      char buf[200];
      snprintf(buf, sizeof buf, " [%d: iters %d, true %.3e, hold %d]", p, hs[p].iters,
               std::sqrt(hs[p].rr / hs[p].bnorm2), hs[p].hold);
      msg += buf;
      ++shown;
    }
    return set_error(c, FEM_ERR_NOCONV, msg);
  }
  return FEM_OK;
}

int32_t fem_batch_problem_stats(const fem_batch* B, double* completion_ms, int32_t* iterations) {
  if (!B) return FEM_ERR_ARG;
  for (size_t p = 0; p < B->last_times_ms.size(); ++p) {
    if (completion_ms) completion_ms[p] = B->last_times_ms[p];
    if (iterations) iterations[p] = B->last_iters[p];
  }
  return FEM_OK;
}

int32_t fem_batch_download(fem_batch* B, int32_t prob, double* nzval, double* f, double* u) {
  if (!B) return FEM_ERR_ARG;
  fem_ctx* c = B->c;
  if (!B->finalized || prob < 0 || prob >= (int)B->probs.size())
    return set_error(c, FEM_ERR_ARG, "fem_batch_download: no such problem");
  FEM_HIP(c, hipSetDevice(c->device));
  int64_t d0 = 0, z0 = 0;
  for (int p = 0; p < prob; ++p) { d0 += B->probs[p].ndofs; z0 += B->probs[p].nnz; }
  const auto& P = B->probs[prob];
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  if (nzval) FEM_HIP(c, hipMemcpy(nzval, c->d_nzval + z0, 8 * P.nnz, hipMemcpyDeviceToHost));
  if (f) FEM_HIP(c, hipMemcpy(f, c->d_f + d0, 8 * P.ndofs, hipMemcpyDeviceToHost));
  if (u) FEM_HIP(c, hipMemcpy(u, c->d_u + d0, 8 * P.ndofs, hipMemcpyDeviceToHost));
  return FEM_OK;
}

int32_t fem_batch_assemble_only(fem_batch* B, int32_t precision, int32_t cells_per_wg, int32_t chunk, double* ms) {
  if (!B) return FEM_ERR_ARG;
  fem_ctx* c = B->c;
  if (!B->finalized) return set_error(c, FEM_ERR_STATE, "fem_batch_assemble_only: finalize first");
  FEM_HIP(c, hipSetDevice(c->device));
  const bool mixed = precision == FEM_MIXED;
  int cpw = cells_per_wg ? cells_per_wg : (B->cells_per_wg ? B->cells_per_wg : 128);
  const int ch = chunk ? chunk : (B->chunk ? B->chunk : 512);
  if (!cells_per_wg && !shape_ok(cpw, ch, mixed)) cpw = 128;
  if (!shape_ok(cpw, ch, mixed)) return set_error(c, FEM_ERR_ARG, "unsupported launch shape");
  float t = 0.f;
  int rc = timed_assemble(B, cpw, ch, mixed, &t);
  if (ms) *ms = t;
  ++c->nzval_version;
  return rc;
}

}  // extern "C"
