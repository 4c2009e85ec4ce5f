// Solvers on the assembled, constrained K (replaces u = K \ f, heat_equation.jl:218;
// the reference has no iterative solver on this path).
//
//  * FEM_SOLVE_CG_FP64   Jacobi-preconditioned CG, FP64 SpMV and vectors.
//  * FEM_SOLVE_IR_MIXED  the same CG with an FP32 operator and FP32 vectors (FP64 dot
//                        products), kept on track by reliable updates: whenever its
//                        residual has fallen by δ the FP32 iterate is folded into the FP64
//                        u and the FP64 residual f − K u replaces the FP32 one.
//
// Jacobi preconditioning is applied as a symmetric scaling: the solver works on
// Â = S K S with S = diag(K)^-1/2, whose values are written once per assembly into the
// solver's own copy of K.  Unpreconditioned CG on Â x̂ = S f is Jacobi-PCG on K u = f
// (u = S x̂) in exact arithmetic, but needs neither the z = D⁻¹ r vector nor D⁻¹ in the
// iteration, so the fused vector step moves 9 instead of 12 vectors' worth of bytes.
//
// K is symmetric after apply! (rows and columns of prescribed dofs are zeroed together),
// so the CSC columns double as rows.  The solver's operator is one of two copies:
//  * lattice planes, when the dofs of a scalar Q1 field sit on a tensor-product lattice
//    (found from the cell topology, fem_lattice.hip) and every entry couples lattice
//    neighbours: the centre and the forward neighbours' values per padded lattice row, no
//    column indices; the backward neighbours read their mirror's plane (checked symmetric
//    at every fill).  The mixed solver's default; the recurrence vectors then live in the
//    padded lattice numbering (map / dofp translate at the hand-overs).
//  * a sliced ELL copy (slices of 64 rows, one wave each) built from the columns, for
//    everything else and for the FP64 solver by default.
// The FP64 residual f − K u uses the CSC arrays themselves (CSR-stream kernel).
//
// Reductions are fence-free and fixed-order: blocks write partials that the next kernel
// sums (by default one block, k_cg_scal), in the same order every time, so results are
// bitwise reproducible and the host never sits in the iteration.  A batch of iterations is
// captured once into a hipGraph and replayed; kernels launched past a hand-over return at
// once on the device flag.
#include <cmath>
#include <cstring>

#include "fem_internal.h"
#include "fem_cell.h"

namespace fem {

constexpr int kBlock = 256;
constexpr int kMaxVecBlocks = 2048;
constexpr int kBatch = 16;           // CG iterations per batch: the first after a reseed, and direct launches
constexpr int kBatchMin = 8;         // captured batches: kBatchMin << i iterations, i = 0, 1, 2


struct SpmvShape {
  int R = 0;          // rows per block
  int blocks = 0;
  int cap = 2048;     // LDS products per chunk
  int L = 1;          // lanes per row in the reduction phase
};

struct GraphCache {
  hipGraphExec_t exec = nullptr;
  const void* key[9] = {};   // operands and the kernel variant the graph was captured with
};

// Sliced ELL copy of K for the solver (slices of 64 rows = one wave): entry k of row
// 64 s + l is stored at off[s] + 64 k + l, so a wave's loads of values and column
// offsets are unit-stride.  Rows shorter than the slice width are padded with value 0
// and the slice's smallest column.  The scaled values s_i K_ij s_j are re-gathered from
// nzval (src map) whenever K changes; the structure is built once per pattern.
struct SellOp {
  int64_t nslices = 0, len = 0;
  int32_t* off = nullptr;      // [nslices + 1] first entry of each slice
  int32_t* base = nullptr;     // [nslices] smallest column index of the slice
  uint16_t* roff = nullptr;    // [len] column − base (when every slice spans < 65536 columns)
  int32_t* col = nullptr;      // [len] column (otherwise)
  int32_t* src = nullptr;      // [len] position in nzval, −1 for padding
  float* val32 = nullptr;
  double* val64 = nullptr;
  uint64_t ver32 = 0, ver64 = 0;   // nzval_version the values were scaled from (+1)
  int jac32 = -1, jac64 = -1;      // the scaling they carry (1 Jacobi, 0 none)
  bool built = false;
};

// Lattice copy of Â, used in place of the sliced ELL when the dofs sit on a tensor-product
// lattice (find_lattice) and every pattern entry couples lattice neighbours.  The lattice is
// padded by one node on each side of every used axis: padded row P = (ix + 1) +
// px ((iy + 1) + py (iz + 1)).  Neighbour codes are (dx + 1) + 3 (dy + 1) + 9 (dz + 1)
// (centre 13); plane s holds Â(P, P + o_s) for the centre (s = 0) and the forward neighbours
// (code 13 + s, s = 1 .. ns − 1: 13 in 3D, 4 in 2D).  Row P's backward neighbour P − o_s is
// read from plane s at row P − o_s (Â is symmetric, which the value fill checks), so the
// operator is half the values and no column indices.  Padding rows hold zero values, so
// the solver's vectors (in padded numbering) stay zero there.
struct LatOffsets {
  int32_t o[14];
};

struct LatOp {
  int state = 0;               // 0 not tried, 1 usable, −1 not a lattice
  int ns = 0;
  int64_t px = 0, py = 0, pxy = 0, npad = 0;
  int64_t r_lo = 0, nrows = 0; // rows the SpMV computes
  LatOffsets off{};
  int32_t* map = nullptr;      // [n] padded row of each dof
  int32_t* dofp = nullptr;     // [npad] dof of each padded row, −1 on padding
  float* val32 = nullptr;
  double* val64 = nullptr;
  float* mir32 = nullptr;      // [ns − 1][npad] mirror planes of the symmetry check
  double* mir64 = nullptr;
  uint64_t ver32 = 0, ver64 = 0;   // nzval_version the planes were filled from (+1)
  int jac32 = -1, jac64 = -1;
  bool sym32 = false, sym64 = false;   // the last fill found Â symmetric
  bool unit32 = false, unit64 = false; // ... and its diagonal all ones
  unsigned long long* cnt = nullptr;   // [4] fill counters
  int grid = 0;
  std::vector<int32_t> hmap;   // host copy of map
};

// Cell-coefficient form of Â on the lattice: the element-assembly operator of
// docs/src/literate-howto/gpu_assembly.jl:266-305 (K x as small dense products with one
// element matrix per cell), here with Dirichlet rows and columns and one number per cell.
// When every cell is a congruent copy of cell 0, K = Σ_c k_c K¹ (K¹ cell 0's element matrix
// for a unit coefficient, k_c the per-cell parameter) with the Dirichlet rows and columns
// zeroed, so
//   (Â x)_P = t_P Σ_{cells c ∋ P} k_c Σ_b K¹[a_c(P)][b] t_{Q_c(b)} x_{Q_c(b)},
// t = S on free rows and 0 on prescribed rows and padding; a prescribed row keeps Â's unit
// diagonal, (Â x)_P = x_P.  The solver uses this form only after k_cell_check found every
// entry of the scaled K equal to it; it then reads one coefficient per cell and the scaling
// instead of the value planes.  K¹ is row-major over Ferrite's local corner order.
struct CellKref {
  double k[64];
};
struct CellOp {
  int state = 0;               // 0 not tried, 1 cell origins placed, −1 not for this mesh
  int dim = 0;
  int32_t* corg = nullptr;     // [ncells] padded row of each cell's first corner
  double* kc64 = nullptr;      // [npad] k_c at the origin row of cell c, 0 where no cell starts
  float* kc32 = nullptr;
  double* t64 = nullptr;       // [npad] s on free rows, 0 on prescribed rows and padding
  float* t32 = nullptr;
  double x0[8][3] = {};        // corners of cell 0
  CellKref kref{};             // K¹ of the last check
  double f[8] = {};            // its values per corner-offset class (box cells)
  uint64_t ver32 = 0, ver64 = 0, aux32 = 0, aux64 = 0;   // versions the last check saw (+1)
  int jac32 = -1, jac64 = -1;
  bool ok32 = false, ok64 = false;   // the last check passed
  unsigned long long* cnt = nullptr; // [4] check counters
  int lds32 = 0, lds64 = 0;    // dynamic LDS of the SpMV (bytes)
  int ne = 0;                  // window entries per thread of the SpMV
  // 3D: the z-marching SpMV, per precision [FP32, FP64] (zn > 0): t·x and coefficient window
  // entries per thread, rows per thread, in-plane tiles; the blocks, planes per chunk, work
  // items and LDS bytes
  int zn[2] = {}, kn[2] = {}, rpt[2] = {}, ntx[2] = {};
  int grid32 = 0, grid64 = 0, lz32 = 0, lz64 = 0, ldsz32 = 0, ldsz64 = 0;
  int64_t nitems32 = 0, nitems64 = 0;
  // FP64 residual f − K u on the cell form: allowed when the last check found every entry
  // within kCellTolR (versions it saw +1); u and K u in the padded numbering, the free-row mask,
  // a scalar record with done = fresh = 0 for the SpMV's early-exit test
  bool okr = false;
  uint64_t verr = 0, auxr = 0, auxm = 0;
  double *up64 = nullptr, *yp64 = nullptr, *tm64 = nullptr;
  Scal* s0 = nullptr;
};

struct SolverWork {
  int64_t n = 0;
  SellOp sell;
  LatOp lat;
  CellOp cells;
  bool use_lat = false;        // the current solve runs on a lattice operator (padded vectors)
  bool use_cells = false;      // ... in its cell-coefficient form
  bool use_elem = false;       // the current solve runs on the stored element matrices
  ElemOp* eop = nullptr;       // their per-solve vectors (fem_elemop.hip)
  int elem_grid = 0;
  int64_t cap32 = 0, cap64 = 0;   // length of the recurrence vectors
  int sell_grid = 0;           // blocks of the SELL SpMV (fixed: its partial count)
  SpmvShape sh64;              // CSR-stream FP64 residual
  // recurrence vectors of the scaled system: x̂, r̂, w = Â r̂, p, s = Â p (FP64 solver / FP32
  // recurrence of the mixed solver)
  double *x = nullptr, *r = nullptr, *wv = nullptr, *p = nullptr, *sv = nullptr;
  float *x32 = nullptr, *r32 = nullptr, *wv32 = nullptr, *p32 = nullptr, *sv32 = nullptr;
  double* sc = nullptr;        // S = diag(K)^-1/2 (1 without Jacobi)
  double* res = nullptr;       // FP64 residual f − K u
  double* part_s = nullptr;    // SpMV partials, one per block
  double* part_v = nullptr;    // vector-kernel partials, two buffers of 2 per block
  Scal* s = nullptr;           // two scalar records (read one, write the other)
  hipEvent_t tev[2 * kBatch] = {};   // SpMV start/end events (time_kernels)
  Scal* hscal = nullptr;       // pinned host copy of scalar record 0 (read_scal)
  GraphCache g64[3], g32[3];   // per batch length kBatchMin << i
};

static void destroy_graph(GraphCache& g) {
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  g = GraphCache{};
}

void free_solver(fem_ctx* c) {
  if (!c->solver) return;
  SolverWork* w = c->solver;
  for (GraphCache& g : w->g64) destroy_graph(g);
  for (GraphCache& g : w->g32) destroy_graph(g);
  for (hipEvent_t& e : w->tev)
    if (e) (void)hipEventDestroy(e);
  if (w->hscal) (void)hipHostFree(w->hscal);
  dev_free(&w->x); dev_free(&w->r); dev_free(&w->wv); dev_free(&w->p); dev_free(&w->sv);
  dev_free(&w->x32); dev_free(&w->r32); dev_free(&w->wv32); dev_free(&w->p32); dev_free(&w->sv32);
  dev_free(&w->sc); dev_free(&w->res);
  dev_free(&w->part_s); dev_free(&w->part_v); dev_free(&w->s);
  SellOp& o = w->sell;
  dev_free(&o.off); dev_free(&o.base); dev_free(&o.roff); dev_free(&o.col); dev_free(&o.src);
  dev_free(&o.val32); dev_free(&o.val64);
  LatOp& l = w->lat;
  dev_free(&l.map); dev_free(&l.dofp); dev_free(&l.val32); dev_free(&l.val64); dev_free(&l.cnt);
  dev_free(&l.mir32); dev_free(&l.mir64);
  elem_op_free(&w->eop);
  CellOp& cl = w->cells;
  dev_free(&cl.corg); dev_free(&cl.kc64); dev_free(&cl.kc32); dev_free(&cl.t64); dev_free(&cl.t32);
  dev_free(&cl.cnt); dev_free(&cl.up64); dev_free(&cl.yp64); dev_free(&cl.tm64); dev_free(&cl.s0);
  delete w;
  c->solver = nullptr;
}

// --------------------------------------------------------------------------
// reductions (no fences: partials are consumed by the next kernel)
// --------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Block sum, valid in thread 0; `sh` holds kBlock/64 doubles.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kBlock / 64; ++i) r += sh[i];
  return r;
}

// Fixed-order sum of part[0..m) (stride apart) by one block; valid in thread 0.
__device__ __forceinline__ double sum_parts(const double* __restrict__ part, int m, int stride, double* sh) {
  double v = 0.0;
  for (int i = threadIdx.x; i < m; i += kBlock) v += part[(int64_t)i * stride];
  return block_sum(v, sh);
}

// Wide variant for the scalar kernels (kWide threads, 8 loads in flight per thread).
constexpr int kWide = 1024;
__device__ __forceinline__ double wide_sum(const double* __restrict__ part, int m, int stride, double* sh) {
  double v = 0.0;
  for (int i0 = 0; i0 < m; i0 += 8 * kWide) {
    double t[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int i = i0 + k * kWide + threadIdx.x;
      t[k] = i < m ? part[(int64_t)i * stride] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) v += t[k];
  }
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kWide / 64; ++i) r += sh[i];
  return r;
}


// --------------------------------------------------------------------------
// CSR-stream SpMV over the CSC arrays (the FP64 residual of the unscaled K):
//   MODE 0: y = K x, partial Σ x·y ;  MODE 1: y = f − K x, partial Σ y²
// A block owns R consecutive rows whose entries it streams once, fully coalesced, into
// LDS products; then L lanes per row sum them (lane k: products k, k+L, ... in order,
// combined by a fixed butterfly: deterministic).  Column indices are the pattern's int32
// rowval (COMP = false) or 16-bit offsets from a per-block base (COMP = true).
// --------------------------------------------------------------------------
// Loads N consecutive elements with 16-byte vector loads (the address is N-element aligned).
template <typename T, int N>
__device__ __forceinline__ void vec_load(const T* __restrict__ p, T (&out)[N]) {
  constexpr int VB = 16 / sizeof(T);       // elements per 16-byte load
  static_assert(N % VB == 0, "vector width");
#pragma unroll
  for (int k = 0; k < N / VB; ++k) {
    const uint4 u = *reinterpret_cast<const uint4*>(p + k * VB);
    const T* e = reinterpret_cast<const T*>(&u);
#pragma unroll
    for (int j = 0; j < VB; ++j) out[k * VB + j] = e[j];
  }
}

template <typename V, typename X, int MODE, int CAP, int L, bool COMP>
__global__ __launch_bounds__(kBlock) void k_spmv(int64_t n, int R, const int32_t* __restrict__ colptr,
                                                 const int32_t* __restrict__ rowval,
                                                 const uint16_t* __restrict__ roff,
                                                 const int32_t* __restrict__ bbase, const V* __restrict__ val,
                                                 const X* __restrict__ x, const double* __restrict__ f,
                                                 X* __restrict__ y, double* __restrict__ part,
                                                 const Scal* __restrict__ s) {
  using A = typename std::conditional<sizeof(X) == 8, double, float>::type;
  if (MODE == 0 && s->done) return;
  __shared__ A prod[CAP];
  __shared__ int32_t sptr[kBlock + 1];
  __shared__ double sh[kBlock / 64];
  const int tid = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * R;
  const int nr = (int)min<int64_t>(R, n - r0);
  for (int t = tid; t <= nr; t += kBlock) sptr[t] = colptr[r0 + t];
  int32_t base = 0;
  if constexpr (COMP) base = bbase[blockIdx.x];
  __syncthreads();
  const int32_t e0 = sptr[0], e1 = sptr[nr];
  const int row = tid / L, sub = tid - (tid / L) * L;
  A acc = A(0);
  // Entries are streamed in windows of CAP starting at an 8-aligned index; thread t owns the
  // PER = CAP/kBlock consecutive entries [w0 + PER t, w0 + PER t + PER), loaded as whole vectors.
  constexpr int PER = CAP / kBlock;
  static_assert(PER % 8 == 0, "window per thread");
  for (int32_t w0 = e0 & ~7; w0 < e1; w0 += CAP) {
    const int32_t t0 = w0 + PER * tid;
    V v[PER];
    int32_t ri[PER];
    if (t0 + PER <= e1 && t0 >= e0) {
      vec_load<V, PER>(val + t0, v);
      if constexpr (COMP) {
        uint16_t o[PER];
        vec_load<uint16_t, PER>(roff + t0, o);
#pragma unroll
        for (int k = 0; k < PER; ++k) ri[k] = base + (int32_t)o[k];
      } else {
        vec_load<int32_t, PER>(rowval + t0, ri);
      }
    } else {
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int32_t e = t0 + k;
        const bool ok = e >= e0 && e < e1;
        v[k] = ok ? val[e] : V(0);
        if constexpr (COMP) ri[k] = ok ? base + (int32_t)roff[e] : 0;
        else ri[k] = ok ? rowval[e] : 0;
      }
    }
    X xv[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) xv[k] = x[ri[k]];
#pragma unroll
    for (int k = 0; k < PER; ++k) prod[PER * tid + k] = (A)v[k] * (A)xv[k];
    __syncthreads();
    if (row < nr) {
      const int32_t a = max(sptr[row], w0), b = min(sptr[row + 1], w0 + CAP);
      for (int32_t e = a + sub; e < b; e += L) acc += prod[e - w0];
    }
    __syncthreads();
  }
#pragma unroll
  for (int o = L / 2; o > 0; o >>= 1) acc += __shfl_down(acc, o, L);
  double d = 0.0;
  if (row < nr && sub == 0) {
    const int64_t j = r0 + row;
    if (MODE == 0) {
      y[j] = (X)acc;
      d = (double)x[j] * (double)(X)acc;
    } else {
      const double rj = f[j] - (double)acc;
      y[j] = (X)rj;
      d = rj * rj;
    }
  }
  const double b = block_sum(d, sh);
  if (tid == 0) part[blockIdx.x] = b;
}


// --------------------------------------------------------------------------
// Sliced ELL (SELL-64) SpMV: one wave per slice of 64 rows, lane l owns row 64 s + l and
// sums its entries in pattern order (deterministic).  Value and offset loads are
// unit-stride across the wave; x is gathered.  U entries per lane are loaded before their
// x gathers are issued (memory-level parallelism).  Fixed grid, so the partial count is
// fixed.  CONTIG: every block owns one contiguous range of slices and the blocks that the
// hardware deals to one XCD (blockIdx % 8) own neighbouring ranges, so an XCD's rows and
// the x window they gather stay in that XCD's L2; otherwise blocks stride over slices.
//   MODE 0: y = K x, partial Σ x·y ;  MODE 1: y = f − K x, partial Σ y²
// --------------------------------------------------------------------------
// Streamed once per launch: loaded non-temporally (NT) when asked, so that they do not
// evict the gathered vector from L2.
template <bool NT, typename T>
__device__ __forceinline__ T stream_load(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

template <typename V, typename X, int MODE, bool COMP, int U, bool CONTIG, bool NT = false>
__global__ __launch_bounds__(kBlock) void k_sell(int64_t n, int64_t nslices, const int32_t* __restrict__ off,
                                                 const int32_t* __restrict__ base,
                                                 const uint16_t* __restrict__ roff,
                                                 const int32_t* __restrict__ col, const V* __restrict__ val,
                                                 const X* __restrict__ x, const double* __restrict__ f,
                                                 X* __restrict__ y, double* __restrict__ part,
                                                 const Scal* __restrict__ s) {
  using A = typename std::conditional<sizeof(X) == 8, double, float>::type;
  // nothing to do once the recurrence has stopped, nor in the step right after a reseed
  // (that step moves x̂ and r̂ along the p, s the reseed built and reads neither y nor δ)
  if (MODE == 0 && (s->done || s->fresh)) return;
  __shared__ double sh[kBlock / 64];
  constexpr int WPB = kBlock / 64;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  int64_t s_begin, s_end, s_step;
  if (CONTIG) {   // gridDim.x is a multiple of 8
    const int64_t per = gridDim.x / 8;
    const int64_t lb = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    s_begin = lb * nslices / gridDim.x + wave;
    s_end = (lb + 1) * nslices / gridDim.x;
    s_step = WPB;
  } else {
    s_begin = (int64_t)blockIdx.x * WPB + wave;
    s_end = nslices;
    s_step = (int64_t)gridDim.x * WPB;
  }
  double d = 0.0;
  for (int64_t sl = s_begin; sl < s_end; sl += s_step) {
    const int32_t o0 = off[sl];
    const int W = (off[sl + 1] - o0) >> 6;
    const int32_t b = COMP ? base[sl] : 0;
    const int64_t row = sl * 64 + lane;
    const int32_t e0 = o0 + lane;
    A acc = A(0);
    int k = 0;
    for (; k + U <= W; k += U) {
      V v[U];
      int32_t ci[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t e = e0 + 64 * (k + u);
        v[u] = stream_load<NT>(val + e);
        ci[u] = COMP ? b + (int32_t)stream_load<NT>(roff + e) : stream_load<NT>(col + e);
      }
      X xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) xv[u] = x[ci[u]];
#pragma unroll
      for (int u = 0; u < U; ++u) acc += (A)v[u] * (A)xv[u];
    }
    for (; k < W; ++k) {
      const int32_t e = e0 + 64 * k;
      const int32_t ci = COMP ? b + (int32_t)stream_load<NT>(roff + e) : stream_load<NT>(col + e);
      acc += (A)stream_load<NT>(val + e) * (A)x[ci];
    }
    if (row < n) {
      if (MODE == 0) {
        y[row] = (X)acc;
        d += (double)x[row] * (double)(X)acc;
      } else {
        const double rj = f[row] - (double)acc;
        y[row] = (X)rj;
        d += rj * rj;
      }
    }
  }
  const double bsum = block_sum(d, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bsum;
}

// Slice widths, bases and column spans (one wave per slice).
__global__ __launch_bounds__(kBlock) void k_sell_shape(int64_t n, int64_t nslices, const int32_t* __restrict__ colptr,
                                                       const int32_t* __restrict__ rowval, int32_t* __restrict__ width,
                                                       int32_t* __restrict__ base, int32_t* __restrict__ span) {
  const int64_t sl = (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (sl >= nslices) return;
  const int64_t row = sl * 64 + (threadIdx.x & 63);
  int len = 0, mn = INT_MAX, mx = INT_MIN;
  if (row < n) {
    const int32_t a = colptr[row], e = colptr[row + 1];
    len = e - a;
    if (len > 0) {
      mn = rowval[a];          // rows are ascending within a column of the pattern
      mx = rowval[e - 1];
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    len = max(len, __shfl_down(len, o, 64));
    mn = min(mn, __shfl_down(mn, o, 64));
    mx = max(mx, __shfl_down(mx, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    width[sl] = len;
    base[sl] = mn == INT_MAX ? 0 : mn;
    span[sl] = mn == INT_MAX ? 0 : mx - mn;
  }
}

template <bool COMP>
__global__ __launch_bounds__(kBlock) void k_sell_fill(int64_t n, int64_t nslices, const int32_t* __restrict__ colptr,
                                                      const int32_t* __restrict__ rowval,
                                                      const int32_t* __restrict__ off,
                                                      const int32_t* __restrict__ base, uint16_t* __restrict__ roff,
                                                      int32_t* __restrict__ col, int32_t* __restrict__ src) {
  const int64_t sl = (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (sl >= nslices) return;
  const int lane = threadIdx.x & 63;
  const int64_t row = sl * 64 + lane;
  const int32_t o0 = off[sl];
  const int W = (off[sl + 1] - o0) >> 6;
  const int32_t b = base[sl];
  const int32_t a = row < n ? colptr[row] : 0;
  const int len = row < n ? colptr[row + 1] - a : 0;
  for (int k = 0; k < W; ++k) {
    const int32_t e = o0 + 64 * k + lane;
    const bool real = k < len;
    const int32_t cidx = real ? rowval[a + k] : b;
    if (COMP) roff[e] = (uint16_t)(cidx - b);
    else col[e] = cidx;
    src[e] = real ? a + k : -1;
  }
}

// Scaled values of the sliced copy: val[e] = s_row K[row, col] s_col (0 on padding); one
// wave per slice like the SpMV, 8 entries per lane in flight.
template <typename V, bool COMP>
__global__ __launch_bounds__(kBlock) void k_sell_vals(int64_t n, int64_t nslices, const int32_t* __restrict__ off,
                                                      const int32_t* __restrict__ base,
                                                      const uint16_t* __restrict__ roff,
                                                      const int32_t* __restrict__ col, const int32_t* __restrict__ src,
                                                      const double* __restrict__ nzval, const double* __restrict__ sc,
                                                      V* __restrict__ val) {
  constexpr int U = 8;
  for (int64_t sl = (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); sl < nslices;
       sl += (int64_t)gridDim.x * (kBlock / 64)) {
    const int lane = threadIdx.x & 63;
    const int64_t row = sl * 64 + lane;
    const int32_t o0 = off[sl];
    const int W = (off[sl + 1] - o0) >> 6;
    const int32_t b = COMP ? base[sl] : 0;
    const double sr = row < n ? sc[row] : 0.0;
    for (int k0 = 0; k0 < W; k0 += U) {
      int32_t p[U], ci[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t e = o0 + 64 * (k0 + u) + lane;
        const bool in = k0 + u < W;
        p[u] = in ? src[e] : -1;
        ci[u] = in ? (COMP ? b + (int32_t)roff[e] : col[e]) : 0;
      }
      double a[U], sj[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        a[u] = p[u] >= 0 ? nzval[p[u]] : 0.0;
        sj[u] = p[u] >= 0 ? sc[ci[u]] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (k0 + u < W) val[o0 + 64 * (k0 + u) + lane] = (V)(a[u] * sr * sj[u]);
    }
  }
}

// --------------------------------------------------------------------------
// Lattice operator (LatOp): plane fill, symmetry check and SpMV.
// --------------------------------------------------------------------------
struct LatGeom {
  uint32_t px, pxy, ofs;       // ofs = 1 + px + pxy: shifts every neighbour offset to [0, 2 ofs]
  int ns;                      // planes: forward codes 13 .. 12 + ns, backward 14 − ns .. 12
  int64_t npad;
};

// Neighbour code (dx + 1) + 3 (dy + 1) + 9 (dz + 1) of the padded row at offset d, or −1 when
// it is not a lattice neighbour.  d + ofs read as the digits (dz + 1, dy + 1, dx + 1) in the
// mixed radix (pxy, px, 1); as every real node lies at least one padding node away from the
// lattice edge, digits ≤ 2 identify a neighbour uniquely.  In 2D (pz = 1) dz + 1 is always 1.
// The digits come from compares, not divisions: for 0 ≤ t ≤ 2 ofs the z digit is at most 2
// (py ≥ 3), a y digit above 2 leaves an x digit ≥ px > 2, and both are rejected alike.
__device__ __forceinline__ int lat_code(int64_t d, const LatGeom& g) {
  const int64_t t = d + g.ofs;
  if (t < 0 || t > 2 * (int64_t)g.ofs) return -1;
  const uint32_t u = (uint32_t)t;
  const uint32_t cz = (u >= g.pxy) + (u >= 2 * g.pxy), rz = u - cz * g.pxy;
  const uint32_t cy = (rz >= g.px) + (rz >= 2 * g.px), cx = rz - cy * g.px;
  if (cx > 2) return -1;
  return (int)(cx + 3 * cy + 9 * cz);
}

// Adds block totals of two counters to cnt[i0], cnt[i0 + 1].
__device__ __forceinline__ void count2(double a, double b, unsigned long long* cnt, int i0, double* sh) {
  a = block_sum(a, sh);
  if (threadIdx.x == 0 && a != 0.0) atomicAdd(cnt + i0, (unsigned long long)a);
  b = block_sum(b, sh);
  if (threadIdx.x == 0 && b != 0.0) atomicAdd(cnt + i0 + 1, (unsigned long long)b);
}

// Plane fill, kLatLanes lanes per padded row P (dofp[P] = its dof j, −1 on padding), one
// entry of column j of the CSC arrays (= row j of the symmetric K) per lane, v = K[i, j] s_j s_i
// (the product order of the sliced copy):
//   i = j or a forward neighbour (code ≥ 13): val[code − 13][P] = v;
//   i a backward neighbour:                     mir[12 − code][P(i)] = v, i.e. where row P(i)
//                                               stores K[j, i] in val[13 − code][P(i)].
// Â is symmetric iff planes 1.. of val equal mir (k_lat_cmp).  cnt[1] counts entries that
// are not lattice neighbours, cnt[0] diagonal values other than 1 (with the Jacobi scaling
// every diagonal value is 1 after rounding, and the SpMV need not read the centre plane).
constexpr int kLatLanes = 32;

template <typename V>
__global__ __launch_bounds__(kBlock) void k_lat_fill(int64_t r_lo, int64_t nrows, const int32_t* __restrict__ dofp,
                                                     const int32_t* __restrict__ colptr,
                                                     const int32_t* __restrict__ rowval,
                                                     const double* __restrict__ nzval, const double* __restrict__ sc,
                                                     const int32_t* __restrict__ map, LatGeom g, V* __restrict__ val,
                                                     V* __restrict__ mir, unsigned long long* __restrict__ cnt) {
  __shared__ double sh[kBlock / 64];
  const int lane = threadIdx.x % kLatLanes;
  const int64_t groups = (int64_t)gridDim.x * (kBlock / kLatLanes);
  double bad = 0.0, nonunit = 0.0;
  for (int64_t q = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kLatLanes; q < nrows; q += groups) {
    const int64_t P = r_lo + q;
    const int32_t j = dofp[P];
    if (j < 0) continue;
    const double sr = sc[j];
    const int32_t p1 = colptr[j + 1];
    for (int32_t p = colptr[j] + lane; p < p1; p += kLatLanes) {
      const int32_t i = rowval[p];
      const int64_t Pi = map[i];
      const int code = lat_code(Pi - P, g);
      // a neighbour outside the used axes (dz ≠ 0 in 2D) cannot occur for lattice rows; it
      // is counted as a non-neighbour rather than written past the planes
      if (code < 0 || code - 13 >= g.ns || 13 - code >= g.ns) {
        bad += 1.0;
        continue;
      }
      const V v = (V)(nzval[p] * sr * sc[i]);
      if (code >= 13) val[(int64_t)(code - 13) * g.npad + P] = v;
      else mir[(int64_t)(12 - code) * g.npad + Pi] = v;
      if (code == 13 && v != V(1)) nonunit += 1.0;
    }
  }
  count2(nonunit, bad, cnt, 0, sh);
}

// Symmetry check: a[k] vs b[k] for k < m (the forward planes against their mirrors) to tol
// relative (the operator's precision); cnt[3] counts mismatches.
template <typename V>
__global__ __launch_bounds__(kBlock) void k_lat_cmp(int64_t m, const V* __restrict__ a, const V* __restrict__ b,
                                                    double tol, unsigned long long* __restrict__ cnt) {
  __shared__ double sh[kBlock / 64];
  double bad = 0.0;
  for (int64_t k = blockIdx.x * (int64_t)kBlock + threadIdx.x; k < m; k += (int64_t)gridDim.x * kBlock) {
    const double x = (double)a[k], y = (double)b[k];
    if (!(fabs(x - y) <= tol * fmax(fabs(x), fabs(y)))) bad += 1.0;
  }
  count2(0.0, bad, cnt, 2, sh);
}

// Element at a 32-bit byte offset from a wave-uniform base (the global-load form with a
// scalar base and a 32-bit vector offset: no 64-bit address per stream in VGPRs).
template <typename T>
__device__ __forceinline__ T at_byte(const T* __restrict__ base, uint32_t boff) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + boff);
}

// w = Â x on the lattice planes, rows [r_lo, r_lo + nrows) of the padded numbering; partial
// x·w per block.  All 2 ns − 1 values and x entries of a row are loaded before the sum (fixed
// order).  Byte offsets are 32-bit (the planes of one operator span < 4 GiB: checked by the
// host).  MAP 0: grid-stride over all rows; MAP 1 (gridDim.x a multiple of 8): the blocks of
// one XCD (b ≡ x mod 8) sweep one contiguous eighth of the rows in order, so the rows a
// backward neighbour reads were loaded by the same XCD shortly before.
template <typename V, typename X, int NS, int MAP, bool UNIT>
__global__ __launch_bounds__(kBlock) void k_lat_spmv(int64_t r_lo, int64_t nrows, int64_t npad, LatOffsets lo,
                                                     const V* __restrict__ val, const X* __restrict__ x,
                                                     X* __restrict__ y, double* __restrict__ part,
                                                     const Scal* __restrict__ s) {
  using A = typename std::conditional<sizeof(X) == 8, double, float>::type;
  if (s->done || s->fresh) return;
  __shared__ double sh[kBlock / 64];
  int64_t i0, iend, step;
  if (MAP == 1) {
    const int64_t per = gridDim.x / 8, xcd = blockIdx.x % 8, k = blockIdx.x / 8;
    i0 = xcd * nrows / 8 + k * kBlock + threadIdx.x;
    iend = (xcd + 1) * nrows / 8;
    step = per * kBlock;
  } else {
    i0 = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    iend = nrows;
    step = (int64_t)gridDim.x * kBlock;
  }
  double d = 0.0;
  for (int64_t i = i0; i < iend; i += step) {
    const uint32_t r = (uint32_t)(r_lo + i);
    V v[2 * NS - 1];
    X xv[2 * NS - 1];
    v[0] = UNIT ? V(1) : at_byte(val, r * (uint32_t)sizeof(V));   // UNIT: every diagonal value is 1
    xv[0] = at_byte(x, r * (uint32_t)sizeof(X));
#pragma unroll
    for (int k = 1; k < NS; ++k) {
      const uint32_t o = (uint32_t)lo.o[k];
      const V* pl = val + k * npad;
      v[2 * k - 1] = at_byte(pl, r * (uint32_t)sizeof(V));
      v[2 * k] = at_byte(pl, (r - o) * (uint32_t)sizeof(V));
      xv[2 * k - 1] = at_byte(x, (r + o) * (uint32_t)sizeof(X));
      xv[2 * k] = at_byte(x, (r - o) * (uint32_t)sizeof(X));
    }
    A acc = (A)v[0] * (A)xv[0];
#pragma unroll
    for (int k = 1; k < 2 * NS - 1; ++k) acc += (A)v[k] * (A)xv[k];
    y[r] = (X)acc;
    d += (double)xv[0] * (double)(X)acc;
  }
  const double b = block_sum(d, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
}

// ---- cell-coefficient form of Â on the lattice ----
// kc[corg[e]] = param[e] (FP64 and FP32 copies); rows where no cell starts keep 0.
__global__ void k_cell_coef(int64_t ncells, const int32_t* __restrict__ corg, const double* __restrict__ param,
                            double* __restrict__ kc64, float* __restrict__ kc32) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ncells; e += (int64_t)gridDim.x * blockDim.x) {
    const double k = param[e];
    kc64[corg[e]] = k;
    kc32[corg[e]] = (float)k;
  }
}

// t[P] = s of the row's dof, 0 on prescribed dofs (isc) and on padding.
template <typename X>
__global__ void k_cell_scale(int64_t npad, const int32_t* __restrict__ dofp, const double* __restrict__ sc,
                             const uint8_t* __restrict__ isc, X* __restrict__ t) {
  for (int64_t P = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; P < npad; P += (int64_t)gridDim.x * blockDim.x) {
    const int32_t j = dofp[P];
    t[P] = (j >= 0 && !(isc && isc[j])) ? (X)sc[j] : X(0);
  }
}

// FP64 residual f − K u on the cell form, allowed when every entry of the scaled K matched it
// to kCellTolR (a few FP64 roundings of Â's unit diagonal): u is spread into the padded
// numbering, the FP64 cell SpMV with the free-row mask for t gives (K u)_P on free rows, and
// prescribed rows use K_jj u_j from nzval.
constexpr double kCellTolR = 1e-14;

// up[map[j]] = u[j]; padding entries of up keep 0.
__global__ void k_to_pad(int64_t n, const int32_t* __restrict__ map, const double* __restrict__ u,
                         double* __restrict__ up) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
    up[map[j]] = u[j];
}

// m[P] = 1 on free rows, 0 on prescribed rows and padding.
__global__ void k_cell_mask(int64_t npad, const int32_t* __restrict__ dofp, const uint8_t* __restrict__ isc,
                            double* __restrict__ m) {
  for (int64_t P = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; P < npad; P += (int64_t)gridDim.x * blockDim.x) {
    const int32_t j = dofp[P];
    m[P] = (j >= 0 && !(isc && isc[j])) ? 1.0 : 0.0;
  }
}

// res_j = f_j − (K u)_j: y[map[j]] on free rows, K_jj u_j on prescribed rows (apply! left only
// the diagonal there); partial Σ res² per block.
__global__ __launch_bounds__(kBlock) void k_res_cells(int64_t n, const int32_t* __restrict__ map,
                                                      const double* __restrict__ yp, const double* __restrict__ f,
                                                      const double* __restrict__ u, const uint8_t* __restrict__ isc,
                                                      const int32_t* __restrict__ diag,
                                                      const double* __restrict__ nzval, double* __restrict__ res,
                                                      double* __restrict__ part) {
  __shared__ double sh[kBlock / 64];
  double acc = 0.0;
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n; j += (int64_t)gridDim.x * kBlock) {
    double ku;
    if (isc && isc[j]) {
      const int32_t p = diag[j];
      ku = (p >= 0 ? nzval[p] : 0.0) * u[j];
    } else {
      ku = yp[map[j]];
    }
    const double r = f[j] - ku;
    res[j] = r;
    acc += r * r;
  }
  const double b = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
}

// Workgroup barrier that orders LDS only: waits for this wave's LDS operations, then meets the
// other waves.  __syncthreads() also fences global memory, which waits for every load in
// flight — including loads issued ahead for later steps, which this barrier leaves in flight.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Entry-by-entry check of the scaled K against the cell form, in FP64, kCheckLanes lanes per
// column j of the CSC arrays (padded row P = map[j], loaded together with the column's
// bounds), each lane up to 32 / kCheckLanes entries whose loads are issued together (a
// lattice column has at most 27), and the row's eight cell coefficients loaded once.  Â's diagonal is 1, so tol is relative to it.  cnt[0] entries that differ by more
// than tol, cnt[1] lattice-neighbour entries in the pattern, cnt[2] lattice neighbours of
// the rows (the cell form couples every two corners of a cell, so each must be a pattern
// entry), cnt[3] entries that differ by more than tolr (the residual's bar); other entries
// must hold zero.  It runs once K¹ has passed cell_classes, so an entry's expected value is
// f[class] times the sum of the coefficients of the cells that hold both corners (selects,
// no table lookups).
// Measured at 128³ (kernel trace): 526 µs with 8 lanes per padded row (its dof looked up
// first), 488 µs with 4, 643 µs with 16; 476 µs with 4 per column.  The per-row loads (the
// coefficients, the neighbour flags, the column bounds) are paid once per lane group, so
// fewer lanes per row amortise them over more rows per wave; summing the coefficients over
// x, y, z instead of the eight-way select measured 491 µs.
constexpr int kCheckLanes = 4;
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_cell_check(int64_t n, const int32_t* __restrict__ dofp,
                                                       const int32_t* __restrict__ colptr,
                                                       const int32_t* __restrict__ rowval,
                                                       const double* __restrict__ nzval, const double* __restrict__ sc,
                                                       const uint8_t* __restrict__ isc, const int32_t* __restrict__ map,
                                                       LatGeom g, const double* __restrict__ kc, CellKref kr,
                                                       double tol, double tolr,
                                                       unsigned long long* __restrict__ cnt) {
  __shared__ double sh[kBlock / 64];
  constexpr int NB = 1 << DIM;
  constexpr int U = 32 / kCheckLanes;   // entries per lane per pass (a pass covers 32)
  const int lane = threadIdx.x % kCheckLanes;
  const int64_t groups = (int64_t)gridDim.x * (kBlock / kCheckLanes);
  const int64_t px = g.px, pxy = g.pxy;
  double bad = 0.0, badr = 0.0, seen = 0.0, want = 0.0;
  // K¹ is in its class form (K¹[a][b] = f[a ⊕ b], set by cell_classes): f = the row of corner 0
  double f8[NB];
#pragma unroll
  for (int m = 0; m < NB; ++m) f8[m] = kr.k[cell_lid(m & 1, (m >> 1) & 1, m >> 2)];
  for (int64_t q = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kCheckLanes; q < n; q += groups) {
    const int32_t j = (int32_t)q;
    const int64_t P = map[j];
    const int32_t p0 = colptr[j], p1 = colptr[j + 1];
    if (lane == 0) {   // the lattice fills its box: the neighbours are a product over the axes
      const double nx = 1.0 + (dofp[P - 1] >= 0) + (dofp[P + 1] >= 0);
      const double ny = 1.0 + (dofp[P - px] >= 0) + (dofp[P + px] >= 0);
      const double nz = DIM == 3 ? 1.0 + (dofp[P - pxy] >= 0) + (dofp[P + pxy] >= 0) : 1.0;
      want += nx * ny * nz;
    }
    double k8[NB];   // cell with origin P − e
#pragma unroll
    for (int e = 0; e < NB; ++e) k8[e] = kc[P - (e & 1) - px * ((e >> 1) & 1) - pxy * (e >> 2)];
    const double sj = sc[j];
    const bool pj = isc && isc[j];
    for (int32_t pb = p0; pb < p1; pb += U * kCheckLanes) {
      int32_t iv[U];
      double a[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t p = pb + lane + u * kCheckLanes;
        iv[u] = p < p1 ? rowval[p] : -1;
        a[u] = p < p1 ? nzval[p] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t i = iv[u];
        if (i < 0) continue;
        const double av = a[u] * sj * sc[i];
        const int code = lat_code((int64_t)map[i] - P, g);
        if (code < 0 || (DIM == 2 && code / 9 != 1)) {
          if (!(fabs(av) <= tol)) bad += 1.0;
          if (!(fabs(av) <= tolr)) badr += 1.0;
          continue;
        }
        seen += 1.0;
        const int dx = code % 3 - 1, dy = (code / 3) % 3 - 1, dz = code / 9 - 1;
        double m = 0.0;
        if (pj || (isc && isc[i])) {
          m = i == j ? 1.0 : 0.0;
        } else {
          // the cells with origin P − e that hold both corners (e_a = 0 where d_a = +1, 1 where
          // d_a = −1), all with the class value f[|dx| + 2 |dy| + 4 |dz|]
          double ks = 0.0;
#pragma unroll
          for (int e = 0; e < NB; ++e) {
            const int ex = e & 1, ey = (e >> 1) & 1, ez = e >> 2;
            const bool both = (dx == 0 || ex == (dx < 0)) && (dy == 0 || ey == (dy < 0)) && (dz == 0 || ez == (dz < 0));
            ks += both ? k8[e] : 0.0;
          }
          const int cls = (dx != 0) + 2 * (dy != 0) + 4 * (dz != 0);
          double fc = f8[0];
#pragma unroll
          for (int c = 1; c < NB; ++c) fc = cls == c ? f8[c] : fc;
          m = fc * ks * (sj * sc[i]);
        }
        if (!(fabs(av - m) <= tol)) bad += 1.0;
        if (!(fabs(av - m) <= tolr)) badr += 1.0;
      }
    }
  }
  count2(bad, seen, cnt, 0, sh);
  count2(want, badr, cnt, 2, sh);
}

// w = Â x in the cell form, rows [r_lo, r_lo + nrows) of the padded numbering, one row per
// thread in tiles of kBlock consecutive rows; partial x·w per block.  Per tile the block
// stages in LDS, for each z-plane offset, the window of t·x its rows' neighbours span
// (kBlock + 2 px + 2 entries) and the windows of the coefficients of the cells around them
// (kBlock + px + 1 entries per cell layer), so every row reads its 27 (9) neighbours and 8 (4)
// cells from LDS.  The windows are one index range of E entries, at most NE per thread; a
// thread issues all its loads of a tile before it stores any of them, and K¹ is read from LDS
// (so that the row arithmetic keeps few registers and many waves stay resident to hide the
// loads' latency).  MAP 1 (gridDim.x a multiple of 8): the blocks of one XCD (b ≡ x mod 8)
// sweep one contiguous eighth of the tiles in order, so the window of the plane below was
// loaded into that XCD's
// L2 shortly before.  Fixed summation order (deterministic).
template <typename X, int DIM, int MAP, int NE>
__global__ __launch_bounds__(kBlock) void k_cell_spmv(int64_t r_lo, int64_t nrows, int64_t npad, int32_t px,
                                                      int32_t pxy, const X* __restrict__ kc, const X* __restrict__ t,
                                                      const X* __restrict__ x, X* __restrict__ y,
                                                      double* __restrict__ part, const Scal* __restrict__ s,
                                                      CellKrefT<X> kr) {
  if (s->done || s->fresh) return;
  extern __shared__ __align__(16) unsigned char cell_lds[];
  __shared__ double sh[kBlock / 64];
  constexpr int NB = 1 << DIM;
  constexpr int NZW = DIM == 3 ? 3 : 1;    // t·x windows: planes z − 1, z, z + 1
  constexpr int NKW = DIM == 3 ? 2 : 1;    // coefficient windows: cell layers z, z − 1
  X* zw = reinterpret_cast<X*>(cell_lds);
  const int W = kBlock + 2 * px + 2, WK = kBlock + px + 1;
  const int EZ = NZW * W, E = EZ + NKW * WK;   // E ≤ NE kBlock (checked by the host)
  X* kw = zw + EZ;
  // offsets from the tile's first row of this thread's window entries (tile independent)
  int32_t rel[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    const int idx = threadIdx.x + k * kBlock;
    if (idx < EZ) {
      const int w = idx / W;
      rel[k] = (DIM == 3 ? (w - 1) * pxy : 0) - px - 1 + (idx - w * W);
    } else {
      const int i2 = idx - EZ, w = i2 / WK;
      rel[k] = -w * pxy - px - 1 + (i2 - w * WK);
    }
  }
  const int64_t ntiles = (nrows + kBlock - 1) / kBlock;
  int64_t t0, t1, tstep;
  if (MAP == 1) {
    const int64_t xcd = blockIdx.x % 8;
    t0 = xcd * ntiles / 8 + blockIdx.x / 8;
    t1 = (xcd + 1) * ntiles / 8;
    tstep = gridDim.x / 8;
  } else {
    t0 = blockIdx.x;
    t1 = ntiles;
    tstep = gridDim.x;
  }
  double d = 0.0;
  // rows and byte offsets are 32-bit (8 npad < 2^32: checked by the host), loads take a
  // scalar base and a 32-bit offset (no 64-bit address per load in VGPRs)
  const uint32_t np32 = (uint32_t)npad, rend32 = (uint32_t)(r_lo + nrows);
  for (int64_t tile = t0; tile < t1; tile += tstep) {
    const uint32_t R0 = (uint32_t)(r_lo + tile * kBlock);   // every window starts at or after row 0
    X va[NE], vb[NE];
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      const int idx = threadIdx.x + k * kBlock;
      const uint32_t q = R0 + (uint32_t)rel[k];
      va[k] = X(0);
      vb[k] = X(1);
      if (idx < E && q < np32) {
        if (idx < EZ) {
          va[k] = at_byte(t, q * (uint32_t)sizeof(X));
          vb[k] = at_byte(x, q * (uint32_t)sizeof(X));
        } else {
          va[k] = at_byte(kc, q * (uint32_t)sizeof(X));
        }
      }
    }
    const uint32_t P = R0 + threadIdx.x;
    const X tc = P < rend32 ? at_byte(t, P * (uint32_t)sizeof(X)) : X(0);
    const X xc = P < rend32 ? at_byte(x, P * (uint32_t)sizeof(X)) : X(0);
    __syncthreads();   // the previous tile's LDS reads are done
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      const int idx = threadIdx.x + k * kBlock;
      if (idx < E) zw[idx] = va[k] * vb[k];
    }
    __syncthreads();
    if (P < rend32) {
      const int l = threadIdx.x + px + 1;   // row P in the windows
      X zn[DIM == 3 ? 27 : 9];              // t·x of the neighbours, (dz, dy, dx) row-major
#pragma unroll
      for (int dz = 0; dz < (DIM == 3 ? 3 : 1); ++dz)
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
          for (int dx = 0; dx < 3; ++dx) zn[dz * 9 + dy * 3 + dx] = zw[dz * W + l + (dx - 1) + px * (dy - 1)];
      X kn[NB];                             // cell with origin P − e: P is its corner e
#pragma unroll
      for (int e = 0; e < NB; ++e) kn[e] = kw[(e >> 2) * WK + l - (e & 1) - px * ((e >> 1) & 1)];
      const X acc = cell_row<X, DIM>(zn, kn, kr.f);
      const X yv = tc != X(0) ? tc * acc : xc;
      y[P] = yv;
      d += (double)xc * (double)yv;
    }
  }
  const double b = block_sum(d, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
}

// 3D cell form marching in z: work item (j, c) = in-plane tile j (T = RPT kBlock consecutive
// rows of a plane) over the planes [z0, z1) of chunk c.  Step p (p = z0 .. z1 − 1) computes
// the rows of plane p; it brings the t·x window of plane p + 1 (T + 2 px + 2 entries, ZN per
// thread) and the coefficient window of cell layer p (T + px + 1 entries, KN per thread)
// into LDS, and each thread reads from them only what is new to its rows: the 9 in-plane
// neighbours in plane p + 1 and the 4 cells around it in layer p.  The neighbours in planes
// p − 1 and p and the cells of layer p − 1 are still in its registers from the two steps
// before (a thread keeps the same rows of every plane), so a row reads 13 LDS words per
// step instead of 35.  Register planes rotate with the step's index mod 3 (layers mod 2)
// over an unrolled group of six steps, so no values move; the LDS windows are double
// buffered (slot = plane − z0 + 1 mod 2, fixed per unrolled step): one barrier per step
// orders a window's stores after the reads of the window two steps back.  D register
// stages hold the global loads of D steps, issued D planes ahead of their use; an item
// issues its first windows and its first D steps together.  A step's loads go through
// three buffer resources (t, x, coefficients) made from one row, the window's first
// coefficient row (a byte offset advanced one plane per step): a thread's byte offsets
// into them are fixed for the whole launch, rows
// outside the vector read 0 — the padding's value — with no compare or select, and a step
// past the item's last plane gets empty resources (its loads read 0 and move no data).
// Rows are stored through a resource that ends at the plane's end, so a tile's rows past it
// are dropped.  A thread computes RPT rows kBlock apart (lanes read consecutive LDS words: no
// bank conflicts).  Window entries move 16 bytes at a time (one buffer load, one LDS write
// of 16 / sizeof(X) consecutive entries), and the LDS windows are padded to whole such
// chunks per thread, so every thread stores all its chunks; the barriers order LDS only, so
// the loads of later steps stay in flight across them.  Padding rows of the marched planes are computed too (t = 0 there, so
// y = x = 0).  Items are ordered chunk-major; MAP 1: the blocks of one XCD sweep one
// contiguous eighth of the items, so neighbouring tiles (which share their halo rows) run on
// the same XCD at the same time.  The host keeps npad · sizeof(X) below 2³¹ (the resources'
// 32-bit extent).
template <typename X, int MAP, int ZN, int KN, int D, int RPT>
__global__ __launch_bounds__(kBlock) void k_cell_spmv_z(int64_t npad, int32_t px, int32_t pxy, int32_t pz,
                                                        int32_t ntx, int32_t lz, int64_t nitems,
                                                        const X* __restrict__ kc, const X* __restrict__ t,
                                                        const X* __restrict__ x, X* __restrict__ y,
                                                        double* __restrict__ part, const Scal* __restrict__ s,
                                                        CellKrefT<X> kr) {
  static_assert(6 % D == 0, "the unrolled group of six steps holds whole stage cycles");
  if (s->done || s->fresh) return;
  extern __shared__ __align__(16) unsigned char cell_lds[];
  __shared__ double sh[kBlock / 64];
  typedef unsigned int U4 __attribute__((ext_vector_type(4)));
  constexpr int VW = 16 / (int)sizeof(X);    // window entries per 16-byte load
  constexpr int ZC = (ZN + VW - 1) / VW, KC = (KN + VW - 1) / VW;   // 16-byte loads per thread
  constexpr int T = RPT * kBlock;            // rows of a tile: RPT per thread, kBlock apart
  constexpr int WZ = ZC * VW * kBlock, WKP = KC * VW * kBlock;   // padded window lengths
  const int W = T + 2 * px + 2, WK = T + px + 1;      // W ≤ ZN kBlock ≤ WZ, WK ≤ WKP (host)
  X* zw = reinterpret_cast<X*>(cell_lds);   // [2][WZ]
  X* kw = zw + 2 * WZ;                       // [2][WKP]
  const int32_t np = (int32_t)npad;
  // Window entries are loaded VW at a time: chunk c of this thread holds the entries
  // (c kBlock + tid) VW .. + VW − 1.  Byte offsets from a step's base row B = (p − 1) pxy + o0
  // − px − 1 (the first row of the coefficient window of layer p − 1): the t·x window of plane
  // p starts pxy rows after it, this thread's rows of plane p − 1 px + 1 rows after it (a
  // chunk that starts past its window: an offset no resource reaches; one that reaches past
  // it loads rows no computed row reads)
  uint32_t oz[ZC], ok[KC], orw[RPT];
#pragma unroll
  for (int c = 0; c < ZC; ++c) {
    const int idx = (c * kBlock + (int)threadIdx.x) * VW;
    oz[c] = idx < W ? (uint32_t)(pxy + idx) * (uint32_t)sizeof(X) : 0x80000000u;
  }
#pragma unroll
  for (int c = 0; c < KC; ++c) {
    const int idx = (c * kBlock + (int)threadIdx.x) * VW;
    ok[c] = idx < WK ? (uint32_t)idx * (uint32_t)sizeof(X) : 0x80000000u;
  }
#pragma unroll
  for (int r = 0; r < RPT; ++r) orw[r] = (uint32_t)(px + 1 + r * kBlock + threadIdx.x) * (uint32_t)sizeof(X);
  int64_t i0, i1, istep;
  if (MAP == 1) {
    const int64_t xcd = blockIdx.x % 8;
    i0 = xcd * nitems / 8 + blockIdx.x / 8;
    i1 = (xcd + 1) * nitems / 8;
    istep = gridDim.x / 8;
  } else {
    i0 = blockIdx.x;
    i1 = nitems;
    istep = gridDim.x;
  }
  U4 vt[D][ZC], vx[D][ZC], vk[D][KC];
  X tn[D][RPT], xn[D][RPT];
  // The item's steps fetch planes p = z0 + 1, z0 + 2, … in order: fetch(u) brings into stage u
  // the t·x window of plane p, the coefficients of layer p − 1 and this thread's t, x of plane
  // p − 1, from the base row B = (p − 1) pxy + o0 − px − 1 (fB: its byte offset, advanced one
  // plane per fetch; 0 ≤ B < npad for 2 ≤ p ≤ pz − 1); nothing is read for p > z1 (the last
  // plane an item's rows read).
  uint32_t fB = 0;
  int32_t fp = 0, fz1 = 0;
  const uint32_t plane_bytes = (uint32_t)pxy * (uint32_t)sizeof(X), end_bytes = (uint32_t)np * (uint32_t)sizeof(X);
  auto fetch = [&](int u) {
    const bool on = fp <= fz1;
    const uint32_t off = on ? fB : 0u, ext = on ? end_bytes - fB : 0u;
    const __amdgpu_buffer_rsrc_t rt = buf_rsrc(reinterpret_cast<const char*>(t) + off, ext);
    const __amdgpu_buffer_rsrc_t rx = buf_rsrc(reinterpret_cast<const char*>(x) + off, ext);
    const __amdgpu_buffer_rsrc_t rk = buf_rsrc(reinterpret_cast<const char*>(kc) + off, ext);
#pragma unroll
    for (int c = 0; c < ZC; ++c) {
      vt[u][c] = __builtin_amdgcn_raw_buffer_load_b128(rt, (int)oz[c], 0, 0);
      vx[u][c] = __builtin_amdgcn_raw_buffer_load_b128(rx, (int)oz[c], 0, 0);
    }
#pragma unroll
    for (int c = 0; c < KC; ++c) vk[u][c] = __builtin_amdgcn_raw_buffer_load_b128(rk, (int)ok[c], 0, 0);
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      tn[u][r] = buf_load<X>(rt, orw[r]);
      xn[u][r] = buf_load<X>(rx, orw[r]);
    }
    fB += plane_bytes;
    ++fp;
  };
  // the products t·x of a chunk into LDS, and a chunk of coefficients
  auto put_tx = [&](X* zs, int c, const U4& tv, const U4& xv) {
    struct A { X v[VW]; };
    const A ta = __builtin_bit_cast(A, tv), xa = __builtin_bit_cast(A, xv);
    A pa;
#pragma unroll
    for (int j = 0; j < VW; ++j) pa.v[j] = ta.v[j] * xa.v[j];
    *reinterpret_cast<U4*>(zs + (c * kBlock + (int)threadIdx.x) * VW) = __builtin_bit_cast(U4, pa);
  };
  auto put_k = [&](X* ks, int c, const U4& kv) {
    *reinterpret_cast<U4*>(ks + (c * kBlock + (int)threadIdx.x) * VW) = kv;
  };
  // a row's 9 in-plane neighbours (dy, dx row-major) from a t·x window; the 4 cells around it
  // (origin P − e, e = ex + 2 ey) from a coefficient window; l = the row's index in the windows
  auto plane9 = [&](X* v, const X* zs, int l) {
#pragma unroll
    for (int dy = 0; dy < 3; ++dy)
#pragma unroll
      for (int dx = 0; dx < 3; ++dx) v[dy * 3 + dx] = zs[l + (dx - 1) + px * (dy - 1)];
  };
  auto layer4 = [&](X* v, const X* ks, int l) {
#pragma unroll
    for (int e = 0; e < 4; ++e) v[e] = ks[l - (e & 1) - px * ((e >> 1) & 1)];
  };
  double d = 0.0;
  for (int64_t item = i0; item < i1; item += istep) {
    const int32_t c = (int32_t)(item / ntx), j = (int32_t)(item - (int64_t)c * ntx);
    const int32_t z0 = 1 + c * lz, z1 = min(z0 + lz, pz - 1), nl = z1 - z0;
    const int32_t o0 = j * T;
    // the item's first windows (t·x of planes z0 − 1 and z0, coefficients of layer z0 − 1)
    // and its first D steps are issued together, so the item waits on one load latency
    // before its first plane.  Plane z0 − 1 may be plane 0, whose window can start before
    // row 0: every entry a computed row reads there is padding, and the whole window reads 0.
    U4 wt0[ZC], wx0[ZC], wt1[ZC], wx1[ZC], wk1[KC];
    {
      const int32_t b0 = (z0 - 1) * pxy + o0 - px - 1;
      const __amdgpu_buffer_rsrc_t rt0 = rows_from(t, b0, np), rx0 = rows_from(x, b0, np);
      const __amdgpu_buffer_rsrc_t rt1 = rows_from(t, b0 + pxy, np), rx1 = rows_from(x, b0 + pxy, np);
      const __amdgpu_buffer_rsrc_t rk1 = rows_from(kc, b0, np);
#pragma unroll
      for (int c = 0; c < ZC; ++c) {
        const int idx = (c * kBlock + (int)threadIdx.x) * VW;   // from the window's first row
        const int o = idx < W ? idx * (int)sizeof(X) : (int)0x80000000u;
        wt0[c] = __builtin_amdgcn_raw_buffer_load_b128(rt0, o, 0, 0);
        wx0[c] = __builtin_amdgcn_raw_buffer_load_b128(rx0, o, 0, 0);
        wt1[c] = __builtin_amdgcn_raw_buffer_load_b128(rt1, o, 0, 0);
        wx1[c] = __builtin_amdgcn_raw_buffer_load_b128(rx1, o, 0, 0);
      }
#pragma unroll
      for (int c = 0; c < KC; ++c) wk1[c] = __builtin_amdgcn_raw_buffer_load_b128(rk1, (int)ok[c], 0, 0);
    }
    fp = z0 + 1;
    fz1 = z1;
    fB = (uint32_t)(z0 * pxy + o0 - px - 1) * (uint32_t)sizeof(X);
#pragma unroll
    for (int u = 0; u < D; ++u) fetch(u);
    lds_barrier();     // the previous item's LDS reads are done
#pragma unroll
    for (int c = 0; c < ZC; ++c) {
      put_tx(zw, c, wt0[c], wx0[c]);        // slot 0: plane z0 − 1
      put_tx(zw + WZ, c, wt1[c], wx1[c]);   // slot 1: plane z0
    }
#pragma unroll
    for (int c = 0; c < KC; ++c) put_k(kw, c, wk1[c]);   // slot 0: layer z0 − 1
    lds_barrier();
    // register planes: slot (p − z0 + 1) mod 3 holds plane p, layer slot (q − z0 + 1) mod 2 layer q
    X zr[3][RPT][9], kq[2][RPT][4];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int l = r * kBlock + threadIdx.x + px + 1;   // the row in the windows
      plane9(zr[0][r], zw, l);
      plane9(zr[1][r], zw + WZ, l);
      layer4(kq[0][r], kw, l);
    }
    bool col[RPT];     // the row lies in the plane (the last tile of a plane is cut)
#pragma unroll
    for (int r = 0; r < RPT; ++r) col[r] = o0 + r * kBlock + (int32_t)threadIdx.x < pxy;
    const uint32_t yext = (uint32_t)(pxy - o0) * (uint32_t)sizeof(X);
    lds_barrier();     // the first step's store overwrites the window of plane z0 − 1
    for (int32_t sb = 0; sb < nl; sb += 6) {
#pragma unroll
      for (int u = 0; u < 6; ++u) {
        if (sb + u >= nl) break;
        const int su = u % D;
        const int32_t zp = z0 + sb + u;
        X tc[RPT], xc[RPT];   // rows of plane zp
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          tc[r] = tn[su][r];
          xc[r] = xn[su][r];
        }
        X* zs = zw + (u & 1) * WZ;              // plane zp + 1: slot (sb + u + 2) mod 2
        X* ks = kw + ((u + 1) & 1) * WKP;       // layer zp: slot (sb + u + 1) mod 2
#pragma unroll
        for (int c = 0; c < ZC; ++c) put_tx(zs, c, vt[su][c], vx[su][c]);
#pragma unroll
        for (int c = 0; c < KC; ++c) put_k(ks, c, vk[su][c]);
        lds_barrier();
        fetch(su);   // plane zp + 1 + D
        const __amdgpu_buffer_rsrc_t ry = buf_rsrc(y + (zp * pxy + o0), yext);
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          const int l = r * kBlock + threadIdx.x + px + 1;
          plane9(zr[(u + 2) % 3][r], zs, l);
          layer4(kq[(u + 1) % 2][r], ks, l);
          X zn[27], kn[8];   // (dz, dy, dx) row-major; cell with origin P − e: P is its corner e
#pragma unroll
          for (int i = 0; i < 9; ++i) {
            zn[i] = zr[u % 3][r][i];
            zn[9 + i] = zr[(u + 1) % 3][r][i];
            zn[18 + i] = zr[(u + 2) % 3][r][i];
          }
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            kn[e] = kq[(u + 1) % 2][r][e];   // layer zp
            kn[4 + e] = kq[u % 2][r][e];     // layer zp − 1
          }
          const X acc = cell_row<X, 3>(zn, kn, kr.f);
          const X yv = tc[r] != X(0) ? tc[r] * acc : xc[r];
          buf_store<X>(yv, ry, (uint32_t)(r * kBlock + threadIdx.x) * (uint32_t)sizeof(X));
          d += col[r] ? (double)(xc[r] * yv) : 0.0;
        }
      }
    }
  }
  const double b = block_sum(d, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
}

// s->rr (which = 0) or s->bnorm2 (which = 1) = Σ part, fixed order
__global__ __launch_bounds__(kWide) void k_sum_to(int m, const double* __restrict__ part, Scal* __restrict__ s,
                                                  int which) {
  __shared__ double sh[kWide / 64];
  const double t = wide_sum(part, m, 1, sh);
  if (threadIdx.x == 0) {
    if (which == 0) s->rr = t; else s->bnorm2 = t;
  }
}

__global__ __launch_bounds__(kBlock) void k_sumsq(int64_t n, const double* __restrict__ v, double* __restrict__ part) {
  __shared__ double sh[kBlock / 64];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    acc += v[i] * v[i];
  const double b = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
}

// S = diag(K)^-1/2 (Jacobi; rows with a zero diagonal keep 1), or 1 without Jacobi.
__global__ void k_scale(int64_t n, const int32_t* __restrict__ diag, const double* __restrict__ nzval, int jacobi,
                        double* __restrict__ sc) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const int32_t p = diag[j];
    const double d = (p >= 0) ? nzval[p] : 0.0;
    sc[j] = (jacobi && d != 0.0) ? 1.0 / sqrt(fabs(d)) : 1.0;
  }
}

__global__ void k_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// Hand-over, step 1: fold the recurrence's iterate into the FP64 u: u += S (x̂ + pend p),
// x̂ = 0 (pend: the last step's α when that step left its x̂ move pending, else 0).
// `map` (lattice operator) gives the recurrence's index of each dof; null: the same index.
template <typename X>
__global__ void k_fold(int64_t n, X* __restrict__ x, const X* __restrict__ p, const Scal* __restrict__ s,
                       const double* __restrict__ sc, double* __restrict__ u, const int32_t* __restrict__ map) {
  const double pend = s->pend;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = map ? (int64_t)map[i] : i;
    const double xk = pend != 0.0 ? (double)x[k] + pend * (double)p[k] : (double)x[k];
    u[i] += sc[i] * xk;
    x[k] = X(0);
  }
}

// --------------------------------------------------------------------------
// Chronopoulos–Gear CG on Â x̂ = b̂ (one reduction phase per iteration).
//   state: x̂, r̂, w = Â r̂, p, s = Â p
//   step i (after the SpMV w = Â r̂, partial δ = r̂·w):
//     γ = r̂·r̂ (partials of the previous step), β = γ/γ_prev,
//     α = γ / (δ − β γ / α_prev), p = r̂ + β p, s = w + β s, x̂ += α p, r̂ −= α s,
//     partials of the next γ.
//   x̂ moves every other step: a step leaves α p pending (x̂ neither read nor written) and the
//   next adds α_prev p_prev + α p, p_prev being read anyway to form p; a hand-over folds the
//   pending term into u (8 instead of 9 vector passes per step on average).
// An iteration is three kernels: SpMV, k_cg_scal (one wave reduces the partials in a fixed
// order and writes the scalars), k_cg_move (the vectors); for small systems two, the vector
// kernel's blocks each doing the scalar step themselves (k_cg_move_scal).  The scalar records
// and the vector partials are double-buffered (read one, write the other).
// --------------------------------------------------------------------------
constexpr int kCgBlocks = 4096;  // most blocks of k_cg_move (its partial count)

// Vector elements per thread per pass of k_cg_move (tuning cg_elems; measured default 8).
static int cg_elems(const fem_ctx* c) { return c->tune.cg_elems ? c->tune.cg_elems : 8; }

// Blocks of k_cg_move (and of the reseed kernels) for n unknowns: one pass over the vectors
// when possible.
static int cg_blocks(const fem_ctx* c, int64_t n) {
  const int64_t per = (int64_t)kBlock * cg_elems(c);
  return (int)std::max<int64_t>(1, std::min<int64_t>(kCgBlocks, (n + per - 1) / per));
}


// Σ a[0..m) and Σ b[0..mb) by one wave in a fixed order: lane l takes i ≡ l (mod 64) into
// eight accumulators (i / 64 mod 8), which are added pairwise, then a butterfly across lanes.
__device__ __forceinline__ void wave_sum2(const double* __restrict__ a, int m, const double* __restrict__ b,
                                          int mb, double& sa, double& sb) {
  const int lane = threadIdx.x & 63;
  double ua[8] = {}, ub[8] = {};
  const int mm = m > mb ? m : mb;
  for (int i0 = 0; i0 < mm; i0 += 8 * 64) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int i = i0 + 64 * k + lane;
      ua[k] += i < m ? a[i] : 0.0;
      ub[k] += i < mb ? b[i] : 0.0;
    }
  }
  sa = ((ua[0] + ua[1]) + (ua[2] + ua[3])) + ((ua[4] + ua[5]) + (ua[6] + ua[7]));
  sb = ((ub[0] + ub[1]) + (ub[2] + ub[3])) + ((ub[4] + ub[5]) + (ub[6] + ub[7]));
  for (int o = 32; o > 0; o >>= 1) {
    sa += __shfl_xor(sa, o, 64);
    sb += __shfl_xor(sb, o, 64);
  }
}

// The step after the SpMV: k_cg_scal reduces the partials once (one wave, fixed order) and
// writes the step's scalars; k_cg_move then updates the vectors with many small blocks (E
// elements per thread), which lets later blocks load while earlier ones store.  (One kernel
// whose every block reduced the partials itself measured slower at 128³: mixed loop 34.9 vs
// 29.7 ms.)
// The scalar step: the record after this step from the record before it and the partials of
// γ (vector step) and δ (SpMV), reduced by the calling wave (threads 0-63) in a fixed order;
// thread 0 writes it to *out.
__device__ __forceinline__ void cg_scalars(int nin, const double* __restrict__ part_in, int na,
                                           const double* __restrict__ part_s, const Scal& S, Scal* out) {
  if (S.done) {
    if (threadIdx.x == 0) *out = S;
    return;
  }
  if (S.fresh) {
    if (threadIdx.x == 0) {
      Scal o = S;
      o.sfresh = 1;
      o.fresh = 0;
      o.rzp = S.rz;
      o.beta = 0.0;
      o.iters = S.iters + 1;
      o.xa0 = 0.0;       // x̂ += α p of the reseed
      o.xa1 = S.alpha;
      o.pend = 0.0;
      *out = o;
    }
    return;
  }
  double gam, del;
  wave_sum2(part_in, nin, part_s, na, gam, del);
  if (threadIdx.x == 0) {
    Scal o = S;
    const double beta = gam / S.rzp;
    const double den = del - beta * gam / S.alpha;
    const bool stop = gam <= S.gstop;
    const bool brk = !(S.rzp > 0.0) || !(S.alpha != 0.0) || !(den > 0.0);
    if (stop || brk) {   // hand over to the host: x̂, r̂ of this step, p and s of the last one
      o.done = 1;
      o.rz = gam;
      if (!stop) o.restart = 1;
    } else {
      o.sfresh = 0;
      o.fresh = 0;
      o.alpha = gam / den;
      o.beta = beta;
      o.rzp = gam;
      o.rz = gam;
      o.iters = S.iters + 1;
      // x̂ moves every other step, by two directions at once (p_old is still read to form p)
      o.xa0 = S.pend;
      o.xa1 = S.pend != 0.0 ? o.alpha : 0.0;
      o.pend = S.pend != 0.0 ? 0.0 : o.alpha;
    }
    *out = o;
  }
}

__global__ __launch_bounds__(64) void k_cg_scal(int nin, const double* __restrict__ part_in, int na,
                                               const double* __restrict__ part_s, const Scal* __restrict__ sin,
                                               Scal* __restrict__ sout) {
  const Scal S = *sin;
  cg_scalars(nin, part_in, na, part_s, S, sout);
}

// The vector step of the record S: p = r̂ + βp, s = w + βs, x̂ moved (every other step), r̂ −=
// αs, the block's partial of the next γ.  With PRE, p, s and r̂ of the block's first pass
// were loaded before S was known (pv0, sw0, rv0).
template <typename X, int E, bool PRE = false>
__device__ __forceinline__ void cg_move(const Scal& S, int64_t n, X* __restrict__ x, X* __restrict__ r,
                                        const X* __restrict__ w, X* __restrict__ p, X* __restrict__ sv,
                                        double* __restrict__ part_out, double* sh, const X* pv0 = nullptr,
                                        const X* sw0 = nullptr, const X* rv0 = nullptr) {
  const bool fresh = S.sfresh != 0;
  const X al = (X)S.alpha, be = (X)S.beta, xa0 = (X)S.xa0, xa1 = (X)S.xa1;
  const bool xs = S.xa0 != 0.0 || S.xa1 != 0.0;   // this step moves x̂
  double ga = 0.0;
  bool first = true;
  for (int64_t base = (int64_t)blockIdx.x * kBlock * E; base < n; base += (int64_t)gridDim.x * kBlock * E) {
    X wv[E], pv[E], sw[E], xv[E], rv[E];
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int64_t i = base + k * kBlock + threadIdx.x;
      const bool in = i < n;
      wv[k] = in && !fresh ? w[i] : X(0);
      if (PRE && first) {
        pv[k] = pv0[k];
        sw[k] = sw0[k];
        rv[k] = rv0[k];
      } else {
        pv[k] = in ? p[i] : X(0);
        sw[k] = in ? sv[i] : X(0);
        rv[k] = in ? r[i] : X(0);
      }
      xv[k] = in && xs ? x[i] : X(0);
    }
    first = false;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int64_t i = base + k * kBlock + threadIdx.x;
      if (i < n) {
        X pi = pv[k], si = sw[k];
        if (!fresh) {
          pi = rv[k] + be * pi;
          si = wv[k] + be * si;
          p[i] = pi;
          sv[i] = si;
        }
        if (xs) x[i] = xv[k] + (xa0 * pv[k] + xa1 * pi);
        const X ri = rv[k] - al * si;
        r[i] = ri;
        ga += (double)ri * (double)ri;
      }
    }
  }
  const double b = block_sum(ga, sh);
  if (threadIdx.x == 0) part_out[blockIdx.x] = b;
}

template <typename X, int E>
__global__ __launch_bounds__(kBlock) void k_cg_move(int64_t n, X* __restrict__ x, X* __restrict__ r,
                                                    const X* __restrict__ w, X* __restrict__ p, X* __restrict__ sv,
                                                    double* __restrict__ part_out, const Scal* __restrict__ sc) {
  __shared__ double sh[kBlock / 64];
  const Scal S = *sc;
  if (S.done) return;
  cg_move<X, E>(S, n, x, r, w, p, sv, part_out, sh);
}

// Systems of at most kCgFuseBlocks vector blocks: the scalar step folded into the vector
// step.  The first wave of every block reduces the partials itself (in k_cg_scal's order, so
// every block holds the same record) and block 0 writes the record for the next step: one
// launch per iteration fewer.  Measured (mixed, cell form): faster at 1,099 blocks (128³:
// 16.9 against 17.4 ms per step), slower at 2,048 (160³: 35.6 against 34.2 ms), where every
// block's sum over ~2,800 partials costs more than the launch it saves.
constexpr int kCgFuseBlocks = 1280;
constexpr int kCgPreBlocks = 512;   // and up to this many, with the vector loads issued first
// PRE: p, s and r̂ of the first pass (needed whatever the record says) are loaded while the
// first wave sums the partials.  It pays where few blocks each move little (up to
// kCgPreBlocks).  At 128³ (1,099 blocks) it measured level with the plain form (16.9 ms per
// step at rtol 1e-8) once the preload was issued after the record's load (96 VGPRs, five
// waves per SIMD: every block resident); issued first it took 108 (four waves per SIMD: 1,024
// resident, 75 blocks in a second round) and the step 18.8 ms.  Preloading in every wave but
// the first (68 VGPRs, the summing wave loading its vectors afterwards) measured 17.6 ms.
template <typename X, int E, bool PRE>
__global__ __launch_bounds__(kBlock) void k_cg_move_scal(int64_t n, X* __restrict__ x, X* __restrict__ r,
                                                         const X* __restrict__ w, X* __restrict__ p,
                                                         X* __restrict__ sv, double* __restrict__ part_out, int nin,
                                                         const double* __restrict__ part_in, int na,
                                                         const double* __restrict__ part_s,
                                                         const Scal* __restrict__ sin, Scal* __restrict__ sout) {
  __shared__ double sh[kBlock / 64];
  __shared__ Scal so;
  X pv0[E], sw0[E], rv0[E];
  auto preload = [&]() {
    const int64_t base0 = (int64_t)blockIdx.x * kBlock * E;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int64_t i = base0 + k * kBlock + threadIdx.x;
      const bool in = i < n;
      pv0[k] = in ? p[i] : X(0);
      sw0[k] = in ? sv[i] : X(0);
      rv0[k] = in ? r[i] : X(0);
    }
  };
  const Scal S0 = *sin;
  if (PRE) preload();
  if (threadIdx.x < 64) cg_scalars(nin, part_in, na, part_s, S0, &so);
  __syncthreads();
  const Scal S = so;
  if (blockIdx.x == 0 && threadIdx.x == 0) *sout = S;
  if (S.done) return;
  cg_move<X, E, PRE>(S, n, x, r, w, p, sv, part_out, sh, pv0, sw0, rv0);
}

// Reseed, step 1: r̂ = S res (rounded to X); partials r̂·r̂ and r̂·p.
// `map` as in k_fold (the recurrence's entries outside the map stay zero).
template <typename X>
__global__ __launch_bounds__(kBlock) void k_cg_reseed(int64_t n, const double* __restrict__ res,
                                                      const double* __restrict__ sc, X* __restrict__ r,
                                                      const X* __restrict__ p, double* __restrict__ part,
                                                      const int32_t* __restrict__ map) {
  __shared__ double sh[kBlock / 64];
  double rr = 0.0, rp = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const int64_t k = map ? (int64_t)map[i] : i;
    const X ri = (X)(sc[i] * res[i]);
    r[k] = ri;
    rr += (double)ri * (double)ri;
    rp += (double)ri * (double)p[k];
  }
  double b = block_sum(rr, sh);
  if (threadIdx.x == 0) part[2 * blockIdx.x] = b;
  b = block_sum(rp, sh);
  if (threadIdx.x == 0) part[2 * blockIdx.x + 1] = b;
}

// Reseed, step 2 (scalars): γ = r̂·r̂, γ' = (r̂·p)/γ, β = γ/γ_prev (0: restart p = r̂).
// s->rr holds the true residual ||f − K u||² when this runs.  The next hand-over is when γ
// has fallen by hold_factor2, or to where the true residual would meet tol2 if ||r||²/γ
// stayed as it is now (the host then checks the true residual itself).
__global__ __launch_bounds__(kBlock) void k_cg_reseed_scal(int m, const double* __restrict__ part,
                                                           Scal* __restrict__ s, double hold_factor2, double tol2,
                                                           int restart) {
  __shared__ double sh[kBlock / 64];
  const double ru = sum_parts(part, m, 2, sh);
  const double rp = sum_parts(part + 1, m, 2, sh);
  if (threadIdx.x == 0) {
    const bool fresh = restart || s->restart || !(s->rzp > 0.0) || !(ru > 0.0);
    s->gam = ru > 0.0 ? rp / ru : 0.0;
    s->beta = fresh ? 0.0 : ru / s->rzp;
    s->rz = ru;
    s->rz0 = ru;
    const double conv = s->rr > 0.0 ? tol2 / s->rr : 0.0;
    s->gstop = ru * (hold_factor2 > conv ? hold_factor2 : conv);
    s->done = 0;
    s->restart = 0;
    s->fresh = 0;
    s->pend = 0.0;     // folded into u
  }
}

// Reseed, step 3 (after w = Â r̂): p = r̂ + β (p − γ' r̂), s = w + β (s − γ' w): the last
// direction made orthogonal to the new residual, then extended (p = r̂, s = w on a
// restart); partial p·s.
template <typename X>
__global__ __launch_bounds__(kBlock) void k_cg_pfix(int64_t n, const X* __restrict__ r, const X* __restrict__ w,
                                                    X* __restrict__ p, X* __restrict__ sv, const Scal* __restrict__ s,
                                                    double* __restrict__ part) {
  __shared__ double sh[kBlock / 64];
  const X be = (X)s->beta, ga = (X)s->gam;
  const bool keep = s->beta != 0.0;
  double ps = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const X ri = r[i], wi = w[i];
    const X pi = keep ? ri + be * (p[i] - ga * ri) : ri;
    const X si = keep ? wi + be * (sv[i] - ga * wi) : wi;
    p[i] = pi;
    sv[i] = si;
    ps += (double)pi * (double)si;
  }
  const double b = block_sum(ps, sh);
  if (threadIdx.x == 0) part[2 * blockIdx.x] = b;
  if (threadIdx.x == 0) part[2 * blockIdx.x + 1] = 0.0;
}

// Reseed, step 4 (scalars): α = γ / p·s; the next step only moves x̂ and r̂.
__global__ __launch_bounds__(kBlock) void k_cg_alpha(int m, const double* __restrict__ part, Scal* __restrict__ s) {
  __shared__ double sh[kBlock / 64];
  const double ps = sum_parts(part, m, 2, sh);
  if (threadIdx.x == 0) {
    s->pq = ps;
    const bool ok = ps > 0.0;
    s->alpha = ok ? s->rz / ps : 0.0;
    s->fresh = ok ? 1 : 0;
    if (!ok) {   // p vanished (r̂ already zero): nothing to step
      s->done = 1;
      s->restart = 1;
    }
  }
}

// After a fold: the pending x̂ term is in u now (the reseed clears it too; a solve that stops
// after a hand-over folds once more).
__global__ void k_cg_unpend(Scal* __restrict__ s) { s->pend = 0.0; }

// Both scalar records cleared; record 0 starts from ||f||² as its (true) residual.
__global__ void k_cg_init_scal(Scal* __restrict__ s, double rr0) {
  Scal z{};
  s[1] = z;
  z.rr = rr0;
  s[0] = z;
}

__global__ void k_zero32(int64_t n, float* __restrict__ a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0.0f;
}

__global__ void k_copy(int64_t n, const double* __restrict__ a, double* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

__global__ void k_zero(int64_t n, double* __restrict__ a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0.0;
}

template <typename X>
static void zero_vec(fem_ctx* c, X* a, int64_t n) {
  if constexpr (sizeof(X) == 4) k_zero32<<<kMaxVecBlocks, kBlock, 0, c->stream>>>(n, a);
  else k_zero<<<kMaxVecBlocks, kBlock, 0, c->stream>>>(n, a);
}

// --------------------------------------------------------------------------
// host driver
// --------------------------------------------------------------------------
static SpmvShape spmv_shape(const fem_ctx* c, int cap, int lanes) {
  SpmvShape s;
  const int longest = std::max(1, c->max_row_nnz);
  s.cap = cap;
  s.L = lanes;
  s.R = std::max(1, std::min(kBlock / lanes, cap / longest));
  s.blocks = (int)((c->n + s.R - 1) / s.R);
  return s;
}

template <typename V, typename X, int MODE, int CAP, int L>
static void launch_spmv_t(fem_ctx* c, const SpmvShape& sh, const V* val, const X* x, const double* f, X* y,
                          double* part, const Scal* s, const uint16_t* roff, const int32_t* bbase) {
  if (roff)
    k_spmv<V, X, MODE, CAP, L, true><<<sh.blocks, kBlock, 0, c->stream>>>(c->n, sh.R, c->d_colptr, nullptr, roff,
                                                                          bbase, val, x, f, y, part, s);
  else
    k_spmv<V, X, MODE, CAP, L, false><<<sh.blocks, kBlock, 0, c->stream>>>(c->n, sh.R, c->d_colptr, c->d_rowval,
                                                                           nullptr, nullptr, val, x, f, y, part, s);
}

template <typename V, typename X, int MODE>
static void launch_spmv(fem_ctx* c, const SpmvShape& sh, const V* val, const X* x, const double* f, X* y,
                        double* part, const Scal* s, const uint16_t* roff = nullptr, const int32_t* bbase = nullptr) {
#define FEM_SPMV_CASE(CAPV, LV)                                                               \
  if (sh.cap == CAPV && sh.L == LV) {                                                        \
    launch_spmv_t<V, X, MODE, CAPV, LV>(c, sh, val, x, f, y, part, s, roff, bbase);          \
    return;                                                                                  \
  }
  FEM_SPMV_CASE(2048, 1) FEM_SPMV_CASE(2048, 2) FEM_SPMV_CASE(2048, 4)
  FEM_SPMV_CASE(4096, 1) FEM_SPMV_CASE(4096, 2) FEM_SPMV_CASE(4096, 4)
#undef FEM_SPMV_CASE
  launch_spmv_t<V, X, MODE, 2048, 1>(c, sh, val, x, f, y, part, s, roff, bbase);
}

static int ensure_work(fem_ctx* c, bool mixed) {
  if (!c->solver) c->solver = new SolverWork();
  SolverWork* w = c->solver;
  const int64_t n = c->n;
  int rc;
  if (w->n != n || !w->res) {
    for (GraphCache& g : w->g64) destroy_graph(g);
    for (GraphCache& g : w->g32) destroy_graph(g);
    w->sh64 = spmv_shape(c, c->tune.spmv_cap64 ? c->tune.spmv_cap64 : 2048, c->tune.spmv_lanes64 ? c->tune.spmv_lanes64 : 2);
    if ((rc = dev_alloc(c, &w->sc, n)) || (rc = dev_alloc(c, &w->res, n)) ||
        (rc = dev_alloc(c, &w->part_s, std::max(w->sh64.blocks, kMaxVecBlocks))) ||
        (rc = dev_alloc(c, &w->part_v, 4 * std::max(kMaxVecBlocks, kCgBlocks))) || (rc = dev_alloc(c, &w->s, 2)))
      return rc;
    if (!w->hscal) FEM_HIP(c, hipHostMalloc((void**)&w->hscal, sizeof(Scal), hipHostMallocDefault));
    w->n = n;
  }
  (void)mixed;
  return FEM_OK;
}

// The five recurrence vectors in X precision, at least nv long (the lattice operator's
// padded numbering is longer than n).
template <typename X>
static int ensure_vecs(fem_ctx* c, SolverWork* w, int64_t nv) {
  int64_t& cap = sizeof(X) == 4 ? w->cap32 : w->cap64;
  if (cap >= nv) return FEM_OK;
  X** v[5];
  if constexpr (sizeof(X) == 4) {
    X** t[5] = {&w->x32, &w->r32, &w->wv32, &w->p32, &w->sv32};
    std::memcpy(v, t, sizeof v);
  } else {
    X** t[5] = {&w->x, &w->r, &w->wv, &w->p, &w->sv};
    std::memcpy(v, t, sizeof v);
  }
  cap = 0;
  for (X** a : v) {
    const int rc = dev_alloc(c, a, nv);
    if (rc) return rc;
  }
  cap = nv;
  return FEM_OK;
}

// Lattice numbering and padded geometry (once per pattern): find_lattice on the host and the
// map to padded rows.  state −1 when the dofs form no lattice; the first plane fill also sets
// it when a pattern entry couples dofs that are not lattice neighbours (the sliced copy is
// used instead).
static int ensure_lattice(fem_ctx* c, SolverWork* w) {
  LatOp& l = w->lat;
  if (l.state != 0) return FEM_OK;
  l.state = -1;
  Lattice L;
  if (!find_lattice(c, &L)) return FEM_OK;
  const int64_t px = L.nx + 2, py = L.ny + 2, pz = L.dim == 3 ? L.nz + 2 : 1;
  const int64_t pxy = px * py, npad = pxy * pz;
  if (8 * npad >= (int64_t)UINT32_MAX) return FEM_OK;   // 32-bit byte offsets within a plane / vector
  std::vector<int32_t> map(c->n), dofp(npad, -1);
  for (int64_t i = 0; i < c->n; ++i) {
    const int64_t q = L.pos[i];
    const int64_t ix = q % L.nx, iy = (q / L.nx) % L.ny, iz = q / (L.nx * L.ny);
    map[i] = (int32_t)((ix + 1) + px * ((iy + 1) + py * (L.dim == 3 ? iz + 1 : 0)));
    dofp[map[i]] = (int32_t)i;
  }
  int rc;
  if ((rc = dev_alloc(c, &l.map, c->n)) || (rc = dev_alloc(c, &l.dofp, npad)) || (rc = dev_alloc(c, &l.cnt, 4)))
    return rc;
  FEM_HIP(c, hipMemcpyAsync(l.map, map.data(), 4 * c->n, hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipMemcpyAsync(l.dofp, dofp.data(), 4 * npad, hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  l.px = px; l.py = py; l.pxy = pxy; l.npad = npad;
  l.ns = L.dim == 3 ? 14 : 5;
  for (int k = 0; k < 14; ++k) l.off.o[k] = 0;
  for (int k = 1; k < l.ns; ++k) {
    const int code = 13 + k;
    const int dx = code % 3 - 1, dy = (code / 3) % 3 - 1, dz = code / 9 - 1;
    l.off.o[k] = (int32_t)(dx + px * dy + pxy * dz);
  }
  const int64_t omax = 1 + px + (L.dim == 3 ? pxy : 0);   // largest offset
  l.r_lo = omax;
  l.nrows = npad - 2 * omax;
  // A multiple of 8 (one share per XCD).  Measured at 128³ (FP32, in the solve): 512 blocks
  // 28.5 µs; 384 34.5, 448 31.1, 536 36.4, 576 35.6, 768 32.8, 2048 36.3 µs.
  int64_t gr = c->tune.lattice_blocks ? c->tune.lattice_blocks : 512;
  gr = std::min<int64_t>(gr, (l.nrows + kBlock - 1) / kBlock);
  if (gr >= 8) gr -= gr % 8;
  l.grid = (int)std::max<int64_t>(1, gr);
  l.hmap = std::move(map);
  l.state = 1;
  return FEM_OK;
}

// Planes of the scaled K in V precision (refilled when nzval or the scaling changed), with the
// symmetry check; *sym = false when it fails (the caller then uses the sliced copy).
template <typename V>
static int ensure_lat_vals(fem_ctx* c, SolverWork* w, int jacobi, bool* sym) {
  LatOp& l = w->lat;
  V*& val = sizeof(V) == 4 ? reinterpret_cast<V*&>(l.val32) : reinterpret_cast<V*&>(l.val64);
  V*& mir = sizeof(V) == 4 ? reinterpret_cast<V*&>(l.mir32) : reinterpret_cast<V*&>(l.mir64);
  uint64_t& ver = sizeof(V) == 4 ? l.ver32 : l.ver64;
  int& jac = sizeof(V) == 4 ? l.jac32 : l.jac64;
  bool& ok = sizeof(V) == 4 ? l.sym32 : l.sym64;
  int rc;
  if (!val) {   // positions never written (padding, missing neighbours) stay zero
    if ((rc = dev_alloc(c, &val, (size_t)l.ns * l.npad)) || (rc = dev_alloc(c, &mir, (size_t)(l.ns - 1) * l.npad)))
      return rc;
    FEM_HIP(c, hipMemsetAsync(val, 0, sizeof(V) * l.ns * l.npad, c->stream));
    FEM_HIP(c, hipMemsetAsync(mir, 0, sizeof(V) * (l.ns - 1) * l.npad, c->stream));
    ver = 0;
  }
  if (ver != c->nzval_version + 1 || jac != jacobi) {
    const LatGeom g{(uint32_t)l.px, (uint32_t)l.pxy, (uint32_t)(1 + l.px + l.pxy), l.ns, l.npad};
    const int64_t rows_per_block = kBlock / kLatLanes;
    const int fb = (int)std::min<int64_t>(16384, (l.nrows + rows_per_block - 1) / rows_per_block);
    FEM_HIP(c, hipMemsetAsync(l.cnt, 0, 4 * sizeof(unsigned long long), c->stream));
    k_lat_fill<V><<<fb, kBlock, 0, c->stream>>>(l.r_lo, l.nrows, l.dofp, c->d_colptr, c->d_rowval, c->d_nzval,
                                                  w->sc, l.map, g, val, mir, l.cnt);
    // FP32 planes: asymmetry below FP32 resolution does not change the operator
    const double tol = sizeof(V) == 4 ? 1.0 / (1 << 22) : 1e-12;
    k_lat_cmp<V><<<kMaxVecBlocks, kBlock, 0, c->stream>>>((int64_t)(l.ns - 1) * l.npad, val + l.npad, mir, tol,
                                                            l.cnt);
    unsigned long long h[4] = {};
    FEM_HIP(c, hipMemcpyAsync(h, l.cnt, sizeof h, hipMemcpyDeviceToHost, c->stream));
    FEM_HIP(c, hipStreamSynchronize(c->stream));
    if (h[1] != 0) {   // not a lattice pattern: never try again for this pattern
      l.state = -1;
      dev_free(&l.val32); dev_free(&l.val64); dev_free(&l.mir32); dev_free(&l.mir64);
      *sym = false;
      return FEM_OK;
    }
    ok = h[3] == 0;    // every forward value equals its mirror (a missing mirror leaves a zero)
    (sizeof(V) == 4 ? l.unit32 : l.unit64) = h[0] == 0;
    ver = c->nzval_version + 1;
    jac = jacobi;
  }
  *sym = ok;
  return FEM_OK;
}

template <typename V, typename X, int NS>
static void launch_lat_ns(fem_ctx* c, const LatOp& l, const V* val, const X* x, X* y, double* part,
                          const Scal* sc) {
  const bool xcd = c->tune.sell_map != 1 && l.grid % 8 == 0;
  const bool unit = sizeof(V) == 4 ? l.unit32 : l.unit64;
#define FEM_LAT(MAP, UNIT) \
  k_lat_spmv<V, X, NS, MAP, UNIT><<<l.grid, kBlock, 0, c->stream>>>(l.r_lo, l.nrows, l.npad, l.off, val, x, y, part, sc)
  if (xcd) {
    if (unit) FEM_LAT(1, true);
    else FEM_LAT(1, false);
  } else {
    if (unit) FEM_LAT(0, true);
    else FEM_LAT(0, false);
  }
#undef FEM_LAT
}

template <typename X>
static void launch_lat(fem_ctx* c, const LatOp& l, const X* x, X* y, double* part, const Scal* sc) {
  if constexpr (sizeof(X) == 4) {
    if (l.ns == 14) launch_lat_ns<float, float, 14>(c, l, l.val32, x, y, part, sc);
    else launch_lat_ns<float, float, 5>(c, l, l.val32, x, y, part, sc);
  } else {
    if (l.ns == 14) launch_lat_ns<double, double, 14>(c, l, l.val64, x, y, part, sc);
    else launch_lat_ns<double, double, 5>(c, l, l.val64, x, y, part, sc);
  }
}

// K¹: cell 0's element matrix of the Poisson element for a unit coefficient, from the
// uploaded CellValues tables, in the element kernels' arithmetic (J_ab = Σ_j x_j[a] dM_j[b],
// ∇N = dN · J⁻¹, K¹_ij = Σ_q ∇N_i·∇N_j det J w_q).  False when the tables do not describe a
// Q1 scalar field on Q1 geometry of the lattice's dimension.
static bool cell_kref(const fem_ctx* c, const double (*x0)[3], int dim, CellKref* out) {
  const int nb = 1 << dim;
  if (!c->have_cv || c->rdim != dim || c->nbf != nb || c->ngeo != nb || c->vdim != 1) return false;
  *out = CellKref{};
  const RefTables& R = c->ref;
  for (int q = 0; q < c->nqp; ++q) {
    double J[3][3] = {}, Ji[3][3] = {};
    for (int j = 0; j < nb; ++j)
      for (int a = 0; a < dim; ++a)
        for (int b = 0; b < dim; ++b) J[a][b] += x0[j][a] * R.dM[q][j][b];
    double det;
    if (dim == 2) {
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      Ji[0][0] = J[1][1] / det; Ji[0][1] = -J[0][1] / det;
      Ji[1][0] = -J[1][0] / det; Ji[1][1] = J[0][0] / det;
    } else {
      const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1], c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2],
                   c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      Ji[0][0] = c00 / det; Ji[1][0] = c01 / det; Ji[2][0] = c02 / det;
      Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
      Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
      Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
      Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
      Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
      Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    }
    if (!(det > 0.0)) return false;
    double gN[8][3] = {};
    for (int i = 0; i < nb; ++i)
      for (int b = 0; b < dim; ++b)
        for (int a = 0; a < dim; ++a) gN[i][b] += R.dN[q][i][a] * Ji[a][b];
    const double dV = det * R.w[q];
    for (int i = 0; i < nb; ++i)
      for (int j = i; j < nb; ++j) {
        double g = 0.0;
        for (int b = 0; b < dim; ++b) g += gN[i][b] * gN[j][b];
        out->k[i * nb + j] += g * dV;
      }
  }
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < i; ++j) out->k[i * nb + j] = out->k[j * nb + i];
  return true;
}

// Steps in flight (register stages) of the z-marching cell SpMV: three for FP32 with one row
// per thread, otherwise two (register pressure; at 128³ three or more stages measured no
// faster with two rows per thread: the loads of two steps ahead already arrive in time).
constexpr int cell_stages(int bytes, int rpt) { return bytes == 4 && rpt == 1 ? 3 : 2; }

// Launch shape of the cell-form SpMV for values of `bytes` bytes.  Measured at 128³ (in the
// solve, z-marching, XCD ranges): FP32 two rows per thread 768 blocks 12.1 µs, 512 12.4, 640
// 15.5, 896 and 1024 14.5; one row per thread 768 12.9, 1024 13.3.  Z-marching: about one
// work item per block (the planes per chunk follow from the block count).
static void cell_shape(const fem_ctx* c, const LatOp& l, CellOp* cl, int bytes) {
  const int64_t want = c->tune.cell_blocks ? c->tune.cell_blocks : 768;
  const int pi = bytes == 4 ? 0 : 1;
  int& grid = bytes == 4 ? cl->grid32 : cl->grid64;
  cl->zn[pi] = 0;
  if (cl->dim == 3) {
    // z-marching: rows per thread (two only in FP32: the FP64 windows would not fit in LDS),
    // then the window entries per thread (ZN, KN) of an instantiated kernel
    int rpt = bytes == 4 ? (c->tune.cell_rows ? c->tune.cell_rows : 2) : 1;
    if (rpt == 2 && l.px > 255) rpt = 1;
    const int64_t T = rpt * kBlock;
    const int zn = (int)((T + 2 * l.px + 2 + kBlock - 1) / kBlock), kn = (int)((T + l.px + 1 + kBlock - 1) / kBlock);
    const bool inst = rpt == 1 ? ((zn == 2 && kn == 2) || (zn == 3 && kn == 2) || (zn == 4 && kn == 3))
                               : ((zn == 3 && kn == 3) || (zn == 4 && kn == 3));
    if (inst) {
      cl->zn[pi] = zn;
      cl->kn[pi] = kn;
      cl->rpt[pi] = rpt;
      cl->ntx[pi] = (int)((l.pxy + T - 1) / T);
    }
  }
  if (cl->zn[pi] > 0) {
    const int64_t pz = l.npad / l.pxy;
    int& lz = bytes == 4 ? cl->lz32 : cl->lz64;
    int64_t& nitems = bytes == 4 ? cl->nitems32 : cl->nitems64;
    lz = (int)std::min<int64_t>(16, std::max<int64_t>(2, (cl->ntx[pi] * (pz - 2) + want - 1) / want));
    nitems = (int64_t)cl->ntx[pi] * ((pz - 2 + lz - 1) / lz);
    grid = (int)std::max<int64_t>(8, std::min<int64_t>(want, (nitems + 7) / 8 * 8));
    const int vw = 16 / bytes;   // the kernel's 16-byte chunks per window entry group
    const int64_t words = (2 * ((cl->zn[pi] + vw - 1) / vw) + 2 * ((cl->kn[pi] + vw - 1) / vw)) * vw *
                          (int64_t)kBlock;   // the windows
    (bytes == 4 ? cl->ldsz32 : cl->ldsz64) = (int)(bytes * words);
    // the (padded) windows and the staged rows must fit one block's LDS, and the vectors the
    // 32-bit extent of a buffer resource; otherwise the tiles
    int maxlds = 0;
    if (hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) != hipSuccess)
      maxlds = 65536;
    if (bytes * words > (int64_t)maxlds || l.npad * bytes >= (int64_t)1 << 31) cl->zn[pi] = 0;
  }
  if (cl->zn[pi] == 0) {
    const int64_t ntiles = (l.nrows + kBlock - 1) / kBlock;
    int64_t gr = std::min<int64_t>(want, ntiles);
    if (gr >= 8) gr -= gr % 8;
    grid = (int)std::max<int64_t>(1, gr);
  }
}

// Box cells: K¹[a][b] depends only on the axes along which corners a and b differ.  f[m] =
// K¹ of corner 0 and the corner at offset bits m; false unless every entry matches its class to
// 1e-12 of the largest.  K¹ is then replaced by the class form the SpMV applies.
static bool cell_classes(int dim, CellKref* kr, double* f) {
  const int nb = 1 << dim;
  double big = 0.0;
  for (int m = 0; m < 8; ++m) f[m] = 0.0;
  for (int m = 0; m < nb; ++m) {
    f[m] = kr->k[cell_lid(m & 1, (m >> 1) & 1, m >> 2)];   // row of corner 0 = offset (0,0,0)
    big = std::max(big, std::fabs(f[m]));
  }
  for (int a = 0; a < nb; ++a)
    for (int b = 0; b < nb; ++b) {
      double& v = kr->k[cell_lid(a & 1, (a >> 1) & 1, a >> 2) * nb + cell_lid(b & 1, (b >> 1) & 1, b >> 2)];
      if (!(std::fabs(v - f[a ^ b]) <= 1e-12 * big)) return false;
      v = f[a ^ b];
    }
  return true;
}

// Cell form of Â for the solver in X precision (needs the lattice operator's numbering and
// the Jacobi scaling): the cell origins once per pattern, then whenever K, the cell
// parameters, the tables or the constraints changed, the coefficients and the entry check.
// *ok = true when the check passed (to 2⁻²² of Â's unit diagonal for the FP32 solver, 1e-12
// for the FP64 one); t, the masked scaling, is refreshed on every call.
template <typename X>
static int ensure_cells(fem_ctx* c, SolverWork* w, int jacobi, bool* ok) {
  *ok = false;
  CellOp& cl = w->cells;
  const LatOp& l = w->lat;
  int rc;
  if (cl.state == -1 || !jacobi || l.state != 1 || !c->have_param || !c->d_param_cell) return FEM_OK;
  if (cl.state == 0) {
    cl.state = -1;
    const int dim = l.ns == 14 ? 3 : 2;
    const int nb = 1 << dim;
    if (c->npc != nb || (int64_t)c->h_cell_dofs.size() != c->ncells * nb) return FEM_OK;
    // LDS windows of the SpMV: at most 24 entries per thread (FP64: 48 KiB per block)
    const int64_t words = (dim == 3 ? 3 : 1) * (kBlock + 2 * l.px + 2) + (dim == 3 ? 2 : 1) * (kBlock + l.px + 1);
    if (words > 24 * kBlock) return FEM_OK;
    std::vector<int32_t> corg(c->ncells);
    for (int64_t e = 0; e < c->ncells; ++e) corg[e] = l.hmap[c->h_cell_dofs[e * nb]];
    for (int a = 0; a < nb; ++a) {
      const int32_t nd = c->h_cell_nodes[a];
      for (int k = 0; k < dim; ++k)
        FEM_HIP(c, hipMemcpy(&cl.x0[a][k], c->d_x[k] + nd, sizeof(double), hipMemcpyDeviceToHost));
    }
    if ((rc = dev_alloc(c, &cl.corg, c->ncells)) || (rc = dev_alloc(c, &cl.kc64, l.npad)) ||
        (rc = dev_alloc(c, &cl.kc32, l.npad)) || (rc = dev_alloc(c, &cl.cnt, 4)))
      return rc;
    FEM_HIP(c, hipMemcpyAsync(cl.corg, corg.data(), 4 * c->ncells, hipMemcpyHostToDevice, c->stream));
    FEM_HIP(c, hipMemsetAsync(cl.kc64, 0, 8 * l.npad, c->stream));
    FEM_HIP(c, hipMemsetAsync(cl.kc32, 0, 4 * l.npad, c->stream));
    FEM_HIP(c, hipStreamSynchronize(c->stream));
    cl.dim = dim;
    cl.ne = (int)((words + kBlock - 1) / kBlock);
    cl.lds32 = (int)(4 * words);
    cl.lds64 = (int)(8 * words);
    cell_shape(c, l, &cl, 4);
    cell_shape(c, l, &cl, 8);
    cl.state = 1;
  }
  X*& t = sizeof(X) == 4 ? reinterpret_cast<X*&>(cl.t32) : reinterpret_cast<X*&>(cl.t64);
  if (!t && (rc = dev_alloc(c, &t, l.npad))) return rc;
  k_cell_scale<X><<<kMaxVecBlocks, kBlock, 0, c->stream>>>(l.npad, l.dofp, w->sc, c->np > 0 ? c->d_isc : nullptr, t);
  uint64_t& ver = sizeof(X) == 4 ? cl.ver32 : cl.ver64;
  uint64_t& aux = sizeof(X) == 4 ? cl.aux32 : cl.aux64;
  int& jac = sizeof(X) == 4 ? cl.jac32 : cl.jac64;
  bool& pass = sizeof(X) == 4 ? cl.ok32 : cl.ok64;
  if (ver != c->nzval_version + 1 || aux != c->aux_version + 1 || jac != jacobi) {
    pass = false;
    if (cell_kref(c, cl.x0, cl.dim, &cl.kref) && cell_classes(cl.dim, &cl.kref, cl.f)) {
      k_cell_coef<<<kMaxVecBlocks, kBlock, 0, c->stream>>>(c->ncells, cl.corg, c->d_param_cell, cl.kc64, cl.kc32);
      const LatGeom g{(uint32_t)l.px, (uint32_t)l.pxy, (uint32_t)(1 + l.px + l.pxy), l.ns, l.npad};
      const int64_t rows_per_block = kBlock / kCheckLanes;
      const int fb = (int)std::min<int64_t>(16384, (c->n + rows_per_block - 1) / rows_per_block);
      const double tol = sizeof(X) == 4 ? 1.0 / (1 << 22) : 1e-12;
      const uint8_t* isc = c->np > 0 ? c->d_isc : nullptr;
      FEM_HIP(c, hipMemsetAsync(cl.cnt, 0, 4 * sizeof(unsigned long long), c->stream));
      if (cl.dim == 3)
        k_cell_check<3><<<fb, kBlock, 0, c->stream>>>(c->n, l.dofp, c->d_colptr, c->d_rowval, c->d_nzval,
                                                      w->sc, isc, l.map, g, cl.kc64, cl.kref, tol, kCellTolR,
                                                      cl.cnt);
      else
        k_cell_check<2><<<fb, kBlock, 0, c->stream>>>(c->n, l.dofp, c->d_colptr, c->d_rowval, c->d_nzval,
                                                      w->sc, isc, l.map, g, cl.kc64, cl.kref, tol, kCellTolR,
                                                      cl.cnt);
      unsigned long long h[4] = {};
      FEM_HIP(c, hipMemcpyAsync(h, cl.cnt, sizeof h, hipMemcpyDeviceToHost, c->stream));
      FEM_HIP(c, hipStreamSynchronize(c->stream));
      pass = h[0] == 0 && h[1] == h[2];
      cl.okr = h[3] == 0 && h[1] == h[2];
      cl.verr = c->nzval_version + 1;
      cl.auxr = c->aux_version + 1;
      if (cl.okr && !cl.up64) {
        if ((rc = dev_alloc(c, &cl.up64, l.npad)) || (rc = dev_alloc(c, &cl.yp64, l.npad)) ||
            (rc = dev_alloc(c, &cl.tm64, l.npad)) || (rc = dev_alloc(c, &cl.s0, 1)))
          return rc;
        FEM_HIP(c, hipMemsetAsync(cl.up64, 0, 8 * l.npad, c->stream));
        FEM_HIP(c, hipMemsetAsync(cl.s0, 0, sizeof(Scal), c->stream));
      }
      if (cl.okr && cl.auxm != c->aux_version + 1) {
        k_cell_mask<<<kMaxVecBlocks, kBlock, 0, c->stream>>>(l.npad, l.dofp, isc, cl.tm64);
        cl.auxm = c->aux_version + 1;
      }
    }
    ver = c->nzval_version + 1;
    aux = c->aux_version + 1;
    jac = jacobi;
  }
  *ok = pass;
  return FEM_OK;
}

template <typename X, int DIM, int NE>
static void launch_cells_ne(fem_ctx* c, const SolverWork* w, const X* x, X* y, double* part, const Scal* sc,
                            const X* tover = nullptr) {
  const LatOp& l = w->lat;
  const CellOp& cl = w->cells;
  CellKrefT<X> kr;
  for (int i = 0; i < 8; ++i) kr.f[i] = (X)cl.f[i];
  const X* kc = sizeof(X) == 4 ? reinterpret_cast<const X*>(cl.kc32) : reinterpret_cast<const X*>(cl.kc64);
  const X* t = tover ? tover
                     : sizeof(X) == 4 ? reinterpret_cast<const X*>(cl.t32) : reinterpret_cast<const X*>(cl.t64);
  const int lds = sizeof(X) == 4 ? cl.lds32 : cl.lds64;
  const int grid = sizeof(X) == 4 ? cl.grid32 : cl.grid64;
  if (c->tune.sell_map != 1 && grid % 8 == 0)
    k_cell_spmv<X, DIM, 1, NE><<<grid, kBlock, lds, c->stream>>>(l.r_lo, l.nrows, l.npad, (int32_t)l.px,
                                                                 (int32_t)l.pxy, kc, t, x, y, part, sc, kr);
  else
    k_cell_spmv<X, DIM, 0, NE><<<grid, kBlock, lds, c->stream>>>(l.r_lo, l.nrows, l.npad, (int32_t)l.px,
                                                                 (int32_t)l.pxy, kc, t, x, y, part, sc, kr);
}

template <typename X, int ZN, int KN, int RPT>
static void launch_cells_z(fem_ctx* c, const SolverWork* w, const X* x, X* y, double* part, const Scal* sc,
                           const X* tover = nullptr) {
  const LatOp& l = w->lat;
  const CellOp& cl = w->cells;
  CellKrefT<X> kr;
  for (int i = 0; i < 8; ++i) kr.f[i] = (X)cl.f[i];
  const X* kc = sizeof(X) == 4 ? reinterpret_cast<const X*>(cl.kc32) : reinterpret_cast<const X*>(cl.kc64);
  const X* t = tover ? tover
                     : sizeof(X) == 4 ? reinterpret_cast<const X*>(cl.t32) : reinterpret_cast<const X*>(cl.t64);
  const int lds = sizeof(X) == 4 ? cl.ldsz32 : cl.ldsz64;
  const int32_t pz = (int32_t)(l.npad / l.pxy);
  constexpr int D = cell_stages(sizeof(X), RPT);   // steps in flight (register stages)
  const int grid = sizeof(X) == 4 ? cl.grid32 : cl.grid64, lz = sizeof(X) == 4 ? cl.lz32 : cl.lz64;
  const int64_t nitems = sizeof(X) == 4 ? cl.nitems32 : cl.nitems64;
  const int ntx = cl.ntx[sizeof(X) == 4 ? 0 : 1];
  if (c->tune.sell_map != 1 && grid % 8 == 0)
    k_cell_spmv_z<X, 1, ZN, KN, D, RPT><<<grid, kBlock, lds, c->stream>>>(l.npad, (int32_t)l.px, (int32_t)l.pxy, pz,
                                                                          ntx, lz, nitems, kc, t, x, y, part, sc, kr);
  else
    k_cell_spmv_z<X, 0, ZN, KN, D, RPT><<<grid, kBlock, lds, c->stream>>>(l.npad, (int32_t)l.px, (int32_t)l.pxy, pz,
                                                                          ntx, lz, nitems, kc, t, x, y, part, sc, kr);
}

// NE: window entries per thread, the smallest instantiated bound of ⌈E / kBlock⌉
template <typename X, int DIM>
static void launch_cells_dim(fem_ctx* c, const SolverWork* w, const X* x, X* y, double* part, const Scal* sc,
                             const X* tover = nullptr) {
  const int pi = sizeof(X) == 4 ? 0 : 1;
  if (DIM == 3 && w->cells.zn[pi] > 0) {
    const int zn = w->cells.zn[pi];
    if (w->cells.rpt[pi] == 2) {   // FP32 only
      if constexpr (sizeof(X) == 4) {
        if (zn == 3) launch_cells_z<X, 3, 3, 2>(c, w, x, y, part, sc, tover);
        else launch_cells_z<X, 4, 3, 2>(c, w, x, y, part, sc, tover);
      }
    } else {
      if (zn == 2) launch_cells_z<X, 2, 2, 1>(c, w, x, y, part, sc, tover);
      else if (zn == 3) launch_cells_z<X, 3, 2, 1>(c, w, x, y, part, sc, tover);
      else launch_cells_z<X, 4, 3, 1>(c, w, x, y, part, sc, tover);
    }
    return;
  }
  const int ne = w->cells.ne;
  if (ne <= 4) launch_cells_ne<X, DIM, 4>(c, w, x, y, part, sc, tover);
  else if (ne <= 6) launch_cells_ne<X, DIM, 6>(c, w, x, y, part, sc, tover);
  else if (ne <= 8) launch_cells_ne<X, DIM, 8>(c, w, x, y, part, sc, tover);
  else if (ne <= 10) launch_cells_ne<X, DIM, 10>(c, w, x, y, part, sc, tover);
  else if (ne <= 12) launch_cells_ne<X, DIM, 12>(c, w, x, y, part, sc, tover);
  else if (ne <= 16) launch_cells_ne<X, DIM, 16>(c, w, x, y, part, sc, tover);
  else launch_cells_ne<X, DIM, 24>(c, w, x, y, part, sc, tover);
}

template <typename X>
static void launch_cells(fem_ctx* c, const SolverWork* w, const X* x, X* y, double* part, const Scal* sc,
                         const X* tover = nullptr) {
  if (w->cells.dim == 3) launch_cells_dim<X, 3>(c, w, x, y, part, sc, tover);
  else launch_cells_dim<X, 2>(c, w, x, y, part, sc, tover);
}

// SELL structure (once per pattern): widths/bases on the device, offsets on the host.
static int ensure_sell(fem_ctx* c, SolverWork* w) {
  SellOp& o = w->sell;
  if (o.built) return FEM_OK;
  const int64_t n = c->n;
  const int64_t ns = (n + 63) / 64;
  int32_t *dw = nullptr, *ds = nullptr;
  int rc;
  if ((rc = dev_alloc(c, &dw, ns)) || (rc = dev_alloc(c, &ds, ns)) || (rc = dev_alloc(c, &o.base, ns))) return rc;
  const int sb = (int)((ns + kBlock / 64 - 1) / (kBlock / 64));
  k_sell_shape<<<sb, kBlock, 0, c->stream>>>(n, ns, c->d_colptr, c->d_rowval, dw, o.base, ds);
  std::vector<int32_t> hw(ns), hs(ns);
  FEM_HIP(c, hipMemcpyAsync(hw.data(), dw, 4 * ns, hipMemcpyDeviceToHost, c->stream));
  FEM_HIP(c, hipMemcpyAsync(hs.data(), ds, 4 * ns, hipMemcpyDeviceToHost, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  dev_free(&dw);
  dev_free(&ds);
  std::vector<int32_t> off(ns + 1);
  int64_t acc = 0;
  int32_t maxspan = 0;
  for (int64_t i = 0; i < ns; ++i) {
    off[i] = (int32_t)acc;
    acc += 64 * (int64_t)hw[i];
    maxspan = std::max(maxspan, hs[i]);
    if (acc >= INT32_MAX) return set_error(c, FEM_ERR_ARG, "matrix too large for the sliced solver operator");
  }
  off[ns] = (int32_t)acc;
  const bool comp = maxspan < 65536 && c->tune.sell_compress >= 0;
  if ((rc = dev_alloc(c, &o.off, ns + 1)) || (rc = dev_alloc(c, &o.src, acc))) return rc;
  if (comp) rc = dev_alloc(c, &o.roff, acc);
  else rc = dev_alloc(c, &o.col, acc);
  if (rc) return rc;
  FEM_HIP(c, hipMemcpyAsync(o.off, off.data(), 4 * (ns + 1), hipMemcpyHostToDevice, c->stream));
  if (comp)
    k_sell_fill<true><<<sb, kBlock, 0, c->stream>>>(n, ns, c->d_colptr, c->d_rowval, o.off, o.base, o.roff, nullptr,
                                                     o.src);
  else
    k_sell_fill<false><<<sb, kBlock, 0, c->stream>>>(n, ns, c->d_colptr, c->d_rowval, o.off, o.base, nullptr, o.col,
                                                      o.src);
  FEM_HIP(c, hipGetLastError());
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  o.nslices = ns;
  o.len = acc;
  o.ver32 = o.ver64 = 0;
  o.jac32 = o.jac64 = -1;
  // a multiple of 8 (one share per XCD) when there are enough slices
  int64_t sg = std::min<int64_t>(kMaxVecBlocks, sb);
  if (sg >= 8) sg -= sg % 8;
  w->sell_grid = (int)std::max<int64_t>(1, sg);
  o.built = true;
  return FEM_OK;
}

// Scaled SELL values of the current K (re-gathered when nzval or the scaling changed).
template <typename V>
static int ensure_sell_vals(fem_ctx* c, SolverWork* w, int jacobi) {
  SellOp& o = w->sell;
  V*& val = sizeof(V) == 4 ? reinterpret_cast<V*&>(o.val32) : reinterpret_cast<V*&>(o.val64);
  uint64_t& ver = sizeof(V) == 4 ? o.ver32 : o.ver64;
  int& jac = sizeof(V) == 4 ? o.jac32 : o.jac64;
  if (!val) {
    int rc = dev_alloc(c, &val, o.len);
    if (rc) return rc;
    ver = 0;
  }
  if (ver != c->nzval_version + 1 || jac != jacobi) {
    const int sb = (int)std::min<int64_t>(8192, (o.nslices + kBlock / 64 - 1) / (kBlock / 64));
    if (o.roff)
      k_sell_vals<V, true><<<sb, kBlock, 0, c->stream>>>(c->n, o.nslices, o.off, o.base, o.roff, nullptr, o.src,
                                                          c->d_nzval, w->sc, val);
    else
      k_sell_vals<V, false><<<sb, kBlock, 0, c->stream>>>(c->n, o.nslices, o.off, o.base, nullptr, o.col, o.src,
                                                           c->d_nzval, w->sc, val);
    FEM_HIP(c, hipGetLastError());
    ver = c->nzval_version + 1;
    jac = jacobi;
  }
  return FEM_OK;
}

// One SELL launch with the unroll depth and slice mapping of the tuning.
template <typename V, typename X, int MODE, bool COMP>
static void launch_sell(fem_ctx* c, const SellOp& o, int g, const V* val, const X* x, const double* f, X* y,
                        double* part, const Scal* sc) {
  // measured at 128³ Q1 hex: 8 entries per lane in flight for FP32 values, 4 for FP64;
  // grid-stride slices beat the XCD-contiguous ranges in both precisions
  const int u = c->tune.sell_unroll ? c->tune.sell_unroll : (sizeof(V) == 4 ? 8 : 4);
  const bool contig = c->tune.sell_map == 2 && g % 8 == 0;
  // non-temporal operator loads by default: measured at 128³, FP64 CG loop 103.9 → 96.5 ms,
  // mixed loop 69.0 → 66.3 ms (the gathered vectors stay in L2)
  const bool nt = c->tune.sell_nt != -1;
  const int32_t* col = COMP ? nullptr : o.col;
  const uint16_t* roff = COMP ? o.roff : nullptr;
#define FEM_SELL(UU, CC)                                                                                 \
  k_sell<V, X, MODE, COMP, UU, CC><<<g, kBlock, 0, c->stream>>>(c->n, o.nslices, o.off, o.base, roff, col, \
                                                                 val, x, f, y, part, sc)
  if (nt && !contig) {
    const int32_t* cl = col;
    if (u == 4)
      k_sell<V, X, MODE, COMP, 4, false, true><<<g, kBlock, 0, c->stream>>>(c->n, o.nslices, o.off, o.base, roff,
                                                                           cl, val, x, f, y, part, sc);
    else
      k_sell<V, X, MODE, COMP, 8, false, true><<<g, kBlock, 0, c->stream>>>(c->n, o.nslices, o.off, o.base, roff,
                                                                           cl, val, x, f, y, part, sc);
    return;
  }
  if (contig) {
    if (u == 4) FEM_SELL(4, true);
    else if (u == 16) FEM_SELL(16, true);
    else FEM_SELL(8, true);
  } else {
    if (u == 4) FEM_SELL(4, false);
    else if (u == 16) FEM_SELL(16, false);
    else FEM_SELL(8, false);
  }
#undef FEM_SELL
}

// y = Â x with the solver's sliced operator (MODE 0, partial x·y per block); returns the
// number of partials written.
template <typename X>
static int spmv_hat(fem_ctx* c, SolverWork* w, const X* x, X* y, const Scal* sc) {
  if (w->use_elem) return elem_op_launch(c, w->eop, x, y, w->part_s, sc);
  if (w->use_cells) {
    launch_cells<X>(c, w, x, y, w->part_s, sc);
    return sizeof(X) == 4 ? w->cells.grid32 : w->cells.grid64;
  }
  if (w->use_lat) {
    launch_lat<X>(c, w->lat, x, y, w->part_s, sc);
    return w->lat.grid;
  }
  const SellOp& o = w->sell;
  const int g = w->sell_grid;
  if constexpr (sizeof(X) == 4) {
    if (o.roff) launch_sell<float, float, 0, true>(c, o, g, o.val32, x, nullptr, y, w->part_s, sc);
    else launch_sell<float, float, 0, false>(c, o, g, o.val32, x, nullptr, y, w->part_s, sc);
  } else {
    if (o.roff) launch_sell<double, double, 0, true>(c, o, g, o.val64, x, nullptr, y, w->part_s, sc);
    else launch_sell<double, double, 0, false>(c, o, g, o.val64, x, nullptr, y, w->part_s, sc);
  }
  return g;
}

// The scaling S and the scaled operator in X precision, ready for spmv_hat: the lattice
// planes when the dofs form a lattice and K is symmetric (the mixed solver by default; the
// FP64 solver with tuning lattice = 1), otherwise the sliced copy.  Returns the length of
// the recurrence vectors in *nv.
template <typename X>
static int prepare_operator(fem_ctx* c, SolverWork* w, int jacobi, int64_t* nv) {
  int rc;
  k_scale<<<kMaxVecBlocks, kBlock, 0, c->stream>>>(c->n, c->d_diag, c->d_nzval, jacobi, w->sc);
  w->use_lat = false;
  w->use_cells = false;
  w->use_elem = false;
  const bool want = sizeof(X) == 4 ? c->tune.lattice != -1 : c->tune.lattice == 1;
  const bool want_cells = sizeof(X) == 4 ? c->tune.cell_operator != -1 : c->tune.cell_operator == 1;
  if (want || want_cells) {
    if ((rc = ensure_lattice(c, w))) return rc;
    if (want_cells && w->lat.state == 1) {
      bool ok = false;
      if ((rc = ensure_cells<X>(c, w, jacobi, &ok))) return rc;
      w->use_cells = w->use_lat = ok;
    }
    if (want && !w->use_cells && w->lat.state == 1) {
      bool sym = false;
      if ((rc = ensure_lat_vals<X>(c, w, jacobi, &sym))) return rc;
      w->use_lat = sym;
    }
  }
  // the stored element matrices where no lattice form applies (any mesh): by default where
  // they beat the sliced copy, i.e. the Q2 node blocks (measured at 48³: FP32 SpMV 0.75 vs
  // 1.40 ms, FP64 1.19 vs 2.11 ms), while the scalar and Q1² planes move more bytes per row
  // than the copy (128³ Q1 hex: 116 vs 74 µs) and run only with tuning elem_operator = 1
  const bool want_elem = c->tune.elem_operator == 1 || (c->tune.elem_operator == 0 && c->ea_kind == 3);
  if (!w->use_lat && want_elem && elem_op_valid(c)) {
    if ((rc = elem_op_prepare(c, &w->eop, w->sc, sizeof(X) == 4, &w->elem_grid))) return rc;
    w->use_elem = true;
  }
  *nv = w->use_lat ? w->lat.npad : c->n;
  if ((rc = ensure_vecs<X>(c, w, *nv))) return rc;
  if (w->use_lat || w->use_elem) return FEM_OK;
  if ((rc = ensure_sell(c, w))) return rc;
  return ensure_sell_vals<X>(c, w, jacobi);
}

// With time_kernels: add the durations of the SpMVs [first, first + count) of the batch just
// run (the stream is synchronized) — the ones that did work, not those that returned at once.
static void add_spmv_times(fem_ctx* c, SolverWork* w, int first, int count) {
  if (!c->tune.time_kernels) return;
  for (int k = first; k < first + count && k < kBatch; ++k) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, w->tev[2 * k], w->tev[2 * k + 1]) == hipSuccess) {
      c->timings.spmv_ms += ms;
      c->timings.spmv_launches += 1;
    }
  }
}

static int read_scal(fem_ctx* c, SolverWork* w, Scal* h) {   // scalar record 0 (the current one)
  FEM_HIP(c, hipMemcpyAsync(w->hscal, w->s, sizeof(Scal), hipMemcpyDeviceToHost, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  *h = *w->hscal;
  return FEM_OK;
}

// The vectors of one CG recurrence (FP64 solver or FP32 recurrence of the mixed solver).
template <typename X>
struct CgVecs {
  X *x, *r, *w, *p, *sv;
  int64_t n;               // their length
  const int32_t* map;      // recurrence index of each dof (lattice operator), or null
};

// One Chronopoulos–Gear iteration: w = Â r̂ (partial r̂·w), then the fused vector step.
// `k` is the position in the captured batch: it picks the scalar/partial buffers.
template <typename X>
static void enqueue_iteration(fem_ctx* c, SolverWork* w, const CgVecs<X>& v, int k, bool timed) {
  Scal* sin = w->s + (k & 1);
  Scal* sout = w->s + ((k + 1) & 1);
  double* pin = w->part_v + (k & 1) * 2 * kCgBlocks;
  double* pout = w->part_v + ((k + 1) & 1) * 2 * kCgBlocks;
  if (timed) (void)hipEventRecord(w->tev[2 * k], c->stream);
  const int na = spmv_hat<X>(c, w, v.r, v.w, sin);
  if (timed) (void)hipEventRecord(w->tev[2 * k + 1], c->stream);
  const int mb = cg_blocks(c, v.n);
  if (mb <= kCgFuseBlocks) {   // the vector step does the scalar step too
    const bool pre = mb <= kCgPreBlocks;
#define FEM_MOVE_SCAL(E_)                                                                                      \
  do {                                                                                                          \
    if (pre)                                                                                                    \
      k_cg_move_scal<X, E_, true><<<mb, kBlock, 0, c->stream>>>(v.n, v.x, v.r, v.w, v.p, v.sv, pout, mb, pin, na, \
                                                                w->part_s, sin, sout);                          \
    else                                                                                                        \
      k_cg_move_scal<X, E_, false><<<mb, kBlock, 0, c->stream>>>(v.n, v.x, v.r, v.w, v.p, v.sv, pout, mb, pin,    \
                                                                 na, w->part_s, sin, sout);                     \
  } while (0)
    switch (cg_elems(c)) {
      case 2: FEM_MOVE_SCAL(2); break;
      case 4: FEM_MOVE_SCAL(4); break;
      case 16: FEM_MOVE_SCAL(16); break;
      default: FEM_MOVE_SCAL(8); break;
    }
#undef FEM_MOVE_SCAL
    return;
  }
  k_cg_scal<<<1, 64, 0, c->stream>>>(mb, pin, na, w->part_s, sin, sout);
  switch (cg_elems(c)) {
    case 2: k_cg_move<X, 2><<<mb, kBlock, 0, c->stream>>>(v.n, v.x, v.r, v.w, v.p, v.sv, pout, sout); break;
    case 4: k_cg_move<X, 4><<<mb, kBlock, 0, c->stream>>>(v.n, v.x, v.r, v.w, v.p, v.sv, pout, sout); break;
    case 16: k_cg_move<X, 16><<<mb, kBlock, 0, c->stream>>>(v.n, v.x, v.r, v.w, v.p, v.sv, pout, sout); break;
    default: k_cg_move<X, 8><<<mb, kBlock, 0, c->stream>>>(v.n, v.x, v.r, v.w, v.p, v.sv, pout, sout); break;
  }
}

// The captured batch of len iterations (rebuilt only when the buffers change).  len is even,
// so the batch starts and ends on scalar record 0 and partial buffer 0.
template <typename X>
static int batch_graph(fem_ctx* c, SolverWork* w, const CgVecs<X>& v, GraphCache& gc, int len) {
  static_assert(kBatch % 2 == 0 && kBatchMin % 2 == 0, "batch must end on scalar record 0");
  const void* val = w->use_elem ? elem_op_key(c, w->eop, sizeof(X) == 4)
                    : w->use_cells ? (sizeof(X) == 4 ? (const void*)w->cells.kc32 : (const void*)w->cells.kc64)
                    : w->use_lat ? (sizeof(X) == 4 ? (const void*)w->lat.val32 : (const void*)w->lat.val64)
                                 : (sizeof(X) == 4 ? (const void*)w->sell.val32 : (const void*)w->sell.val64);
  // the lattice SpMV variant depends on the values (a unit diagonal), not only on the operands;
  // the cell form's K¹ class values are kernel arguments, so its graph is keyed on their bits
  // (a new K with new k_e reuses the graph: k_e is read from memory)
  uintptr_t variant = w->use_lat ? 1 + (sizeof(X) == 4 ? w->lat.unit32 : w->lat.unit64)
                      : w->use_elem ? 8 + c->ea_kind : 0;
  if (w->use_cells) {
    uint64_t h = 1469598103934665603ull;   // FNV-1a over the eight values in X precision
    for (int i = 0; i < 8; ++i) {
      const X v = (X)w->cells.f[i];
      uint64_t bits = 0;
      std::memcpy(&bits, &v, sizeof(X));
      h = (h ^ bits) * 1099511628211ull;
    }
    variant = (uintptr_t)(h << 2) | 3;
  }
  const void* key[9] = {val, v.x, v.r, v.p, v.w, v.sv, w->s,
                        w->use_lat ? (const void*)w->lat.map
                        : w->use_elem ? (c->ea_kind == 3 ? (const void*)c->d_n2c : (const void*)c->d_d2c)
                                      : w->sell.off,
                        reinterpret_cast<const void*>(variant)};
  if (c->tune.time_kernels) {   // kernels launched directly (launch_batch), SpMVs bracketed by events
    for (hipEvent_t& e : w->tev)
      // timing-only events: no system-scope fence (its L2 writeback and invalidation would
      // slow the kernel that follows each event and inflate the measured durations)
      if (!e) FEM_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    return FEM_OK;
  }
  bool same = gc.exec != nullptr;
  for (int i = 0; i < 9 && same; ++i) same = gc.key[i] == key[i];
  if (same) return FEM_OK;
  destroy_graph(gc);
  hipGraph_t graph = nullptr;
  FEM_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < len; ++k) enqueue_iteration<X>(c, w, v, k, false);
  FEM_HIP(c, hipStreamEndCapture(c->stream, &graph));
  hipError_t e = hipGraphInstantiate(&gc.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return check_hip(c, e, "hipGraphInstantiate");
  for (int i = 0; i < 9; ++i) gc.key[i] = key[i];
  return FEM_OK;
}

// One batch: the captured graph, or (time_kernels) kBatch direct launches with every SpMV
// bracketed by HIP events.
template <typename X>
static int launch_batch(fem_ctx* c, SolverWork* w, const CgVecs<X>& v, GraphCache& gc) {
  if (c->tune.time_kernels) {
    for (int k = 0; k < kBatch; ++k) enqueue_iteration<X>(c, w, v, k, true);
    FEM_HIP(c, hipGetLastError());
    return FEM_OK;
  }
  FEM_HIP(c, hipGraphLaunch(gc.exec, c->stream));
  return FEM_OK;
}

// (Re)start the recurrence from the FP64 residual `res` (= f − K u, s->rr = its norm²):
// r̂ = S res, w = Â r̂, p and s = Â p rebuilt (orthogonalised old direction, or p = r̂ on
// restart), α = γ / p·s.  Works on scalar record 0 / partial buffer 0.
template <typename X>
static void reseed(fem_ctx* c, SolverWork* w, const CgVecs<X>& v, const double* res, double hold_factor2,
                   double tol2, int restart) {
  const int nb = cg_blocks(c, v.n);
  k_cg_reseed<X><<<nb, kBlock, 0, c->stream>>>(c->n, res, w->sc, v.r, v.p, w->part_v, v.map);
  k_cg_reseed_scal<<<1, kBlock, 0, c->stream>>>(nb, w->part_v, w->s, hold_factor2, tol2, restart);
  spmv_hat<X>(c, w, v.r, v.w, w->s);
  k_cg_pfix<X><<<nb, kBlock, 0, c->stream>>>(v.n, v.r, v.w, v.p, v.sv, w->s, w->part_v);
  k_cg_alpha<<<1, kBlock, 0, c->stream>>>(nb, w->part_v, w->s);
}

// res = f − K u and s->rr = ||res||² with the unscaled FP64 K (CSC arrays, fixed order).
static void residual(fem_ctx* c, SolverWork* w) {
  const CellOp& cl = w->cells;
  if (w->use_cells && cl.okr && cl.up64 && cl.verr == c->nzval_version + 1 && cl.auxr == c->aux_version + 1 &&
      cl.auxm == c->aux_version + 1) {   // on the cell form (k_cell_check passed at kCellTolR)
    const uint8_t* isc = c->np > 0 ? c->d_isc : nullptr;
    k_to_pad<<<kMaxVecBlocks, kBlock, 0, c->stream>>>(c->n, w->lat.map, c->d_u, cl.up64);
    launch_cells<double>(c, w, cl.up64, cl.yp64, w->part_s, cl.s0, cl.tm64);
    k_res_cells<<<kMaxVecBlocks, kBlock, 0, c->stream>>>(c->n, w->lat.map, cl.yp64, c->d_f, c->d_u, isc, c->d_diag,
                                                         c->d_nzval, w->res, w->part_s);
    k_sum_to<<<1, kWide, 0, c->stream>>>(kMaxVecBlocks, w->part_s, w->s, 0);
    return;
  }
  launch_spmv<double, double, 1>(c, w->sh64, c->d_nzval, c->d_u, c->d_f, w->res, w->part_s, w->s);
  k_sum_to<<<1, kWide, 0, c->stream>>>(w->sh64.blocks, w->part_s, w->s, 0);
}

// CG on Â x̂ = S f in X precision from u = 0, with hand-overs: the device stops the
// recurrence when γ has fallen by hold_factor2 (0: never) or claims convergence; the host
// folds x̂ into u, computes the true residual with the FP64 K, stops if it meets tol2 and
// otherwise reseeds from it (residual replacement / reliable update).  It gives up when
// the true residual stalls (improves by less than 10% in norm) at kMaxStall hand-overs in
// a row, or after maxit iterations.  On an ill-conditioned K the FP32 recurrence can claim
// its δ drop for a few hand-overs while the true residual hardly moves and then resume
// (measured on a plane-strain cantilever, cond(K) ~ 1e6): six hand-overs in a row, not two.
constexpr int kMaxStall = 6;

// Index of the next batch's length (kBatchMin << i): the recurrence's remaining iterations
// predicted from γ's geometric rate since the reseed (rz0 → rz in `it` iterations, hand-over
// at gstop), so a batch seldom runs past the hand-over, where its launches return at once.
static int batch_index(const Scal& h, int it) {
  if (it <= 0 || !(h.rz > 0.0) || !(h.rz < h.rz0) || !(h.gstop > 0.0) || !(h.gstop < h.rz)) return 1;
  const double m = it * std::log(h.gstop / h.rz) / std::log(h.rz / h.rz0);
  return m >= 5.0 * kBatchMin ? 2 : m >= 2.5 * kBatchMin ? 1 : 0;
}

template <typename X>
static int cg_run(fem_ctx* c, SolverWork* w, const CgVecs<X>& v, double bnorm2, double tol2, double hold_factor2,
                  int maxit, GraphCache* gcs, int* iters, int* handovers, double* ms) {
  zero_vec<double>(c, c->d_u, c->n);
  // all five: the lattice numbering's padding entries must start (and so stay) zero
  for (X* a : {v.x, v.r, v.w, v.p, v.sv}) zero_vec<X>(c, a, v.n);
  k_cg_init_scal<<<1, 1, 0, c->stream>>>(w->s, bnorm2);
  reseed<X>(c, w, v, c->d_f, hold_factor2, tol2, 1);
  FEM_HIP(c, hipGetLastError());
  int rc;
  int launched = 0, nho = 0, stall = 0, iters_before = 0, iters_seed = 0;
  int li = 1;   // the next batch is kBatchMin << li iterations (kBatch with time_kernels)
  bool converged = false, fresh_start = true;
  double prev_true = bnorm2;
  FEM_HIP(c, hipEventRecord(c->ev0, c->stream));
  while (launched < maxit) {
    if (c->tune.time_kernels) li = 1;
    GraphCache& gc = gcs[li];
    if ((rc = batch_graph<X>(c, w, v, gc, kBatchMin << li))) return rc;   // at once when captured
    if ((rc = launch_batch<X>(c, w, v, gc))) return rc;
    launched += kBatchMin << li;
    Scal h{};
    if ((rc = read_scal(c, w, &h))) return rc;
    // SpMVs that did work: every executed step's except a fresh one's (it skips the SpMV),
    // plus the SpMV of the step that found the hand-over
    const int first = fresh_start ? 1 : 0;
    const int ran = h.iters - iters_before + (h.done ? 1 : 0);
    add_spmv_times(c, w, first, std::max(0, ran - first));
    iters_before = h.iters;
    fresh_start = false;
    if (!h.done) {
      li = batch_index(h, h.iters - iters_seed);
      continue;
    }
    fresh_start = true;
    li = 1;
    iters_seed = h.iters;
    ++nho;
    k_fold<X><<<kMaxVecBlocks, kBlock, 0, c->stream>>>(c->n, v.x, v.p, w->s, w->sc, c->d_u, v.map);
    k_cg_unpend<<<1, 1, 0, c->stream>>>(w->s);
    residual(c, w);
    Scal ht{};
    if ((rc = read_scal(c, w, &ht))) return rc;
    if (ht.rr <= tol2) { converged = true; break; }
    stall = ht.rr > 0.81 * prev_true ? stall + 1 : 0;
    if (stall >= kMaxStall) break;
    prev_true = std::min(prev_true, ht.rr);
    reseed<X>(c, w, v, w->res, hold_factor2, tol2, 0);
  }
  if (!converged) {   // fold what the recurrence holds and leave the true residual in s->rr
    k_fold<X><<<kMaxVecBlocks, kBlock, 0, c->stream>>>(c->n, v.x, v.p, w->s, w->sc, c->d_u, v.map);
    k_cg_unpend<<<1, 1, 0, c->stream>>>(w->s);
    residual(c, w);
  }
  FEM_HIP(c, hipEventRecord(c->ev1, c->stream));
  FEM_HIP(c, hipGetLastError());
  *ms += event_ms(c);
  Scal h{};
  if ((rc = read_scal(c, w, &h))) return rc;
  *iters = h.iters;
  *handovers = nho;
  return converged ? FEM_OK : FEM_ERR_NOCONV;
}

}  // namespace fem

using namespace fem;

extern "C" int32_t fem_solve(fem_ctx* c, const fem_solve_cfg* cfg, fem_solve_info* info) {
  if (!c || !cfg) return FEM_ERR_ARG;
  if (c->nnz == 0 || !c->d_f) return set_error(c, FEM_ERR_STATE, "assemble a system first");
  if (!(cfg->precond == 0 || cfg->precond == 1)) return set_error(c, FEM_ERR_ARG, "precond must be 0 or 1");
  FEM_HIP(c, hipSetDevice(c->device));
  const bool mixed = cfg->mode == FEM_SOLVE_IR_MIXED;
  int rc = ensure_work(c, mixed);
  if (rc) return rc;
  SolverWork* w = c->solver;
  const int64_t n = c->n;
  const int maxit = cfg->maxit > 0 ? cfg->maxit : 100000;
  fem_solve_info out{};
  double cg_ms = 0.0;
  c->timings.spmv_ms = 0.0;
  c->timings.spmv_launches = 0;

  k_sumsq<<<kMaxVecBlocks, kBlock, 0, c->stream>>>(n, c->d_f, w->part_s);
  k_sum_to<<<1, kWide, 0, c->stream>>>(kMaxVecBlocks, w->part_s, w->s, 1);
  Scal h{};
  if ((rc = read_scal(c, w, &h))) return rc;
  const double bnorm2 = h.bnorm2;
  if (bnorm2 == 0.0) {
    k_zero<<<kMaxVecBlocks, kBlock, 0, c->stream>>>(n, c->d_u);
    FEM_HIP(c, hipStreamSynchronize(c->stream));
    out.outer_iterations = 1;
    if (info) *info = out;
    return FEM_OK;
  }
  const double tol2 = cfg->rtol * cfg->rtol * bnorm2;
  int iters = 0, handovers = 0;
  int64_t nv = 0;
  if (!mixed) {
    if ((rc = prepare_operator<double>(c, w, cfg->precond, &nv))) return rc;
    const CgVecs<double> v{w->x, w->r, w->wv, w->p, w->sv, nv, w->use_lat ? w->lat.map : nullptr};
    rc = cg_run<double>(c, w, v, bnorm2, tol2, 0.0, maxit, w->g64, &iters, &handovers, &cg_ms);
  } else {
    // the FP32 recurrence hands over whenever its residual has fallen by δ (inner_rtol)
    if ((rc = prepare_operator<float>(c, w, cfg->precond, &nv))) return rc;
    const double delta = cfg->inner_rtol > 0 ? cfg->inner_rtol : 0.1;
    const CgVecs<float> v{w->x32, w->r32, w->wv32, w->p32, w->sv32, nv, w->use_lat ? w->lat.map : nullptr};
    rc = cg_run<float>(c, w, v, bnorm2, tol2, delta * delta, maxit, w->g32, &iters, &handovers, &cg_ms);
  }
  if (rc && rc != FEM_ERR_NOCONV) return rc;
  out.iterations = iters;
  out.outer_iterations = handovers;
  out.operator_format = w->use_elem ? 4 : w->use_cells ? 3 : !w->use_lat ? 0
                        : (mixed ? w->lat.unit32 : w->lat.unit64) ? 2 : 1;
  if (w->use_cells && w->cells.dim == 3) {   // the z-marching template that ran
    const int pi = mixed ? 0 : 1;
    const CellOp& cl = w->cells;
    const int grid = mixed ? cl.grid32 : cl.grid64;
    if (cl.zn[pi] > 0)
      out.operator_shape = (c->tune.sell_map != 1 && grid % 8 == 0 ? 10000 : 0) + 1000 * cl.zn[pi] +
                           100 * cl.kn[pi] + 10 * cell_stages(mixed ? 4 : 8, cl.rpt[pi]) + cl.rpt[pi];
  }
  if ((rc = read_scal(c, w, &h))) return rc;   // s->rr: the true residual of the returned u
  out.relres = std::sqrt(h.rr / bnorm2);
  c->timings.solve_ms = cg_ms;
  if (info) *info = out;
  if (out.relres > cfg->rtol * 1.0000001) {
    char buf[160];
    snprintf(buf, sizeof buf, "CG did not converge: relres %.3e after %d iterations", out.relres, out.iterations);
    return set_error(c, FEM_ERR_NOCONV, buf);
  }
  return FEM_OK;
}

// Average duration (HIP events) of `reps` back-to-back launches of the solver's SpMV
// w = Â x (sliced operator, FP64 or FP32 values and vectors).
extern "C" int32_t fem_spmv_bench(fem_ctx* c, int32_t mixed, int32_t reps, double* avg_ms) {
  if (!c || reps <= 0 || !avg_ms) return FEM_ERR_ARG;
  if (c->nnz == 0) return set_error(c, FEM_ERR_STATE, "no matrix");
  FEM_HIP(c, hipSetDevice(c->device));
  int rc = ensure_work(c, mixed != 0);
  if (rc) return rc;
  SolverWork* w = c->solver;
  Scal zero{};
  FEM_HIP(c, hipMemcpyAsync(w->s, &zero, sizeof(Scal), hipMemcpyHostToDevice, c->stream));
  int64_t nv = 0;
  if (mixed) {   // x = f (in the operator's numbering when it is the lattice's: padding zero)
    if ((rc = prepare_operator<float>(c, w, 1, &nv))) return rc;
    zero_vec<float>(c, w->p32, nv);
    k_to_f32<<<kMaxVecBlocks, kBlock, 0, c->stream>>>(c->n, c->d_f, w->p32);
  } else {
    if ((rc = prepare_operator<double>(c, w, 1, &nv))) return rc;
    zero_vec<double>(c, w->p, nv);
    k_copy<<<kMaxVecBlocks, kBlock, 0, c->stream>>>(c->n, c->d_f, w->p);
  }
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  FEM_HIP(c, hipEventRecord(c->ev0, c->stream));
  for (int i = 0; i < reps; ++i) {
    if (mixed) spmv_hat<float>(c, w, w->p32, w->wv32, w->s);
    else spmv_hat<double>(c, w, w->p, w->wv, w->s);
  }
  FEM_HIP(c, hipGetLastError());
  FEM_HIP(c, hipEventRecord(c->ev1, c->stream));
  *avg_ms = event_ms(c) / reps;
  return FEM_OK;
}

extern "C" int32_t fem_set_tuning(fem_ctx* c, const fem_tuning* t) {
  if (!c || !t) return FEM_ERR_ARG;
  auto ok_cap = [](int v) { return v == 0 || v == 2048 || v == 4096; };
  auto ok_l = [](int v) { return v == 0 || v == 1 || v == 2 || v == 4; };
  if (!ok_cap(t->spmv_cap64) || !ok_l(t->spmv_lanes64) || !(t->sell_compress == 0 || t->sell_compress == -1) ||
      !(t->asm_block == 0 || t->asm_block == 64 || t->asm_block == 128 || t->asm_block == 256) ||
      !(t->asm_mode >= 0 && t->asm_mode <= 2) || !(t->time_kernels == 0 || t->time_kernels == 1) ||
      !(t->sell_unroll == 0 || t->sell_unroll == 4 || t->sell_unroll == 8 || t->sell_unroll == 16) ||
      !(t->sell_map >= 0 && t->sell_map <= 2) ||
      !(t->gather_group == 0 || t->gather_group == 1 || t->gather_group == 2 || t->gather_group == 4 ||
        t->gather_group == 8) ||
      !(t->cg_elems == 0 || t->cg_elems == 2 || t->cg_elems == 4 || t->cg_elems == 8 || t->cg_elems == 16) ||
      !(t->sell_nt == 0 || t->sell_nt == 1 || t->sell_nt == -1) || !(t->affine == 0 || t->affine == -1) ||
      !(t->lattice == 0 || t->lattice == 1 || t->lattice == -1) ||
      !(t->lattice_blocks == 0 || (t->lattice_blocks >= 8 && t->lattice_blocks <= kMaxVecBlocks &&
                                   t->lattice_blocks % 8 == 0)) ||
      !(t->cell_operator == 0 || t->cell_operator == 1 || t->cell_operator == -1) ||
      !(t->cell_blocks == 0 || (t->cell_blocks >= 8 && t->cell_blocks <= kMaxVecBlocks && t->cell_blocks % 8 == 0)) ||
      !(t->cell_rows >= 0 && t->cell_rows <= 2) ||
      !(t->elem_operator >= -1 && t->elem_operator <= 1))
    return set_error(c, FEM_ERR_ARG, "unsupported tuning value");
  (void)hipStreamSynchronize(c->stream);
  c->tune = *t;
  free_solver(c);
  return FEM_OK;
}

extern "C" int32_t fem_get_tuning(const fem_ctx* c, fem_tuning* t) {
  if (!c || !t) return FEM_ERR_ARG;
  *t = c->tune;
  return FEM_OK;
}
