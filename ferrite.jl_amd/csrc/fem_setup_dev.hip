// Setup on the device: dof distribution (close!(dh)) and the sparsity pattern
// (allocate_matrix(dh)) from the uploaded mesh, with the same results as the reference's
// serial algorithms, which the host restatement (fem_setup.cpp) and the oracle follow.
//
// Dof distribution, src/Dofs/DofHandler.jl:576-800: cells are visited in ascending id and, in
// each cell, its entities in local order (vertices, then edges keyed by the sorted node pair,
// faces keyed by their nodes, then the interior); an entity gets the next vdim dofs at its
// first visit.  The visit sequence is cell-major with S slots per cell (one per local node of
// the Lagrange element), so the entity numbers are the ranks of the entities' first-visit
// positions:
//   k_first_visit  for every (cell, slot): the first cell holding the same entity (the
//                  smallest cell id in the node -> cell list of one of its nodes that holds all
//                  of them) and the entity's slot there; marks the slot that is the first visit
//   scan           exclusive prefix sum of the marks = entity number of each first visit
//   k_cell_dofs    cell_dofs[c][s vdim + d] = vdim · number(first visit of (c, s)) + d
// Only the order of the first visits matters, so the result is the serial loop's.
//
// Pattern, src/Dofs/sparsity_pattern.jl:223-253, :1136-1249, :951-1016: column j holds the
// dofs of every cell that holds j (the diagonal among them), ascending.  One wave per column
// gathers those candidates in LDS, sorts them (bitonic), and counts (pass 1) or writes (pass 2)
// the distinct ones; colptr is the scan of the counts.
#include <algorithm>
#include <climits>
#include <vector>

#include "fem_internal.h"
#include "ferrite_mi355x_setup.h"

namespace fem {

constexpr int kScanBlock = 1024;   // elements per block of the scan's first level
constexpr int kMaxCand = 1024;     // candidate rows per column (cells holding j x dofs per cell)

// ---- exclusive scan of int32 (counts up to 2^31 − 1) -------------------------------------
__global__ __launch_bounds__(256) void k_scan_blocks(int64_t n, const int32_t* __restrict__ in,
                                                     int32_t* __restrict__ out, int32_t* __restrict__ bsum) {
  __shared__ int32_t sh[kScanBlock];
  const int64_t b0 = (int64_t)blockIdx.x * kScanBlock;
  for (int i = threadIdx.x; i < kScanBlock; i += 256) sh[i] = b0 + i < n ? in[b0 + i] : 0;
  __syncthreads();
  for (int o = 1; o < kScanBlock; o <<= 1) {   // Hillis–Steele inclusive scan in LDS
    int32_t v[kScanBlock / 256];
#pragma unroll
    for (int k = 0; k < kScanBlock / 256; ++k) {
      const int i = threadIdx.x + 256 * k;
      v[k] = i >= o ? sh[i - o] : 0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanBlock / 256; ++k) sh[threadIdx.x + 256 * k] += v[k];
    __syncthreads();
  }
  for (int i = threadIdx.x; i < kScanBlock; i += 256)
    if (b0 + i < n) out[b0 + i] = sh[i] - (b0 + i < n ? in[b0 + i] : 0);   // exclusive
  if (threadIdx.x == 0) bsum[blockIdx.x] = sh[kScanBlock - 1];
}

// One block: exclusive scan of the block sums in place (sequential chunks), total in *tot.
__global__ __launch_bounds__(256) void k_scan_sums(int64_t nb, int32_t* __restrict__ bsum, int64_t* __restrict__ tot) {
  __shared__ int64_t carry;
  __shared__ int64_t sh[256];
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < nb; base += 256) {
    const int64_t i = base + threadIdx.x;
    const int64_t v = i < nb ? bsum[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
      const int64_t t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) bsum[i] = (int32_t)(carry + sh[threadIdx.x] - v);
    __syncthreads();
    if (threadIdx.x == 255) carry += sh[255];
    __syncthreads();
  }
  if (threadIdx.x == 0) *tot = carry;
}

__global__ __launch_bounds__(256) void k_scan_add(int64_t n, const int32_t* __restrict__ bsum, int32_t* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] += bsum[i / kScanBlock];
}

static int gridd(int64_t n, int b = 256) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + b - 1) / b, 16384)); }

// out[i] = Σ_{k<i} in[k]; returns the total.  in and out may not alias.
static int scan_exclusive(fem_ctx* c, const int32_t* in, int32_t* out, int64_t n, int64_t* total) {
  const int64_t nb = (n + kScanBlock - 1) / kScanBlock;
  int32_t* bsum = nullptr;
  int64_t* dt = nullptr;
  int rc;
  if ((rc = dev_alloc(c, &bsum, nb)) || (rc = dev_alloc(c, &dt, 1))) { dev_free(&bsum); return rc; }
  k_scan_blocks<<<(unsigned)nb, 256, 0, c->stream>>>(n, in, out, bsum);
  k_scan_sums<<<1, 256, 0, c->stream>>>(nb, bsum, dt);
  k_scan_add<<<gridd(n), 256, 0, c->stream>>>(n, bsum, out);
  FEM_HIP(c, hipMemcpyAsync(total, dt, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  dev_free(&bsum);
  dev_free(&dt);
  return FEM_OK;
}

// ---- lists item -> cells (counting sort; the order inside a list is not used) --------------
__global__ void k_count_items(int64_t m, const int32_t* __restrict__ item, int32_t* __restrict__ cnt) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[item[i]], 1);
}

__global__ void k_fill_items(int64_t m, int per, const int32_t* __restrict__ item, int32_t* __restrict__ cursor,
                             int32_t* __restrict__ list) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    list[atomicAdd(&cursor[item[i]], 1)] = (int32_t)(i / per);
}

// ptr[0..nitems] and list of the cells holding each item (item[c * per + k], k < per).
static int build_lists(fem_ctx* c, int64_t ncells, int per, const int32_t* item, int64_t nitems, int32_t** ptr,
                       int32_t** list) {
  const int64_t m = ncells * per;
  int32_t* cnt = nullptr;
  int rc;
  if ((rc = dev_alloc(c, &cnt, nitems + 1)) || (rc = dev_alloc(c, ptr, nitems + 1)) || (rc = dev_alloc(c, list, m))) {
    dev_free(&cnt);
    return rc;
  }
  FEM_HIP(c, hipMemsetAsync(cnt, 0, sizeof(int32_t) * (nitems + 1), c->stream));
  k_count_items<<<gridd(m), 256, 0, c->stream>>>(m, item, cnt);
  int64_t tot = 0;
  if ((rc = scan_exclusive(c, cnt, *ptr, nitems + 1, &tot))) { dev_free(&cnt); return rc; }
  FEM_HIP(c, hipMemcpyAsync(cnt, *ptr, sizeof(int32_t) * (nitems + 1), hipMemcpyDeviceToDevice, c->stream));
  k_fill_items<<<gridd(m), 256, 0, c->stream>>>(m, per, item, cnt, *list);
  FEM_HIP(c, hipGetLastError());
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  dev_free(&cnt);
  return FEM_OK;
}

// ---- dof distribution ---------------------------------------------------------------------
// Reference entities (0-based local vertices), the host restatement's tables
// (src/Grid/grid.jl:218-229 edges and faces of Hexahedron; Quadrilateral edges).
__constant__ int8_t c_hex_edge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                         {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
__constant__ int8_t c_hex_face[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5},
                                        {2, 3, 7, 6}, {0, 4, 7, 3}, {4, 5, 6, 7}};
__constant__ int8_t c_quad_edge[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};

// DIM 2 / 3, ORDER 1 / 2.  Slots: NV vertices, then (order 2) NE edges, then (3D) NF faces, then
// one interior slot (the quad's face, the hex's volume), never shared.
template <int DIM, int ORDER>
struct Slots {
  static constexpr int NV = DIM == 3 ? 8 : 4;
  static constexpr int NE = ORDER == 2 ? (DIM == 3 ? 12 : 4) : 0;
  static constexpr int NF = ORDER == 2 && DIM == 3 ? 6 : 0;
  static constexpr int NI = ORDER == 2 ? 1 : 0;
  static constexpr int S = NV + NE + NF + NI;
};

template <int NV>
__device__ __forceinline__ bool holds(const int32_t* __restrict__ cn, int32_t node) {
  bool h = false;
#pragma unroll
  for (int k = 0; k < NV; ++k) h |= cn[k] == node;
  return h;
}

template <int DIM, int ORDER>
__global__ __launch_bounds__(256) void k_first_visit(int64_t ncells, const int32_t* __restrict__ cnodes,
                                                     const int32_t* __restrict__ n2c_ptr,
                                                     const int32_t* __restrict__ n2c, int32_t* __restrict__ fpos,
                                                     int32_t* __restrict__ mark) {
  using SL = Slots<DIM, ORDER>;
  constexpr int NV = SL::NV, S = SL::S;
  const int64_t total = ncells * S;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = i / S;
    const int s = (int)(i - c * S);
    const int32_t* cn = cnodes + c * NV;
    int64_t fc = c;
    int fs = s;
    if (s < NV) {   // vertex: the smallest cell holding the node, and the node's slot there
      const int32_t node = cn[s];
      int32_t best = INT_MAX;
      for (int32_t e = n2c_ptr[node]; e < n2c_ptr[node + 1]; ++e) best = min(best, n2c[e]);
      fc = best;
      const int32_t* fn = cnodes + fc * NV;
      for (int k = 0; k < NV; ++k)
        if (fn[k] == node) fs = k;
    } else if (s < NV + SL::NE) {   // edge {a, b}: the smallest cell holding both
      const int k = s - NV;
      const int32_t a = cn[DIM == 3 ? c_hex_edge[k][0] : c_quad_edge[k][0]];
      const int32_t b = cn[DIM == 3 ? c_hex_edge[k][1] : c_quad_edge[k][1]];
      int32_t best = INT_MAX;
      for (int32_t e = n2c_ptr[a]; e < n2c_ptr[a + 1]; ++e) {
        const int32_t x = n2c[e];
        if (x < best && holds<NV>(cnodes + (int64_t)x * NV, b)) best = x;
      }
      fc = best;
      const int32_t* fn = cnodes + fc * NV;
      for (int q = 0; q < SL::NE; ++q) {
        const int32_t u = fn[DIM == 3 ? c_hex_edge[q][0] : c_quad_edge[q][0]];
        const int32_t v = fn[DIM == 3 ? c_hex_edge[q][1] : c_quad_edge[q][1]];
        if ((u == a && v == b) || (u == b && v == a)) fs = NV + q;
      }
    } else if (s < NV + SL::NE + SL::NF) {   // face: the smallest cell holding its four nodes
      const int k = s - NV - SL::NE;
      int32_t fnode[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) fnode[q] = cn[c_hex_face[k][q]];
      int32_t best = INT_MAX;
      for (int32_t e = n2c_ptr[fnode[0]]; e < n2c_ptr[fnode[0] + 1]; ++e) {
        const int32_t x = n2c[e];
        if (x < best && holds<NV>(cnodes + (int64_t)x * NV, fnode[1]) &&
            holds<NV>(cnodes + (int64_t)x * NV, fnode[2]) && holds<NV>(cnodes + (int64_t)x * NV, fnode[3]))
          best = x;
      }
      fc = best;
      const int32_t* fn = cnodes + fc * NV;
      for (int q = 0; q < SL::NF; ++q) {
        bool all = true;
#pragma unroll
        for (int t = 0; t < 4; ++t) all = all && holds<4>(fnode, fn[c_hex_face[q][t]]);
        if (all) fs = NV + SL::NE + q;
      }
    }
    const int64_t p = fc * S + fs;
    fpos[i] = (int32_t)p;
    mark[i] = p == i ? 1 : 0;
  }
}

__global__ void k_cell_dofs(int64_t m, int vdim, const int32_t* __restrict__ fpos, const int32_t* __restrict__ rank,
                            int32_t* __restrict__ cd) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t ent = rank[fpos[i]];
    for (int d = 0; d < vdim; ++d) cd[i * vdim + d] = ent * vdim + d;
  }
}

// ---- pattern ------------------------------------------------------------------------------
// One wave per column j: the dofs of the cells holding j, sorted in LDS; COUNT: cnt[j] = the
// number of distinct ones; otherwise they are written from rowval[colptr[j]].  *err = 1 when a
// column has more than kMaxCand candidates.
template <bool COUNT>
__global__ __launch_bounds__(64) void k_pattern_cols(int64_t n, int ndpc, const int32_t* __restrict__ cd,
                                                     const int32_t* __restrict__ d2c_ptr,
                                                     const int32_t* __restrict__ d2c, int32_t* __restrict__ cnt,
                                                     const int32_t* __restrict__ colptr, int32_t* __restrict__ rowval,
                                                     int* __restrict__ err) {
  __shared__ int32_t v[kMaxCand];
  const int lane = threadIdx.x;
  for (int64_t j = blockIdx.x; j < n; j += gridDim.x) {
    const int32_t e0 = d2c_ptr[j], e1 = d2c_ptr[j + 1];
    const int m = (e1 - e0) * ndpc;
    if (m > kMaxCand) {
      if (lane == 0) *err = 1;
      continue;
    }
    int np2 = 64;
    while (np2 < m) np2 <<= 1;
    for (int t = lane; t < np2; t += 64) {
      int32_t r = INT_MAX;
      if (t < m) {
        const int q = t / ndpc;
        r = cd[(int64_t)d2c[e0 + q] * ndpc + (t - q * ndpc)];
      }
      v[t] = r;
    }
    __syncthreads();
    for (int k = 2; k <= np2; k <<= 1)   // bitonic sort, ascending
      for (int h = k >> 1; h > 0; h >>= 1) {
        for (int t = lane; t < np2; t += 64) {
          const int p = t ^ h;
          if (p > t) {
            const int32_t a = v[t], b = v[p];
            const bool up = (t & k) == 0;
            if ((a > b) == up) { v[t] = b; v[p] = a; }
          }
        }
        __syncthreads();
      }
    int32_t base = COUNT ? 0 : colptr[j];
    for (int t0 = 0; t0 < np2; t0 += 64) {
      const int t = t0 + lane;
      const bool first = v[t] != INT_MAX && (t == 0 || v[t] != v[t - 1]);
      const uint64_t bal = __ballot(first);
      if (!COUNT && first) rowval[base + __popcll(bal & ((1ull << lane) - 1))] = v[t];
      base += __popcll(bal);
    }
    if (COUNT && lane == 0) cnt[j] = base;
    __syncthreads();
  }
}

__global__ void k_to_1based64(int64_t m, const int32_t* __restrict__ a, int64_t* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (int64_t)a[i] + 1;
}

}  // namespace fem

using namespace fem;

extern "C" {

int32_t fem_dofs_distribute(fem_ctx* c, int32_t order, int32_t vdim, int64_t* ndofs_out, int32_t* ndpc_out) {
  if (!c || !ndofs_out || !ndpc_out || (order != 1 && order != 2) || vdim < 1 || vdim > 3)
    return set_error(c, FEM_ERR_ARG, "fem_dofs_distribute: Lagrange order 1 or 2, 1 to 3 components");
  if (c->ncells == 0) return set_error(c, FEM_ERR_STATE, "upload the mesh first");
  const bool hex = c->npc == 8 && c->sdim == 3, quad = c->npc == 4 && c->sdim == 2;
  if (!hex && !quad) return set_error(c, FEM_ERR_ARG, "fem_dofs_distribute: quadrilateral or hexahedral cells");
  FEM_HIP(c, hipSetDevice(c->device));
  const int S = hex ? (order == 2 ? 27 : 8) : (order == 2 ? 9 : 4);
  const int64_t nc = c->ncells, m = nc * S;
  if (m * vdim >= INT32_MAX) return set_error(c, FEM_ERR_ARG, "mesh too large for int32 dof tables");
  int32_t *cn = nullptr, *n2c_ptr = nullptr, *n2c = nullptr, *fpos = nullptr, *mark = nullptr, *rank = nullptr;
  int32_t* cd = nullptr;
  auto cleanup = [&]() {
    dev_free(&cn); dev_free(&n2c_ptr); dev_free(&n2c); dev_free(&fpos); dev_free(&mark); dev_free(&rank);
    dev_free(&cd);
  };
  int rc;
  if ((rc = dev_alloc(c, &cn, nc * c->npc))) { cleanup(); return rc; }
  FEM_HIP(c, hipMemcpyAsync(cn, c->h_cell_nodes.data(), 4 * nc * c->npc, hipMemcpyHostToDevice, c->stream));
  if ((rc = build_lists(c, nc, c->npc, cn, c->nnodes, &n2c_ptr, &n2c))) { cleanup(); return rc; }
  if ((rc = dev_alloc(c, &fpos, m)) || (rc = dev_alloc(c, &mark, m)) || (rc = dev_alloc(c, &rank, m)) ||
      (rc = dev_alloc(c, &cd, m * vdim))) {
    cleanup();
    return rc;
  }
  if (hex && order == 2)
    k_first_visit<3, 2><<<gridd(m), 256, 0, c->stream>>>(nc, cn, n2c_ptr, n2c, fpos, mark);
  else if (hex)
    k_first_visit<3, 1><<<gridd(m), 256, 0, c->stream>>>(nc, cn, n2c_ptr, n2c, fpos, mark);
  else if (order == 2)
    k_first_visit<2, 2><<<gridd(m), 256, 0, c->stream>>>(nc, cn, n2c_ptr, n2c, fpos, mark);
  else
    k_first_visit<2, 1><<<gridd(m), 256, 0, c->stream>>>(nc, cn, n2c_ptr, n2c, fpos, mark);
  FEM_HIP(c, hipGetLastError());
  int64_t nent = 0;
  if ((rc = scan_exclusive(c, mark, rank, m, &nent))) { cleanup(); return rc; }
  k_cell_dofs<<<gridd(m), 256, 0, c->stream>>>(m, vdim, fpos, rank, cd);
  FEM_HIP(c, hipGetLastError());
  // the context's dofs, as fem_dofs_upload sets them (host copy included: host-side table
  // builders and error reports read it)
  const int ndpc = S * vdim;
  c->h_cell_dofs.resize(m * vdim);
  FEM_HIP(c, hipMemcpyAsync(c->h_cell_dofs.data(), cd, 4 * m * vdim, hipMemcpyDeviceToHost, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  cleanup();
  const int64_t ndofs = nent * vdim;
  if ((rc = dev_alloc(c, &c->d_f, ndofs)) || (rc = dev_alloc(c, &c->d_u, ndofs))) return rc;
  FEM_HIP(c, hipMemsetAsync(c->d_f, 0, sizeof(double) * ndofs, c->stream));
  FEM_HIP(c, hipMemsetAsync(c->d_u, 0, sizeof(double) * ndofs, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  c->ndofs = ndofs;
  c->ndpc = ndpc;
  free_solver(c);
  free_q2_tables(c);
  dev_free(&c->d_d2c_ptr); dev_free(&c->d_d2c); dev_free(&c->d_gcdofs);
  c->ea_kind = 0;
  c->tables_ready = false;
  c->gather_ready = false;
  c->affine_mesh = -1;
  *ndofs_out = ndofs;
  *ndpc_out = ndpc;
  return FEM_OK;
}

int32_t fem_dofs_download(fem_ctx* c, int64_t* cell_dofs) {
  if (!c || !cell_dofs) return set_error(c, FEM_ERR_ARG, "fem_dofs_download: invalid arguments");
  if (c->h_cell_dofs.empty()) return set_error(c, FEM_ERR_STATE, "no dofs");
  const int64_t m = (int64_t)c->h_cell_dofs.size();
  for (int64_t i = 0; i < m; ++i) cell_dofs[i] = (int64_t)c->h_cell_dofs[i] + 1;
  return FEM_OK;
}

int32_t fem_pattern_build(fem_ctx* c, int64_t* nnz_out) {
  if (!c || !nnz_out) return set_error(c, FEM_ERR_ARG, "fem_pattern_build: invalid arguments");
  if (c->ndofs == 0 || c->h_cell_dofs.empty()) return set_error(c, FEM_ERR_STATE, "distribute or upload the dofs first");
  FEM_HIP(c, hipSetDevice(c->device));
  const int64_t nc = c->ncells, n = c->ndofs;
  const int ndpc = c->ndpc;
  int32_t *cd = nullptr, *d2c_ptr = nullptr, *d2c = nullptr, *cnt = nullptr, *colptr = nullptr, *rowval = nullptr;
  int* derr = nullptr;
  auto cleanup = [&]() {
    dev_free(&cd); dev_free(&d2c_ptr); dev_free(&d2c); dev_free(&cnt); dev_free(&derr);
  };
  int rc;
  if ((rc = dev_alloc(c, &cd, nc * ndpc)) || (rc = dev_alloc(c, &derr, 1))) { cleanup(); return rc; }
  FEM_HIP(c, hipMemcpyAsync(cd, c->h_cell_dofs.data(), 4 * nc * ndpc, hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipMemsetAsync(derr, 0, sizeof(int), c->stream));
  if ((rc = build_lists(c, nc, ndpc, cd, n, &d2c_ptr, &d2c))) { cleanup(); return rc; }
  if ((rc = dev_alloc(c, &cnt, n + 1)) || (rc = dev_alloc(c, &colptr, n + 1))) { cleanup(); return rc; }
  FEM_HIP(c, hipMemsetAsync(cnt, 0, sizeof(int32_t) * (n + 1), c->stream));
  const unsigned gcols = (unsigned)std::min<int64_t>(n, 262144);
  k_pattern_cols<true><<<gcols, 64, 0, c->stream>>>(n, ndpc, cd, d2c_ptr, d2c, cnt, nullptr, nullptr, derr);
  FEM_HIP(c, hipGetLastError());
  int herr = 0;
  FEM_HIP(c, hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  if (herr) {
    cleanup();
    dev_free(&colptr);
    return set_error(c, FEM_ERR_ARG, "fem_pattern_build: a column gathers more than 1024 candidate rows");
  }
  int64_t nnz = 0;
  if ((rc = scan_exclusive(c, cnt, colptr, n + 1, &nnz))) { cleanup(); dev_free(&colptr); return rc; }
  if (nnz >= INT32_MAX) { cleanup(); dev_free(&colptr); return set_error(c, FEM_ERR_ARG, "pattern too large for int32"); }
  if ((rc = dev_alloc(c, &rowval, nnz))) { cleanup(); dev_free(&colptr); return rc; }
  k_pattern_cols<false><<<gcols, 64, 0, c->stream>>>(n, ndpc, cd, d2c_ptr, d2c, nullptr, colptr, rowval, derr);
  FEM_HIP(c, hipGetLastError());
  // install it as fem_pattern_upload does
  std::vector<int32_t> hcp(n + 1);
  FEM_HIP(c, hipMemcpyAsync(hcp.data(), colptr, 4 * (n + 1), hipMemcpyDeviceToHost, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  cleanup();
  dev_free(&c->d_colptr);
  dev_free(&c->d_rowval);
  c->d_colptr = colptr;
  c->d_rowval = rowval;
  if ((rc = dev_alloc(c, &c->d_nzval, nnz)) || (rc = dev_alloc(c, &c->d_diag, n))) return rc;
  FEM_HIP(c, hipMemsetAsync(c->d_nzval, 0, sizeof(double) * nnz, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  if ((rc = pattern_finish(c, n, nnz, hcp.data()))) return rc;
  *nnz_out = nnz;
  return FEM_OK;
}

int32_t fem_pattern_download(fem_ctx* c, int64_t* colptr, int64_t* rowval) {
  if (!c || !colptr || !rowval) return set_error(c, FEM_ERR_ARG, "fem_pattern_download: invalid arguments");
  if (c->nnz == 0) return set_error(c, FEM_ERR_STATE, "no pattern");
  FEM_HIP(c, hipSetDevice(c->device));
  // 1-based Int64 formed on the device, in chunks of a bounded staging buffer
  const int64_t chunk = std::min<int64_t>(std::max(c->nnz, c->n + 1), int64_t(1) << 26);
  int64_t* st = nullptr;
  int rc = dev_alloc(c, &st, chunk);
  if (rc) return rc;
  auto out = [&](const int32_t* src, int64_t m, int64_t* dst) -> int {
    for (int64_t o = 0; o < m; o += chunk) {
      const int64_t k = std::min(chunk, m - o);
      k_to_1based64<<<gridd(k), 256, 0, c->stream>>>(k, src + o, st);
      FEM_HIP(c, hipMemcpyAsync(dst + o, st, sizeof(int64_t) * k, hipMemcpyDeviceToHost, c->stream));
      FEM_HIP(c, hipStreamSynchronize(c->stream));
    }
    return FEM_OK;
  };
  if ((rc = out(c->d_colptr, c->n + 1, colptr)) || (rc = out(c->d_rowval, c->nnz, rowval))) {
    dev_free(&st);
    return rc;
  }
  dev_free(&st);
  return FEM_OK;
}

}  // extern "C"
