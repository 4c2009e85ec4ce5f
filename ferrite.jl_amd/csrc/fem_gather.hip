// Two-phase, owner-computes assembly of the scalar Q1 Poisson element (hex and quad).
//
// Replaces the same cell loop as the colored scatter (threaded_assembly.jl:232-265:
// reinit! + element routine + assemble!), but writes every K.nzval and f entry exactly
// once instead of read-modify-writing it once per incident cell:
//
//   phase 1  k_elem_poisson   one lane per cell (cell-id order): gather the cell's node
//                             coordinates and k, reinit! + quadrature loop, store the strict
//                             upper triangle of Ke and fe as planes Ke[t][cell], fe[a][cell]
//                             (unit-stride stores across a wave).  The diagonal is not
//                             stored: every row of a Lagrange Poisson Ke sums to zero
//                             (∇ of the partition of unity Σ_b N_b = 1 vanishes), so
//                             Ke[b,b] = −Σ_{a≠b} Ke[a,b], which phase 2 forms from the values
//                             it loads anyway (equal to the directly integrated diagonal to
//                             rounding).
//   phase 2  k_gather_cols    one lane per column j of K (= dof j): for every cell that
//                             holds dof j (ascending cell id) and every local row a, add
//                             Ke[a, b_j] into the slot of row dofs[a] of column j, kept in
//                             LDS; then the block writes its contiguous range of nzval (and
//                             f) with unit-stride stores.
//
// Each nzval entry is summed in ascending cell order — the order of Ferrite's serial
// loop — so the result is deterministic and independent of launch shape and colouring.
// A row missing from column j's pattern is recorded (first cell) like assemble!'s error.
//
// Tables (built once per mesh/dofs/pattern): cell-order node ids, the dof → (cell, local
// index) lists d2c (packed cell*NB + b), and for each list entry and local row a the slot
// of dofs[a] inside column j (one byte; 255 = absent).
#include <algorithm>
#include <cstring>
#include <vector>

#include "fem_element.h"

namespace fem {

// Index of (lo, hi), lo < hi, in the row-major strict upper triangle of an nb x nb matrix.
constexpr __host__ __device__ int off_idx(int lo, int hi, int nb) {
  return lo * (2 * nb - lo - 1) / 2 + (hi - lo - 1);
}

constexpr int kGatherBlock = 128;
constexpr uint8_t kNoSlot = 255;

// GEO: 0 = the uploaded dM tables (any geometry), 1 = trilinear hex (coefficient form of J),
// 2 = affine hex (every cell of the mesh a parallelepiped; constant-J form)
template <typename T, int DIM, int NG, int NB, int NQ, int GEO = 0>
__global__ __launch_bounds__(256) void k_elem_poisson(int64_t ncells, int64_t cbeg, int64_t cend,
                                                      const T* __restrict__ qc, const T* __restrict__ aff,
                                                      const double* __restrict__ X0,
                                                      const double* __restrict__ X1, const double* __restrict__ X2,
                                                      const int32_t* __restrict__ cnodes,
                                                      const double* __restrict__ param, const T* __restrict__ tab,
                                                      T src, T* __restrict__ keb, T* __restrict__ feb,
                                                      DevError* __restrict__ err) {
  constexpr int TRI = NB * (NB + 1) / 2;
  using L = TabLayout<DIM, NG, NB, NQ>;
  constexpr int NOFF = NB * (NB - 1) / 2;
  constexpr int NAFF = 6 * NOFF + NB;
  __shared__ T stab[L::size + 6 * NQ + NAFF];
  if constexpr (GEO == 1) {
    for (int i = threadIdx.x; i < 6 * NQ; i += blockDim.x) stab[L::size + i] = qc[i];
  }
  if constexpr (GEO == 2) {
    for (int i = threadIdx.x; i < NAFF; i += blockDim.x) stab[L::size + 6 * NQ + i] = aff[i];
  }
  stage_table<T, L::size>(stab, tab);
  for (int64_t cell = cbeg + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; cell < cend;
       cell += (int64_t)gridDim.x * blockDim.x) {
    int32_t nd[NG];
    if constexpr (NG == 8) {
      const int4* p = reinterpret_cast<const int4*>(cnodes + cell * 8);
      const int4 u0 = p[0], u1 = p[1];
      nd[0] = u0.x; nd[1] = u0.y; nd[2] = u0.z; nd[3] = u0.w;
      nd[4] = u1.x; nd[5] = u1.y; nd[6] = u1.z; nd[7] = u1.w;
    } else if constexpr (NG == 4) {
      const int4 u0 = *reinterpret_cast<const int4*>(cnodes + cell * 4);
      nd[0] = u0.x; nd[1] = u0.y; nd[2] = u0.z; nd[3] = u0.w;
    } else {
#pragma unroll
      for (int a = 0; a < NG; ++a) nd[a] = cnodes[cell * NG + a];
    }
    T x[NG][DIM];
#pragma unroll
    for (int a = 0; a < NG; ++a) {
      x[a][0] = (T)X0[nd[a]];
      x[a][1] = (T)X1[nd[a]];
      if constexpr (DIM == 3) x[a][2] = (T)X2[nd[a]];
    }
    if constexpr (GEO == 2) {
      // constant-J form, each plane value stored as soon as it is formed (no Ke array live)
      const T* P = stab + L::size + 6 * NQ;
      T A[6];
      const T detJ = affine_metric<T>(x, (T)param[cell], A);
      if (!(detJ > T(0))) atomicMin(&err->detj_cell, (int)cell);
#pragma unroll 4
      for (int t = 0; t < NOFF; ++t) {
        T acc = A[0] * P[6 * t];
#pragma unroll
        for (int m = 1; m < 6; ++m) acc += A[m] * P[6 * t + m];
        keb[t * ncells + cell] = acc;
      }
      const T sd = src * detJ;
#pragma unroll
      for (int a = 0; a < NB; ++a) feb[a * ncells + cell] = sd * P[6 * NOFF + a];
      continue;
    }
    T ke[TRI], fe[NB];
    bool ok;
    if constexpr (GEO == 1) ok = poisson_cell_tri<T, NB, NQ>(x, (T)param[cell], src, stab, stab + L::size, ke, fe);
    else ok = poisson_cell<T, DIM, NG, NB, NQ>(x, (T)param[cell], src, stab, ke, fe);
    if (!ok) atomicMin(&err->detj_cell, (int)cell);
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int b = a + 1; b < NB; ++b) keb[off_idx(a, b, NB) * ncells + cell] = ke[tri_idx(a, b, NB)];
#pragma unroll
    for (int a = 0; a < NB; ++a) feb[a * ncells + cell] = fe[a];
  }
}

// One lane per column; LDS slots acc[k][lane] for the column's entries k < MAXL.  The
// lane walks its column's cells G at a time: the list entries, slot bytes and element
// values of G cells are loaded before any of them is added (loads in flight), then added
// in ascending cell order.  ecol[q] names the column of the block's q-th nzval entry, so
// the coalesced read-in / write-out of the block's nzval range needs no search.
template <typename T, int NB, int MAXL, int G>
__global__ __launch_bounds__(kGatherBlock) void k_gather_cols(
    int64_t n, int64_t ncells, const int32_t* __restrict__ colptr, const int32_t* __restrict__ d2c_ptr,
    const int32_t* __restrict__ d2c, const uint8_t* __restrict__ slot, const T* __restrict__ keb,
    const T* __restrict__ feb, double* __restrict__ nzval, double* __restrict__ f, int fillzero,
    DevError* __restrict__ err) {
  static_assert(kGatherBlock <= 256, "column index of a block entry is one byte");
  __shared__ double acc[MAXL * kGatherBlock];
  __shared__ int32_t scp[kGatherBlock + 1];
  __shared__ uint8_t ecol[MAXL * kGatherBlock];
  const int tid = threadIdx.x;
  // Workgroups are dealt round-robin to the 8 XCDs (each with its own L2): give every XCD
  // one contiguous range of column blocks, so neighbouring columns, which read the same
  // cells' element planes, share an L2.  gridDim.x is a multiple of 8.
  const int64_t per = gridDim.x / 8;
  const int64_t cb = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
  const int64_t j0 = cb * kGatherBlock;
  if (j0 >= n) return;
  const int ncol = (int)min<int64_t>(kGatherBlock, n - j0);
  for (int t = tid; t <= ncol; t += kGatherBlock) scp[t] = colptr[j0 + t];
  __syncthreads();
  const int32_t base0 = scp[0], end0 = scp[ncol];
  if (tid < ncol) {   // every column holds at most MAXL entries (the pattern's longest column)
    const int32_t c0 = scp[tid] - base0, len = scp[tid + 1] - scp[tid];
    for (int k = 0; k < len; ++k) ecol[c0 + k] = (uint8_t)tid;
  }
  if (fillzero) {
#pragma unroll
    for (int k = 0; k < MAXL; ++k) acc[k * kGatherBlock + tid] = 0.0;
  }
  __syncthreads();
  // start values: the current K when accumulating (start_assemble(...; fillzero=false))
  if (!fillzero) {
    for (int32_t q = base0 + tid; q < end0; q += kGatherBlock) {
      const int cl = ecol[q - base0];
      acc[(q - scp[cl]) * kGatherBlock + cl] = nzval[q];
    }
    __syncthreads();
  }
  if (tid < ncol) {
    const int64_t j = j0 + tid;
    double fj = fillzero ? 0.0 : f[j];
    const int32_t e0 = d2c_ptr[j], e1 = d2c_ptr[j + 1];
    for (int32_t eb = e0; eb < e1; eb += G) {
      int32_t pk[G];
      uint8_t sl[G][NB];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int32_t e = eb + g;
        const bool in = e < e1;
        pk[g] = in ? d2c[e] : -1;
        if constexpr (NB == 8) {
          const uint2 u = in ? *reinterpret_cast<const uint2*>(slot + (int64_t)e * 8) : make_uint2(~0u, ~0u);
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            sl[g][a] = (uint8_t)(u.x >> (8 * a));
            sl[g][a + 4] = (uint8_t)(u.y >> (8 * a));
          }
        } else if constexpr (NB == 4) {
          const uint32_t u = in ? *reinterpret_cast<const uint32_t*>(slot + (int64_t)e * 4) : ~0u;
#pragma unroll
          for (int a = 0; a < 4; ++a) sl[g][a] = (uint8_t)(u >> (8 * a));
        } else {
#pragma unroll
          for (int a = 0; a < NB; ++a) sl[g][a] = in ? slot[(int64_t)e * NB + a] : kNoSlot;
        }
      }
      T v[G][NB], fv[G];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const bool in = pk[g] >= 0;
        const int64_t cell = in ? pk[g] / NB : 0;
        const int b = in ? pk[g] - (int)cell * NB : 0;
        T diag = T(0);
#pragma unroll
        for (int a = 0; a < NB; ++a) {
          const int lo = a < b ? a : b, hi = a < b ? b : a;
          v[g][a] = (in && a != b) ? keb[(int64_t)off_idx(lo, hi, NB) * ncells + cell] : T(0);
        }
#pragma unroll
        for (int a = 0; a < NB; ++a) diag -= v[g][a];   // v[g][b] is 0 here; ascending a
#pragma unroll
        for (int a = 0; a < NB; ++a)
          if (a == b) v[g][a] = diag;
        fv[g] = in ? feb[(int64_t)b * ncells + cell] : T(0);
      }
#pragma unroll
      for (int g = 0; g < G; ++g) {
        if (pk[g] < 0) break;
#pragma unroll
        for (int a = 0; a < NB; ++a) {
          if (sl[g][a] != kNoSlot) {
            acc[sl[g][a] * kGatherBlock + tid] += (double)v[g][a];
          } else if (v[g][a] != T(0)) {
            atomicMin(&err->pattern_cell, (int)(pk[g] / NB));
          }
        }
        fj += (double)fv[g];
      }
    }
    f[j] = fj;
  }
  __syncthreads();
  // unit-stride write of the block's nzval range
  for (int32_t q = base0 + tid; q < end0; q += kGatherBlock) {
    const int cl = ecol[q - base0];
    nzval[q] = acc[(q - scp[cl]) * kGatherBlock + cl];
  }
}

// slot[e][a] = position of dofs(cell_e)[a] in column j (0-based within the column), or 255.
template <int NB>
__global__ void k_build_slots(int64_t n, const int32_t* __restrict__ colptr, const int32_t* __restrict__ rowval,
                              const int32_t* __restrict__ d2c_ptr, const int32_t* __restrict__ d2c,
                              const int32_t* __restrict__ cdofs, uint8_t* __restrict__ slot) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c0 = colptr[j], c1 = colptr[j + 1];
    for (int32_t e = d2c_ptr[j]; e < d2c_ptr[j + 1]; ++e) {
      const int64_t cell = d2c[e] / NB;
      for (int a = 0; a < NB; ++a) {
        const int32_t row = cdofs[cell * NB + a];
        int32_t lo = c0, hi = c1;
        while (lo < hi) {
          const int32_t mid = (lo + hi) >> 1;
          if (rowval[mid] < row) lo = mid + 1; else hi = mid;
        }
        const bool found = lo < c1 && rowval[lo] == row;
        slot[(int64_t)e * NB + a] = found ? (uint8_t)(lo - c0) : kNoSlot;
      }
    }
  }
}

static int gridn(int64_t n, int b = 256) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + b - 1) / b, 16384)); }

// Cell-order node table, dof -> cell lists and the slot table (once per mesh/dofs/pattern).
int build_gather_tables(fem_ctx* c) {
  if (c->gather_ready) return FEM_OK;
  const int64_t nc = c->ncells, nd = c->ndofs;
  const int nb = c->ndpc;
  if (nc == 0 || nd == 0 || c->n != nd) return set_error(c, FEM_ERR_STATE, "mesh, dofs and pattern must be uploaded");
  if (nc * nb >= INT32_MAX || nc >= INT32_MAX / nb) return set_error(c, FEM_ERR_ARG, "mesh too large for the gather tables");
  int rc;
  if ((rc = dev_alloc(c, &c->d_gcnodes, nc * c->npc))) return rc;
  FEM_HIP(c, hipMemcpyAsync(c->d_gcnodes, c->h_cell_nodes.data(), 4 * nc * c->npc, hipMemcpyHostToDevice, c->stream));
  // dof -> (cell, local index), ascending cell id per dof (counting sort over the cell loop)
  std::vector<int32_t> ptr(nd + 1, 0), ent(nc * nb);
  for (int64_t i = 0; i < nc * nb; ++i) ++ptr[c->h_cell_dofs[i] + 1];
  for (int64_t j = 0; j < nd; ++j) ptr[j + 1] += ptr[j];
  std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
  for (int64_t cell = 0; cell < nc; ++cell)
    for (int b = 0; b < nb; ++b) {
      const int32_t d = c->h_cell_dofs[cell * nb + b];
      ent[fill[d]++] = (int32_t)(cell * nb + b);
    }
  // cell-order dofs stay resident: the element-assembly solver operator reads them
  if ((rc = dev_alloc(c, &c->d_d2c_ptr, nd + 1)) || (rc = dev_alloc(c, &c->d_d2c, nc * nb)) ||
      (rc = dev_alloc(c, &c->d_gslot, nc * (int64_t)nb * nb)) || (rc = dev_alloc(c, &c->d_gcdofs, nc * nb)))
    return rc;
  int32_t* dcd = c->d_gcdofs;
  FEM_HIP(c, hipMemcpyAsync(c->d_d2c_ptr, ptr.data(), 4 * (nd + 1), hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipMemcpyAsync(c->d_d2c, ent.data(), 4 * nc * nb, hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipMemcpyAsync(dcd, c->h_cell_dofs.data(), 4 * nc * nb, hipMemcpyHostToDevice, c->stream));
  if (nb == 8)
    k_build_slots<8><<<gridn(nd), 256, 0, c->stream>>>(nd, c->d_colptr, c->d_rowval, c->d_d2c_ptr, c->d_d2c, dcd,
                                                        c->d_gslot);
  else if (nb == 4)
    k_build_slots<4><<<gridn(nd), 256, 0, c->stream>>>(nd, c->d_colptr, c->d_rowval, c->d_d2c_ptr, c->d_d2c, dcd,
                                                        c->d_gslot);
  else {
    return set_error(c, FEM_ERR_ARG, "gather assembly supports 4 or 8 dofs per cell");
  }
  FEM_HIP(c, hipGetLastError());
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  c->gather_ready = true;
  return FEM_OK;
}

// flag = 0 when some cell's trilinear map is not a parallelepiped (FP64, tol 1e-14).
__global__ __launch_bounds__(256) void k_check_affine(int64_t ncells, const double* __restrict__ X0,
                                                      const double* __restrict__ X1, const double* __restrict__ X2,
                                                      const int32_t* __restrict__ cnodes, int* __restrict__ flag) {
  for (int64_t cell = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; cell < ncells;
       cell += (int64_t)gridDim.x * blockDim.x) {
    double x[8][3];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int32_t nd = cnodes[cell * 8 + a];
      x[a][0] = X0[nd];
      x[a][1] = X1[nd];
      x[a][2] = X2[nd];
    }
    double Y[8][3];
    trilinear_coeffs<double>(x, Y);
    if (!is_affine<double>(Y, 1e-14)) *flag = 0;   // every writer stores the same 0
  }
}

// Whether every cell is affine (cached until the mesh changes).
static int check_affine_mesh(fem_ctx* c) {
  if (c->affine_mesh >= 0) return FEM_OK;
  int* d = nullptr;
  int rc = dev_alloc(c, &d, 1);
  if (rc) return rc;
  const int one = 1;
  FEM_HIP(c, hipMemcpyAsync(d, &one, sizeof(int), hipMemcpyHostToDevice, c->stream));
  k_check_affine<<<gridn(c->ncells), 256, 0, c->stream>>>(c->ncells, c->d_x[0], c->d_x[1], c->d_x[2], c->d_gcnodes, d);
  int h = 0;
  FEM_HIP(c, hipMemcpyAsync(&h, d, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  dev_free(&d);
  c->affine_mesh = h;
  return FEM_OK;
}

// True when the gather path handles this assembly (scalar Q1 Poisson with short columns).
bool gather_supported(const fem_ctx* c) {
  const bool hex = c->rdim == 3 && c->ngeo == 8 && c->nbf == 8 && c->nqp == 8 && c->npc == 8 && c->sdim == 3;
  const bool quad = c->rdim == 2 && c->ngeo == 4 && c->nbf == 4 && c->nqp == 4 && c->npc == 4 && c->sdim == 2;
  return (hex || quad) && c->vdim == 1 && c->ndpc == c->nbf && c->max_row_nnz <= 32;
}

template <typename T, int DIM, int NG, int NB, int NQ>
static int run_gather(fem_ctx* c, const T* tab, double src, bool fillzero) {
  constexpr int NOFF = NB * (NB - 1) / 2;   // stored planes: the strict upper triangle
  const int64_t nc = c->ncells;
  int rc;
  T*& keb = sizeof(T) == 8 ? reinterpret_cast<T*&>(c->d_keb64) : reinterpret_cast<T*&>(c->d_keb32);
  T*& feb = sizeof(T) == 8 ? reinterpret_cast<T*&>(c->d_feb64) : reinterpret_cast<T*&>(c->d_feb32);
  int64_t& cap = sizeof(T) == 8 ? c->keb64_cells : c->keb32_cells;
  if (!keb || cap != nc) {
    if ((rc = dev_alloc(c, &keb, nc * NOFF)) || (rc = dev_alloc(c, &feb, nc * NB))) return rc;
    cap = nc;
  }
  const int grp = c->tune.gather_group ? c->tune.gather_group : 4;
#define FEM_GATHER(L, GG, NBLK)                                                                                \
  k_gather_cols<T, NB, L, GG><<<NBLK, kGatherBlock, 0, c->stream>>>(c->n, nc, c->d_colptr, c->d_d2c_ptr,       \
                                                                    c->d_d2c, c->d_gslot, keb, feb, c->d_nzval, \
                                                                    c->d_f, fillzero ? 1 : 0, c->d_err)
#define FEM_GATHER_G(L, NBLK)                          \
  do {                                                 \
    if (grp == 1) FEM_GATHER(L, 1, NBLK);              \
    else if (grp == 2) FEM_GATHER(L, 2, NBLK);         \
    else if (grp == 8) FEM_GATHER(L, 8, NBLK);         \
    else FEM_GATHER(L, 4, NBLK);                       \
  } while (0)
  auto gather = [&](int nblk) {
    if (c->max_row_nnz <= 9) FEM_GATHER_G(9, nblk);
    else if (c->max_row_nnz <= 27) FEM_GATHER_G(27, nblk);
    else FEM_GATHER_G(32, nblk);
  };
  const T* qc = sizeof(T) == 8 ? (const T*)c->d_qc : (const T*)c->d_qcf;
  const T* aff = sizeof(T) == 8 ? (const T*)c->d_aff : (const T*)c->d_afff;
  int geo = 0;
  if constexpr (DIM == 3 && NG == 8) {
    if (c->trilinear) {
      geo = 1;
      if (c->tune.affine != -1) {
        if ((rc = check_affine_mesh(c))) return rc;
        if (c->affine_mesh == 1) geo = 2;
      }
    }
  }
  auto elem = [&](int64_t c0, int64_t c1, hipStream_t st) {
    if constexpr (DIM == 3 && NG == 8) {
      if (geo == 2) {
        k_elem_poisson<T, DIM, NG, NB, NQ, 2><<<gridn(c1 - c0), 256, 0, st>>>(
            nc, c0, c1, qc, aff, c->d_x[0], c->d_x[1], c->d_x[2], c->d_gcnodes, c->d_param_cell, tab, (T)src, keb, feb,
            c->d_err);
        return;
      }
      if (geo == 1) {
        k_elem_poisson<T, DIM, NG, NB, NQ, 1><<<gridn(c1 - c0), 256, 0, st>>>(
            nc, c0, c1, qc, aff, c->d_x[0], c->d_x[1], c->d_x[2], c->d_gcnodes, c->d_param_cell, tab, (T)src, keb, feb,
            c->d_err);
        return;
      }
    }
    k_elem_poisson<T, DIM, NG, NB, NQ><<<gridn(c1 - c0), 256, 0, st>>>(
        nc, c0, c1, qc, nullptr, c->d_x[0], c->d_x[1], c->d_x[2], c->d_gcnodes, c->d_param_cell, tab, (T)src, keb, feb,
        c->d_err);
  };
  FEM_HIP(c, hipEventRecord(c->ev0, c->stream));
  elem(0, nc, c->stream);
  FEM_HIP(c, hipEventRecord(c->evm, c->stream));
  gather((int)(((c->n + kGatherBlock - 1) / kGatherBlock + 7) / 8 * 8));
  c->timings.assemble_launches = 2;
#undef FEM_GATHER_G
#undef FEM_GATHER
  FEM_HIP(c, hipGetLastError());
  FEM_HIP(c, hipEventRecord(c->ev1, c->stream));
  c->timings.assemble_kernel_ms = event_ms(c);
  float e1 = 0.f, e2 = 0.f;
  FEM_HIP(c, hipEventElapsedTime(&e1, c->ev0, c->evm));
  FEM_HIP(c, hipEventElapsedTime(&e2, c->evm, c->ev1));
  c->timings.element_ms = e1;
  c->timings.scatter_ms = e2;
  return FEM_OK;
}

// Entry from fem_assemble (tables uploaded by the caller, error word reset).
int gather_poisson(fem_ctx* c, bool mixed, double src, bool fillzero, const double* tab64, const float* tab32) {
  int rc = build_gather_tables(c);
  if (rc) return rc;
  const bool hex = c->rdim == 3;
  if (hex) {
    return mixed ? run_gather<float, 3, 8, 8, 8>(c, tab32, src, fillzero)
                 : run_gather<double, 3, 8, 8, 8>(c, tab64, src, fillzero);
  }
  return mixed ? run_gather<float, 2, 4, 4, 4>(c, tab32, src, fillzero)
               : run_gather<double, 2, 4, 4, 4>(c, tab64, src, fillzero);
}

}  // namespace fem
