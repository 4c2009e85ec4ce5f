// Linear elasticity element kernels (vector Lagrange, small strain, isotropic C):
//   Ke[(a,i),(b,j)] += E_e ∫ ε(N_a e_i) : C(ν) : ε(N_b e_j) dΩ
//                    = E_e Σ_q dV_q (λ g_a,i g_b,j + μ δ_ij g_a·g_b + μ g_a,j g_b,i)
// with g = ∇N at the quadrature point, λ = ν/((1+ν)(1−2ν)), μ = 1/(2(1+ν)) for E = 1
// (docs/src/literate-howto/threaded_assembly.jl:105-132 element routine and material,
// plane strain in 2D).  Local dofs are node-major, component-minor (ip^dim,
// src/interpolations.jl:1815-1825).  Scatter: same colored, atomic-free CSC
// read-modify-write as the Poisson kernel.
//
// Q2 hex (27 nodes, 81 dofs, 3x3x3 Gauss): one wavefront per cell.  The 81x81 Ke is
// assembled from six 27x27 products P_ij = A_i B_j (i <= j) with A_i[a][q] = g_q[a,i]
// and B_j[q][b] = dV_q g_q[b,j]:
//   Ke_ij = E (λ P_ij + μ P_ji + μ δ_ij (P_00 + P_11 + P_22)),   P_ji = P_ijᵀ.  That is 6·27³ multiply-adds per cell instead of the 81·81·162 of a
// dense BᵀDB.  FP32 products run on MFMA (v_mfma_f32_16x16x4f32, 27 padded to 32/28);
// the FP64 path uses register-blocked VALU FMAs.
#include "fem_internal.h"

#include <algorithm>

namespace fem {

template <int DIM, int NG, int NB, int NQ>
struct TabL {
  static constexpr int oN = 0;
  static constexpr int odN = oN + NQ * NB;
  static constexpr int odM = odN + NQ * NB * DIM;
  static constexpr int ow = odM + NQ * NG * DIM;
  static constexpr int size = ow + NQ;
};

template <typename T, int DIM>
__device__ __forceinline__ T det_inv_e(const T (&J)[DIM][DIM], T (&Ji)[DIM][DIM]) {
  if constexpr (DIM == 2) {
    const T d = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    const T id = T(1) / d;
    Ji[0][0] = J[1][1] * id;  Ji[0][1] = -J[0][1] * id;
    Ji[1][0] = -J[1][0] * id; Ji[1][1] = J[0][0] * id;
    return d;
  } else {
    const T a00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const T a01 = J[0][2] * J[2][1] - J[0][1] * J[2][2];
    const T a02 = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    const T d = J[0][0] * a00 + J[1][0] * a01 + J[2][0] * a02;
    const T id = T(1) / d;
    Ji[0][0] = a00 * id; Ji[0][1] = a01 * id; Ji[0][2] = a02 * id;
    Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    return d;
  }
}

constexpr __host__ __device__ int tri_e(int lo, int hi, int nb) { return lo * (2 * nb - lo + 1) / 2 + (hi - lo); }

// ---------------------------------------------------------------------------
// Small vector elements (Q1 quad^2: 8 dofs): one lane per cell, Ke in registers.
// ---------------------------------------------------------------------------
template <typename T, int DIM, int NG, int NQ, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_elast_small(
    int64_t s0, int64_t s1, const double* __restrict__ X0, const double* __restrict__ X1,
    const double* __restrict__ X2, const int32_t* __restrict__ cnodes, const int32_t* __restrict__ cdofs,
    const int32_t* __restrict__ cnz, const int32_t* __restrict__ ccell, const double* __restrict__ cparam,
    const T* __restrict__ tab, T lam, T mu, double* __restrict__ nzval, DevError* __restrict__ err,
    T* __restrict__ planes, int64_t ncells) {
  constexpr int NB = NG * DIM;  // dofs (function ip == geometry ip, vectorised)
  constexpr int TRI = NB * (NB + 1) / 2;
  using L = TabL<DIM, NG, NG, NQ>;
  __shared__ T s_ke[BLOCK * TRI];
  const int64_t sb = s0 + (int64_t)blockIdx.x * BLOCK;
  const int ncell = (int)min<int64_t>(BLOCK, s1 - sb);
  const int lc = threadIdx.x;
  if (lc < ncell) {
    const int64_t s = sb + lc;
    T x[NG][DIM];
#pragma unroll
    for (int a = 0; a < NG; ++a) {
      const int32_t nd = cnodes[s * NG + a];
      x[a][0] = (T)X0[nd];
      x[a][1] = (T)X1[nd];
      if constexpr (DIM == 3) x[a][2] = (T)X2[nd];
    }
    const T E = (T)cparam[s];
    T ke[TRI];
#pragma unroll
    for (int i = 0; i < TRI; ++i) ke[i] = T(0);
    bool bad = false;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      T J[DIM][DIM];
#pragma unroll
      for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j) J[i][j] = T(0);
#pragma unroll
      for (int a = 0; a < NG; ++a)
#pragma unroll
        for (int i = 0; i < DIM; ++i)
#pragma unroll
          for (int j = 0; j < DIM; ++j) J[i][j] += x[a][i] * tab[L::odM + (q * NG + a) * DIM + j];
      T Ji[DIM][DIM];
      const T detJ = det_inv_e<T, DIM>(J, Ji);
      bad |= !(detJ > T(0));
      const T dV = detJ * tab[L::ow + q] * E;
      T g[NG][DIM];
#pragma unroll
      for (int b = 0; b < NG; ++b)
#pragma unroll
        for (int j = 0; j < DIM; ++j) {
          T acc = T(0);
#pragma unroll
          for (int i = 0; i < DIM; ++i) acc += tab[L::odN + (q * NG + b) * DIM + i] * Ji[i][j];
          g[b][j] = acc;
        }
#pragma unroll
      for (int I = 0; I < NB; ++I) {
#pragma unroll
        for (int J2 = I; J2 < NB; ++J2) {
          const int a = I / DIM, i = I % DIM, b = J2 / DIM, j = J2 % DIM;
          T gg = g[a][0] * g[b][0];
#pragma unroll
          for (int k = 1; k < DIM; ++k) gg += g[a][k] * g[b][k];
          T v = lam * g[a][i] * g[b][j] + mu * g[a][j] * g[b][i];
          if (i == j) v += mu * gg;
          ke[tri_e(I, J2, NB)] += v * dV;
        }
      }
    }
    if (bad) atomicMin(&err->detj_cell, ccell[s]);
#pragma unroll
    for (int i = 0; i < TRI; ++i) s_ke[lc * TRI + i] = ke[i];
    if (planes) {   // the element matrix for the element-assembly solver operator: Ke[t][cell]
      const int64_t cell = ccell[s];
#pragma unroll
      for (int i = 0; i < TRI; ++i) planes[i * ncells + cell] = ke[i];
    }
  }
  __syncthreads();
  constexpr int E2 = NB * NB;
  constexpr int CH = 16;
  const int32_t* __restrict__ nzt = cnz + sb * E2;
  const int tot = ncell * E2;
  for (int base = 0; base < tot; base += BLOCK * CH) {
    int32_t pp[CH];
    double old[CH];
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int t = base + k * BLOCK + threadIdx.x;
      pp[k] = t < tot ? nzt[t] : -1;
    }
#pragma unroll
    for (int k = 0; k < CH; ++k) old[k] = pp[k] >= 0 ? nzval[pp[k]] : 0.0;
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int t = base + k * BLOCK + threadIdx.x;
      if (t < tot) {
        const int lcell = t / E2;
        const int e = t - lcell * E2;
        const int b = e / NB, a = e - b * NB;
        const T v = s_ke[lcell * TRI + tri_e(a < b ? a : b, a < b ? b : a, NB)];
        if (pp[k] >= 0) {
          if (v != T(0)) nzval[pp[k]] = old[k] + (double)v;
        } else if (v != T(0)) {
          atomicMin(&err->pattern_cell, ccell[sb + lcell]);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Q2 hex^3: one wavefront (64 lanes) per cell.
// ---------------------------------------------------------------------------
constexpr int kQN = 27;   // nodes / scalar basis functions
constexpr int kQQ = 27;   // quadrature points
constexpr int kQD = 81;   // dofs
constexpr int kKP = 28;   // padded K (quadrature) dimension, multiple of 4
constexpr int kMP = 32;   // padded M/N dimension
constexpr int kPS = 28;   // row stride of the stored products

typedef float floatx4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Q2Smem {
  // A_i[a][q] and B_j[q][b]; the products P_ij alias them in the FP32 (MFMA) path, whose
  // accumulators stay in registers until every product is done, and follow them in the
  // FP64 path, which writes each product as soon as it is formed.
  static constexpr int nA = 3 * kMP * kKP;
  static constexpr int nB = 3 * kKP * kMP;
  static constexpr int nP = 6 * kQN * kPS;
  static constexpr int oP = sizeof(T) == 4 ? 0 : nA + nB;
  static constexpr int n = sizeof(T) == 4 ? ((nA + nB) > nP ? (nA + nB) : nP) : nA + nB + nP;
};

__device__ __forceinline__ int pidx(int i, int j) {  // (i <= j) -> 0..5
  return i == 0 ? j : (i == 1 ? 2 + j : 5);
}

// Stored element matrices of the owner-computes path: per cell and local node b the 81 x 3
// column block Ke[(a, i), (b, j)] as [(i, j)][a] with 28 rows a (row 27 is padding), so the
// element kernel stores runs of four a straight from its MFMA accumulators and the gather
// reads one contiguous 252-value block per (cell, b).
constexpr int kKS = 9 * kKP;        // 252 values per (cell, b)
constexpr int kKC = kQN * kKS;      // 6804 values per cell

// Colored scatter: s = colour slot s0 + block (cnodes/cparam in colour order, ccell maps the
// slot to its cell).
template <typename T>
__global__ __launch_bounds__(64) void k_elast_q2hex(
    int64_t s0, const double* __restrict__ X0, const double* __restrict__ X1, const double* __restrict__ X2,
    const int32_t* __restrict__ cnodes, const int32_t* __restrict__ cnz, const int32_t* __restrict__ ccell,
    const double* __restrict__ cparam, const T* __restrict__ tab, T lam, T mu, double* __restrict__ nzval,
    DevError* __restrict__ err) {
  using L = TabL<3, 8, kQN, kQQ>;
  __shared__ T sm[Q2Smem<T>::n];
  T* sA = sm;                        // [3][kMP][kKP]
  T* sB = sm + Q2Smem<T>::nA;        // [3][kKP][kMP]
  const int lane = threadIdx.x;
  const int64_t s = s0 + blockIdx.x;

  for (int t = lane; t < Q2Smem<T>::nA + Q2Smem<T>::nB; t += 64) sm[t] = T(0);
  T x[8][3];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int32_t nd = cnodes[s * 8 + a];
    x[a][0] = (T)X0[nd];
    x[a][1] = (T)X1[nd];
    x[a][2] = (T)X2[nd];
  }
  __syncthreads();
  if (lane < kQQ) {
    const int q = lane;
    T J[3][3] = {{T(0), T(0), T(0)}, {T(0), T(0), T(0)}, {T(0), T(0), T(0)}};
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) J[i][j] += x[a][i] * tab[L::odM + (q * 8 + a) * 3 + j];
    T Ji[3][3];
    const T detJ = det_inv_e<T, 3>(J, Ji);
    if (!(detJ > T(0))) atomicMin(&err->detj_cell, ccell[s]);
    const T dV = detJ * tab[L::ow + q];
    for (int a = 0; a < kQN; ++a) {
      const T* dn = tab + L::odN + (q * kQN + a) * 3;
      const T d0 = dn[0], d1 = dn[1], d2 = dn[2];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const T gi = d0 * Ji[0][i] + d1 * Ji[1][i] + d2 * Ji[2][i];
        sA[(i * kMP + a) * kKP + q] = gi;
        sB[(i * kKP + q) * kMP + a] = dV * gi;
      }
    }
  }
  __syncthreads();

  // P_ij = A_i B_j for the six (i <= j) pairs; results go to registers, then to LDS.
  if constexpr (sizeof(T) == 4) {
    floatx4 acc[6][2][2];
    const int r = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int pi = 0; pi < 6; ++pi) {
      const int i = pi < 3 ? 0 : (pi < 5 ? 1 : 2);
      const int j = pi < 3 ? pi : (pi < 5 ? pi - 2 : 2);
#pragma unroll
      for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) {
          floatx4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int kk = 0; kk < kKP / 4; ++kk) {
            const float av = sA[(i * kMP + 16 * ti + r) * kKP + 4 * kk + kq];
            const float bv = sB[(j * kKP + 4 * kk + kq) * kMP + 16 * tj + r];
            c = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, c, 0, 0, 0);
          }
          acc[pi][ti][tj] = c;
        }
    }
    __syncthreads();
    T* sP = sm;
#pragma unroll
    for (int pi = 0; pi < 6; ++pi)
#pragma unroll
      for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int row = 16 * ti + 4 * kq + e, col = 16 * tj + r;
            if (row < kQN && col < kQN) sP[(pi * kQN + row) * kPS + col] = acc[pi][ti][tj][e];
          }
  } else {
    // FP64: lane owns a 4x4 output tile of the padded 32x32 product (8x8 lanes).
    const int tr = (lane >> 3) * 4, tc = (lane & 7) * 4;
    T* sP = sm + Q2Smem<T>::oP;
    for (int pi = 0; pi < 6; ++pi) {
      const int i = pi < 3 ? 0 : (pi < 5 ? 1 : 2);
      const int j = pi < 3 ? pi : (pi < 5 ? pi - 2 : 2);
      T acc[4][4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = T(0);
      for (int q = 0; q < kQQ; ++q) {
        T av[4], bv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) av[u] = sA[(i * kMP + tr + u) * kKP + q];
#pragma unroll
        for (int v = 0; v < 4; ++v) bv[v] = sB[(j * kKP + q) * kMP + tc + v];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[u][v] += av[u] * bv[v];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = tr + u, col = tc + v;
          if (row < kQN && col < kQN) sP[(pi * kQN + row) * kPS + col] = acc[u][v];
        }
    }
  }
  __syncthreads();

  const T* sP = sm + Q2Smem<T>::oP;
  const T E = (T)cparam[s];
  const T lE = lam * E, mE = mu * E;
  // Ke entries in CSC order (entry e = J*81 + I: column J, row I) and the scatter.
  const int32_t* __restrict__ nzt = cnz + s * (kQD * kQD);
  constexpr int CH = 8;
  constexpr int TOT = kQD * kQD;
  for (int base = 0; base < TOT; base += 64 * CH) {
    int32_t pp[CH];
    double old[CH];
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int t = base + k * 64 + lane;
      pp[k] = t < TOT ? nzt[t] : -1;
    }
#pragma unroll
    for (int k = 0; k < CH; ++k) old[k] = pp[k] >= 0 ? nzval[pp[k]] : 0.0;
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int t = base + k * 64 + lane;
      if (t >= TOT) continue;
      const int Jc = t / kQD, Ir = t - Jc * kQD;
      const int a = Ir / 3, i = Ir - 3 * a, b = Jc / 3, j = Jc - 3 * b;
      const T pij = i <= j ? sP[(pidx(i, j) * kQN + a) * kPS + b] : sP[(pidx(j, i) * kQN + b) * kPS + a];
      const T pji = j <= i ? sP[(pidx(j, i) * kQN + a) * kPS + b] : sP[(pidx(i, j) * kQN + b) * kPS + a];
      T v = lE * pij + mE * pji;
      if (i == j) v += mE * (sP[(0 * kQN + a) * kPS + b] + sP[(3 * kQN + a) * kPS + b] + sP[(5 * kQN + a) * kPS + b]);
      if (pp[k] >= 0) {
        if (v != T(0)) nzval[pp[k]] = old[k] + (double)v;
      } else if (v != T(0)) {
        atomicMin(&err->pattern_cell, ccell[s]);
      }
    }
  }
}

__global__ void k_fill_zero_f(double* __restrict__ p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Q2 hex^3, owner-computes path.
//   phase 1  k_elast_q2hex_ke<float> (FP32 MFMA) / k_elast_q2hex_ke<double> (FP64 MFMA):
//            one 4-wave workgroup per cell in cell-id order, the element matrix stored as
//            above.
//   phase 2  k_gather_q2: one wave per dof node m (dofs 3m, 3m+1, 3m+2).  The three columns
//            share one row set made of triples (checked once per pattern), so the wave keeps
//            the three columns in LDS, adds the 81 x 3 block of every cell holding m in
//            ascending cell id (the order of Ferrite's serial loop), and writes the three
//            columns, which are contiguous in nzval, once.  No nzval entry is read-modify-
//            written per cell and no per-entry position table is read: a cell's local node a
//            lands at row triple nrank[e][a] of the node's columns.
// ---------------------------------------------------------------------------

// MFMA 16x16x4 in FP32 (v_mfma_f32_16x16x4f32) or FP64 (v_mfma_f64_16x16x4f64).  Lane l
// supplies A[l % 16][l / 16] and B[l / 16][l % 16]; it holds D[row(l / 16, e)][l % 16], with
// row = 4 (l / 16) + e for FP32 and (l / 16) + 4 e for FP64 (measured on gfx950).
template <typename T> struct Mfma16;
template <> struct Mfma16<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static constexpr __device__ int row(int kq, int e) { return 4 * kq + e; }
};
template <> struct Mfma16<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static constexpr __device__ int row(int kq, int e) { return kq + 4 * e; }
};

// Element matrices of the owner-computes path, one cell per 256-thread block.  All nine
// products P_ij = A_i A_jᵀ (not only i <= j) are formed, so a lane holds P_ij[a][b] and
// P_ji[a][b] for the same (a, b) and forms Ke_ij = E (λ P_ij + μ P_ji + μ δ_ij Σ_k P_kk) in
// registers.  A_i[a][q] = √dV_q g_q[a, i] carries the quadrature weight on both operands, so
// P_ji[b][a] multiplies the same two factors as P_ij[a][b] in the same order: Ke is symmetric
// bit for bit, as Ferrite's Ke is, and CG sees a symmetric K.  Wave w computes the 16 x 16
// output tile (ti, tj) = (w / 2, w % 2) of the 32 x 32 padded products; the 27 x 27 (q, a)
// gradient pairs are spread over the block.
template <typename T>
__global__ __launch_bounds__(256) void k_elast_q2hex_ke(
    const double* __restrict__ X0, const double* __restrict__ X1, const double* __restrict__ X2,
    const int32_t* __restrict__ cnodes, const double* __restrict__ param, const T* __restrict__ tab, T lam, T mu,
    T* __restrict__ keb, DevError* __restrict__ err) {
  using L = TabL<3, 8, kQN, kQQ>;
  using V = typename Mfma16<T>::v4;
  __shared__ T sA[3 * kMP * kKP];   // [i][a][q]; rows a >= 27 and column q = 27 stay zero
  __shared__ T sJi[kKP][9];
  __shared__ T sdq[kKP];            // √(det J w_q)
  const int tid = threadIdx.x;
  const int64_t cell = blockIdx.x;
  // this thread's (q, a) pairs t = tid + 256 m of the gradient phase: their table entries are
  // loaded first, so the loads overlap the Jacobian phase
  constexpr int NQA = (kQQ * kQN + 255) / 256;
  T dnv[NQA][3];
#pragma unroll
  for (int m = 0; m < NQA; ++m) {
    const int t = tid + 256 * m;
#pragma unroll
    for (int d = 0; d < 3; ++d) dnv[m][d] = t < kQQ * kQN ? tab[L::odN + t * 3 + d] : T(0);
  }
  for (int t = tid; t < 3 * kMP * kKP; t += 256) sA[t] = T(0);
  if (tid < kQQ) {   // Jacobian at point q from the 8 corners
    const int q = tid;
    T J[3][3] = {{T(0), T(0), T(0)}, {T(0), T(0), T(0)}, {T(0), T(0), T(0)}};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int32_t nd = cnodes[cell * 8 + a];
      const T xa[3] = {(T)X0[nd], (T)X1[nd], (T)X2[nd]};
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) J[i][j] += xa[i] * tab[L::odM + (q * 8 + a) * 3 + j];
    }
    T Ji[3][3];
    const T detJ = det_inv_e<T, 3>(J, Ji);
    if (!(detJ > T(0))) atomicMin(&err->detj_cell, (int)cell);
    sdq[q] = sqrt(detJ * tab[L::ow + q]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) sJi[q][3 * i + j] = Ji[i][j];
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < NQA; ++m) {   // √dV ∇N_a at point q: g_i = Σ_d dN_d Ji[d][i]
    const int t = tid + 256 * m;
    if (t >= kQQ * kQN) break;
    const int q = t / kQN, a = t - q * kQN;
    const T d0 = dnv[m][0], d1 = dnv[m][1], d2 = dnv[m][2], sq = sdq[q];
#pragma unroll
    for (int i = 0; i < 3; ++i)
      sA[(i * kMP + a) * kKP + q] = sq * (d0 * sJi[q][i] + d1 * sJi[q][3 + i] + d2 * sJi[q][6 + i]);
  }
  __syncthreads();
  const T E = (T)param[cell];
  const T lE = lam * E, mE = mu * E;
  const int lane = tid & 63, w = tid >> 6;
  const int ti = w >> 1, tj = w & 1;
  const int r = lane & 15, kq = lane >> 4;
  V P[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) P[i][j] = V{T(0), T(0), T(0), T(0)};
#pragma unroll
  for (int kk = 0; kk < kKP / 4; ++kk) {
    const int k = 4 * kk + kq;
    T av[3], bv[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      av[i] = sA[(i * kMP + 16 * ti + r) * kKP + k];
      bv[i] = sA[(i * kMP + 16 * tj + r) * kKP + k];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) P[i][j] = Mfma16<T>::op(av[i], bv[j], P[i][j]);
  }
  // lane holds rows a = 16 ti + row(kq, e) (e = 0..3) of column b = 16 tj + r; rows up to 27
  // are stored (27 is the zero padding row)
  const int b = 16 * tj + r;
  if (b < kQN) {
    T* __restrict__ out = keb + cell * (int64_t)kKC + b * kKS + 16 * ti;
    const V tr = P[0][0] + P[1][1] + P[2][2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        V v = lE * P[i][j] + mE * P[j][i];
        if (i == j) v += mE * tr;
        T* o = out + (i * 3 + j) * kKP;
        if constexpr (sizeof(T) == 4) {   // four consecutive rows: one 16-byte store
          if (16 * ti + 4 * kq < kKP) *reinterpret_cast<V*>(o + 4 * kq) = v;
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (16 * ti + Mfma16<T>::row(kq, e) < kKP) o[Mfma16<T>::row(kq, e)] = v[e];
        }
      }
  }
}

// flag = 0 unless every dof node m has three equal-length columns 3m..3m+2 whose rows are
// the same triples (3k, 3k+1, 3k+2), at most 254 of them.
__global__ void k_check_node_blocks(int64_t nnode, const int32_t* __restrict__ colptr,
                                    const int32_t* __restrict__ rowval, int* __restrict__ flag) {
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < nnode; m += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c0 = colptr[3 * m], L = colptr[3 * m + 1] - c0;
    bool ok = L % 3 == 0 && L / 3 <= 254 && colptr[3 * m + 2] - colptr[3 * m + 1] == L &&
              colptr[3 * m + 3] - colptr[3 * m + 2] == L;
    for (int32_t k = 0; ok && k < L; k += 3) {
      const int32_t r0 = rowval[c0 + k];
      ok = r0 % 3 == 0;
      for (int col = 0; col < 3; ++col)
        for (int i = 0; i < 3; ++i) ok = ok && rowval[c0 + col * L + k + i] == r0 + i;
    }
    if (!ok) *flag = 0;   // every writer stores the same 0
  }
}

// nrank[e][a] = row triple of local node a of list entry e's cell inside its node's columns
// (255 when absent from the pattern).
__global__ void k_build_nrank(int64_t nnode, const int32_t* __restrict__ colptr, const int32_t* __restrict__ rowval,
                              const int32_t* __restrict__ n2c_ptr, const int32_t* __restrict__ n2c,
                              const int32_t* __restrict__ cdofs, uint8_t* __restrict__ nrank) {
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < nnode; m += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c0 = colptr[3 * m], c1 = colptr[3 * m + 1];
    for (int32_t e = n2c_ptr[m]; e < n2c_ptr[m + 1]; ++e) {
      const int64_t cell = n2c[e] / kQN;
      for (int a = 0; a < kQN; ++a) {
        const int32_t row = cdofs[cell * kQD + 3 * a];
        int32_t lo = c0, hi = c1;
        while (lo < hi) {
          const int32_t mid = (lo + hi) >> 1;
          if (rowval[mid] < row) lo = mid + 1; else hi = mid;
        }
        const bool found = lo < c1 && rowval[lo] == row;
        nrank[(int64_t)e * kQN + a] = found ? (uint8_t)((lo - c0) / 3) : (uint8_t)255;
      }
    }
  }
}

// One wave per dof node; G list entries have their loads in flight before they are added.
template <typename T, int G>
__global__ __launch_bounds__(64) void k_gather_q2(int64_t nnode, const int32_t* __restrict__ colptr,
                                                  const int32_t* __restrict__ n2c_ptr, const int32_t* __restrict__ n2c,
                                                  const uint8_t* __restrict__ nrank, const T* __restrict__ keb,
                                                  double* __restrict__ nzval, int fillzero,
                                                  DevError* __restrict__ err) {
  extern __shared__ double acc[];   // the node's three columns, 3 L values
  constexpr int S = (kKS + 63) / 64;   // 4 values of the 252-value block per lane
  const int lane = threadIdx.x;
  // workgroups are dealt round-robin to the 8 XCDs: each XCD takes one contiguous range of
  // nodes, whose cells (numbered close together by the first-visit dof numbering) then
  // share that XCD's L2.  gridDim.x is a multiple of 8.
  const int64_t per = gridDim.x / 8;
  const int64_t m = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (m >= nnode) return;
  const int32_t c0 = colptr[3 * m], L = colptr[3 * m + 1] - c0, tot = 3 * L;
  for (int q = lane; q < tot; q += 64) acc[q] = fillzero ? 0.0 : nzval[c0 + q];
  // block slot t = lane + 64 s -> (i, j) = t / 28, row a = t % 28 (a = 27: padding)
  int dst[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int t = lane + 64 * s;
    const int ij = t / kKP, a = t - ij * kKP;
    const int i = ij / 3, j = ij - 3 * i;
    dst[s] = (t < kKS && a < kQN) ? (j * L + i) : -1;   // + 3 rank
  }
  __syncthreads();
  const int32_t e0 = n2c_ptr[m], e1 = n2c_ptr[m + 1];
  for (int32_t eb = e0; eb < e1; eb += G) {
    T v[G][S];
    uint8_t rk[G][S];
    int32_t cl[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int32_t e = eb + g;
      const bool in = e < e1;
      const int32_t ent = in ? n2c[e] : 0;
      const int64_t cell = ent / kQN;
      const int b = ent - (int)cell * kQN;
      cl[g] = (int32_t)cell;
      const T* blk = keb + cell * (int64_t)kKC + b * kKS;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const int t = lane + 64 * s;
        const bool ok = in && dst[s] >= 0;
        v[g][s] = ok ? blk[t] : T(0);
        rk[g][s] = ok ? nrank[(int64_t)e * kQN + (t % kKP)] : (uint8_t)255;
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (eb + g >= e1) break;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        if (dst[s] < 0) continue;
        if (rk[g][s] != 255) {
          // LDS add without a return value: the adds of one instruction hit distinct slots and a
          // wave's LDS operations complete in order, so each slot still sums its entries in
          // ascending cell order, rounded as acc + v
          atomicAdd(&acc[dst[s] + 3 * rk[g][s]], (double)v[g][s]);
        } else if (v[g][s] != T(0)) {
          atomicMin(&err->pattern_cell, cl[g]);
        }
      }
    }
  }
  __syncthreads();
  for (int q = lane; q < tot; q += 64) nzval[c0 + q] = acc[q];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <typename T, int DIM, int NG, int NB, int NQ>
static void pack_tab(const RefTables& r, std::vector<T>& out) {
  using L = TabL<DIM, NG, NB, NQ>;
  out.assign(L::size, T(0));
  for (int q = 0; q < NQ; ++q) {
    out[L::ow + q] = (T)r.w[q];
    for (int b = 0; b < NB; ++b) {
      out[L::oN + q * NB + b] = (T)r.N[q][b];
      for (int d = 0; d < DIM; ++d) out[L::odN + (q * NB + b) * DIM + d] = (T)r.dN[q][b][d];
    }
    for (int a = 0; a < NG; ++a)
      for (int d = 0; d < DIM; ++d) out[L::odM + (q * NG + a) * DIM + d] = (T)r.dM[q][a][d];
  }
}

template <int DIM, int NG, int NB, int NQ>
static int upload_tab(fem_ctx* c) {
  std::vector<double> td;
  std::vector<float> tf;
  pack_tab<double, DIM, NG, NB, NQ>(c->ref, td);
  pack_tab<float, DIM, NG, NB, NQ>(c->ref, tf);
  int rc;
  if ((rc = dev_alloc(c, &c->d_tab, td.size())) || (rc = dev_alloc(c, &c->d_tabf, tf.size()))) return rc;
  FEM_HIP(c, hipMemcpyAsync(c->d_tab, td.data(), td.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipMemcpyAsync(c->d_tabf, tf.data(), tf.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  c->tab_dirty = true;  // packed for this kernel family; the Poisson path repacks
  return FEM_OK;
}

void free_q2_tables(fem_ctx* c) {
  dev_free(&c->d_n2c_ptr); dev_free(&c->d_n2c); dev_free(&c->d_nrank); dev_free(&c->d_q2cnode);
  dev_free(&c->d_keq32); dev_free(&c->d_keq64);
  c->keq32_cells = c->keq64_cells = 0;
  c->q2_ready = -1;
}

static int gridq(int64_t n, int b = 256) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + b - 1) / b, 16384)); }

// Whether this is vector Q2 hex elasticity (27 nodes, 3x3x3 Gauss, Q1 geometry).
static bool is_q2hex(const fem_ctx* c) {
  return c->rdim == 3 && c->sdim == 3 && c->ngeo == 8 && c->nbf == kQN && c->nqp == kQQ && c->npc == 8 &&
         c->vdim == 3 && c->ndpc == kQD;
}

// Node lists and rank table of the owner-computes Q2 path, once per mesh/dofs/pattern.
// Sets q2_ready = 0 (colored scatter instead) when a cell's dofs are not node triples
// 3m, 3m+1, 3m+2 in local order or the pattern is not made of such triples.
static int build_q2_tables(fem_ctx* c) {
  if (c->q2_ready >= 0) return FEM_OK;
  c->q2_ready = 0;
  const int64_t nc = c->ncells, nd = c->ndofs;
  if (!is_q2hex(c) || nd % 3 || c->n != nd || nc * kQN >= INT32_MAX) return FEM_OK;
  const int64_t nnode = nd / 3;
  const int32_t* cd = c->h_cell_dofs.data();
  for (int64_t i = 0; i < nc * kQN; ++i) {
    const int32_t d = cd[3 * i];
    if (d % 3 || cd[3 * i + 1] != d + 1 || cd[3 * i + 2] != d + 2) return FEM_OK;
  }
  // node -> (cell, local node), ascending cell id per node (counting sort over the cell loop)
  std::vector<int32_t> ptr(nnode + 1, 0), ent(nc * kQN);
  for (int64_t i = 0; i < nc * kQN; ++i) ++ptr[cd[3 * i] / 3 + 1];
  for (int64_t m = 0; m < nnode; ++m) ptr[m + 1] += ptr[m];
  std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
  for (int64_t cell = 0; cell < nc; ++cell)
    for (int b = 0; b < kQN; ++b) ent[fill[cd[(cell * kQN + b) * 3] / 3]++] = (int32_t)(cell * kQN + b);
  int rc;
  int* flag = nullptr;
  int32_t* dcd = nullptr;
  if ((rc = dev_alloc(c, &flag, 1))) return rc;
  const int one = 1;
  int h = 0;
  FEM_HIP(c, hipMemcpyAsync(flag, &one, sizeof(int), hipMemcpyHostToDevice, c->stream));
  k_check_node_blocks<<<gridq(nnode), 256, 0, c->stream>>>(nnode, c->d_colptr, c->d_rowval, flag);
  FEM_HIP(c, hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  dev_free(&flag);
  if (!h) return FEM_OK;
  if (!c->d_gcnodes) {
    if ((rc = dev_alloc(c, &c->d_gcnodes, nc * c->npc))) return rc;
    FEM_HIP(c, hipMemcpyAsync(c->d_gcnodes, c->h_cell_nodes.data(), 4 * nc * c->npc, hipMemcpyHostToDevice,
                              c->stream));
  }
  if ((rc = dev_alloc(c, &c->d_n2c_ptr, nnode + 1)) || (rc = dev_alloc(c, &c->d_n2c, nc * kQN)) ||
      (rc = dev_alloc(c, &c->d_nrank, nc * kQN * (int64_t)kQN)) || (rc = dev_alloc(c, &dcd, nc * kQD)))
    return rc;
  FEM_HIP(c, hipMemcpyAsync(c->d_n2c_ptr, ptr.data(), 4 * (nnode + 1), hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipMemcpyAsync(c->d_n2c, ent.data(), 4 * nc * kQN, hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipMemcpyAsync(dcd, cd, 4 * nc * kQD, hipMemcpyHostToDevice, c->stream));
  k_build_nrank<<<gridq(nnode), 256, 0, c->stream>>>(nnode, c->d_colptr, c->d_rowval, c->d_n2c_ptr, c->d_n2c, dcd,
                                                      c->d_nrank);
  FEM_HIP(c, hipGetLastError());
  // dof node of every local node (element-assembly solver operator)
  std::vector<int32_t> cn(nc * kQN);
  for (int64_t i = 0; i < nc * kQN; ++i) cn[i] = cd[3 * i] / 3;
  if ((rc = dev_alloc(c, &c->d_q2cnode, nc * kQN))) return rc;
  FEM_HIP(c, hipMemcpyAsync(c->d_q2cnode, cn.data(), 4 * nc * kQN, hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  dev_free(&dcd);
  c->q2_ready = 1;
  return FEM_OK;
}

int assemble_elasticity_gather(fem_ctx* c, const fem_assemble_cfg* cfg, bool* done) {
  *done = false;
  if (!is_q2hex(c) || c->tune.asm_mode == 1) return FEM_OK;
  const double nu = cfg->nu;
  if (!(nu > -1.0 && nu < 0.5)) return set_error(c, FEM_ERR_ARG, "Poisson's ratio must be in (-1, 0.5)");
  int rc;
  if ((rc = build_q2_tables(c))) return rc;
  if (c->q2_ready != 1) return FEM_OK;
  const double lam = nu / ((1 + nu) * (1 - 2 * nu)), mu = 1.0 / (2 * (1 + nu));
  const bool mixed = cfg->precision == FEM_MIXED;
  const int64_t nc = c->ncells, nnode = c->ndofs / 3;
  if (mixed && (!c->d_keq32 || c->keq32_cells != nc)) {
    if ((rc = dev_alloc(c, &c->d_keq32, nc * (int64_t)kKC))) return rc;
    c->keq32_cells = nc;
  }
  if (!mixed && (!c->d_keq64 || c->keq64_cells != nc)) {
    if ((rc = dev_alloc(c, &c->d_keq64, nc * (int64_t)kKC))) return rc;
    c->keq64_cells = nc;
  }
  if ((rc = upload_tab<3, 8, 27, 27>(c))) return rc;
  const size_t lds = sizeof(double) * 3 * (size_t)c->max_row_nnz;
  if (lds > 64 * 1024) return FEM_OK;   // longest column beyond one workgroup's LDS: colored scatter
  *done = true;
  FEM_HIP(c, hipEventRecord(c->ev0, c->stream));
  if (cfg->fillzero) k_fill_zero_f<<<512, 256, 0, c->stream>>>(c->d_f, c->ndofs);   // nzval is written whole
  if (mixed)
    k_elast_q2hex_ke<float><<<(unsigned)nc, 256, 0, c->stream>>>(c->d_x[0], c->d_x[1], c->d_x[2], c->d_gcnodes,
                                                                 c->d_param_cell, c->d_tabf, (float)lam, (float)mu,
                                                                 c->d_keq32, c->d_err);
  else
    k_elast_q2hex_ke<double><<<(unsigned)nc, 256, 0, c->stream>>>(c->d_x[0], c->d_x[1], c->d_x[2], c->d_gcnodes,
                                                                  c->d_param_cell, c->d_tab, lam, mu, c->d_keq64,
                                                                  c->d_err);
  FEM_HIP(c, hipEventRecord(c->evm, c->stream));
  const unsigned nblk = (unsigned)((nnode + 7) / 8 * 8);
  const int fz = cfg->fillzero ? 1 : 0;
  if (mixed)
    k_gather_q2<float, 8><<<nblk, 64, lds, c->stream>>>(nnode, c->d_colptr, c->d_n2c_ptr, c->d_n2c, c->d_nrank,
                                                        c->d_keq32, c->d_nzval, fz, c->d_err);
  else
    k_gather_q2<double, 8><<<nblk, 64, lds, c->stream>>>(nnode, c->d_colptr, c->d_n2c_ptr, c->d_n2c, c->d_nrank,
                                                         c->d_keq64, c->d_nzval, fz, c->d_err);
  FEM_HIP(c, hipGetLastError());
  FEM_HIP(c, hipEventRecord(c->ev1, c->stream));
  c->timings.assemble_kernel_ms = event_ms(c);
  float e1 = 0.f, e2 = 0.f;
  FEM_HIP(c, hipEventElapsedTime(&e1, c->ev0, c->evm));
  FEM_HIP(c, hipEventElapsedTime(&e2, c->evm, c->ev1));
  c->timings.element_ms = e1;
  c->timings.scatter_ms = e2;
  c->timings.fill_ms = 0.0;
  c->timings.assemble_launches = 2 + (cfg->fillzero ? 1 : 0);
  return FEM_OK;
}

int assemble_elasticity(fem_ctx* c, const fem_assemble_cfg* cfg) {
  const double nu = cfg->nu;
  if (!(nu > -1.0 && nu < 0.5)) return set_error(c, FEM_ERR_ARG, "Poisson's ratio must be in (-1, 0.5)");
  const double lam = nu / ((1 + nu) * (1 - 2 * nu)), mu = 1.0 / (2 * (1 + nu));
  const bool mixed = cfg->precision == FEM_MIXED;
  if (c->vdim != c->rdim || c->sdim != c->rdim || c->ndpc != c->nbf * c->vdim)
    return set_error(c, FEM_ERR_ARG, "elasticity needs a vector field ip^dim on a dim-dimensional grid");
  int rc;
  FEM_HIP(c, hipEventRecord(c->ev0, c->stream));
  if (c->rdim == 2 && c->ngeo == 4 && c->nbf == 4 && c->nqp == 4 && c->npc == 4) {
    if ((rc = upload_tab<2, 4, 4, 4>(c))) return rc;
    // element matrices kept as planes (upper triangle, 36 per cell) for the element-assembly
    // solver operator, with the row lists it walks
    constexpr int TRI = 36;
    if (c->ep_cells != c->ncells) {
      dev_free(&c->d_ep64);
      dev_free(&c->d_ep32);
    }
    if (mixed ? !c->d_ep32 : !c->d_ep64) {
      if ((rc = mixed ? dev_alloc(c, &c->d_ep32, TRI * c->ncells) : dev_alloc(c, &c->d_ep64, TRI * c->ncells)))
        return rc;
    }
    c->ep_cells = c->ncells;
    if ((rc = build_ea_rows_tables(c))) return rc;
    FEM_HIP(c, hipEventRecord(c->ev0, c->stream));
    for (int col = 0; col < c->ncolors; ++col) {
      const int64_t s0 = c->color_ptr[col], s1 = c->color_ptr[col + 1];
      if (s1 <= s0) continue;
      constexpr int B = 128;
      const unsigned nb = (unsigned)((s1 - s0 + B - 1) / B);
      if (mixed)
        k_elast_small<float, 2, 4, 4, B><<<nb, B, 0, c->stream>>>(s0, s1, c->d_x[0], c->d_x[1], c->d_x[2], c->d_cnodes,
            c->d_cdofs, c->d_cnz, c->d_ccell, c->d_cparam, c->d_tabf, (float)lam, (float)mu, c->d_nzval, c->d_err,
            c->d_ep32, c->ncells);
      else
        k_elast_small<double, 2, 4, 4, B><<<nb, B, 0, c->stream>>>(s0, s1, c->d_x[0], c->d_x[1], c->d_x[2], c->d_cnodes,
            c->d_cdofs, c->d_cnz, c->d_ccell, c->d_cparam, c->d_tab, lam, mu, c->d_nzval, c->d_err, c->d_ep64,
            c->ncells);
      c->timings.assemble_launches++;
    }
  } else if (c->rdim == 3 && c->ngeo == 8 && c->nbf == 27 && c->nqp == 27 && c->npc == 8) {
    if ((rc = upload_tab<3, 8, 27, 27>(c))) return rc;
    FEM_HIP(c, hipEventRecord(c->ev0, c->stream));
    for (int col = 0; col < c->ncolors; ++col) {
      const int64_t s0 = c->color_ptr[col], s1 = c->color_ptr[col + 1];
      if (s1 <= s0) continue;
      const unsigned nb = (unsigned)(s1 - s0);
      if (mixed)
        k_elast_q2hex<float><<<nb, 64, 0, c->stream>>>(s0, c->d_x[0], c->d_x[1], c->d_x[2], c->d_cnodes, c->d_cnz,
            c->d_ccell, c->d_cparam, c->d_tabf, (float)lam, (float)mu, c->d_nzval, c->d_err);
      else
        k_elast_q2hex<double><<<nb, 64, 0, c->stream>>>(s0, c->d_x[0], c->d_x[1], c->d_x[2], c->d_cnodes, c->d_cnz,
            c->d_ccell, c->d_cparam, c->d_tab, lam, mu, c->d_nzval, c->d_err);
      c->timings.assemble_launches++;
    }
  } else {
    return set_error(c, FEM_ERR_ARG, "no elasticity kernel for this element/quadrature combination "
                                     "(supported: Q1 quad^2 with 2x2 Gauss, Q2 hex^3 with 3x3x3 Gauss)");
  }
  FEM_HIP(c, hipGetLastError());
  FEM_HIP(c, hipEventRecord(c->ev1, c->stream));
  c->timings.assemble_kernel_ms = event_ms(c);
  return FEM_OK;
}

__global__ void k_point_load(int64_t n, const int32_t* __restrict__ dofs, const double* __restrict__ v,
                             double* __restrict__ f) {
  // serial in one thread: loads may repeat a dof, and their order must be the caller's
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int64_t i = 0; i < n; ++i) f[dofs[i]] += v[i];
}

}  // namespace fem

using namespace fem;

extern "C" int32_t fem_point_load(fem_ctx* c, int64_t n, const int64_t* dofs, const double* values) {
  if (!c || n < 0 || (n > 0 && (!dofs || !values)) || c->ndofs == 0)
    return set_error(c, FEM_ERR_ARG, "fem_point_load: invalid arguments");
  if (n == 0) return FEM_OK;
  std::vector<int32_t> d(n);
  for (int64_t i = 0; i < n; ++i) {
    if (dofs[i] < 1 || dofs[i] > c->ndofs) return set_error(c, FEM_ERR_ARG, "point load dof out of range");
    d[i] = (int32_t)(dofs[i] - 1);
  }
  int32_t* dd = nullptr;
  double* dv = nullptr;
  int rc;
  if ((rc = dev_alloc(c, &dd, n)) || (rc = dev_alloc(c, &dv, n))) { dev_free(&dd); return rc; }
  FEM_HIP(c, hipMemcpyAsync(dd, d.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipMemcpyAsync(dv, values, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  k_point_load<<<1, 64, 0, c->stream>>>(n, dd, dv, c->d_f);
  FEM_HIP(c, hipGetLastError());
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  dev_free(&dd);
  dev_free(&dv);
  return FEM_OK;
}
