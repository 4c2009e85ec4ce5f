// Element math shared by the assembly kernels: packed reference tables, det/inverse of
// the Jacobian, and the scalar Poisson element (CellValues reinit! + quadrature loop).
#pragma once

#include "fem_internal.h"

namespace fem {

// Packed reference tables: N[NQ][NB], dN[NQ][NB][DIM], dM[NQ][NG][DIM], w[NQ].
template <int DIM, int NG, int NB, int NQ>
struct TabLayout {
  static constexpr int oN = 0;
  static constexpr int odN = oN + NQ * NB;
  static constexpr int odM = odN + NQ * NB * DIM;
  static constexpr int ow = odM + NQ * NG * DIM;
  static constexpr int size = ow + NQ;
};

template <typename T, int DIM>
__device__ __forceinline__ T det_inv(const T (&J)[DIM][DIM], T (&Ji)[DIM][DIM]) {
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

// Index of (lo, hi), lo <= hi, in the row-major upper triangle of an nb x nb matrix.
constexpr __host__ __device__ int tri_idx(int lo, int hi, int nb) {
  return lo * (2 * nb - lo + 1) / 2 + (hi - lo);
}

// Scalar Poisson/heat element with Q1 geometry == function interpolation
// (heat_equation.jl:143-164 with a per-cell conductivity k):
//   per quadrature point: J = Σ_j x_j ⊗ dMdξ_j (GeometryMapping.jl:124-130), det J > 0
//   (CellValues.jl:110-116), dNdx = dNdξ ⋅ J⁻¹ (FunctionValues.jl:186-192), then
//   Ke[a,b] += k (∇N_a·∇N_b) dΩ (upper triangle) and fe[a] += src N_a dΩ.
// ∇N_a·∇N_b dΩ is formed as dNdξ_aᵀ A dNdξ_b with A = dΩ J⁻¹J⁻ᵀ (symmetric), which needs
// no per-point array of pushed-forward gradients.  `tab` is best an LDS copy of the table.
// Returns false if some det J <= 0.
template <typename T, int DIM, int NG, int NB, int NQ>
__device__ __forceinline__ bool poisson_cell(const T (&x)[NG][DIM], T k, T src, const T* __restrict__ tab,
                                             T (&ke)[NB * (NB + 1) / 2], T (&fe)[NB]) {
  using L = TabLayout<DIM, NG, NB, NQ>;
  constexpr int TRI = NB * (NB + 1) / 2;
#pragma unroll
  for (int i = 0; i < TRI; ++i) ke[i] = T(0);
#pragma unroll
  for (int i = 0; i < NB; ++i) fe[i] = T(0);
  bool ok = true;
  // one quadrature point at a time (not unrolled): only that point's table entries are live
#pragma unroll 1
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
    const T detJ = det_inv<T, DIM>(J, Ji);
    ok &= detJ > T(0);
    const T dV = detJ * tab[L::ow + q];
    const T kdV = k * dV;
    const T sdV = src * dV;
    T A[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int l = i; l < DIM; ++l) {
        T acc = Ji[i][0] * Ji[l][0];
#pragma unroll
        for (int m = 1; m < DIM; ++m) acc += Ji[i][m] * Ji[l][m];
        A[i][l] = kdV * acc;
        A[l][i] = A[i][l];
      }
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      fe[a] += tab[L::oN + q * NB + a] * sdV;
      T wa[DIM];
#pragma unroll
      for (int i = 0; i < DIM; ++i) {
        T acc = A[i][0] * tab[L::odN + (q * NB + a) * DIM];
#pragma unroll
        for (int j = 1; j < DIM; ++j) acc += A[i][j] * tab[L::odN + (q * NB + a) * DIM + j];
        wa[i] = acc;
      }
#pragma unroll
      for (int b = a; b < NB; ++b) {
        T dot = tab[L::odN + (q * NB + b) * DIM] * wa[0];
#pragma unroll
        for (int i = 1; i < DIM; ++i) dot += tab[L::odN + (q * NB + b) * DIM + i] * wa[i];
        ke[tri_idx(a, b, NB)] += dot;
      }
    }
  }
  return ok;
}

// The same element on a trilinear hex (Ferrite's node order, geometry == 2x2x2 function
// interpolation checked on the host).  With Y[m] = (1/8) Σ_a Π_{d∈m} s_a,d x_a (a
// Walsh–Hadamard transform of the node coordinates, m a subset of {ξ, η, ζ} as a bit mask),
// x(ξ) = Σ_m Y[m] Π_{d∈m} ξ_d, so the Jacobian columns at a point are
//   ∂x/∂ξ = Y₁ + Y₃η + Y₅ζ + Y₇ηζ,  ∂x/∂η = Y₂ + Y₃ξ + Y₆ζ + Y₇ξζ,  ∂x/∂ζ = Y₄ + Y₅ξ + Y₆η + Y₇ξη:
// 27 multiply-adds per point instead of the 72 of Σ_a x_a ⊗ ∇M_a.  qc[q] = ξ, η, ζ, ηζ, ξζ, ξη.
// Y[m] (m ≥ 1, scaled by 1/8) of the trilinear map of a hex with Ferrite's node order.
template <typename T>
__device__ __forceinline__ void trilinear_coeffs(const T (&x)[8][3], T (&Y)[8][3]) {
  // node a -> bit index (bit 0: ξ sign, bit 1: η, bit 2: ζ)
  constexpr int kBit[8] = {0, 1, 3, 2, 4, 5, 7, 6};
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) Y[kBit[a]][i] = x[a][i];
#pragma unroll
  for (int bit = 1; bit < 8; bit <<= 1)
#pragma unroll
    for (int m = 0; m < 8; ++m)
      if (!(m & bit)) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const T lo = Y[m][i], hi = Y[m | bit][i];
          Y[m][i] = lo + hi;
          Y[m | bit][i] = hi - lo;
        }
      }
#pragma unroll
  for (int m = 1; m < 8; ++m)
#pragma unroll
    for (int i = 0; i < 3; ++i) Y[m][i] *= T(0.125);
}

// True when the cell is a parallelepiped up to rounding: |Y₃,Y₅,Y₆,Y₇| ≤ tol |Y₁,Y₂,Y₄|.
template <typename T>
__device__ __forceinline__ bool is_affine(const T (&Y)[8][3], T tol) {
  T lo = T(0), hi = T(0);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    lo = fmax(lo, fmax(fabs(Y[1][i]), fmax(fabs(Y[2][i]), fabs(Y[4][i]))));
    hi = fmax(hi, fmax(fmax(fabs(Y[3][i]), fabs(Y[5][i])), fmax(fabs(Y[6][i]), fabs(Y[7][i]))));
  }
  return hi <= tol * lo;
}

template <typename T, int NB, int NQ>
__device__ __forceinline__ bool poisson_cell_tri(const T (&x)[8][3], T k, T src, const T* __restrict__ tab,
                                                 const T* __restrict__ qc, T (&ke)[NB * (NB + 1) / 2],
                                                 T (&fe)[NB]) {
  using L = TabLayout<3, 8, NB, NQ>;
  constexpr int TRI = NB * (NB + 1) / 2;
  T Y[8][3];
  trilinear_coeffs<T>(x, Y);
#pragma unroll
  for (int i = 0; i < TRI; ++i) ke[i] = T(0);
#pragma unroll
  for (int i = 0; i < NB; ++i) fe[i] = T(0);
  bool ok = true;
#pragma unroll 1
  for (int q = 0; q < NQ; ++q) {
    const T xi = qc[6 * q], eta = qc[6 * q + 1], zeta = qc[6 * q + 2];
    const T ez = qc[6 * q + 3], xz = qc[6 * q + 4], xe = qc[6 * q + 5];
    T J[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      J[i][0] = Y[1][i] + Y[3][i] * eta + Y[5][i] * zeta + Y[7][i] * ez;
      J[i][1] = Y[2][i] + Y[3][i] * xi + Y[6][i] * zeta + Y[7][i] * xz;
      J[i][2] = Y[4][i] + Y[5][i] * xi + Y[6][i] * eta + Y[7][i] * xe;
    }
    T Ji[3][3];
    const T detJ = det_inv<T, 3>(J, Ji);
    ok &= detJ > T(0);
    const T dV = detJ * tab[L::ow + q];
    const T kdV = k * dV;
    const T sdV = src * dV;
    T A[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int l = i; l < 3; ++l) {
        T acc = Ji[i][0] * Ji[l][0];
#pragma unroll
        for (int m = 1; m < 3; ++m) acc += Ji[i][m] * Ji[l][m];
        A[i][l] = kdV * acc;
        A[l][i] = A[i][l];
      }
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      fe[a] += tab[L::oN + q * NB + a] * sdV;
      T wa[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        T acc = A[i][0] * tab[L::odN + (q * NB + a) * 3];
#pragma unroll
        for (int j = 1; j < 3; ++j) acc += A[i][j] * tab[L::odN + (q * NB + a) * 3 + j];
        wa[i] = acc;
      }
#pragma unroll
      for (int b = a; b < NB; ++b) {
        T dot = tab[L::odN + (q * NB + b) * 3] * wa[0];
#pragma unroll
        for (int i = 1; i < 3; ++i) dot += tab[L::odN + (q * NB + b) * 3 + i] * wa[i];
        ke[tri_idx(a, b, NB)] += dot;
      }
    }
  }
  return ok;
}

// Affine cells (parallelepipeds: every cell of the mesh passed is_affine with tol 1e-14,
// checked once per mesh) have one J for every point, so with A = k det J J⁻¹J⁻ᵀ the
// quadrature sum factors: Ke[a,b] = Σ_m A_m P[(a,b)][m] and fe[a] = src det J Σ_q w_q N_a, with
// the reference integrals P (m = 00, 11, 22, 01, 02, 12; off-diagonal m with both orders)
// taken over the same quadrature on the host.  This returns det J and A.
template <typename T>
__device__ __forceinline__ T affine_metric(const T (&x)[8][3], T k, T (&A)[6]) {
  T Y[8][3];
  trilinear_coeffs<T>(x, Y);
  T J[3][3], Ji[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    J[i][0] = Y[1][i];
    J[i][1] = Y[2][i];
    J[i][2] = Y[4][i];
  }
  const T detJ = det_inv<T, 3>(J, Ji);
  const T kd = k * detJ;
  constexpr int mi[6] = {0, 1, 2, 0, 0, 1}, mj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    T acc = Ji[mi[m]][0] * Ji[mj[m]][0];
#pragma unroll
    for (int l = 1; l < 3; ++l) acc += Ji[mi[m]][l] * Ji[mj[m]][l];
    A[m] = kd * acc;
  }
  return detJ;
}

// Copies the packed reference table into LDS (all threads of the block take part).
template <typename T, int N>
__device__ __forceinline__ void stage_table(T* __restrict__ dst, const T* __restrict__ src) {
  for (int i = threadIdx.x; i < N; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
}

}  // namespace fem
