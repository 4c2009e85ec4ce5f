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

// Copies the packed reference table into LDS (all threads of the block take part).
template <typename T, int N>
__device__ __forceinline__ void stage_table(T* __restrict__ dst, const T* __restrict__ src) {
  for (int i = threadIdx.x; i < N; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
}

}  // namespace fem
