// Pieces of the cell-coefficient form of the scaled operator (fem_solve.hip): Ferrite's corner
// order, the row arithmetic of a box-cell lattice, and raw buffer loads whose range check
// returns 0.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace fem {

// Local index of the corner at offset (ex, ey, ez) ∈ {0,1}³ from a cell's first corner, in
// Ferrite's corner order ((0,0,0) (1,0,0) (1,1,0) (0,1,0), then the same at z = 1).
__host__ __device__ constexpr int cell_lid(int ex, int ey, int ez) {
  return ez * 4 + (ey ? 3 - ex : ex);
}

// The SpMV's K¹ for box cells: K¹[a][b] = f[|Δx| + 2 |Δy| + 4 |Δz|] of the corners' offsets
// (eight values in 3D, four in 2D; the host checks K¹ has this form).
template <typename X>
struct CellKrefT {
  X f[8];
};

// One row of the cell form from the t·x of its neighbours zn ((dz, dy, dx) row-major, offsets
// −1..1) and the coefficients k[e] of the cells around it (cell with origin P − e, e = ex + 2 ey
// + 4 ez): the neighbour at offset d lies in the cells whose e has e_a = 0 where d_a = +1 and
// e_a = 1 where d_a = −1 (either where d_a = 0), with K¹ value f[|dx| + 2 |dy| + 4 |dz|], so
//   y = Σ_m f[m] Σ_{|d| = m} (Σ of k over those cells) zn[d];
// the sums of k over cell pairs and quadruples are shared between neighbours (54 operations
// instead of 8 × 8 + 8).
template <typename X, int DIM>
__device__ __forceinline__ X cell_row(const X* zn, const X* k, const X* f) {
  if constexpr (DIM == 3) {
    auto Z = [&](int dx, int dy, int dz) { return zn[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)]; };
    const X z00 = k[0] + k[4], z10 = k[1] + k[5], z01 = k[2] + k[6], z11 = k[3] + k[7];   // over ez: [ex][ey]
    const X y00 = k[0] + k[2], y10 = k[1] + k[3], y01 = k[4] + k[6], y11 = k[5] + k[7];   // over ey: [ex][ez]
    const X x00 = k[0] + k[1], x10 = k[2] + k[3], x01 = k[4] + k[5], x11 = k[6] + k[7];   // over ex: [ey][ez]
    const X yz0 = z00 + z01, yz1 = z10 + z11;   // over ey, ez: [ex]
    const X xz0 = z00 + z10, xz1 = z01 + z11;   // over ex, ez: [ey]
    const X xy0 = x00 + x10, xy1 = x01 + x11;   // over ex, ey: [ez]
    const X t0 = (yz0 + yz1) * Z(0, 0, 0);
    const X t1 = yz0 * Z(1, 0, 0) + yz1 * Z(-1, 0, 0);
    const X t2 = xz0 * Z(0, 1, 0) + xz1 * Z(0, -1, 0);
    const X t4 = xy0 * Z(0, 0, 1) + xy1 * Z(0, 0, -1);
    const X t3 = z00 * Z(1, 1, 0) + z10 * Z(-1, 1, 0) + z01 * Z(1, -1, 0) + z11 * Z(-1, -1, 0);
    const X t5 = y00 * Z(1, 0, 1) + y10 * Z(-1, 0, 1) + y01 * Z(1, 0, -1) + y11 * Z(-1, 0, -1);
    const X t6 = x00 * Z(0, 1, 1) + x10 * Z(0, -1, 1) + x01 * Z(0, 1, -1) + x11 * Z(0, -1, -1);
    X t7 = X(0);
#pragma unroll
    for (int e = 0; e < 8; ++e) t7 += k[e] * Z(1 - 2 * (e & 1), 1 - 2 * ((e >> 1) & 1), 1 - 2 * (e >> 2));
    return f[0] * t0 + f[1] * t1 + f[2] * t2 + f[3] * t3 + f[4] * t4 + f[5] * t5 + f[6] * t6 + f[7] * t7;
  } else {
    auto Z = [&](int dx, int dy) { return zn[(dy + 1) * 3 + (dx + 1)]; };
    const X y0 = k[0] + k[2], y1 = k[1] + k[3];   // over ey: [ex]
    const X x0 = k[0] + k[1], x1 = k[2] + k[3];   // over ex: [ey]
    const X t0 = (y0 + y1) * Z(0, 0);
    const X t1 = y0 * Z(1, 0) + y1 * Z(-1, 0);
    const X t2 = x0 * Z(0, 1) + x1 * Z(0, -1);
    X t3 = X(0);
#pragma unroll
    for (int e = 0; e < 4; ++e) t3 += k[e] * Z(1 - 2 * (e & 1), 1 - 2 * (e >> 1));
    return f[0] * t0 + f[1] * t1 + f[2] * t2 + f[3] * t3;
  }
}

// Raw buffer resource over `bytes` bytes from `base`: a load at a byte offset at or beyond
// `bytes` (the offset is unsigned, so a negative one too) returns 0 without touching memory.
template <typename X>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const X* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<X*>(base), (short)0, (int)bytes, 0x00020000);
}

template <typename X>
__device__ __forceinline__ X buf_load(__amdgpu_buffer_rsrc_t r, uint32_t boff) {
  if constexpr (sizeof(X) == 4) return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)boff, 0, 0));
  else return __longlong_as_double((long long)__builtin_amdgcn_raw_buffer_load_b64(r, (int)boff, 0, 0));
}

// Store of one value at a byte offset; at or beyond the resource's extent it is dropped.
template <typename X>
__device__ __forceinline__ void buf_store(X v, __amdgpu_buffer_rsrc_t r, uint32_t boff) {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  if constexpr (sizeof(X) == 4) __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)boff, 0, 0);
  else __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)boff, 0, 0);
}

// The resource for the rows [row0, npad) of a vector (none when row0 is outside [0, npad)).
// 32-bit row arithmetic (npad · sizeof(X) < 2³¹), which stays on the scalar unit.
template <typename X>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rows_from(const X* v, int32_t row0, int32_t npad) {
  const bool in = row0 >= 0 && row0 < npad;
  return buf_rsrc(in ? v + row0 : v, in ? (uint32_t)(npad - row0) * (uint32_t)sizeof(X) : 0u);
}

}  // namespace fem
