// Bricked layout of the 3-D half spectrum [L][L][L/2+1] (complex64) of a
// length-L transform (L = 2N for the padded volume of a box-N model, L = N
// unpadded), shared by the volume transform (writer) and the central-slice
// gather (reader): 2 (z) x 2 (y) x 4 (x) values per 128-byte brick, bricks in
// (z, y, x) order, x padded to 4 * XB = L/2 + 4 values.
// index = z-part + y-part + x-part.
#pragma once

#include <stddef.h>

template <int L>
struct VdftBricks {
  static constexpr int XB = L / 8 + 1;                          // bricks along x
  static constexpr size_t YSTRIDE = (size_t)XB * 16;            // one y-brick row
  static constexpr size_t ZSTRIDE = (size_t)(L / 2) * XB * 16;  // one z-brick plane (L/2 y-bricks)
  static constexpr size_t ELEMS = (size_t)L * L * XB * 4;
};

template <int L>
__device__ __forceinline__ size_t vdft_brick_z(int kz) {
  return (size_t)(kz >> 1) * VdftBricks<L>::ZSTRIDE + (size_t)(kz & 1) * 8;
}
template <int L>
__device__ __forceinline__ size_t vdft_brick_y(int ky) {
  return (size_t)(ky >> 1) * VdftBricks<L>::YSTRIDE + (size_t)(ky & 1) * 4;
}
__device__ __forceinline__ size_t vdft_brick_x(int kx) { return (size_t)(kx >> 2) * 16 + (kx & 3); }

// 32-bit forms of the same offsets (the spectrum holds < 2^32 values for every
// supported length), so a gather can address D with a scalar base and a
// 32-bit vector offset.
// (24-bit multiplies: the strides and brick counts fit, and v_mul_u32_u24 is
// full rate where a 32-bit multiply is not.)
template <int L>
__device__ __forceinline__ unsigned vdft_brick_z32(unsigned kz) {
  static_assert(VdftBricks<L>::ZSTRIDE < (1u << 24), "24-bit stride");
  return __umul24(kz >> 1, (unsigned)VdftBricks<L>::ZSTRIDE) + (kz & 1u) * 8u;
}
template <int L>
__device__ __forceinline__ unsigned vdft_brick_y32(unsigned ky) {
  static_assert(VdftBricks<L>::YSTRIDE < (1u << 24), "24-bit stride");
  return __umul24(ky >> 1, (unsigned)VdftBricks<L>::YSTRIDE) + (ky & 1u) * 4u;
}
__device__ __forceinline__ unsigned vdft_brick_x32(unsigned kx) { return (kx >> 2) * 16u + (kx & 3u); }
static_assert(VdftBricks<512>::ELEMS < (1ull << 32), "32-bit brick offsets");
