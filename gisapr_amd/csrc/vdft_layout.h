// Bricked layout of the 3-D half spectrum [2N][2N][N+1] (complex64) shared by
// the volume transform (writer) and the central-slice gather (reader):
// 2 (z) x 2 (y) x 4 (x) values per 128-byte brick, bricks in (z, y, x) order,
// x padded to 4 * XB = N + 4 values.  index = z-part + y-part + x-part.
#pragma once

#include <stddef.h>

template <int N>
struct VdftBricks {
  static constexpr int XB = N / 4 + 1;                         // bricks along x
  static constexpr size_t YSTRIDE = (size_t)XB * 16;            // one y-brick row
  static constexpr size_t ZSTRIDE = (size_t)N * XB * 16;        // one z-brick plane (N y-bricks)
  static constexpr size_t ELEMS = (size_t)2 * N * 2 * N * XB * 4;
};

template <int N>
__device__ __forceinline__ size_t vdft_brick_z(int kz) {
  return (size_t)(kz >> 1) * VdftBricks<N>::ZSTRIDE + (size_t)(kz & 1) * 8;
}
template <int N>
__device__ __forceinline__ size_t vdft_brick_y(int ky) {
  return (size_t)(ky >> 1) * VdftBricks<N>::YSTRIDE + (size_t)(ky & 1) * 4;
}
__device__ __forceinline__ size_t vdft_brick_x(int kx) { return (size_t)(kx >> 2) * 16 + (kx & 3); }

// 32-bit forms of the same offsets (the spectrum holds < 2^32 values for every
// supported box), so a gather can address D with a scalar base and a 32-bit
// vector offset.
template <int N>
__device__ __forceinline__ unsigned vdft_brick_z32(unsigned kz) {
  return (kz >> 1) * (unsigned)VdftBricks<N>::ZSTRIDE + (kz & 1u) * 8u;
}
template <int N>
__device__ __forceinline__ unsigned vdft_brick_y32(unsigned ky) {
  return (ky >> 1) * (unsigned)VdftBricks<N>::YSTRIDE + (ky & 1u) * 4u;
}
__device__ __forceinline__ unsigned vdft_brick_x32(unsigned kx) { return (kx >> 2) * 16u + (kx & 3u); }
static_assert(VdftBricks<256>::ELEMS < (1ull << 32), "32-bit brick offsets");
