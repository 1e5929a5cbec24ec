// Element-wise helpers of the grid_sample-based projectors that callers use
// directly:
//
//   gis_sinc2_filter      volume[z,y,x] *= sinc^2(|f|), f = fftshift(fftfreq(n))
//                         per axis (v606:837-853; project3d_and_whiten_cuda_nowrite.py:42-81)
//   gis_slice_sample_grid rotated central-slice sampling grid for a 5-D
//                         F.grid_sample (align_corners=True) over an rfft
//                         volume, plus the Friedel-conjugation mask (v606:866-897)
//   gis_sample_dft_3d     trilinear, border-clamped samples of a complex volume
//                         at array coordinates (project3d_and_whiten_cuda.py:612-650)
//
// All three are HBM-bound streams: the filter reads and writes 8 bytes per
// voxel; the grid reads one base point per thread and writes 13 bytes per
// sample; the sampler reads 12 bytes of coordinates and eight 8-byte taps and
// writes 8 bytes per sample.  Coordinates are clamped before use, so NaN or
// infinite ones read the border, never outside the volume.
// Float arithmetic follows the reference op by op (float32 products, sums and
// divisions in the same order, a correctly rounded sqrt and division), so the
// results match torch's float32 values to a few ulp.
#include "common.h"

namespace {

constexpr int kGridThreads = 256;

// fftshift(fftfreq(n, dtype=float32))[i] = float(i - n/2) * float(1/n): torch
// forms fftfreq as arange * (1/n) in float32 and fftshift only reorders.
__device__ __forceinline__ float shifted_freq(int i, int n, float inv_n) { return (float)(i - n / 2) * inv_n; }

// torch.sinc(r)^2 in float32: sin(pi r) / (pi r), 1 at r = 0.
__device__ __forceinline__ float sinc_squared(float r) {
  if (r == 0.0f) return 1.0f;
  const float p = 3.14159265358979323846f * r;
  const float s = sinf(p) / p;
  return s * s;
}

// One thread per 4 consecutive x of one (z, y) row when W % 4 == 0 (16-byte
// loads and stores), one per voxel otherwise; grid-stride over the volume.
template <int VEC>
__global__ __launch_bounds__(kGridThreads) void sinc2_filter_kernel(float* __restrict__ vol, int D, int H, int W,
                                                                    float inv_d, float inv_h, float inv_w) {
  const size_t nvec = (size_t)D * H * (W / VEC);
  const int wv = W / VEC;
  for (size_t t = (size_t)blockIdx.x * kGridThreads + threadIdx.x; t < nvec;
       t += (size_t)gridDim.x * kGridThreads) {
    const size_t row = t / wv;
    const int x0 = (int)(t - row * wv) * VEC;
    const int z = (int)(row / H), y = (int)(row - (size_t)z * H);
    const float fz = shifted_freq(z, D, inv_d), fy = shifted_freq(y, H, inv_h);
    const float fzy = fz * fz + fy * fy;  // (fz^2 + fy^2) + fx^2, torch's broadcast order
    float* p = vol + row * W + x0;
    if constexpr (VEC == 4) {
      float4 v = *reinterpret_cast<const float4*>(p);
      float fx;
      fx = shifted_freq(x0 + 0, W, inv_w); v.x *= sinc_squared(sqrtf(fzy + fx * fx));
      fx = shifted_freq(x0 + 1, W, inv_w); v.y *= sinc_squared(sqrtf(fzy + fx * fx));
      fx = shifted_freq(x0 + 2, W, inv_w); v.z *= sinc_squared(sqrtf(fzy + fx * fx));
      fx = shifted_freq(x0 + 3, W, inv_w); v.w *= sinc_squared(sqrtf(fzy + fx * fx));
      *reinterpret_cast<float4*>(p) = v;
    } else {
      const float fx = shifted_freq(x0, W, inv_w);
      p[0] *= sinc_squared(sqrtf(fzy + fx * fx));
    }
  }
}

// One thread per (orientation m, base point q).  base_zyx[q] = (z, y, x)
// frequencies; R[m] acts on xyz column vectors.
__global__ __launch_bounds__(kGridThreads) void slice_sample_grid_kernel(
    const float* __restrict__ base_zyx, int P, const float* __restrict__ R, int M, float img_d, float img_h,
    float img_w, float half_d, float half_h, float den_d, float den_h, float den_w, float* __restrict__ grid_xyz,
    unsigned char* __restrict__ conj) {
  const size_t total = (size_t)M * P;
  for (size_t t = (size_t)blockIdx.x * kGridThreads + threadIdx.x; t < total;
       t += (size_t)gridDim.x * kGridThreads) {
    const int m = (int)(t / P), q = (int)(t - (size_t)m * P);
    const float bz = base_zyx[3 * (size_t)q + 0], by = base_zyx[3 * (size_t)q + 1],
                bx = base_zyx[3 * (size_t)q + 2];
    const float* r = R + 9 * (size_t)m;
    // rotated xyz = R @ (x, y, z)
    const float rx = fmaf(r[2], bz, fmaf(r[1], by, r[0] * bx));
    const float ry = fmaf(r[5], bz, fmaf(r[4], by, r[3] * bx));
    const float rz = fmaf(r[8], bz, fmaf(r[7], by, r[6] * bx));
    // sample the Friedel mate where the rotated x frequency is negative
    const bool c = rx < 0.0f;
    const float gz = c ? -rz : rz, gy = c ? -ry : ry, gx = c ? -rx : rx;
    // array coordinates (z and y centred, x from 0), then align_corners=True
    // normalisation by (size/2 - 1/2), written in grid_sample's x, y, z order
    const float cz = gz * img_d + half_d, cy = gy * img_h + half_h, cx = gx * img_w;
    float* g = grid_xyz + 3 * t;
    g[0] = cx / den_w - 1.0f;
    g[1] = cy / den_h - 1.0f;
    g[2] = cz / den_d - 1.0f;
    conj[t] = c ? 1 : 0;
  }
}

// One thread per sample: trilinear interpolation of a natural-layout complex
// [D][H][W] volume at array coordinates (z, y, x), each clamped to
// [0, size - 1] first (border padding); taps beyond the last index carry
// weight 0 and are read clamped.
__global__ __launch_bounds__(kGridThreads) void sample_dft_3d_kernel(const float2* __restrict__ dft, int D, int H,
                                                                     int W, const float* __restrict__ coords,
                                                                     size_t P, float2* __restrict__ out) {
  for (size_t t = (size_t)blockIdx.x * kGridThreads + threadIdx.x; t < P; t += (size_t)gridDim.x * kGridThreads) {
    const float cz = fminf(fmaxf(coords[3 * t + 0], 0.0f), (float)(D - 1));
    const float cy = fminf(fmaxf(coords[3 * t + 1], 0.0f), (float)(H - 1));
    const float cx = fminf(fmaxf(coords[3 * t + 2], 0.0f), (float)(W - 1));
    const float fz = floorf(cz), fy = floorf(cy), fx = floorf(cx);
    const float tz = cz - fz, ty = cy - fy, tx = cx - fx;
    const int z0 = (int)fz, y0 = (int)fy, x0 = (int)fx;
    const int z1 = min(z0 + 1, D - 1), y1 = min(y0 + 1, H - 1), x1 = min(x0 + 1, W - 1);
    auto at = [&](int z, int y, int x) { return dft[((size_t)z * H + y) * W + x]; };
    const float2 v000 = at(z0, y0, x0), v001 = at(z0, y0, x1), v010 = at(z0, y1, x0), v011 = at(z0, y1, x1);
    const float2 v100 = at(z1, y0, x0), v101 = at(z1, y0, x1), v110 = at(z1, y1, x0), v111 = at(z1, y1, x1);
    const float wz0 = 1.0f - tz, wy0 = 1.0f - ty, wx0 = 1.0f - tx;
    float2 acc = make_float2(0.f, 0.f);
    auto tap = [&](float2 v, float w) {
      acc.x += w * v.x;
      acc.y += w * v.y;
    };
    tap(v000, (wz0 * wy0) * wx0);
    tap(v001, (wz0 * wy0) * tx);
    tap(v010, (wz0 * ty) * wx0);
    tap(v011, (wz0 * ty) * tx);
    tap(v100, (tz * wy0) * wx0);
    tap(v101, (tz * wy0) * tx);
    tap(v110, (tz * ty) * wx0);
    tap(v111, (tz * ty) * tx);
    out[t] = acc;
  }
}

int grid_blocks(size_t work) {
  const size_t want = (work + kGridThreads - 1) / kGridThreads;
  const size_t cap = (size_t)gis_cu_count() * 32;
  return (int)std::max<size_t>(1, std::min(want, cap));
}

}  // namespace

// In-place sinc^2 window of a [D][H][W] float32 volume (C order).
GIS_API int gis_sinc2_filter(float* vol, int D, int H, int W, hipStream_t stream) {
  GIS_REQUIRE(vol != nullptr, "gis_sinc2_filter: null volume");
  GIS_REQUIRE(D > 0 && H > 0 && W > 0, "gis_sinc2_filter: shape (%d, %d, %d) must be positive", D, H, W);
  GIS_REQUIRE((size_t)D * H * W < (1ull << 40), "gis_sinc2_filter: volume too large");
  const float id = 1.0f / (float)D, ih = 1.0f / (float)H, iw = 1.0f / (float)W;
  const bool vec = (W % 4 == 0) && ((uintptr_t)vol % 16 == 0);
  if (vec) {
    hipLaunchKernelGGL(sinc2_filter_kernel<4>, dim3(grid_blocks((size_t)D * H * (W / 4))), dim3(kGridThreads), 0,
                       stream, vol, D, H, W, id, ih, iw);
  } else {
    hipLaunchKernelGGL(sinc2_filter_kernel<1>, dim3(grid_blocks((size_t)D * H * W)), dim3(kGridThreads), 0, stream,
                       vol, D, H, W, id, ih, iw);
  }
  GIS_LAUNCH_CHECK("gis_sinc2_filter");
  return GIS_OK;
}

// Sampling grid of M rotated copies of P base points for an rfft volume of
// shape dft_shape, sampled as the real image of shape image_shape.
GIS_API int gis_slice_sample_grid(const float* base_zyx, int P, const float* R, int M, int img_d, int img_h,
                                  int img_w, int dft_d, int dft_h, int dft_w, float* grid_xyz, unsigned char* conj,
                                  hipStream_t stream) {
  GIS_REQUIRE(base_zyx && R && grid_xyz && conj, "gis_slice_sample_grid: null pointer");
  GIS_REQUIRE(P > 0 && M > 0, "gis_slice_sample_grid: P = %d and M = %d must be positive", P, M);
  GIS_REQUIRE(img_d > 0 && img_h > 0 && img_w > 0, "gis_slice_sample_grid: image shape must be positive");
  GIS_REQUIRE(dft_d > 1 && dft_h > 1 && dft_w > 1, "gis_slice_sample_grid: every dft dimension must exceed 1");
  GIS_REQUIRE((size_t)M * P < (1ull << 40), "gis_slice_sample_grid: grid too large");
  const float den_d = 0.5f * (float)dft_d - 0.5f, den_h = 0.5f * (float)dft_h - 0.5f,
              den_w = 0.5f * (float)dft_w - 0.5f;
  hipLaunchKernelGGL(slice_sample_grid_kernel, dim3(grid_blocks((size_t)M * P)), dim3(kGridThreads), 0, stream,
                     base_zyx, P, R, M, (float)img_d, (float)img_h, (float)img_w, (float)(img_d / 2),
                     (float)(img_h / 2), den_d, den_h, den_w, grid_xyz, conj);
  GIS_LAUNCH_CHECK("gis_slice_sample_grid");
  return GIS_OK;
}

// Trilinear samples of a complex [D][H][W] volume (natural layout, complex64)
// at P array coordinates coords [P][3] (z, y, x; border-clamped), into out
// [P] complex64: sample_dft_3d of project3d_and_whiten_cuda.py:612-650 (its
// grid_sample with align_corners=True and border padding on the real and
// imaginary channels).
GIS_API int gis_sample_dft_3d(const void* dft, int D, int H, int W, const float* coords, size_t P, void* out,
                              hipStream_t stream) {
  GIS_REQUIRE(dft && coords && out, "gis_sample_dft_3d: null pointer");
  GIS_REQUIRE(D > 0 && H > 0 && W > 0, "gis_sample_dft_3d: shape (%d, %d, %d) must be positive", D, H, W);
  if (P == 0) return GIS_OK;
  hipLaunchKernelGGL(sample_dft_3d_kernel, dim3(grid_blocks(P)), dim3(kGridThreads), 0, stream, (const float2*)dft, D,
                     H, W, coords, P, (float2*)out);
  GIS_LAUNCH_CHECK("gis_sample_dft_3d");
  return GIS_OK;
}
