// Geometric restraint on the splat: threshold masks, an exact Euclidean
// distance transform, and the overlap / minimum-separation reduction of
// GeometryRestraint (func_check_boundary_for_testing_v8.py:62-126).
//
// Distance transform: d2[v] = min over mask voxels u of |v - u|^2 (squared
// voxel units) on an [N][N][N] ([z][y][x]) grid, separable in the axes:
//   pass x: along each x line, distance to the nearest mask voxel (forward
//           and backward sweep), squared;
//   pass y, pass z: along each line, d2'[i] = min_j (i - j)^2 + d2[j]
//           (the exact 1-D transform of the previous pass), one block per
//           line, the line staged in LDS, one output per thread.
// Squared distances are integers below 3 N^2 and exact in float32; lines
// without a mask voxel carry a large sentinel.  The result equals
// scipy.ndimage.distance_transform_edt(~mask)**2 (distance to the nearest
// True voxel).
#include "common.h"

namespace {

constexpr float kFar = 1.0e30f;

__global__ void threshold_kernel(const float* __restrict__ vol, size_t nvox, float thr,
                                 unsigned char* __restrict__ mask) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvox; i += (size_t)gridDim.x * blockDim.x)
    mask[i] = vol[i] > thr ? 1 : 0;
}

// one thread per x line: nearest mask voxel along the line, squared
__global__ void edt_x_kernel(const unsigned char* __restrict__ mask, int N, float* __restrict__ d2) {
  const size_t line = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= (size_t)N * N) return;
  const unsigned char* m = mask + line * N;
  float* o = d2 + line * N;
  int last = -1;
  for (int x = 0; x < N; ++x) {  // distance to the nearest mask voxel at or before x
    if (m[x]) last = x;
    o[x] = last < 0 ? kFar : (float)(x - last);
  }
  last = -1;
  for (int x = N - 1; x >= 0; --x) {  // ... or after x; then squared
    if (m[x]) last = x;
    float d = o[x];
    if (last >= 0) d = fminf(d, (float)(last - x));
    o[x] = d >= kFar ? kFar : d * d;
  }
}

// one block per line of the axis with element stride `stride`; lines are
// indexed by (outer, inner) with line base = outer * outer_stride + inner
__global__ void edt_line_kernel(float* __restrict__ d2, int N, int stride, int inner_count, size_t outer_stride) {
  extern __shared__ float f[];
  const int line = blockIdx.x;
  const int outer = line / inner_count, inner = line - outer * inner_count;
  float* base = d2 + (size_t)outer * outer_stride + inner;
  for (int i = threadIdx.x; i < N; i += blockDim.x) f[i] = base[(size_t)i * stride];
  __syncthreads();
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    float best = kFar;
    for (int j = 0; j < N; ++j) {
      const float fj = f[j];
      if (fj < kFar) {
        const float dj = (float)(i - j);
        best = fminf(best, dj * dj + fj);
      }
    }
    base[(size_t)i * stride] = best;
  }
}

__global__ void restraint_init_kernel(unsigned long long* overlap, unsigned int* min_d2, int* any) {
  *overlap = 0ull;
  *min_d2 = __float_as_uint(kFar);
  *any = 0;
}

// overlap = #(chain & main); min_d2 = min d2 over chain voxels; any = chain nonempty.
// d2 >= 0, so the float order equals the unsigned order of its bits.
__global__ void __launch_bounds__(256)
restraint_eval_kernel(const float* __restrict__ chain_vol, float thr, const unsigned char* __restrict__ main_mask,
                      const float* __restrict__ main_d2, size_t nvox, unsigned long long* __restrict__ overlap,
                      unsigned int* __restrict__ min_d2, int* __restrict__ any) {
  __shared__ unsigned int s_min[4];
  __shared__ unsigned long long s_cnt[4];
  __shared__ int s_any[4];
  unsigned long long cnt = 0;
  unsigned int mn = __float_as_uint(kFar);
  int seen = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvox; i += (size_t)gridDim.x * blockDim.x) {
    if (chain_vol[i] > thr) {
      seen = 1;
      cnt += main_mask[i] ? 1ull : 0ull;
      mn = min(mn, __float_as_uint(main_d2[i]));
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    cnt += __shfl_xor(cnt, o, 64);
    mn = min(mn, (unsigned int)__shfl_xor((int)mn, o, 64));
    seen |= __shfl_xor(seen, o, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_cnt[w] = cnt;
    s_min[w] = mn;
    s_any[w] = seen;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
      cnt += s_cnt[k];
      mn = min(mn, s_min[k]);
      seen |= s_any[k];
    }
    if (cnt) atomicAdd(overlap, cnt);
    if (seen) {
      atomicMin(min_d2, mn);
      atomicOr(any, 1);
    }
  }
}

int grid_for(size_t n, int threads) {
  const size_t cap = (size_t)gis_cu_count() * 16;
  return (int)std::max<size_t>(1, std::min<size_t>((n + threads - 1) / threads, cap));
}

}  // namespace

// mask = vol > thr (vol [n] float32, mask [n] uint8): the density masks of
// GeometryRestraint (func_check_boundary_for_testing_v8.py:84, :112).
GIS_API int gis_threshold_mask(const float* vol, size_t n, float thr, unsigned char* mask, void* stream) {
  if (n == 0) return GIS_OK;
  GIS_REQUIRE(vol && mask, "gis_threshold_mask: null pointer");
  hipLaunchKernelGGL(threshold_kernel, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, vol, n, thr,
                     mask);
  GIS_LAUNCH_CHECK("gis_threshold_mask");
  return GIS_OK;
}

// Squared Euclidean distance (voxels^2) from every voxel of an [N]^3 grid to
// the nearest voxel with mask != 0; 1e30 everywhere when the mask is empty.
// Replaces scipy.ndimage.distance_transform_edt(~mask)
// (func_check_boundary_for_testing_v8.py:90-92) up to the square root.
GIS_API int gis_edt_squared(const unsigned char* mask, int N, float* d2, void* stream) {
  GIS_REQUIRE(N >= 1 && N <= 1024, "gis_edt_squared: N=%d not in [1, 1024]", N);
  GIS_REQUIRE(mask && d2, "gis_edt_squared: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const size_t lines = (size_t)N * N;
  hipLaunchKernelGGL(edt_x_kernel, dim3((unsigned)((lines + 255) / 256)), dim3(256), 0, st, mask, N, d2);
  GIS_LAUNCH_CHECK("gis_edt_squared x");
  const int threads = std::min(N, 256);
  const size_t lds = (size_t)N * sizeof(float);
  // y lines: (z, x) -> base z * N^2 + x, stride N
  hipLaunchKernelGGL(edt_line_kernel, dim3((unsigned)lines), dim3(threads), lds, st, d2, N, N, N, (size_t)N * N);
  GIS_LAUNCH_CHECK("gis_edt_squared y");
  // z lines: (y, x) -> base y * N + x, stride N^2
  hipLaunchKernelGGL(edt_line_kernel, dim3((unsigned)lines), dim3(threads), lds, st, d2, N, N * N, N * N,
                     (size_t)0);
  GIS_LAUNCH_CHECK("gis_edt_squared z");
  return GIS_OK;
}

// One pose of GeometryRestraint.evaluate (func_check_boundary_for_testing_v8.py:96-126)
// on device data: chain_vol [n] float32 (the moved chain's splat), main_mask
// [n] uint8 and main_d2 [n] float32 (gis_edt_squared of it).  Writes
// overlap[0] = #(chain_vol > thr and main_mask), min_d2[0] = the smallest
// main_d2 over voxels with chain_vol > thr (as float32 bits; 1e30 if none),
// any[0] = 1 if some voxel has chain_vol > thr.
GIS_API int gis_restraint_eval(const float* chain_vol, float thr, const unsigned char* main_mask, const float* main_d2,
                               size_t n, unsigned long long* overlap, unsigned int* min_d2, int* any, void* stream) {
  GIS_REQUIRE(chain_vol && main_mask && main_d2 && overlap && min_d2 && any, "gis_restraint_eval: null pointer");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(restraint_init_kernel, dim3(1), dim3(1), 0, st, overlap, min_d2, any);
  GIS_LAUNCH_CHECK("gis_restraint_eval init");
  if (n == 0) return GIS_OK;
  hipLaunchKernelGGL(restraint_eval_kernel, dim3(grid_for(n, 256)), dim3(256), 0, st, chain_vol, thr, main_mask,
                     main_d2, n, overlap, min_d2, any);
  GIS_LAUNCH_CHECK("gis_restraint_eval");
  return GIS_OK;
}
