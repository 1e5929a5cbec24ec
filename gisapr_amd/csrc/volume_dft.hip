// 3-D transform of the model volume for central-slice projection.
//
// Computes, for a box-N volume, the half-spectrum
//     D = rfftn( fftshift( pad_{N/2}(vol) * sinc^2(|f|) ) )     [2N][2N][N+1]
// in natural frequency order (v606:1068-1073: pad N/2 zeros per side,
// apply_sinc2_filter_inplace_3d v606:837-853, fftshift, rfftn).  The final
// fftshift over (z, y) of v606:1073 is not materialised: the slice gather
// indexes D with the shift folded in.
//
// Three LDS Stockham passes (x: real-to-complex, then y, then z).  The zero
// padding is never read or written: the x pass touches only the N^2 rows
// that hold data, the y pass only the N data planes and synthesises the zero
// rows, and the z pass synthesises the zero planes.  Traffic per transform
// is ~ 4N^3 + 2*8*N^2(N+1) + 2*8*2N^2(N+1) bytes.
//
// Output layout (bricked): the x and y passes work on a natural-order
// workspace [2N][2N][N+1]; the z pass writes the result in bricks of
// 2 (z) x 2 (y) x 4 (x) complex values = one 128-byte cache line, so the 2x2x2
// trilinear cell of a central-slice sample spans ~2.8 lines instead of four
// 16-element x rows, whatever the slice orientation:
//   D[((((kz/2) * N + ky/2) * XB + kx/4) * 2 + kz%2) * 2 + ky%2) * 4 + kx%4],
//   XB = N/4 + 1 (x padded to N+4; the padding is never read).
#include "common.h"
#include "vdft_layout.h"

// LDS per workgroup (float2): 52.8 KB lets three workgroups share a CU.
// Sequences per workgroup are a multiple of the 4 waves (each wave owns
// whole sequences; see fft_lds_waves).
constexpr int kVdftLds = 6600;
constexpr int vdft_round4(int x) { return x >= 4 ? x & ~3 : x; }

// Row x' of the shifted padded volume holds padded column (x'+N) mod 2N,
// which is data column (x'+N) mod 2N - N/2 when that lies in [0, N).
__device__ __forceinline__ int shifted_to_data(int xs, int N) {
  const int p = (xs + N) % (2 * N);
  const int v = p - N / 2;
  return (v >= 0 && v < N) ? v : -1;
}

__device__ __forceinline__ float sinc2_axis_freq(int padded_index, int N) {
  // fftshift(fftfreq(2N))[p] = (p - N) / (2N)
  return (float)((double)(padded_index - N) / (double)(2 * N));
}

// x pass: two data rows per complex FFT (real a + i*b), split afterwards.
template <int N>
__global__ void __launch_bounds__(256)
vdft_x_pass(const float* __restrict__ vol, float2* __restrict__ D) {
  constexpr int L = 2 * N;
  constexpr int M = N + 1;
  constexpr int RB = vdft_round4((kVdftLds - L) / (2 * L));  // complex sequences per block
  extern __shared__ float2 lds[];
  float2* tw = lds;
  float2* bufA = tw + L;
  float2* bufB = bufA + RB * L;
  const int tid = threadIdx.x;
  constexpr int nthr = 256;  // launched with exactly 256 threads
  fill_twiddles(tw, L, tid, nthr);
  const int row_pair0 = blockIdx.x * RB;  // data row index = 2 * pair (+1)
  const int total_pairs = (N * N) / 2;
  const int nseq = min(RB, total_pairs - row_pair0);
  for (int t = tid; t < nseq * L; t += nthr) {
    const int q = t / L, xs = t - q * L;
    float v2[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int row = 2 * (row_pair0 + q) + h;  // row = vz * N + vy
      const int vz = row / N, vy = row - vz * N;
      const int vx = shifted_to_data(xs, N);
      float v = 0.0f;
      if (vx >= 0) {
        const float fz = sinc2_axis_freq(vz + N / 2, N);
        const float fy = sinc2_axis_freq(vy + N / 2, N);
        const float fx = sinc2_axis_freq(vx + N / 2, N);
        const float r = sqrtf((fz * fz + fy * fy) + fx * fx);
        float s = 1.0f;
        if (r != 0.0f) s = sinpif(r) / (3.14159265358979323846f * r);
        v = vol[((size_t)vz * N + vy) * N + vx] * (s * s);
      }
      v2[h] = v;
    }
    bufA[q * L + xs] = make_float2(v2[0], v2[1]);
  }
  __syncthreads();
  float2* res = fft_lds_waves<L, false>(bufA, bufB, nseq, L, tw, tid, nthr);
  // split: A[k] = (Z[k] + conj Z[L-k]) / 2 ; B[k] = (Z[k] - conj Z[L-k]) / (2i)
  for (int t = tid; t < nseq * M; t += nthr) {
    const int q = t / M, k = t - q * M;
    const float2 z = res[q * L + k];
    const float2 zc = cconj(res[q * L + ((L - k) % L)]);
    const float2 a = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y + zc.y));
    const float2 d = make_float2(0.5f * (z.x - zc.x), 0.5f * (z.y - zc.y));
    const float2 b = make_float2(d.y, -d.x);  // d / i
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int row = 2 * (row_pair0 + q) + h;
      const int vz = row / N, vy = row - vz * N;
      const int zs = (vz + N / 2 + N) % L;  // shifted plane index
      const int ys = (vy + N / 2 + N) % L;
      D[((size_t)zs * L + ys) * M + k] = h ? b : a;
    }
  }
}

// y pass over the N data planes: column FFTs along y (stride M).
template <int N>
__global__ void __launch_bounds__(256)
vdft_y_pass(float2* __restrict__ D) {
  constexpr int L = 2 * N;
  constexpr int M = N + 1;
  constexpr int TK = vdft_round4((kVdftLds - L) / (2 * (L + 1)));
  constexpr int SP = L + 1;
  extern __shared__ float2 lds[];
  float2* tw = lds;
  float2* bufA = tw + L;
  float2* bufB = bufA + TK * SP;
  const int tid = threadIdx.x;
  constexpr int nthr = 256;  // launched with exactly 256 threads
  fill_twiddles(tw, L, tid, nthr);
  const int vz = blockIdx.y;
  const int zs = (vz + N / 2 + N) % L;
  const int k0 = blockIdx.x * TK;
  const int nk = min(TK, M - k0);
  float2* P = D + (size_t)zs * L * M;
  for (int t = tid; t < L * TK; t += nthr) {
    const int ys = t / TK, q = t - ys * TK;
    float2 v = make_float2(0.f, 0.f);
    if (q < nk && shifted_to_data(ys, N) >= 0) v = P[(size_t)ys * M + k0 + q];
    bufA[q * SP + ys] = v;
  }
  __syncthreads();
  float2* res = fft_lds_waves<L, false>(bufA, bufB, nk, SP, tw, tid, nthr);
  for (int t = tid; t < L * TK; t += nthr) {
    const int ky = t / TK, q = t - ky * TK;
    if (q < nk) P[(size_t)ky * M + k0 + q] = res[q * SP + ky];
  }
}

// z pass: column FFTs along z (stride 2N*M) of the natural-order workspace W;
// zero planes are synthesised; the result goes to D in the bricked layout.
template <int N>
__global__ void __launch_bounds__(256)
vdft_z_pass(const float2* __restrict__ W, float2* __restrict__ D) {
  constexpr int L = 2 * N;
  constexpr int M = N + 1;
  constexpr int TK = vdft_round4((kVdftLds - L) / (2 * (L + 1)));
  constexpr int SP = L + 1;
  extern __shared__ float2 lds[];
  float2* tw = lds;
  float2* bufA = tw + L;
  float2* bufB = bufA + TK * SP;
  const int tid = threadIdx.x;
  constexpr int nthr = 256;  // launched with exactly 256 threads
  fill_twiddles(tw, L, tid, nthr);
  const int ky = blockIdx.y;
  const int k0 = blockIdx.x * TK;
  const int nk = min(TK, M - k0);
  const size_t plane = (size_t)L * M;
  for (int t = tid; t < L * TK; t += nthr) {
    const int zs = t / TK, q = t - zs * TK;
    float2 v = make_float2(0.f, 0.f);
    if (q < nk && shifted_to_data(zs, N) >= 0) v = W[zs * plane + (size_t)ky * M + k0 + q];
    bufA[q * SP + zs] = v;
  }
  __syncthreads();
  float2* res = fft_lds_waves<L, false>(bufA, bufB, nk, SP, tw, tid, nthr);
  const size_t ypart = vdft_brick_y<N>(ky);
  for (int t = tid; t < L * TK; t += nthr) {
    const int kz = t / TK, q = t - kz * TK;
    if (q < nk) D[vdft_brick_z<N>(kz) + ypart + vdft_brick_x(k0 + q)] = res[q * SP + kz];
  }
}

template <int N>
static int launch_volume_dft(const float* vol, float2* D, float2* Wk, hipStream_t st) {
  constexpr int L = 2 * N;
  constexpr int M = N + 1;
  constexpr int RB = vdft_round4((kVdftLds - L) / (2 * L));
  constexpr int TK = vdft_round4((kVdftLds - L) / (2 * (L + 1)));
  const int pairs = N * N / 2;
  const size_t lx = (size_t)(L + 2 * RB * L) * sizeof(float2);
  hipLaunchKernelGGL(vdft_x_pass<N>, dim3((pairs + RB - 1) / RB), dim3(256), lx, st, vol, Wk);
  GIS_LAUNCH_CHECK("gis_volume_dft x");
  const size_t lyz = (size_t)(L + 2 * TK * (L + 1)) * sizeof(float2);
  const int ktiles = (M + TK - 1) / TK;
  hipLaunchKernelGGL(vdft_y_pass<N>, dim3(ktiles, N), dim3(256), lyz, st, Wk);
  GIS_LAUNCH_CHECK("gis_volume_dft y");
  hipLaunchKernelGGL(vdft_z_pass<N>, dim3(ktiles, L), dim3(256), lyz, st, (const float2*)Wk, D);
  GIS_LAUNCH_CHECK("gis_volume_dft z");
  return GIS_OK;
}

static bool vdft_box_supported(int N) {
  return N == 32 || N == 64 || N == 96 || N == 128 || N == 192 || N == 256;
}

// Bytes of the bricked spectrum and of the natural-order workspace at box N.
GIS_API size_t gis_volume_dft_bytes(int N) {
  return N > 0 ? (size_t)4 * N * N * (N / 4 + 1) * 4 * sizeof(float2) : 0;
}
GIS_API size_t gis_volume_dft_workspace(int N) {
  return N > 0 ? (size_t)4 * N * N * (N + 1) * sizeof(float2) : 0;
}

// Replaces the pad / sinc^2 / fftshift / rfftn prefix of
// IntegratedProjector.build_templates (v606:1068-1073).
// `volume` is [N][N][N] float32 ([z][y][x]); `dft` receives the [2N][2N][N+1]
// complex64 half spectrum in natural frequency order, stored in 2x2x4 bricks
// (gis_volume_dft_bytes(N) bytes); `ws` holds gis_volume_dft_workspace(N) bytes.
GIS_API int gis_volume_dft(const float* volume, int N, void* dft, void* ws, size_t ws_bytes, void* stream) {
  GIS_REQUIRE(volume != nullptr && dft != nullptr && ws != nullptr, "gis_volume_dft: null pointer");
  if (!vdft_box_supported(N)) {
    gis_set_error("gis_volume_dft: box %d not in {32, 64, 96, 128, 192, 256}", N);
    return GIS_EUNSUPPORTED;
  }
  GIS_REQUIRE(ws_bytes >= gis_volume_dft_workspace(N), "gis_volume_dft: workspace %zu < %zu bytes", ws_bytes,
              gis_volume_dft_workspace(N));
  hipStream_t st = (hipStream_t)stream;
  float2* D = (float2*)dft;
  float2* W = (float2*)ws;
  switch (N) {
    case 32: return launch_volume_dft<32>(volume, D, W, st);
    case 64: return launch_volume_dft<64>(volume, D, W, st);
    case 96: return launch_volume_dft<96>(volume, D, W, st);
    case 128: return launch_volume_dft<128>(volume, D, W, st);
    case 192: return launch_volume_dft<192>(volume, D, W, st);
    default: return launch_volume_dft<256>(volume, D, W, st);
  }
}
