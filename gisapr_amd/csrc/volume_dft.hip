// 3-D transform of the model volume for central-slice projection.
//
// Computes, for a box-N volume and transform length L (L = 2N: the volume
// padded with N/2 zeros per side; L = N: unpadded), the half spectrum
//     D = rfftn( fftshift( pad(vol) * sinc^2(|f|) ) )     [L][L][L/2+1]
// in natural frequency order, f = fftshift(fftfreq(L)) per axis
// (v606:1068-1073: pad N/2, apply_sinc2_filter_inplace_3d v606:837-853,
// fftshift, rfftn; project3d_and_whiten_cuda.py:103-117, whose --skip_padding
// gives L = N).  The final fftshift over (z, y) is not materialised: the
// slice gather indexes D with the shift folded in.
//
// Three LDS Stockham passes (x: real-to-complex, then y, then z).  The zero
// padding is never read or written: the x pass touches only the N^2 rows
// that hold data, the y pass only the N data planes and synthesises the zero
// rows, and the z pass synthesises the zero planes.  Every pass moves whole
// 128-byte lines: a block transforms 16 columns (two groups of 8 kx, or a
// pair of ky rows x 8 kx) in place in one LDS buffer.  Traffic per
// transform ~ 4N^3 + 8*N*(L/2)*G*16*(N/L + N/L + 1 + 1) + 8*L*L*(L/2+4) bytes
// (G = ceil((L/2+1)/8)): 1.44 GB at box 256, L = 512.
//
// Output layout (bricked, vdft_layout.h): the x and y passes work on a
// workspace of the N data planes (see VdftWork); the z pass writes the
// result in bricks of 2 (z) x 2 (y) x 4 (x) complex values = one 128-byte
// cache line, so the 2x2x2 trilinear cell of a central-slice sample spans
// ~2.8 lines instead of four 16-element x rows, whatever the orientation:
//   D[((((kz/2) * L/2 + ky/2) * XB + kx/4) * 2 + kz%2) * 2 + ky%2) * 4 + kx%4],
//   XB = L/8 + 1 (x padded to L/2+4; the padding is never read).
#include "common.h"
#include "vdft_layout.h"

// Position xs of the fftshifted length-L axis holds padded position
// (xs + L/2) mod L, which is data position (xs + L/2) mod L - (L-N)/2 when
// that lies in [0, N).
template <int N, int L>
__device__ __forceinline__ int shifted_to_data(int xs) {
  const int v = (xs + L / 2) % L - (L - N) / 2;
  return (v >= 0 && v < N) ? v : -1;
}
// Shifted position of data position v.
template <int N, int L>
__device__ __forceinline__ int data_to_shifted(int v) { return (v + (L - N) / 2 + L / 2) % L; }

template <int N, int L>
__device__ __forceinline__ float sinc2_axis_freq(int v) {
  // fftshift(fftfreq(L))[p] = (p - L/2) / L at padded position p = v + (L-N)/2
  return (float)((double)(v + (L - N) / 2 - L / 2) / (double)L);
}

// Workspace of the x and y passes: the N data planes (vz) of the
// natural-order spectrum, all L rows (ky, shifted order), kx in groups of
// eight; a pair of rows of one group is one 128-byte line:
//   W[(((vz * L/2 + ky/2) * G + kx/8) * 2 + ky%2) * 8 + kx%8],  G = ceil((L/2+1)/8).
template <int N, int L>
struct VdftWork {
  static constexpr int G = (L / 2 + 8) / 8;
  static constexpr size_t ELEMS = (size_t)N * (L / 2) * G * 16;
};

// Every pass: 512 threads (8 waves), 16 sequences of length L per block in
// one LDS buffer (each wave transforms its sequences in place, registers
// holding a stage's butterflies), ~78 KB at L = 512 so two blocks share a CU.
constexpr int kVdftThreads = 512;
constexpr int kVdftSeq = 16;
constexpr int kVdftPad = 3;  // one LDS slot of padding per 8 values
template <int L>
constexpr int vdft_sp() { return lds_padded_len<L, kVdftPad>() + 1; }
template <int L>
constexpr size_t vdft_lds_bytes() { return (size_t)(L + kVdftSeq * vdft_sp<L>()) * sizeof(float2); }

// x pass: two data rows per complex FFT (real a + i*b), split afterwards.
template <int N, int L>
__global__ void __launch_bounds__(kVdftThreads)
vdft_x_pass(const float* __restrict__ vol, float2* __restrict__ Wk) {
  constexpr int SP = vdft_sp<L>();
  constexpr int G = VdftWork<N, L>::G;
  static_assert(L % 32 == 0 && N % 32 == 0 && (L == N || L == 2 * N), "transform length");
  extern __shared__ float2 lds[];
  float2* tw = lds;
  float2* buf = tw + L;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int nthr = kVdftThreads;
  fill_twiddles(tw, L, tid, nthr);
  const int pair0 = blockIdx.x * kVdftSeq;  // data row index = 2 * pair (+1)
  constexpr int nseq = kVdftSeq;             // N^2 / 2 is a multiple of 16
  static_assert((N * N / 2) % kVdftSeq == 0, "row pairs per block");
  // data columns (q, h, vx), vx fastest: unconditional loads, unrolled, so
  // all of a thread's loads are in flight together
#pragma unroll
  for (int it = 0; it < kVdftSeq * 2 * N / nthr; ++it) {
    const int t = tid + it * nthr;
    const int vx = t % N, qh = t / N, q = qh >> 1, h = qh & 1;
    const int row = 2 * (pair0 + q) + h;  // row = vz * N + vy
    const int vz = row / N, vy = row - vz * N;
    const float raw = vol[((size_t)vz * N + vy) * N + vx];
    const float fz = sinc2_axis_freq<N, L>(vz);
    const float fy = sinc2_axis_freq<N, L>(vy);
    const float fx = sinc2_axis_freq<N, L>(vx);
    const float r = sqrtf((fz * fz + fy * fy) + fx * fx);
    float sn = 1.0f;
    if (r != 0.0f) sn = sinpif(r) / (3.14159265358979323846f * r);
    float* dst = (float*)&buf[q * SP + lds_pad<kVdftPad>(data_to_shifted<N, L>(vx))];
    dst[h] = raw * (sn * sn);
  }
  // zero columns: the L - N padded positions outside the data
  if constexpr (L > N) {
    constexpr int OFF = (L - N) / 2;
    for (int t = tid; t < kVdftSeq * (L - N); t += nthr) {
      const int q = t / (L - N), j = t - q * (L - N);
      const int pp = j < OFF ? j : j + N;
      buf[q * SP + lds_pad<kVdftPad>((pp + L / 2) % L)] = make_float2(0.f, 0.f);
    }
  }
  __syncthreads();
  for (int q = wave; q < nseq; q += nthr / 64) fft_wave_inplace<L, false, kVdftPad>(buf + q * SP, tw, lane);
  __syncthreads();
  // split: A[k] = (Z[k] + conj Z[L-k]) / 2 ; B[k] = (Z[k] - conj Z[L-k]) / (2i).
  // Rows 2p and 2p+1 are workspace rows ys and ys+1 (ys even): 16
  // consecutive threads fill one line (row parity h, 8 kx).
  for (int t = tid; t < nseq * G * 16; t += nthr) {
    const int q = t / (G * 16), r = t - q * (G * 16);
    const int g = r >> 4, h = (r >> 3) & 1, k = g * 8 + (r & 7);
    if (k > L / 2) continue;
    const float2 z = buf[q * SP + lds_pad<kVdftPad>(k)];
    const float2 zc = cconj(buf[q * SP + lds_pad<kVdftPad>((L - k) % L)]);
    float2 val;
    if (h == 0) {
      val = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y + zc.y));
    } else {
      const float2 d = make_float2(0.5f * (z.x - zc.x), 0.5f * (z.y - zc.y));
      val = make_float2(d.y, -d.x);  // d / i
    }
    const int row = 2 * (pair0 + q);
    const int vz = row / N, vy = row - vz * N;
    const int ys = data_to_shifted<N, L>(vy);
    Wk[(((size_t)vz * (L / 2) + ys / 2) * G + g) * 16 + h * 8 + (k & 7)] = val;
  }
}

// y pass: column FFTs along ky over the N data rows of one data plane, 16
// kx columns (two groups) per block; the zero rows are synthesised.
template <int N, int L>
__global__ void __launch_bounds__(kVdftThreads)
vdft_y_pass(float2* __restrict__ Wk) {
  constexpr int SP = vdft_sp<L>();
  constexpr int G = VdftWork<N, L>::G;
  extern __shared__ float2 lds[];
  float2* tw = lds;
  float2* buf = tw + L;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int nthr = kVdftThreads;
  fill_twiddles(tw, L, tid, nthr);
  const int vz = blockIdx.y;
  const int g0 = blockIdx.x * 2;
  const int ng = min(2, G - g0);
  const int ncol = min(kVdftSeq, L / 2 + 1 - g0 * 8);  // valid kx columns of the tile
  float2* P = Wk + (size_t)vz * (L / 2) * G * 16;
  // 32 threads per row pair: (group, row parity, kx % 8) = two adjacent lines
#pragma unroll
  for (int it = 0; it < (L / 2) * 32 / nthr; ++it) {
    const int t = tid + it * nthr;
    const int p = t >> 5, r = t & 31;
    const int gl = r >> 4, h = (r >> 3) & 1, c = gl * 8 + (r & 7);
    const int ys = 2 * p + h;
    float2 v = make_float2(0.f, 0.f);
    if (gl < ng && shifted_to_data<N, L>(ys) >= 0) v = P[((size_t)p * G + g0 + gl) * 16 + (r & 15)];
    buf[c * SP + lds_pad<kVdftPad>(ys)] = v;
  }
  __syncthreads();
  for (int q = wave; q < ncol; q += nthr / 64) fft_wave_inplace<L, false, kVdftPad>(buf + q * SP, tw, lane);
  __syncthreads();
#pragma unroll
  for (int it = 0; it < (L / 2) * 32 / nthr; ++it) {
    const int t = tid + it * nthr;
    const int p = t >> 5, r = t & 31;
    const int gl = r >> 4, h = (r >> 3) & 1, c = gl * 8 + (r & 7);
    if (c < ncol) P[((size_t)p * G + g0 + gl) * 16 + (r & 15)] = buf[c * SP + lds_pad<kVdftPad>(2 * p + h)];
  }
}

// z pass: column FFTs along kz for one pair of ky rows and one group of 8 kx
// (16 columns); the zero planes are synthesised; the result goes to D in
// the bricked layout, two whole bricks per kz pair.
template <int N, int L>
__global__ void __launch_bounds__(kVdftThreads)
vdft_z_pass(const float2* __restrict__ Wk, float2* __restrict__ D) {
  constexpr int SP = vdft_sp<L>();
  constexpr int G = VdftWork<N, L>::G;
  using BR = VdftBricks<L>;
  extern __shared__ float2 lds[];
  float2* tw = lds;
  float2* buf = tw + L;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int nthr = kVdftThreads;
  fill_twiddles(tw, L, tid, nthr);
  const int g = blockIdx.x, p = blockIdx.y;  // kx group, ky pair
  const int nx = min(8, L / 2 + 1 - g * 8);  // valid kx of the group
  // column c = (ky % 2) * 8 + kx % 8
#pragma unroll
  for (int it = 0; it < L * 16 / nthr; ++it) {
    const int t = tid + it * nthr;
    const int zs = t >> 4, c = t & 15;
    const int vz = shifted_to_data<N, L>(zs);
    float2 v = make_float2(0.f, 0.f);
    if (vz >= 0 && (c & 7) < nx) v = Wk[(((size_t)vz * (L / 2) + p) * G + g) * 16 + c];
    buf[c * SP + lds_pad<kVdftPad>(zs)] = v;
  }
  __syncthreads();
  for (int q = wave; q < 16; q += nthr / 64)
    if ((q & 7) < nx) fft_wave_inplace<L, false, kVdftPad>(buf + q * SP, tw, lane);
  __syncthreads();
  // 32 threads per kz pair: (x brick 2g or 2g+1, kz % 2, ky % 2, kx % 4)
#pragma unroll
  for (int it = 0; it < (L / 2) * 32 / nthr; ++it) {
    const int t = tid + it * nthr;
    const int kp = t >> 5, r = t & 31;
    const int xb = r >> 4, kzp = (r >> 3) & 1, h = (r >> 2) & 1, x4 = r & 3;
    const int x8 = xb * 4 + x4;
    if (x8 >= nx) continue;
    float2* dst = D + (size_t)kp * BR::ZSTRIDE + (size_t)p * BR::YSTRIDE + (size_t)(2 * g + xb) * 16 + (r & 15);
    const float2 v = buf[(h * 8 + x8) * SP + lds_pad<kVdftPad>(2 * kp + kzp)];
    // streaming store: the spectrum is written once and read by a later
    // launch, so it need not displace cached lines (6 % faster at box 256)
    typedef float f2v __attribute__((ext_vector_type(2)));
    const f2v w = {v.x, v.y};
    __builtin_nontemporal_store(w, (f2v*)dst);
  }
}

template <int N, int L>
static int launch_volume_dft(const float* vol, float2* D, float2* Wk, hipStream_t st) {
  constexpr int G = VdftWork<N, L>::G;
  const size_t lds = vdft_lds_bytes<L>();
  static bool attr_set = false;
  if (lds > 65536 && !attr_set) {
    const void* fns[3] = {(const void*)vdft_x_pass<N, L>, (const void*)vdft_y_pass<N, L>,
                          (const void*)vdft_z_pass<N, L>};
    for (const void* f : fns) {
      hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        gis_set_error("gis_volume_dft: cannot reserve %zu bytes of LDS: %s", lds, hipGetErrorString(e));
        return (int)e;
      }
    }
    attr_set = true;
  }
  const int pairs = N * N / 2;
  hipLaunchKernelGGL((vdft_x_pass<N, L>), dim3((pairs + kVdftSeq - 1) / kVdftSeq), dim3(kVdftThreads), lds, st,
                     vol, Wk);
  GIS_LAUNCH_CHECK("gis_volume_dft x");
  hipLaunchKernelGGL((vdft_y_pass<N, L>), dim3((G + 1) / 2, N), dim3(kVdftThreads), lds, st, Wk);
  GIS_LAUNCH_CHECK("gis_volume_dft y");
  hipLaunchKernelGGL((vdft_z_pass<N, L>), dim3(G, L / 2), dim3(kVdftThreads), lds, st, (const float2*)Wk, D);
  GIS_LAUNCH_CHECK("gis_volume_dft z");
  return GIS_OK;
}

static bool vdft_box_supported(int N) {
  return N == 32 || N == 64 || N == 96 || N == 128 || N == 192 || N == 256;
}

static size_t vdft_len(int N, int pad) { return pad ? (size_t)2 * N : (size_t)N; }

// Bytes of the bricked spectrum and of the workspace at box N, padded (pad 1,
// L = 2N) or not (pad 0, L = N).
GIS_API size_t gis_volume_dft_bytes_ex(int N, int pad) {
  const size_t L = vdft_len(N, pad);
  return N > 0 ? L * L * (L / 8 + 1) * 4 * sizeof(float2) : 0;
}
GIS_API size_t gis_volume_dft_workspace_ex(int N, int pad) {
  const size_t L = vdft_len(N, pad);
  return N > 0 ? (size_t)N * (L / 2) * ((L / 2 + 8) / 8) * 16 * sizeof(float2) : 0;
}
GIS_API size_t gis_volume_dft_bytes(int N) { return gis_volume_dft_bytes_ex(N, 1); }
GIS_API size_t gis_volume_dft_workspace(int N) { return gis_volume_dft_workspace_ex(N, 1); }

// The pad / sinc^2 / fftshift / rfftn prefix of the central-slice projector
// (IntegratedProjector.build_templates v606:1068-1073; project3d_and_whiten_cuda.py
// :103-117, pad 0 = its --skip_padding).  `volume` is [N][N][N] float32
// ([z][y][x]); `dft` receives the [L][L][L/2+1] complex64 half spectrum in
// natural frequency order, stored in 2x2x4 bricks (gis_volume_dft_bytes_ex
// bytes); `ws` holds gis_volume_dft_workspace_ex bytes.
GIS_API int gis_volume_dft_ex(const float* volume, int N, int pad, void* dft, void* ws, size_t ws_bytes,
                              void* stream) {
  GIS_REQUIRE(volume != nullptr && dft != nullptr && ws != nullptr, "gis_volume_dft: null pointer");
  if (!vdft_box_supported(N)) {
    gis_set_error("gis_volume_dft: box %d not in {32, 64, 96, 128, 192, 256}", N);
    return GIS_EUNSUPPORTED;
  }
  GIS_REQUIRE(pad == 0 || pad == 1, "gis_volume_dft: pad %d not 0 or 1", pad);
  GIS_REQUIRE(ws_bytes >= gis_volume_dft_workspace_ex(N, pad), "gis_volume_dft: workspace %zu < %zu bytes",
              ws_bytes, gis_volume_dft_workspace_ex(N, pad));
  hipStream_t st = (hipStream_t)stream;
  float2* D = (float2*)dft;
  float2* W = (float2*)ws;
#define GIS_VDFT_CASE(NV)                                                                       \
  case NV:                                                                                      \
    return pad ? launch_volume_dft<NV, 2 * NV>(volume, D, W, st) : launch_volume_dft<NV, NV>(volume, D, W, st);
  switch (N) {
    GIS_VDFT_CASE(32)
    GIS_VDFT_CASE(64)
    GIS_VDFT_CASE(96)
    GIS_VDFT_CASE(128)
    GIS_VDFT_CASE(192)
    default: return pad ? launch_volume_dft<256, 512>(volume, D, W, st) : launch_volume_dft<256, 256>(volume, D, W, st);
  }
#undef GIS_VDFT_CASE
}

GIS_API int gis_volume_dft(const float* volume, int N, void* dft, void* ws, size_t ws_bytes, void* stream) {
  return gis_volume_dft_ex(volume, N, 1, dft, ws, ws_bytes, stream);
}
