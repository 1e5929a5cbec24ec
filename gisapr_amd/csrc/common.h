// Shared device helpers for the gfx950 candidate-scoring kernels:
// complex arithmetic, the LDS-resident mixed-radix Stockham FFT, wave
// reductions, and the C-ABI error plumbing.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include <algorithm>

#define GIS_API extern "C" __attribute__((visibility("default")))

// ---------------------------------------------------------------------------
// Error reporting across the C ABI: every entry point returns 0 on success,
// a positive hipError_t on a runtime failure, or a negative code for invalid
// arguments; gis_last_error() returns the message of the last failure on the
// calling thread.
// ---------------------------------------------------------------------------
enum {
  GIS_OK = 0,
  GIS_EINVAL = -1,      // invalid argument (shape, size, pointer)
  GIS_EUNSUPPORTED = -2 // box size / crop size outside the compiled set
};

// 16-bit operand type of the split small-shift search.
enum { GIS_OPERAND_BF16 = 0, GIS_OPERAND_F16 = 1 };

void gis_set_error(const char* fmt, ...);

#define GIS_REQUIRE(cond, ...)              \
  do {                                      \
    if (!(cond)) {                          \
      gis_set_error(__VA_ARGS__);           \
      return GIS_EINVAL;                    \
    }                                       \
  } while (0)

#define GIS_LAUNCH_CHECK(what)                                              \
  do {                                                                      \
    hipError_t e_ = hipGetLastError();                                      \
    if (e_ != hipSuccess) {                                                 \
      gis_set_error("%s: launch failed: %s", what, hipGetErrorString(e_));  \
      return (int)e_;                                                       \
    }                                                                       \
  } while (0)

// ---------------------------------------------------------------------------
// Complex helpers (float2 = re, im)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// ---------------------------------------------------------------------------
// Twiddle table: tw[k] = exp(-2 pi i k / L), k in [0, L), computed in double.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void fill_twiddles(float2* tw, int L, int tid, int nthr) {
  for (int k = tid; k < L; k += nthr) {
    double s, c;
    sincospi(-2.0 * (double)k / (double)L, &s, &c);
    tw[k] = make_float2((float)c, (float)s);
  }
}

// ---------------------------------------------------------------------------
// LDS Stockham FFT, L = 2^a * 3^b (b <= 1): radix-8 stages, then one radix-4
// or radix-2 stage for the rest of 2^a, one radix-3 stage if 3 | L.
// Natural-order input and output.
//
// nseq sequences live in LDS at buf[q * sp + i]; the result ends in one of
// the two ping-pong buffers, whose address is returned.  The caller must
// __syncthreads() before the call (data ready); the function ends with a
// barrier.  Forward: exp(-2 pi i / L); inverse: exp(+...), unscaled.
// ---------------------------------------------------------------------------
// In-register R-point DFT (R in {2, 3, 4, 8}); forward e^{-2 pi i/R}, inverse e^{+...}.
template <int R, bool INV>
__device__ __forceinline__ void dft_radix(float2 (&v)[R]) {
  if constexpr (R == 2) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 8) {
    // two 4-point DFTs (even / odd inputs) and the twiddles w^m, w = e^{-+2 pi i / 8}
    float2 a = cadd(v[0], v[4]), b = csub(v[0], v[4]);
    float2 c = cadd(v[2], v[6]), d = mul_mi<INV>(csub(v[2], v[6]));
    const float2 A0 = cadd(a, c), A2 = csub(a, c), A1 = cadd(b, d), A3 = csub(b, d);
    a = cadd(v[1], v[5]); b = csub(v[1], v[5]);
    c = cadd(v[3], v[7]); d = mul_mi<INV>(csub(v[3], v[7]));
    const float2 B0 = cadd(a, c), B2 = csub(a, c), B1 = cadd(b, d), B3 = csub(b, d);
    constexpr float rh = 0.70710678118654752440f;
    const float2 t1 = INV ? make_float2((B1.x - B1.y) * rh, (B1.x + B1.y) * rh)
                          : make_float2((B1.x + B1.y) * rh, (B1.y - B1.x) * rh);
    const float2 t2 = mul_mi<INV>(B2);
    const float2 t3 = INV ? make_float2((-B3.x - B3.y) * rh, (B3.x - B3.y) * rh)
                          : make_float2((B3.y - B3.x) * rh, (-B3.x - B3.y) * rh);
    v[0] = cadd(A0, B0);
    v[4] = csub(A0, B0);
    v[1] = cadd(A1, t1);
    v[5] = csub(A1, t1);
    v[2] = cadd(A2, t2);
    v[6] = csub(A2, t2);
    v[3] = cadd(A3, t3);
    v[7] = csub(A3, t3);
  } else if constexpr (R == 4) {
    float2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    float2 c = cadd(v[1], v[3]), d = mul_mi<INV>(csub(v[1], v[3]));
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
  } else if constexpr (R == 3) {
    const float c1 = -0.5f;
    const float s1 = INV ? 0.86602540378443864676f : -0.86602540378443864676f;
    float2 a = v[0], b = v[1], c = v[2];
    float2 t1 = cadd(b, c);
    float2 t2 = make_float2(a.x + c1 * t1.x, a.y + c1 * t1.y);
    float2 t3 = csub(b, c);  // (b - c) * i*s1
    float2 t4 = make_float2(-s1 * t3.y, s1 * t3.x);
    v[0] = cadd(a, t1);
    v[1] = cadd(t2, t4);
    v[2] = csub(t2, t4);
  }
}

// One Stockham radix-R stage: butterfly j (of L/R) reads src[j + r*L/R],
// applies the twiddles w^(r*k), k = j mod Ns, and writes
// dst[(j/Ns)*Ns*R + k + r*Ns].
template <int L, int R, bool INV>
__device__ __forceinline__ void stockham_stage(const float2* __restrict__ src, float2* __restrict__ dst,
                                               int Ns, int nseq, int sp, const float2* __restrict__ tw,
                                               int tid, int nthr) {
  constexpr int LR = L / R;
  const int total = nseq * LR;
  const int step = L / (Ns * R);
  for (int t = tid; t < total; t += nthr) {
    const int q = t / LR;
    const int j = t - q * LR;
    const float2* S = src + q * sp;
    float2* D = dst + q * sp;
    const int k = j % Ns;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = S[j + r * LR];
    if (Ns > 1)  // the first stage's twiddles are all 1
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = tw[k * r * step];
        if (INV) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    dft_radix<R, INV>(v);
    const int base = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) D[base + r * Ns] = v[r];
  }
}

template <int L>
struct FftShape {
  static constexpr int three = (L % 3 == 0) ? 1 : 0;
  static constexpr int p2 = three ? L / 3 : L;
  static constexpr int lg = (p2 >= 512) ? 9 : (p2 >= 256) ? 8 : (p2 >= 128) ? 7 : (p2 >= 64) ? 6
                          : (p2 >= 32) ? 5 : (p2 >= 16) ? 4 : (p2 >= 8) ? 3 : (p2 >= 4) ? 2 : 1;
  static_assert((1 << lg) == p2, "FFT length must be 2^a or 3*2^a");
  static constexpr int n8 = lg / 3;                  // radix-8 stages
  static constexpr int n4 = (lg % 3 == 2) ? 1 : 0;   // then one radix-4 or one radix-2 stage
  static constexpr int n2 = (lg % 3 == 1) ? 1 : 0;
  static constexpr int stages = n8 + n4 + n2 + three;
};

template <int L, bool INV>
__device__ float2* fft_lds(float2* a, float2* b, int nseq, int sp, const float2* tw, int tid, int nthr) {
  using S = FftShape<L>;
  float2* src = a;
  float2* dst = b;
  int Ns = 1;
#pragma unroll
  for (int s = 0; s < S::n8; ++s) {
    stockham_stage<L, 8, INV>(src, dst, Ns, nseq, sp, tw, tid, nthr);
    __syncthreads();
    float2* t = src; src = dst; dst = t;
    Ns *= 8;
  }
  if constexpr (S::n4) {
    stockham_stage<L, 4, INV>(src, dst, Ns, nseq, sp, tw, tid, nthr);
    __syncthreads();
    float2* t = src; src = dst; dst = t;
    Ns *= 4;
  }
  if constexpr (S::n2) {
    stockham_stage<L, 2, INV>(src, dst, Ns, nseq, sp, tw, tid, nthr);
    __syncthreads();
    float2* t = src; src = dst; dst = t;
    Ns *= 2;
  }
  if constexpr (S::three) {
    stockham_stage<L, 3, INV>(src, dst, Ns, nseq, sp, tw, tid, nthr);
    __syncthreads();
    float2* t = src; src = dst; dst = t;
    Ns *= 3;
  }
  return src;
}

// Orders LDS accesses of the lanes of one wave (a wave executes its LDS
// operations in order; this keeps the compiler from moving them across).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Same transform as fft_lds, but each wave owns whole sequences (q = wave,
// wave + nwaves, ...) and runs all their stages with wave-level ordering
// only: no workgroup barrier between stages.  Ends with one __syncthreads.
template <int L, bool INV>
__device__ float2* fft_lds_waves(float2* a, float2* b, int nseq, int sp, const float2* tw, int tid, int nthr) {
  using S = FftShape<L>;
  const int wave = tid >> 6, lane = tid & 63, nw = nthr >> 6;
  constexpr int parity = S::stages & 1;
  for (int q = wave; q < nseq; q += nw) {
    float2* src = a + q * sp;
    float2* dst = b + q * sp;
    int Ns = 1;
#pragma unroll
    for (int s = 0; s < S::n8; ++s) {
      stockham_stage<L, 8, INV>(src, dst, Ns, 1, 0, tw, lane, 64);
      wave_lds_sync();
      float2* t = src; src = dst; dst = t;
      Ns *= 8;
    }
    if constexpr (S::n4) {
      stockham_stage<L, 4, INV>(src, dst, Ns, 1, 0, tw, lane, 64);
      wave_lds_sync();
      float2* t = src; src = dst; dst = t;
      Ns *= 4;
    }
    if constexpr (S::n2) {
      stockham_stage<L, 2, INV>(src, dst, Ns, 1, 0, tw, lane, 64);
      wave_lds_sync();
      float2* t = src; src = dst; dst = t;
      Ns *= 2;
    }
    if constexpr (S::three) {
      stockham_stage<L, 3, INV>(src, dst, Ns, 1, 0, tw, lane, 64);
      wave_lds_sync();
      float2* t = src; src = dst; dst = t;
      Ns *= 3;
    }
  }
  __syncthreads();
  return parity ? b : a;
}

// In-place variant for one sequence owned by one wave: every stage first
// loads all of its butterflies' inputs into registers (a lane holds
// ceil(L/R/64) butterflies), orders the wave's LDS accesses, then writes the
// outputs over the same buffer.  Same arithmetic, in the same order, as
// fft_lds / fft_lds_waves with half the LDS.  Element i of the sequence
// lives at s[lds_pad<PAD>(i)]: PAD > 0 inserts one unused slot after every
// 2^PAD elements, which spreads the strided stores of the early stages over
// the LDS banks.  The caller orders the wave's (or the block's) accesses
// before and after.
template <int PAD>
__device__ __forceinline__ int lds_pad(int i) { return PAD > 0 ? i + (i >> PAD) : i; }
template <int L, int PAD>
constexpr int lds_padded_len() { return PAD > 0 ? L + (L >> PAD) : L; }

template <int L, int R, bool INV, int PAD>
__device__ __forceinline__ void stockham_stage_inplace(float2* s, int Ns, const float2* __restrict__ tw, int lane) {
  constexpr int LR = L / R;
  constexpr int NB = (LR + 63) / 64;
  const int step = L / (Ns * R);
  float2 v[NB][R];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = lane + 64 * b;
    if (LR % 64 == 0 || j < LR)
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = s[lds_pad<PAD>(j + r * LR)];
  }
  wave_lds_sync();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = lane + 64 * b;
    if (LR % 64 == 0 || j < LR) {
      const int k = j % Ns;
      if (Ns > 1)  // the first stage's twiddles are all 1
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[k * r * step];
          if (INV) w.y = -w.y;
          v[b][r] = cmul(v[b][r], w);
        }
      dft_radix<R, INV>(v[b]);
      const int base = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) s[lds_pad<PAD>(base + r * Ns)] = v[b][r];
    }
  }
  wave_lds_sync();
}

template <int L, bool INV, int PAD = 0>
__device__ __forceinline__ void fft_wave_inplace(float2* s, const float2* __restrict__ tw, int lane) {
  using S = FftShape<L>;
  int Ns = 1;
#pragma unroll
  for (int st = 0; st < S::n8; ++st) {
    stockham_stage_inplace<L, 8, INV, PAD>(s, Ns, tw, lane);
    Ns *= 8;
  }
  if constexpr (S::n4) {
    stockham_stage_inplace<L, 4, INV, PAD>(s, Ns, tw, lane);
    Ns *= 4;
  }
  if constexpr (S::n2) {
    stockham_stage_inplace<L, 2, INV, PAD>(s, Ns, tw, lane);
    Ns *= 2;
  }
  if constexpr (S::three) stockham_stage_inplace<L, 3, INV, PAD>(s, Ns, tw, lane);
}

// nseq sequences at buf[q * sp + lds_pad<PAD>(i)], each transformed in place
// by one wave (q = wave, wave + nwaves, ...).  The caller must
// __syncthreads() before the call; the function ends with a barrier.
template <int L, bool INV, int PAD>
__device__ __forceinline__ void fft_lds_inplace(float2* buf, int nseq, int sp, const float2* __restrict__ tw,
                                                int tid, int nthr) {
  const int wave = tid >> 6, lane = tid & 63, nw = nthr >> 6;
  for (int q = wave; q < nseq; q += nw) fft_wave_inplace<L, INV, PAD>(buf + q * sp, tw, lane);
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Block-wide reductions (blockDim.x a multiple of 64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sums `nv` doubles per thread over the block; result valid in all threads.
// `scratch` must hold nv * (blockDim.x / 64) doubles of LDS.
template <int NV>
__device__ __forceinline__ void block_sum_f64(double (&v)[NV], double* scratch) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum_f64(v[i]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) scratch[i * nw + wid] = v[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += scratch[i * nw + w];
    v[i] = s;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Per-particle tail of a search (the numerics of make_result_records,
// v606:1466-1491): one wave per particle over its [M_b][P] pair scores:
// first maximum in (local direction, psi) order, mean and population std in
// float64.  Internal linkage: each search source gets its own copy.
// ---------------------------------------------------------------------------
// (bv, bi) := the first maximum of (bv, bi) and (v, i): the larger value,
// the smaller index on a tie; selects rather than branches.
__device__ __forceinline__ void first_max(float& bv, int& bi, float v, int i) {
  const bool take = (v > bv) | ((v == bv) & (i < bi));
  bv = take ? v : bv;
  bi = take ? i : bi;
}

namespace {
__global__ void __launch_bounds__(64)
particle_tail_kernel(const int* __restrict__ dir_off, int P, const float* __restrict__ scores,
                     int* __restrict__ best, float* __restrict__ peak, double* __restrict__ mean,
                     double* __restrict__ stdv) {
  const int b = blockIdx.x, lane = threadIdx.x;
  const int d0 = dir_off[b];
  const int np = (dir_off[b + 1] - d0) * P;
  const float* s = scores + (size_t)d0 * P;
  float bv = -INFINITY;
  int bi = 0x7fffffff;
  double sum = 0.0;
  for (int k = lane; k < np; k += 64) {
    const float v = s[k];
    sum += (double)v;
    if (v > bv || (v == bv && k < bi)) { bv = v; bi = k; }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const float ov = __shfl_xor(bv, o, 64);
    const int oi = __shfl_xor(bi, o, 64);
    first_max(bv, bi, ov, oi);
    sum += __shfl_xor(sum, o, 64);
  }
  const double mu = np > 0 ? sum / (double)np : 0.0;
  double ss = 0.0;
  for (int k = lane; k < np; k += 64) {
    const double dv = (double)s[k] - mu;
    ss += dv * dv;
  }
  ss = wave_sum_f64(ss);
  if (lane == 0) {
    best[b] = np > 0 ? bi : -1;
    peak[b] = np > 0 ? bv : 0.0f;
    mean[b] = mu;
    stdv[b] = np > 0 ? sqrt(ss / (double)np) : 0.0;
  }
}
}  // namespace

// Number of compute units of the current device (cached per device).
// Fragment order of a packed plane (n rows, m = n/2+1 columns), the search
// operand layout read with 16-byte loads: tile (ks, ct) of rows
// [16 ks, 16 ks + 16) x columns [16 ct, 16 ct + 16) is 256 complex values at
// (ks * nct + ct) * 256, value (row 16 ks + 4g + 2h + e, column 16 ct + i) at
// ((h * 4 + g) * 16 + i) * 2 + e; when the last column is alone in its tile
// (m % 16 == 1) nct counts the tiles before it and its values follow the
// tiles, row 16 ks + 4g + j at nks * nct * 256 + 16 ks + 4g + j.  Rows >= n
// and columns >= m hold zeros.  nks = the searches' even-padded row steps.
__host__ __device__ inline int frag_nks(int n) { return (((n + 15) / 16) + 1) & ~1; }
__host__ __device__ inline int frag_nct(int n) {
  const int m = n / 2 + 1;
  return (m & 15) == 1 ? (m >> 4) : (m + 15) >> 4;
}
__host__ __device__ inline int frag_plane(int n) {
  const int m = n / 2 + 1;
  return frag_nks(n) * frag_nct(n) * 256 + ((m & 15) == 1 ? frag_nks(n) * 16 : 0);
}

// Position of (row r, column c) in a fragment-order plane (nct = frag_nct(n)).
__host__ __device__ inline int frag_pos(int n, int nct, int r, int c) {
  const int m = n / 2 + 1;
  if ((m & 15) == 1 && c == m - 1) return frag_nks(n) * nct * 256 + r;
  return ((r >> 4) * nct + (c >> 4)) * 256 + ((((r >> 1) & 1) * 4 + ((r >> 2) & 3)) * 16 + (c & 15)) * 2 + (r & 1);
}

// (row, column) of position e of a fragment-order plane (tiles = frag_nks(n)
// * nct * 256); a row >= n or column >= n/2+1 marks a zero pad.
__host__ __device__ inline void frag_rc(int n, int nct, int tiles, int e, int& r, int& c) {
  if (e < tiles) {
    const int tl = e >> 8, w = e & 255;
    const int ks = tl / nct, hg = w >> 5;
    r = 16 * ks + 4 * (hg & 3) + 2 * (hg >> 2) + (w & 1);
    c = 16 * (tl - ks * nct) + ((w >> 1) & 15);
  } else {
    r = e - tiles;  // the lone last column
    c = n / 2;
  }
}

int gis_cu_count();
// Looks up the allocation-free search workspace of the current device.
int gis_split_scratch_setup();
// The FP32 search on the split search's work decomposition (split.hip);
// ws holds gis_correlate_split_workspace(n, C, B) bytes.
int gis_correlate_units(const void* bank, const void* spectra, const float* inv_norm, const float* tinv,
                        const int* dir_off, const int* dir_idx, const int* pmap, const int* order, int B, int P, int n,
                        int C, int nrows, int ldt, int ldp, int dt, int pt, float* scores, int* peak_y, int* peak_x,
                        int* best, float* peak, double* mean, double* stdv, void* ws, size_t ws_bytes,
                        hipStream_t st);
GIS_API size_t gis_correlate_split_workspace(int n, int C, int B);
