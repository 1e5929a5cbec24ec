// Full-plane correlation search on split-BF16 MFMA.
//
// The v606 score of (particle, direction, psi) at crop lag (y, x) is the
// fixed linear functional of Q = T * P (correlate_templates_and_particle,
// v606:1418-1449, for UNshifted operands):
//     ccg[y][x] = 1/n^2 sum_{r,c} Re( Q[r][c] W[r][y] E[c][x] )
//     W[r][y] = (-1)^l_y exp(+2 pi i r l_y / n)                      (row roll)
//     E[c][x] = w(s_c) exp(+2 pi i s_c l_x / n),  s_c = (c + (n/2+1)//2) mod (n/2+1)
//     w(s) = 1 for s in {0, n/2}, else 2                             (C2R rule)
// over every packed coefficient and every lag of the C x C crop: the same
// quantity gis_correlate computes on FP32 MFMA.  Here it runs on
// v_mfma_f32_16x16x32_bf16 as two chained products per 16-column tile of the
// packed plane:
//     stage 1 (rows):     G[c][y]  = sum_r Q[r][c] W[r][y]        M = 16 columns, K = 16 rows (re, im) per
//                                                                  step, N = 16 lags, real and imaginary part
//     stage 2 (columns):  S[x][y] += sum_c Re(E[c][x] G[c][y])    M = 16 lags, K = 16 columns (re, im), N = 16 lags
// Every operand is split into BF16 hi + lo and each product takes three
// MFMAs (hi*hi + hi*lo + lo*hi, FP32 accumulation): ~16 significant bits per
// product.  The tables are split once per launch (the row table, read every
// step, is held in LDS); Q is formed from the FP32 spectra and split in
// registers per step; the stage-1 accumulators, re-split, are the stage-2 B
// operand register for register (their column index is the K index, their
// lag index the N index), so G never leaves the wave.
//
// Lane layout (16x16x32 fragments: A lane l holds A[l%16][8(l/16) .. +7],
// B lane l holds B[8(l/16) .. +7][l%16], C lane l holds C[4(l/16) .. +3][l%16]),
// g = l / 16, i = l % 16:
//   stage-1 A: column 16 ct + i, slot 2j + ri <-> row 16 ks + 4g + j (j = 0..3):
//              one load instruction j reads 4 rows x 16 consecutive columns
//   stage-1 B (table W): lag i, the same row slots; tile "re" holds
//              (Re W, -Im W), tile "im" holds (Im W, Re W)
//   stage-2 A (table E): lag i, slot 2v + ri <-> column 16 ct + 4g + v
//   stage-2 B: G_re / G_im accumulator element v of lane (g, i) = column 16 ct + 4g + v
// A crop C <= 16 uses one lag tile per axis (LT = 1), C <= 32 two (LT = 2).
//
// Work decomposition: one wave per (particle, DT local directions, PT psi
// values); a particle's waves are consecutive, four per block, and blocks are
// mapped so consecutive ones share an XCD (and its L2, where a particle's psi
// spectra are reused by its directions).  Each wave keeps two operand buffers
// in registers and loads a step's operands two steps ahead.  The pair tail
// (first maximum over the crop) runs in the wave's epilogue; the particle
// tail is particle_tail_kernel.
#include "common.h"

#include <mutex>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kWaves = 4;
constexpr int kXcds = 8;

__device__ __forceinline__ floatx4 mfma_bf16(bf16x8 a, bf16x8 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ floatx4 mfma_f32(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// x = hi + lo, hi = bf16(x), lo = bf16(x - hi), round to nearest even; two
// values per packed conversion.
__device__ __forceinline__ void split8(const float (&x)[8], bf16x8& hi, bf16x8& lo) {
#pragma unroll
  for (int e = 0; e < 8; e += 2) {
    const f32x2 v = {x[e], x[e + 1]};
    const bf16x2 h = __builtin_convertvector(v, bf16x2);
    const f32x2 hf = __builtin_convertvector(h, f32x2);
    const bf16x2 l = __builtin_convertvector(v - hf, bf16x2);
    hi[e] = h[0];
    hi[e + 1] = h[1];
    lo[e] = l[0];
    lo[e + 1] = l[1];
  }
}

// a * q in float32 with a fixed operation order (one product rounded, then a
// fused multiply-add), whatever the tile shape: results are bit-identical
// across the kernel's variants
__device__ __forceinline__ void cprod(float2 a, float2 q, float& re, float& im) {
  re = __builtin_fmaf(a.x, q.x, -(a.y * q.y));
  im = __builtin_fmaf(a.x, q.y, a.y * q.x);
}

__host__ __device__ inline int crop_lag(int i, int n, int C) { return i < C / 2 ? n - C / 2 + i : i - C / 2; }

// One table value split from float64: hi = bf16(v), lo = bf16(v - hi).
__device__ __forceinline__ void split_d(double v, __bf16& hi, __bf16& lo) {
  hi = (__bf16)(float)v;
  lo = (__bf16)(float)(v - (double)(float)hi);
}

// dst [rows][frag_plane(n)] from src [rows][n][ld] (row major, ld >= n/2+1).
__global__ void pack_fragments_kernel(const float2* __restrict__ src, int rows, int n, int ld,
                                      float2* __restrict__ dst) {
  const int m = n / 2 + 1, nct = frag_nct(n), F = frag_plane(n);
  const int tiles = frag_nks(n) * nct * 256;
  const size_t total = (size_t)rows * F;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const size_t row = t / F;
    int r, c;
    frag_rc(n, nct, tiles, (int)(t - row * F), r, c);
    dst[t] = (r < n && c < m) ? src[row * (size_t)n * ld + (size_t)r * ld + c] : make_float2(0.f, 0.f);
  }
}

// Split tables in fragment order (hi at o, lo at o + 64).
//   Wt [nks][2 LT tiles][hi, lo][64 lanes] x 8: tile 2 ty + part, lane (g, i) -> lag 16 ty + i,
//      slot 2j + ri -> row 16 ks + 4g + j; part 0: (Re W, -Im W), part 1: (Im W, Re W)
//   Et [nct][LT][hi, lo][64 lanes] x 8: lane (g, i) -> lag 16 tx + i, slot 2v + ri -> column 16 ct + 4g + v,
//      (Re E, -Im E)
// Lags beyond C and rows / columns beyond the plane are zero.
//   mode kTabF32: the same eight values per lane and fragment, rounded to
//   float and stored [fragment][64 lanes][8] (one 32-byte run per lane).
//   mode kTabF16: Wt split into fp16 hi + lo (same layout), Et as bf16.
enum { kTabBF16 = 0, kTabF32 = 1, kTabF16 = 2 };
__global__ void split_tables_kernel(int n, int C, int LT, int nks, int nct, int mode, void* __restrict__ Wt,
                                    void* __restrict__ Et) {
  const int f32 = mode == kTabF32;
  const int m = n / 2 + 1, s = m / 2;
  const int nw = nks * 2 * LT * 64, ne = nct * LT * 64;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nw + ne; t += gridDim.x * blockDim.x) {
    const int lane = t & 63, g = lane >> 4, i = lane & 15;
    double v[8];
    size_t o;
    if (t < nw) {
      const int tile = (t >> 6) % (2 * LT), ks = (t >> 6) / (2 * LT);
      const int y = (tile >> 1) * 16 + i, part = tile & 1;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = 16 * ks + 4 * g + j;
        double wr = 0.0, wi = 0.0;
        if (r < n && y < C) {
          const int l = crop_lag(y, n, C);
          const long k = ((long)r * l) % n;
          double sn, cs;
          sincospi(2.0 * (double)k / (double)n, &sn, &cs);
          const double sg = (l & 1) ? -1.0 : 1.0;
          wr = sg * cs;
          wi = sg * sn;
        }
        v[2 * j] = part == 0 ? wr : wi;
        v[2 * j + 1] = part == 0 ? -wi : wr;
      }
      o = ((size_t)(ks * 2 * LT + tile) * 2) * 64 + lane;
    } else {
      const int u = t - nw;
      const int tx = (u >> 6) % LT, ct = (u >> 6) / LT;
      const int x = tx * 16 + i;
#pragma unroll
      for (int vv = 0; vv < 4; ++vv) {
        const int c = 16 * ct + 4 * g + vv;
        double er = 0.0, ei = 0.0;
        if (c < m && x < C) {
          const int l = crop_lag(x, n, C);
          const int pos = (c + s) % m;
          const double w = (pos == 0 || pos == n / 2) ? 1.0 : 2.0;
          const long k = ((long)pos * l) % n;
          double sn, cs;
          sincospi(2.0 * (double)k / (double)n, &sn, &cs);
          er = w * cs;
          ei = w * sn;
        }
        v[2 * vv] = er;
        v[2 * vv + 1] = -ei;
      }
      o = ((size_t)(ct * LT + tx) * 2) * 64 + lane;
    }
    if (f32) {
      float* dst = (float*)(t < nw ? Wt : Et);
      const size_t blk = (o - lane) / 128;  // fragment index
#pragma unroll
      for (int e = 0; e < 8; ++e) dst[(blk * 64 + lane) * 8 + e] = (float)v[e];
      continue;
    }
    if (mode == kTabF16 && t < nw) {
      f16x8* dst = (f16x8*)Wt;
      f16x8 hi, lo;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        hi[e] = (_Float16)(float)v[e];
        lo[e] = (_Float16)(float)(v[e] - (double)(float)hi[e]);
      }
      dst[o] = hi;
      dst[o + 64] = lo;
      continue;
    }
    bf16x8* dst = (bf16x8*)(t < nw ? Wt : Et);
    bf16x8 hi, lo;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      __bf16 h, l;
      split_d(v[e], h, l);
      hi[e] = h;
      lo[e] = l;
    }
    dst[o] = hi;
    dst[o + 64] = lo;
  }
}

// Exclusive scans over the particles in processing order (order[k], or k
// when order is NULL), in two launches of ceil(B / 1024) blocks of 1024
// threads, one entry per thread.  Values: kScanUnits, the search units of
// particle b, ceil((src[b+1] - src[b]) / DT) for src = dir_off, written at
// the position k (unit_off[k]); kScanCounts, src[b] written at b (off[b]).
// out[B] = the total.  WRITE = false stores each tile's total in
// tile_sum[tile]; WRITE = true adds the totals of the earlier tiles to the
// tile's own scan.  B = 0 takes one block and writes out[0] = 0.
enum ScanValue { kScanUnits = 0, kScanCounts = 1 };

template <int MODE, bool WRITE>
__global__ void __launch_bounds__(1024)
order_scan_kernel(const int* __restrict__ src, const int* __restrict__ order, int B, int DT,
                  int* __restrict__ tile_sum, int* __restrict__ out) {
  __shared__ int wave_sum[16];
  __shared__ int base_s;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int k = (int)blockIdx.x * 1024 + t;
  int b = k, v = 0;
  if (k < B) {
    b = order ? order[k] : k;
    v = MODE == kScanUnits ? (src[b + 1] - src[b] + DT - 1) / DT : src[b];
  }
  int x = v;  // inclusive scan within the wave
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) wave_sum[w] = x;
  if (WRITE && w == 0) {  // the totals of the earlier tiles
    int s = 0;
    for (int i = lane; i < (int)blockIdx.x; i += 64) s += tile_sum[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) base_s = s;
  }
  __syncthreads();
  int before = 0;  // the earlier waves of this tile
  for (int i = 0; i < w; ++i) before += wave_sum[i];
  if (!WRITE) {
    if (t == 1023) tile_sum[blockIdx.x] = before + x;
    return;
  }
  const int excl = base_s + before + x - v;
  if (k < B) out[MODE == kScanUnits ? k : b] = excl;
  if (k == B - 1) out[B] = excl + v;
  if (B == 0 && t == 0) out[0] = 0;
}

// The two launches of order_scan_kernel; tile_sum holds scan_tiles(B) ints.
int scan_tiles(int B) { return std::max(1, (B + 1023) / 1024); }

template <int MODE>
void launch_order_scan(const int* src, const int* order, int B, int DT, int* tile_sum, int* out, hipStream_t st) {
  const dim3 grid((unsigned)scan_tiles(B));
  hipLaunchKernelGGL((order_scan_kernel<MODE, false>), grid, dim3(1024), 0, st, src, order, B, DT, tile_sum, out);
  hipLaunchKernelGGL((order_scan_kernel<MODE, true>), grid, dim3(1024), 0, st, src, order, B, DT, tile_sum, out);
}

// The step operands of one lane: 4 complex values (rows 16 ks + 4g + j of
// column 16 ct + i) of each of the wave's DT template rows and PT psi spectra.
template <int DT, int PT>
struct Frags {
  float2 t[DT][4];
  float2 p[PT][4];
};

template <int LT, int DT, int PT, bool F32, bool FR, bool LIST = false, int PF = 2>
__global__ void __launch_bounds__(64 * kWaves)
split_correlate_kernel(const float2* __restrict__ bank, const float2* __restrict__ spectra,
                       const float* __restrict__ inv_norm, const float* __restrict__ tinv,
                       const int* __restrict__ dir_off, const int* __restrict__ dir_idx,
                       const int* __restrict__ pmap, const int* __restrict__ order,
                       const int* __restrict__ unit_off, int B, int P, int n, int C, int ldt, int ldp, int nblk,
                       const void* __restrict__ Wtv, const void* __restrict__ Etv, float* __restrict__ scores,
                       int* __restrict__ peak_y, int* __restrict__ peak_x, const int2* __restrict__ list = nullptr,
                       const int* __restrict__ list_count = nullptr) {
  // the row table (read every step) lives in LDS; the column table (read
  // once per column tile) is read from global memory (cached)
  extern __shared__ bf16x8 tab[];
  const bf16x8* Wt = (const bf16x8*)Wtv;
  const bf16x8* Et = (const bf16x8*)Etv;
  const float4* Wf = (const float4*)tab;   // FP32 tables: [fragment][lane][2] float4
  const float4* Ef = (const float4*)Etv;
  const int m = n / 2 + 1;
  // when the packed plane's last column is alone in its tile (n a multiple
  // of 32), it is searched separately: the wave's pairs share one tile
  const bool lone = (m & 15) == 1;
  const int nct = lone ? (m >> 4) : (m + 15) >> 4;
  const int nks = (((n + 15) >> 4) + 1) & ~1;  // row steps, padded to even (zero table rows)
  static_assert(!(LIST && F32), "the pair list mode is split-bf16");
  const bf16x8* Wl = tab;
  {
    const int nw = nks * 4 * LT * 64;
    for (int e = threadIdx.x; e < nw; e += 64 * kWaves) tab[e] = Wt[e];
    __syncthreads();
  }

  // consecutive logical blocks on one XCD (blocks are dealt to the XCDs round robin)
  const int per = (int)(gridDim.x / kXcds);
  const int lb = (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
  const int lane = threadIdx.x & 63;
  // LIST (pair list mode, DT = PT = 1): each wave scores the listed pairs
  // (particle b, pair k = local direction * P + psi) wave, wave + waves, ...
  // of the list_count[0] entries; otherwise one unit per wave
  const int wave = lb * kWaves + (int)(threadIdx.x >> 6);
  const int nlist = LIST ? list_count[0] : 0;
  for (int it = wave;; it += (int)gridDim.x * kWaves) {
    int b, j0, p0;
    if constexpr (LIST) {
      if (it >= nlist) break;  // no barrier below
      const int2 e = list[__builtin_amdgcn_readfirstlane(it)];
      b = __builtin_amdgcn_readfirstlane(e.x);
      const int k = __builtin_amdgcn_readfirstlane(e.y);
      j0 = k / P;
      p0 = k - j0 * P;
    } else {
      // wave-uniform from here on: unit, particle and row pointers live in SGPRs
      const int u = __builtin_amdgcn_readfirstlane(it);
      const int total = unit_off[B];
      if (lb >= nblk || u >= total) return;  // no barrier below
      int lo = 0, hi = B;  // unit_off[lo] <= u < unit_off[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (unit_off[mid] <= u) lo = mid; else hi = mid;
      }
      b = order ? order[lo] : lo;
      j0 = (u - unit_off[lo]) * DT;
      p0 = (int)blockIdx.y * PT;
    }
    const int d0 = dir_off[b], Mi = dir_off[b + 1] - d0;
    const int pb = pmap ? pmap[b] : b;
    // template rows are ldt complex values apart, psi spectrum rows ldp
    // (ld > n/2+1: zero padding, so 16-column segments start on 128-byte
    // lines); ld = 0: the operand is in fragment order (frag_plane), each
    // step's tile two contiguous 1 KiB runs read with 16-byte loads
    const size_t fplane = (size_t)frag_plane(n);
    const float2* tp[DT];
    const float2* pp[PT];
#pragma unroll
    for (int d = 0; d < DT; ++d)
      tp[d] = bank + (size_t)dir_idx[d0 + min(j0 + d, Mi - 1)] * (FR ? fplane : (size_t)n * ldt);
#pragma unroll
    for (int p = 0; p < PT; ++p)
      pp[p] = spectra + ((size_t)pb * P + min(p0 + p, P - 1)) * (FR ? fplane : (size_t)n * ldp);
    const int g = lane >> 4, i = lane & 15;

    // Operands of step (ct, ks).  Loads are unconditional: an offset past the
    // plane (rows >= n) is clamped to its last element, columns >= n/2+1 read
    // the next row or the zero padding, and the tables hold zeros at those rows
    // and columns, so the (finite) values read there contribute exactly
    // nothing.
    const int lastt = n * ldt - 1, lastp = n * ldp - 1;
    auto load = [&](int ct, int ks, Frags<DT, PT>& f) {
      const int r0 = 16 * ks + 4 * g, c = 16 * ct + i;
      const int tile = (ks * nct + ct) * 256 + 2 * lane;  // fragment order: [h][lane][2 rows]
      if constexpr (FR) {
#pragma unroll
        for (int d = 0; d < DT; ++d) {
          const float4 q0 = *(const float4*)(tp[d] + tile), q1 = *(const float4*)(tp[d] + tile + 128);
          f.t[d][0] = make_float2(q0.x, q0.y);
          f.t[d][1] = make_float2(q0.z, q0.w);
          f.t[d][2] = make_float2(q1.x, q1.y);
          f.t[d][3] = make_float2(q1.z, q1.w);
        }
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int ot = min((r0 + j) * ldt + c, lastt);
#pragma unroll
          for (int d = 0; d < DT; ++d) f.t[d][j] = tp[d][ot];
        }
      }
      if constexpr (FR) {
#pragma unroll
        for (int p = 0; p < PT; ++p) {
          const float4 q0 = *(const float4*)(pp[p] + tile), q1 = *(const float4*)(pp[p] + tile + 128);
          f.p[p][0] = make_float2(q0.x, q0.y);
          f.p[p][1] = make_float2(q0.z, q0.w);
          f.p[p][2] = make_float2(q1.x, q1.y);
          f.p[p][3] = make_float2(q1.z, q1.w);
        }
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int op = min((r0 + j) * ldp + c, lastp);
#pragma unroll
          for (int p = 0; p < PT; ++p) f.p[p][j] = pp[p][op];
        }
      }
    };

    floatx4 out[DT][PT][LT * LT], gacc[DT][PT][2 * LT];
#pragma unroll
    for (int d = 0; d < DT; ++d)
#pragma unroll
      for (int p = 0; p < PT; ++p)
#pragma unroll
        for (int t = 0; t < LT * LT; ++t) out[d][p][t] = floatx4{0.f, 0.f, 0.f, 0.f};

    // stage 1 of one step from buffer f, then f is refilled with step (ct2, ks2)
    auto step = [&](int ks, Frags<DT, PT>& f, int ct2, int ks2) {
      if constexpr (F32) {
        float w[2 * LT][8];
#pragma unroll
        for (int t = 0; t < 2 * LT; ++t) {
          const float4 w0 = Wf[((ks * 2 * LT + t) * 64 + lane) * 2], w1 = Wf[((ks * 2 * LT + t) * 64 + lane) * 2 + 1];
          w[t][0] = w0.x; w[t][1] = w0.y; w[t][2] = w0.z; w[t][3] = w0.w;
          w[t][4] = w1.x; w[t][5] = w1.y; w[t][6] = w1.z; w[t][7] = w1.w;
        }
#pragma unroll
        for (int d = 0; d < DT; ++d) {
#pragma unroll
          for (int p = 0; p < PT; ++p) {
            float x[8];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              cprod(f.t[d][j], f.p[p][j], x[2 * j], x[2 * j + 1]);
            }
#pragma unroll
            for (int t = 0; t < 2 * LT; ++t)
#pragma unroll
              for (int kk = 0; kk < 8; ++kk) gacc[d][p][t] = mfma_f32(x[kk], w[t][kk], gacc[d][p][t]);
          }
        }
      } else {
        bf16x8 bh[2 * LT], bl[2 * LT];
#pragma unroll
        for (int t = 0; t < 2 * LT; ++t) {
          bh[t] = Wl[((ks * 2 * LT + t) * 2) * 64 + lane];
          bl[t] = Wl[((ks * 2 * LT + t) * 2 + 1) * 64 + lane];
        }
#pragma unroll
        for (int d = 0; d < DT; ++d) {
#pragma unroll
          for (int p = 0; p < PT; ++p) {
            float x[8];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              cprod(f.t[d][j], f.p[p][j], x[2 * j], x[2 * j + 1]);
            }
            bf16x8 ahi, alo;
            split8(x, ahi, alo);
#pragma unroll
            for (int t = 0; t < 2 * LT; ++t) {
              gacc[d][p][t] = mfma_bf16(ahi, bh[t], gacc[d][p][t]);
              gacc[d][p][t] = mfma_bf16(ahi, bl[t], gacc[d][p][t]);
              gacc[d][p][t] = mfma_bf16(alo, bh[t], gacc[d][p][t]);
            }
          }
        }
      }
      load(ct2, ks2, f);
    };

    // PF buffers, each step's operands loaded PF steps ahead; nks is a
    // multiple of PF (even always; the host launches PF = 4, for the
    // latency-bound list pass, only where nks % 4 == 0), so buffer q always
    // holds the steps ks = q (mod PF), across column tiles too.
    static_assert(PF == 2 || PF == 4, "prefetch depth");
    Frags<DT, PT> fr[PF];
#pragma unroll
    for (int q = 0; q < PF; ++q) load(0, q, fr[q]);
    for (int ct = 0; ct < nct; ++ct) {
      const int ctn = min(ct + 1, nct - 1);
#pragma unroll
      for (int d = 0; d < DT; ++d)
#pragma unroll
        for (int p = 0; p < PT; ++p)
#pragma unroll
          for (int t = 0; t < 2 * LT; ++t) gacc[d][p][t] = floatx4{0.f, 0.f, 0.f, 0.f};
      for (int ks = 0; ks < nks; ks += PF) {
        const bool wrap = ks + PF >= nks;
#pragma unroll
        for (int q = 0; q < PF; ++q) step(ks + q, fr[q], wrap ? ctn : ct, wrap ? q : ks + PF + q);
      }
      // stage 2: S += Re(E^T G) over this column tile
      if constexpr (F32) {
        float e[LT][8];
#pragma unroll
        for (int tx = 0; tx < LT; ++tx) {
          const float4 e0 = Ef[((ct * LT + tx) * 64 + lane) * 2], e1 = Ef[((ct * LT + tx) * 64 + lane) * 2 + 1];
          e[tx][0] = e0.x; e[tx][1] = e0.y; e[tx][2] = e0.z; e[tx][3] = e0.w;
          e[tx][4] = e1.x; e[tx][5] = e1.y; e[tx][6] = e1.z; e[tx][7] = e1.w;
        }
#pragma unroll
        for (int d = 0; d < DT; ++d)
#pragma unroll
          for (int p = 0; p < PT; ++p)
#pragma unroll
            for (int ty = 0; ty < LT; ++ty)
#pragma unroll
              for (int tx = 0; tx < LT; ++tx) {
                floatx4 o = out[d][p][tx * LT + ty];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                  o = mfma_f32(e[tx][2 * v], gacc[d][p][2 * ty][v], o);
                  o = mfma_f32(e[tx][2 * v + 1], gacc[d][p][2 * ty + 1][v], o);
                }
                out[d][p][tx * LT + ty] = o;
              }
        continue;
      }
      bf16x8 eh[LT], el[LT];
#pragma unroll
      for (int tx = 0; tx < LT; ++tx) {
        eh[tx] = Et[((ct * LT + tx) * 2) * 64 + lane];
        el[tx] = Et[((ct * LT + tx) * 2 + 1) * 64 + lane];
      }
#pragma unroll
      for (int d = 0; d < DT; ++d) {
#pragma unroll
        for (int p = 0; p < PT; ++p) {
#pragma unroll
          for (int ty = 0; ty < LT; ++ty) {
            float x[8];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              x[2 * v] = gacc[d][p][2 * ty][v];
              x[2 * v + 1] = gacc[d][p][2 * ty + 1][v];
            }
            bf16x8 gh, gl;
            split8(x, gh, gl);
#pragma unroll
            for (int tx = 0; tx < LT; ++tx) {
              floatx4 o = out[d][p][tx * LT + ty];
              o = mfma_bf16(eh[tx], gh, o);
              o = mfma_bf16(eh[tx], gl, o);
              o = mfma_bf16(el[tx], gh, o);
              out[d][p][tx * LT + ty] = o;
            }
          }
        }
      }
    }

    if (lone) {
      // The last column c = m - 1 of the wave's pairs, sixteen pairs per tile:
      // M slot i of group q holds pair k = 16 q + i = d * PT + p (slots past the
      // last pair hold a copy of pair 0, never read).  Stage 1 as above on
      // split-bf16 MFMA; stage 2 is the rank-one update
      // S[x][y] += Re(E[c][x] G[c][y]) in FP32, G[c][y] of pair k fetched from
      // lane ((k % 16) / 4, y) of the group's accumulator.
      // E[m - 1][x] for this lane's lags x = 16 tx + 4g + v, in float64 -> float
      float er[LT][4], ei[LT][4];
      {
        const int sh = (m - 1 + m / 2) % m;
        const double w = (sh == 0 || sh == n / 2) ? 1.0 : 2.0;
#pragma unroll
        for (int tx = 0; tx < LT; ++tx)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int x = 16 * tx + 4 * g + v;
            double sn = 0.0, cs = 0.0;
            if (x < C) {
              const long kk = ((long)sh * crop_lag(x, n, C)) % n;
              sincospi(2.0 * (double)kk / (double)n, &sn, &cs);
            }
            er[tx][v] = (float)(w * cs);
            ei[tx][v] = (float)(w * sn);
          }
      }
      constexpr int NG = (DT * PT + 15) / 16;
#pragma unroll
      for (int q = 0; q < NG; ++q) {
        const int k = 16 * q + i < DT * PT ? 16 * q + i : 0;
        const float2* lt = tp[0];
        const float2* lq = pp[0];
#pragma unroll
        for (int d = 1; d < DT; ++d)
          if (k / PT == d) lt = tp[d];
#pragma unroll
        for (int p = 1; p < PT; ++p)
          if (k % PT == p) lq = pp[p];
        floatx4 gn[2 * LT];
#pragma unroll
        for (int t = 0; t < 2 * LT; ++t) gn[t] = floatx4{0.f, 0.f, 0.f, 0.f};
        float2 na[4], nq[4];
        auto nload = [&](int ks) {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int r = 16 * ks + 4 * g + j;
            const int lone_at = nks * nct * 256 + 16 * ks + 4 * g + j;  // fragment order: the lone column
            na[j] = FR ? lt[lone_at] : lt[min(r * ldt + (m - 1), lastt)];
            nq[j] = FR ? lq[lone_at] : lq[min(r * ldp + (m - 1), lastp)];
          }
        };
        nload(0);
        for (int ks = 0; ks < nks; ++ks) {
          float x[8];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            cprod(na[j], nq[j], x[2 * j], x[2 * j + 1]);
          }
          if (ks + 1 < nks) nload(ks + 1);
          if constexpr (F32) {
#pragma unroll
            for (int t = 0; t < 2 * LT; ++t) {
              const float4 w0 = Wf[((ks * 2 * LT + t) * 64 + lane) * 2];
              const float4 w1 = Wf[((ks * 2 * LT + t) * 64 + lane) * 2 + 1];
              const float w[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
              for (int kk = 0; kk < 8; ++kk) gn[t] = mfma_f32(x[kk], w[kk], gn[t]);
            }
          } else {
            bf16x8 ahi, alo;
            split8(x, ahi, alo);
#pragma unroll
            for (int t = 0; t < 2 * LT; ++t) {
              const bf16x8 bh = Wl[((ks * 2 * LT + t) * 2) * 64 + lane];
              const bf16x8 bl = Wl[((ks * 2 * LT + t) * 2 + 1) * 64 + lane];
              gn[t] = mfma_bf16(ahi, bh, gn[t]);
              gn[t] = mfma_bf16(ahi, bl, gn[t]);
              gn[t] = mfma_bf16(alo, bh, gn[t]);
            }
          }
        }
#pragma unroll
        for (int d = 0; d < DT; ++d)
#pragma unroll
          for (int p = 0; p < PT; ++p) {
            const int kp = d * PT + p;
            if (kp / 16 != q) continue;  // compile-time after unrolling
            const int src = ((kp & 15) >> 2) * 16 + i;
#pragma unroll
            for (int ty = 0; ty < LT; ++ty) {
              const float gr = __shfl(gn[2 * ty][kp & 3], src, 64);
              const float gi = __shfl(gn[2 * ty + 1][kp & 3], src, 64);
#pragma unroll
              for (int tx = 0; tx < LT; ++tx)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                  out[d][p][tx * LT + ty][v] += __builtin_fmaf(er[tx][v], gr, -(ei[tx][v] * gi));
            }
          }
      }
    }

    // pair tail: first maximum over the crop in y * C + x order; lane (g, i)
    // holds S[x = 16 tx + 4g + v][y = 16 ty + i]
#pragma unroll
    for (int d = 0; d < DT; ++d) {
#pragma unroll
      for (int p = 0; p < PT; ++p) {
        if (j0 + d >= Mi || p0 + p >= P) continue;  // wave-uniform
        const int dir = dir_idx[d0 + j0 + d];
        const float sc = 1024.0f * (inv_norm ? inv_norm[(size_t)pb * P + p0 + p] : 1.0f) * (tinv ? tinv[dir] : 1.0f);
        float bv = -INFINITY;
        int bi = 0x7fffffff;
#pragma unroll
        for (int tx = 0; tx < LT; ++tx)
#pragma unroll
          for (int ty = 0; ty < LT; ++ty)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int x = 16 * tx + 4 * g + v, y = 16 * ty + i;
              const bool in = (y < C) & (x < C);
              first_max(bv, bi, in ? out[d][p][tx * LT + ty][v] * sc : -INFINITY, in ? y * C + x : 0x7fffffff);
            }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
          const float ov = __shfl_xor(bv, o, 64);
          const int oi = __shfl_xor(bi, o, 64);
          first_max(bv, bi, ov, oi);
        }
        if (lane == 0) {
          const size_t gidx = (size_t)(d0 + j0 + d) * P + p0 + p;
          scores[gidx] = bv;
          peak_y[gidx] = bi / C;
          peak_x[gidx] = bi % C;
        }
      }
    }
    if constexpr (!LIST) break;
  }
}

// ---------------------------------------------------------------------------
// FP16-operand variant of the search.  The psi spectra are stored as fp16
// pairs, normalised and scaled by 2^7 (pack_fragments16_kernel), in the fp16
// fragment order: tile value (row 16 ks + 4g + j, column 16 ct + i) at
// tile * 256 + 4 (16 g + i) + j, one 16-byte run per lane and tile (the lone
// last column as in fragment order).  The template rows stay float32 in
// fragment order; each step scales a lane's four values by 2^7 / norm and
// rounds them to fp16.  Q' = T' P' (= 2^14 Q / (|T| |P|)) is formed with
// packed fp16 multiply-adds and is the stage-1 A operand as it is; the row
// table W is split into fp16 hi + lo (WS: two MFMAs per tile) or used as fp16
// hi alone.  Stage 2 and the pair tail are those of the split-bf16 search;
// the pair score is 1024 * 2^-14 * S'.
// ---------------------------------------------------------------------------
constexpr float kH16Scale = 128.0f;          // 2^7 per operand
constexpr float kH16Unscale = 1024.0f / 16384.0f;  // 1024 * 2^-14

__device__ __forceinline__ floatx4 mfma_f16(f16x8 a, f16x8 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// s t as the fp16 pairs th = (tr, ti) and tn = (-ti, tr) (i s t)
__device__ __forceinline__ void t_pair(float2 t, float s, f16x2& th, f16x2& tn) {
  const float tr = t.x * s, ti = t.y * s;
  const f32x2 v0 = {tr, ti}, v1 = {-ti, tr};
  th = __builtin_convertvector(v0, f16x2);
  tn = __builtin_convertvector(v1, f16x2);
}

// t (pr + i pi) = pr t + pi (i t): one packed fp16 multiply (pi (i t),
// rounded to fp16) and one packed fused multiply-add, in this order whatever
// the tile shape (bit-identical results across the kernel's variants)
__device__ __forceinline__ f16x2 cmul_h(f16x2 th, f16x2 tn, _Float16 pr, _Float16 pi) {
  const f16x2 r = {pr, pr}, i = {pi, pi};
  return __builtin_elementwise_fma(r, th, i * tn);
}

__device__ __forceinline__ f16x2 to_h2(float2 v, float s) {
  const f32x2 x = {v.x * s, v.y * s};
  return __builtin_convertvector(x, f16x2);
}

// dst [rows][frag_plane(n)] fp16 pairs: src [rows][frag_plane(n)] complex64
// (fragment order) times scale * inv_norm[row] (inv_norm may be NULL), in the
// fp16 fragment order.
__global__ void pack_fragments16_kernel(const float2* __restrict__ src, const float* __restrict__ inv_norm, int rows,
                                        int n, float scale, f16x2* __restrict__ dst) {
  const int F = frag_plane(n), tiles = frag_nks(n) * frag_nct(n) * 256;
  const size_t total = (size_t)rows * F;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const size_t row = t / F;
    const int e = (int)(t - row * F);
    int s = e;
    if (e < tiles) {
      const int w = e & 255, lane = w >> 2, j = w & 3;
      s = (e & ~255) + (j >> 1) * 128 + 2 * lane + (j & 1);
    }
    const float sc = scale * (inv_norm ? inv_norm[row] : 1.0f);
    dst[t] = to_h2(src[row * F + s], sc);
  }
}

// Near-tie selection after an fp16-operand search: particle b's pairs whose
// score is within max(rel |peak_b|, abs_margin) of its best go on the list
// (b, pair index) for the split-bf16 rescoring; the best pair always does.
// Two passes without atomics: WRITE = false counts each particle's pairs
// (cnt[b]); order_scan_kernel turns the counts into offsets in the
// particles' processing order; WRITE = true writes each particle's entries
// there, so a particle's entries (and, in a batched search, the slots'
// copies of one particle) are adjacent on the list.
template <bool WRITE>
__global__ void __launch_bounds__(256)
split_select_kernel(const int* __restrict__ dir_off, int B, int P, const float* __restrict__ scores,
                    const int* __restrict__ best, const float* __restrict__ peak, float rel, float abs_margin,
                    int* __restrict__ cnt, const int* __restrict__ off, int2* __restrict__ list) {
  const int b = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;  // one wave per particle
  if (b >= B) return;  // no barrier below
  if (best[b] < 0) {
    if (!WRITE && lane == 0) cnt[b] = 0;
    return;
  }
  const float pk = peak[b];
  const float thr = pk - fmaxf(rel * fabsf(pk), abs_margin);
  const int d0 = dir_off[b], np = (dir_off[b + 1] - d0) * P;
  const float* s = scores + (size_t)d0 * P;
  int base = WRITE ? off[b] : 0;
  for (int k0 = 0; k0 < np; k0 += 64) {
    const int k = k0 + lane;
    const bool keep = k < np && s[k] >= thr;
    const unsigned long long m = __ballot(keep);
    if (WRITE && keep) list[base + __popcll(m & ((1ull << lane) - 1))] = make_int2(b, k);
    base += __popcll(m);
  }
  if (!WRITE && lane == 0) cnt[b] = base;
}

template <int DT, int PT>
struct Frags16 {
  float2 t[DT][4];  // float32 template values (T16 = false)
  f16x8 t16[DT];    // fp16 template pairs (T16 = true)
  f16x8 p[PT];      // four fp16 pairs: rows 16 ks + 4g + j, j = 0..3
};

// i t = (-ti, tr) of an fp16 pair t = (tr, ti): one packed multiply
__device__ __forceinline__ f16x2 times_i(f16x2 t) {
  const f16x2 sw = {t[1], t[0]};
  const f16x2 ng = {(_Float16)-1.0f, (_Float16)1.0f};
  return sw * ng;
}

// T16: the bank is fp16 pairs too (gis_pack_fragments_f16 of the float32
// bank with scale 128 and the rows' 1/norms), read like the spectra.
template <int LT, int DT, int PT, bool WS, bool T16>
__global__ void __launch_bounds__(64 * kWaves)
// one direction per wave fits 168 registers (three waves per SIMD), two
// directions and five psi 256 (two waves)
__attribute__((amdgpu_waves_per_eu(DT * PT * LT * LT <= 5 ? 3 : DT * PT * LT * LT <= 10 ? 2 : 1)))
split16_correlate_kernel(const void* __restrict__ bankv, const f16x2* __restrict__ spectra,
                         const float* __restrict__ tinv, const int* __restrict__ dir_off,
                         const int* __restrict__ dir_idx, const int* __restrict__ pmap,
                         const int* __restrict__ order, const int* __restrict__ unit_off, int B, int P, int n, int C,
                         int nblk, const void* __restrict__ Wtv, const void* __restrict__ Etv,
                         float* __restrict__ scores, int* __restrict__ peak_y, int* __restrict__ peak_x) {
  // row table in LDS: [ks][2 LT tiles][hi, lo][64 lanes] (WS) or hi alone
  extern __shared__ f16x8 tab16[];
  const f16x8* Wt = (const f16x8*)Wtv;
  const bf16x8* Et = (const bf16x8*)Etv;
  const int m = n / 2 + 1;
  const bool lone = (m & 15) == 1;
  const int nct = lone ? (m >> 4) : (m + 15) >> 4;
  const int nks = (((n + 15) >> 4) + 1) & ~1;
  constexpr int NH = WS ? 2 : 1;
  const int nw = nks * 2 * LT * NH * 64;
  for (int e = threadIdx.x; e < nw; e += 64 * kWaves) {
    const int lane = e & 63, f = e >> 6;  // f = (ks * 2 LT + t) * NH + h
    tab16[e] = Wt[((f / NH) * 2 + (f % NH)) * 64 + lane];
  }
  __syncthreads();

  const int per = (int)(gridDim.x / kXcds);
  const int lb = (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
  const int lane = threadIdx.x & 63;
  const int u = __builtin_amdgcn_readfirstlane(lb * kWaves + (int)(threadIdx.x >> 6));
  const int total = unit_off[B];
  if (lb >= nblk || u >= total) return;  // no barrier below
  int lo = 0, hi = B;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (unit_off[mid] <= u) lo = mid; else hi = mid;
  }
  const int b = order ? order[lo] : lo;
  const int d0 = dir_off[b], Mi = dir_off[b + 1] - d0;
  const int j0 = (u - unit_off[lo]) * DT;
  const int p0 = (int)blockIdx.y * PT;
  const int pb = pmap ? pmap[b] : b;
  const size_t fplane = (size_t)frag_plane(n);
  const float2* tp[DT];
  const f16x2* tp16[DT];
  float ts[DT];
  const f16x2* pp[PT];
#pragma unroll
  for (int d = 0; d < DT; ++d) {
    const int dir = dir_idx[d0 + min(j0 + d, Mi - 1)];
    tp[d] = (const float2*)bankv + (size_t)dir * fplane;
    tp16[d] = (const f16x2*)bankv + (size_t)dir * fplane;
    ts[d] = kH16Scale * (tinv && !T16 ? tinv[dir] : 1.0f);
  }
#pragma unroll
  for (int p = 0; p < PT; ++p) pp[p] = spectra + ((size_t)pb * P + min(p0 + p, P - 1)) * fplane;
  const int g = lane >> 4, i = lane & 15;

  auto load_tpl = [&](int ct, int ks, Frags16<DT, PT>& f) {
    const int tile = (ks * nct + ct) * 256;
#pragma unroll
    for (int d = 0; d < DT; ++d) {
      if constexpr (T16) {
        f.t16[d] = *(const f16x8*)(tp16[d] + tile + 4 * lane);
      } else {
        const float4 q0 = *(const float4*)(tp[d] + tile + 2 * lane);
        const float4 q1 = *(const float4*)(tp[d] + tile + 128 + 2 * lane);
        f.t[d][0] = make_float2(q0.x, q0.y);
        f.t[d][1] = make_float2(q0.z, q0.w);
        f.t[d][2] = make_float2(q1.x, q1.y);
        f.t[d][3] = make_float2(q1.z, q1.w);
      }
    }
  };
  auto load_psi = [&](int ct, int ks, int p, Frags16<DT, PT>& f) {
    f.p[p] = *(const f16x8*)(pp[p] + (ks * nct + ct) * 256 + 4 * lane);
  };
  auto load = [&](int ct, int ks, Frags16<DT, PT>& f) {
    load_tpl(ct, ks, f);
#pragma unroll
    for (int p = 0; p < PT; ++p) load_psi(ct, ks, p, f);
  };

  floatx4 out[DT][PT][LT * LT], gacc[DT][PT][2 * LT];
#pragma unroll
  for (int d = 0; d < DT; ++d)
#pragma unroll
    for (int p = 0; p < PT; ++p)
#pragma unroll
      for (int t = 0; t < LT * LT; ++t) out[d][p][t] = floatx4{0.f, 0.f, 0.f, 0.f};

  auto step = [&](int ks, Frags16<DT, PT>& f, int ct2, int ks2, bool first) {
    f16x8 wh[2 * LT], wl[2 * LT];
#pragma unroll
    for (int t = 0; t < 2 * LT; ++t) {
      wh[t] = tab16[((ks * 2 * LT + t) * NH) * 64 + lane];
      if constexpr (WS) wl[t] = tab16[((ks * 2 * LT + t) * NH + 1) * 64 + lane];
    }
    // the template values as fp16 (Tr, Ti) and (-Ti, Tr): Q = Pr T + Pi (i T)
    f16x2 th[DT][4], tn[DT][4];
#pragma unroll
    for (int d = 0; d < DT; ++d)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if constexpr (T16) {
          th[d][j] = f16x2{f.t16[d][2 * j], f.t16[d][2 * j + 1]};
          tn[d][j] = times_i(th[d][j]);
        } else {
          t_pair(f.t[d][j], ts[d], th[d][j], tn[d][j]);
        }
      }
    // psi outermost: once a psi fragment has met every direction its
    // registers take the fragment of step ks2, so the psi loads (the bulk of
    // the step's memory traffic) leave as early as the products allow
#pragma unroll
    for (int p = 0; p < PT; ++p) {
#pragma unroll
      for (int d = 0; d < DT; ++d) {
        f16x8 a;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const f16x2 q = cmul_h(th[d][j], tn[d][j], f.p[p][2 * j], f.p[p][2 * j + 1]);
          a[2 * j] = q[0];
          a[2 * j + 1] = q[1];
        }
        // the first step of a column tile starts the accumulators from zero
        // (an explicit operand rather than a cleared register keeps the
        // accumulators in place across the loop)
#pragma unroll
        for (int t = 0; t < 2 * LT; ++t) {
          gacc[d][p][t] = mfma_f16(a, wh[t], first ? floatx4{0.f, 0.f, 0.f, 0.f} : gacc[d][p][t]);
          if constexpr (WS) gacc[d][p][t] = mfma_f16(a, wl[t], gacc[d][p][t]);
        }
      }
      load_psi(ct2, ks2, p, f);
    }
    load_tpl(ct2, ks2, f);
  };

  Frags16<DT, PT> fa, fb;
  load(0, 0, fa);
  load(0, 1, fb);
  for (int ct = 0; ct < nct; ++ct) {
    const int ctn = min(ct + 1, nct - 1);
    {
      const bool wrap = 2 >= nks;  // nks >= 2 and even
      step(0, fa, wrap ? ctn : ct, wrap ? 0 : 2, true);
      step(1, fb, wrap ? ctn : ct, wrap ? 1 : 3, false);
    }
    for (int ks = 2; ks < nks; ks += 2) {
      const bool wrap = ks + 2 >= nks;
      step(ks, fa, wrap ? ctn : ct, wrap ? 0 : ks + 2, false);
      step(ks + 1, fb, wrap ? ctn : ct, wrap ? 1 : ks + 3, false);
      // keeps the scheduler from moving work across the loop edge, which
      // otherwise makes the register allocator copy every accumulator per trip
      __builtin_amdgcn_sched_barrier(0);
    }
    // stage 2 (split-bf16, as in split_correlate_kernel)
    bf16x8 eh[LT], el[LT];
#pragma unroll
    for (int tx = 0; tx < LT; ++tx) {
      eh[tx] = Et[((ct * LT + tx) * 2) * 64 + lane];
      el[tx] = Et[((ct * LT + tx) * 2 + 1) * 64 + lane];
    }
#pragma unroll
    for (int d = 0; d < DT; ++d) {
#pragma unroll
      for (int p = 0; p < PT; ++p) {
#pragma unroll
        for (int ty = 0; ty < LT; ++ty) {
          float x[8];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            x[2 * v] = gacc[d][p][2 * ty][v];
            x[2 * v + 1] = gacc[d][p][2 * ty + 1][v];
          }
          bf16x8 gh, gl;
          split8(x, gh, gl);
#pragma unroll
          for (int tx = 0; tx < LT; ++tx) {
            floatx4 o = out[d][p][tx * LT + ty];
            o = mfma_bf16(eh[tx], gh, o);
            o = mfma_bf16(eh[tx], gl, o);
            o = mfma_bf16(el[tx], gh, o);
            out[d][p][tx * LT + ty] = o;
          }
        }
      }
    }
  }

  if (lone) {
    // the last column c = m - 1 of the wave's pairs, sixteen pairs per tile
    // (as in split_correlate_kernel): stage 1 on fp16 MFMA, stage 2 the
    // rank-one update in FP32
    float er[LT][4], ei[LT][4];
    {
      const int sh = (m - 1 + m / 2) % m;
      const double w = (sh == 0 || sh == n / 2) ? 1.0 : 2.0;
#pragma unroll
      for (int tx = 0; tx < LT; ++tx)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int x = 16 * tx + 4 * g + v;
          double sn = 0.0, cs = 0.0;
          if (x < C) {
            const long kk = ((long)sh * crop_lag(x, n, C)) % n;
            sincospi(2.0 * (double)kk / (double)n, &sn, &cs);
          }
          er[tx][v] = (float)(w * cs);
          ei[tx][v] = (float)(w * sn);
        }
    }
    constexpr int NG = (DT * PT + 15) / 16;
#pragma unroll
    for (int q = 0; q < NG; ++q) {
      const int k = 16 * q + i < DT * PT ? 16 * q + i : 0;
      const float2* lt = tp[0];
      const f16x2* lt16 = tp16[0];
      float lts = ts[0];
      const f16x2* lq = pp[0];
#pragma unroll
      for (int d = 1; d < DT; ++d)
        if (k / PT == d) { lt = tp[d]; lt16 = tp16[d]; lts = ts[d]; }
#pragma unroll
      for (int p = 1; p < PT; ++p)
        if (k % PT == p) lq = pp[p];
      floatx4 gn[2 * LT];
#pragma unroll
      for (int t = 0; t < 2 * LT; ++t) gn[t] = floatx4{0.f, 0.f, 0.f, 0.f};
      float2 na[4];
      f16x2 na16[4], nq[4];
      auto nload = [&](int ks) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int lone_at = nks * nct * 256 + 16 * ks + 4 * g + j;
          if constexpr (T16) na16[j] = lt16[lone_at]; else na[j] = lt[lone_at];
          nq[j] = lq[lone_at];
        }
      };
      nload(0);
      for (int ks = 0; ks < nks; ++ks) {
        f16x8 a;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          f16x2 th, tn;
          if constexpr (T16) {
            th = na16[j];
            tn = times_i(th);
          } else {
            t_pair(na[j], lts, th, tn);
          }
          const f16x2 qv = cmul_h(th, tn, nq[j][0], nq[j][1]);
          a[2 * j] = qv[0];
          a[2 * j + 1] = qv[1];
        }
        if (ks + 1 < nks) nload(ks + 1);
#pragma unroll
        for (int t = 0; t < 2 * LT; ++t) {
          gn[t] = mfma_f16(a, tab16[((ks * 2 * LT + t) * NH) * 64 + lane], gn[t]);
          if constexpr (WS) gn[t] = mfma_f16(a, tab16[((ks * 2 * LT + t) * NH + 1) * 64 + lane], gn[t]);
        }
      }
#pragma unroll
      for (int d = 0; d < DT; ++d)
#pragma unroll
        for (int p = 0; p < PT; ++p) {
          const int kp = d * PT + p;
          if (kp / 16 != q) continue;  // compile-time after unrolling
          const int src = ((kp & 15) >> 2) * 16 + i;
#pragma unroll
          for (int ty = 0; ty < LT; ++ty) {
            const float gr = __shfl(gn[2 * ty][kp & 3], src, 64);
            const float gi = __shfl(gn[2 * ty + 1][kp & 3], src, 64);
#pragma unroll
            for (int tx = 0; tx < LT; ++tx)
#pragma unroll
              for (int v = 0; v < 4; ++v)
                out[d][p][tx * LT + ty][v] += __builtin_fmaf(er[tx][v], gr, -(ei[tx][v] * gi));
          }
        }
    }
  }

  // pair tail: first maximum over the crop in y * C + x order
#pragma unroll
  for (int d = 0; d < DT; ++d) {
#pragma unroll
    for (int p = 0; p < PT; ++p) {
      if (j0 + d >= Mi || p0 + p >= P) continue;  // wave-uniform
      float bv = -INFINITY;
      int bi = 0x7fffffff;
#pragma unroll
      for (int tx = 0; tx < LT; ++tx)
#pragma unroll
        for (int ty = 0; ty < LT; ++ty)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int x = 16 * tx + 4 * g + v, y = 16 * ty + i;
            const bool in = (y < C) & (x < C);
            first_max(bv, bi, in ? out[d][p][tx * LT + ty][v] * kH16Unscale : -INFINITY, in ? y * C + x : 0x7fffffff);
          }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const float ov = __shfl_xor(bv, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        first_max(bv, bi, ov, oi);
      }
      if (lane == 0) {
        const size_t gidx = (size_t)(d0 + j0 + d) * P + p0 + p;
        scores[gidx] = bv;
        peak_y[gidx] = bi / C;
        peak_x[gidx] = bi % C;
      }
    }
  }
}

inline int split_lt(int C) { return C <= 16 ? 1 : 2; }
inline int split_nct(int n) { return (n / 2 + 1 + 15) / 16; }
inline int split_nks(int n) { return (((n + 15) / 16) + 1) & ~1; }
inline size_t split_w_frags(int n, int C) { return (size_t)split_nks(n) * 4 * split_lt(C) * 64; }
inline size_t split_e_frags(int n, int C) { return (size_t)split_nct(n) * 2 * split_lt(C) * 64; }
inline size_t split_lds_bytes(int n, int C) { return split_w_frags(n, C) * sizeof(bf16x8); }
inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

template <int LT, int DT, int PT, bool F32, bool FR, bool LIST = false, int PF = 2>
int launch_split(dim3 grid, size_t lds, hipStream_t st, const float2* bank, const float2* spectra,
                 const float* inv_norm, const float* tinv, const int* dir_off, const int* dir_idx, const int* pmap,
                 const int* order, const int* unit_off, int B, int P, int n, int C, int ldt, int ldp, int nblk,
                 const void* Wt, const void* Et, float* scores, int* peak_y, int* peak_x,
                 const int2* list = nullptr, const int* list_count = nullptr) {
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute((const void*)split_correlate_kernel<LT, DT, PT, F32, FR, LIST, PF>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)split_lds_bytes(256, LT == 1 ? 16 : 32));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      gis_set_error("gis_correlate_split: cannot reserve LDS: %s", hipGetErrorString(e));
      return (int)e;
    }
    attr_set = true;
  }
  hipLaunchKernelGGL((split_correlate_kernel<LT, DT, PT, F32, FR, LIST, PF>), grid, dim3(64 * kWaves), lds, st, bank,
                     spectra, inv_norm, tinv, dir_off, dir_idx, pmap, order, unit_off, B, P, n, C, ldt, ldp, nblk, Wt,
                     Et, scores, peak_y, peak_x, list, list_count);
  return GIS_OK;
}

template <int LT, int DT, int PT, bool WS, bool T16>
int launch_split16(dim3 grid, size_t lds, hipStream_t st, const void* bank, const f16x2* spectra, const float* tinv,
                   const int* dir_off, const int* dir_idx, const int* pmap, const int* order, const int* unit_off,
                   int B, int P, int n, int C, int nblk, const void* Wt, const void* Et, float* scores, int* peak_y,
                   int* peak_x) {
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute((const void*)split16_correlate_kernel<LT, DT, PT, WS, T16>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)split_lds_bytes(256, LT == 1 ? 16 : 32));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      gis_set_error("gis_correlate_split_f16: cannot reserve LDS: %s", hipGetErrorString(e));
      return (int)e;
    }
    attr_set = true;
  }
  hipLaunchKernelGGL((split16_correlate_kernel<LT, DT, PT, WS, T16>), grid, dim3(64 * kWaves), lds, st, bank,
                     spectra, tinv, dir_off, dir_idx, pmap, order, unit_off, B, P, n, C, nblk, Wt, Et, scores, peak_y,
                     peak_x);
  return GIS_OK;
}

}  // namespace

// Complex values of one plane in fragment order at newbox n.
GIS_API size_t gis_fragment_plane(int n) { return n >= 2 && n % 2 == 0 ? (size_t)frag_plane(n) : 0; }

// dst [rows][gis_fragment_plane(n)] complex64, the planes of src
// [rows][n][ld] (ld >= n/2+1) in fragment order: the searches' operand
// layout read with 16-byte loads (pass ld = 0 to them).
GIS_API int gis_pack_fragments(const void* src, int rows, int n, int ld, void* dst, void* stream) {
  GIS_REQUIRE(rows >= 0 && n >= 2 && n % 2 == 0 && ld >= n / 2 + 1, "gis_pack_fragments: rows %d n %d ld %d", rows,
              n, ld);
  if (rows == 0) return GIS_OK;
  GIS_REQUIRE(src && dst, "gis_pack_fragments: null pointer");
  const size_t total = (size_t)rows * frag_plane(n);
  const int blocks = (int)std::min<size_t>((total + 255) / 256, (size_t)gis_cu_count() * 32);
  hipLaunchKernelGGL(pack_fragments_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (const float2*)src, rows,
                     n, ld, (float2*)dst);
  GIS_LAUNCH_CHECK("gis_pack_fragments");
  return GIS_OK;
}

// dst [rows][gis_fragment_plane(n)] fp16 pairs (real, imaginary): the
// fragment-order planes src [rows][gis_fragment_plane(n)] complex64 times
// scale * inv_norm[row] (inv_norm may be NULL), in the fp16 fragment order
// gis_correlate_split_f16 reads (tile value (row 16 ks + 4g + j, column
// 16 ct + i) at tile * 256 + 4 (16 g + i) + j).
GIS_API int gis_pack_fragments_f16(const void* src, const float* inv_norm, int rows, int n, float scale, void* dst,
                                   void* stream) {
  GIS_REQUIRE(rows >= 0 && n >= 2 && n % 2 == 0 && scale > 0.0f, "gis_pack_fragments_f16: rows %d n %d scale %g",
              rows, n, (double)scale);
  if (rows == 0) return GIS_OK;
  GIS_REQUIRE(src && dst, "gis_pack_fragments_f16: null pointer");
  const size_t total = (size_t)rows * frag_plane(n);
  const int blocks = (int)std::min<size_t>((total + 255) / 256, (size_t)gis_cu_count() * 32);
  hipLaunchKernelGGL(pack_fragments16_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (const float2*)src,
                     inv_norm, rows, n, scale, (f16x2*)dst);
  GIS_LAUNCH_CHECK("gis_pack_fragments_f16");
  return GIS_OK;
}

GIS_API size_t gis_correlate_split_workspace(int n, int C, int B) {
  if (n < 2 || C < 2 || B < 0) return 0;
  // tables, unit offsets [B + 1], the scans' tile totals
  return align256((split_w_frags(n, C) + split_e_frags(n, C)) * sizeof(bf16x8)) +
         align256((size_t)(B + 1 + scan_tiles(B)) * sizeof(int));
}

namespace {

// Argument checks shared by both entry points; GIS_OK when the call may go on.
int split_check(const char* who, int B, int P, int n, int C, int nrows, int ldt, int ldp, int dt, int pt) {
  GIS_REQUIRE((ldt == 0 && ldp == 0) || (ldt >= n / 2 + 1 && ldp >= n / 2 + 1),
              "%s: row strides %d, %d: both 0 (fragment order) or both >= n/2+1 = %d", who, ldt, ldp, n / 2 + 1);
  GIS_REQUIRE(B >= 0 && P >= 1 && nrows >= 0, "%s: B=%d P=%d nrows=%d", who, B, P, nrows);
  GIS_REQUIRE(n >= 4 && n % 2 == 0 && n <= 256, "%s: newbox %d must be even, in [4, 256]", who, n);
  if (C > 32 || C < 2 || C % 2 != 0 || C > n) {
    gis_set_error("%s: crop %d must be even, in [2, 32] and <= newbox", who, C);
    return GIS_EUNSUPPORTED;
  }
  GIS_REQUIRE(dt == 0 || dt == 1 || dt == 2 || dt == 4, "%s: dt %d not in {0, 1, 2, 4}", who, dt);
  GIS_REQUIRE(pt >= 0 && pt <= 5, "%s: pt %d not in [0, 5]", who, pt);
  return GIS_OK;
}

// The launches of one search; ws holds gis_correlate_split_workspace(n, C, B)
// bytes of device memory (tables, then the unit offsets), all of which the
// first two kernels write before the search reads them.
int split_run(const void* bank, const void* spectra, const float* inv_norm, const float* tinv, const int* dir_off,
              const int* dir_idx, const int* pmap, const int* order, int B, int P, int n, int C, int nrows, int ldt,
              int ldp, int dt, int pt, float* scores, int* peak_y, int* peak_x, int* best, float* peak, double* mean,
              double* stdv, void* ws, hipStream_t st, bool f32 = false) {
  const int LT = split_lt(C);
  if (pt == 0) pt = std::min(P, 5);
  if (dt == 0) dt = 1;  // measured fastest at box 256
  if (LT == 2 || (f32 && dt > 2)) dt = 1;  // registers: one direction per wave with two lag tiles
  bf16x8* Wt = (bf16x8*)ws;
  bf16x8* Et = Wt + split_w_frags(n, C);
  int* unit_off = (int*)((char*)ws + align256((split_w_frags(n, C) + split_e_frags(n, C)) * sizeof(bf16x8)));
  hipLaunchKernelGGL(split_tables_kernel, dim3(64), dim3(256), 0, st, n, C, LT, split_nks(n), split_nct(n),
                     f32 ? kTabF32 : kTabBF16, (void*)Wt, (void*)Et);
  GIS_LAUNCH_CHECK("gis_correlate_split tables");
  launch_order_scan<kScanUnits>(dir_off, order, B, dt, unit_off + B + 1, unit_off, st);
  GIS_LAUNCH_CHECK("gis_correlate_split units");
  if (nrows > 0) {
    // units <= sum_b ceil(M_b / dt) <= (nrows + B (dt - 1)) / dt
    const long bound = ((long)nrows + (long)B * (dt - 1) + dt - 1) / dt;
    const int nblk = (int)((bound + kWaves - 1) / kWaves);
    const int grid_x = (nblk + kXcds - 1) / kXcds * kXcds;
    const dim3 grid((unsigned)grid_x, (unsigned)((P + pt - 1) / pt));
    const size_t lds = split_lds_bytes(n, C);
    int rc = GIS_OK;
#define GIS_SPLIT_ARGS                                                                                          \
  grid, lds, st, (const float2*)bank, (const float2*)spectra, inv_norm, tinv, dir_off, dir_idx, pmap, order,    \
      unit_off, B, P, n, C, ldt, ldp, nblk, Wt, Et, scores, peak_y, peak_x
#define GIS_SPLIT_PT(LTV, DTV, F, R)                                         \
  switch (pt) {                                                              \
    case 1: rc = launch_split<LTV, DTV, 1, F, R>(GIS_SPLIT_ARGS); break;     \
    case 2: rc = launch_split<LTV, DTV, 2, F, R>(GIS_SPLIT_ARGS); break;     \
    case 3: rc = launch_split<LTV, DTV, 3, F, R>(GIS_SPLIT_ARGS); break;     \
    case 4: rc = launch_split<LTV, DTV, 4, F, R>(GIS_SPLIT_ARGS); break;     \
    default: rc = launch_split<LTV, DTV, 5, F, R>(GIS_SPLIT_ARGS); break;    \
  }
#define GIS_SPLIT_PT5(LTV, DTV, F)          \
  if (ldt == 0) {                           \
    GIS_SPLIT_PT(LTV, DTV, F, true)         \
  } else {                                  \
    GIS_SPLIT_PT(LTV, DTV, F, false)        \
  }
    if (f32) {
      if (LT == 2) {
        GIS_SPLIT_PT5(2, 1, true)
      } else if (dt == 1) {
        GIS_SPLIT_PT5(1, 1, true)
      } else {
        GIS_SPLIT_PT5(1, 2, true)
      }
    } else if (LT == 2) {
      GIS_SPLIT_PT5(2, 1, false)
    } else if (dt == 1) {
      GIS_SPLIT_PT5(1, 1, false)
    } else if (dt == 2) {
      GIS_SPLIT_PT5(1, 2, false)
    } else {
      GIS_SPLIT_PT5(1, 4, false)
    }
#undef GIS_SPLIT_PT5
#undef GIS_SPLIT_PT
#undef GIS_SPLIT_ARGS
    if (rc != GIS_OK) return rc;
    GIS_LAUNCH_CHECK("gis_correlate_split");
  }
  hipLaunchKernelGGL(particle_tail_kernel, dim3(B), dim3(64), 0, st, dir_off, P, scores, best, peak, mean, stdv);
  GIS_LAUNCH_CHECK("gis_correlate_split tail");
  return GIS_OK;
}

}  // namespace

namespace {

// Near-tie margin of the refined fp16 search (rawCC units): a pair is
// rescored on split-bf16 MFMA when its fp16 score is within
// max(kRefineRel |best|, kRefineAbs / sqrt(n (n/2+1))) of its particle's best
// fp16 score.  The fp16 products carry ~11 significant bits: at box 256 the
// fp16 scores deviate from the FP32 search's by at most 1.6e-4 of the peak
// (0.04 rawCC), so the margin (2e-3 of the peak, >= 0.35 rawCC at n = 256)
// keeps every pair that could be the best in exact arithmetic.
constexpr float kRefineRel = 1.0f / 512.0f;
constexpr float kRefineAbs = 64.0f;
constexpr int kRefineBlocks = 512;  // persistent grid of the rescoring pass: the resident blocks (a multiple of kXcds)

size_t split16_list_offset(int n, int C, int B) { return align256(gis_correlate_split_workspace(n, C, B)); }
// the selection's counts and offsets: 64 spare ints, cnt[B], off[B+1]
size_t split16_count_bytes(int B) { return align256((size_t)(64 + 2 * B + 1) * sizeof(int)); }

// The launches of one fp16-operand search (ws as in split_run, plus the pair
// list when `spectra32` is given: the refining pass rescoring each particle's
// best pair and its near ties with the split-bf16 operator from the float32
// spectra).
int split16_run(const void* bank, const void* bank16, const void* spectra, const void* spectra32,
                const float* inv_norm, const float* tinv, const int* dir_off, const int* dir_idx, const int* pmap,
                const int* order, int B, int P, int n, int C, int nrows, int dt, int pt, int wsplit, float* scores,
                int* peak_y, int* peak_x, int* best, float* peak, double* mean, double* stdv, void* ws,
                hipStream_t st) {
  const int LT = split_lt(C);
  if (pt == 0) pt = std::min(P, 5);
  if (dt == 0) dt = 2;  // measured fastest at box 256
  if (LT == 2) dt = 1;
  if (dt > 2) dt = 2;  // four directions per wave exceed the registers
  bf16x8* Wt = (bf16x8*)ws;
  bf16x8* Et = Wt + split_w_frags(n, C);
  int* unit_off = (int*)((char*)ws + align256((split_w_frags(n, C) + split_e_frags(n, C)) * sizeof(bf16x8)));
  int* count = (int*)((char*)ws + split16_list_offset(n, C, B));
  int2* list = (int2*)((char*)count + split16_count_bytes(B));
  hipLaunchKernelGGL(split_tables_kernel, dim3(64), dim3(256), 0, st, n, C, LT, split_nks(n), split_nct(n), kTabF16,
                     (void*)Wt, (void*)Et);
  GIS_LAUNCH_CHECK("gis_correlate_split_f16 tables");
  launch_order_scan<kScanUnits>(dir_off, order, B, dt, unit_off + B + 1, unit_off, st);
  GIS_LAUNCH_CHECK("gis_correlate_split_f16 units");
  if (nrows > 0) {
    const long bound = ((long)nrows + (long)B * (dt - 1) + dt - 1) / dt;
    const int nblk = (int)((bound + kWaves - 1) / kWaves);
    const int grid_x = (nblk + kXcds - 1) / kXcds * kXcds;
    const dim3 grid((unsigned)grid_x, (unsigned)((P + pt - 1) / pt));
    const size_t lds = split_lds_bytes(n, C) / (wsplit ? 1 : 2);
    int rc = GIS_OK;
#define GIS_S16_ARGS                                                                                          \
  grid, lds, st, bank16 ? bank16 : bank, (const f16x2*)spectra, tinv, dir_off, dir_idx, pmap, order, unit_off, \
      B, P, n, C, nblk, (const void*)Wt, (const void*)Et, scores, peak_y, peak_x
#define GIS_S16_PTT(LTV, DTV, WSV, TV)                                           \
  switch (pt) {                                                                  \
    case 1: rc = launch_split16<LTV, DTV, 1, WSV, TV>(GIS_S16_ARGS); break;      \
    case 2: rc = launch_split16<LTV, DTV, 2, WSV, TV>(GIS_S16_ARGS); break;      \
    case 3: rc = launch_split16<LTV, DTV, 3, WSV, TV>(GIS_S16_ARGS); break;      \
    case 4: rc = launch_split16<LTV, DTV, 4, WSV, TV>(GIS_S16_ARGS); break;      \
    default: rc = launch_split16<LTV, DTV, 5, WSV, TV>(GIS_S16_ARGS); break;     \
  }
#define GIS_S16_PT(LTV, DTV, WSV)        \
  if (bank16) {                          \
    GIS_S16_PTT(LTV, DTV, WSV, true)     \
  } else {                               \
    GIS_S16_PTT(LTV, DTV, WSV, false)    \
  }
#define GIS_S16_WS(LTV, DTV)    \
  if (wsplit) {                 \
    GIS_S16_PT(LTV, DTV, true)  \
  } else {                      \
    GIS_S16_PT(LTV, DTV, false) \
  }
    if (LT == 2) {
      GIS_S16_WS(2, 1)
    } else if (dt == 1) {
      GIS_S16_WS(1, 1)
    } else {
      GIS_S16_WS(1, 2)
    }
#undef GIS_S16_WS
#undef GIS_S16_PT
#undef GIS_S16_PTT
#undef GIS_S16_ARGS
    if (rc != GIS_OK) return rc;
    GIS_LAUNCH_CHECK("gis_correlate_split_f16");
  }
  hipLaunchKernelGGL(particle_tail_kernel, dim3(B), dim3(64), 0, st, dir_off, P, scores, best, peak, mean, stdv);
  GIS_LAUNCH_CHECK("gis_correlate_split_f16 tail");
  if (!spectra32 || nrows == 0) return GIS_OK;

  // refine: the best pair and its near ties of every particle on split-bf16
  const float abs_margin = kRefineAbs / sqrtf((float)n * (float)(n / 2 + 1));
  int* cnt = count + 64;  // [B] pair counts, then [B+1] offsets (off[B] = the list length)
  int* off = cnt + B;
  hipLaunchKernelGGL(split_select_kernel<false>, dim3((B + 3) / 4), dim3(256), 0, st, dir_off, B, P, scores, best,
                     peak, kRefineRel, abs_margin, cnt, (const int*)nullptr, (int2*)nullptr);
  GIS_LAUNCH_CHECK("gis_correlate_split_f16 select");
  launch_order_scan<kScanCounts>(cnt, order, B, 1, unit_off + B + 1, off, st);  // the units' tile totals are spent
  GIS_LAUNCH_CHECK("gis_correlate_split_f16 select scan");
  hipLaunchKernelGGL(split_select_kernel<true>, dim3((B + 3) / 4), dim3(256), 0, st, dir_off, B, P, scores, best,
                     peak, kRefineRel, abs_margin, cnt, off, list);
  GIS_LAUNCH_CHECK("gis_correlate_split_f16 select write");
  hipLaunchKernelGGL(split_tables_kernel, dim3(64), dim3(256), 0, st, n, C, LT, split_nks(n), split_nct(n), kTabBF16,
                     (void*)Wt, (void*)Et);
  GIS_LAUNCH_CHECK("gis_correlate_split_f16 refine tables");
  {
    int rc = GIS_OK;
    const dim3 grid(kRefineBlocks, 1);
    const size_t lds = split_lds_bytes(n, C);
#define GIS_REFINE(LTV, PFV)                                                                                      \
  rc = launch_split<LTV, 1, 1, false, true, true, PFV>(grid, lds, st, (const float2*)bank,                         \
                                                       (const float2*)spectra32, inv_norm, tinv, dir_off, dir_idx, \
                                                       pmap, order, unit_off, B, P, n, C, 0, 0, 0, (const void*)Wt, \
                                                       (const void*)Et, scores, peak_y, peak_x, list, off + B)
    const bool deep = split_nks(n) % 4 == 0;  // four steps ahead where the row steps allow it
    if (LT == 2) {
      if (deep) GIS_REFINE(2, 4); else GIS_REFINE(2, 2);
    } else {
      if (deep) GIS_REFINE(1, 4); else GIS_REFINE(1, 2);
    }
#undef GIS_REFINE
    if (rc != GIS_OK) return rc;
    GIS_LAUNCH_CHECK("gis_correlate_split_f16 refine");
  }
  hipLaunchKernelGGL(particle_tail_kernel, dim3(B), dim3(64), 0, st, dir_off, P, scores, best, peak, mean, stdv);
  GIS_LAUNCH_CHECK("gis_correlate_split_f16 refine tail");
  return GIS_OK;
}

}  // namespace

// Workspace bytes of gis_correlate_split_f16 at newbox n, crop C, B
// particles and npairs = dir_off[B] * P searched pairs.
GIS_API size_t gis_correlate_split_f16_workspace(int n, int C, int B, long npairs) {
  if (n < 2 || C < 2 || B < 0 || npairs < 0) return 0;
  return split16_list_offset(n, C, B) + split16_count_bytes(B) + align256((size_t)npairs * sizeof(int2));
}

// The search of gis_correlate_split with fp16 operands: the psi spectra from
// gis_pack_fragments_f16 (scale 128, with their 1/norms: unit norm times
// 2^7), the template bank float32 in fragment order (tinv: 1/norm per row,
// or NULL for a unit-norm bank) or, when bank16 is given, that bank packed
// the same way (gis_pack_fragments_f16 with scale 128 and tinv), products in
// packed fp16, the row table split into fp16 hi + lo (wsplit = 1) or fp16
// alone (wsplit = 0), FP32 accumulation.  With spectra32 (the float32 fragment-order spectra, and
// their inv_norm) the best pair of every particle and the pairs within the
// near-tie margin are rescored with the split-bf16 operator of
// gis_correlate_split; without it the fp16 scores stand.  Outputs as
// gis_correlate_split; ws: gis_correlate_split_f16_workspace bytes.
GIS_API int gis_correlate_split_f16(const void* bank, const void* bank16, const void* spectra16,
                                    const void* spectra32, const float* inv_norm, const float* tinv,
                                    const int* dir_off, const int* dir_idx, const int* pmap, const int* order, int B,
                                    int P, int n, int C, int nrows, int dt, int pt, int wsplit, float* scores,
                                    int* peak_y, int* peak_x, int* best, float* peak, double* mean, double* stdv,
                                    void* ws, size_t ws_bytes, void* stream) {
  const int rc = split_check("gis_correlate_split_f16", B, P, n, C, nrows, 0, 0, dt, pt);
  if (rc != GIS_OK) return rc;
  if (B == 0) return GIS_OK;
  GIS_REQUIRE(bank && spectra16 && dir_off && dir_idx && scores && peak_y && peak_x && best && peak && mean &&
                  stdv && ws,
              "gis_correlate_split_f16: null pointer");
  GIS_REQUIRE(wsplit == 0 || wsplit == 1, "gis_correlate_split_f16: wsplit %d not 0 or 1", wsplit);
  GIS_REQUIRE(ws_bytes >= gis_correlate_split_f16_workspace(n, C, B, (long)nrows * P),
              "gis_correlate_split_f16: workspace too small");
  return split16_run(bank, bank16, spectra16, spectra32, inv_norm, tinv, dir_off, dir_idx, pmap, order, B, P, n, C,
                     nrows, dt, pt, wsplit, scores, peak_y, peak_x, best, peak, mean, stdv, ws, (hipStream_t)stream);
}

// The FP32 search of gis_correlate with the work decomposition of the split
// search (one wave per particle, direction and psi group; no padded pair
// slots): FP32 MFMA, the same two-stage functional, zero-padded rows.
int gis_correlate_units(const void* bank, const void* spectra, const float* inv_norm, const float* tinv,
                        const int* dir_off, const int* dir_idx, const int* pmap, const int* order, int B, int P, int n,
                        int C, int nrows, int ldt, int ldp, int dt, int pt, float* scores, int* peak_y, int* peak_x,
                        int* best, float* peak, double* mean, double* stdv, void* ws, size_t ws_bytes,
                        hipStream_t st) {
  const int rc = split_check("gis_correlate", B, P, n, C, nrows, ldt, ldp, dt, pt);
  if (rc != GIS_OK) return rc;
  if (B == 0) return GIS_OK;
  GIS_REQUIRE(ws_bytes >= gis_correlate_split_workspace(n, C, B), "gis_correlate: workspace too small");
  return split_run(bank, spectra, inv_norm, tinv, dir_off, dir_idx, pmap, order, B, P, n, C, nrows, ldt, ldp, dt, pt,
                   scores, peak_y, peak_x, best, peak, mean, stdv, ws, st, true);
}

// Workspace of the allocation-free entry point: one array per GPU, resident
// until the process ends.  Host code must not use the array's name as a
// device address; split_scratch() returns the device's own address of it.
constexpr size_t kSplitScratchBytes = (size_t)1 << 20;
static __device__ __align__(256) unsigned char scratch[kSplitScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device with that device current; NULL (and the error set) on failure.
void* split_scratch(const char* who) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    gis_set_error("%s: no usable current device (%d): %s", who, dev, hipGetErrorString(e));
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      gis_set_error("%s: hipGetSymbolAddress(scratch) on device %d: %s", who, dev, hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

}  // namespace

// Looks up the allocation-free workspace of the current device (called by
// gis_init; the entry point also does it on its first call on a device).
int gis_split_scratch_setup() { return split_scratch("gis_init") ? GIS_OK : GIS_EINVAL; }

// Full-plane correlate_templates_and_particle (v606:1418-1449) plus the
// numerics of make_result_records (v606:1452-1491) on split-BF16 MFMA; the
// arguments and outputs of gis_correlate, plus
//   order  [B] processing order of the particles (a permutation) or NULL
//   nrows  dir_off[B], the number of dir_idx entries
//   ldt    complex values between template rows (n/2+1, or more: zero
//          padding), or 0: the bank's planes are in fragment order
//          (gis_pack_fragments, gis_fragment_plane(n) values per plane)
//   ldp    the same for the spectra
//   dt     directions per wave (1, 2 or 4; 0 = automatic)
//   pt     psi values per wave (1..5; 0 = automatic)
GIS_API int gis_correlate_split(const void* bank, const void* spectra, const float* inv_norm, const float* tinv,
                                const int* dir_off, const int* dir_idx, const int* pmap, const int* order, int B,
                                int P, int n, int C, int nrows, int ldt, int ldp, int dt, int pt, float* scores,
                                int* peak_y, int* peak_x, int* best, float* peak, double* mean, double* stdv, void* ws,
                                size_t ws_bytes, void* stream) {
  const int rc = split_check("gis_correlate_split", B, P, n, C, nrows, ldt, ldp, dt, pt);
  if (rc != GIS_OK) return rc;
  if (B == 0) return GIS_OK;
  GIS_REQUIRE(bank && spectra && dir_off && dir_idx && scores && peak_y && peak_x && best && peak && mean && stdv &&
                  ws,
              "gis_correlate_split: null pointer");
  GIS_REQUIRE(ws_bytes >= gis_correlate_split_workspace(n, C, B), "gis_correlate_split: workspace too small");
  return split_run(bank, spectra, inv_norm, tinv, dir_off, dir_idx, pmap, order, B, P, n, C, nrows, ldt, ldp, dt, pt,
                   scores, peak_y, peak_x, best, peak, mean, stdv, ws, (hipStream_t)stream);
}

// gis_correlate_split without a workspace argument and without allocating:
// the workspace is a 1 MiB array per GPU inside the library, so calls with
// gis_correlate_split_workspace(n, C, B) > 1 MiB (about 220000 particles at
// newbox 256, crop 32) are refused.  Calls that overlap on one GPU (other
// streams or threads) share the array and race.
GIS_API int gis_correlate_split_noalloc(const void* bank, const void* spectra, const float* inv_norm,
                                        const float* tinv, const int* dir_off, const int* dir_idx, const int* pmap,
                                        const int* order, int B, int P, int n, int C, int nrows, int ldt, int ldp,
                                        int dt, int pt, float* scores, int* peak_y, int* peak_x, int* best,
                                        float* peak, double* mean, double* stdv, void* stream) {
  const int rc = split_check("gis_correlate_split_noalloc", B, P, n, C, nrows, ldt, ldp, dt, pt);
  if (rc != GIS_OK) return rc;
  if (B == 0) return GIS_OK;
  GIS_REQUIRE(bank && spectra && dir_off && dir_idx && scores && peak_y && peak_x && best && peak && mean && stdv,
              "gis_correlate_split_noalloc: null pointer");
  const size_t need = gis_correlate_split_workspace(n, C, B);
  GIS_REQUIRE(need <= kSplitScratchBytes,
              "gis_correlate_split_noalloc: %d particles at newbox %d, crop %d need %zu bytes of workspace, above "
              "the %zu-byte limit of the allocation-free path",
              B, n, C, need, kSplitScratchBytes);
  void* ws = split_scratch("gis_correlate_split_noalloc");
  if (!ws) return GIS_EINVAL;
  return split_run(bank, spectra, inv_norm, tinv, dir_off, dir_idx, pmap, order, B, P, n, C, nrows, ldt, ldp, dt, pt,
                   scores, peak_y, peak_x, best, peak, mean, stdv, ws, (hipStream_t)stream);
}
