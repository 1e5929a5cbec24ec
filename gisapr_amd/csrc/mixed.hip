// Small-shift correlation on split-BF16 (or split-FP16) MFMA (precision variant).
//
// The v606 score at an integer lag l = (ly, lx) is a fixed linear functional
// of the packed product Q = T * P (correlate_templates_and_particle,
// v606:1418-1449):  score(l) = 1024 * sum_k Re(Q_k * Phi[k][l]).  Keeping the
// K coefficients inside a resolution radius and the lags |ly|, |lx| <= 3
// turns the search of one particle into a dense product
//     S[pairs][lags] = A[pairs][2K] * B[2K][lags],
//     A[p][2k] = Re Q_pk, A[p][2k+1] = Im Q_pk;  B[2k][l] = Re Phi, B[2k+1][l] = -Im Phi
// evaluated on v_mfma_f32_16x16x32_bf16 with both operands split into BF16
// hi + lo parts (hi*hi + hi*lo + lo*hi, FP32 accumulation): ~16 significant
// bits per product instead of BF16's 8.  The FP16 variant
// (v_mfma_f32_16x16x32_f16, same operand layout) splits into FP16 hi + lo,
// ~22 bits per product; FP16's range needs A scaled by 2^14 before the split
// (|A| <= 1 for unit-norm spectra, so |2^14 A| < 65504 and typical products
// stay normal) and the accumulator scaled back by 2^-14, both exact.
//
// gfx950 mapping: one wave per (particle, tile of 16 direction x psi pairs).
// Each lane forms 4 complex Q values per k-step from the FP32 template and
// particle coefficients (two 32-byte loads each), splits them in registers
// (v_cvt_pk_bf16_f32, round to nearest even) and feeds 4 lag tiles x 3
// MFMAs; the B operand is pre-split on the host and laid out so each lane's
// 8 BF16 values are one 16-byte load.  The pair tail (argmax over lags) runs
// in the epilogue with 16-lane shuffles; the per-particle tail (argmax over
// pairs, mean / std in FP64) is a separate one-wave-per-particle kernel.
#include "common.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kLagTiles = 4;  // 64 lag columns (49 used)

__global__ void __launch_bounds__(256)
gather_coeffs_kernel(const float2* __restrict__ src, size_t plane, int rows, const int* __restrict__ kidx, int K,
                     int Kpad, const float* __restrict__ scale, float2* __restrict__ dst) {
  const size_t total = (size_t)rows * Kpad;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const size_t row = t / Kpad;
    const int k = (int)(t - row * Kpad);
    float2 v = make_float2(0.f, 0.f);
    if (k < K) {
      v = src[row * plane + kidx[k]];
      if (scale) {
        const float s = scale[row];
        v.x *= s;
        v.y *= s;
      }
    }
    dst[t] = v;
  }
}

// x = hi + lo with hi = h(x) and lo = h(x - hi), both round to nearest even
// (h = the 16-bit operand type); pairs of values go through one packed
// conversion each way.
template <typename V8, typename V2>
__device__ __forceinline__ void split8(const float (&x)[8], V8& hi, V8& lo) {
#pragma unroll
  for (int e = 0; e < 8; e += 2) {
    const f32x2 v = {x[e], x[e + 1]};
    const V2 h = __builtin_convertvector(v, V2);
    const f32x2 hf = __builtin_convertvector(h, f32x2);
    const V2 l = __builtin_convertvector(v - hf, V2);
    hi[e] = h[0];
    hi[e + 1] = h[1];
    lo[e] = l[0];
    lo[e + 1] = l[1];
  }
}

// Operand types of the split product: 8-element fragments, the MFMA, the A
// scale exponent and the factor from the accumulator to the v606 score
// (1024 x 2^-scale).
struct OpBF16 {
  typedef bf16x8 v8;
  typedef bf16x2 v2;
  static constexpr int kScaleExp = 0;
  static constexpr float kOut = 1024.0f;
  static __device__ __forceinline__ floatx4 mfma(v8 a, v8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};
struct OpF16 {
  typedef f16x8 v8;
  typedef f16x2 v2;
  static constexpr int kScaleExp = 14;
  static constexpr float kOut = 1024.0f / 16384.0f;
  static __device__ __forceinline__ floatx4 mfma(v8 a, v8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};

// One wave per (group of 16 direction rows, chunk of PT psi values).  Row r
// is the r-th entry of dir_idx (particle b with dir_off[b] <= r <
// dir_off[b+1]); M-tile p holds the pairs (row, psi p0 + p).  The template
// coefficients of a row are loaded once per k-step and multiplied by the PT
// particle spectra, and every lag-table fragment feeds PT M-tiles.  The lag
// table streams through LDS in chunks of KC k-steps shared by the WPB waves
// of a block (register-staged, double-buffered).  NT lag tiles of 16 go to
// the MFMAs; with VL the last lag (49 = 3 x 16 + 1) is a float32 dot product
// on the VALU instead of a fourth, almost empty, tile.
template <typename OP, int PT, int WPB, int NT, bool VL>
__global__ void __launch_bounds__(64 * WPB)
mixed_correlate_kernel(const float2* __restrict__ Tk, const float2* __restrict__ Pk,
                       const int* __restrict__ dir_off, const int* __restrict__ dir_idx,
                       const int* __restrict__ pmap, const int* __restrict__ rows, int B, int nrows,
                       int P, int Kpad, const typename OP::v8* __restrict__ Bhi,
                       const typename OP::v8* __restrict__ Blo, const float2* __restrict__ Plast, int half,
                       float* __restrict__ scores, int* __restrict__ peak_y, int* __restrict__ peak_x) {
  constexpr int KC = 4;               // k-steps per LDS chunk
  constexpr int FRAG = NT * 2 * 64;   // lag-table fragments per k-step (hi and lo)
  constexpr int NTH = 64 * WPB;
  constexpr int SPT = (KC * FRAG + NTH - 1) / NTH;  // staged fragments per thread
  typedef typename OP::v8 h8;
  __shared__ h8 blds[2][KC * FRAG];
  const int tid = threadIdx.x, lane = tid & 63;
  int group = blockIdx.x * WPB + (tid >> 6);
  const bool live = group * 16 < nrows;
  if (!live) group = (nrows - 1) / 16;  // keeps the block's barriers uniform; writes nothing
  const int p0 = blockIdx.y * PT;
  const int i = lane & 15, g = lane >> 4;
  const int rslot = min(group * 16 + i, nrows - 1);
  const int r = rows ? rows[rslot] : rslot;  // processing order -> CSR row
  // particle of row r: last b with dir_off[b] <= r (empty particles share offsets)
  int lo = 0, hi = B;  // invariant dir_off[lo] <= r < dir_off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (dir_off[mid] <= r) lo = mid; else hi = mid;
  }
  const int pb = pmap ? pmap[lo] : lo;  // particle row of Pk
  const float4* trow = (const float4*)(Tk + (size_t)dir_idx[r] * Kpad);
  const float4* prow[PT];
#pragma unroll
  for (int p = 0; p < PT; ++p) prow[p] = (const float4*)(Pk + ((size_t)pb * P + min(p0 + p, P - 1)) * Kpad);
  const float4* plast = (const float4*)Plast;
  const int ksteps = Kpad / 16;
  const int nchunk = (ksteps + KC - 1) / KC;
  floatx4 acc[PT][NT];
  float accl[PT];
#pragma unroll
  for (int p = 0; p < PT; ++p) {
    accl[p] = 0.f;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[p][nt] = floatx4{0.f, 0.f, 0.f, 0.f};
  }

  h8 stg[SPT];
  auto fetch = [&](int ch) {
#pragma unroll
    for (int s = 0; s < SPT; ++s) {
      const int e = tid + s * NTH;
      const int kk = e / FRAG, f = e - kk * FRAG;
      const int ks = ch * KC + kk;
      const int nt = f >> 7, h = (f >> 6) & 1, ln = f & 63;
      h8 v = {};
      if (e < KC * FRAG && ks < ksteps) v = (h ? Blo : Bhi)[((size_t)nt * ksteps + ks) * 64 + ln];
      stg[s] = v;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int s = 0; s < SPT; ++s) {
      const int e = tid + s * NTH;
      if (e < KC * FRAG) blds[buf][e] = stg[s];
    }
  };
  fetch(0);
  store(0);
  __syncthreads();

#ifndef GIS_MIXED_PREFETCH
#define GIS_MIXED_PREFETCH 1
#endif
  // This lane's template / particle / last-lag coefficients of one k-step.
  struct KStep {
    float4 t0, t1, f0, f1;
    float4 u[PT][2];
  };
  auto load_ks = [&](int ks, KStep& o) {
    const int q4 = ks * 8 + 2 * g;  // float4 index of this lane's 4 coefficients
    o.t0 = trow[q4];
    o.t1 = trow[q4 + 1];
#pragma unroll
    for (int p = 0; p < PT; ++p) {
      o.u[p][0] = prow[p][q4];
      o.u[p][1] = prow[p][q4 + 1];
    }
    if (VL) {
      o.f0 = plast[q4];
      o.f1 = plast[q4 + 1];
    } else {
      o.f0 = o.f1 = float4{};
    }
  };
  // software pipeline: the next k-step's loads are in flight while this one computes
  KStep nxt;
  if (GIS_MIXED_PREFETCH) load_ks(0, nxt);

  for (int ch = 0; ch < nchunk; ++ch) {
    const int cur = ch & 1;
    if (ch + 1 < nchunk) fetch(ch + 1);  // in flight during this chunk's MFMAs
    const int kend = min(KC, ksteps - ch * KC);
    for (int kk = 0; kk < kend; ++kk) {
      const int ks = ch * KC + kk;
      KStep now;
      if (GIS_MIXED_PREFETCH) {
        now = nxt;
        load_ks(min(ks + 1, ksteps - 1), nxt);
      } else {
        load_ks(ks, now);
      }
      const float4 t0 = now.t0, t1 = now.t1, f0 = now.f0, f1 = now.f1;
      float4 uc[PT][2];
#pragma unroll
      for (int p = 0; p < PT; ++p) { uc[p][0] = now.u[p][0]; uc[p][1] = now.u[p][1]; }
      // the A scale (a power of two, 1 for bf16) goes on the template values
      // once per k-step; every product, and the last-lag sum, carry it exactly
      constexpr float as = (float)(1 << OP::kScaleExp);
      const float tr[4] = {t0.x * as, t0.z * as, t1.x * as, t1.z * as};
      const float ti[4] = {t0.y * as, t0.w * as, t1.y * as, t1.w * as};
      h8 ahi[PT], alo[PT];
#pragma unroll
      for (int p = 0; p < PT; ++p) {
        const float4 u0 = uc[p][0], u1 = uc[p][1];
        const float pr[4] = {u0.x, u0.z, u1.x, u1.z}, pim[4] = {u0.y, u0.w, u1.y, u1.w};
        float x[8];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          x[2 * c] = tr[c] * pr[c] - ti[c] * pim[c];
          x[2 * c + 1] = tr[c] * pim[c] + ti[c] * pr[c];
        }
        if (VL)
          accl[p] += ((x[0] * f0.x + x[1] * f0.y) + (x[2] * f0.z + x[3] * f0.w)) +
                     ((x[4] * f1.x + x[5] * f1.y) + (x[6] * f1.z + x[7] * f1.w));
        split8<h8, typename OP::v2>(x, ahi[p], alo[p]);
      }
      const h8* fr = blds[cur] + kk * FRAG + lane;
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const h8 bh = fr[(nt * 2) * 64], bl = fr[(nt * 2 + 1) * 64];
#pragma unroll
        for (int p = 0; p < PT; ++p) {
          acc[p][nt] = OP::mfma(ahi[p], bh, acc[p][nt]);
          acc[p][nt] = OP::mfma(ahi[p], bl, acc[p][nt]);
          acc[p][nt] = OP::mfma(alo[p], bh, acc[p][nt]);
        }
      }
    }
    if (ch + 1 < nchunk) store(cur ^ 1);
    __syncthreads();
  }

  // epilogue: lane (g, i) holds S[row 4g+v][lag column nt*16+i] of each psi tile; first maximum over lags
  const int side = 2 * half + 1, nlag = side * side;
#pragma unroll
  for (int p = 0; p < PT; ++p) {
    float tl = 0.f;
    if (VL) {  // row i's float32 last-lag sum, over the 4 k-groups of lanes i, i+16, i+32, i+48
      tl = accl[p];
      tl += __shfl_xor(tl, 16, 64);
      tl += __shfl_xor(tl, 32, 64);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      float bv = -INFINITY;
      int bs = 0x7fffffff;
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const int s = nt * 16 + i;
        if (s < (VL ? nlag - 1 : nlag)) {
          const float val = acc[p][nt][v] * OP::kOut;
          if (val > bv) { bv = val; bs = s; }  // s increases with nt: strict > keeps the first
        }
      }
#pragma unroll
      for (int o = 8; o > 0; o >>= 1) {
        const float ov = __shfl_xor(bv, o, 16);
        const int os = __shfl_xor(bs, o, 16);
        if (ov > bv || (ov == bv && os < bs)) { bv = ov; bs = os; }
      }
      const int slot = group * 16 + 4 * g + v;
      const int row = slot < nrows ? (rows ? rows[slot] : slot) : nrows;
      if (VL) {
        const float vl = __shfl(tl, 4 * g + v, 64) * OP::kOut;  // lane 4g+v holds that row's sum
        if (vl > bv) { bv = vl; bs = nlag - 1; }  // the last lag: strict > keeps earlier ties
      }
      if (live && i == 0 && row < nrows && p0 + p < P) {
        const size_t out = (size_t)row * P + p0 + p;
        scores[out] = bv;
        peak_y[out] = bs / side;  // crop index: lag + half
        peak_x[out] = bs % side;
      }
    }
  }
}

}  // namespace

// Gathers the K listed packed coefficients of `rows` spectra into a dense
// [rows][Kpad] block (zero padded), optionally scaled per row (1/norm).
GIS_API int gis_gather_coefficients(const void* src, size_t plane, int rows, const int* kidx, int K, int Kpad,
                                    const float* scale, void* dst, void* stream) {
  GIS_REQUIRE(rows >= 0 && K >= 0 && Kpad >= K, "gis_gather_coefficients: rows=%d K=%d Kpad=%d", rows, K, Kpad);
  if (rows == 0 || Kpad == 0) return GIS_OK;
  GIS_REQUIRE(src && dst && (K == 0 || kidx), "gis_gather_coefficients: null pointer");
  const size_t total = (size_t)rows * Kpad;
  const int blocks = (int)std::min<size_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(gather_coeffs_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (const float2*)src,
                     plane, rows, kidx, K, Kpad, scale, (float2*)dst);
  GIS_LAUNCH_CHECK("gis_gather_coefficients");
  return GIS_OK;
}

namespace {

// Launches the tile variant (pt, wpb, lag tiles) of the search for operand type OP.
template <typename OP>
void launch_mixed(dim3 grid, int pt, int wpb, int nt, bool vl, hipStream_t st, const void* Tk, const void* Pk,
                  const int* dir_off, const int* dir_idx, const int* pmap, const int* rows, int B, int nrows, int P,
                  int Kpad, const void* Bhi, const void* Blo, const float2* pl, int half, float* scores,
                  int* peak_y, int* peak_x) {
  typedef const typename OP::v8* BP;
#define GIS_MIXED_LAUNCH(PTV, WV, NTV, VLV)                                                                   \
  hipLaunchKernelGGL((mixed_correlate_kernel<OP, PTV, WV, NTV, VLV>), grid, dim3(64 * WV), 0, st,              \
                     (const float2*)Tk, (const float2*)Pk, dir_off, dir_idx, pmap, rows, B, nrows, P, Kpad,    \
                     (BP)Bhi, (BP)Blo, pl, half, scores, peak_y, peak_x)
#define GIS_MIXED_NT(PTV, WV)                                 \
  if (vl && nt == 3) GIS_MIXED_LAUNCH(PTV, WV, 3, true);      \
  else if (nt == 1) GIS_MIXED_LAUNCH(PTV, WV, 1, false);      \
  else if (nt == 2) GIS_MIXED_LAUNCH(PTV, WV, 2, false);      \
  else if (nt == 3) GIS_MIXED_LAUNCH(PTV, WV, 3, false);      \
  else GIS_MIXED_LAUNCH(PTV, WV, 4, false);
#define GIS_MIXED_PT(WV)                     \
  switch (pt) {                              \
    case 1: { GIS_MIXED_NT(1, WV) } break;   \
    case 2: { GIS_MIXED_NT(2, WV) } break;   \
    case 3: { GIS_MIXED_NT(3, WV) } break;   \
    case 4: { GIS_MIXED_NT(4, WV) } break;   \
    default: { GIS_MIXED_NT(5, WV) } break;  \
  }
  if (wpb == 2) {
    GIS_MIXED_PT(2)
  } else {
    GIS_MIXED_PT(4)
  }
#undef GIS_MIXED_PT
#undef GIS_MIXED_NT
#undef GIS_MIXED_LAUNCH
}

}  // namespace

// Precision variant of correlate_templates_and_particle (v606:1418-1449)
// over lags [-half, half]^2 (half <= 3) and a truncated coefficient set.
//   Tk [bank rows][Kpad], Pk [particles*P][Kpad] complex64 (unit-norm, unshifted)
//   pmap [B] particle row of each searched particle in Pk, or NULL (identity)
//   rows [nrows] processing order of the dir_idx rows (a permutation), or NULL:
//        rows of the same particle processed together share its Pk rows in
//        cache (several candidates' rows of one particle in a batched launch)
//   nrows = dir_off[B] = number of dir_idx entries (passed so no device read is needed)
//   pt = psi values per wave (1..5, 0 = min(P, 5)); wpb = waves per block (2 or 4; 0 = 4)
//   Bhi, Blo: split lag table, [4][Kpad/16][64] x 8 16-bit values (host layout)
//   Plast: float2 [Kpad] (Re Phi, -Im Phi) of the last lag, used when the lag
//          count is 16 t + 1 (t tiles on MFMA, the last lag on the VALU); else NULL
//   scores / peak_y / peak_x [dir_off[B]][P] (peak indices are lag + half)
//   best / peak / mean / std [B] as gis_correlate
//   operand: GIS_OPERAND_BF16 (Bhi / Blo are bf16) or GIS_OPERAND_F16 (fp16)
GIS_API int gis_correlate_mixed_ex(const void* Tk, const void* Pk, const int* dir_off, const int* dir_idx,
                                   const int* pmap, const int* rows, int B, int P, int nrows, int Kpad,
                                   const void* Bhi, const void* Blo, const void* Plast, int half, int pt, int wpb,
                                   float* scores, int* peak_y, int* peak_x, int* best, float* peak, double* mean,
                                   double* stdv, int operand, void* stream) {
  GIS_REQUIRE(B >= 0 && P >= 1 && nrows >= 0, "gis_correlate_mixed: B=%d P=%d nrows=%d", B, P, nrows);
  GIS_REQUIRE(Kpad % 16 == 0 && Kpad >= 16, "gis_correlate_mixed: Kpad %d must be a positive multiple of 16", Kpad);
  GIS_REQUIRE(operand == GIS_OPERAND_BF16 || operand == GIS_OPERAND_F16,
              "gis_correlate_mixed: operand %d not bf16 (%d) or fp16 (%d)", operand, GIS_OPERAND_BF16,
              GIS_OPERAND_F16);
  if (half < 0 || (2 * half + 1) * (2 * half + 1) > 16 * kLagTiles) {
    gis_set_error("gis_correlate_mixed: half %d gives more than %d lags", half, 16 * kLagTiles);
    return GIS_EUNSUPPORTED;
  }
  if (B == 0) return GIS_OK;
  GIS_REQUIRE(Tk && Pk && dir_off && dir_idx && Bhi && Blo && scores && peak_y && peak_x && best && peak && mean &&
                  stdv,
              "gis_correlate_mixed: null pointer");
  hipStream_t st = (hipStream_t)stream;
  GIS_REQUIRE(pt >= 0 && pt <= 5, "gis_correlate_mixed: pt %d not in [0, 5]", pt);
  const int nlag = (2 * half + 1) * (2 * half + 1);
  const bool vl = nlag % 16 == 1 && nlag > 1;
  const int nt = vl ? nlag / 16 : (nlag + 15) / 16;
  GIS_REQUIRE(!vl || Plast, "gis_correlate_mixed: %d lags need the last-lag table", nlag);
  if (nrows > 0) {
    if (pt == 0) pt = std::min(P, 5);
    if (wpb == 0) wpb = 4;
    GIS_REQUIRE(wpb == 2 || wpb == 4, "gis_correlate_mixed: wpb %d not in {0, 2, 4}", wpb);
    const int rows_per_block = 16 * wpb;
    const dim3 grid((unsigned)((nrows + rows_per_block - 1) / rows_per_block), (unsigned)((P + pt - 1) / pt));
    const float2* pl = (const float2*)Plast;
    if (operand == GIS_OPERAND_F16)
      launch_mixed<OpF16>(grid, pt, wpb, nt, vl, st, Tk, Pk, dir_off, dir_idx, pmap, rows, B, nrows, P, Kpad, Bhi,
                          Blo, pl, half, scores, peak_y, peak_x);
    else
      launch_mixed<OpBF16>(grid, pt, wpb, nt, vl, st, Tk, Pk, dir_off, dir_idx, pmap, rows, B, nrows, P, Kpad, Bhi,
                           Blo, pl, half, scores, peak_y, peak_x);
    GIS_LAUNCH_CHECK("gis_correlate_mixed");
  }
  hipLaunchKernelGGL(particle_tail_kernel, dim3(B), dim3(64), 0, st, dir_off, P, scores, best, peak, mean, stdv);
  GIS_LAUNCH_CHECK("gis_correlate_mixed tail");
  return GIS_OK;
}

GIS_API int gis_correlate_mixed(const void* Tk, const void* Pk, const int* dir_off, const int* dir_idx,
                                const int* pmap, const int* rows, int B, int P, int nrows, int Kpad, const void* Bhi,
                                const void* Blo, const void* Plast, int half, int pt, int wpb, float* scores,
                                int* peak_y, int* peak_x, int* best, float* peak, double* mean, double* stdv,
                                void* stream) {
  return gis_correlate_mixed_ex(Tk, Pk, dir_off, dir_idx, pmap, rows, B, P, nrows, Kpad, Bhi, Blo, Plast, half, pt,
                                wpb, scores, peak_y, peak_x, best, peak, mean, stdv, GIS_OPERAND_BF16, stream);
}
