// Fused FFT cross-correlation search and per-particle reduction.
//
// Reference operator (correlate_templates_and_particle, v606:1418-1449):
// for pre-shifted packed operands T~ (template) and P~ (conjugated
// particle), ccg = irfft2(T~ * P~, s=(n, n)); keep the C x C corner block at
// rows/cols [n-C/2 .. n-1, 0 .. C/2-1]; score = max * 1024 n^2,
// (y, x) = divmod(argmax, C).
//
// That block is a fixed linear functional of Q = T * P (UNshifted
// operands):
//     ccg[y][x] = 1/n^2 * sum_{r,c} Re( Q[r][c] * W[r][y] * E[c][x] )
//     W[r][y]   = (-1)^l_y * exp(+2 pi i r l_y / n)       (row roll by n/2)
//     E[c][x]   = w(s_c) * exp(+2 pi i s_c l_x / n),  s_c = (c + (n/2+1)//2) mod (n/2+1)
//     w(s)      = 1 for s in {0, n/2} else 2             (C2R: Im of columns 0, n/2 dropped)
// with l_i the crop lags.  Per (particle, direction, psi) it is evaluated
// as two chained FP32 MFMA products on gfx950:
//     G^T[c][y]   = sum_r  Q^T[c][r] * W[r][y]     (16 columns x 16 lags, K = n rows)
//     out^T[x][y] = sum_c  Re(E^T[x][c] * G^T[c][y]) (K = 16 columns per tile)
// The 16x16x4 accumulator of the first product is, register by register,
// the B operand of the second (its K index is the column), so G never
// leaves registers.  Q = T*P is formed on the fly from LDS-staged tiles.
//
// Work decomposition: one 256-thread workgroup per particle; a sweep covers
// up to 32 (direction, psi) pairs, 8 per wave.  The W and E tables live in
// LDS for the whole workgroup; the spectra stream through LDS in
// (16-row x 16-column) tiles, double-buffered: the global loads of chunk
// r+1 are in flight (in registers) while the MFMAs consume chunk r.
// The per-particle tail (make_result_records numerics, v606:1452-1491) runs
// in the same workgroup: first-maximum argmax over (direction, psi), mean
// and population std in float64.  A crop of up to 32 runs as four launches,
// one per 16 x 16 lag window, each merged into the stored maxima.
#include "common.h"

typedef float floatx4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = 4;
constexpr int kPairsPerWave = 8;
constexpr int kRC = 16;                 // rows per staged chunk
constexpr int kTileElems = kRC * 16;    // one float2 per thread per tile
constexpr int kMaxTiles = 16;           // templates + psi images per sweep

__device__ __forceinline__ floatx4 mfma(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// W / E tables of the 16 x 16 lag window whose first crop row / column are
// (y0, x0): table column y holds crop index y0 + y (zero beyond the crop).
__global__ void build_corr_tables(int n, int C, int y0, int x0, int nrow_pad, int mpad, float2* __restrict__ Wt,
                                  float2* __restrict__ Et) {
  const int m = n / 2 + 1;
  const int s = m / 2;
  const int total = nrow_pad * 16 + mpad * 16;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    if (t < nrow_pad * 16) {
      const int r = t / 16, y = y0 + t % 16;
      float2 v = make_float2(0.f, 0.f);
      if (r < n && y < C) {
        const int l = y < C / 2 ? n - C / 2 + y : y - C / 2;
        const long k = ((long)r * l) % n;
        double sn, cs;
        sincospi(2.0 * (double)k / (double)n, &sn, &cs);
        const double sg = (l & 1) ? -1.0 : 1.0;
        v = make_float2((float)(sg * cs), (float)(sg * sn));
      }
      Wt[t] = v;
    } else {
      const int u = t - nrow_pad * 16;
      const int c = u / 16, x = x0 + u % 16;
      float2 v = make_float2(0.f, 0.f);
      if (c < m && x < C) {
        const int l = x < C / 2 ? n - C / 2 + x : x - C / 2;
        const int pos = (c + s) % m;
        const double w = (pos == 0 || pos == n / 2) ? 1.0 : 2.0;
        const long k = ((long)pos * l) % n;
        double sn, cs;
        sincospi(2.0 * (double)k / (double)n, &sn, &cs);
        v = make_float2((float)(w * cs), (float)(w * sn));
      }
      Et[u] = v;
    }
  }
}

__global__ void __launch_bounds__(kThreads)
correlate_kernel(const float2* __restrict__ bank, const float2* __restrict__ spectra,
                 const float* __restrict__ inv_norm, const float* __restrict__ tinv,
                 const int* __restrict__ dir_off, const int* __restrict__ dir_idx, const int* __restrict__ pmap,
                 int P, int n, int C, int JT, int PT, int y0, int x0, int merge, int tail,
                 const float2* __restrict__ Wt, const float2* __restrict__ Et,
                 float* __restrict__ scores, int* __restrict__ peak_y, int* __restrict__ peak_x,
                 int* __restrict__ best, float* __restrict__ peak, double* __restrict__ mean,
                 double* __restrict__ stdv) {
  extern __shared__ float2 smem[];
  __shared__ const float2* tptr[kMaxTiles];
  const int b = blockIdx.x;
  const int pb = pmap ? pmap[b] : b;  // particle row of spectra / inv_norm
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m = n / 2 + 1;
  const int nct = (m + 15) / 16;
  const int nrc = (n + kRC - 1) / kRC;
  const int nrow_pad = nrc * kRC;
  float2* Wl = smem;                          // [nrow_pad][16]
  float2* El = Wl + nrow_pad * 16;            // [nct*16][16]
  float2* tiles = El + nct * 16 * 16;         // [2][kMaxTiles][kTileElems]
  const int d0 = dir_off[b];
  const int Mi = dir_off[b + 1] - d0;
  const size_t plane = (size_t)n * m;
  const int lr = lane >> 4, lc = lane & 15;
  const int srow = tid >> 4, scol = tid & 15;  // staging element of this thread

  for (int i = tid; i < nrow_pad * 16; i += kThreads) Wl[i] = Wt[i];
  for (int i = tid; i < nct * 256; i += kThreads) El[i] = Et[i];

  for (int j0 = 0; j0 < Mi; j0 += JT) {
    const int jt = min(JT, Mi - j0);
    for (int p0 = 0; p0 < P; p0 += PT) {
      const int pt = min(PT, P - p0);
      const int sweep = jt * pt;
      const int ntile = jt + pt;
      __syncthreads();
      if (tid < ntile)
        tptr[tid] = tid < jt ? bank + (size_t)dir_idx[d0 + j0 + tid] * plane
                             : spectra + ((size_t)pb * P + p0 + (tid - jt)) * plane;
      __syncthreads();
      // Pair slots beyond the sweep alias its last pair, so the MFMA loop has
      // no per-pair branches (branches force accumulator copies between
      // AGPRs and VGPRs); their results are never written.
      int toff[kPairsPerWave], poff[kPairsPerWave];
#pragma unroll
      for (int i = 0; i < kPairsPerWave; ++i) {
        const int k = min(wave + kWaves * i, sweep - 1);
        const int j = k / pt, p = k - j * pt;
        toff[i] = j * kTileElems;
        poff[i] = (jt + p) * kTileElems;
      }
      floatx4 outv[kPairsPerWave];
#pragma unroll
      for (int i = 0; i < kPairsPerWave; ++i) outv[i] = floatx4{0.f, 0.f, 0.f, 0.f};

      for (int ct = 0; ct < nct; ++ct) {
        const int c = ct * 16 + scol;
        float2 pre[kMaxTiles];
        // prologue: chunk 0 -> buffer 0
        {
          const bool ok = srow < n && c < m;
#pragma unroll
          for (int t = 0; t < kMaxTiles; ++t)
            pre[t] = (t < ntile && ok) ? tptr[t][(size_t)srow * m + c] : make_float2(0.f, 0.f);
#pragma unroll
          for (int t = 0; t < kMaxTiles; ++t)
            if (t < ntile) tiles[t * kTileElems + tid] = pre[t];
        }
        __syncthreads();
        floatx4 gr[kPairsPerWave], gi[kPairsPerWave];
#pragma unroll
        for (int i = 0; i < kPairsPerWave; ++i) {
          gr[i] = floatx4{0.f, 0.f, 0.f, 0.f};
          gi[i] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
        for (int rc = 0; rc < nrc; ++rc) {
          const int cur = rc & 1;
          const bool more = rc + 1 < nrc;
          if (more) {  // issue the next chunk's global loads; consumed after the MFMAs
            const int r = (rc + 1) * kRC + srow;
            const bool ok = r < n && c < m;
#pragma unroll
            for (int t = 0; t < kMaxTiles; ++t)
              pre[t] = (t < ntile && ok) ? tptr[t][(size_t)r * m + c] : make_float2(0.f, 0.f);
          }
          const float2* tb = tiles + cur * (kMaxTiles * kTileElems);
#pragma unroll 1
          for (int ks = 0; ks < kRC / 4; ++ks) {
            const int base = (ks * 4 + lr) * 16 + lc;
            const float2 wv = Wl[(rc * kRC) * 16 + base];
            const float nwi = -wv.y;
            float qr[kPairsPerWave], qi[kPairsPerWave];
#pragma unroll
            for (int i = 0; i < kPairsPerWave; ++i) {
              const float2 t = tb[toff[i] + base];
              const float2 pp = tb[poff[i] + base];
              qr[i] = t.x * pp.x - t.y * pp.y;
              qi[i] = t.x * pp.y + t.y * pp.x;
            }
#pragma unroll
            for (int i = 0; i < kPairsPerWave; ++i) gr[i] = mfma(qr[i], wv.x, gr[i]);
#pragma unroll
            for (int i = 0; i < kPairsPerWave; ++i) gi[i] = mfma(qr[i], wv.y, gi[i]);
#pragma unroll
            for (int i = 0; i < kPairsPerWave; ++i) gr[i] = mfma(qi[i], nwi, gr[i]);
#pragma unroll
            for (int i = 0; i < kPairsPerWave; ++i) gi[i] = mfma(qi[i], wv.x, gi[i]);
          }
          if (more) {
            float2* nb = tiles + (cur ^ 1) * (kMaxTiles * kTileElems);
#pragma unroll
            for (int t = 0; t < kMaxTiles; ++t)
              if (t < ntile) nb[t * kTileElems + tid] = pre[t];
          }
          __syncthreads();
        }
        // column-to-lag step: out^T += Re(E^T G^T) for this 16-column tile
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const float2 ev = El[(ct * 16 + 4 * lr + v) * 16 + lc];
          const float ner = ev.x, nei = -ev.y;
#pragma unroll
          for (int i = 0; i < kPairsPerWave; ++i) {
            outv[i] = mfma(ner, gr[i][v], outv[i]);
            outv[i] = mfma(nei, gi[i][v], outv[i]);
          }
        }
      }

      // per-pair max / argmax over the C x C lags (first maximum in y*C + x order)
#pragma unroll
      for (int i = 0; i < kPairsPerWave; ++i) {
        const int k = wave + kWaves * i;
        if (k >= sweep) continue;
        const int j = k / pt, p = k - j * pt;
        const float sc = 1024.0f * (inv_norm ? inv_norm[(size_t)pb * P + p0 + p] : 1.0f) *
                         (tinv ? tinv[dir_idx[d0 + j0 + j]] : 1.0f);
        float bv = -INFINITY;
        int bi = 0x7fffffff;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int x = x0 + 4 * lr + v, y = y0 + lc;
          if (x < C && y < C) {
            const float val = outv[i][v] * sc;
            const int idx = y * C + x;
            if (val > bv || (val == bv && idx < bi)) { bv = val; bi = idx; }
          }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
          const float ov = __shfl_xor(bv, o, 64);
          const int oi = __shfl_xor(bi, o, 64);
          if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
        }
        if (lane == 0) {
          const size_t g = (size_t)(d0 + j0 + j) * P + p0 + p;
          // a later lag window keeps the earlier windows' maximum unless it
          // is larger, or equal at a smaller crop index (first maximum)
          const bool keep_old =
              merge && (scores[g] > bv || (scores[g] == bv && peak_y[g] * C + peak_x[g] < bi));
          if (!keep_old) {
            scores[g] = bv;
            peak_y[g] = bi / C;
            peak_x[g] = bi % C;
          }
        }
      }
    }
  }

  // ---- per-particle tail -------------------------------------------------
  __syncthreads();
  if (tail && wave == 0) {
    const int np = Mi * P;
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
      if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
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
}

__global__ void score_sum_kernel(const float* __restrict__ peak, const int* __restrict__ best, int B,
                                 double* __restrict__ out) {
  __shared__ double red[kThreads / 64];
  double s = 0.0;
  for (int i = threadIdx.x; i < B; i += blockDim.x)
    if (best == nullptr || best[i] >= 0) s += exp((double)peak[i] / 1024.0);
  s = wave_sum_f64(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    out[0] = t;
  }
}

// One block per segment: out[s] = sum over i in [off[s], off[s+1]) with
// best_i >= 0 of exp(peak_i / 1024), float64, fixed reduction order.
__global__ void score_sum_segments_kernel(const float* __restrict__ peak, const int* __restrict__ best,
                                          const int* __restrict__ off, double* __restrict__ out) {
  __shared__ double red[kThreads / 64];
  const int sgm = blockIdx.x;
  const int i0 = off[sgm], i1 = off[sgm + 1];
  double s = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x)
    if (best == nullptr || best[i] >= 0) s += exp((double)peak[i] / 1024.0);
  s = wave_sum_f64(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    out[sgm] = t;
  }
}

inline int row_pad(int n) { return ((n + kRC - 1) / kRC) * kRC; }
inline int col_pad(int n) { return ((n / 2 + 1 + 15) / 16) * 16; }

size_t corr_lds_bytes(int n) {
  return (size_t)(row_pad(n) * 16 + col_pad(n) * 16 + 2 * kMaxTiles * kTileElems) * sizeof(float2);
}

}  // namespace

GIS_API size_t gis_correlate_workspace(int n) {
  if (n < 2) return 0;
  return (size_t)(row_pad(n) * 16 + col_pad(n) * 16) * sizeof(float2);
}

// Replaces correlate_templates_and_particle (v606:1418-1449) and the numeric
// part of make_result_records (v606:1466-1491), batched over particles.
//   bank      [Mtot][n][n/2+1] complex64, unit-norm templates (UNshifted)
//   spectra   [B][P][n][n/2+1] complex64 conjugated particle spectra (UNshifted)
//   inv_norm  [B][P] float32 per-spectrum 1/L2 norm, or NULL (already unit norm)
//   tinv      [Mtot] float32 per-template 1/L2 norm, or NULL (bank already unit norm)
//   dir_off   [B+1] int32 offsets into dir_idx; dir_idx global template indices
//   pmap      [B] int32 spectra row of each searched particle, or NULL (identity);
//             lets one launch search several candidates' banks against shared spectra
//   jt        template tiles per sweep (0 = automatic)
//   scores/peak_y/peak_x  [dir_off[B]][P] outputs (score = max * 1024 n^2)
//   best [B] flat argmax (local_dir * P + psi) or -1; peak [B]; mean/std [B] float64
GIS_API int gis_correlate(const void* bank, const void* spectra, const float* inv_norm, const float* tinv,
                          const int* dir_off, const int* dir_idx, const int* pmap, int B, int P, int n, int C,
                          int jt, float* scores,
                          int* peak_y, int* peak_x, int* best, float* peak, double* mean, double* stdv, void* ws,
                          size_t ws_bytes, void* stream) {
  GIS_REQUIRE(B >= 0 && P >= 1, "gis_correlate: B=%d P=%d", B, P);
  GIS_REQUIRE(n >= 4 && n % 2 == 0 && n <= 256, "gis_correlate: newbox %d must be even, in [4, 256]", n);
  if (C > 32 || C < 2 || C % 2 != 0 || C > n) {
    gis_set_error("gis_correlate: crop %d must be even, in [2, 32] and <= newbox", C);
    return GIS_EUNSUPPORTED;
  }
  if (B == 0) return GIS_OK;
  GIS_REQUIRE(bank && spectra && dir_off && dir_idx && scores && peak_y && peak_x && best && peak && mean &&
                  stdv && ws,
              "gis_correlate: null pointer");
  GIS_REQUIRE(ws_bytes >= gis_correlate_workspace(n), "gis_correlate: workspace too small");
  hipStream_t st = (hipStream_t)stream;
  float2* Wt = (float2*)ws;
  float2* Et = Wt + (size_t)row_pad(n) * 16;
  const int PT = std::min(P, 8);
  const int jt_max = kMaxTiles - PT;
  GIS_REQUIRE(jt >= 0 && jt <= jt_max, "gis_correlate: jt %d not in [0, %d]", jt, jt_max);
  const int JT = jt > 0 ? jt : std::max(1, std::min(jt_max, (kWaves * kPairsPerWave) / PT));
  const size_t lds = corr_lds_bytes(n);
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute((const void*)correlate_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)corr_lds_bytes(256));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      gis_set_error("gis_correlate: cannot reserve %zu bytes of LDS: %s", corr_lds_bytes(256), hipGetErrorString(e));
      return (int)e;
    }
    attr_set = true;
  }
  // one 16 x 16 lag window per launch: a crop above 16 takes four, each
  // merged into the stored pair maxima; the last one runs the particle tail
  const int nw = C > 16 ? 2 : 1;
  for (int q = 0; q < nw * nw; ++q) {
    const int y0 = 16 * (q / nw), x0 = 16 * (q % nw);
    hipLaunchKernelGGL(build_corr_tables, dim3(64), dim3(256), 0, st, n, C, y0, x0, row_pad(n), col_pad(n), Wt,
                       Et);
    GIS_LAUNCH_CHECK("gis_correlate tables");
    hipLaunchKernelGGL(correlate_kernel, dim3(B), dim3(kThreads), lds, st, (const float2*)bank,
                       (const float2*)spectra, inv_norm, tinv, dir_off, dir_idx, pmap, P, n, C, JT, PT, y0, x0,
                       q > 0 ? 1 : 0, q == nw * nw - 1 ? 1 : 0, Wt, Et, scores, peak_y, peak_x, best, peak, mean,
                       stdv);
    GIS_LAUNCH_CHECK("gis_correlate");
  }
  return GIS_OK;
}

// sum_i exp(peak_i / 1024) in float64 over particles with best_i >= 0
// (new_method_to_fit_the_2nd_Gaussian_PEAK_v4.py:133-134).
GIS_API int gis_score_sum(const float* peak, const int* best, int B, double* out, void* stream) {
  GIS_REQUIRE(B >= 0 && out, "gis_score_sum: bad arguments");
  GIS_REQUIRE(B == 0 || peak, "gis_score_sum: null peak");
  hipLaunchKernelGGL(score_sum_kernel, dim3(1), dim3(kThreads), 0, (hipStream_t)stream, peak, best, B, out);
  GIS_LAUNCH_CHECK("gis_score_sum");
  return GIS_OK;
}

// Per-candidate conformation scores in one launch: out[s] = sum over particles
// i in [off[s], off[s+1]) with best_i >= 0 of exp(peak_i / 1024), float64
// (new_method_to_fit_the_2nd_Gaussian_PEAK_v4.py:133-134 per candidate;
// ZL:gisapr_batch_search.py:361).  off [nseg+1] device int32.
GIS_API int gis_score_sum_segments(const float* peak, const int* best, const int* off, int nseg, double* out,
                                   void* stream) {
  GIS_REQUIRE(nseg >= 0, "gis_score_sum_segments: nseg %d", nseg);
  if (nseg == 0) return GIS_OK;
  GIS_REQUIRE(peak && off && out, "gis_score_sum_segments: null pointer");
  hipLaunchKernelGGL(score_sum_segments_kernel, dim3(nseg), dim3(kThreads), 0, (hipStream_t)stream, peak, best,
                     off, out);
  GIS_LAUNCH_CHECK("gis_score_sum_segments");
  return GIS_OK;
}
