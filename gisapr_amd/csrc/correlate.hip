// The v606 FP32 search (correlate_templates_and_particle, v606:1418-1449,
// and the numerics of make_result_records, v606:1452-1491) and the
// cross-particle score sums.
//
// The search evaluates the reference operator -- irfft2 of the pre-shifted
// product cropped to the C x C corner -- as the fixed linear functional of
// the unshifted product Q = T * P written out in split.hip, on FP32 MFMA
// (v_mfma_f32_16x16x4_f32) with the split search's work decomposition: one
// wave per (particle, directions, psi group), two chained products per
// 16-column tile (rows to lags, then columns to lags), the first product's
// accumulators reused register for register as the second's operand, the
// packed plane's lone last column of all of a wave's pairs in one tile,
// operand rows optionally zero padded to 128-byte lines.  C <= 32.
#include "common.h"

namespace {

constexpr int kThreads = 256;

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

}  // namespace

// Workspace bytes gis_correlate needs at newbox n, crop C, B particles.
GIS_API size_t gis_correlate_workspace(int n, int C, int B) { return gis_correlate_split_workspace(n, C, B); }

// Replaces correlate_templates_and_particle (v606:1418-1449) and the numeric
// part of make_result_records (v606:1466-1491), batched over particles.
//   bank      [Mtot][n][ldt] complex64 templates (UNshifted rfft2 layout,
//             rows ldt >= n/2+1 apart, columns past n/2 zero), or with
//             ldt = 0 [Mtot][gis_fragment_plane(n)] in fragment order
//   spectra   [B'][P][n][ldp] complex64 conjugated particle spectra
//             (UNshifted; ldp = 0: fragment order, as the bank)
//   inv_norm  [B'][P] float32 per-spectrum 1/L2 norm, or NULL (already unit norm)
//   tinv      [Mtot] float32 per-template 1/L2 norm, or NULL (bank already unit norm)
//   dir_off   [B+1] int32 offsets into dir_idx; dir_idx global template indices
//   pmap      [B] int32 spectra row of each searched particle, or NULL (identity);
//             lets one launch search several candidates' banks against shared spectra
//   order     [B] processing order of the particles (a permutation) or NULL
//   nrows     dir_off[B]
//   dt, pt    directions (1, 2) / psi values (1..5) per wave, 0 = automatic
//   scores/peak_y/peak_x  [dir_off[B]][P] outputs (score = max * 1024 n^2)
//   best [B] flat argmax (local_dir * P + psi) or -1; peak [B]; mean/std [B] float64
GIS_API int gis_correlate(const void* bank, const void* spectra, const float* inv_norm, const float* tinv,
                          const int* dir_off, const int* dir_idx, const int* pmap, const int* order, int B, int P,
                          int n, int C, int nrows, int ldt, int ldp, int dt, int pt, float* scores, int* peak_y,
                          int* peak_x, int* best, float* peak, double* mean, double* stdv, void* ws, size_t ws_bytes,
                          void* stream) {
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
  GIS_REQUIRE(dt == 0 || dt == 1 || dt == 2, "gis_correlate: dt %d not in {0, 1, 2}", dt);
  return gis_correlate_units(bank, spectra, inv_norm, tinv, dir_off, dir_idx, pmap, order, B, P, n, C, nrows, ldt,
                             ldp, dt, pt, scores, peak_y, peak_x, best, peak, mean, stdv, ws, ws_bytes,
                             (hipStream_t)stream);
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
