// Particle preparation: in-plane rotation, mask, FFT, pack, CTF, conj, norm.
//
// For every (particle b, psi p) this computes the packed particle spectrum of
// ParticleFourierPreparer.prepare (v606:1340-1374):
//   rotated = bilinear sample of the image at M (y, x) + c - M c
//             (grid_sample, zeros padding, align_corners; the psi grids of
//             PsiGridBank v606:1193-1256)
//   x mask -> fft2 -> centre -> crop to n -> pack (rfft2 layout)
//   x CTF (PackedCTFTorch.batch, v606:778-829, evaluated in the kernel from
//          the static per-layout tables and the particle's defocus)
//   -> conj
// and its L2 norm.  The spectrum is stored UNnormalised and UNshifted
// ([B][P][n][n/2+1]); inv_norm[b][p] = 1 / ||spectrum|| is applied by the
// correlation kernel (the reference divides before correlating).
//
// gfx950 mapping: one 256-thread workgroup per (particle, psi) image on a
// persistent grid; the row pass samples two rotated rows per complex FFT;
// the column pass runs only over the n/2+1 columns that survive the crop.
#include "common.h"

#pragma clang fp contract(off)

namespace {

constexpr int kThreads = 256;
constexpr int kLdsFloat2 = 8192;

template <int N>
struct PrepLayout {
  static constexpr int MH = N / 2 + 1;
  static constexpr int FIXED = N + 16;
  static constexpr int CAP = (kLdsFloat2 - FIXED) / 2;
  static constexpr int RB = CAP / N;          // complex rows per batch (2 real rows each)
  static constexpr int TK = CAP / (N + 1);    // columns per tile
  static constexpr size_t SLOT = (size_t)N * MH;  // float2
};

// ctf_tab: [4][n][n/2+1] = x, y, u4, ncurve (float32, PackedCTFTorch.__init__)
// defocus: [B][3] = dfU, dfV, angle (deg) or NULL (no CTF columns -> all ones)
// rotp:    [B][P][6] = c, -s, s, c, off_y, off_x of the psi sampling grid
template <int N>
__global__ void __launch_bounds__(kThreads)
prepare_particles_kernel(const float* __restrict__ images, int B, int P, int n,
                         const float* __restrict__ rotp, const float* __restrict__ mask,
                         const float* __restrict__ ctf_tab, const float* __restrict__ defocus,
                         float k1, float k2, float k3, float k5, float kk,
                         float2* __restrict__ spectra, float* __restrict__ inv_norm,
                         float2* __restrict__ scratch) {
  using T = PrepLayout<N>;
  constexpr int MH = T::MH;
  __shared__ float2 lds[kLdsFloat2];
  float2* tw = lds;
  double* red = (double*)(tw + N);
  float2* bufA = lds + T::FIXED;
  float2* bufB = bufA + T::CAP;
  const int tid = threadIdx.x, nthr = blockDim.x;
  fill_twiddles(tw, N, tid, nthr);
  float2* W = scratch + (size_t)blockIdx.x * T::SLOT;
  const int mh_n = n / 2 + 1;
  const size_t plane_n = (size_t)n * mh_n;

  for (int img = blockIdx.x; img < B * P; img += gridDim.x) {
    const int b = img / P;
    const float* src = images + (size_t)b * N * N;
    const float* rp = rotp + (size_t)img * 6;
    const float m00 = rp[0], m01 = rp[1], m10 = rp[2], m11 = rp[3], oy = rp[4], ox = rp[5];
    __syncthreads();

    // ---- rows: bilinear rotation (zeros padding), mask, forward FFT -------
    for (int y0 = 0; y0 < N; y0 += 2 * T::RB) {
      const int nseq = min(T::RB, (N - y0 + 1) / 2);
      for (int t = tid; t < nseq * N; t += nthr) {
        const int q = t / N, X = t - q * N;
        float v2[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int Y = y0 + 2 * q + h;
          float v = 0.0f;
          if (Y < N) {
            const float sy = (float)Y * m00 + (float)X * m01 + oy;
            const float sx = (float)Y * m10 + (float)X * m11 + ox;
            const float fy0 = floorf(sy), fx0 = floorf(sx);
            const int iy = (int)fy0, ix = (int)fx0;
            const float ty = sy - fy0, tx = sx - fx0;
            float a00 = 0.f, a01 = 0.f, a10 = 0.f, a11 = 0.f;
            const bool y0ok = iy >= 0 && iy < N, y1ok = iy + 1 >= 0 && iy + 1 < N;
            const bool x0ok = ix >= 0 && ix < N, x1ok = ix + 1 >= 0 && ix + 1 < N;
            if (y0ok && x0ok) a00 = src[(size_t)iy * N + ix];
            if (y0ok && x1ok) a01 = src[(size_t)iy * N + ix + 1];
            if (y1ok && x0ok) a10 = src[(size_t)(iy + 1) * N + ix];
            if (y1ok && x1ok) a11 = src[(size_t)(iy + 1) * N + ix + 1];
            v = ((1.0f - ty) * (1.0f - tx)) * a00 + ((1.0f - ty) * tx) * a01 +
                (ty * (1.0f - tx)) * a10 + (ty * tx) * a11;
            if (mask) v *= mask[(size_t)Y * N + X];
          }
          v2[h] = v;
        }
        bufA[q * N + X] = make_float2(v2[0], v2[1]);
      }
      __syncthreads();
      float2* res = fft_lds<N, false>(bufA, bufB, nseq, N, tw, tid, nthr);
      for (int t = tid; t < nseq * MH; t += nthr) {
        const int q = t / MH, k = t - q * MH;
        const float2 z = res[q * N + k];
        const float2 zc = cconj(res[q * N + (N - k) % N]);
        const int ya = y0 + 2 * q;
        W[(size_t)ya * MH + k] = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y + zc.y));
        if (ya + 1 < N) W[(size_t)(ya + 1) * MH + k] = make_float2(0.5f * (z.y - zc.y), -0.5f * (z.x - zc.x));
      }
      __syncthreads();
    }

    // ---- columns: forward FFT, crop, pack, CTF, conj ----------------------
    float2* out = spectra + (size_t)img * plane_n;
    float sin_az = 0.f, cos_az = 1.f, dfu = 0.f, dfv = 0.f;
    const bool has_ctf = defocus != nullptr;
    if (has_ctf) {
      dfu = defocus[3 * b + 0];
      dfv = defocus[3 * b + 1];
      const float ang = defocus[3 * b + 2] * 0.017453292519943295f;
      sin_az = sinf(ang);
      cos_az = cosf(ang);
    }
    const float axx = (-dfu * cos_az) * cos_az - (dfv * sin_az) * sin_az;
    const float ayy = (-dfu * sin_az) * sin_az - (dfv * cos_az) * cos_az;
    const float axy = (-dfu * sin_az) * cos_az + (dfv * sin_az) * cos_az;
    double nrm[1] = {0.0};
    const int kmax = n / 2;
    for (int k0 = 0; k0 <= kmax; k0 += T::TK) {
      const int nk = min(T::TK, kmax + 1 - k0);
      for (int t = tid; t < nk * N; t += nthr) {
        const int Y = t / nk, q = t - Y * nk;
        bufA[q * (N + 1) + Y] = W[(size_t)Y * MH + k0 + q];
      }
      __syncthreads();
      float2* res = fft_lds<N, false>(bufA, bufB, nk, N + 1, tw, tid, nthr);
      for (int t = tid; t < nk * n; t += nthr) {
        const int r = t / nk, q = t - r * nk;
        const int c = k0 + q;
        const int fy = r < n / 2 ? r : r - n;
        float2 v;
        if (c < n / 2 || n == N) {
          v = res[q * (N + 1) + (fy < 0 ? fy + N : fy)];
        } else {
          const int my = -fy;
          v = cconj(res[q * (N + 1) + (my < 0 ? my + N : my)]);
        }
        if (has_ctf) {
          const size_t e = (size_t)r * mh_n + c;
          const float x = ctf_tab[e], y = ctf_tab[plane_n + e];
          const float u4 = ctf_tab[2 * plane_n + e], nc = ctf_tab[3 * plane_n + e];
          const float gamma = k1 * ((axx * x * x + 2.0f * axy * x * y) + ayy * y * y) + k2 * u4 - k5 - k3;
          float cv = -sinf(gamma);
          if (fabsf(cv) < 1e-8f) cv = cv > 0.f ? 1e-8f : (cv < 0.f ? -1e-8f : 0.f);
          const float w = cv * sqrtf(1.0f / (nc + kk * cv * cv));
          v.x *= w;
          v.y *= w;
        } else if (ctf_tab) {  // explicit per-particle packed CTF [B][n][n/2+1]
          const float w = ctf_tab[(size_t)b * plane_n + (size_t)r * mh_n + c];
          v.x *= w;
          v.y *= w;
        }
        v.y = -v.y;  // conj
        out[(size_t)r * mh_n + c] = v;
        nrm[0] += (double)v.x * v.x + (double)v.y * v.y;
      }
      __syncthreads();
    }
    block_sum_f64<1>(nrm, red);
    if (tid == 0) inv_norm[img] = 1.0f / (float)sqrt(nrm[0]);
  }
}

template <int N>
int launch_prepare(const float* images, int B, int P, int n, const float* rotp, const float* mask,
                   const float* ctf_tab, const float* defocus, const float* k, float2* spectra,
                   float* inv_norm, void* ws, size_t ws_bytes, hipStream_t st) {
  using T = PrepLayout<N>;
  const size_t slot = T::SLOT * sizeof(float2);
  int grid = (int)std::min<size_t>((size_t)B * P, ws_bytes / slot);
  grid = std::min(grid, 8 * gis_cu_count());
  GIS_REQUIRE(grid >= 1, "gis_prepare_particles: workspace of %zu bytes is below one slot (%zu)", ws_bytes, slot);
  hipLaunchKernelGGL(prepare_particles_kernel<N>, dim3(grid), dim3(kThreads), 0, st, images, B, P, n, rotp,
                     mask, ctf_tab, defocus, k[0], k[1], k[2], k[3], k[4], spectra, inv_norm, (float2*)ws);
  GIS_LAUNCH_CHECK("gis_prepare_particles");
  return GIS_OK;
}

template <int N>
size_t prep_slot_bytes() { return PrepLayout<N>::SLOT * sizeof(float2); }

}  // namespace

GIS_API size_t gis_prepare_particles_workspace(int N, int images) {
  size_t s = 0;
  switch (N) {
    case 32: s = prep_slot_bytes<32>(); break;
    case 64: s = prep_slot_bytes<64>(); break;
    case 96: s = prep_slot_bytes<96>(); break;
    case 128: s = prep_slot_bytes<128>(); break;
    case 192: s = prep_slot_bytes<192>(); break;
    case 256: s = prep_slot_bytes<256>(); break;
    default: return 0;
  }
  if (images <= 0) return 0;
  return std::min<size_t>((size_t)images, (size_t)8 * gis_cu_count()) * s;
}

// Replaces ParticleFourierPreparer.prepare (v606:1340-1374) together with the
// packed CTF of PackedCTFTorch.batch (v606:778-829), batched over particles.
//   images   [B][N][N] float32
//   rotp     [B][P][6] float32 psi grids: (c, -s, s, c, off_y, off_x)
//   mask     [N][N] float32 or NULL
//   ctf_tab  [4][n][n/2+1] float32 (x, y, u4, ncurve) or NULL with defocus NULL
//   defocus  [B][3] float32 (dfU, dfV, angle deg) or NULL (CTF = 1)
//   kconst   [5] float32: k1, k2, k3, k5, kk
//   spectra  [B][P][n][n/2+1] complex64 output (conjugated, unnormalised, unshifted)
//   inv_norm [B][P] float32 output
GIS_API int gis_prepare_particles(const float* images, int B, int N, int n, int P, const float* rotp,
                                  const float* mask, const float* ctf_tab, const float* defocus,
                                  const float* kconst, void* spectra, float* inv_norm, void* ws,
                                  size_t ws_bytes, void* stream) {
  GIS_REQUIRE(B >= 0 && P >= 1, "gis_prepare_particles: B=%d P=%d", B, P);
  GIS_REQUIRE(n >= 2 && n <= N && n % 2 == 0, "gis_prepare_particles: newbox %d must be even and <= box %d", n, N);
  if (B == 0) return GIS_OK;
  GIS_REQUIRE(images && rotp && spectra && inv_norm && ws && kconst, "gis_prepare_particles: null pointer");
  GIS_REQUIRE(defocus == nullptr || ctf_tab != nullptr, "gis_prepare_particles: defocus without CTF tables");
  hipStream_t st = (hipStream_t)stream;
  float2* S = (float2*)spectra;
  switch (N) {
    case 32: return launch_prepare<32>(images, B, P, n, rotp, mask, ctf_tab, defocus, kconst, S, inv_norm, ws, ws_bytes, st);
    case 64: return launch_prepare<64>(images, B, P, n, rotp, mask, ctf_tab, defocus, kconst, S, inv_norm, ws, ws_bytes, st);
    case 96: return launch_prepare<96>(images, B, P, n, rotp, mask, ctf_tab, defocus, kconst, S, inv_norm, ws, ws_bytes, st);
    case 128: return launch_prepare<128>(images, B, P, n, rotp, mask, ctf_tab, defocus, kconst, S, inv_norm, ws, ws_bytes, st);
    case 192: return launch_prepare<192>(images, B, P, n, rotp, mask, ctf_tab, defocus, kconst, S, inv_norm, ws, ws_bytes, st);
    case 256: return launch_prepare<256>(images, B, P, n, rotp, mask, ctf_tab, defocus, kconst, S, inv_norm, ws, ws_bytes, st);
    default:
      gis_set_error("gis_prepare_particles: box %d not in {32, 64, 96, 128, 192, 256}", N);
      return GIS_EUNSUPPORTED;
  }
}
