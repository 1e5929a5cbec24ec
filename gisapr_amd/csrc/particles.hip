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
// gfx950 mapping: one 256-thread workgroup per particle (persistent grid).
// Row pass: for each psi, two rotated rows per complex LDS FFT, written to
// the workgroup's scratch slot.  Column pass: per 8-column tile the CTF is
// evaluated once into LDS and applied to all psi images of the particle.
// Tile widths are compile-time powers of two (no runtime divisions).
#include "common.h"

#pragma clang fp contract(off)

namespace {

constexpr int kThreads = 256;
constexpr int kMaxPsiGroup = 8;  // psi images per pass (scratch slots per workgroup)
constexpr int kTK = 8;           // columns per column tile
constexpr int kRB = 8;           // complex rows (= 2 real rows) per row batch

// Scratch layout of one row-transformed image: tile-major [NT][N][kTK] with
// NT = ceil((N/2+1)/kTK), so a column tile is one contiguous 16 KiB block.
template <int N>
struct PrepLayout {
  static constexpr int MH = N / 2 + 1;
  static constexpr int SP = N + 1;                 // padded column stride in LDS
  static constexpr int NT = (MH + kTK - 1) / kTK;  // column tiles
  static constexpr size_t IMG = (size_t)NT * N * kTK;            // float2 per image
  static constexpr size_t SLOT = (size_t)kMaxPsiGroup * IMG;     // float2 per workgroup
};

typedef float f2a4 __attribute__((ext_vector_type(2), aligned(4)));

__device__ __forceinline__ float ctf_weight(float x, float y, float u4, float nc, float axx, float axy, float ayy,
                                            float k1, float k2, float k3, float k5, float kk) {
  const float gamma = k1 * ((axx * x * x + 2.0f * axy * x * y) + ayy * y * y) + k2 * u4 - k5 - k3;
  float cv = -sinf(gamma);
  if (fabsf(cv) < 1e-8f) cv = cv > 0.f ? 1e-8f : (cv < 0.f ? -1e-8f : 0.f);
  return cv * sqrtf(1.0f / (nc + kk * cv * cv));
}

// ctf_tab with defocus:  [4][n][n/2+1] = x, y, u4, ncurve (PackedCTFTorch.__init__)
// ctf_tab without:       [B][n][n/2+1] explicit packed CTF, or NULL (ones)
// rotp:                  [B][P][6] = c, -s, s, c, off_y, off_x of the psi sampling grid
template <int N>
__global__ void __launch_bounds__(kThreads)
prepare_particles_kernel(const float* __restrict__ images, int B, int P, int n,
                         const float* __restrict__ rotp, const float* __restrict__ mask,
                         const float* __restrict__ ctf_tab, const float* __restrict__ defocus,
                         float k1, float k2, float k3, float k5, float kk,
                         float2* __restrict__ spectra, float* __restrict__ inv_norm,
                         float2* __restrict__ scratch) {
  using T = PrepLayout<N>;
  constexpr int MH = T::MH, SP = T::SP;
  __shared__ float2 tw[N];
  __shared__ float2 bufA[kTK * SP > kRB * N ? kTK * SP : kRB * N];
  __shared__ float2 bufB[kTK * SP > kRB * N ? kTK * SP : kRB * N];
  __shared__ float ctf_l[256 * kTK];  // [row][q], rows < n <= N <= 256
  __shared__ double red[kMaxPsiGroup * (kThreads / 64)];
  const int tid = threadIdx.x;
  constexpr int nthr = kThreads;  // launched with exactly kThreads threads
  fill_twiddles(tw, N, tid, nthr);
  float2* Wbase = scratch + (size_t)blockIdx.x * T::SLOT;
  const int mh_n = n / 2 + 1;
  const size_t plane_n = (size_t)n * mh_n;
  const bool has_ctf = defocus != nullptr;
  const bool explicit_ctf = !has_ctf && ctf_tab != nullptr;

  for (int b = blockIdx.x; b < B; b += gridDim.x) {
    const float* src = images + (size_t)b * N * N;
    float axx = 0.f, ayy = 0.f, axy = 0.f;
    if (has_ctf) {
      const float dfu = defocus[3 * b + 0], dfv = defocus[3 * b + 1];
      const float ang = defocus[3 * b + 2] * 0.017453292519943295f;
      const float sa = sinf(ang), ca = cosf(ang);
      axx = (-dfu * ca) * ca - (dfv * sa) * sa;
      ayy = (-dfu * sa) * sa - (dfv * ca) * ca;
      axy = (-dfu * sa) * ca + (dfv * sa) * ca;
    }
    for (int pg = 0; pg < P; pg += kMaxPsiGroup) {
      const int np = min(kMaxPsiGroup, P - pg);

      // ---- rows: bilinear rotation (zeros padding), mask, forward FFT ---------
      for (int pi = 0; pi < np; ++pi) {
        const float* rp = rotp + ((size_t)b * P + pg + pi) * 6;
        const float m00 = rp[0], m01 = rp[1], m10 = rp[2], m11 = rp[3], oy = rp[4], ox = rp[5];
        float2* W = Wbase + (size_t)pi * T::IMG;
        for (int y0 = 0; y0 < N; y0 += 2 * kRB) {
          __syncthreads();
          // Each wave samples 8x8 output blocks, so one load instruction's
          // addresses stay within a compact rotated source patch; the two
          // horizontal taps of a row come from one 8-byte load.
          // Branch-free: the 2x2 footprint is read at clamped coordinates (always
          // in bounds) and taps outside the image are zeroed by selects, so
          // all ITER iterations' loads can be in flight together.
          constexpr int BPR = N / 8;
          constexpr int ITER = 2 * BPR / (kThreads / 64);
          float* bufA_f = (float*)bufA;
          const int lane = tid & 63;
#pragma unroll
          for (int it = 0; it < ITER; ++it) {
            const int blk = (tid >> 6) + it * (kThreads / 64);
            const int band = blk / BPR, bc = blk - band * BPR;
            const int yl = band * 8 + (lane >> 3);
            const int Y = y0 + yl, X = bc * 8 + (lane & 7);
            const float sy = (float)Y * m00 + (float)X * m01 + oy;
            const float sx = (float)Y * m10 + (float)X * m11 + ox;
            const float fy0 = floorf(sy), fx0 = floorf(sx);
            const float ty = sy - fy0, tx = sx - fx0;
            // clamp before converting so far-outside samples stay finite ints
            const int iy = (int)fminf(fmaxf(fy0, -2.0f), (float)N);
            const int ix = (int)fminf(fmaxf(fx0, -2.0f), (float)N);
            const int cy = min(max(iy, 0), N - 2), cx = min(max(ix, 0), N - 2);
            const float* s0 = src + (size_t)cy * N + cx;
            const f2a4 r0 = *(const f2a4*)s0;
            const f2a4 r1 = *(const f2a4*)(s0 + N);
            const bool y0ok = iy >= 0 && iy < N, y1ok = iy + 1 >= 0 && iy + 1 < N;
            const bool x0ok = ix >= 0 && ix < N, x1ok = ix + 1 >= 0 && ix + 1 < N;
            // offsets of each valid tap inside the loaded 2x2 block are 0 or 1
            const int dy0 = iy - cy, dy1 = iy + 1 - cy, dx0 = ix - cx, dx1 = ix + 1 - cx;
            auto pick = [&](int dy, int dx) { return dy ? (dx ? r1.y : r1.x) : (dx ? r0.y : r0.x); };
            const float a00 = (y0ok && x0ok) ? pick(dy0, dx0) : 0.f;
            const float a01 = (y0ok && x1ok) ? pick(dy0, dx1) : 0.f;
            const float a10 = (y1ok && x0ok) ? pick(dy1, dx0) : 0.f;
            const float a11 = (y1ok && x1ok) ? pick(dy1, dx1) : 0.f;
            float v = ((1.0f - ty) * (1.0f - tx)) * a00 + ((1.0f - ty) * tx) * a01 +
                      (ty * (1.0f - tx)) * a10 + (ty * tx) * a11;
            if (mask) v *= mask[(size_t)Y * N + X];
            bufA_f[2 * ((yl >> 1) * N + X) + (yl & 1)] = v;  // row pair -> (re, im)
          }
          __syncthreads();
          float2* res = fft_lds<N, false>(bufA, bufB, kRB, N, tw, tid, nthr);
          for (int t = tid; t < kRB * MH; t += nthr) {
            const int q = t / MH, k = t - q * MH;
            const float2 z = res[q * N + k];
            const float2 zc = cconj(res[q * N + (N - k) % N]);
            const int ya = y0 + 2 * q;
            float2* Wk = W + ((size_t)(k / kTK) * N) * kTK + (k % kTK);
            Wk[(size_t)ya * kTK] = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y + zc.y));
            Wk[(size_t)(ya + 1) * kTK] = make_float2(0.5f * (z.y - zc.y), -0.5f * (z.x - zc.x));
          }
        }
      }

      // ---- columns: CTF tile once, then FFT/crop/pack/CTF/conj per psi ------
      double nrm[kMaxPsiGroup];
#pragma unroll
      for (int i = 0; i < kMaxPsiGroup; ++i) nrm[i] = 0.0;
      const int kmax = n / 2;
      for (int k0 = 0; k0 <= kmax; k0 += kTK) {
        __syncthreads();
        for (int t = tid; t < n * kTK; t += nthr) {
          const int r = t / kTK, q = t - r * kTK;
          const int c = k0 + q;
          float w = 1.0f;
          if (c <= kmax) {
            const size_t e = (size_t)r * mh_n + c;
            if (has_ctf)
              w = ctf_weight(ctf_tab[e], ctf_tab[plane_n + e], ctf_tab[2 * plane_n + e], ctf_tab[3 * plane_n + e],
                             axx, axy, ayy, k1, k2, k3, k5, kk);
            else if (explicit_ctf)
              w = ctf_tab[(size_t)b * plane_n + e];
          }
          ctf_l[t] = w;
        }
#pragma unroll 1
        for (int pi = 0; pi < np; ++pi) {
          const float2* Wt = Wbase + (size_t)pi * T::IMG + (size_t)(k0 / kTK) * N * kTK;
          for (int t = tid; t < kTK * N; t += nthr) {  // contiguous tile block
            const int Y = t / kTK, q = t - Y * kTK;
            const int c = k0 + q;
            bufA[q * SP + Y] = c < MH ? Wt[t] : make_float2(0.f, 0.f);
          }
          __syncthreads();
          float2* res = fft_lds<N, false>(bufA, bufB, kTK, SP, tw, tid, nthr);
          float2* out = spectra + ((size_t)b * P + pg + pi) * plane_n;
          double acc = 0.0;
          for (int t = tid; t < n * kTK; t += nthr) {
            const int r = t / kTK, q = t - r * kTK;
            const int c = k0 + q;
            if (c > kmax) continue;
            const int fy = r < n / 2 ? r : r - n;
            float2 v;
            if (c < n / 2 || n == N) {
              v = res[q * SP + (fy < 0 ? fy + N : fy)];
            } else {  // packed column n/2 holds frequency -n/2 = conj of (+n/2) at -fy
              const int my = -fy;
              v = cconj(res[q * SP + (my < 0 ? my + N : my)]);
            }
            const float w = ctf_l[t];
            v.x *= w;
            v.y = -(v.y * w);  // conj
            out[(size_t)r * mh_n + c] = v;
            acc += (double)v.x * v.x + (double)v.y * v.y;
          }
#pragma unroll
          for (int i = 0; i < kMaxPsiGroup; ++i)
            if (i == pi) nrm[i] += acc;
          __syncthreads();
        }
      }
      block_sum_f64<kMaxPsiGroup>(nrm, red);
      if (tid < np) {
        double v = 0.0;
#pragma unroll
        for (int i = 0; i < kMaxPsiGroup; ++i)
          if (i == tid) v = nrm[i];
        inv_norm[(size_t)b * P + pg + tid] = 1.0f / (float)sqrt(v);
      }
    }
  }
}

template <int N>
int launch_prepare(const float* images, int B, int P, int n, const float* rotp, const float* mask,
                   const float* ctf_tab, const float* defocus, const float* k, float2* spectra,
                   float* inv_norm, void* ws, size_t ws_bytes, hipStream_t st) {
  using T = PrepLayout<N>;
  const size_t slot = T::SLOT * sizeof(float2);
  int grid = (int)std::min<size_t>((size_t)B, ws_bytes / slot);
  grid = std::min(grid, 4 * gis_cu_count());
  GIS_REQUIRE(grid >= 1, "gis_prepare_particles: workspace of %zu bytes is below one slot (%zu)", ws_bytes, slot);
  hipLaunchKernelGGL(prepare_particles_kernel<N>, dim3(grid), dim3(kThreads), 0, st, images, B, P, n, rotp,
                     mask, ctf_tab, defocus, k[0], k[1], k[2], k[3], k[4], spectra, inv_norm, (float2*)ws);
  GIS_LAUNCH_CHECK("gis_prepare_particles");
  return GIS_OK;
}

template <int N>
size_t prep_slot_bytes() { return PrepLayout<N>::SLOT * sizeof(float2); }

__global__ void __launch_bounds__(256)
ctf_packed_kernel(const float* __restrict__ tab, const float* __restrict__ defocus, int B, int n, float k1,
                  float k2, float k3, float k5, float kk, float* __restrict__ out) {
  const size_t plane = (size_t)n * (n / 2 + 1);
  const size_t total = plane * (size_t)B;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const int b = (int)(t / plane);
    const size_t e = t - (size_t)b * plane;
    const float dfu = defocus[3 * b + 0], dfv = defocus[3 * b + 1];
    const float ang = defocus[3 * b + 2] * 0.017453292519943295f;
    const float sa = sinf(ang), ca = cosf(ang);
    const float axx = (-dfu * ca) * ca - (dfv * sa) * sa;
    const float ayy = (-dfu * sa) * sa - (dfv * ca) * ca;
    const float axy = (-dfu * sa) * ca + (dfv * sa) * ca;
    out[t] = ctf_weight(tab[e], tab[plane + e], tab[2 * plane + e], tab[3 * plane + e], axx, axy, ayy, k1, k2, k3,
                        k5, kk);
  }
}

}  // namespace

// Packed CTF of B particles (PackedCTFTorch.batch, v606:778-829).
//   ctf_tab [4][n][n/2+1] float32 (x, y, u4, ncurve); defocus [B][3] (dfU, dfV, angle deg)
//   kconst  [5] float32 (host): k1, k2, k3, k5, kk;  out [B][n][n/2+1] float32
GIS_API int gis_ctf_packed(const float* ctf_tab, const float* defocus, int B, int n, const float* kconst, float* out,
                           void* stream) {
  GIS_REQUIRE(B >= 0 && n >= 2 && n % 2 == 0, "gis_ctf_packed: B=%d n=%d", B, n);
  if (B == 0) return GIS_OK;
  GIS_REQUIRE(ctf_tab && defocus && kconst && out, "gis_ctf_packed: null pointer");
  const size_t total = (size_t)B * n * (n / 2 + 1);
  const int blocks = (int)std::min<size_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(ctf_packed_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, ctf_tab, defocus, B, n,
                     kconst[0], kconst[1], kconst[2], kconst[3], kconst[4], out);
  GIS_LAUNCH_CHECK("gis_ctf_packed");
  return GIS_OK;
}

// Workspace bytes for gis_prepare_particles over `particles` particles.
GIS_API size_t gis_prepare_particles_workspace(int N, int particles) {
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
  if (particles <= 0) return 0;
  return std::min<size_t>((size_t)particles, (size_t)4 * gis_cu_count()) * s;
}

// Replaces ParticleFourierPreparer.prepare (v606:1340-1374) together with the
// packed CTF of PackedCTFTorch.batch (v606:778-829), batched over particles.
//   images   [B][N][N] float32
//   rotp     [B][P][6] float32 psi grids: (c, -s, s, c, off_y, off_x)
//   mask     [N][N] float32 or NULL
//   ctf_tab  with defocus: [4][n][n/2+1] float32 (x, y, u4, ncurve);
//            without: [B][n][n/2+1] explicit packed CTF, or NULL (CTF = 1)
//   defocus  [B][3] float32 (dfU, dfV, angle deg) or NULL
//   kconst   [5] float32 (host): k1, k2, k3, k5, kk
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
