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
// gfx950 mapping: two kernels per chunk of particles, the row-transformed
// images staged in a workspace small enough to stay in the last-level cache
// (a chunk of at most 512 images):
//   rows    one 256-thread block per (image, 16-row band): bilinear
//           rotation of the band (8x8 blocked lanes, all loads in flight
//           together), mask, two real rows per complex FFT (one wave per two
//           sequences, in place in LDS), the half-spectrum split, written
//           column-tile-major ([tile][row][8], one contiguous 16 KiB block
//           per tile) so the column pass reads whole tiles;
//   columns one 256-thread block per (particle, 8-column tile): the tile's
//           CTF is evaluated once into LDS, then for each psi image the tile
//           is loaded (the next image's tile is in flight during the
//           current one's FFT), transformed along the rows, cropped,
//           packed, multiplied by the CTF, conjugated and written; the
//           block's sum of |v|^2 per image goes to a partial slot;
//   norm    inv_norm = 1 / sqrt(sum of the partials in tile order)
//           (deterministic).
// Many small independent blocks keep every CU busy where one block per
// particle waited on its own chain of barriers and scratch round trips.
#include "common.h"

#pragma clang fp contract(off)

namespace {

constexpr int kThreads = 256;
constexpr int kTK = 8;            // columns per column tile
constexpr int kRows = 16;         // real rows per rows-kernel block (8 complex sequences)
constexpr int kPad = 3;           // one LDS pad slot per 8 elements
constexpr int kMaxChunkImages = 512;

// Stage layout of one row-transformed image: tile-major [NT][N][kTK],
// NT = ceil((N/2+1)/kTK), a column tile one contiguous block.  In LDS a
// block's sequences start SL float2 apart, SL = 8 mod 32, so the eight
// sequences' words of one row fall on different banks.
template <int N>
struct PrepLayout {
  static constexpr int MH = N / 2 + 1;
  static constexpr int NT = (MH + kTK - 1) / kTK;               // column tiles
  static constexpr size_t IMG = (size_t)NT * N * kTK;           // float2 per image
  static constexpr int PL = lds_padded_len<N, kPad>();          // padded sequence length
  static constexpr int SL = PL + ((8 - PL % 32) + 32) % 32;     // sequence stride
};

// One stage of fft_wave_inplace over two sequences of one wave at once:
// butterfly j2 < 2 L/R of the stage belongs to sequence j2 / (L/R), so a
// stage with L/R = 32 butterflies keeps all 64 lanes busy.  Each butterfly
// is computed exactly as in stockham_stage_inplace (same results).
template <int L, int R, bool INV, int PAD>
__device__ __forceinline__ void stage_inplace2(float2* s0, float2* s1, int Ns, const float2* __restrict__ tw,
                                               int lane) {
  constexpr int LR = L / R;
  constexpr int NB = (2 * LR + 63) / 64;
  const int step = L / (Ns * R);
  float2 v[NB][R];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j2 = lane + 64 * b;
    if ((2 * LR) % 64 == 0 || j2 < 2 * LR) {
      const int sq = j2 >= LR ? 1 : 0;
      const float2* s = sq ? s1 : s0;
      const int j = j2 - sq * LR;
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = s[lds_pad<PAD>(j + r * LR)];
    }
  }
  wave_lds_sync();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j2 = lane + 64 * b;
    if ((2 * LR) % 64 == 0 || j2 < 2 * LR) {
      const int sq = j2 >= LR ? 1 : 0;
      float2* s = sq ? s1 : s0;
      const int j = j2 - sq * LR;
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

template <int L, bool INV, int PAD>
__device__ __forceinline__ void fft_wave_inplace2(float2* s0, float2* s1, const float2* __restrict__ tw, int lane) {
  using S = FftShape<L>;
  int Ns = 1;
#pragma unroll
  for (int st = 0; st < S::n8; ++st) {
    stage_inplace2<L, 8, INV, PAD>(s0, s1, Ns, tw, lane);
    Ns *= 8;
  }
  if constexpr (S::n4) {
    stage_inplace2<L, 4, INV, PAD>(s0, s1, Ns, tw, lane);
    Ns *= 4;
  }
  if constexpr (S::n2) {
    stage_inplace2<L, 2, INV, PAD>(s0, s1, Ns, tw, lane);
    Ns *= 2;
  }
  if constexpr (S::three) stage_inplace2<L, 3, INV, PAD>(s0, s1, Ns, tw, lane);
}

typedef float f2a4 __attribute__((ext_vector_type(2), aligned(4)));

__device__ __forceinline__ float ctf_weight(float x, float y, float u4, float nc, float axx, float axy, float ayy,
                                            float k1, float k2, float k3, float k5, float kk) {
  const float gamma = k1 * ((axx * x * x + 2.0f * axy * x * y) + ayy * y * y) + k2 * u4 - k5 - k3;
  float cv = -sinf(gamma);
  if (fabsf(cv) < 1e-8f) cv = cv > 0.f ? 1e-8f : (cv < 0.f ? -1e-8f : 0.f);
  return cv * sqrtf(1.0f / (nc + kk * cv * cv));
}

// rotp: [B][P][6] = c, -s, s, c, off_y, off_x of the psi sampling grid.
// Block = (image i0 + blockIdx.x / bands, band of kRows rows); X holds the
// chunk's images from image i0 on.
template <int N, bool MASK>
__global__ void __launch_bounds__(kThreads)
prep_rows_kernel(const float* __restrict__ images, int P, int n, int i0, const float* __restrict__ rotp,
                 const float* __restrict__ mask, const float2* __restrict__ twg, float2* __restrict__ X) {
  using T = PrepLayout<N>;
  constexpr int MH = T::MH, SL = T::SL, bands = N / kRows;
  __shared__ float2 tw[N];
  __shared__ float2 seq[(kRows / 2) * SL];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = blockIdx.x / bands, band = blockIdx.x - li * bands;
  const int img = i0 + li, b = img / P;
  const int y0 = band * kRows;
  for (int k = tid; k < N; k += kThreads) tw[k] = twg[k];
  const float* src = images + (size_t)b * N * N;
  const float* rp = rotp + (size_t)img * 6;
  const float m00 = rp[0], m01 = rp[1], m10 = rp[2], m11 = rp[3], oy = rp[4], ox = rp[5];
  // Each wave samples 8x8 output blocks, so one load instruction's addresses
  // stay within a compact rotated source patch; the two horizontal taps of a
  // row come from one 8-byte load.  Branch-free: the 2x2 footprint is read at
  // clamped coordinates (always in bounds) and taps outside the image are
  // zeroed by selects.  The loads of a group of GI pixels are all issued
  // before any is used.
  constexpr int BPR = N / 8;
  constexpr int ITER = 2 * BPR / (kThreads / 64);
  constexpr int GI = (ITER % 2 == 0 && ITER > 8) ? ITER / 2 : ITER;
  float* seq_f = (float*)seq;
  auto coords = [&](int it, int& yl, int& Y, int& Xc, float& sy, float& sx) {
    const int blk = wave + it * (kThreads / 64);
    const int hb = blk / BPR, bc = blk - hb * BPR;
    yl = hb * 8 + (lane >> 3);
    Y = y0 + yl;
    Xc = bc * 8 + (lane & 7);
    sy = (float)Y * m00 + (float)Xc * m01 + oy;
    sx = (float)Y * m10 + (float)Xc * m11 + ox;
  };
#pragma unroll
  for (int g0 = 0; g0 < ITER; g0 += GI) {
    f2a4 r0[GI], r1[GI];
    float mv[GI];
#pragma unroll
    for (int u = 0; u < GI; ++u) {
      int yl, Y, Xc;
      float sy, sx;
      coords(g0 + u, yl, Y, Xc, sy, sx);
      // clamp before converting so far-outside samples stay finite ints
      const int iy = (int)fminf(fmaxf(floorf(sy), -2.0f), (float)N);
      const int ix = (int)fminf(fmaxf(floorf(sx), -2.0f), (float)N);
      const int cy = min(max(iy, 0), N - 2), cx = min(max(ix, 0), N - 2);
      const float* s0 = src + (size_t)cy * N + cx;
      r0[u] = *(const f2a4*)s0;
      r1[u] = *(const f2a4*)(s0 + N);
      if constexpr (MASK) mv[u] = mask[(size_t)Y * N + Xc];
    }
#pragma unroll
    for (int u = 0; u < GI; ++u) {
      int yl, Y, Xc;
      float sy, sx;
      coords(g0 + u, yl, Y, Xc, sy, sx);
      const float fy0 = floorf(sy), fx0 = floorf(sx);
      const float ty = sy - fy0, tx = sx - fx0;
      const int iy = (int)fminf(fmaxf(fy0, -2.0f), (float)N);
      const int ix = (int)fminf(fmaxf(fx0, -2.0f), (float)N);
      // the loaded 2x2 block starts at (clamp(iy), clamp(ix)); inside the
      // image tap row iy is its row 0 except at the bottom edge (iy = N-1,
      // block at N-2), and tap row iy+1 its row 1 except at the top edge
      // (iy = -1, block at 0); likewise for columns; taps outside are zero
      const bool y0ok = (unsigned)iy < (unsigned)N, y1ok = (unsigned)(iy + 1) < (unsigned)N;
      const bool x0ok = (unsigned)ix < (unsigned)N, x1ok = (unsigned)(ix + 1) < (unsigned)N;
      const bool top = iy < 0, bottom = iy >= N - 1, left = ix < 0, right = ix >= N - 1;
      const f2a4 a0 = r0[u], a1 = r1[u];
      const f2a4 row0 = bottom ? a1 : a0;   // tap row iy
      const f2a4 row1 = top ? a0 : a1;      // tap row iy + 1
      const float a00 = (y0ok && x0ok) ? (right ? row0.y : row0.x) : 0.f;
      const float a01 = (y0ok && x1ok) ? (left ? row0.x : row0.y) : 0.f;
      const float a10 = (y1ok && x0ok) ? (right ? row1.y : row1.x) : 0.f;
      const float a11 = (y1ok && x1ok) ? (left ? row1.x : row1.y) : 0.f;
      float v = ((1.0f - ty) * (1.0f - tx)) * a00 + ((1.0f - ty) * tx) * a01 +
                (ty * (1.0f - tx)) * a10 + (ty * tx) * a11;
      if constexpr (MASK) v *= mv[u];
      seq_f[2 * ((yl >> 1) * SL + lds_pad<kPad>(Xc)) + (yl & 1)] = v;  // row pair -> (re, im)
    }
  }
  __syncthreads();
  static_assert(kRows / 2 == 2 * (kThreads / 64), "two sequences per wave");
  fft_wave_inplace2<N, false, kPad>(seq + (2 * wave) * SL, seq + (2 * wave + 1) * SL, tw, lane);
  __syncthreads();
  // split the two real rows' spectra; write the tiles the crop uses
  const int tiles = n / 2 / kTK + 1;
  float2* Xi = X + (size_t)li * T::IMG;
  constexpr int KW = T::NT * kTK;
  for (int t = tid; t < (kRows / 2) * KW; t += kThreads) {
    const int q = t / KW, k = t - q * KW;
    if (k >= MH || k / kTK >= tiles) continue;
    const float2 z = seq[q * SL + lds_pad<kPad>(k)];
    const float2 zc = cconj(seq[q * SL + lds_pad<kPad>((N - k) % N)]);
    const int ya = y0 + 2 * q;
    float2* Wk = Xi + ((size_t)(k / kTK) * N) * kTK + (k % kTK);
    Wk[(size_t)ya * kTK] = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y + zc.y));
    Wk[(size_t)(ya + 1) * kTK] = make_float2(0.5f * (z.y - zc.y), -0.5f * (z.x - zc.x));
  }
}

// ctf_tab with defocus:  [4][n][n/2+1] = x, y, u4, ncurve (PackedCTFTorch.__init__)
// ctf_tab without:       [B][n][n/2+1] explicit packed CTF, or NULL (ones)
// Block = (particle b0 + blockIdx.x / tiles, column tile); the chunk starts
// at particle b0, whose first image is X's image 0.
template <int N>
__global__ void __launch_bounds__(kThreads)
prep_cols_kernel(int P, int n, int b0, const float* __restrict__ ctf_tab, const float* __restrict__ defocus,
                 float k1, float k2, float k3, float k5, float kk, const float2* __restrict__ twg,
                 const float2* __restrict__ X, float2* __restrict__ spectra, int ldo,
                 double* __restrict__ partial) {
  using T = PrepLayout<N>;
  constexpr int MH = T::MH, SL = T::SL;
  __shared__ float2 tw[N];
  __shared__ float2 seq[kTK * SL];
  __shared__ float ctf_l[256 * kTK];  // [row][q], rows < n <= N <= 256
  __shared__ double red[kThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tiles = n / 2 / kTK + 1;
  const int lb = blockIdx.x / tiles, tile = blockIdx.x - lb * tiles;
  const int b = b0 + lb;
  const int k0 = tile * kTK, kmax = n / 2;
  const int mh_n = n / 2 + 1;
  const size_t plane_n = (size_t)n * mh_n;
  const int n16 = (n + 15) & ~15, fnct = frag_nct(n);
  for (int k = tid; k < N; k += kThreads) tw[k] = twg[k];
  const bool has_ctf = defocus != nullptr;
  const bool explicit_ctf = !has_ctf && ctf_tab != nullptr;
  float axx = 0.f, ayy = 0.f, axy = 0.f;
  if (has_ctf) {
    const float dfu = defocus[3 * b + 0], dfv = defocus[3 * b + 1];
    const float ang = defocus[3 * b + 2] * 0.017453292519943295f;
    const float sa = sinf(ang), ca = cosf(ang);
    axx = (-dfu * ca) * ca - (dfv * sa) * sa;
    ayy = (-dfu * sa) * sa - (dfv * ca) * ca;
    axy = (-dfu * sa) * ca + (dfv * sa) * ca;
  }
  for (int t = tid; t < n * kTK; t += kThreads) {
    const int r = t / kTK, q = t - r * kTK;
    const int c = k0 + q;
    float w = 1.0f;
    if (c <= kmax) {
      const size_t e = (size_t)r * mh_n + c;
      if (has_ctf)
        w = ctf_weight(ctf_tab[e], ctf_tab[plane_n + e], ctf_tab[2 * plane_n + e], ctf_tab[3 * plane_n + e], axx,
                       axy, ayy, k1, k2, k3, k5, kk);
      else if (explicit_ctf)
        w = ctf_tab[(size_t)b * plane_n + e];
    }
    ctf_l[t] = w;
  }
  constexpr int PER = N * kTK / kThreads;  // tile elements per thread
  static_assert(PER * kThreads == N * kTK, "tile not divisible by the block");
  float2 nxt[PER];
  auto load_tile = [&](int p) {
    const float2* Xt = X + ((size_t)lb * P + p) * T::IMG + (size_t)tile * N * kTK;
#pragma unroll
    for (int u = 0; u < PER; ++u) nxt[u] = Xt[tid + u * kThreads];
  };
  load_tile(0);
  for (int p = 0; p < P; ++p) {
    __syncthreads();  // the previous image's output reads are done
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int t = tid + u * kThreads;
      const int Y = t / kTK, q = t - Y * kTK;
      seq[q * SL + lds_pad<kPad>(Y)] = k0 + q < MH ? nxt[u] : make_float2(0.f, 0.f);
    }
    if (p + 1 < P) load_tile(p + 1);  // in flight during this image's transform
    __syncthreads();
    static_assert(kTK == 2 * (kThreads / 64), "two sequences per wave");
    fft_wave_inplace2<N, false, kPad>(seq + (2 * wave) * SL, seq + (2 * wave + 1) * SL, tw, lane);
    __syncthreads();
    // output: rows ldo apart, or (ldo = 0) fragment order
    float2* out = spectra + ((size_t)b * P + p) * (ldo ? (size_t)n * ldo : (size_t)frag_plane(n));
    double acc = 0.0;
    // element t of the tile is (row 16 ks + 4g + 2h + e, column k0 + i) for
    // t = (((ks * 2 + h) * 4 + g) * 8 + i) * 2 + e: sixteen consecutive t
    // are one 128-byte run of a fragment-order tile
    for (int t = tid; t < n16 * kTK; t += kThreads) {
      const int q = (t >> 1) & 7, rest = t >> 4;
      const int r = 16 * (rest >> 3) + 4 * (rest & 3) + 2 * ((rest >> 2) & 1) + (t & 1);
      const int c = k0 + q;
      if (c > kmax || r >= n) continue;
      const int fy = r < n / 2 ? r : r - n;
      float2 v;
      if (c < n / 2 || n == N) {
        v = seq[q * SL + lds_pad<kPad>(fy < 0 ? fy + N : fy)];
      } else {  // packed column n/2 holds frequency -n/2 = conj of (+n/2) at -fy
        const int my = -fy;
        v = cconj(seq[q * SL + lds_pad<kPad>(my < 0 ? my + N : my)]);
      }
      const float w = ctf_l[r * kTK + q];
      v.x *= w;
      v.y = -(v.y * w);  // conj
      out[ldo ? (size_t)r * ldo + c : (size_t)frag_pos(n, fnct, r, c)] = v;
      acc += (double)v.x * v.x + (double)v.y * v.y;
    }
    acc = wave_sum_f64(acc);
    if (lane == 0) red[wave] = acc;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      for (int w = 0; w < kThreads / 64; ++w) s += red[w];
      partial[((size_t)lb * P + p) * T::NT + tile] = s;
    }
  }
}

// Zeros at the pads of fragment-order planes (rows n .. 16 nks, and columns
// n/2+1 .. 16 nct when the last column is not kept apart): the positions no
// preparation block writes.
__global__ void frag_zero_pads_kernel(float2* __restrict__ dst, size_t rows, int n) {
  const int m = n / 2 + 1, nct = frag_nct(n), nks = frag_nks(n);
  const bool lone = (m & 15) == 1;
  const int rpad = 16 * nks - n, wcol = 16 * nct + (lone ? 1 : 0);
  const int padA = rpad * wcol, padB = lone ? 0 : n * (16 * nct - m);
  const size_t per = (size_t)(padA + padB), total = rows * per;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const size_t row = t / per;
    const int e = (int)(t - row * per);
    int r, c;
    if (e < padA) {
      r = n + e / wcol;
      c = e % wcol;
      if (lone && c == 16 * nct) c = m - 1;
    } else {
      r = (e - padA) / (16 * nct - m);
      c = m + (e - padA) % (16 * nct - m);
    }
    dst[row * (size_t)frag_plane(n) + frag_pos(n, nct, r, c)] = make_float2(0.f, 0.f);
  }
}

// The forward twiddle table of length N (the same values every FFT here uses).
__global__ void prep_twiddle_kernel(float2* __restrict__ tw, int N) {
  fill_twiddles(tw, N, threadIdx.x, blockDim.x);
}

// inv_norm[b][p] = 1 / sqrt(sum over tiles of partial), tiles in order.
__global__ void prep_norm_kernel(const double* __restrict__ partial, int nimg, int stride, int tiles, int i0,
                                 float* __restrict__ inv_norm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nimg) return;
  double s = 0.0;
  for (int t = 0; t < tiles; ++t) s += partial[(size_t)i * stride + t];
  inv_norm[i0 + i] = 1.0f / (float)sqrt(s);
}

template <int N>
size_t prep_image_bytes() { return PrepLayout<N>::IMG * sizeof(float2) + PrepLayout<N>::NT * sizeof(double); }
constexpr size_t kTwiddleBytes = 256 * sizeof(float2);  // table at the workspace's head (N <= 256)

template <int N>
int launch_prepare(const float* images, int B, int P, int n, const float* rotp, const float* mask,
                   const float* ctf_tab, const float* defocus, const float* k, float2* spectra, int ldo,
                   float* inv_norm, void* ws, size_t ws_bytes, hipStream_t st) {
  using T = PrepLayout<N>;
  const size_t cap = ws_bytes > kTwiddleBytes ? (ws_bytes - kTwiddleBytes) / prep_image_bytes<N>() : 0;
  const int chunk = (int)std::min<size_t>((size_t)B, cap / (size_t)P);
  GIS_REQUIRE(chunk >= 1, "gis_prepare_particles: workspace of %zu bytes holds %zu images, below one particle (%d)",
              ws_bytes, cap, P);
  float2* twg = (float2*)ws;
  float2* X = (float2*)((char*)ws + kTwiddleBytes);
  double* partial = (double*)((char*)X + (size_t)chunk * P * T::IMG * sizeof(float2));
  const int tiles = n / 2 / kTK + 1;
  hipLaunchKernelGGL(prep_twiddle_kernel, dim3(1), dim3(256), 0, st, twg, N);
  GIS_LAUNCH_CHECK("gis_prepare_particles twiddles");
  if (ldo == 0 && (size_t)frag_plane(n) > (size_t)n * (n / 2 + 1)) {
    hipLaunchKernelGGL(frag_zero_pads_kernel, dim3(std::max(1, std::min(B * P, 4096))), dim3(256), 0, st, spectra,
                       (size_t)B * P, n);
    GIS_LAUNCH_CHECK("gis_prepare_particles pads");
  }
  for (int b0 = 0; b0 < B; b0 += chunk) {
    const int nb = std::min(chunk, B - b0);
    const int i0 = b0 * P, nimg = nb * P;
    if (mask)
      hipLaunchKernelGGL((prep_rows_kernel<N, true>), dim3((unsigned)(nimg * (N / kRows))), dim3(kThreads), 0, st,
                         images, P, n, i0, rotp, mask, twg, X);
    else
      hipLaunchKernelGGL((prep_rows_kernel<N, false>), dim3((unsigned)(nimg * (N / kRows))), dim3(kThreads), 0, st,
                         images, P, n, i0, rotp, mask, twg, X);
    GIS_LAUNCH_CHECK("gis_prepare_particles rows");
    hipLaunchKernelGGL(prep_cols_kernel<N>, dim3((unsigned)(nb * tiles)), dim3(kThreads), 0, st, P, n, b0,
                       ctf_tab, defocus, k[0], k[1], k[2], k[3], k[4], twg, X, spectra, ldo, partial);
    GIS_LAUNCH_CHECK("gis_prepare_particles columns");
    hipLaunchKernelGGL(prep_norm_kernel, dim3((unsigned)((nimg + 255) / 256)), dim3(256), 0, st, partial, nimg,
                       T::NT, tiles, i0, inv_norm);
    GIS_LAUNCH_CHECK("gis_prepare_particles norm");
  }
  return GIS_OK;
}

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

// Workspace bytes for gis_prepare_particles over `particles` particles of P
// psi images each: a chunk of min(particles * P, max(512, P)) row-transformed
// images and their norm partials (the call processes the particles in
// chunks that fit, at least one particle).
GIS_API size_t gis_prepare_particles_workspace_ex(int N, int particles, int P) {
  size_t s = 0;
  switch (N) {
    case 32: s = prep_image_bytes<32>(); break;
    case 64: s = prep_image_bytes<64>(); break;
    case 96: s = prep_image_bytes<96>(); break;
    case 128: s = prep_image_bytes<128>(); break;
    case 192: s = prep_image_bytes<192>(); break;
    case 256: s = prep_image_bytes<256>(); break;
    default: return 0;
  }
  if (particles <= 0 || P <= 0) return 0;
  return kTwiddleBytes + std::min<size_t>((size_t)particles * P, (size_t)std::max(kMaxChunkImages, P)) * s;
}

// The same for up to 64 psi images per particle.
GIS_API size_t gis_prepare_particles_workspace(int N, int particles) {
  return gis_prepare_particles_workspace_ex(N, particles, 64);
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
//   spectra  [B][P][n][ld] complex64 output (conjugated, unnormalised, unshifted;
//            ld >= n/2+1, columns past n/2 untouched), or with ld = 0
//            [B][P][frag_plane(n)] in fragment order (the searches' operand
//            layout, pads zero); gis_prepare_particles: ld = n/2+1
//   inv_norm [B][P] float32 output
GIS_API int gis_prepare_particles_ld(const float* images, int B, int N, int n, int P, const float* rotp,
                                     const float* mask, const float* ctf_tab, const float* defocus,
                                     const float* kconst, void* spectra, int ld, float* inv_norm, void* ws,
                                     size_t ws_bytes, void* stream) {
  GIS_REQUIRE(B >= 0 && P >= 1, "gis_prepare_particles: B=%d P=%d", B, P);
  GIS_REQUIRE(n >= 2 && n <= N && n % 2 == 0, "gis_prepare_particles: newbox %d must be even and <= box %d", n, N);
  GIS_REQUIRE(ld == 0 || ld >= n / 2 + 1, "gis_prepare_particles: row stride %d below n/2+1 = %d", ld, n / 2 + 1);
  if (B == 0) return GIS_OK;
  GIS_REQUIRE(images && rotp && spectra && inv_norm && ws && kconst, "gis_prepare_particles: null pointer");
  GIS_REQUIRE(defocus == nullptr || ctf_tab != nullptr, "gis_prepare_particles: defocus without CTF tables");
  hipStream_t st = (hipStream_t)stream;
  float2* S = (float2*)spectra;
#define GIS_PREP(NV) \
  launch_prepare<NV>(images, B, P, n, rotp, mask, ctf_tab, defocus, kconst, S, ld, inv_norm, ws, ws_bytes, st)
  switch (N) {
    case 32: return GIS_PREP(32);
    case 64: return GIS_PREP(64);
    case 96: return GIS_PREP(96);
    case 128: return GIS_PREP(128);
    case 192: return GIS_PREP(192);
    case 256: return GIS_PREP(256);
    default:
      gis_set_error("gis_prepare_particles: box %d not in {32, 64, 96, 128, 192, 256}", N);
      return GIS_EUNSUPPORTED;
  }
#undef GIS_PREP
}

GIS_API int gis_prepare_particles(const float* images, int B, int N, int n, int P, const float* rotp,
                                  const float* mask, const float* ctf_tab, const float* defocus,
                                  const float* kconst, void* spectra, float* inv_norm, void* ws,
                                  size_t ws_bytes, void* stream) {
  return gis_prepare_particles_ld(images, B, N, n, P, rotp, mask, ctf_tab, defocus, kconst, spectra, n / 2 + 1,
                                  inv_norm, ws, ws_bytes, stream);
}
