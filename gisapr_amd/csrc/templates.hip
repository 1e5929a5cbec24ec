// Template bank: central-slice projection, whitening and packing.
//
// For every direction d (rotation R_d = Euler(rot, tilt, 0)^T) this computes
// the unit-norm packed template of IntegratedProjector.build_templates
// (v606:1077-1147):
//   1. trilinear central-slice gather from the 3-D half spectrum (bricked
//      layout of gis_volume_dft) with the Friedel conjugation where the
//      rotated x < 0 (v606:856-896, 1099-1113)
//   2. ifftshift rows, irfft2 (complex-to-real rule: the imaginary parts of
//      columns 0 and N are dropped after the row transform), ifftshift, crop
//      the padding (v606:1122-1130)
//   3. whitening (v606:938-992): fft2, ring mean of |F| -> divide (DC ring
//      kept), inverse; least-squares plane with x = row and y = column
//      (v606:916-918), subtract, (x - mean) / std
//   4. x mask, fft2, crop to n, pack to the rfft2 layout (v606:494-541),
//      x FSC, / L2 norm (v606:1131-1146)
// The bank is stored UNshifted ([M][n][n/2+1], rfft2 layout); the
// correlation kernel folds the reference's pre-shift roll into its tables.
//
// gfx950 mapping: one 256-thread workgroup per direction (persistent grid),
// every 1-D transform an LDS Stockham FFT, the 2-D passes staged through a
// per-workgroup scratch slot in HBM (mostly L2/MALL resident).  Two real
// rows travel in one complex FFT (real + i*imag) in the row passes.
#include "common.h"
#include "vdft_layout.h"

#include <type_traits>

#pragma clang fp contract(off)

namespace {

// Workgroup shape (build-time tunables): 256 threads and 52.8 KB of LDS let
// three workgroups share a CU, so a 768-direction bank runs in one round.
#ifndef GIS_TPL_THREADS
#define GIS_TPL_THREADS 256
#endif
#ifndef GIS_TPL_LDS
#define GIS_TPL_LDS 6600
#endif
constexpr int kThreads = GIS_TPL_THREADS;
constexpr int kWavesT = kThreads / 64;
constexpr int kLdsFloat2 = GIS_TPL_LDS;

__device__ __forceinline__ int ring_of(int fy, int kx) {
  return (int)rintf(sqrtf((float)(fy * fy + kx * kx)));
}

// Sequences live in LDS at buf[q * SP + lds_pad<kPad>(i)] and are
// transformed in place, one wave per sequence (one buffer, no ping-pong).
constexpr int kPad = 3;
// Sequence strides in LDS (float2): the padded length plus a skew chosen so
// the transposing tile loops and the gather's LDS stores spread over banks.
#ifndef GIS_TPL_SP2X
#define GIS_TPL_SP2X 1
#endif
#ifndef GIS_TPL_SP1X
#define GIS_TPL_SP1X 1
#endif
template <int L>
constexpr int tpl_sp(int skew) { return lds_padded_len<L, kPad>() + skew; }

// A strided tile loop of (at least) IT iterations whose loads are issued in
// batches of UB before any of them is used: load(it) returns the iteration's
// values and must stay in bounds for every it < round_up(IT, UB) (callers
// clamp indices); store(it, v) consumes them and drops iterations past the
// tile.  A scheduling barrier between a batch's loads and its uses keeps a
// batch's loads in flight together whatever the register budget.
template <int IT, int UB, typename LoadF, typename StoreF>
__device__ __forceinline__ void batched_tile_loop(LoadF load, StoreF store) {
  constexpr int ITP = (IT + UB - 1) / UB * UB;
#pragma unroll 1
  for (int i0 = 0; i0 < ITP; i0 += UB) {
    decltype(load(0)) v[UB];
#pragma unroll
    for (int u = 0; u < UB; ++u) v[u] = load(i0 + u);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < UB; ++u) store(i0 + u, v[u]);
  }
}

// x / D for 0 <= x < XMAX as one 24-bit multiply and a shift: with
// M = ceil(2^S / D) = (2^S + e) / D, floor(x M / 2^S) = floor(x / D) whenever
// (XMAX - 1) e < 2^S (the added x e / (D 2^S) stays below 1/D).  The
// compiler's 32-bit division by a constant costs a quarter-rate multiply-high,
// and the tile loops divide every element index.
template <unsigned D, unsigned XMAX>
struct SmallDiv {
  static constexpr unsigned long long m_of(unsigned s) { return ((1ull << s) + D - 1) / D; }
  static constexpr bool ok(unsigned s) {
    return m_of(s) < (1ull << 24) && (unsigned long long)(XMAX - 1) * m_of(s) < (1ull << 32) &&
           (unsigned long long)(XMAX - 1) * (m_of(s) * D - (1ull << s)) < (1ull << s);
  }
  static constexpr unsigned pick() {
    for (unsigned s = 1; s < 32; ++s)
      if (ok(s)) return s;
    return 0;
  }
  static constexpr unsigned S = pick();
  static constexpr unsigned M = (unsigned)m_of(S);
  static_assert(S != 0, "no 24-bit reciprocal is exact over the range");
  static __device__ __forceinline__ unsigned div(unsigned x) { return __umul24(x, M) >> S; }
};

// (x / D, x % D) for 0 <= x < XMAX
template <unsigned D, unsigned XMAX>
__device__ __forceinline__ void divmod_small(int x, int& q, int& r) {
  q = (int)SmallDiv<D, XMAX>::div((unsigned)x);
  r = x - (int)__umul24((unsigned)q, D);
}

constexpr int largest_divisor_upto(int n, int cap) {
  int d = cap < n ? cap : n;
  while (d > 1 && n % d != 0) --d;
  return d < 1 ? 1 : d;
}

// Optional coefficient output: instead of the packed bank row, direction d
// writes the listed coefficients, coeffs[d * Kpad + kslot[packed index]] for
// the packed indices with kslot >= 0 (K of them, slots [0, K)), scaled by
// 1/norm once the norm is known, and zeros in slots [K, Kpad).
struct CoefOut {
  const int* kslot;  // [n * (n/2+1)], -1 = not kept; null = write the bank
  int K, Kpad;
  float2* coeffs;    // [M][Kpad]
  int ldb = 0;       // bank row stride in complex values (0: n/2+1); rows are n of them
  bool frag = false; // bank planes in fragment order (frag_pos; ldb unused)
};

// Real-space modes (the standalone projector's path): `img` replaces the
// central slices by [M][N][N] real images; `proj` receives [M][N][N] real
// images instead of the bank -- the cropped projection, or after the ring
// division (whiten) and / or the plane fit + (x - mean) / std (normz, std
// threshold eps).
struct TplIO {
  const float* img;
  float* proj;
  int whiten, normz;
  float eps;
  float2* slices;  // [M][L][L/2+1] sampled central slices (rows fftshifted) instead of any transform, or null
};

template <int N, int L>
struct TplLayout {
  static constexpr int L2 = L;           // slice transform length (2N padded, N unpadded)
  static constexpr int M2 = L / 2 + 1;   // slice half-spectrum width
  static constexpr int MH = N / 2 + 1;   // box half-spectrum width
  static constexpr int NR = (int)(0.7072 * N) + 3;  // rings (upper bound)
  static constexpr int FIXED = L2 + N + NR / 2 + 1 + 32;  // twiddles, ring divisors, reductions
  static constexpr int CAP = kLdsFloat2 - FIXED;           // sequence buffer
  static constexpr int SP2 = tpl_sp<L2>(GIS_TPL_SP2X);     // stride of a 2N sequence
  static constexpr int SP1 = tpl_sp<N>(GIS_TPL_SP1X);      // stride of an N sequence
  // sequence counts per tile are rounded down to a multiple of the waves
  // (each wave owns whole sequences of a tile)
  static constexpr int roundw(int x) { return x >= kWavesT ? x - x % kWavesT : x; }
  static constexpr int TK1 = roundw(CAP / SP2);           // phase 1 columns per tile
  static constexpr int RB2 = roundw(CAP / SP2);           // phase 2 complex rows per batch
  static constexpr int TK3 = CAP / SP1;                   // box-length column tiles
  static constexpr int RB6 = CAP / SP1;                   // box-length complex rows
  static_assert(TK1 >= 1 && RB2 >= 1, "LDS budget");
  // Scratch images are stored tile-major (the columns of one tile of TK
  // columns form one contiguous [rows][TK] block), so the column passes
  // read and write contiguous blocks and the row passes full 32-byte runs.
  static constexpr int UT = (M2 + TK1 - 1) / TK1;  // column tiles of U
  static constexpr int WT = (MH + TK3 - 1) / TK3;  // column tiles of W
  // U also holds the ring partials ([MH][2][N/2+1] sums + counts) and the
  // real image (N x N floats) once its columns are consumed
  static constexpr size_t cmax(size_t a, size_t b) { return a > b ? a : b; }
  static constexpr size_t U_ELEMS = cmax(cmax((size_t)N * UT * TK1, (size_t)MH * 2 * (N / 2 + 1)),
                                         ((size_t)N * N + 1) / 2);  // float2
  static constexpr size_t W_ELEMS = (size_t)N * WT * TK3;  // float2
  static constexpr size_t SLOT = U_ELEMS + W_ELEMS;
  static_assert(SLOT < (1u << 31), "32-bit scratch offsets");
};

// C spectra at D + c * stride (c < C, stride in complex64 values): work item
// w = c * M + d is direction d of spectrum c, and writes output row w.
struct VolSet {
  int C;
  size_t stride;
};

template <int N, int L>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3)))
build_templates_kernel(const float2* __restrict__ D0, VolSet vs, const float* __restrict__ rts, int M, int n,
                       const float* __restrict__ mask, const float* __restrict__ fsc,
                       float2* __restrict__ bank, float* __restrict__ norms, float2* __restrict__ scratch,
                       int normalize, CoefOut co, TplIO io) {
  using T = TplLayout<N, L>;
  constexpr int L2 = T::L2, M2 = T::M2, MH = T::MH;
  extern __shared__ float2 lds[];  // kLdsFloat2 values
  float2* tw2 = lds;                       // L2 twiddles
  float2* tw1 = tw2 + L2;                  // N twiddles
  float* div = (float*)(tw1 + N);          // NR ring divisors
  double* red = (double*)(div + T::NR + (T::NR & 1));  // reduction scratch (16 doubles)
  float2* buf = lds + T::FIXED;
  constexpr int SP2 = T::SP2, SP1 = T::SP1;
  auto pd = [](int i) { return lds_pad<kPad>(i); };
  const int tid = threadIdx.x;
  constexpr int nthr = kThreads;  // launched with exactly kThreads threads
  fill_twiddles(tw2, L2, tid, nthr);
  fill_twiddles(tw1, N, tid, nthr);
  float2* U = scratch + (size_t)blockIdx.x * T::SLOT;
  float2* W = U + T::U_ELEMS;
  // 32-bit element offsets within the workgroup's scratch slot
  auto uidx = [](int row, int col) -> unsigned {
    const unsigned c = (unsigned)col;
    const unsigned ct = SmallDiv<T::TK1, (unsigned)(T::UT * T::TK1)>::div(c);
    return (ct * N + (unsigned)row) * T::TK1 + (c - __umul24(ct, (unsigned)T::TK1));
  };
  using ColDiv = SmallDiv<T::TK3, (unsigned)(T::MH + T::TK3)>;
  auto widx = [](int row, int col) -> unsigned {
    const unsigned c = (unsigned)col, ct = ColDiv::div(c);
    return (ct * N + (unsigned)row) * T::TK3 + (c - __umul24(ct, (unsigned)T::TK3));
  };
  // A column tile of W (k0 a multiple of TK3) is the contiguous block
  // W[k0 * N + t], t = Y * TK3 + q: tile loops index it linearly and split t
  // into (row Y, column q) for the LDS side.
  using TileDiv = SmallDiv<T::TK3, (unsigned)(((T::TK3 * N + kThreads - 1) / kThreads + 3) / 4 * 4 * kThreads)>;
  auto tile_yq = [](int t, int& Y, int& q) {
    Y = (int)TileDiv::div((unsigned)t);
    q = t - (int)__umul24((unsigned)Y, (unsigned)T::TK3);
  };
  // Element i of a per-workgroup array addressed as a 32-bit byte offset from
  // a uniform base (the slot and the tables are < 4 GiB), which the loads and
  // stores take directly instead of a 64-bit address per element.
  auto at = [](auto* base, unsigned i) -> decltype(*base)& {
    using E = std::remove_reference_t<decltype(*base)>;
    return *reinterpret_cast<E*>(reinterpret_cast<std::conditional_t<std::is_const_v<E>, const char*, char*>>(base) +
                                 (size_t)(i * (unsigned)sizeof(E)));
  };
  float* Z = (float*)U;  // real image, aliases U once U is consumed
  const int mh_n = n / 2 + 1;
  const float inv_l2 = 1.0f / (float)L2;

#ifndef GIS_TPL_XCD_ORDER
#define GIS_TPL_XCD_ORDER 1
#endif
  // Workgroups are dispatched to the 8 XCDs round robin; with the XCD order
  // each round gives each XCD a contiguous range of work items.
  int w_first = blockIdx.x;
  if (GIS_TPL_XCD_ORDER) {
    const int g8 = gridDim.x / 8;
    if ((int)blockIdx.x < 8 * g8) w_first = (blockIdx.x % 8) * g8 + blockIdx.x / 8;
  }
  const int W_items = vs.C * M;
  for (int w = w_first; w < W_items; w += gridDim.x) {
    const int cand = w / M, d = w - cand * M;
    const float2* __restrict__ D = D0 + (size_t)cand * vs.stride;
    float r00 = 0.f, r01 = 0.f, r10 = 0.f, r11 = 0.f, r20 = 0.f, r21 = 0.f;
    if (!io.img) {  // image mode has no rotations
      const float* R = rts + 9 * d;
      r00 = R[0], r01 = R[1], r10 = R[3], r11 = R[4], r20 = R[6], r21 = R[7];
    }
    __syncthreads();

    // ---- phase 1: gather + inverse column FFT (length L) -----------------
    for (int k0 = 0; k0 < (io.img ? 0 : M2); k0 += T::TK1) {
      const int nk = min(T::TK1, M2 - k0);
      // Each wave samples blocks of (64/CW) slice rows x CW slice columns,
      // so one load instruction touches a compact patch of the 3-D spectrum,
      // which is stored in 2x2x4 bricks (one cache line each).  Branch-free:
      // every tap index is clamped into the array; a clamped upper tap
      // always carries weight 0 (the coordinate was clamped to the last
      // index first), so all eight loads of a sample are in flight together.
      constexpr int CW = T::TK1 <= 4 ? 4 : 8;  // columns per lane patch
      constexpr int RW = 64 / CW;              // rows per lane patch
#ifndef GIS_TPL_SPS
#define GIS_TPL_SPS 2
#endif
      // samples per thread and step (all their taps are issued together)
      constexpr int SPS = largest_divisor_upto((CW * L2) / kThreads, GIS_TPL_SPS);
      static_assert((CW * L2) % (SPS * kThreads) == 0, "whole steps per column group");
      const int ngroups = (nk + CW - 1) / CW;
      // (z0 + L/2) mod L for z0 in [0, L): the fftshift of the spectrum's z / y axes
      auto wrap = [](unsigned z) -> unsigned { return z >= (unsigned)(L2 / 2) ? z - L2 / 2 : z + L2 / 2; };
      for (int cg = 0; cg < ngroups; ++cg) {
        for (int t0 = tid; t0 < CW * L2; t0 += SPS * kThreads) {
          // SPS samples per step: all their taps are issued before any is used
          unsigned ad[SPS][8];
          float wt[SPS][3];  // tz, ty, tx
          int slot[SPS];
          bool cjs[SPS];
#pragma unroll
          for (int u = 0; u < SPS; ++u) {
            const int t = t0 + u * kThreads;
            const int lane = t & 63, rb = t >> 6;
            const int q = cg * CW + lane / RW, ip = rb * RW + lane % RW;
            slot[u] = q < nk ? q * SP2 + pd(ip) : -1;
            const int kx = k0 + min(q, nk - 1);
            const int f = ip < L2 / 2 ? ip : ip - L2;
            const float fy = (float)f * inv_l2;
            const float fx = (float)kx * inv_l2;
            float gx = r00 * fx + r01 * fy;
            float gy = r10 * fx + r11 * fy;
            float gz = r20 * fx + r21 * fy;
            const bool cj = gx < 0.0f;
            if (cj) { gx = -gx; gy = -gy; gz = -gz; }
            cjs[u] = cj;
            float cz = gz * (float)L2 + (float)(L2 / 2);
            float cy = gy * (float)L2 + (float)(L2 / 2);
            float cx = gx * (float)L2;
            cz = fminf(fmaxf(cz, 0.0f), (float)(L2 - 1));
            cy = fminf(fmaxf(cy, 0.0f), (float)(L2 - 1));
            cx = fminf(fmaxf(cx, 0.0f), (float)(M2 - 1));
            const float z0f = floorf(cz), y0f = floorf(cy), x0f = floorf(cx);
            wt[u][0] = cz - z0f;
            wt[u][1] = cy - y0f;
            wt[u][2] = cx - x0f;
            const unsigned z0 = (unsigned)(int)z0f, y0 = (unsigned)(int)y0f, x0 = (unsigned)(int)x0f;
            // Branch-free: a clamped upper tap always carries weight 0 (the
            // coordinate was clamped to the last index first)
            const unsigned pz0 = vdft_brick_z32<L2>(wrap(z0)), pz1 = vdft_brick_z32<L2>(wrap(min(z0 + 1, (unsigned)L2 - 1)));
            const unsigned py0 = vdft_brick_y32<L2>(wrap(y0)), py1 = vdft_brick_y32<L2>(wrap(min(y0 + 1, (unsigned)L2 - 1)));
            const unsigned px0 = vdft_brick_x32(x0), px1 = vdft_brick_x32(min(x0 + 1, (unsigned)M2 - 1));
            ad[u][0] = pz0 + py0 + px0; ad[u][1] = pz0 + py0 + px1;
            ad[u][2] = pz0 + py1 + px0; ad[u][3] = pz0 + py1 + px1;
            ad[u][4] = pz1 + py0 + px0; ad[u][5] = pz1 + py0 + px1;
            ad[u][6] = pz1 + py1 + px0; ad[u][7] = pz1 + py1 + px1;
          }
          // 32-bit byte offsets from the (scalar) base: the spectrum is < 4 GiB
          const char* Db = reinterpret_cast<const char*>(D);
          float2 v[SPS][8];
#pragma unroll
          for (int u = 0; u < SPS; ++u)
#pragma unroll
            for (int i = 0; i < 8; ++i) v[u][i] = *reinterpret_cast<const float2*>(Db + (size_t)(ad[u][i] * 8u));
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int u = 0; u < SPS; ++u) {
            const float tz = wt[u][0], ty = wt[u][1], tx = wt[u][2];
            const float wx0 = 1.0f - tx, wx1 = tx;
            const float w00 = (1.0f - tz) * (1.0f - ty), w01 = (1.0f - tz) * ty;
            const float w10 = tz * (1.0f - ty), w11 = tz * ty;
            float2 acc = make_float2(0.f, 0.f);
            auto tap = [&](const float2& a0, const float2& a1, float wzy) {
              acc.x += (wzy * wx0) * a0.x;
              acc.y += (wzy * wx0) * a0.y;
              acc.x += (wzy * wx1) * a1.x;
              acc.y += (wzy * wx1) * a1.y;
            };
            tap(v[u][0], v[u][1], w00);
            tap(v[u][2], v[u][3], w01);
            tap(v[u][4], v[u][5], w10);
            tap(v[u][6], v[u][7], w11);
            if (cjs[u]) acc.y = -acc.y;
            if (slot[u] >= 0) buf[slot[u]] = acc;
          }
        }
      }
      __syncthreads();
      if (io.slices) {
        // the sampled slice itself: row (ip + L/2) mod L (fftshifted order), column kx
        for (int t = tid; t < nk * L2; t += nthr) {
          int q, ip;
          divmod_small<L2, (unsigned)(T::TK1 * L2)>(t, q, ip);
          io.slices[((size_t)d * L2 + (ip + L2 / 2) % L2) * M2 + k0 + q] = buf[q * SP2 + pd(ip)];
        }
        __syncthreads();
        continue;
      }
      fft_lds_inplace<L2, true, kPad>(buf, nk, SP2, tw2, tid, nthr);
      // ifftshift + crop: projection row Y is spatial row (Y + (L-N)/2 + L/2) mod L
      for (int t = tid; t < T::TK1 * N; t += nthr) {
        int Y, q;
        divmod_small<T::TK1, (unsigned)(T::TK1 * N)>(t, Y, q);
        if (q >= nk) continue;
        const int y = (Y + (L2 - N) / 2 + L2 / 2) % L2;
        at(U, uidx(Y, k0 + q)) = buf[q * SP2 + pd(y)];
      }
      __syncthreads();
    }

    if (io.slices) continue;  // slices only: done with this direction

    // ---- phase 2: complex-to-real rows (2N) -> crop -> forward FFT (N) ----
    for (int y0 = 0; y0 < N; y0 += 2 * T::RB2) {
      const int nrow = min(2 * T::RB2, N - y0);
      const int nseq = (nrow + 1) / 2;
      if (io.img) {
        // image mode: rows ya (re) and yb (im) of the input image to the slot fronts
        const float* im = io.img + (size_t)d * N * N;
        constexpr int IT2I = (T::RB2 * N + kThreads - 1) / kThreads;
        batched_tile_loop<IT2I, 4>(
            [&](int it) {
              const int t = tid + it * kThreads;
              int q, X;
              divmod_small<N, (unsigned)(((T::RB2 * N + kThreads - 1) / kThreads + 3) / 4 * 4 * kThreads)>(t, q, X);
              const int ya = min(y0 + 2 * q, N - 1), yb = min(y0 + 2 * q + 1, N - 1);
              return make_float2(at(im, (unsigned)(ya * N + X)), at(im, (unsigned)(yb * N + X)));
            },
            [&](int it, float2 ab) {
              const int t = tid + it * kThreads;
              int q, X;
              divmod_small<N, (unsigned)(((T::RB2 * N + kThreads - 1) / kThreads + 3) / 4 * 4 * kThreads)>(t, q, X);
              if (2 * q + 1 >= nrow) ab.y = 0.f;
              if (q < nseq) buf[q * SP2 + pd(X)] = ab;
            });
        __syncthreads();
      } else {
      // Tile loads: unconditional (rows clamped into the image, unused
      // values dropped at the LDS store) and batched (batched_tile_loop).
      constexpr int IT2 = (T::RB2 * L2 + kThreads - 1) / kThreads;
      batched_tile_loop<IT2, 4>(
          [&](int it) {
            const int t = tid + it * kThreads;
            int q, k;
            divmod_small<L2, (unsigned)(((T::RB2 * L2 + kThreads - 1) / kThreads + 3) / 4 * 4 * kThreads)>(t, q, k);
            const int ya = min(y0 + 2 * q, N - 1), yb = min(y0 + 2 * q + 1, N - 1);
            const int ks = k <= L2 / 2 ? k : L2 - k;
            const float2 a = at(U, uidx(ya, ks)), b = at(U, uidx(yb, ks));
            return make_float4(a.x, a.y, b.x, b.y);
          },
          [&](int it, float4 ab) {
            const int t = tid + it * kThreads;
            int q, k;
            divmod_small<L2, (unsigned)(((T::RB2 * L2 + kThreads - 1) / kThreads + 3) / 4 * 4 * kThreads)>(t, q, k);
            // Hermitian extension of the half spectrum, Im of columns 0 and L/2 dropped
            const int ks = k <= L2 / 2 ? k : L2 - k;
            float2 a = make_float2(ab.x, ab.y), b = make_float2(ab.z, ab.w);
            if (2 * q + 1 >= nrow) b = make_float2(0.f, 0.f);
            if (ks == 0 || ks == L2 / 2) { a.y = 0.f; b.y = 0.f; }
            if (k > L2 / 2) { a.y = -a.y; b.y = -b.y; }
            if (q < nseq) buf[q * SP2 + pd(k)] = make_float2(a.x - b.y, a.y + b.x);  // a + i*b
          });
      __syncthreads();
      // inverse 2N transform; each wave then crops the centre N (scaled) of
      // its own sequences to the front of their slots; forward N transform
      fft_lds_inplace<L2, true, kPad>(buf, nseq, SP2, tw2, tid, nthr);
      {
        const int wave = tid >> 6, lane = tid & 63;
        const float sc = 1.0f / ((float)L2 * (float)L2);
        constexpr int KP = (N + 63) / 64;
        for (int q = wave; q < nseq; q += kWavesT) {
          float2* sq = buf + q * SP2;
          float2 keep[KP];
#pragma unroll
          for (int i = 0; i < KP; ++i) {
            const int X = lane + 64 * i;
            if (X < N) {
              const float2 v = sq[pd((X + (L2 - N) / 2 + L2 / 2) % L2)];
              keep[i] = make_float2(v.x * sc, v.y * sc);  // projection rows ya (re), yb (im)
            }
          }
          wave_lds_sync();
#pragma unroll
          for (int i = 0; i < KP; ++i)
            if (lane + 64 * i < N) sq[pd(lane + 64 * i)] = keep[i];
        }
        __syncthreads();
      }
      }
      if (io.proj && !io.whiten && !io.normz) {
        // the cropped projection itself: rows ya (re) and yb (im) of each slot front
        for (int t = tid; t < nseq * N; t += nthr) {
          int q, X;
          divmod_small<N, (unsigned)(T::RB2 * N)>(t, q, X);
          const float2 v = buf[q * SP2 + pd(X)];
          const int ya = y0 + 2 * q;
          io.proj[((size_t)d * N + ya) * N + X] = v.x;
          if (ya + 1 < y0 + nrow) io.proj[((size_t)d * N + ya + 1) * N + X] = v.y;
        }
        __syncthreads();
        continue;
      }
      fft_lds_inplace<N, false, kPad>(buf, nseq, SP2, tw1, tid, nthr);
      for (int t = tid; t < nseq * MH; t += nthr) {
        int q, k;
        divmod_small<MH, (unsigned)(T::RB2 * MH)>(t, q, k);
        const float2 z = buf[q * SP2 + pd(k)];
        const float2 zc = cconj(buf[q * SP2 + pd((N - k) % N)]);
        const int ya = y0 + 2 * q;
        at(W, widx(ya, k)) = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y + zc.y));
        if (ya + 1 < y0 + nrow)
          at(W, widx(ya + 1, k)) = make_float2(0.5f * (z.y - zc.y), -0.5f * (z.x - zc.x));
      }
      __syncthreads();
    }

    if (io.proj && !io.whiten && !io.normz) continue;  // raw projection: done with this direction
    const bool ring_divide = io.whiten || !io.proj;   // the bank is always whitened

    // ---- phase 3: forward column FFT (N) -> projection spectrum F, and the
    // ring sums of |F| over the full plane.  Along one half column (fy >= 0
    // or fy < 0 at fixed kx) the ring index never decreases with |fy| and
    // steps by at most one, so a wave runs a segmented scan over |fy| and
    // stores one partial (sum, count) per ring present in that half column;
    // phase 4 adds the partials of a ring in a fixed order (deterministic).
    // Half-plane columns 1..N/2-1 stand for two full-plane elements.
    if (ring_divide) {
    const int nr = ring_of(N / 2, N / 2) + 1;
    constexpr int RS = N / 2 + 1;  // rings per half column, upper bound
    float* rsum = (float*)U;       // [MH][2][RS] (U is free after phase 2)
    int* rcnt = (int*)(rsum + (size_t)MH * 2 * RS);
    for (int k0 = 0; k0 < MH; k0 += T::TK3) {
      const int nk = min(T::TK3, MH - k0);
      constexpr int IT3 = (T::TK3 * N + kThreads - 1) / kThreads;
      batched_tile_loop<IT3, 4>(
          [&](int it) {
            // W's last column tile is padded to TK3 columns: the load is in bounds
            return at(W, (unsigned)k0 * N + (unsigned)min(tid + it * kThreads, T::TK3 * N - 1));
          },
          [&](int it, float2 v) {
            int Y, q;
            tile_yq(tid + it * kThreads, Y, q);
            if (Y < N && q < nk) buf[q * SP1 + pd(Y)] = v;
          });
      __syncthreads();
      fft_lds_inplace<N, false, kPad>(buf, nk, SP1, tw1, tid, nthr);
      const float2* res = buf;
      for (int t = tid; t < T::TK3 * N; t += nthr) {
        int ky, q;
        tile_yq(t, ky, q);
        if (q >= nk) continue;
        at(W, (unsigned)k0 * N + (unsigned)t) = res[q * SP1 + pd(ky)];
      }
      const int lane = tid & 63;
      for (int task = tid >> 6; task < 2 * nk; task += nthr / 64) {
        const int q = task >> 1, h = task & 1;
        const int kx = k0 + q;
        const float wgt = (kx == 0 || kx == N / 2) ? 1.0f : 2.0f;
        const int iw = (kx == 0 || kx == N / 2) ? 1 : 2;
        float* ps = rsum + (size_t)(kx * 2 + h) * RS - kx;  // indexed by ring (>= kx)
        int* pc = rcnt + (size_t)(kx * 2 + h) * RS - kx;
        int carry_key = -1;
        float carry_val = 0.0f;
        int carry_cnt = 0;
        for (int c0 = 0; c0 < N / 2; c0 += 64) {
          const int a = c0 + lane;  // |fy| - h
          const bool ok = a < N / 2;
          const int fy = h ? -(a + 1) : a;
          int key = -2 - lane;  // unique sentinel above the valid lanes
          float val = 0.0f;
          int cnt = 0;
          if (ok) {
            const float2 v = res[q * SP1 + pd(fy < 0 ? fy + N : fy)];
            val = sqrtf(v.x * v.x + v.y * v.y) * wgt;
            cnt = iw;
            key = ring_of(fy, kx);
          }
#pragma unroll
          for (int d = 1; d < 64; d <<= 1) {
            const float vu = __shfl_up(val, d, 64);
            const int cu = __shfl_up(cnt, d, 64);
            const int ku = __shfl_up(key, d, 64);
            if (lane >= d && ku == key) { val += vu; cnt += cu; }
          }
          if (key == carry_key) { val += carry_val; cnt += carry_cnt; }
          const int fyn = h ? fy - 1 : fy + 1;
          const bool more = a + 1 < N / 2;
          const bool last = ok && (!more || ring_of(fyn, kx) != key);
          if (last) { ps[key] = val; pc[key] = cnt; }
          carry_key = __shfl(key, 63, 64);
          carry_val = __shfl(val, 63, 64);
          carry_cnt = __shfl(cnt, 63, 64);
        }
      }
      __syncthreads();
    }

    // ---- phase 4: ring divisors ------------------------------------------
    // One thread per ring adds the ring's partials in a fixed order (kx
    // ascending, then the half column).  The partials are in scratch, so the
    // loads go out unconditionally in batches of KB columns (offsets clamped
    // into the table, unused values masked afterwards) rather than as one
    // dependent load per step.
    if (tid < nr) {
      const int ring = tid;
      float ring_sum = 0.0f;
      int ring_cnt = 0;
      constexpr int KB = 8;
      const int kxmax = min(ring, N / 2);
      for (int kx0 = 0; kx0 <= kxmax; kx0 += KB) {
        float sv[KB][2];
        int cv[KB][2];
#pragma unroll
        for (int u = 0; u < KB; ++u) {
          const int kx = min(kx0 + u, N / 2);
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int off = (kx * 2 + h) * RS + min(max(ring - kx, 0), RS - 1);
            sv[u][h] = rsum[off];
            cv[u][h] = rcnt[off];
          }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < KB; ++u) {
          const int kx = kx0 + u;
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int lo = ring_of(h ? -1 : 0, kx), hi = ring_of(h ? -N / 2 : N / 2 - 1, kx);
            if (kx <= kxmax && ring >= lo && ring <= hi) {
              ring_sum += sv[u][h];
              ring_cnt += cv[u][h];
            }
          }
        }
      }
      div[ring] = (ring == 0 || ring_cnt == 0) ? 1.0f : (float)ring_cnt / ring_sum;
    }
    __syncthreads();

    // ---- phase 5: scale by ring divisor, inverse column FFT (N) -----------
    for (int k0 = 0; k0 < MH; k0 += T::TK3) {
      const int nk = min(T::TK3, MH - k0);
      constexpr int IT5 = (T::TK3 * N + kThreads - 1) / kThreads;
      batched_tile_loop<IT5, 4>(
          [&](int it) {
            return at(W, (unsigned)k0 * N + (unsigned)min(tid + it * kThreads, T::TK3 * N - 1));
          },
          [&](int it, float2 v) {
            int ky, q;
            tile_yq(tid + it * kThreads, ky, q);
            if (ky < N && q < nk) {
              const int fy = ky < N / 2 ? ky : ky - N;
              const float g = div[ring_of(fy, k0 + q)];
              buf[q * SP1 + pd(ky)] = make_float2(v.x * g, v.y * g);
            }
          });
      __syncthreads();
      fft_lds_inplace<N, true, kPad>(buf, nk, SP1, tw1, tid, nthr);
      for (int t = tid; t < T::TK3 * N; t += nthr) {
        int Y, q;
        tile_yq(t, Y, q);
        if (q >= nk) continue;
        at(W, (unsigned)k0 * N + (unsigned)t) = buf[q * SP1 + pd(Y)];
      }
      __syncthreads();
    }
    }

    // ---- phase 6: complex-to-real rows (N) -> real image + plane sums ----
    double acc[4] = {0.0, 0.0, 0.0, 0.0};  // sum z, sum z^2, sum Y*z, sum X*z
    for (int y0 = 0; y0 < N; y0 += 2 * T::RB6) {
      const int nrow = min(2 * T::RB6, N - y0);
      const int nseq = (nrow + 1) / 2;
      constexpr int IT6 = (T::RB6 * N + kThreads - 1) / kThreads;
      batched_tile_loop<IT6, 4>(
          [&](int it) {
            const int t = tid + it * kThreads;
            int q, k;
            divmod_small<N, (unsigned)(((T::RB6 * N + kThreads - 1) / kThreads + 3) / 4 * 4 * kThreads)>(t, q, k);
            const int ya = min(y0 + 2 * q, N - 1), yb = min(y0 + 2 * q + 1, N - 1);
            const int ks = k <= N / 2 ? k : N - k;
            const float2 a = at(W, widx(ya, ks)), b = at(W, widx(yb, ks));
            return make_float4(a.x, a.y, b.x, b.y);
          },
          [&](int it, float4 ab) {
            const int t = tid + it * kThreads;
            int q, k;
            divmod_small<N, (unsigned)(((T::RB6 * N + kThreads - 1) / kThreads + 3) / 4 * 4 * kThreads)>(t, q, k);
            const int ks = k <= N / 2 ? k : N - k;
            float2 a = make_float2(ab.x, ab.y), b = make_float2(ab.z, ab.w);
            if (2 * q + 1 >= nrow) b = make_float2(0.f, 0.f);
            if (ks == 0 || ks == N / 2) { a.y = 0.f; b.y = 0.f; }
            if (k > N / 2) { a.y = -a.y; b.y = -b.y; }
            if (q < nseq) buf[q * SP1 + pd(k)] = make_float2(a.x - b.y, a.y + b.x);
          });
      __syncthreads();
      fft_lds_inplace<N, true, kPad>(buf, nseq, SP1, tw1, tid, nthr);
      const float sc = 1.0f / ((float)N * (float)N);
      for (int t = tid; t < nseq * N; t += nthr) {
        int q, X;
        divmod_small<N, (unsigned)(T::RB6 * N)>(t, q, X);
        const float2 v = buf[q * SP1 + pd(X)];
        const int ya = y0 + 2 * q;
        const float za = v.x * sc;
        at(Z, (unsigned)(ya * N + X)) = za;
        acc[0] += za; acc[1] += (double)za * za; acc[2] += (double)ya * za; acc[3] += (double)X * za;
        if (ya + 1 < y0 + nrow) {
          const float zb = v.y * sc;
          at(Z, (unsigned)((ya + 1) * N + X)) = zb;
          acc[0] += zb; acc[1] += (double)zb * zb; acc[2] += (double)(ya + 1) * zb; acc[3] += (double)X * zb;
        }
      }
      __syncthreads();
    }

    // ---- phase 7: least-squares plane, mean, std (float64) ---------------
    block_sum_f64<4>(acc, red);
    double pa, pb, pc, mean, inv_std;
    {
      const double nn = (double)N;
      const double s1 = nn * (nn * (nn - 1.0) / 2.0);                       // sum of a coordinate
      const double s2 = nn * ((nn - 1.0) * nn * (2.0 * nn - 1.0) / 6.0);   // sum of its square
      const double sxy = (nn * (nn - 1.0) / 2.0) * (nn * (nn - 1.0) / 2.0);
      const double cnt = nn * nn;
      // normal equations [[s2, sxy, s1], [sxy, s2, s1], [s1, s1, cnt]] * (a, b, c) = (Sxz, Syz, Sz)
      double A[3][4] = {{s2, sxy, s1, acc[2]}, {sxy, s2, s1, acc[3]}, {s1, s1, cnt, acc[0]}};
      for (int c = 0; c < 3; ++c) {
        int p = c;
        for (int r = c + 1; r < 3; ++r)
          if (fabs(A[r][c]) > fabs(A[p][c])) p = r;
        if (p != c)
          for (int j = 0; j < 4; ++j) { double tt = A[c][j]; A[c][j] = A[p][j]; A[p][j] = tt; }
        for (int r = c + 1; r < 3; ++r) {
          const double f = A[r][c] / A[c][c];
          for (int j = c; j < 4; ++j) A[r][j] -= f * A[c][j];
        }
      }
      pc = A[2][3] / A[2][2];
      pb = (A[1][3] - A[1][2] * pc) / A[1][1];
      pa = (A[0][3] - A[0][1] * pb - A[0][2] * pc) / A[0][0];
      // residual statistics from the moments
      const double sz = acc[0], szz = acc[1], sxz = acc[2], syz = acc[3];
      const double sum_res = sz - pa * s1 - pb * s1 - pc * cnt;
      const double sum_plane2 = pa * pa * s2 + pb * pb * s2 + pc * pc * cnt + 2.0 * pa * pb * sxy +
                                2.0 * pa * pc * s1 + 2.0 * pb * pc * s1;
      const double sum_res2 = szz - 2.0 * (pa * sxz + pb * syz + pc * sz) + sum_plane2;
      mean = sum_res / cnt;
      double var = sum_res2 / cnt - mean * mean;
      var = var > 0.0 ? var : 0.0;
      const double sd = sqrt(var);
      inv_std = sd > (double)io.eps ? 1.0 / sd : 1.0;
    }
    const float fa = (float)pa, fb = (float)pb, fc = (float)pc, fmean = (float)mean;
    const float fis = (float)inv_std;

    if (io.proj) {
      // real-space output: Z, or (Z - plane - mean) / std
      float* po = io.proj + (size_t)d * N * N;
      for (int t = tid; t < N * N; t += nthr) {
        int y, x;
        divmod_small<N, (unsigned)(N * N)>(t, y, x);
        const float z = Z[t];
        po[t] = io.normz ? ((z - ((fa * (float)y + fb * (float)x) + fc)) - fmean) * fis : z;
      }
      continue;
    }

    // ---- phase 8: normalise, mask, forward row FFT (N) -------------------
    for (int y0 = 0; y0 < N; y0 += 2 * T::RB6) {
      const int nrow = min(2 * T::RB6, N - y0);
      const int nseq = (nrow + 1) / 2;
      constexpr int IT8 = (T::RB6 * N + kThreads - 1) / kThreads;
      batched_tile_loop<IT8, 4>(
          [&](int it) {
            const int t = tid + it * kThreads;
            int q, X;
            divmod_small<N, (unsigned)(((T::RB6 * N + kThreads - 1) / kThreads + 3) / 4 * 4 * kThreads)>(t, q, X);
            const int ya = min(y0 + 2 * q, N - 1), yb = min(y0 + 2 * q + 1, N - 1);
            return make_float4(at(Z, (unsigned)(ya * N + X)), at(Z, (unsigned)(yb * N + X)),
                               mask ? at(mask, (unsigned)(ya * N + X)) : 1.0f,
                               mask ? at(mask, (unsigned)(yb * N + X)) : 1.0f);
          },
          [&](int it, float4 zm) {
            const int t = tid + it * kThreads;
            int q, X;
            divmod_small<N, (unsigned)(((T::RB6 * N + kThreads - 1) / kThreads + 3) / 4 * 4 * kThreads)>(t, q, X);
            const int ya = y0 + 2 * q, yb = ya + 1;
            float va = ((zm.x - ((fa * (float)ya + fb * (float)X) + fc)) - fmean) * fis;
            if (mask) va *= zm.z;
            float vb = 0.f;
            if (2 * q + 1 < nrow) {
              vb = ((zm.y - ((fa * (float)yb + fb * (float)X) + fc)) - fmean) * fis;
              if (mask) vb *= zm.w;
            }
            if (q < nseq) buf[q * SP1 + pd(X)] = make_float2(va, vb);
          });
      __syncthreads();
      fft_lds_inplace<N, false, kPad>(buf, nseq, SP1, tw1, tid, nthr);
      for (int t = tid; t < nseq * MH; t += nthr) {
        int q, k;
        divmod_small<MH, (unsigned)(T::RB6 * MH)>(t, q, k);
        const float2 z = buf[q * SP1 + pd(k)];
        const float2 zc = cconj(buf[q * SP1 + pd((N - k) % N)]);
        const int ya = y0 + 2 * q;
        at(W, widx(ya, k)) = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y + zc.y));
        if (ya + 1 < y0 + nrow)
          at(W, widx(ya + 1, k)) = make_float2(0.5f * (z.y - zc.y), -0.5f * (z.x - zc.x));
      }
      __syncthreads();
    }

    // ---- phase 9: forward column FFT (N), crop to n, pack, FSC -----------
    const int ldb = co.ldb > 0 ? co.ldb : mh_n;
    const int fnct = frag_nct(n);
    const size_t plane = co.frag ? (size_t)frag_plane(n) : (size_t)n * ldb;
    float2* out = co.kslot ? nullptr : bank + (size_t)w * plane;
    auto bidx = [&](int r, int c) -> size_t { return co.frag ? (size_t)frag_pos(n, fnct, r, c) : (size_t)r * ldb + c; };
    double nrm[1] = {0.0};
    const int kmax = n / 2;  // columns 0..n/2 are needed
    for (int k0 = 0; k0 <= kmax; k0 += T::TK3) {
      const int nk = min(T::TK3, kmax + 1 - k0);
      constexpr int IT9 = (T::TK3 * N + kThreads - 1) / kThreads;
      batched_tile_loop<IT9, 4>(
          [&](int it) {
            return at(W, (unsigned)k0 * N + (unsigned)min(tid + it * kThreads, T::TK3 * N - 1));
          },
          [&](int it, float2 v) {
            int Y, q;
            tile_yq(tid + it * kThreads, Y, q);
            if (Y < N && q < nk) buf[q * SP1 + pd(Y)] = v;
          });
      __syncthreads();
      fft_lds_inplace<N, false, kPad>(buf, nk, SP1, tw1, tid, nthr);
      const float2* res = buf;
      // the FSC weight (and, for coefficient output, the slot) of each packed
      // entry is read ahead in batches; the per-thread norm adds in t order
      constexpr int IT9S = (T::TK3 * N + kThreads - 1) / kThreads;  // covers TK3 * n (n <= N)
      batched_tile_loop<IT9S, 4>(
          [&](int it) {
            int r, c;
            tile_yq(tid + it * kThreads, r, c);
            r = min(r, n - 1);
            c = min(k0 + c, mh_n - 1);
            const float g = fsc ? at(fsc, (unsigned)(r * mh_n + c)) : 1.0f;
            const int j = co.kslot ? at(co.kslot, (unsigned)(r * mh_n + c)) : -1;
            return make_float2(g, __int_as_float(j));
          },
          [&](int it, float2 gj) {
            const int t = tid + it * kThreads;
            int r, q;
            tile_yq(t, r, q);
            if (t >= T::TK3 * n || q >= nk) return;
            const int c = k0 + q;
            const int fy = r < n / 2 ? r : r - n;
            float2 v;
            if (c < n / 2 || n == N) {
              v = res[q * SP1 + pd(fy < 0 ? fy + N : fy)];
            } else {  // packed column n/2 holds frequency -n/2 = conj of (+n/2) at -fy
              const int my = -fy;
              v = cconj(res[q * SP1 + pd(my < 0 ? my + N : my)]);
            }
            v.x *= gj.x;
            v.y *= gj.x;
            if (co.kslot) {
              const int j = __float_as_int(gj.y);
              if (j >= 0) co.coeffs[(size_t)w * co.Kpad + j] = v;
            } else {
              out[bidx(r, c)] = v;
            }
            nrm[0] += (double)v.x * v.x + (double)v.y * v.y;
          });
      __syncthreads();
    }

    if (out && co.frag && plane > (size_t)n * mh_n) {  // fragment order: zeros past row n / column n/2
      const int tiles = frag_nks(n) * fnct * 256;
      for (int e = tid; e < (int)plane; e += nthr) {
        int r, c;
        frag_rc(n, fnct, tiles, e, r, c);
        if (r >= n || c >= mh_n) out[e] = make_float2(0.f, 0.f);
      }
    } else if (out && !co.frag && ldb > mh_n) {  // zero padding past column n/2 (read by searches as a clamp target)
      const int pw = ldb - mh_n;
      for (int t = tid; t < n * pw; t += nthr) {
        const int r = t / pw, c = t - r * pw;
        out[(size_t)r * ldb + mh_n + c] = make_float2(0.f, 0.f);
      }
    }

    // ---- phase 10: unit L2 norm over the packed half plane (or, when the
    // caller applies it later, only the norm) ------------------------------
    block_sum_f64<1>(nrm, red);
    const float norm = (float)sqrt(nrm[0]);
    if (co.kslot) {
      // the phase-9 stores of every thread are visible after block_sum_f64's
      // barriers; 1/norm rounded once, as the reciprocal a gather would apply
      const float sc = 1.0f / norm;
      float2* cr = co.coeffs + (size_t)w * co.Kpad;
      for (int j0 = tid; j0 < co.Kpad; j0 += 4 * nthr) {  // loads in batches of 4, clamped into the row
        float2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = cr[co.K > 0 ? min(j0 + u * nthr, co.K - 1) : 0];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int j = j0 + u * nthr;
          if (j < co.Kpad) cr[j] = j < co.K ? make_float2(v[u].x * sc, v[u].y * sc) : make_float2(0.f, 0.f);
        }
      }
    } else if (normalize) {
      for (int t = tid; t < n * mh_n; t += nthr) {
        const int r = t / mh_n, c = t - r * mh_n;
        float2 v = out[bidx(r, c)];
        out[bidx(r, c)] = make_float2(v.x / norm, v.y / norm);
      }
    }
    if (norms && tid == 0) norms[w] = norm;
  }
}

// Workgroups of build_templates_kernel<N, L> resident on the current device
// at once (occupancy per CU x CUs; cached per device).
template <int N, int L>
int resident_workgroups() {
  static int per_cu[64] = {0};
  int dev = 0;
  (void)hipGetDevice(&dev);
  int* pc = (dev >= 0 && dev < 64) ? &per_cu[dev] : nullptr;
  if (pc && *pc == 0) {
    int nb = 0;
    const size_t lds = (size_t)kLdsFloat2 * sizeof(float2);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, build_templates_kernel<N, L>, kThreads, lds) != hipSuccess ||
        nb <= 0) {
      (void)hipGetLastError();
      nb = 4;
    }
    *pc = nb;
  }
  return (pc ? *pc : 4) * gis_cu_count();
}

template <int N, int L>
int launch_build_templates(const float2* D, VolSet vs, const float* rts, int M, int n, const float* mask,
                           const float* fsc, float2* bank, float* norms, int normalize, void* ws, size_t ws_bytes,
                           hipStream_t st, CoefOut co, TplIO io) {
  using T = TplLayout<N, L>;
  const size_t slot = T::SLOT * sizeof(float2);
  const size_t lds = (size_t)kLdsFloat2 * sizeof(float2);
  // a persistent grid of exactly the resident workgroups: a workgroup beyond
  // them would start a full direction late and finish a round after the rest
  const int resident = resident_workgroups<N, L>();
  int grid = (int)std::min<size_t>((size_t)vs.C * (size_t)M, ws_bytes / slot);
  grid = std::min(grid, resident);
  GIS_REQUIRE(grid >= 1, "gis_build_templates: workspace of %zu bytes is below one slot (%zu)", ws_bytes, slot);
  static bool attr_set = false;
  if (lds > 65536 && !attr_set) {
    hipError_t e = hipFuncSetAttribute((const void*)build_templates_kernel<N, L>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      gis_set_error("gis_build_templates: cannot reserve %zu bytes of LDS: %s", lds, hipGetErrorString(e));
      return (int)e;
    }
    attr_set = true;
  }
  hipLaunchKernelGGL((build_templates_kernel<N, L>), dim3(grid), dim3(kThreads), lds, st, D, vs, rts, M, n, mask,
                     fsc, bank, norms, (float2*)ws, normalize, co, io);
  GIS_LAUNCH_CHECK("gis_build_templates");
  return GIS_OK;
}

template <int N, int L>
size_t tpl_slot_bytes() { return TplLayout<N, L>::SLOT * sizeof(float2); }

}  // namespace

static size_t slot_bytes_for(int N, int pad) {
#define GIS_TPL_SLOT(NV) \
  case NV: return pad ? tpl_slot_bytes<NV, 2 * NV>() : tpl_slot_bytes<NV, NV>();
  switch (N) {
    GIS_TPL_SLOT(32)
    GIS_TPL_SLOT(64)
    GIS_TPL_SLOT(96)
    GIS_TPL_SLOT(128)
    GIS_TPL_SLOT(192)
    GIS_TPL_SLOT(256)
    default: return 0;
  }
#undef GIS_TPL_SLOT
}

// Workspace bytes that let the template kernel run `M` directions (or images)
// with full parallelism (one scratch slot per resident workgroup), for the
// padded (pad 1) or unpadded (pad 0) slice transform.
GIS_API size_t gis_build_templates_workspace_ex(int N, int M, int pad) {
  const size_t s = slot_bytes_for(N, pad);
  if (s == 0 || M <= 0) return 0;
  const size_t g = std::min<size_t>((size_t)M, (size_t)4 * gis_cu_count());
  return g * s;
}
GIS_API size_t gis_build_templates_workspace(int N, int M) { return gis_build_templates_workspace_ex(N, M, 1); }

// Template workgroups the current device runs at once for box N (pad 1: the
// padded transform): a launch with fewer directions (C * M) leaves CUs idle.
GIS_API int gis_build_templates_resident(int N, int pad) {
#define GIS_TPL_RES(NV) \
  case NV: return pad ? resident_workgroups<NV, 2 * NV>() : resident_workgroups<NV, NV>();
  switch (N) {
    GIS_TPL_RES(32)
    GIS_TPL_RES(64)
    GIS_TPL_RES(96)
    GIS_TPL_RES(128)
    GIS_TPL_RES(192)
    GIS_TPL_RES(256)
    default:
      gis_set_error("gis_build_templates_resident: box %d not in {32, 64, 96, 128, 192, 256}", N);
      return GIS_EUNSUPPORTED;
  }
#undef GIS_TPL_RES
}

// Replaces the per-direction loop of IntegratedProjector.build_templates
// (v606:1077-1147).
//   dft   [2N][2N][N+1] complex64 from gis_volume_dft
//   rts   [M][3][3] float32 rotation matrices R = Euler(rot, tilt, 0)^T
//   mask  [N][N] float32 real-space mask, or NULL (--discardMask)
//   fsc   [n][n/2+1] float32 packed FSC weights, or NULL (all ones)
//   bank  [M][n][n/2+1] complex64 output (unit L2 norm, UNshifted rfft2 layout)
//   norms [M] float32 output of the pre-normalisation norms, or NULL
static int build_templates_dispatch(const void* dft, VolSet vs, int N, int pad, int n, const float* rts, int M,
                                    const float* mask, const float* fsc, void* bank, float* norms, int normalize,
                                    void* ws, size_t ws_bytes, void* stream, CoefOut co, TplIO io) {
  hipStream_t st = (hipStream_t)stream;
  const float2* D = (const float2*)dft;
  float2* B = (float2*)bank;
#define GIS_TPL_CASE(NV)                                                                                         \
  case NV:                                                                                                       \
    return pad ? launch_build_templates<NV, 2 * NV>(D, vs, rts, M, n, mask, fsc, B, norms, normalize, ws,        \
                                                    ws_bytes, st, co, io)                                        \
               : launch_build_templates<NV, NV>(D, vs, rts, M, n, mask, fsc, B, norms, normalize, ws, ws_bytes, \
                                                st, co, io);
  switch (N) {
    GIS_TPL_CASE(32)
    GIS_TPL_CASE(64)
    GIS_TPL_CASE(96)
    GIS_TPL_CASE(128)
    GIS_TPL_CASE(192)
    GIS_TPL_CASE(256)
    default:
      gis_set_error("gis_build_templates: box %d not in {32, 64, 96, 128, 192, 256}", N);
      return GIS_EUNSUPPORTED;
  }
#undef GIS_TPL_CASE
}

static const TplIO kBankIO = {nullptr, nullptr, 1, 1, 1e-12f, nullptr};
static const VolSet kOneVolume = {1, 0};

GIS_API int gis_build_templates(const void* dft, int N, int n, const float* rts, int M, const float* mask,
                                const float* fsc, void* bank, float* norms, int normalize, void* ws,
                                size_t ws_bytes, void* stream) {
  GIS_REQUIRE(dft && rts && bank && ws, "gis_build_templates: null pointer");
  GIS_REQUIRE(normalize || norms, "gis_build_templates: an unnormalised bank needs the norms output");
  GIS_REQUIRE(M >= 0, "gis_build_templates: M < 0");
  GIS_REQUIRE(n >= 2 && n <= N && n % 2 == 0, "gis_build_templates: newbox %d must be even and <= box %d", n, N);
  if (M == 0) return GIS_OK;
  return build_templates_dispatch(dft, kOneVolume, N, 1, n, rts, M, mask, fsc, bank, norms, normalize, ws, ws_bytes, stream,
                                  CoefOut{nullptr, 0, 0, nullptr}, kBankIO);
}

// The template bank reduced to the listed coefficients of every direction
// (the small-shift search's template operand): same transform as
// gis_build_templates, but direction d writes only
//   coeffs[d * Kpad + kslot[r * (n/2+1) + c]] = bank[d][r][c] / norm_d
// for the packed entries with kslot >= 0 (slots [0, K)), zeros in [K, Kpad),
// and norms[d] (optional), without storing the bank.
GIS_API int gis_build_template_coefficients(const void* dft, int N, int n, const float* rts, int M,
                                            const float* mask, const float* fsc, const int* kslot, int K, int Kpad,
                                            void* coeffs, float* norms, void* ws, size_t ws_bytes, void* stream) {
  GIS_REQUIRE(dft && rts && kslot && coeffs && ws, "gis_build_template_coefficients: null pointer");
  GIS_REQUIRE(M >= 0, "gis_build_template_coefficients: M < 0");
  GIS_REQUIRE(K >= 0 && Kpad >= K, "gis_build_template_coefficients: K=%d Kpad=%d", K, Kpad);
  GIS_REQUIRE(n >= 2 && n <= N && n % 2 == 0, "gis_build_template_coefficients: newbox %d must be even and <= box %d",
              n, N);
  if (M == 0) return GIS_OK;
  return build_templates_dispatch(dft, kOneVolume, N, 1, n, rts, M, mask, fsc, nullptr, norms, 0, ws, ws_bytes, stream,
                                  CoefOut{kslot, K, Kpad, (float2*)coeffs}, kBankIO);
}

// Several candidates' banks in one launch: C spectra at dft + c * dft_stride
// (complex64 values, each as from gis_volume_dft), the same M rotations, mask
// and FSC; bank row c * M + d (and norms[c * M + d]) for direction d of
// spectrum c.  One launch keeps every workgroup busy across the candidates
// (no per-launch tail) and lets one workgroup's slice gather overlap another's
// transforms.  ws: gis_build_templates_workspace(N, C * M).
GIS_API int gis_build_templates_batched(const void* dft, size_t dft_stride, int C, int N, int n, const float* rts,
                                        int M, const float* mask, const float* fsc, void* bank, float* norms,
                                        int normalize, void* ws, size_t ws_bytes, void* stream) {
  GIS_REQUIRE(dft && rts && bank && ws, "gis_build_templates_batched: null pointer");
  GIS_REQUIRE(normalize || norms, "gis_build_templates_batched: an unnormalised bank needs the norms output");
  GIS_REQUIRE(M >= 0 && C >= 0, "gis_build_templates_batched: M = %d, C = %d", M, C);
  GIS_REQUIRE((size_t)C * (size_t)M < (1u << 31), "gis_build_templates_batched: C * M too large");
  GIS_REQUIRE(n >= 2 && n <= N && n % 2 == 0, "gis_build_templates_batched: newbox %d must be even and <= box %d", n,
              N);
  if (M == 0 || C == 0) return GIS_OK;
  return build_templates_dispatch(dft, VolSet{C, dft_stride}, N, 1, n, rts, M, mask, fsc, bank, norms, normalize, ws,
                                  ws_bytes, stream, CoefOut{nullptr, 0, 0, nullptr}, kBankIO);
}

// gis_build_templates_batched with bank rows ldb >= n/2+1 complex values
// apart (columns past n/2 zero): the padded row layout the searches read in
// aligned 128-byte segments; ldb = 0: the planes in fragment order
// (frag_pos, frag_plane(n) complex values each), read with 16-byte loads.
GIS_API int gis_build_templates_batched_ld(const void* dft, size_t dft_stride, int C, int N, int n,
                                           const float* rts, int M, const float* mask, const float* fsc, void* bank,
                                           int ldb, float* norms, int normalize, void* ws, size_t ws_bytes,
                                           void* stream) {
  GIS_REQUIRE(dft && rts && bank && ws, "gis_build_templates_batched_ld: null pointer");
  GIS_REQUIRE(normalize || norms, "gis_build_templates_batched_ld: an unnormalised bank needs the norms output");
  GIS_REQUIRE(M >= 0 && C >= 0, "gis_build_templates_batched_ld: M = %d, C = %d", M, C);
  GIS_REQUIRE((size_t)C * (size_t)M < (1u << 31), "gis_build_templates_batched_ld: C * M too large");
  GIS_REQUIRE(n >= 2 && n <= N && n % 2 == 0, "gis_build_templates_batched_ld: newbox %d must be even and <= box %d",
              n, N);
  GIS_REQUIRE(ldb == 0 || ldb >= n / 2 + 1, "gis_build_templates_batched_ld: row stride %d below n/2+1 = %d", ldb,
              n / 2 + 1);
  if (M == 0 || C == 0) return GIS_OK;
  CoefOut co{nullptr, 0, 0, nullptr};
  co.ldb = ldb;
  co.frag = ldb == 0;
  return build_templates_dispatch(dft, VolSet{C, dft_stride}, N, 1, n, rts, M, mask, fsc, bank, norms, normalize, ws,
                                  ws_bytes, stream, co, kBankIO);
}

// gis_build_template_coefficients over C spectra in one launch (layout as
// gis_build_templates_batched): coeffs row c * M + d.
GIS_API int gis_build_template_coefficients_batched(const void* dft, size_t dft_stride, int C, int N, int n,
                                                    const float* rts, int M, const float* mask, const float* fsc,
                                                    const int* kslot, int K, int Kpad, void* coeffs, float* norms,
                                                    void* ws, size_t ws_bytes, void* stream) {
  GIS_REQUIRE(dft && rts && kslot && coeffs && ws, "gis_build_template_coefficients_batched: null pointer");
  GIS_REQUIRE(M >= 0 && C >= 0, "gis_build_template_coefficients_batched: M = %d, C = %d", M, C);
  GIS_REQUIRE((size_t)C * (size_t)M < (1u << 31), "gis_build_template_coefficients_batched: C * M too large");
  GIS_REQUIRE(K >= 0 && Kpad >= K, "gis_build_template_coefficients_batched: K=%d Kpad=%d", K, Kpad);
  GIS_REQUIRE(n >= 2 && n <= N && n % 2 == 0,
              "gis_build_template_coefficients_batched: newbox %d must be even and <= box %d", n, N);
  if (M == 0 || C == 0) return GIS_OK;
  return build_templates_dispatch(dft, VolSet{C, dft_stride}, N, 1, n, rts, M, mask, fsc, nullptr, norms, 0, ws,
                                  ws_bytes, stream, CoefOut{kslot, K, Kpad, (float2*)coeffs}, kBankIO);
}

// Real-space central-slice projections of a volume (the standalone
// projector, project3d_and_whiten_cuda.py:119-147): for every direction the
// cropped projection [N][N]; with `whiten` the ring division of its spectrum
// (getSpectrum_divideBySpectrum_torch, :196-230) and with `normz` the plane
// fit + (x - mean) / std (normalize_torch, :231-291, std threshold eps).
//   dft  [L][L][L/2+1] from gis_volume_dft_ex(.., pad, ..), L = pad ? 2N : N
//   proj [M][N][N] float32 out;  ws: gis_build_templates_workspace_ex(N, M, pad)
GIS_API int gis_project_slices(const void* dft, int N, int pad, const float* rts, int M, int whiten, int normz,
                               float eps, float* proj, void* ws, size_t ws_bytes, void* stream) {
  GIS_REQUIRE(dft && rts && proj && ws, "gis_project_slices: null pointer");
  GIS_REQUIRE(M >= 0, "gis_project_slices: M < 0");
  GIS_REQUIRE(pad == 0 || pad == 1, "gis_project_slices: pad %d not 0 or 1", pad);
  if (M == 0) return GIS_OK;
  return build_templates_dispatch(dft, kOneVolume, N, pad, N, rts, M, nullptr, nullptr, nullptr, nullptr, 0, ws, ws_bytes,
                                  stream, CoefOut{nullptr, 0, 0, nullptr}, TplIO{nullptr, proj, whiten, normz, eps, nullptr});
}

// The same ring division and / or plane normalisation applied to M real
// [N][N] images (getSpectrum_divideBySpectrum_torch and normalize_torch,
// project3d_and_whiten_cuda.py:196-291).  At least one of whiten / normz.
//   img, out [M][N][N] float32;  ws: gis_build_templates_workspace(N, M)
GIS_API int gis_whiten_images(const float* img, int N, int M, int whiten, int normz, float eps, float* out,
                              void* ws, size_t ws_bytes, void* stream) {
  GIS_REQUIRE(img && out && ws, "gis_whiten_images: null pointer");
  GIS_REQUIRE(M >= 0, "gis_whiten_images: M < 0");
  GIS_REQUIRE(whiten || normz, "gis_whiten_images: nothing to do (whiten and normz both 0)");
  if (M == 0) return GIS_OK;
  return build_templates_dispatch(nullptr, kOneVolume, N, 1, N, nullptr, M, nullptr, nullptr, nullptr, nullptr, 0, ws, ws_bytes,
                                  stream, CoefOut{nullptr, 0, 0, nullptr}, TplIO{img, out, whiten, normz, eps, nullptr});
}

// The (N, pad) instantiation that runs slices of transform length L.
static bool slice_config(int L, int* N, int* pad) {
  const int boxes[6] = {32, 64, 96, 128, 192, 256};
  for (int b : boxes)
    if (2 * b == L) { *N = b; *pad = 1; return true; }
  for (int b : boxes)
    if (b == L) { *N = b; *pad = 0; return true; }
  return false;
}

GIS_API size_t gis_extract_slices_workspace(int L, int M) {
  int N = 0, pad = 0;
  return slice_config(L, &N, &pad) ? gis_build_templates_workspace_ex(N, M, pad) : 0;
}

// extract_central_slices_rfft (project3d_and_whiten_cuda.py:397-435; the
// gather of v606:1099-1113): trilinear, border-clamped samples of the half
// spectrum at M rotated central slices, conjugated where the rotated x < 0.
//   dft    [L][L][L/2+1] bricked (gis_volume_dft_ex, or a natural-order
//          spectrum rearranged into the same bricks), L in {32..512}
//   rts    [M][3][3] float32, rotation applied to (x, y, z) frequency vectors
//   slices [M][L][L/2+1] complex64 out, rows in fftshifted order (the
//          reference's output layout);  ws: gis_extract_slices_workspace(L, M)
GIS_API int gis_extract_slices(const void* dft, int L, const float* rts, int M, void* slices, void* ws,
                               size_t ws_bytes, void* stream) {
  GIS_REQUIRE(dft && rts && slices && ws, "gis_extract_slices: null pointer");
  GIS_REQUIRE(M >= 0, "gis_extract_slices: M < 0");
  int N = 0, pad = 0;
  if (!slice_config(L, &N, &pad)) {
    gis_set_error("gis_extract_slices: transform length %d not a supported box or twice one", L);
    return GIS_EUNSUPPORTED;
  }
  if (M == 0) return GIS_OK;
  return build_templates_dispatch(dft, kOneVolume, N, pad, N, rts, M, nullptr, nullptr, nullptr, nullptr, 0, ws, ws_bytes, stream,
                                  CoefOut{nullptr, 0, 0, nullptr},
                                  TplIO{nullptr, nullptr, 0, 0, 0.f, (float2*)slices});
}
