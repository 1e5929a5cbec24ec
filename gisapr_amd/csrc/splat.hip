// Atom -> voxel radial-exponential splat (density_from_coords).
//
// Rule (pdb2mrc_gpu_ver_fp32_v3.py:41-98; the zip release's CUDA deposit
// kernel, gisapr_cuda_density.py:20-56):
//   pos   = float(double(xyz) / apix + box/2)
//   lower = round_half_away(double(pos - width)), upper likewise with +width
//   voxels v in [max(lower,0), min(upper, box-1))   (exclusive upper)
//   r     = sqrtf((dx*dx + dy*dy) + dz*dz),  dx = float(v) - pos
//   value = float( double(Z)*kn * double(expf(float(-double(r)*rp))) )
// The volume is indexed x + box*(y + box*z), i.e. [z][y][x].
//
// gfx950 mapping: one 64-lane wave per atom, the (2 width)^3 voxel window is
// strided over the lanes and accumulated with hardware float atomics
// (global_atomic_add_f32) into the caller's volume, which the caller either
// zeroes or pre-fills with a resident fixed-body density.
#include "common.h"

#pragma clang fp contract(off)

__global__ void __launch_bounds__(256)
splat_wave_per_atom(const float* __restrict__ xyz, const float* __restrict__ numbers, int n_atoms,
                    int box, int width, double apix, double rp, double kn, float* __restrict__ vol) {
  const int atom = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (atom >= n_atoms) return;
  const float zn = numbers[atom];
  if (zn == 0.0f) return;
  float pos[3];
  int lo[3], span[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float p = (float)((double)xyz[3 * atom + k] / apix + (double)(box / 2));
    pos[k] = p;
    const float l = p - (float)width, u = p + (float)width;
    const double lower = copysign(floor(fabs((double)l) + 0.5), (double)l);
    const double upper = copysign(floor(fabs((double)u) + 0.5), (double)u);
    const double a = fmax(0.0, fmin((double)box, lower));
    const double b = fmax(0.0, fmin((double)(box - 1), upper));
    lo[k] = (int)a;
    span[k] = (int)fmax(0.0, b - a);
  }
  const int count = span[0] * span[1] * span[2];
  const double zkn = (double)zn * kn;
  for (int i = lane; i < count; i += 64) {
    const int x = lo[0] + i % span[0];
    const int yz = i / span[0];
    const int y = lo[1] + yz % span[1];
    const int z = lo[2] + yz / span[1];
    const float dx = (float)x - pos[0], dy = (float)y - pos[1], dz = (float)z - pos[2];
    const float dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
    const float r = sqrtf((dx2 + dy2) + dz2);
    const float e = (float)(-(double)r * rp);
    const float v = (float)(zkn * (double)expf(e));
    unsafeAtomicAdd(vol + (size_t)x + (size_t)box * ((size_t)y + (size_t)box * (size_t)z), v);
  }
}

// Replaces deposit(xyz, numbers, volume, n, box, width, apix, rp, kn)
// (gisapr_cuda_density.py:21-22) and the index_add_ loop of
// density_from_coords (pdb2mrc_gpu_ver_fp32_v3.py:84-97).
GIS_API int gis_splat(const float* xyz, const float* numbers, int n_atoms, int box, int width,
                      double apix, double rp, double kn, float* volume, void* stream) {
  GIS_REQUIRE(n_atoms >= 0, "gis_splat: n_atoms < 0");
  GIS_REQUIRE(box >= 1, "gis_splat: box < 1");
  GIS_REQUIRE(width >= 3 && width <= 64, "gis_splat: width %d outside [3, 64]", width);
  GIS_REQUIRE(apix > 0 && rp > 0 && kn > 0, "gis_splat: non-positive apix/rp/kn");
  GIS_REQUIRE(volume != nullptr, "gis_splat: null volume");
  if (n_atoms == 0) return GIS_OK;
  GIS_REQUIRE(xyz != nullptr && numbers != nullptr, "gis_splat: null atom arrays");
  const int blocks = (n_atoms + 3) / 4;
  hipLaunchKernelGGL(splat_wave_per_atom, dim3(blocks), dim3(256), 0, (hipStream_t)stream, xyz, numbers,
                     n_atoms, box, width, apix, rp, kn, volume);
  GIS_LAUNCH_CHECK("gis_splat");
  return GIS_OK;
}

// Moving-chain coordinates of `poses` rigid poses (pose p, atom a, axis j),
// computed as the host does (pdb_text.py:123-142, then the PDB 8.3 write /
// re-read round trip, Z3 pdb_text.py:179-198):
//   d   = xyz[a] - centre                                    (float32)
//   v_j = fma(d2, R[2][j], fma(d1, R[1][j], d0 R[0][j])) + centre_j + t_j
//   out = float(k / 1000), k = the integer nearest v_j * 1000 (ties to even)
// The product chain is the one NumPy's float64 matmul of (x - c) @ R gives
// (OpenBLAS: one multiply, then fused multiply-adds along the 3-term dot);
// the nearest integer is decided on the exact product v_j * 1000, as
// float(f"{v:8.3f}") does: rint of the rounded product, then the exact sign
// of v * 1000 - (k +- 1/2) from one fma each (a single rounding keeps the
// sign, and the result is zero only at an exact tie).
__device__ __forceinline__ double round_8_3(double v) {
#pragma clang fp contract(off)
  double k = rint(v * 1000.0);
  const double up = fma(v, 1000.0, -(k + 0.5));
  const double dn = fma(v, 1000.0, -(k - 0.5));
  if (up > 0.0) {
    k += 1.0;
  } else if (up == 0.0) {
    if (fmod(k, 2.0) != 0.0) k += 1.0;  // tie between k and k + 1
  } else if (dn < 0.0) {
    k -= 1.0;
  } else if (dn == 0.0) {
    if (fmod(k, 2.0) != 0.0) k -= 1.0;  // tie between k - 1 and k
  }
  return k / 1000.0;
}

__global__ void __launch_bounds__(256)
move_chain_kernel(const float* __restrict__ xyz, int n_atoms, float cx, float cy, float cz,
                  const double* __restrict__ rt, int poses, float* __restrict__ out) {
#pragma clang fp contract(off)
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)n_atoms * poses) return;
  const int p = (int)(i / n_atoms), a = (int)(i - (long)p * n_atoms);
  const double* m = rt + 12 * (size_t)p;
  const float c[3] = {cx, cy, cz};
  const double d0 = (double)(xyz[3 * (size_t)a] - cx);
  const double d1 = (double)(xyz[3 * (size_t)a + 1] - cy);
  const double d2 = (double)(xyz[3 * (size_t)a + 2] - cz);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double v = d0 * m[j];
    v = fma(d1, m[3 + j], v);
    v = fma(d2, m[6 + j], v);
    v = v + (double)c[j];
    v = v + m[9 + j];
    out[3 * (size_t)i + j] = (float)round_8_3(v);
  }
}

// out [poses][n_atoms][3] float32: the coordinates xyz [n_atoms][3] moved by
// each pose and 8.3-rounded.  rt [poses][12] float64 device array: the
// rotation matrix R row-major (row vectors: (x - centre) R) then the
// translation t; centre: 3 floats (host), the chain's mean.  Replaces
// transform_coordinates (pdb_text.py:123-142) and the 8.3 round trip of the
// moved chain (Z3 pdb_text.py:179-198) for a batch of candidate poses; the
// caller checks that the results fit the 8.3 field (|R| = 1 bounds them).
GIS_API int gis_move_chain(const float* xyz, int n_atoms, const float* centre, const double* rt, int poses,
                           float* out, void* stream) {
  GIS_REQUIRE(n_atoms >= 0 && poses >= 0, "gis_move_chain: n_atoms %d poses %d", n_atoms, poses);
  if (n_atoms == 0 || poses == 0) return GIS_OK;
  GIS_REQUIRE(xyz && centre && rt && out, "gis_move_chain: null pointer");
  const long total = (long)n_atoms * poses;
  GIS_REQUIRE(total <= (long)INT32_MAX - 255, "gis_move_chain: %ld atom poses", total);
  hipLaunchKernelGGL(move_chain_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     xyz, n_atoms, centre[0], centre[1], centre[2], rt, poses, out);
  GIS_LAUNCH_CHECK("gis_move_chain");
  return GIS_OK;
}
