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
// gfx950 mapping: one 64-lane wave per atom, the (<= 13^3) voxel window is
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
