// Instantiations of the fused z + y plane kernels (gfft_plane.h): planes
// whose Ny x K half spectra (plus the y exchange's tail) fit in the 160 KB of
// LDS of one workgroup, with even Nz (the z exchange stays inside a wavefront
// only when Nz / (points per thread) divides 64).  PlanePass<T, z schedule,
// z points per thread, y schedule, y points per thread, columns per y step>.
#include "gfft_plane.h"
#include "gfft_schedules.h"

namespace gfft {

int registry_plane(const PlaneEntry** out) {
  static const PlaneEntry table[] = {
      // measured faster than the separate passes (whole 3D round trip,
      // cubes): FP32 64 -20 %, 96 -4 %, 128 -15 %, 160 -16 %; FP64 64 -4 %,
      // 96 -11 %, 128 -9 %.  72, 80, 120, 144, 192 (FP32) and 72, 80 (FP64)
      // were compiled and measured slower or level, and are left out.
      // In-place y exchange with 1024-thread workgroups against the earlier
      // column-scratch kernels (256-512 threads), whole round trips: FP32 64
      // -15 %, 96 -22 %, 128 -4 % (512 threads; 1024 was 4 % slower), 160
      // -8 %; FP64 64 -11 %, 96 -15 %, 128 -12 %.  FP32 192 fits only with the
      // in-place exchange (150.5 KB): 78.9 -> 72.3 us against separate passes;
      // FP32 144 72.5 -> 68.9 us (in place; FP32 80 and FP64 80 in place were
      // slower or level and are left out).
      entry_plane<float, Sc64, 4, Sc64, 4, 64>(),
      entry_plane<float, Sc96, 6, Sc96, 6, 64>(),
      entry_plane<float, Sc128, 8, Sc128, 4, 32>(),
      entry_plane<float, Sc160, 10, Sc160, 5, 32>(),
      entry_plane<float, Sc192, 24, Sc192, 24, 64>(),
      entry_plane<float, Sc144, 36, Sc144, 36, 80>(),
      entry_plane<double, Sc64, 4, Sc64, 4, 64>(),
      entry_plane<double, Sc96, 6, Sc96, 6, 64>(),
      entry_plane<double, Sc128, 8, Sc128, 4, 32>(),
  };
  *out = table;
  return (int)(sizeof(table) / sizeof(table[0]));
}

}  // namespace gfft
