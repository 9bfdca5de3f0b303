// Compiled register-engine variants, FP32, for the 2^a 3^b 5^c sizes of
// the serving range [64, 512] that have no power-of-two schedule (N = 72 .. 225).
// Per size: points per thread P (TPT = N/P threads per transform) keeps the
// idle butterfly slots of stages whose radix does not divide P small and the
// register slots at <= 16; column tiles are 16 columns (128 B per row
// segment) or wider so that a workgroup has >= 128 threads; z tiles hold G
// row pairs with G * TPT threads and <= 64 KB of LDS.
#include "gfft_schedules.h"

namespace gfft {

int registry_smooth_f32_a(const KernelEntry** out) {
  static const KernelEntry table[] = {
      //        type   schedule  P_rows G  P_cols CW
      GFFT_SIZE(float, Sc72, 6, 16, 6, 16),
      GFFT_SIZE(float, Sc75, 15, 64, 15, 32),
      GFFT_SIZE(float, Sc80, 4, 16, 4, 16),
      GFFT_SIZE(float, Sc81, 9, 56, 9, 16),
      GFFT_SIZE(float, Sc90, 15, 32, 15, 32),
      GFFT_SIZE(float, Sc100, 10, 32, 10, 16),
      GFFT_SIZE(float, Sc108, 12, 56, 12, 16),
      GFFT_SIZE(float, Sc120, 15, 16, 15, 16),
      GFFT_SIZE(float, Sc125, 5, 20, 5, 16),
      GFFT_SIZE(float, Sc135, 15, 53, 15, 16),
      GFFT_SIZE(float, Sc144, 12, 16, 12, 16),
      GFFT_SIZE(float, Sc150, 15, 32, 15, 16),
      GFFT_SIZE(float, Sc162, 6, 18, 6, 16),
      GFFT_SIZE(float, Sc180, 15, 16, 15, 16),
      GFFT_SIZE(float, Sc200, 5, 8, 5, 16),
      GFFT_SIZE(float, Sc216, 6, 14, 6, 16),
      GFFT_SIZE(float, Sc225, 15, 32, 15, 16),
  };
  *out = table;
  return (int)(sizeof(table) / sizeof(table[0]));
}

}  // namespace gfft
