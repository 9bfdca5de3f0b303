// Compiled register-engine variants, FP64, for the 2^a 3^b 5^c sizes of
// the serving range [64, 512] that have no power-of-two schedule (N = 72 .. 225).
// Per size: points per thread P (TPT = N/P threads per transform) keeps the
// idle butterfly slots of stages whose radix does not divide P small and the
// register slots at <= 16; column tiles are 8 columns (128 B per row
// segment) or wider so that a workgroup has >= 128 threads; z tiles hold G
// row pairs with G * TPT threads and <= 64 KB of LDS.
#include "gfft_schedules.h"

namespace gfft {

int registry_smooth_f64_a(const KernelEntry** out) {
  static const KernelEntry table[] = {
      //        type   schedule  P_rows G  P_cols CW
      GFFT_SIZE(double, Sc72, 6, 16, 6, 16),
      GFFT_SIZE(double, Sc75, 15, 48, 15, 32),
      GFFT_SIZE(double, Sc80, 4, 16, 4, 8),
      GFFT_SIZE(double, Sc81, 9, 44, 9, 16),
      GFFT_SIZE(double, Sc90, 15, 32, 15, 32),
      GFFT_SIZE(double, Sc100, 10, 32, 10, 16),
      GFFT_SIZE(double, Sc108, 12, 33, 12, 16),
      GFFT_SIZE(double, Sc120, 15, 16, 15, 16),
      GFFT_SIZE(double, Sc125, 5, 20, 5, 8),
      GFFT_SIZE(double, Sc135, 9, 26, 9, 16),
      GFFT_SIZE(double, Sc144, 12, 16, 12, 16),
      GFFT_SIZE(double, Sc150, 15, 24, 15, 16),
      GFFT_SIZE(double, Sc162, 6, 18, 6, 8),
      GFFT_SIZE(double, Sc180, 15, 16, 15, 16),
      GFFT_SIZE(double, Sc200, 5, 8, 5, 8),
      GFFT_SIZE(double, Sc216, 6, 14, 6, 8),
      GFFT_SIZE(double, Sc225, 15, 16, 15, 16),
  };
  *out = table;
  return (int)(sizeof(table) / sizeof(table[0]));
}

}  // namespace gfft
