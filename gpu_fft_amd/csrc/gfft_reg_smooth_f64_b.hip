// Compiled register-engine variants, FP64, for the 2^a 3^b 5^c sizes of
// the serving range [64, 512] that have no power-of-two schedule (N = 240 .. 500).
// Per size: points per thread P (TPT = N/P threads per transform) keeps the
// idle butterfly slots of stages whose radix does not divide P small and the
// register slots at <= 16; column tiles are 8 columns (128 B per row
// segment) or wider so that a workgroup has >= 128 threads; z tiles hold G
// row pairs with G * TPT threads and <= 64 KB of LDS.
#include "gfft_schedules.h"

namespace gfft {

int registry_smooth_f64_b(const KernelEntry** out) {
  static const KernelEntry table[] = {
      //        type   schedule  P_rows G  P_cols CW
      GFFT_SIZE(double, Sc240, 15, 8, 15, 8),
      GFFT_SIZE(double, Sc243, 9, 14, 9, 8),
      GFFT_SIZE(double, Sc250, 10, 14, 10, 8),
      GFFT_SIZE(double, Sc270, 9, 13, 9, 8),
      GFFT_SIZE(double, Sc288, 12, 8, 12, 8),
      GFFT_SIZE(double, Sc300, 15, 12, 15, 8),
      GFFT_SIZE(double, Sc320, 8, 8, 8, 8),
      GFFT_SIZE(double, Sc324, 12, 11, 12, 8),
      GFFT_SIZE(double, Sc360, 15, 8, 15, 8),
      GFFT_SIZE(double, Sc375, 15, 9, 15, 8),
      GFFT_SIZE(double, Sc400, 10, 8, 10, 8),
      GFFT_SIZE(double, Sc405, 9, 8, 9, 8),
      GFFT_SIZE(double, Sc432, 12, 8, 12, 8),
      GFFT_SIZE(double, Sc450, 15, 8, 15, 8),
      GFFT_SIZE(double, Sc480, 15, 4, 15, 8),
      GFFT_SIZE(double, Sc486, 6, 6, 6, 8),
      GFFT_SIZE(double, Sc500, 10, 7, 10, 8),
  };
  *out = table;
  return (int)(sizeof(table) / sizeof(table[0]));
}

}  // namespace gfft
