// Compiled register-engine variants, FP32, for the 2^a 3^b 5^c sizes of
// the serving range [64, 512] that have no power-of-two schedule (N = 240 .. 500).
// Per size: points per thread P (TPT = N/P threads per transform) keeps the
// idle butterfly slots of stages whose radix does not divide P small and the
// register slots at <= 16; column tiles are 16 columns (128 B per row
// segment) or wider so that a workgroup has >= 128 threads; z tiles hold G
// row pairs with G * TPT threads and <= 64 KB of LDS.
#include "gfft_schedules.h"

namespace gfft {

int registry_smooth_f32_b(const KernelEntry** out) {
  static const KernelEntry table[] = {
      //        type   schedule  P_rows G  P_cols CW
      GFFT_SIZE(float, Sc240, 15, 8, 15, 16),
      GFFT_SIZE(float, Sc243, 9, 18, 9, 16),
      GFFT_SIZE(float, Sc250, 10, 20, 10, 16),
      GFFT_SIZE(float, Sc270, 15, 26, 15, 16),
      GFFT_SIZE(float, Sc288, 12, 8, 12, 16),
      GFFT_SIZE(float, Sc300, 15, 16, 15, 16),
      GFFT_SIZE(float, Sc320, 8, 8, 8, 16),
      GFFT_SIZE(float, Sc324, 12, 18, 12, 16),
      GFFT_SIZE(float, Sc360, 15, 8, 15, 16),
      GFFT_SIZE(float, Sc375, 15, 19, 15, 16),
      GFFT_SIZE(float, Sc400, 10, 8, 10, 16),
      GFFT_SIZE(float, Sc405, 15, 17, 15, 16),
      GFFT_SIZE(float, Sc432, 12, 14, 12, 16),
      GFFT_SIZE(float, Sc450, 15, 16, 15, 16),
      GFFT_SIZE(float, Sc480, 15, 4, 15, 16),
      GFFT_SIZE(float, Sc486, 9, 9, 9, 16),
      GFFT_SIZE(float, Sc500, 10, 10, 10, 16),
  };
  *out = table;
  return (int)(sizeof(table) / sizeof(table[0]));
}

}  // namespace gfft
