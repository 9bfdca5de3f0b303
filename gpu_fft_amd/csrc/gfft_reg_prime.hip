// Compiled register-engine variants for edges with a factor 7, 11 or 13
// beyond 7 x 2^k: 3 x 7 x 2^k (84 .. 672), 11 x 2^k (88 .. 704) and
// 13 x 2^k (104 .. 832), FP32 and FP64.  P = 21 / 11 / 13 points per thread
// keeps TPT = N/P a power of two (wave-local z exchanges); column tiles are
// 16 (FP32) / 8 (FP64) columns, z tiles G row pairs within 64 KB of LDS.
#include "gfft_schedules.h"

namespace gfft {

int registry_prime_f32(const KernelEntry** out) {
  static const KernelEntry table[] = {
      //        type   schedule  P_rows G  P_cols CW
      GFFT_SIZE(float, Sc84, 21, 32, 21, 32),
      GFFT_SIZE(float, Sc168, 21, 16, 21, 16),
      GFFT_SIZE(float, Sc336, 21, 16, 21, 16),
      GFFT_SIZE(float, Sc672, 21, 8, 21, 16),
      GFFT_SIZE(float, Sc88, 11, 32, 11, 32),
      GFFT_SIZE(float, Sc176, 11, 16, 11, 16),
      GFFT_SIZE(float, Sc352, 11, 8, 11, 16),
      GFFT_SIZE(float, Sc704, 11, 4, 11, 16),
      GFFT_SIZE(float, Sc104, 13, 32, 13, 32),
      GFFT_SIZE(float, Sc208, 13, 16, 13, 16),
      GFFT_SIZE(float, Sc416, 13, 8, 13, 16),
      GFFT_SIZE(float, Sc832, 13, 4, 13, 16),
  };
  *out = table;
  return (int)(sizeof(table) / sizeof(table[0]));
}

int registry_prime_f64(const KernelEntry** out) {
  static const KernelEntry table[] = {
      GFFT_SIZE(double, Sc84, 21, 32, 21, 16),
      GFFT_SIZE(double, Sc168, 21, 16, 21, 8),
      GFFT_SIZE(double, Sc336, 21, 8, 21, 8),
      GFFT_SIZE(double, Sc672, 21, 4, 21, 8),
      GFFT_SIZE(double, Sc88, 11, 32, 11, 16),
      GFFT_SIZE(double, Sc176, 11, 16, 11, 8),
      GFFT_SIZE(double, Sc352, 11, 8, 11, 8),
      GFFT_SIZE(double, Sc704, 11, 4, 11, 8),
      GFFT_SIZE(double, Sc104, 13, 32, 13, 16),
      GFFT_SIZE(double, Sc208, 13, 16, 13, 8),
      GFFT_SIZE(double, Sc416, 13, 8, 13, 8),
      GFFT_SIZE(double, Sc832, 13, 4, 13, 8),
  };
  *out = table;
  return (int)(sizeof(table) / sizeof(table[0]));
}

}  // namespace gfft
