// Compiled register-engine variants, FP64.  Column tiles are 8 columns wide
// (8 x 16 B = 128 B per row segment) except where the tile would not fit LDS.
#include "gfft_schedules.h"

namespace gfft {

int registry_f64(const KernelEntry** out) {
  static const KernelEntry table[] = {
      //        type    schedule P_rows G  P_cols CW
      GFFT_SIZE(double, Sc32, 8, 64, 8, 8),
      GFFT_SIZE(double, Sc48, 12, 32, 12, 8),
      GFFT_SIZE(double, Sc64, 8, 32, 8, 8),
      GFFT_SIZE(double, Sc96, 24, 16, 24, 8),
      GFFT_SIZE(double, Sc128, 8, 16, 8, 8),
      GFFT_SIZE(double, Sc160, 40, 16, 40, 8),
      GFFT_SIZE(double, Sc192, 24, 8, 24, 8),
      GFFT_SIZE(double, Sc256, 8, 8, 8, 8),
      GFFT_SIZE(double, Sc384, 24, 4, 24, 8),
      GFFT_SIZE(double, Sc512, 8, 4, 8, 8),
      GFFT_SIZE(double, Sc768, 24, 2, 24, 8),
      GFFT_SIZE(double, Sc1024, 8, 2, 16, 8),
      GFFT_SIZE(double, Sc2048, 8, 1, 16, 4),
  };
  *out = table;
  return (int)(sizeof(table) / sizeof(table[0]));
}

}  // namespace gfft
