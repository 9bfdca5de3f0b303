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
      GFFT_SIZE(double, Sc160, 10, 16, 10, 16),  // P=10: 16 threads per transform
      GFFT_SIZE(double, Sc192, 24, 8, 24, 8),
      GFFT_SIZE(double, Sc256, 8, 8, 8, 8),
      GFFT_SIZE(double, Sc384, 24, 4, 24, 8),
      // 512: measured default columns first (16 points per thread)
      GFFT_ROWS(double, Sc512, 16, 4),
      GFFT_COLS(double, Sc512, 16, 8),
      GFFT_ROWS(double, Sc512, 8, 4),
      GFFT_COLS(double, Sc512, 8, 8),
      GFFT_SIZE(double, Sc768, 24, 2, 24, 8),
      GFFT_SIZE(double, Sc1024, 8, 2, 16, 8),
      GFFT_SIZE(double, Sc2048, 8, 1, 16, 4),
      // tuning alternatives
      GFFT_SIZE(double, Sc160, 40, 16, 40, 8),
      GFFT_SIZE(double, Sc160, 20, 16, 20, 16),
      GFFT_SIZE(double, Sc160, 16, 24, 16, 16),
      GFFT_COLS(double, Sc160, 40, 16),
      GFFT_COLS(double, Sc512, 8, 4),
      GFFT_COLS(double, Sc512, 16, 4),
      GFFT_ROWS(double, Sc512, 8, 2),
      GFFT_COLS(double, Sc768, 24, 4),
      GFFT_COLS_PIPE(double, Sc512, 8, 8, 2, 0),
      GFFT_COLS_PIPE(double, Sc512, 16, 8, 2, 0),
      GFFT_COLS_PIPE(double, Sc512, 16, 8, 1, 0),
      GFFT_COLS_PIPE(double, Sc512, 8, 8, 2, 1),
      GFFT_COLS_PIPE(double, Sc512, 16, 8, 2, 1),
      GFFT_ROWS_PIPE(double, Sc512, 8, 4, 2, 1),
      GFFT_ROWS_PIPE(double, Sc512, 8, 2, 4, 1),
      GFFT_COLS_PIPE_F(double, Sc512, 16, 8, 1, 0, 0),
      GFFT_COLS_PIPE_F(double, Sc512, 16, 8, 1, 1, 0),
      GFFT_COLS_PIPE_F(double, Sc512, 8, 8, 2, 0, 0),
      GFFT_COLS_PIPE_F(double, Sc512, 16, 8, 2, 0, 1),
      GFFT_COLS_PIPE_F(double, Sc512, 16, 8, 2, 1, 1),
      GFFT_COLS_PIPE_RA(double, Sc1024, 16, 8, 1, 0),
      GFFT_COLS_PIPE_RA(double, Sc1024, 16, 8, 1, 1),
      GFFT_COLS_PIPE_RA(double, Sc1024, 16, 4, 2, 0),
      GFFT_COLS_PIPE_RA(double, Sc1024, 16, 4, 2, 1),
  };
  *out = table;
  return (int)(sizeof(table) / sizeof(table[0]));
}

}  // namespace gfft
