// Compiled register-engine variants, FP32.  Column tiles are 16 columns wide
// (16 x 8 B = 128 B per row segment) except where the tile would not fit LDS.
#include "gfft_schedules.h"

namespace gfft {

int registry_f32(const KernelEntry** out) {
  static const KernelEntry table[] = {
      //        type   schedule  P_rows G  P_cols CW
      GFFT_SIZE(float, Sc32, 8, 64, 8, 16),
      GFFT_SIZE(float, Sc48, 12, 64, 12, 16),
      GFFT_SIZE(float, Sc64, 8, 32, 8, 16),
      GFFT_SIZE(float, Sc96, 24, 32, 24, 16),
      GFFT_SIZE(float, Sc128, 8, 16, 8, 16),
      GFFT_SIZE(float, Sc160, 10, 16, 10, 16),  // P=10: 16 threads per transform
      GFFT_SIZE(float, Sc192, 24, 16, 24, 16),
      GFFT_SIZE(float, Sc256, 8, 8, 8, 16),
      GFFT_SIZE(float, Sc384, 24, 8, 24, 16),
      // 512: measured defaults first (z: 8 pairs per workgroup; y and x: persistent,
      // pipelined, one workgroup per CU, round-robin tiles)
      GFFT_ROWS(float, Sc512, 8, 8),
      GFFT_COLS_PIPE(float, Sc512, 16, 16, 1, 1),
      GFFT_SIZE(float, Sc512, 8, 4, 8, 16),
      // 768, 1024, 2048: measured defaults first (z in the radix order 8,4,8,3 /
      // 8,4,4,8 / 8,8,8,4; 4 / 4 / 2 row pairs per tile; strided passes
      // persistent, pipelined, 16 columns per tile, round-robin tiles,
      // element offsets recomputed per tile)
      GFFT_ROWS(float, Sc768c, 24, 4),
      GFFT_COLS_PIPE_RA(float, Sc768, 24, 16, 1, 1),
      GFFT_ROWS(float, Sc1024c, 8, 4),
      GFFT_COLS_PIPE_RA(float, Sc1024, 16, 16, 1, 1),
      GFFT_COLS(float, Sc1024, 16, 8),
      GFFT_SIZE(float, Sc768, 24, 4, 24, 16),
      GFFT_SIZE(float, Sc1024, 8, 2, 16, 16),
      GFFT_ROWS(float, Sc2048, 16, 2),
      GFFT_SIZE(float, Sc2048, 8, 1, 16, 8),
      // tuning alternatives
      GFFT_SIZE(float, Sc160, 40, 32, 40, 16),
      GFFT_SIZE(float, Sc160, 20, 32, 20, 32),
      GFFT_SIZE(float, Sc160, 16, 24, 16, 16),
      GFFT_COLS(float, Sc512, 16, 16),
      GFFT_COLS(float, Sc512, 8, 8),
      GFFT_COLS(float, Sc512, 16, 8),
      GFFT_COLS(float, Sc512, 32, 16),
      GFFT_COLS(float, Sc512, 16, 32),
      GFFT_ROWS(float, Sc512, 8, 2),
      GFFT_ROWS(float, Sc512, 16, 4),
      GFFT_ROWS(float, Sc512, 16, 8),
      GFFT_COLS(float, Sc1024, 32, 16),
      GFFT_COLS(float, Sc256, 16, 16),
      GFFT_COLS(float, Sc256, 8, 8),
      GFFT_COLS(float, Sc768, 24, 8),
      GFFT_COLS_PIPE(float, Sc1024, 16, 8, 2, 0),
      GFFT_COLS_PIPE(float, Sc1024, 16, 8, 2, 1),
      GFFT_COLS_PIPE_RA(float, Sc1024, 16, 16, 1, 0),
      GFFT_COLS_PIPE_RA(float, Sc1024, 32, 16, 1, 0),
      GFFT_COLS_PIPE_RA(float, Sc1024, 32, 16, 1, 1),
      GFFT_COLS_PIPE(float, Sc768, 24, 8, 2, 0),
      GFFT_COLS_PIPE(float, Sc768, 24, 8, 2, 1),
      GFFT_COLS_PIPE_RA(float, Sc768, 24, 16, 1, 0),
            GFFT_COLS_PIPE(float, Sc256, 8, 16, 2, 0),
      GFFT_COLS_PIPE(float, Sc256, 8, 16, 2, 1),
      GFFT_COLS_PIPE(float, Sc512, 32, 32, 1, 0),
      GFFT_COLS_PIPE(float, Sc512, 16, 32, 1, 0),
      GFFT_ROWS_PIPE(float, Sc1024, 8, 2, 2, 1),
      GFFT_ROWS_PIPE(float, Sc2048, 8, 1, 4, 1),
      GFFT_ROWS_PIPE(float, Sc768, 24, 4, 2, 1),
      // linear tile dispatch
      GFFT_COLS_LIN(float, Sc512, 16, 16),
      GFFT_COLS_LIN(float, Sc512, 8, 16),
      GFFT_COLS_LIN(float, Sc768, 24, 16),
      GFFT_COLS_LIN(float, Sc768, 24, 8),
      GFFT_COLS_LIN(float, Sc1024, 16, 8),
      GFFT_COLS_LIN(float, Sc1024, 16, 16),
      GFFT_COLS_LIN(float, Sc1024, 32, 16),
      GFFT_COLS_LIN(float, Sc256, 8, 16),
      // z-pass alternatives for the long rows (768, 1024, 2048)
      GFFT_ROWS(float, Sc1024c, 16, 2),
      GFFT_ROWS(float, Sc1024c, 8, 2),
      GFFT_ROWS(float, Sc1024c, 16, 4),
      GFFT_ROWS_PIPE(float, Sc1024c, 8, 2, 4, 1),
      GFFT_ROWS_PIPE(float, Sc1024c, 16, 4, 2, 1),
      GFFT_ROWS(float, Sc2048, 16, 1),
      GFFT_ROWS(float, Sc2048, 8, 2),
      GFFT_ROWS_PIPE(float, Sc2048, 16, 2, 2, 1),
      GFFT_ROWS(float, Sc768c, 24, 8),
      GFFT_ROWS(float, Sc768c, 16, 4),
      GFFT_ROWS_PIPE(float, Sc768c, 24, 8, 2, 1),
      // other radix orders
      GFFT_ROWS(float, Sc1024b, 8, 2),
      GFFT_COLS_PIPE_RA(float, Sc1024b, 16, 16, 1, 0),
      GFFT_COLS_PIPE_RA(float, Sc1024c, 16, 16, 1, 0),
      GFFT_ROWS(float, Sc768b, 24, 4),
      GFFT_COLS_PIPE_RA(float, Sc768b, 24, 16, 1, 0),
      GFFT_COLS_PIPE_RA(float, Sc768c, 24, 16, 1, 0),
      GFFT_COLS_PIPE(float, Sc512, 8, 16, 2, 0),
      GFFT_COLS_PIPE(float, Sc512, 16, 16, 2, 0),
      GFFT_COLS_PIPE(float, Sc512, 16, 16, 1, 0),
      GFFT_COLS_PIPE(float, Sc512, 32, 16, 2, 0),
      GFFT_COLS_PIPE(float, Sc512, 8, 16, 2, 1),
      GFFT_COLS_PIPE(float, Sc512, 16, 16, 2, 1),
      GFFT_COLS_PIPE(float, Sc512, 32, 16, 2, 1),
      GFFT_COLS_PIPE_F(float, Sc512, 16, 16, 1, 0, 0),
      GFFT_COLS_PIPE_F(float, Sc512, 16, 16, 1, 1, 0),
      GFFT_COLS_PIPE_F(float, Sc768, 24, 16, 1, 0, 1),
      GFFT_COLS_PIPE_F(float, Sc768, 24, 16, 1, 1, 1),
      GFFT_COLS_PIPE_F(float, Sc1024, 16, 16, 1, 0, 1),
      GFFT_COLS_PIPE_F(float, Sc1024, 16, 16, 1, 1, 1),
      GFFT_COLS_PIPE_F(float, Sc1024c, 16, 16, 1, 0, 1),
      GFFT_ROWS_PIPE(float, Sc512, 8, 4, 2, 0),
      GFFT_ROWS_PIPE(float, Sc512, 8, 4, 4, 0),
      GFFT_ROWS_PIPE(float, Sc512, 8, 4, 2, 1),
      GFFT_ROWS_PIPE(float, Sc512, 8, 4, 4, 1),
      GFFT_ROWS_PIPE(float, Sc512, 8, 8, 2, 1),
      GFFT_ROWS_PIPE(float, Sc512, 16, 8, 2, 1),
  };
  *out = table;
  return (int)(sizeof(table) / sizeof(table[0]));
}

}  // namespace gfft
