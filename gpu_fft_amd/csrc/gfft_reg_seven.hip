// Compiled register-engine variants for the edges 7 x 2^k (56 .. 896), FP32
// and FP64 (other edges with a factor 7, 11 or 13 use the run-time-schedule
// kernels).  P = 7 or 14 points per thread keeps TPT = N/P a power of two
// (wave-local z exchanges) with few idle butterfly slots.
#include "gfft_schedules.h"

namespace gfft {

int registry_seven_f32(const KernelEntry** out) {
  static const KernelEntry table[] = {
      //        type   schedule  P_rows G  P_cols CW
      GFFT_SIZE(float, Sc56, 7, 32, 7, 32),
      GFFT_SIZE(float, Sc112, 7, 16, 7, 16),
      GFFT_SIZE(float, Sc224, 14, 16, 14, 16),
      GFFT_SIZE(float, Sc448, 14, 8, 14, 16),
      GFFT_SIZE(float, Sc896, 14, 4, 14, 16),
      GFFT_COLS_PIPE(float, Sc448, 14, 16, 1, 1),
      GFFT_COLS_PIPE(float, Sc448, 14, 16, 1, 0),
      GFFT_COLS_PIPE_RA(float, Sc896, 14, 16, 1, 1),
      GFFT_COLS_PIPE_RA(float, Sc896, 14, 16, 1, 0),
  };
  *out = table;
  return (int)(sizeof(table) / sizeof(table[0]));
}

int registry_seven_f64(const KernelEntry** out) {
  static const KernelEntry table[] = {
      GFFT_SIZE(double, Sc56, 7, 32, 7, 16),
      GFFT_SIZE(double, Sc112, 7, 16, 7, 16),
      GFFT_SIZE(double, Sc224, 14, 16, 14, 8),
      GFFT_SIZE(double, Sc448, 14, 8, 14, 8),
      GFFT_SIZE(double, Sc896, 14, 4, 14, 8),
  };
  *out = table;
  return (int)(sizeof(table) / sizeof(table[0]));
}

}  // namespace gfft
