// Radix orders of the compiled register-engine sizes (shared by the FP32 and
// FP64 instantiation units).  Larger radices first: the first stage needs no
// twiddles and the last stage writes straight to global memory.
#pragma once
#include "gfft_kernels.h"

namespace gfft {

using Sc32 = Sched<8, 4>;
using Sc48 = Sched<4, 4, 3>;
using Sc64 = Sched<8, 8>;
using Sc96 = Sched<8, 4, 3>;
using Sc128 = Sched<8, 4, 4>;
using Sc160 = Sched<8, 4, 5>;
using Sc192 = Sched<8, 8, 3>;
using Sc256 = Sched<8, 8, 4>;
using Sc384 = Sched<8, 4, 4, 3>;
using Sc512 = Sched<8, 8, 8>;
using Sc768 = Sched<8, 8, 4, 3>;
using Sc1024 = Sched<8, 8, 4, 4>;
using Sc2048 = Sched<8, 8, 8, 4>;

// Four entries per size: z R2C, z C2R, strided forward, strided inverse.
#define GFFT_SIZE(T, SC, PR, G, PC, CW)                                                          \
  entry_rows_r2c<T, SC, PR, G>(), entry_rows_c2r<T, SC, PR, G>(), entry_cols<T, false, SC, PC, CW>(), \
      entry_cols<T, true, SC, PC, CW>()

}  // namespace gfft
