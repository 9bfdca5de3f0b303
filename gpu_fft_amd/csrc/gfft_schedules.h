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

// Default variant of a size: z R2C, z C2R, strided forward, strided inverse.
// The first variant registered for a (precision, kind, N) is the default;
// later ones are tuning alternatives (gfft_plan_tune / gfft_plan_set_variant).
#define GFFT_SIZE(T, SC, PR, G, PC, CW) GFFT_ROWS(T, SC, PR, G), GFFT_COLS(T, SC, PC, CW)
#define GFFT_ROWS(T, SC, PR, G) entry_rows_r2c<T, SC, PR, G>(), entry_rows_c2r<T, SC, PR, G>()
#define GFFT_COLS(T, SC, PC, CW) entry_cols<T, false, SC, PC, CW>(), entry_cols<T, true, SC, PC, CW>()
// persistent software-pipelined column pass, OCC workgroups per CU
#define GFFT_COLS_PIPE(T, SC, PC, CW, OCC, ORD) \
  entry_cols_pipe<T, false, SC, PC, CW, OCC, ORD>(), entry_cols_pipe<T, true, SC, PC, CW, OCC, ORD>()
// persistent software-pipelined z passes, G row pairs per tile
#define GFFT_ROWS_PIPE(T, SC, PR, G, OCC, ORD) \
  entry_rows_r2c_pipe<T, SC, PR, G, OCC, ORD>(), entry_rows_c2r_pipe<T, SC, PR, G, OCC, ORD>()

}  // namespace gfft
