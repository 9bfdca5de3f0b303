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

// Alternative radix orders (tuning candidates; each has its own twiddle table).
using Sc768b = Sched<3, 8, 8, 4>;
using Sc768c = Sched<8, 4, 8, 3>;
using Sc1024b = Sched<4, 4, 8, 8>;
using Sc1024c = Sched<8, 4, 4, 8>;

// 2^a 3^b 5^c sizes of the serving range [64, 512] without a power-of-two
// schedule above: 8s first (a 2^4 factor as 4 x 4), then 5s, then 3s.
using Sc72 = Sched<8, 3, 3>;
using Sc75 = Sched<5, 5, 3>;
using Sc80 = Sched<4, 4, 5>;
using Sc81 = Sched<3, 3, 3, 3>;
using Sc90 = Sched<2, 5, 3, 3>;
using Sc100 = Sched<4, 5, 5>;
using Sc108 = Sched<4, 3, 3, 3>;
using Sc120 = Sched<8, 5, 3>;
using Sc125 = Sched<5, 5, 5>;
using Sc135 = Sched<5, 3, 3, 3>;
using Sc144 = Sched<4, 4, 3, 3>;
using Sc150 = Sched<2, 5, 5, 3>;
using Sc162 = Sched<2, 3, 3, 3, 3>;
using Sc180 = Sched<4, 5, 3, 3>;
using Sc200 = Sched<8, 5, 5>;
using Sc216 = Sched<8, 3, 3, 3>;
using Sc225 = Sched<5, 5, 3, 3>;
using Sc240 = Sched<4, 4, 5, 3>;
using Sc243 = Sched<3, 3, 3, 3, 3>;
using Sc250 = Sched<2, 5, 5, 5>;
using Sc270 = Sched<2, 5, 3, 3, 3>;
using Sc288 = Sched<8, 4, 3, 3>;
using Sc300 = Sched<4, 5, 5, 3>;
using Sc320 = Sched<8, 8, 5>;
using Sc324 = Sched<4, 3, 3, 3, 3>;
using Sc360 = Sched<8, 5, 3, 3>;
using Sc375 = Sched<5, 5, 5, 3>;
using Sc400 = Sched<4, 4, 5, 5>;
using Sc405 = Sched<5, 3, 3, 3, 3>;
using Sc432 = Sched<4, 4, 3, 3, 3>;
using Sc450 = Sched<2, 5, 5, 3, 3>;
using Sc480 = Sched<8, 4, 5, 3>;
using Sc486 = Sched<2, 3, 3, 3, 3, 3>;
using Sc500 = Sched<4, 5, 5, 5>;

// 7 x 2^k: 8s first, then 4s, radix 7 last.
using Sc56 = Sched<8, 7>;
using Sc112 = Sched<4, 4, 7>;
using Sc224 = Sched<8, 4, 7>;
using Sc448 = Sched<8, 8, 7>;
using Sc896 = Sched<8, 4, 4, 7>;

// 3 x 7 x 2^k, 11 x 2^k, 13 x 2^k: 8s and 4s first, the odd primes last.
using Sc84 = Sched<4, 3, 7>;
using Sc168 = Sched<8, 3, 7>;
using Sc336 = Sched<4, 4, 3, 7>;
using Sc672 = Sched<8, 4, 3, 7>;
using Sc88 = Sched<8, 11>;
using Sc176 = Sched<4, 4, 11>;
using Sc352 = Sched<8, 4, 11>;
using Sc704 = Sched<8, 8, 11>;
using Sc104 = Sched<8, 13>;
using Sc208 = Sched<4, 4, 13>;
using Sc416 = Sched<8, 4, 13>;
using Sc832 = Sched<8, 8, 13>;

// Default variant of a size: z R2C, z C2R, strided forward, strided inverse.
// The first variant registered for a (precision, kind, N) is the default;
// later ones are tuning alternatives (gfft_plan_tune / gfft_plan_set_variant).
#define GFFT_SIZE(T, SC, PR, G, PC, CW) GFFT_ROWS(T, SC, PR, G), GFFT_COLS(T, SC, PC, CW)
#define GFFT_ROWS(T, SC, PR, G) entry_rows_r2c<T, SC, PR, G>(), entry_rows_c2r<T, SC, PR, G>()
#define GFFT_COLS(T, SC, PC, CW) entry_cols<T, false, SC, PC, CW>(), entry_cols<T, true, SC, PC, CW>()
// the same with linear tile dispatch (no XCD-aware remap)
#define GFFT_COLS_LIN(T, SC, PC, CW) \
  entry_cols<T, false, SC, PC, CW, 2>(), entry_cols<T, true, SC, PC, CW, 2>()
// persistent software-pipelined column pass, OCC workgroups per CU
#define GFFT_COLS_PIPE(T, SC, PC, CW, OCC, ORD) \
  entry_cols_pipe<T, false, SC, PC, CW, OCC, ORD>(), entry_cols_pipe<T, true, SC, PC, CW, OCC, ORD>()
// the same with per-tile recomputed element offsets (RA 1 in k_cols_pipe)
#define GFFT_COLS_PIPE_RA(T, SC, PC, CW, OCC, ORD) \
  entry_cols_pipe<T, false, SC, PC, CW, OCC, ORD, 1>(), entry_cols_pipe<T, true, SC, PC, CW, OCC, ORD, 1>()
// the same with the straight-line loop for full tiles (FL 1 in launch_cols_pipe)
#define GFFT_COLS_PIPE_F(T, SC, PC, CW, OCC, ORD, RA)                 \
  entry_cols_pipe<T, false, SC, PC, CW, OCC, ORD, RA, 1>(),          \
      entry_cols_pipe<T, true, SC, PC, CW, OCC, ORD, RA, 1>()
// persistent software-pipelined z passes, G row pairs per tile
#define GFFT_ROWS_PIPE(T, SC, PR, G, OCC, ORD) \
  entry_rows_r2c_pipe<T, SC, PR, G, OCC, ORD>(), entry_rows_c2r_pipe<T, SC, PR, G, OCC, ORD>()

}  // namespace gfft
