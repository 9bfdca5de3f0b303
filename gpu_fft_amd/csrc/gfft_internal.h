// Host-side internal interfaces shared by the planner, the C ABI and the
// kernel instantiation units.  Nothing here is part of the public ABI.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <stddef.h>

namespace gfft {

enum Prec { F32 = 0, F64 = 1 };

// Axis-pass kinds.
enum Kind {
  ROWS_R2C = 0,  // z axis, real rows (pitch 2K reals) -> half spectra (pitch K)
  ROWS_C2R = 1,  // z axis, half spectra -> real rows, unnormalised
  COLS_FWD = 2,  // strided C2C, forward
  COLS_INV = 3,  // strided C2C, inverse (unnormalised)
  NUM_KINDS = 4
};

// Column geometry of a strided pass: column c (0 <= c < ncols) starts at
// element (c / inner) * outer + (c % inner); its point n sits rstride
// elements further per step.  y pass: inner = K, outer = Ny*K, rstride = K.
// x pass: inner = ncols = Ny*K, outer = 0, rstride = Ny*K.
struct ColGeom {
  uint32_t ncols;
  uint32_t inner;
  int64_t outer;
  int64_t rstride;
  double scale;  // applied to the outputs (1 = unnormalised)
};

// Geometry of a z pass: nrows real rows of Nz points; row r's reals start at
// r * 2K and its K half-spectrum bins at complex index r * K (same bytes).
struct RowGeom {
  int64_t nrows;
  int32_t K;
  double scale;  // applied to the outputs (1 = unnormalised)
};

constexpr int kMaxRadices = 16;

struct Schedule {
  int N;
  int S;
  int R[kMaxRadices];
};

// Launch signatures.  `tw` is the stage twiddle table of the kernel's
// schedule (see gfft_twiddles.cpp), in the transform's precision.
typedef hipError_t (*RowsLaunchFn)(void* data, RowGeom g, const void* tw, hipStream_t s);
typedef hipError_t (*ColsLaunchFn)(void* data, ColGeom g, const void* tw, hipStream_t s);

// A compiled (register-resident) kernel variant.
struct KernelEntry {
  int prec;
  int kind;
  int N;
  Schedule sched;   // radix order the kernel was compiled for
  int P;            // points per thread
  int tile;         // pairs (rows) or columns (cols) per workgroup
  int threads;      // threads per workgroup
  void* fn;         // RowsLaunchFn or ColsLaunchFn
  int persist;      // 0: one tile per workgroup; k: persistent, k workgroups per CU
  int order;        // persistent tile order: 0 contiguous ranges, 1 round robin
};

// Registry of compiled variants; returns the number of entries.
int registry_f32(const KernelEntry** out);
int registry_f64(const KernelEntry** out);

// Run-time-schedule (generic) launchers.
struct RtSchedHost {
  int N;
  int S;
  int R[kMaxRadices];
  int p[kMaxRadices];
  int twoff[kMaxRadices];
};
hipError_t launch_generic_rows(int prec, int kind, void* data, RowGeom g, const RtSchedHost& sc,
                               const void* tw, hipStream_t s);
hipError_t launch_generic_cols(int prec, int kind, void* data, ColGeom g, const RtSchedHost& sc,
                               const void* tw, hipStream_t s);
// LDS bytes the generic kernels would need for a size (0 if unsupported).
size_t generic_rows_lds(int prec, int N, int K);
size_t generic_cols_lds(int prec, int N);

}  // namespace gfft
