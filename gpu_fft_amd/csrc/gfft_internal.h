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

// Address map of one side (input or output) of a strided pass: column c
// starts at element (c / inner) * outer + (c % inner); its point n sits at
// + (n % nsub) * rlo + (n / nsub) * rhi (nsub >= N: a single stride rlo).
//   y pass, in place:  inner = K, outer = Ny*K, rlo = K
//   x pass, in place:  inner = Ny*K, outer = 0, rlo = Ny*K
// The two-level form addresses the p destination chunks of a slab exchange.
struct ColMap {
  uint32_t inner;
  int32_t nsub;
  int64_t outer;
  int64_t rlo;
  int64_t rhi;
};

struct ColGeom {
  uint32_t ncols;
  ColMap in, out;
  double scale;  // applied to the outputs (1 = unnormalised)
  int32_t rev = 0;  // 1: tiles run in reverse order (last tile first)
};

inline ColMap col_map(uint32_t inner, int64_t outer, int64_t rlo) {
  ColMap m;
  m.inner = inner;
  m.nsub = 0x7fffffff;
  m.outer = outer;
  m.rlo = rlo;
  m.rhi = 0;
  return m;
}

// Geometry of a z pass: nrows real rows of Nz points; row r's reals start at
// r * 2K and its K half-spectrum bins at complex index r * K (same bytes).
struct RowGeom {
  int64_t nrows;
  int32_t K;
  double scale;  // applied to the outputs (1 = unnormalised)
  int32_t rev = 0;  // 1: tiles run in reverse order (last tile first)
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
// Column passes read `in` and write `out` (the same pointer when in place; a
// workgroup reads all of its tile before writing it, tiles are disjoint).
typedef hipError_t (*ColsLaunchFn)(const void* in, void* out, ColGeom g, const void* tw, hipStream_t s);

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
  int order;        // tile order: persistent 0 contiguous ranges, 1 round robin; one tile per
                    // workgroup 0 XCD-aware ranges, 2 linear
  int readdr = 0;   // persistent columns: 1 = element offsets recomputed per tile (not hoisted)
  int straight = 0; // persistent columns: 1 = straight-line loop for full tiles, single-stride maps
};

// Fused z + y kernel pair for x-planes that fit in LDS (gfft_plane.h).
struct PlaneEntry {
  int prec;
  int nz, ny;
  Schedule schedz, schedy;  // radix orders of the z and y engines
  int threads;
  void* fn_r2c;  // PlaneLaunchFn
  void* fn_c2r;
};
typedef hipError_t (*PlaneLaunchFn)(void* data, int64_t nplanes, double scale, const void* twz, const void* twy,
                                    hipStream_t s);
int registry_plane(const PlaneEntry** out);

// Registry units of compiled variants; each returns its number of entries.
int registry_f32(const KernelEntry** out);
int registry_f64(const KernelEntry** out);
int registry_smooth_f32_a(const KernelEntry** out);
int registry_smooth_f32_b(const KernelEntry** out);
int registry_smooth_f64_a(const KernelEntry** out);
int registry_smooth_f64_b(const KernelEntry** out);
int registry_seven_f32(const KernelEntry** out);
int registry_seven_f64(const KernelEntry** out);
int registry_prime_f32(const KernelEntry** out);
int registry_prime_f64(const KernelEntry** out);

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
hipError_t launch_generic_cols(int prec, int kind, const void* in, void* out, ColGeom g, const RtSchedHost& sc,
                               const void* tw, hipStream_t s);
// LDS bytes the generic kernels would need for a size (0 if unsupported).
size_t generic_rows_lds(int prec, int N, int K);
size_t generic_cols_lds(int prec, int N);

}  // namespace gfft
