// Plan structures shared by the planner (gfft_plan.cpp) and the C ABI
// (gfft_abi.cpp).  Internal.
#pragma once
#include <string>
#include <utility>
#include <vector>

#include "gfft_internal.h"
#include "gpufft/gpufft.h"

namespace gfft {

struct AxisPlan {
  int N = 0;
  bool rows = false;     // z pass (two-for-one real rows) vs strided columns
  bool active = false;   // false for a length-1 strided axis (identity)
  const KernelEntry* reg_fwd = nullptr;  // chosen compiled variant (both or neither;
  const KernelEntry* reg_inv = nullptr;  //  null = the generic run-time-schedule kernel)
  // candidate (forward, inverse) pairs sharing `sched`; {null, null} = generic
  std::vector<std::pair<const KernelEntry*, const KernelEntry*>> cands;
  int variant = 0;       // index into cands
  double tuned_ms = -1;  // fwd+inv time of the chosen variant when tuned
  Schedule sched{};      // radix order of the current variant
  RtSchedHost rt{};      // radix order of the generic kernel (scheds[0])
  // distinct radix orders among the candidates (scheds[0] = the default's)
  // and, per candidate, the index of its order; one twiddle table per order
  std::vector<Schedule> scheds;
  std::vector<int> cand_sched;
  std::vector<void*> d_tws;
  void* d_tw = nullptr;  // device twiddle table of the current variant
  size_t tw_bytes = 0;   // all tables of this axis
};

struct PlanCore {
  int64_t nx = 0, ny = 0, nz = 0, K = 0;
  int procs = 1, rank = 0, prec = 0;
  double r2c_scale = 1.0, c2r_scale = 1.0;  // fused output scales (1 = unnormalised)
  int64_t chunk_planes = 0;  // > 0: z and y passes alternate per block of x-planes
  // Reversed tile walks, bit (inverse * 3 + axis): the default (42) is the
  // alternating rule of gfft_plan_set_walk mode 1 (R2C y, C2R z and x)
  int walk_mask = 42;
  // Order of the strided passes: bit 0 = R2C runs x before y, bit 1 = C2R
  // runs y before x (the z pass is always the first R2C / last C2R pass)
  int pass_order = 0;
  AxisPlan z, y, x;
  // fused z + y kernels over LDS-resident x-planes (null: none compiled for
  // this geometry); `fused` selects them for full transforms
  const PlaneEntry* plane = nullptr;
  void* d_twz_plane = nullptr;
  void* d_twy_plane = nullptr;
  size_t plane_tw_bytes = 0;
  int fused = 0;
  double fused_ms = -1, unfused_ms = -1;  // z + y round-trip times when tuned
};

int validate_dims(int64_t nx, int64_t ny, int64_t nz, int procs, int rank, int prec);
int plan_host(PlanCore& c, int64_t nx, int64_t ny, int64_t nz, int procs, int rank, int prec);
std::string plan_json(const PlanCore& c);
int upload_twiddles(PlanCore& c);
void free_twiddles(PlanCore& c);
int exec_r2c(const PlanCore& c, void* data, hipStream_t s);
int exec_c2r(const PlanCore& c, void* data, hipStream_t s);
int exec_axis(const PlanCore& c, int axis, bool inverse, void* data, hipStream_t s);
int set_variant(PlanCore& c, int axis, int variant);
int set_fused(PlanCore& c, int on);
// process-wide tuning cache as text lines "<device>|<geometry> <choice>"
std::string tune_cache_text();
int tune_cache_load(const char* text);
void tune_cache_clear();
size_t plan_device_bytes(const PlanCore& c);
size_t local_bytes(const PlanCore& c);
int device_cu_count();
// Stage twiddle table of a schedule (layout of gfft_device.h), as bytes.
std::vector<unsigned char> twiddle_table(const Schedule& s, int prec);
int exec_r2c_local(const PlanCore& c, void* data, void* sendbuf, hipStream_t s);
int exec_r2c_finish(const PlanCore& c, const void* recvbuf, void* out, hipStream_t s);
int exec_c2r_start(const PlanCore& c, const void* data, void* sendbuf, hipStream_t s);
int exec_c2r_local(const PlanCore& c, const void* recvbuf, void* data, hipStream_t s);
int tune(PlanCore& c, void* scratch, hipStream_t s, int reps);
// Set the variants (and fused mode) the process-wide tune cache holds for
// this plan's geometry on the current device; call after upload_twiddles.
// Returns the number of choices applied.
int apply_tune_cache(PlanCore& c);
ColGeom y_geom(const PlanCore& c, double scale);
ColGeom x_geom(const PlanCore& c, double scale);
RowGeom z_geom(const PlanCore& c, double scale);

}  // namespace gfft
