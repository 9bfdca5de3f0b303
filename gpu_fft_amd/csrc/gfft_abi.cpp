// extern "C" entry points declared in include/gpufft/gpufft.h.
#include <cstdint>
#include <cstring>
#include <new>
#include <string>

#include "gfft_plan.h"

struct gfft_plan_s {
  gfft::PlanCore core;
  int device = 0;
};

namespace {

int copy_text(const std::string& s, char* buf, size_t len) {
  if (!buf) return GFFT_ERR_INVALID_ARG;
  if (s.size() + 1 > len) return GFFT_ERR_BUFFER_TOO_SMALL;
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return GFFT_OK;
}

// every kernel moves whole complex elements: 8 B (FP32) or 16 B (FP64)
bool misaligned(const gfft_plan_s* p, const void* ptr) {
  const uintptr_t a = p->core.prec == gfft::F32 ? 8 : 16;
  return (reinterpret_cast<uintptr_t>(ptr) & (a - 1)) != 0;
}

int check_device(const gfft_plan_s* p) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return GFFT_ERR_HIP;
  return dev == p->device ? GFFT_OK : GFFT_ERR_DEVICE_MISMATCH;
}

}  // namespace

extern "C" {

int gfft_version(void) { return GFFT_VERSION; }

const char* gfft_strerror(int status) {
  switch (status) {
    case GFFT_OK: return "success";
    case GFFT_ERR_INVALID_ARG:
      return "invalid argument (null handle or pointer, data pointer not aligned to a complex element, or bad flag)";
    case GFFT_ERR_INVALID_DIMS: return "invalid dimensions (an edge < 1, procs does not divide Nx and Ny, or overflow)";
    case GFFT_ERR_UNSUPPORTED_SIZE: return "unsupported edge length (prime factor > 13 or above the LDS-resident limit)";
    case GFFT_ERR_UNSUPPORTED_PROCS: return "full transform with procs > 1 needs an inter-rank exchange (not provided)";
    case GFFT_ERR_INVALID_RANK: return "rank outside [0, procs)";
    case GFFT_ERR_OUT_OF_MEMORY: return "device allocation failed";
    case GFFT_ERR_HIP: return "HIP runtime or kernel launch error";
    case GFFT_ERR_DEVICE_MISMATCH: return "current HIP device differs from the plan's device";
    case GFFT_ERR_BUFFER_TOO_SMALL: return "text buffer too small";
    case GFFT_ERR_OVER_BUDGET: return "request larger than the server's whole HBM budget";
    case GFFT_ERR_TIMEOUT: return "timed out (server drain, or a bounded wait of a persistent batch launch)";
    case GFFT_ERR_ABORTED: return "a persistent batch launch gave up its waits on the host's abort request";
    case GFFT_ERR_SCRATCH_TOO_SMALL: return "the plan needs more counter words than the library's static scratch array holds";
    default: return "unknown status";
  }
}

int gfft_query_plan(int64_t nx, int64_t ny, int64_t nz, int procs, int rank, int prec, char* buf, size_t len) {
  gfft::PlanCore c;
  const int st = gfft::plan_host(c, nx, ny, nz, procs, rank, prec);
  if (st != GFFT_OK) return st;
  return copy_text(gfft::plan_json(c), buf, len);
}

int gfft_plan_create(gfft_plan* plan, int64_t nx, int64_t ny, int64_t nz, int procs, int rank, int prec) {
  if (!plan) return GFFT_ERR_INVALID_ARG;
  *plan = nullptr;
  gfft_plan_s* p = new (std::nothrow) gfft_plan_s;
  if (!p) return GFFT_ERR_OUT_OF_MEMORY;
  int st = gfft::plan_host(p->core, nx, ny, nz, procs, rank, prec);
  if (st == GFFT_OK && hipGetDevice(&p->device) != hipSuccess) st = GFFT_ERR_HIP;
  if (st == GFFT_OK) st = gfft::upload_twiddles(p->core);
  if (st != GFFT_OK) {
    gfft::free_twiddles(p->core);
    delete p;
    return st;
  }
  (void)gfft::apply_tune_cache(p->core);  // tuned choices of this geometry, if any
  *plan = p;
  return GFFT_OK;
}

int gfft_plan_destroy(gfft_plan plan) {
  if (!plan) return GFFT_ERR_INVALID_ARG;
  gfft::free_twiddles(plan->core);
  delete plan;
  return GFFT_OK;
}

int gfft_exec_r2c(gfft_plan plan, void* data, gfft_stream_t stream) {
  if (!plan || !data || misaligned(plan, data)) return GFFT_ERR_INVALID_ARG;
  if (plan->core.procs != 1) return GFFT_ERR_UNSUPPORTED_PROCS;
  const int st = check_device(plan);
  if (st != GFFT_OK) return st;
  return gfft::exec_r2c(plan->core, data, reinterpret_cast<hipStream_t>(stream));
}

int gfft_exec_c2r(gfft_plan plan, void* data, gfft_stream_t stream) {
  if (!plan || !data || misaligned(plan, data)) return GFFT_ERR_INVALID_ARG;
  if (plan->core.procs != 1) return GFFT_ERR_UNSUPPORTED_PROCS;
  const int st = check_device(plan);
  if (st != GFFT_OK) return st;
  return gfft::exec_c2r(plan->core, data, reinterpret_cast<hipStream_t>(stream));
}

int gfft_exec_axis(gfft_plan plan, int axis, int inverse, void* data, gfft_stream_t stream) {
  if (!plan || !data || axis < 0 || axis > 3 || misaligned(plan, data)) return GFFT_ERR_INVALID_ARG;
  const int st = check_device(plan);
  if (st != GFFT_OK) return st;
  return gfft::exec_axis(plan->core, axis, inverse != 0, data, reinterpret_cast<hipStream_t>(stream));
}

int gfft_exec_r2c_local(gfft_plan plan, void* data, void* sendbuf, gfft_stream_t stream) {
  if (!plan || !data || !sendbuf || data == sendbuf || misaligned(plan, data) || misaligned(plan, sendbuf))
    return GFFT_ERR_INVALID_ARG;
  const int st = check_device(plan);
  if (st != GFFT_OK) return st;
  return gfft::exec_r2c_local(plan->core, data, sendbuf, reinterpret_cast<hipStream_t>(stream));
}

int gfft_exec_r2c_finish(gfft_plan plan, const void* recvbuf, void* out, gfft_stream_t stream) {
  if (!plan || !recvbuf || !out || recvbuf == out || misaligned(plan, recvbuf) || misaligned(plan, out))
    return GFFT_ERR_INVALID_ARG;
  const int st = check_device(plan);
  if (st != GFFT_OK) return st;
  return gfft::exec_r2c_finish(plan->core, recvbuf, out, reinterpret_cast<hipStream_t>(stream));
}

int gfft_exec_c2r_start(gfft_plan plan, const void* data, void* sendbuf, gfft_stream_t stream) {
  if (!plan || !data || !sendbuf || data == sendbuf || misaligned(plan, data) || misaligned(plan, sendbuf))
    return GFFT_ERR_INVALID_ARG;
  const int st = check_device(plan);
  if (st != GFFT_OK) return st;
  return gfft::exec_c2r_start(plan->core, data, sendbuf, reinterpret_cast<hipStream_t>(stream));
}

int gfft_exec_c2r_local(gfft_plan plan, const void* recvbuf, void* data, gfft_stream_t stream) {
  if (!plan || !recvbuf || !data || recvbuf == data || misaligned(plan, recvbuf) || misaligned(plan, data))
    return GFFT_ERR_INVALID_ARG;
  const int st = check_device(plan);
  if (st != GFFT_OK) return st;
  return gfft::exec_c2r_local(plan->core, recvbuf, data, reinterpret_cast<hipStream_t>(stream));
}

int gfft_plan_set_variant(gfft_plan plan, int axis, int variant) {
  if (!plan) return GFFT_ERR_INVALID_ARG;
  return gfft::set_variant(plan->core, axis, variant);
}

int gfft_plan_tune(gfft_plan plan, void* scratch, gfft_stream_t stream, int reps) {
  if (!plan || !scratch || misaligned(plan, scratch)) return GFFT_ERR_INVALID_ARG;
  const int st = check_device(plan);
  if (st != GFFT_OK) return st;
  return gfft::tune(plan->core, scratch, reinterpret_cast<hipStream_t>(stream), reps);
}

int gfft_plan_set_chunking(gfft_plan plan, int64_t planes) {
  if (!plan || planes < 0) return GFFT_ERR_INVALID_ARG;
  plan->core.chunk_planes = planes;
  return GFFT_OK;
}

int gfft_tune_cache_export(char* buf, size_t len, size_t* needed) {
  const std::string t = gfft::tune_cache_text();
  if (needed) *needed = t.size() + 1;
  if (!buf) return needed ? GFFT_OK : GFFT_ERR_INVALID_ARG;
  return copy_text(t, buf, len);
}

int gfft_tune_cache_import(const char* text) {
  if (!text) return GFFT_ERR_INVALID_ARG;
  return gfft::tune_cache_load(text);
}

int gfft_tune_cache_clear(void) {
  gfft::tune_cache_clear();
  return GFFT_OK;
}

int gfft_plan_set_fused(gfft_plan plan, int on) {
  if (!plan) return GFFT_ERR_INVALID_ARG;
  return gfft::set_fused(plan->core, on);
}

int gfft_plan_set_walk(gfft_plan plan, int mode) {
  if (!plan || mode < 0 || mode > 1) return GFFT_ERR_INVALID_ARG;
  plan->core.walk_mask = mode ? 42 : 0;
  return GFFT_OK;
}

int gfft_plan_set_walk_mask(gfft_plan plan, int mask) {
  if (!plan || mask < 0 || mask > 63) return GFFT_ERR_INVALID_ARG;
  plan->core.walk_mask = mask;
  return GFFT_OK;
}

int gfft_plan_set_pass_order(gfft_plan plan, int order) {
  if (!plan || order < 0 || order > 3) return GFFT_ERR_INVALID_ARG;
  plan->core.pass_order = order;
  return GFFT_OK;
}

int gfft_plan_set_scale(gfft_plan plan, double r2c_scale, double c2r_scale) {
  if (!plan) return GFFT_ERR_INVALID_ARG;
  plan->core.r2c_scale = r2c_scale;
  plan->core.c2r_scale = c2r_scale;
  return GFFT_OK;
}

int gfft_plan_device_bytes(gfft_plan plan, size_t* bytes) {
  if (!plan || !bytes) return GFFT_ERR_INVALID_ARG;
  *bytes = gfft::plan_device_bytes(plan->core);
  return GFFT_OK;
}

int gfft_plan_data_bytes(gfft_plan plan, size_t* bytes) {
  if (!plan || !bytes) return GFFT_ERR_INVALID_ARG;
  const gfft::PlanCore& c = plan->core;
  *bytes = (size_t)(c.nx / c.procs) * (size_t)c.ny * (size_t)c.K * (c.prec == gfft::F32 ? 8u : 16u);
  return GFFT_OK;
}

int gfft_plan_describe(gfft_plan plan, char* buf, size_t len) {
  if (!plan) return GFFT_ERR_INVALID_ARG;
  return copy_text(gfft::plan_json(plan->core), buf, len);
}

}  // extern "C"
