// Planner and executor of the 3D R2C/C2R engine (host C++).
//
// Replaces the reference's INIT_GPU_FFT (GPU_FFT_init.cu:41-93), which stored
// dimensions in process globals, allocated an Nx*(Ny/p)*K work buffer and
// built three hipFFT plans, and the two drivers FFT_R2C.cu:41-70 /
// FFT_C2R.cu:41-70.  At p = 1 the reference's transpose / exchange /
// transpose sequence is the identity permutation plus a full-array copy, so
// the plan here is three in-place axis passes per direction:
//   R2C: z rows (two-for-one real FFT) -> y columns -> x columns
//   C2R: x columns -> y columns -> z rows
#include "gfft_plan.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace gfft {

namespace {

bool factor_235(int64_t n, std::vector<int>& radices) {
  radices.clear();
  if (n < 1) return false;
  int64_t m = n;
  int c2 = 0;
  while (m % 2 == 0) {
    m /= 2;
    ++c2;
  }
  int c3 = 0;
  while (m % 3 == 0) {
    m /= 3;
    ++c3;
  }
  int c5 = 0;
  while (m % 5 == 0) {
    m /= 5;
    ++c5;
  }
  if (m != 1) return false;
  while (c2 >= 3) {
    radices.push_back(8);
    c2 -= 3;
  }
  if (c2 == 2) radices.push_back(4);
  if (c2 == 1) radices.push_back(2);
  for (int i = 0; i < c5; ++i) radices.push_back(5);
  for (int i = 0; i < c3; ++i) radices.push_back(3);
  return true;
}

const KernelEntry* find_entry(int prec, int kind, int n) {
  const KernelEntry* tab = nullptr;
  const int cnt = prec == F32 ? registry_f32(&tab) : registry_f64(&tab);
  for (int i = 0; i < cnt; ++i)
    if (tab[i].kind == kind && tab[i].N == n) return &tab[i];
  return nullptr;
}

// Stage twiddles W_{pR}^{rk} = exp(-2 pi i r k / (pR)), stage s >= 1,
// k-major, r = 1..R-1.  Evaluated in long double (64-bit significand) from
// the integer phase r*k mod pR, then rounded once to the target precision.
template <typename T>
std::vector<T> make_twiddles(const Schedule& s) {
  std::vector<T> out;
  int p = s.S > 0 ? s.R[0] : 1;
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int st = 1; st < s.S; ++st) {
    const int R = s.R[st];
    const int64_t L = (int64_t)p * R;
    for (int k = 0; k < p; ++k)
      for (int r = 1; r < R; ++r) {
        const int64_t m = ((int64_t)r * k) % L;
        const long double a = two_pi * (long double)m / (long double)L;
        out.push_back((T)cosl(a));
        out.push_back((T)-sinl(a));
      }
    p *= R;
  }
  if (out.empty()) {  // keep a valid (unused) pointer for single-stage schedules
    out.push_back((T)1);
    out.push_back((T)0);
  }
  return out;
}

RtSchedHost to_rt(const Schedule& s) {
  RtSchedHost h{};
  h.N = s.N;
  h.S = s.S;
  int p = 1, off = 0;
  for (int i = 0; i < s.S; ++i) {
    h.R[i] = s.R[i];
    h.p[i] = p;
    h.twoff[i] = off;
    if (i >= 1) off += p * (s.R[i] - 1);
    p *= s.R[i];
  }
  // twoff is relative to the table start; stage 0 has none.
  return h;
}

int plan_axis(AxisPlan& a, int prec, int64_t n, bool rows, int64_t K) {
  a = AxisPlan{};
  a.N = (int)n;
  a.rows = rows;
  a.active = rows || n > 1;
  if (!a.active) return GFFT_OK;
  std::vector<int> rad;
  if (n > (1 << 20) || !factor_235(n, rad)) return GFFT_ERR_UNSUPPORTED_SIZE;
  const int kf = rows ? ROWS_R2C : COLS_FWD;
  const int ki = rows ? ROWS_C2R : COLS_INV;
  a.reg_fwd = find_entry(prec, kf, (int)n);
  a.reg_inv = find_entry(prec, ki, (int)n);
  if (a.reg_fwd && a.reg_inv) {
    a.sched = a.reg_fwd->sched;
  } else {
    a.reg_fwd = a.reg_inv = nullptr;
    if ((int)rad.size() > kMaxRadices) return GFFT_ERR_UNSUPPORTED_SIZE;
    a.sched.N = (int)n;
    a.sched.S = (int)rad.size();
    for (size_t i = 0; i < rad.size(); ++i) a.sched.R[i] = rad[i];
    const size_t lds = rows ? generic_rows_lds(prec, (int)n, (int)K) : generic_cols_lds(prec, (int)n);
    if (lds == 0) return GFFT_ERR_UNSUPPORTED_SIZE;
  }
  a.rt = to_rt(a.sched);
  return GFFT_OK;
}

std::string sched_str(const Schedule& s) {
  std::string r = "[";
  for (int i = 0; i < s.S; ++i) {
    if (i) r += ",";
    r += std::to_string(s.R[i]);
  }
  return r + "]";
}

std::string axis_json(const char* name, const AxisPlan& a) {
  char buf[512];
  if (!a.active) {
    std::snprintf(buf, sizeof buf, "\"%s\":{\"n\":%d,\"pass\":\"identity\"}", name, a.N);
    return buf;
  }
  const KernelEntry* e = a.reg_fwd;
  std::snprintf(buf, sizeof buf,
                "\"%s\":{\"n\":%d,\"pass\":\"%s\",\"engine\":\"%s\",\"radices\":%s,\"points_per_thread\":%d,"
                "\"tile\":%d,\"threads\":%d,\"twiddle_bytes\":%zu}",
                name, a.N, a.rows ? "rows_two_for_one" : "strided_columns", e ? "register" : "generic",
                sched_str(a.sched).c_str(), e ? e->P : 0, e ? e->tile : 0, e ? e->threads : 256, a.tw_bytes);
  return buf;
}

}  // namespace

int validate_dims(int64_t nx, int64_t ny, int64_t nz, int procs, int rank, int prec) {
  if (prec != F32 && prec != F64) return GFFT_ERR_INVALID_ARG;
  if (nx < 1 || ny < 1 || nz < 1 || procs < 1) return GFFT_ERR_INVALID_DIMS;
  if (rank < 0 || rank >= procs) return GFFT_ERR_INVALID_RANK;
  if (nx % procs != 0 || ny % procs != 0) return GFFT_ERR_INVALID_DIMS;
  const int64_t K = nz / 2 + 1;
  // column counts of the strided passes are 32-bit
  if ((double)nx * (double)K >= 4294967295.0 || (double)ny * (double)K >= 4294967295.0) return GFFT_ERR_INVALID_DIMS;
  // total complex elements of the local share must fit comfortably in 63 bits
  if ((double)(nx / procs) * (double)ny * (double)K * 16.0 > 9.0e18) return GFFT_ERR_INVALID_DIMS;
  return GFFT_OK;
}

int plan_host(PlanCore& c, int64_t nx, int64_t ny, int64_t nz, int procs, int rank, int prec) {
  c = PlanCore{};
  int st = validate_dims(nx, ny, nz, procs, rank, prec);
  if (st != GFFT_OK) return st;
  c.nx = nx;
  c.ny = ny;
  c.nz = nz;
  c.procs = procs;
  c.rank = rank;
  c.prec = prec;
  c.K = nz / 2 + 1;
  if ((st = plan_axis(c.z, prec, nz, true, c.K)) != GFFT_OK) return st;
  if ((st = plan_axis(c.y, prec, ny, false, c.K)) != GFFT_OK) return st;
  if ((st = plan_axis(c.x, prec, nx, false, c.K)) != GFFT_OK) return st;
  const size_t cs = prec == F32 ? 8 : 16;
  for (AxisPlan* a : {&c.z, &c.y, &c.x}) {
    if (!a->active) continue;
    const size_t n = prec == F32 ? make_twiddles<float>(a->sched).size() : make_twiddles<double>(a->sched).size();
    a->tw_bytes = n / 2 * cs;
  }
  return GFFT_OK;
}

std::string plan_json(const PlanCore& c) {
  char head[256];
  std::snprintf(head, sizeof head,
                "{\"nx\":%lld,\"ny\":%lld,\"nz\":%lld,\"procs\":%d,\"rank\":%d,\"precision\":\"%s\",\"K\":%lld,",
                (long long)c.nx, (long long)c.ny, (long long)c.nz, c.procs, c.rank, c.prec == F32 ? "f32" : "f64",
                (long long)c.K);
  return std::string(head) + axis_json("z", c.z) + "," + axis_json("y", c.y) + "," + axis_json("x", c.x) + "}";
}

int upload_twiddles(PlanCore& c) {
  for (AxisPlan* a : {&c.z, &c.y, &c.x}) {
    if (!a->active) continue;
    size_t bytes = 0;
    const void* src = nullptr;
    std::vector<float> tf;
    std::vector<double> td;
    if (c.prec == F32) {
      tf = make_twiddles<float>(a->sched);
      src = tf.data();
      bytes = tf.size() * sizeof(float);
    } else {
      td = make_twiddles<double>(a->sched);
      src = td.data();
      bytes = td.size() * sizeof(double);
    }
    void* d = nullptr;
    if (hipMalloc(&d, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return GFFT_ERR_OUT_OF_MEMORY;
    }
    a->d_tw = d;
    a->tw_bytes = bytes;
    if (hipMemcpy(d, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return GFFT_ERR_HIP;
  }
  return GFFT_OK;
}

void free_twiddles(PlanCore& c) {
  for (AxisPlan* a : {&c.z, &c.y, &c.x}) {
    if (a->d_tw) (void)hipFree(a->d_tw);
    a->d_tw = nullptr;
  }
}

ColGeom y_geom(const PlanCore& c, double scale) {
  const int64_t nxl = c.nx / c.procs;
  ColGeom g;
  g.ncols = (uint32_t)(nxl * c.K);
  g.inner = (uint32_t)c.K;
  g.outer = c.ny * c.K;
  g.rstride = c.K;
  g.scale = scale;
  return g;
}

ColGeom x_geom(const PlanCore& c, double scale) {
  const int64_t nyl = c.ny / c.procs;
  ColGeom g;
  g.ncols = (uint32_t)(nyl * c.K);
  g.inner = g.ncols;
  g.outer = 0;
  g.rstride = nyl * c.K;
  g.scale = scale;
  return g;
}

RowGeom z_geom(const PlanCore& c, double scale) {
  RowGeom g;
  g.nrows = (c.nx / c.procs) * c.ny;
  g.K = (int32_t)c.K;
  g.scale = scale;
  return g;
}

hipError_t run_rows(const PlanCore& c, const AxisPlan& a, bool inverse, void* data, RowGeom g, hipStream_t s) {
  const KernelEntry* e = inverse ? a.reg_inv : a.reg_fwd;
  if (e) return reinterpret_cast<RowsLaunchFn>(e->fn)(data, g, a.d_tw, s);
  return launch_generic_rows(c.prec, inverse ? ROWS_C2R : ROWS_R2C, data, g, a.rt, a.d_tw, s);
}

hipError_t run_cols(const PlanCore& c, const AxisPlan& a, bool inverse, void* data, ColGeom g, hipStream_t s) {
  if (!a.active) return hipSuccess;
  const KernelEntry* e = inverse ? a.reg_inv : a.reg_fwd;
  if (e) return reinterpret_cast<ColsLaunchFn>(e->fn)(data, g, a.d_tw, s);
  return launch_generic_cols(c.prec, inverse ? COLS_INV : COLS_FWD, data, g, a.rt, a.d_tw, s);
}

// The output scale of a direction is fused into its last active pass:
// R2C ends with x (or y, or z when both strided axes have length 1), C2R
// always ends with z.
double axis_scale(const PlanCore& c, int axis, bool inverse) {
  if (inverse) return axis == 0 ? c.c2r_scale : 1.0;
  const int last = c.x.active ? 2 : (c.y.active ? 1 : 0);
  return axis == last ? c.r2c_scale : 1.0;
}

int exec_axis(const PlanCore& c, int axis, bool inverse, void* data, hipStream_t s) {
  hipError_t e = hipSuccess;
  const double sc = axis_scale(c, axis, inverse);
  switch (axis) {
    case 0: e = run_rows(c, c.z, inverse, data, z_geom(c, sc), s); break;
    case 1: e = run_cols(c, c.y, inverse, data, y_geom(c, sc), s); break;
    case 2: e = run_cols(c, c.x, inverse, data, x_geom(c, sc), s); break;
    default: return GFFT_ERR_INVALID_ARG;
  }
  return e == hipSuccess ? GFFT_OK : GFFT_ERR_HIP;
}

int exec_r2c(const PlanCore& c, void* data, hipStream_t s) {
  for (int axis = 0; axis < 3; ++axis) {
    const int st = exec_axis(c, axis, false, data, s);
    if (st != GFFT_OK) return st;
  }
  return GFFT_OK;
}

int exec_c2r(const PlanCore& c, void* data, hipStream_t s) {
  for (int axis = 2; axis >= 0; --axis) {
    const int st = exec_axis(c, axis, true, data, s);
    if (st != GFFT_OK) return st;
  }
  return GFFT_OK;
}

}  // namespace gfft
