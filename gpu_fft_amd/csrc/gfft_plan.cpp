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

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace gfft {

namespace {

bool factor_smooth(int64_t n, std::vector<int>& radices) {
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
  std::vector<int> big;  // 7, 11, 13 go last
  for (int q : {7, 11, 13})
    while (m % q == 0) {
      m /= q;
      big.push_back(q);
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
  radices.insert(radices.end(), big.begin(), big.end());
  return true;
}

// All compiled variants of a precision, in registration order (units in the
// order below, so a unit's first entry for a size is that size's default).
const std::vector<KernelEntry>& all_entries(int prec) {
  using Unit = int (*)(const KernelEntry**);
  auto gather = [](std::initializer_list<Unit> units) {
    std::vector<KernelEntry> v;
    for (Unit u : units) {
      const KernelEntry* t = nullptr;
      const int n = u(&t);
      v.insert(v.end(), t, t + n);
    }
    return v;
  };
  static const std::vector<KernelEntry> f32 =
      gather({registry_f32, registry_smooth_f32_a, registry_smooth_f32_b, registry_seven_f32, registry_prime_f32});
  static const std::vector<KernelEntry> f64 =
      gather({registry_f64, registry_smooth_f64_a, registry_smooth_f64_b, registry_seven_f64, registry_prime_f64});
  return prec == F32 ? f32 : f64;
}

int registry(int prec, const KernelEntry** tab) {
  const std::vector<KernelEntry>& v = all_entries(prec);
  *tab = v.data();
  return (int)v.size();
}

bool same_sched(const Schedule& a, const Schedule& b) {
  if (a.N != b.N || a.S != b.S) return false;
  for (int i = 0; i < a.S; ++i)
    if (a.R[i] != b.R[i]) return false;
  return true;
}

// All compiled (forward, inverse) pairs for a size, default first (any radix
// order: the plan keeps one twiddle table per distinct order).
std::vector<std::pair<const KernelEntry*, const KernelEntry*>> find_pairs(int prec, int kf, int ki, int n) {
  std::vector<std::pair<const KernelEntry*, const KernelEntry*>> out;
  const KernelEntry* tab = nullptr;
  const int cnt = registry(prec, &tab);
  for (int i = 0; i < cnt; ++i) {
    if (tab[i].kind != kf || tab[i].N != n) continue;
    for (int j = 0; j < cnt; ++j)
      if (tab[j].kind == ki && tab[j].N == n && tab[j].P == tab[i].P && tab[j].tile == tab[i].tile &&
          tab[j].persist == tab[i].persist && tab[j].order == tab[i].order && tab[j].readdr == tab[i].readdr &&
          tab[j].straight == tab[i].straight &&
          same_sched(tab[j].sched, tab[i].sched)) {
        out.emplace_back(&tab[i], &tab[j]);
        break;
      }
  }
  return out;
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
  if (n > (1 << 20) || !factor_smooth(n, rad)) return GFFT_ERR_UNSUPPORTED_SIZE;
  const int kf = rows ? ROWS_R2C : COLS_FWD;
  const int ki = rows ? ROWS_C2R : COLS_INV;
  a.cands = find_pairs(prec, kf, ki, (int)n);
  Schedule def{};
  if (!a.cands.empty()) {
    def = a.cands[0].first->sched;
  } else {
    if ((int)rad.size() > kMaxRadices) return GFFT_ERR_UNSUPPORTED_SIZE;
    def.N = (int)n;
    def.S = (int)rad.size();
    for (size_t i = 0; i < rad.size(); ++i) def.R[i] = rad[i];
  }
  a.scheds.push_back(def);
  for (const auto& cd : a.cands) {
    int k = 0;
    while (k < (int)a.scheds.size() && !same_sched(a.scheds[k], cd.first->sched)) ++k;
    if (k == (int)a.scheds.size()) a.scheds.push_back(cd.first->sched);
    a.cand_sched.push_back(k);
  }
  const size_t lds = rows ? generic_rows_lds(prec, (int)n, (int)K) : generic_cols_lds(prec, (int)n);
  if (lds != 0) {
    a.cands.emplace_back(nullptr, nullptr);  // the generic kernel runs the default order
    a.cand_sched.push_back(0);
  }
  if (a.cands.empty()) return GFFT_ERR_UNSUPPORTED_SIZE;
  a.variant = 0;
  a.reg_fwd = a.cands[0].first;
  a.reg_inv = a.cands[0].second;
  a.sched = def;
  a.rt = to_rt(def);
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
                "\"tile\":%d,\"threads\":%d,\"twiddle_bytes\":%zu,\"variant\":%d,\"candidates\":%d,"
                "\"persist\":%d,\"order\":%d,\"readdr\":%d,\"straight\":%d,\"tuned_ms\":%.6f}",
                name, a.N, a.rows ? "rows_two_for_one" : "strided_columns", e ? "register" : "generic",
                sched_str(a.sched).c_str(), e ? e->P : 0, e ? e->tile : 0, e ? e->threads : 256, a.tw_bytes,
                a.variant, (int)a.cands.size(), e ? e->persist : 0, e ? e->order : 0, e ? e->readdr : 0,
                e ? e->straight : 0, a.tuned_ms);
  return buf;
}

}  // namespace

// Compute units of the current device (cached per device id).
int device_cu_count() {
  static std::atomic<int> cache[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int n = cache[dev].load(std::memory_order_relaxed);
  if (n == 0) {
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cache[dev].store(n, std::memory_order_relaxed);
  }
  return n;
}

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
  if (procs == 1 && c.y.active) {
    const PlaneEntry* tab = nullptr;
    const int cnt = registry_plane(&tab);
    for (int i = 0; i < cnt && !c.plane; ++i)
      if (tab[i].prec == prec && tab[i].nz == nz && tab[i].ny == ny) c.plane = &tab[i];
    c.fused = c.plane ? 1 : 0;
  }
  const size_t cs = prec == F32 ? 8 : 16;
  for (AxisPlan* a : {&c.z, &c.y, &c.x}) {
    if (!a->active) continue;
    a->tw_bytes = 0;
    for (const Schedule& sc : a->scheds) {
      const size_t n = prec == F32 ? make_twiddles<float>(sc).size() : make_twiddles<double>(sc).size();
      a->tw_bytes += n / 2 * cs;
    }
  }
  if (c.plane) {
    for (const Schedule* sc : {&c.plane->schedz, &c.plane->schedy}) {
      const size_t n = prec == F32 ? make_twiddles<float>(*sc).size() : make_twiddles<double>(*sc).size();
      c.plane_tw_bytes += n / 2 * cs;
    }
  }
  return GFFT_OK;
}

size_t plan_device_bytes(const PlanCore& c) {
  return c.z.tw_bytes + c.y.tw_bytes + c.x.tw_bytes + c.plane_tw_bytes;
}

std::string plan_json(const PlanCore& c) {
  char head[256];
  std::snprintf(head, sizeof head,
                "{\"nx\":%lld,\"ny\":%lld,\"nz\":%lld,\"procs\":%d,\"rank\":%d,\"precision\":\"%s\",\"K\":%lld,",
                (long long)c.nx, (long long)c.ny, (long long)c.nz, c.procs, c.rank, c.prec == F32 ? "f32" : "f64",
                (long long)c.K);
  char tail[160];
  std::snprintf(tail, sizeof tail, ",\"zy_plane_kernel\":%d,\"zy_fused\":%d,\"fused_ms\":%.6f,\"unfused_ms\":%.6f}",
                c.plane ? 1 : 0, (c.plane && c.fused) ? 1 : 0, c.fused_ms, c.unfused_ms);
  return std::string(head) + axis_json("z", c.z) + "," + axis_json("y", c.y) + "," + axis_json("x", c.x) + tail;
}

int upload_twiddles(PlanCore& c) {
  for (AxisPlan* a : {&c.z, &c.y, &c.x}) {
    if (!a->active) continue;
    a->tw_bytes = 0;
    for (const Schedule& sc : a->scheds) {
      const std::vector<unsigned char> t = twiddle_table(sc, c.prec);
      void* d = nullptr;
      if (hipMalloc(&d, t.size()) != hipSuccess) {
        (void)hipGetLastError();
        return GFFT_ERR_OUT_OF_MEMORY;
      }
      a->d_tws.push_back(d);
      a->tw_bytes += t.size();
      if (hipMemcpy(d, t.data(), t.size(), hipMemcpyHostToDevice) != hipSuccess) return GFFT_ERR_HIP;
    }
    a->d_tw = a->d_tws[a->cand_sched[a->variant]];
  }
  if (c.plane) {
    c.plane_tw_bytes = 0;
    for (int i = 0; i < 2; ++i) {
      const std::vector<unsigned char> t = twiddle_table(i == 0 ? c.plane->schedz : c.plane->schedy, c.prec);
      void*& d = i == 0 ? c.d_twz_plane : c.d_twy_plane;
      if (hipMalloc(&d, t.size()) != hipSuccess) {
        d = nullptr;
        (void)hipGetLastError();
        return GFFT_ERR_OUT_OF_MEMORY;
      }
      c.plane_tw_bytes += t.size();
      if (hipMemcpy(d, t.data(), t.size(), hipMemcpyHostToDevice) != hipSuccess) return GFFT_ERR_HIP;
    }
  }
  return GFFT_OK;
}

void free_twiddles(PlanCore& c) {
  for (AxisPlan* a : {&c.z, &c.y, &c.x}) {
    for (void* d : a->d_tws)
      if (d) (void)hipFree(d);
    a->d_tws.clear();
    a->d_tw = nullptr;
  }
  for (void** d : {&c.d_twz_plane, &c.d_twy_plane}) {
    if (*d) (void)hipFree(*d);
    *d = nullptr;
  }
}

// y pass over the local x-slab (Nx/p, Ny, K), in place: column c = (x, kz).
ColGeom y_geom(const PlanCore& c, double scale) {
  const int64_t nxl = c.nx / c.procs;
  ColGeom g;
  g.ncols = (uint32_t)(nxl * c.K);
  g.in = g.out = col_map((uint32_t)c.K, c.ny * c.K, c.K);
  g.scale = scale;
  return g;
}

// x pass over the y-slab (Nx, Ny/p, K), in place: column c = (y, kz).
ColGeom x_geom(const PlanCore& c, double scale) {
  const int64_t nyl = c.ny / c.procs;
  ColGeom g;
  g.ncols = (uint32_t)(nyl * c.K);
  g.in = g.out = col_map(g.ncols, 0, nyl * c.K);
  g.scale = scale;
  return g;
}

// The slab-exchange buffers hold p chunks of (Ny/p, Nx/p, K), chunk s for
// (or from) rank s: the reference's __buffer__ after __transpose_slab__
// (transpose.cu:41-54, FFT_R2C.cu:48) and the receive layout that
// __chunk_transpose__ unpacks (transpose.cu:56-73, FFT_R2C.cu:66).
//
// y pass with the pack fused into its stores: x-slab (Nx/p, Ny, K) ->
// (Ny, Nx/p, K); column c = (xl, kz) lands at c + y * (Nx/p) * K.
ColGeom y_pack_geom(const PlanCore& c, double scale) {
  ColGeom g = y_geom(c, scale);
  g.out = col_map(g.ncols, 0, (c.nx / c.procs) * c.K);
  return g;
}

// The inverse: y pass reading (Ny, Nx/p, K), writing the x-slab in place
// order (FFT_C2R.cu:66, __transpose_slab__ back).
ColGeom y_unpack_geom(const PlanCore& c, double scale) {
  ColGeom g = y_geom(c, scale);
  g.in = col_map(g.ncols, 0, (c.nx / c.procs) * c.K);
  return g;
}

// x pass reading the p received chunks: point x = s * (Nx/p) + xl of column
// (yl, kz) sits at s * chunk + (yl * (Nx/p) + xl) * K + kz.
ColMap chunk_map(const PlanCore& c) {
  const int64_t nxl = c.nx / c.procs, nyl = c.ny / c.procs;
  ColMap m = col_map((uint32_t)c.K, nxl * c.K, c.K);
  m.nsub = (int32_t)nxl;
  m.rhi = nxl * nyl * c.K;
  return m;
}

ColGeom x_unpack_geom(const PlanCore& c, double scale) {
  ColGeom g = x_geom(c, scale);
  g.in = chunk_map(c);
  return g;
}

// x inverse with the pack into p chunks fused into its stores
// (__chunk_transpose_inverse__, transpose.cu:75-91, FFT_C2R.cu:48).
ColGeom x_pack_geom(const PlanCore& c, double scale) {
  ColGeom g = x_geom(c, scale);
  g.out = chunk_map(c);
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

hipError_t run_cols(const PlanCore& c, const AxisPlan& a, bool inverse, const void* in, void* out, ColGeom g,
                    hipStream_t s) {
  const KernelEntry* e = inverse ? a.reg_inv : a.reg_fwd;
  if (e) return reinterpret_cast<ColsLaunchFn>(e->fn)(in, out, g, a.d_tw, s);
  return launch_generic_cols(c.prec, inverse ? COLS_INV : COLS_FWD, in, out, g, a.rt, a.d_tw, s);
}

hipError_t run_cols(const PlanCore& c, const AxisPlan& a, bool inverse, void* data, ColGeom g, hipStream_t s) {
  if (!a.active) return hipSuccess;
  return run_cols(c, a, inverse, data, data, g, s);
}

// The output scale of a direction is fused into its last active pass:
// R2C ends with x (or y, or z when both strided axes have length 1), C2R
// always ends with z.
double axis_scale(const PlanCore& c, int axis, bool inverse) {
  if (inverse) return axis == 0 ? c.c2r_scale : 1.0;
  const int last = c.x.active ? 2 : (c.y.active ? 1 : 0);
  return axis == last ? c.r2c_scale : 1.0;
}

// Walk direction of a pass.  With alternating walks (the default mask) the
// passes of a round trip run R2C z forward, y reversed, x forward; C2R x
// reversed, y forward, z reversed: each pass starts on the tiles the
// previous pass wrote last, which are still in the memory-side cache.
bool reversed_walk(const PlanCore& c, int axis, bool inverse) {
  return (c.walk_mask >> ((inverse ? 3 : 0) + axis)) & 1;
}

bool use_plane(const PlanCore& c) { return c.fused && c.plane && c.d_twz_plane && c.d_twy_plane; }

hipError_t run_plane(const PlanCore& c, bool inverse, void* data, hipStream_t s) {
  const PlaneLaunchFn f = reinterpret_cast<PlaneLaunchFn>(inverse ? c.plane->fn_c2r : c.plane->fn_r2c);
  return f(data, c.nx, inverse ? axis_scale(c, 0, true) : axis_scale(c, 1, false), c.d_twz_plane, c.d_twy_plane,
           s);
}

int exec_axis(const PlanCore& c, int axis, bool inverse, void* data, hipStream_t s) {
  hipError_t e = hipSuccess;
  const double sc = axis_scale(c, axis, inverse);
  const int32_t rev = reversed_walk(c, axis, inverse) ? 1 : 0;
  switch (axis) {
    case 0: {
      RowGeom g = z_geom(c, sc);
      g.rev = rev;
      e = run_rows(c, c.z, inverse, data, g, s);
      break;
    }
    case 1: {
      ColGeom g = y_geom(c, sc);
      g.rev = rev;
      e = run_cols(c, c.y, inverse, data, g, s);
      break;
    }
    case 2: {
      ColGeom g = x_geom(c, sc);
      g.rev = rev;
      e = run_cols(c, c.x, inverse, data, g, s);
      break;
    }
    case 3: {
      if (!use_plane(c)) return GFFT_ERR_INVALID_ARG;
      e = run_plane(c, inverse, data, s);
      break;
    }
    default: return GFFT_ERR_INVALID_ARG;
  }
  return e == hipSuccess ? GFFT_OK : GFFT_ERR_HIP;
}

AxisPlan* axis_ptr(PlanCore& c, int axis) {
  return axis == 0 ? &c.z : axis == 1 ? &c.y : axis == 2 ? &c.x : nullptr;
}

int set_variant(PlanCore& c, int axis, int variant) {
  AxisPlan* a = axis_ptr(c, axis);
  if (!a || !a->active || variant < 0 || variant >= (int)a->cands.size()) return GFFT_ERR_INVALID_ARG;
  a->variant = variant;
  a->reg_fwd = a->cands[variant].first;
  a->reg_inv = a->cands[variant].second;
  a->sched = a->scheds[a->cand_sched[variant]];
  if (!a->d_tws.empty()) a->d_tw = a->d_tws[a->cand_sched[variant]];
  a->tuned_ms = -1;
  return GFFT_OK;
}

int set_fused(PlanCore& c, int on) {
  if (on && !c.plane) return GFFT_ERR_INVALID_ARG;
  c.fused = on ? 1 : 0;
  return GFFT_OK;
}

namespace {
// Tuning results, process-wide: key = device tag | geometry, value = the
// chosen candidate's description (not its index, so an exported cache stays
// meaningful for a library with a different variant list).
std::mutex g_tune_mu;
std::map<std::string, std::string> g_tune_cache;

// "<arch>/<CUs>" of the current device (cached per device ordinal)
std::string device_tag() {
  static std::mutex mu;
  static std::map<int, std::string> tags;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return "unknown";
  std::lock_guard<std::mutex> lk(mu);
  auto it = tags.find(dev);
  if (it != tags.end()) return it->second;
  hipDeviceProp_t p{};
  std::string t = "unknown";
  if (hipGetDeviceProperties(&p, dev) == hipSuccess) t = std::string(p.gcnArchName) + "/" + std::to_string(p.multiProcessorCount);
  for (char& ch : t)
    if (ch == ' ' || ch == '|') ch = '_';
  tags[dev] = t;
  return t;
}

std::string tune_key(const PlanCore& c, int axis) {
  char b[256];
  if (axis == 0) {
    const RowGeom g = z_geom(c, 1.0);
    std::snprintf(b, sizeof b, "p%d:rows:%d:%lld", c.prec, c.z.N, (long long)g.nrows);
  } else {
    const ColGeom g = axis == 1 ? y_geom(c, 1.0) : x_geom(c, 1.0);
    std::snprintf(b, sizeof b, "p%d:cols:%d:%u:%u:%lld:%lld", c.prec, axis == 1 ? c.y.N : c.x.N, g.ncols,
                  g.in.inner, (long long)g.in.outer, (long long)g.in.rlo);
  }
  return device_tag() + "|" + b;
}

std::string plane_key(const PlanCore& c) {
  char b[128];
  std::snprintf(b, sizeof b, "p%d:plane:%lld:%lld:%lld", c.prec, (long long)c.nx, (long long)c.ny, (long long)c.nz);
  return device_tag() + "|" + b;
}

// description of candidate v of an axis: kernel shape and radix order
std::string cand_desc(const AxisPlan& a, int v) {
  const KernelEntry* e = a.cands[v].first;
  if (!e) return "generic";
  std::string r;
  for (int i = 0; i < e->sched.S; ++i) r += (i ? "." : "") + std::to_string(e->sched.R[i]);
  char b[160];
  std::snprintf(b, sizeof b, "P%d:T%d:W%d:K%d:O%d:A%d:S%d:R%s", e->P, e->tile, e->threads, e->persist, e->order,
                e->readdr, e->straight, r.c_str());
  return b;
}
}  // namespace

int apply_tune_cache(PlanCore& c) {
  std::string keys[3];
  for (int axis = 0; axis < 3; ++axis) {
    const AxisPlan* a = axis_ptr(c, axis);
    if (a->active && a->cands.size() >= 2) keys[axis] = tune_key(c, axis);
  }
  const std::string kb = c.plane ? plane_key(c) : std::string();
  int applied = 0;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (int axis = 0; axis < 3; ++axis) {
    if (keys[axis].empty()) continue;
    auto it = g_tune_cache.find(keys[axis]);
    if (it == g_tune_cache.end()) continue;
    AxisPlan* a = axis_ptr(c, axis);
    for (int v = 0; v < (int)a->cands.size(); ++v)
      if (cand_desc(*a, v) == it->second) {
        set_variant(c, axis, v);
        ++applied;
        break;
      }
  }
  if (!kb.empty()) {
    auto it = g_tune_cache.find(kb);
    if (it != g_tune_cache.end() && (it->second == "fused:0" || it->second == "fused:1")) {
      c.fused = it->second == "fused:1" ? 1 : 0;
      ++applied;
    }
  }
  return applied;
}

std::string tune_cache_text() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  std::string out;
  for (const auto& kv : g_tune_cache) out += kv.first + " " + kv.second + "\n";
  return out;
}

int tune_cache_load(const char* text) {
  std::map<std::string, std::string> add;
  const char* p = text;
  while (*p) {
    const char* e = p;
    while (*e && *e != '\n') ++e;
    std::string line(p, e);
    p = *e ? e + 1 : e;
    while (!line.empty() && (line.back() == '\r' || line.back() == ' ')) line.pop_back();
    if (line.empty() || line[0] == '#') continue;
    const size_t sp = line.find(' ');
    if (sp == std::string::npos || sp == 0 || sp + 1 >= line.size() ||
        line.find(' ', sp + 1) != std::string::npos || line.find('|') == std::string::npos)
      return GFFT_ERR_INVALID_ARG;
    add[line.substr(0, sp)] = line.substr(sp + 1);
  }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (auto& kv : add) g_tune_cache[kv.first] = kv.second;
  return GFFT_OK;
}

void tune_cache_clear() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache.clear();
}

// Time every candidate of every axis (forward + inverse pass on `scratch`,
// whose contents are destroyed) and keep the fastest.  Results are cached per
// geometry for the life of the process.
int tune(PlanCore& c, void* scratch, hipStream_t s, int reps) {
  if (reps < 1) reps = 1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return GFFT_ERR_HIP;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return GFFT_ERR_HIP;
  }
  int st = GFFT_OK;
  std::vector<float> iso[3];  // isolated pair times per candidate, axes timed here
  for (int axis = 0; axis < 3 && st == GFFT_OK; ++axis) {
    AxisPlan* a = axis_ptr(c, axis);
    if (!a->active || a->cands.size() < 2) continue;
    const std::string key = tune_key(c, axis);
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto it = g_tune_cache.find(key);
      int hit = -1;
      if (it != g_tune_cache.end())
        for (int v = 0; v < (int)a->cands.size() && hit < 0; ++v)
          if (cand_desc(*a, v) == it->second) hit = v;
      if (hit >= 0) {
        set_variant(c, axis, hit);
        continue;
      }
    }
    int best = 0;
    float best_ms = 1e30f;
    for (int v = 0; v < (int)a->cands.size() && st == GFFT_OK; ++v) {
      set_variant(c, axis, v);
      // warm-up pair, then `reps` timed pairs
      if ((st = exec_axis(c, axis, false, scratch, s)) != GFFT_OK) break;
      if ((st = exec_axis(c, axis, true, scratch, s)) != GFFT_OK) break;
      if (hipEventRecord(e0, s) != hipSuccess) st = GFFT_ERR_HIP;
      for (int r = 0; r < reps && st == GFFT_OK; ++r) {
        st = exec_axis(c, axis, false, scratch, s);
        if (st == GFFT_OK) st = exec_axis(c, axis, true, scratch, s);
      }
      if (st != GFFT_OK) break;
      float ms = 0;
      if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        st = GFFT_ERR_HIP;
        break;
      }
      ms /= reps;
      iso[axis].push_back(ms);
      if (ms < best_ms) {
        best_ms = ms;
        best = v;
      }
    }
    if (st != GFFT_OK) break;
    set_variant(c, axis, best);
    a->tuned_ms = best_ms;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = cand_desc(*a, best);
  }
  // In-context refinement: an axis timed alone misses how its walk meets the
  // neighbouring passes' (the memory-side cache still holds the tiles a pass
  // wrote last), so the three fastest candidates of each axis timed above
  // are timed again inside whole R2C + C2R round trips.
  if (st == GFFT_OK && c.procs == 1) {
    const int fused_keep = c.fused;
    c.fused = 0;  // the z and y choices matter only without the plane kernels
    auto round_trip_ms = [&](float* ms) -> int {
      int r = exec_r2c(c, scratch, s);
      if (r == GFFT_OK) r = exec_c2r(c, scratch, s);
      if (r != GFFT_OK) return r;
      if (hipEventRecord(e0, s) != hipSuccess) return GFFT_ERR_HIP;
      for (int k = 0; k < reps && r == GFFT_OK; ++k) {
        r = exec_r2c(c, scratch, s);
        if (r == GFFT_OK) r = exec_c2r(c, scratch, s);
      }
      if (r != GFFT_OK) return r;
      if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(ms, e0, e1) != hipSuccess)
        return GFFT_ERR_HIP;
      *ms /= reps;
      return GFFT_OK;
    };
    for (int axis = 0; axis < 3 && st == GFFT_OK; ++axis) {
      const std::vector<float>& t = iso[axis];
      if (t.size() < 2) continue;
      AxisPlan* a = axis_ptr(c, axis);
      std::vector<int> order(t.size());
      for (int v = 0; v < (int)t.size(); ++v) order[v] = v;
      std::sort(order.begin(), order.end(), [&](int x, int y) { return t[x] < t[y]; });
      int best = a->variant;
      float best_rt = 0;
      if ((st = round_trip_ms(&best_rt)) != GFFT_OK) break;
      for (int k = 0; k < 3 && k < (int)order.size() && st == GFFT_OK; ++k) {
        if (order[k] == best) continue;
        set_variant(c, axis, order[k]);
        float rt = 0;
        if ((st = round_trip_ms(&rt)) != GFFT_OK) break;
        if (rt < 0.997f * best_rt) {  // a margin against timing noise
          best_rt = rt;
          best = order[k];
        }
      }
      if (st != GFFT_OK) break;
      set_variant(c, axis, best);
      a->tuned_ms = t[best];
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_tune_cache[tune_key(c, axis)] = cand_desc(*a, best);
    }
    c.fused = fused_keep;
  }
  // fused z + y plane kernels against the (tuned) separate passes
  if (st == GFFT_OK && c.plane && c.d_twz_plane) {
    const std::string kb = plane_key(c);
    int cached = -1;
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto it = g_tune_cache.find(kb);
      if (it != g_tune_cache.end() && (it->second == "fused:0" || it->second == "fused:1"))
        cached = it->second == "fused:1" ? 1 : 0;
    }
    if (cached >= 0) {
      c.fused = cached;
    } else {
      float ms[2] = {0, 0};
      for (int mode = 0; mode < 2 && st == GFFT_OK; ++mode) {
        // whole round trips (the x pass included: its launch overlaps
        // differently with one or two z/y launches before it)
        c.fused = mode;
        auto once = [&]() -> int {
          const int r = exec_r2c(c, scratch, s);
          return r == GFFT_OK ? exec_c2r(c, scratch, s) : r;
        };
        if ((st = once()) != GFFT_OK) break;
        if (hipEventRecord(e0, s) != hipSuccess) st = GFFT_ERR_HIP;
        for (int r = 0; r < reps && st == GFFT_OK; ++r) st = once();
        if (st != GFFT_OK) break;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms[mode], e0, e1) != hipSuccess) {
          st = GFFT_ERR_HIP;
          break;
        }
        ms[mode] /= reps;
      }
      c.fused = 1;
      if (st == GFFT_OK) {
        c.unfused_ms = ms[0];
        c.fused_ms = ms[1];
        c.fused = ms[1] < ms[0] ? 1 : 0;
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[kb] = c.fused ? "fused:1" : "fused:0";
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return st;
}

// z and y passes over a block of `planes` consecutive local x-planes,
// alternating per block so that a block's z output is still in the
// memory-side cache (MALL) when its y pass reads it.
int exec_zy_chunked(const PlanCore& c, bool inverse, void* data, hipStream_t s) {
  const int64_t nxl = c.nx / c.procs;
  const size_t rs = c.prec == F32 ? 4 : 8;
  const double zs = axis_scale(c, 0, inverse), ys = axis_scale(c, 1, inverse);
  for (int64_t x0 = 0; x0 < nxl; x0 += c.chunk_planes) {
    const int64_t B = std::min<int64_t>(c.chunk_planes, nxl - x0);
    char* base = static_cast<char*>(data) + (size_t)x0 * c.ny * 2 * c.K * rs;
    RowGeom zg = z_geom(c, zs);
    zg.nrows = B * c.ny;
    ColGeom yg = y_geom(c, ys);
    yg.ncols = (uint32_t)(B * c.K);
    hipError_t e;
    if (!inverse) {
      if ((e = run_rows(c, c.z, false, base, zg, s)) != hipSuccess) return GFFT_ERR_HIP;
      if ((e = run_cols(c, c.y, false, base, yg, s)) != hipSuccess) return GFFT_ERR_HIP;
    } else {
      if ((e = run_cols(c, c.y, true, base, yg, s)) != hipSuccess) return GFFT_ERR_HIP;
      if ((e = run_rows(c, c.z, true, base, zg, s)) != hipSuccess) return GFFT_ERR_HIP;
    }
  }
  return GFFT_OK;
}

bool use_chunks(const PlanCore& c) { return c.chunk_planes > 0 && c.chunk_planes < c.nx / c.procs; }

int exec_r2c(const PlanCore& c, void* data, hipStream_t s) {
  if (use_plane(c)) {
    if (run_plane(c, false, data, s) != hipSuccess) return GFFT_ERR_HIP;
    return exec_axis(c, 2, false, data, s);
  }
  if (use_chunks(c)) {
    const int st = exec_zy_chunked(c, false, data, s);
    if (st != GFFT_OK) return st;
    return exec_axis(c, 2, false, data, s);
  }
  const int order[3] = {0, (c.pass_order & 1) ? 2 : 1, (c.pass_order & 1) ? 1 : 2};
  for (int axis : order) {
    const int st = exec_axis(c, axis, false, data, s);
    if (st != GFFT_OK) return st;
  }
  return GFFT_OK;
}

int exec_c2r(const PlanCore& c, void* data, hipStream_t s) {
  if (use_plane(c)) {
    const int st = exec_axis(c, 2, true, data, s);
    if (st != GFFT_OK) return st;
    return run_plane(c, true, data, s) == hipSuccess ? GFFT_OK : GFFT_ERR_HIP;
  }
  if (use_chunks(c)) {
    const int st = exec_axis(c, 2, true, data, s);
    if (st != GFFT_OK) return st;
    return exec_zy_chunked(c, true, data, s);
  }
  const int order[3] = {(c.pass_order & 2) ? 1 : 2, (c.pass_order & 2) ? 2 : 1, 0};
  for (int axis : order) {
    const int st = exec_axis(c, axis, true, data, s);
    if (st != GFFT_OK) return st;
  }
  return GFFT_OK;
}

std::vector<unsigned char> twiddle_table(const Schedule& s, int prec) {
  std::vector<unsigned char> out;
  if (prec == F32) {
    const std::vector<float> t = make_twiddles<float>(s);
    out.resize(t.size() * sizeof(float));
    std::memcpy(out.data(), t.data(), out.size());
  } else {
    const std::vector<double> t = make_twiddles<double>(s);
    out.resize(t.size() * sizeof(double));
    std::memcpy(out.data(), t.data(), out.size());
  }
  return out;
}

// ---------------------------------------------------------------------------
// The two halves of each direction around a slab exchange (any p, incl. 1).
// ---------------------------------------------------------------------------
size_t local_bytes(const PlanCore& c) {
  return (size_t)(c.nx / c.procs) * (size_t)c.ny * (size_t)c.K * (c.prec == F32 ? 8u : 16u);
}

namespace {
int copy_d2d(void* dst, const void* src, size_t bytes, hipStream_t s) {
  return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s) == hipSuccess ? GFFT_OK : GFFT_ERR_HIP;
}
}  // namespace

// FFT_R2C.cu:45-48: z R2C and y C2C on the local x-slab, the slab transpose
// fused into the y pass's stores (data -> sendbuf).
int exec_r2c_local(const PlanCore& c, void* data, void* sendbuf, hipStream_t s) {
  if (run_rows(c, c.z, false, data, z_geom(c, axis_scale(c, 0, false)), s) != hipSuccess) return GFFT_ERR_HIP;
  if (!c.y.active) return copy_d2d(sendbuf, data, local_bytes(c), s);  // Ny = 1: same order
  return run_cols(c, c.y, false, data, sendbuf, y_pack_geom(c, axis_scale(c, 1, false)), s) == hipSuccess
             ? GFFT_OK
             : GFFT_ERR_HIP;
}

// FFT_R2C.cu:66-69: the chunk transpose fused into the x pass's loads
// (recvbuf -> data, the y-slab output).
int exec_r2c_finish(const PlanCore& c, const void* recvbuf, void* out, hipStream_t s) {
  if (!c.x.active) return copy_d2d(out, recvbuf, local_bytes(c), s);  // Nx = 1 (then p = 1)
  return run_cols(c, c.x, false, recvbuf, out, x_unpack_geom(c, axis_scale(c, 2, false)), s) == hipSuccess
             ? GFFT_OK
             : GFFT_ERR_HIP;
}

// FFT_C2R.cu:45-48: x inverse with the chunk pack fused into its stores.
int exec_c2r_start(const PlanCore& c, const void* data, void* sendbuf, hipStream_t s) {
  if (!c.x.active) return copy_d2d(sendbuf, data, local_bytes(c), s);
  return run_cols(c, c.x, true, data, sendbuf, x_pack_geom(c, axis_scale(c, 2, true)), s) == hipSuccess
             ? GFFT_OK
             : GFFT_ERR_HIP;
}

// FFT_C2R.cu:66-69: y inverse reading the received slab-transposed layout,
// then z C2R in place on the x-slab.
int exec_c2r_local(const PlanCore& c, const void* recvbuf, void* data, hipStream_t s) {
  int st = GFFT_OK;
  if (!c.y.active)
    st = copy_d2d(data, recvbuf, local_bytes(c), s);
  else if (run_cols(c, c.y, true, recvbuf, data, y_unpack_geom(c, axis_scale(c, 1, true)), s) != hipSuccess)
    st = GFFT_ERR_HIP;
  if (st != GFFT_OK) return st;
  return run_rows(c, c.z, true, data, z_geom(c, axis_scale(c, 0, true)), s) == hipSuccess ? GFFT_OK : GFFT_ERR_HIP;
}

}  // namespace gfft
