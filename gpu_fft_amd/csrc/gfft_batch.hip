// Batched execution: many independent 3D R2C / C2R transforms of uneven
// sizes (edges 32..256) as one unit of work, eagerly or as one hipGraph.
//
// Work is cut into tiles of axis passes.  Transforms are grouped by size
// class (all three edges of a transform belong to one class).  For a class,
// one "level" is one axis pass (R2C: z, y, x; C2R: x, y, z) of EVERY
// transform of that class, run as a single launch whose workgroups stride
// over the concatenated tile lists of those transforms.  The levels of a
// class form a chain (stream order / graph edges); the chains of different
// classes are independent and run side by side (eager: up to kLanes internal
// streams forked from and joined back into the caller's stream; graph: one
// branch per class).  There are no waits between workgroups, so a launch
// always drains.
//
// Each class has kVariants compiled variants (radix order, and points per
// thread, which sets the pencils per 256-thread tile) plus one more choice,
// kPlainVariant: the class's transforms run one after another through
// single-transform plans (persistent pass kernels, fused z+y plane kernels),
// which beats level launches for a few large transforms.  gfft_batch_tune
// times every choice on the batch's own data and keeps the fastest per
// class.  Every tile runs the same register-engine tile code as the
// single-transform kernels.
#include <algorithm>
#include <map>
#include <memory>
#include <new>
#include <tuple>
#include <vector>

#include "gfft_batch_tiles.h"
#include "gfft_plan.h"
#include "gpufft/gpufft.h"

namespace gfft {
namespace {

using namespace batch;

// Workgroups stride over the level's tiles; tile k belongs to the transform
// t with prefix[t] <= k < prefix[t + 1].
template <int C, int V>
__global__ __launch_bounds__(kBT) void k_batch_level(LevelArgs a) {
  using F = BCfg<C, V>;
  __shared__ cx<float> smem[F::lds_entries()];
  for (uint32_t k = blockIdx.x; k < a.total; k += gridDim.x) {
    int lo = 0, hi = a.count;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (a.prefix[mid] <= k)
        lo = mid;
      else
        hi = mid;
    }
    run_tile<F>(a.xf[lo], a.axis, a.inv != 0, k - a.prefix[lo], a.tw, smem);
    __syncthreads();  // the tile's LDS reads are done before the next tile's writes
  }
}

// Fused z + y level of a plane class: one x-plane per tile, in LDS.
template <int C, int V>
__global__ __launch_bounds__(kBT) void k_batch_plane(LevelArgs a) {
  using PP = BPlane<C, V>;
  static_assert(PP::NT == kBT, "plane pass shape");
  __shared__ cx<float> smem[PP::PLANE + PP::SCRATCH];
  for (uint32_t k = blockIdx.x; k < a.total; k += gridDim.x) {
    int lo = 0, hi = a.count;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (a.prefix[mid] <= k)
        lo = mid;
      else
        hi = mid;
    }
    plane_tile<PP>(a.xf[lo], a.inv != 0, k - a.prefix[lo], a.tw, smem);
  }
}

// Fused z + y level of the 32 and 48 classes: one x-plane per wavefront
// (four per workgroup), each wavefront striding over the level's planes on
// its own and prefetching its next one (wave_planes).
template <int C, int V>
__global__ __launch_bounds__(kBT) void k_batch_plane_w(LevelArgs a) {
  using PP = BPlaneW<C, V>;
  constexpr uint32_t NW = kBT / 64;
  __shared__ cx<float> smem[NW * (PP::PLANE + PP::SCRATCH)];
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t k0 = blockIdx.x * NW + w, dk = gridDim.x * NW;
  wave_planes<PP>(
      [&](uint32_t i, const BXform** px, uint32_t* plane) {
        const uint32_t k = k0 + i * dk;
        if (k >= a.total) return false;
        int lo = 0, hi = a.count;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (a.prefix[mid] <= k)
            lo = mid;
          else
            hi = mid;
        }
        *px = a.xf + lo;
        *plane = k - a.prefix[lo];
        return true;
      },
      a.inv != 0, a.tw, smem + w * (PP::PLANE + PP::SCRATCH));
}

// Planes per workgroup of a plane level's kernel.
template <int C>
constexpr uint32_t planes_per_wg() {
  return C <= 1 ? kBT / 64 : 1u;
}

using LaunchFn = hipError_t (*)(dim3, const LevelArgs&, hipStream_t);

// Launches report their own status (hipLaunchKernel's return), not the
// thread's last error, which an earlier tolerated failure may have left.
template <int C, int V>
hipError_t launch_cv(dim3 grid, const LevelArgs& a, hipStream_t s) {
  void* args[] = {const_cast<LevelArgs*>(&a)};
  return hipLaunchKernel(reinterpret_cast<const void*>(&k_batch_level<C, V>), grid, dim3(kBT), args, 0, s);
}

template <int C, int V>
hipError_t launch_plane_cv(dim3 grid, const LevelArgs& a, hipStream_t s) {
  void* args[] = {const_cast<LevelArgs*>(&a)};
  const void* fn = nullptr;
  if constexpr (C <= 1)
    fn = reinterpret_cast<const void*>(&k_batch_plane_w<C, V>);
  else
    fn = reinterpret_cast<const void*>(&k_batch_plane<C, V>);
  return hipLaunchKernel(fn, grid, dim3(kBT), args, 0, s);
}

struct VarInfo {
  Schedule sched;
  int P, W;
  const void* kernel;
  LaunchFn launch;
  const void* plane_kernel;  // fused z + y level (plane classes), else null
  LaunchFn plane_launch;
  uint32_t plane_tiles;      // planes per workgroup of the plane kernel
};

template <int C, int V>
VarInfo var_info() {
  using F = BCfg<C, V>;
  VarInfo vi{make_sched<typename F::SC>(), F::P, F::W, reinterpret_cast<const void*>(&k_batch_level<C, V>),
             &launch_cv<C, V>, nullptr, nullptr, 1u};
  if constexpr (C < kPlaneClasses) {
    if constexpr (C <= 1)
      vi.plane_kernel = reinterpret_cast<const void*>(&k_batch_plane_w<C, V>);
    else
      vi.plane_kernel = reinterpret_cast<const void*>(&k_batch_plane<C, V>);
    vi.plane_launch = &launch_plane_cv<C, V>;
    vi.plane_tiles = planes_per_wg<C>();
  }
  return vi;
}

template <int C>
struct VarRow {
  static void fill(VarInfo* row) {
    row[0] = var_info<C, 0>();
    row[1] = var_info<C, 1>();
    row[2] = var_info<C, 2>();
  }
};

const VarInfo& var(int c, int v) {
  static const auto tab = [] {
    std::vector<VarInfo> t(kClasses * kVariants);
    VarRow<0>::fill(&t[0 * kVariants]);
    VarRow<1>::fill(&t[1 * kVariants]);
    VarRow<2>::fill(&t[2 * kVariants]);
    VarRow<3>::fill(&t[3 * kVariants]);
    VarRow<4>::fill(&t[4 * kVariants]);
    VarRow<5>::fill(&t[5 * kVariants]);
    VarRow<6>::fill(&t[6 * kVariants]);
    VarRow<7>::fill(&t[7 * kVariants]);
    return t;
  }();
  return tab[c * kVariants + v];
}

}  // namespace
}  // namespace gfft

using gfft::batch::BXform;

// Transforms are grouped by size class; group c occupies positions
// [off[c], off[c] + n[c]) of the device descriptor array.
struct gfft_batch_s {
  int count = 0;
  int device = 0;
  std::vector<int64_t> dims;   // 3 per transform, caller order
  std::vector<size_t> bytes;   // data bytes per transform, caller order
  std::vector<int> pos;        // caller index -> sorted position
  std::vector<BXform> hx;      // sorted by class
  int off[gfft::batch::kClasses] = {}, n[gfft::batch::kClasses] = {};
  int var[gfft::batch::kClasses] = {};  // chosen variant per class
  BXform* d_xf = nullptr;
  uint32_t* d_prefix = nullptr;  // [class][variant][mode][level][n_c + 1], packed
  size_t prefix_off[gfft::batch::kClasses][gfft::batch::kVariants][3] = {};
  uint32_t level_tiles[gfft::batch::kClasses][gfft::batch::kVariants][3][gfft::batch::kMaxLevels] = {};
  void* d_tw[gfft::batch::kClasses][gfft::batch::kVariants] = {};
  int occ[gfft::batch::kClasses][gfft::batch::kVariants] = {};        // resident workgroups per CU
  int occ_plane[gfft::batch::kClasses][gfft::batch::kVariants] = {};  // the same for the plane levels
  float tuned_ms[2] = {-1.0f, -1.0f};  // last tune: ms per execution, persistent and levels
  int lane[gfft::batch::kClasses] = {};  // internal stream of the class's chain (eager path)
  int cus = 0;
  bool data_set = false;
  hipStream_t ws[gfft::batch::kLanes] = {};
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_join[gfft::batch::kLanes] = {};
  hipGraph_t graph[3] = {};
  hipGraphExec_t gexec[3] = {};
  // single-transform plans (one per distinct geometry) of the plain variant
  std::map<std::tuple<int64_t, int64_t, int64_t>, std::unique_ptr<gfft::PlanCore>> plans;
  std::vector<const gfft::PlanCore*> xplan;  // per sorted position
  // persistent execution (default): one launch per mode following a plan
  int sched = GFFT_BATCH_PERSISTENT;
  int pset = 0, pteam = 2;            // variant set, team factor (a team above pteam slot shares of
                                      // work; 0: global-level plans, every workgroup in every transform)
  int pspan = 2;                      // team plans: XCDs a team may span
  struct DevPlan {
    gfft::PSeg* seg = nullptr;
    gfft::PTeam* team = nullptr;
    gfft::PPart* parts = nullptr;
    gfft::PEntry* local = nullptr;
    uint32_t* solo = nullptr;
    uint32_t* team_ids = nullptr;
    uint32_t* dom_off = nullptr;
    uint32_t* dom_teams = nullptr;
    uint32_t nslots = 0;  // workgroups of the launch (domains * per_dom; 0: nothing to run)
    uint32_t domains = 1, per_dom = 1, nsolo = 0, claim_words = 0, tiles = 0, teams_k2 = 0, members = 0;
    // global-level plan (team factor 0)
    gfft::WShare* wshare = nullptr;
    uint32_t* lvl_off = nullptr;
    uint32_t* own_off = nullptr;
    uint32_t* own = nullptr;
    uint32_t nlevels = 0, nshares = 0, level_base = 0;
    size_t ctr_words = 0;  // counter words the launch uses
  } plan[3];
  uint32_t* d_ctr = nullptr;          // counter words, sized for the largest plan
  size_t ctr_words = 0;
  uint32_t* h_abort = nullptr;        // pinned host word read by the kernel's waits
  double wait_limit_s = 2.0;
  int wall_khz = 0;                   // device wall-clock rate
  size_t plds = 0;                    // dynamic LDS of the current set for the classes present
  int pocc = 1;                       // resident persistent workgroups per CU
  int xccs = 1;                       // L2 domains (XCDs) of the device
};

namespace {

// Levels of each mode: R2C z, y, (z + y fused), x; C2R the reverse.  A
// transform has either the z and y levels or the fused one (BXform::plane).
const int kModeAxis[3][gfft::batch::kMaxLevels] = {{0, 1, 3, 2}, {2, 3, 1, 0}, {0, 1, 3, 2, 2, 3, 1, 0}};
const int kModeInv[3][gfft::batch::kMaxLevels] = {{0, 0, 0, 0}, {1, 1, 1, 1}, {0, 0, 0, 0, 1, 1, 1, 1}};
const int kModeLevels[3] = {4, 4, 8};
// Shares per team member and level in the persistent plans: a member runs
// its own in turn, so a slow member's later ones can be taken by the others
// (and by helpers).  Three measured ~1 % faster than two on the bench mix
// with helpers and span teams; four was slower (up to +15 % with spans).
constexpr uint32_t kSharesPerMember = 3;
// Global-level plans: about three shares per slot and level (the level's
// tail is then short), and the transforms below a quarter of a slot share of work in
// the solo queue, run whole while workgroups wait at level hand-overs.
constexpr uint32_t kWideSharesPerSlot = 3;
constexpr double kWideSoloFactor = 0.25;
// plane tiles' cost per value moved relative to row and column tiles (bench
// mix, global-level plan: weight 1 0.737 ms, 1.4 0.74, 1.8 0.71, 2.5 0.72)
constexpr double kWidePlaneWeight = 1.8;

// L2 domains (XCDs) of the current device; 1 where the count is unknown or
// above what a plan holds (the kernel then gives team roles only to the
// workgroups of XCC 0, which share one L2).
int device_xcc_count() {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess) {
    (void)hipGetLastError();
    return 1;
  }
  return n >= 1 && n <= (int)gfft::kMaxDomains ? n : 1;
}

uint32_t tiles_of(const BXform& x, int axis, int w, int max_plane = gfft::batch::kPlaneClasses) {
  if (gfft::batch::level_cls(x, axis, max_plane) < 0) return 0;
  if (axis == gfft::batch::kAxisPlane) return (uint32_t)(x.nrows / x.ny);  // one tile per x-plane
  if (axis == 0) return (uint32_t)(((x.nrows + 1) / 2 + w - 1) / w);
  const uint32_t nc = axis == 1 ? x.ncols_y : x.ncols_x;
  return (nc + w - 1) / w;
}

// An executable may still run on any stream the caller launched it on (or
// be queued behind other work there), so the whole device is synchronised
// before any executable or graph is destroyed.
void drop_graphs(gfft_batch_s* b) {
  bool any = false;
  for (int m = 0; m < 3; ++m) any = any || b->gexec[m] || b->graph[m];
  if (!any) return;
  (void)hipDeviceSynchronize();
  for (int m = 0; m < 3; ++m) {
    if (b->gexec[m]) (void)hipGraphExecDestroy(b->gexec[m]);
    if (b->graph[m]) (void)hipGraphDestroy(b->graph[m]);
    b->gexec[m] = nullptr;
    b->graph[m] = nullptr;
  }
}

bool any_plain(const gfft_batch_s* b) {
  for (int c = 0; c < gfft::batch::kClasses; ++c)
    if (b->n[c] && b->var[c] == gfft::batch::kPlainVariant) return true;
  return false;
}

void free_plans(gfft_batch_s* b) {
  for (auto& p : b->plan) {
    if (p.seg) (void)hipFree(p.seg);
    if (p.team) (void)hipFree(p.team);
    if (p.parts) (void)hipFree(p.parts);
    if (p.local) (void)hipFree(p.local);
    if (p.solo) (void)hipFree(p.solo);
    if (p.team_ids) (void)hipFree(p.team_ids);
    if (p.dom_off) (void)hipFree(p.dom_off);
    if (p.dom_teams) (void)hipFree(p.dom_teams);
    if (p.wshare) (void)hipFree(p.wshare);
    if (p.lvl_off) (void)hipFree(p.lvl_off);
    if (p.own_off) (void)hipFree(p.own_off);
    if (p.own) (void)hipFree(p.own);
    p = gfft_batch_s::DevPlan{};
  }
}

void release(gfft_batch_s* b) {
  drop_graphs(b);
  if (b->d_ctr || b->plan[0].seg || b->plan[1].seg || b->plan[2].seg) (void)hipDeviceSynchronize();  // launches read them
  free_plans(b);
  if (b->d_ctr) (void)hipFree(b->d_ctr);
  if (b->h_abort) (void)hipHostFree(b->h_abort);
  for (auto& kv : b->plans) gfft::free_twiddles(*kv.second);
  b->plans.clear();
  for (int i = 0; i < gfft::batch::kLanes; ++i) {
    if (b->ev_join[i]) (void)hipEventDestroy(b->ev_join[i]);
    if (b->ws[i]) (void)hipStreamDestroy(b->ws[i]);
  }
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->d_xf) (void)hipFree(b->d_xf);
  if (b->d_prefix) (void)hipFree(b->d_prefix);
  for (int c = 0; c < gfft::batch::kClasses; ++c)
    for (int v = 0; v < gfft::batch::kVariants; ++v)
      if (b->d_tw[c][v]) (void)hipFree(b->d_tw[c][v]);
}

gfft::batch::LevelArgs level_args(const gfft_batch_s* b, int c, int mode, int l) {
  const int v = b->var[c];
  gfft::batch::LevelArgs a{};
  a.xf = b->d_xf + b->off[c];
  a.prefix = b->d_prefix + b->prefix_off[c][v][mode] + (size_t)l * (b->n[c] + 1);
  a.tw = static_cast<const gfft::cx<float>*>(b->d_tw[c][v]);
  a.count = b->n[c];
  a.axis = kModeAxis[mode][l];
  a.inv = kModeInv[mode][l];
  a.total = b->level_tiles[c][v][mode][l];
  return a;
}

uint32_t level_grid(const gfft_batch_s* b, int c, int axis, uint32_t total) {
  const bool plane = axis == gfft::batch::kAxisPlane;
  const int occ = plane ? b->occ_plane[c][b->var[c]] : b->occ[c][b->var[c]];
  const uint32_t per = plane ? gfft::var(c, b->var[c]).plane_tiles : 1u;  // planes one workgroup runs at once
  return std::min<uint32_t>((total + per - 1) / per, (uint32_t)(b->cus * occ));
}

// The plain variant: the class's transforms one after another through their
// single-transform plans, each with its own output scales.
int launch_plain(const gfft_batch_s* b, int c, int mode, hipStream_t s) {
  for (int k = 0; k < b->n[c]; ++k) {
    const BXform& x = b->hx[b->off[c] + k];
    gfft::PlanCore pc = *b->xplan[b->off[c] + k];  // shares the plan's device tables
    pc.r2c_scale = x.scale_r2c;
    pc.c2r_scale = x.scale_c2r;
    int st = GFFT_OK;
    if (mode != 1) st = gfft::exec_r2c(pc, x.data, s);
    if (st == GFFT_OK && mode != 0) st = gfft::exec_c2r(pc, x.data, s);
    if (st != GFFT_OK) return st;
  }
  return GFFT_OK;
}

// The level chain of one class on one stream.
int launch_chain(const gfft_batch_s* b, int c, int mode, hipStream_t s) {
  if (b->var[c] == gfft::batch::kPlainVariant) return launch_plain(b, c, mode, s);
  for (int l = 0; l < kModeLevels[mode]; ++l) {
    const gfft::batch::LevelArgs a = level_args(b, c, mode, l);
    if (a.total == 0) continue;
    const gfft::VarInfo& vi = gfft::var(c, b->var[c]);
    const gfft::LaunchFn fn = a.axis == gfft::batch::kAxisPlane ? vi.plane_launch : vi.launch;
    if (!fn || fn(dim3(level_grid(b, c, a.axis, a.total)), a, s) != hipSuccess) return GFFT_ERR_HIP;
  }
  return GFFT_OK;
}

// Eager path: fork the class chains onto the internal lanes, join back.
int launch_eager(gfft_batch_s* b, int mode, hipStream_t s) {
  using namespace gfft;
  bool used[kLanes] = {};
  for (int c = 0; c < kClasses; ++c)
    if (b->n[c]) used[b->lane[c]] = true;
  if (hipEventRecord(b->ev_fork, s) != hipSuccess) return GFFT_ERR_HIP;
  for (int i = 0; i < kLanes; ++i)
    if (used[i] && hipStreamWaitEvent(b->ws[i], b->ev_fork, 0) != hipSuccess) return GFFT_ERR_HIP;
  for (int c = 0; c < kClasses; ++c) {
    if (b->n[c] == 0) continue;
    const int st = launch_chain(b, c, mode, b->ws[b->lane[c]]);
    if (st != GFFT_OK) return st;
  }
  for (int i = 0; i < kLanes; ++i) {
    if (!used[i]) continue;
    if (hipEventRecord(b->ev_join[i], b->ws[i]) != hipSuccess) return GFFT_ERR_HIP;
    if (hipStreamWaitEvent(s, b->ev_join[i], 0) != hipSuccess) return GFFT_ERR_HIP;
  }
  return GFFT_OK;
}

// ---------------------------------------------------------------------------
// Persistent execution: the plans of the three modes for the current set and
// team factor, for as many slots as the device holds workgroups of the
// set's kernel at once; uploaded once, read by every launch.
// ---------------------------------------------------------------------------
template <typename T>
int upload_vec(T** d, const std::vector<T>& v) {
  if (hipMalloc(d, sizeof(T) * std::max<size_t>(v.size(), 1)) != hipSuccess) {
    (void)hipGetLastError();
    *d = nullptr;
    return GFFT_ERR_OUT_OF_MEMORY;
  }
  if (!v.empty() && hipMemcpy(*d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice) != hipSuccess)
    return GFFT_ERR_HIP;
  return GFFT_OK;
}

int build_plans(gfft_batch_s* b) {
  using namespace gfft;
  using namespace gfft::batch;
  if (hipDeviceSynchronize() != hipSuccess) return GFFT_ERR_HIP;  // earlier launches read the old plans
  free_plans(b);
  int w[kClasses];
  bool present[kClasses] = {};
  for (int c = 0; c < kClasses; ++c) {
    w[c] = var(c, kSetVariant[b->pset][c]).W;
    present[c] = b->n[c] > 0;
  }
  b->plds = persist_lds_bytes(b->pset, present);
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, persist_kernel(b->pset), kBT, b->plds) != hipSuccess ||
      occ < 1) {
    (void)hipGetLastError();  // tolerated: one workgroup per CU
    occ = 1;
  }
  b->pocc = occ;
  const uint32_t domains = (uint32_t)b->xccs;
  const uint32_t per_dom = std::max(1u, (uint32_t)std::max(1, b->cus * occ) / domains);
  size_t need = 0;
  for (int m = 0; m < 3; ++m) {
    PLevelSpec lv[kMaxLevels];
    for (int l = 0; l < kModeLevels[m]; ++l) lv[l] = PLevelSpec{kModeAxis[m][l], kModeInv[m][l]};
    auto tiles = [&](uint32_t t, int l) {
      const BXform& x = b->hx[t];
      const int c = level_cls(x, kModeAxis[m][l], kPersistPlaneClasses);
      return c < 0 ? 0u : tiles_of(x, kModeAxis[m][l], w[c], kPersistPlaneClasses);
    };
    gfft_batch_s::DevPlan& dp = b->plan[m];
    if (b->pteam == 0) {
      // relative cost of one tile: the complex values it moves, a fused
      // plane's weighted by kWidePlaneWeight (its two stages and its load
      // latency cost more per value than a row or column tile: per-workgroup
      // stamps showed the slots heavy in planes ending each level's tail)
      auto cost = [&](uint32_t t, int l) {
        const BXform& x = b->hx[t];
        const int axis = kModeAxis[m][l];
        const int c = level_cls(x, axis, kPersistPlaneClasses);
        if (c < 0) return 0.0;
        if (axis == kAxisPlane) return kWidePlaneWeight * x.ny * x.K;
        if (axis == 0) return 2.0 * w[c] * x.K;
        return (double)w[c] * (axis == 1 ? x.ny : x.nrows / x.ny);
      };
      const WPlan wp = build_wide_plan(
          (uint32_t)b->count, kModeLevels[m], lv, tiles,
          [&](uint32_t t, int l) { return level_cls(b->hx[t], kModeAxis[m][l], kPersistPlaneClasses); }, cost,
          domains * per_dom, kWideSharesPerSlot, kWideSoloFactor);
      if (check_wide_plan(wp) != 0) return GFFT_ERR_INVALID_ARG;  // cannot happen: build_wide_plan's construction
      std::vector<PTeam> xt(b->count);  // each transform's segs, for the solo transforms' whole runs
      for (int t = 0; t < b->count; ++t) xt[t] = PTeam{wp.xseg0[t], wp.xnseg[t], 1u, 1u, 0u, 0u, 0u, 0u};
      int st = upload_vec(&dp.seg, wp.segs);
      if (st == GFFT_OK) st = upload_vec(&dp.team, xt);
      if (st == GFFT_OK) st = upload_vec(&dp.solo, wp.solo);
      if (st == GFFT_OK) st = upload_vec(&dp.wshare, wp.shares);
      if (st == GFFT_OK) st = upload_vec(&dp.lvl_off, wp.lvl_off);
      if (st == GFFT_OK) st = upload_vec(&dp.own_off, wp.own_off);
      if (st == GFFT_OK) st = upload_vec(&dp.own, wp.own);
      if (st != GFFT_OK) {
        free_plans(b);
        return st;
      }
      dp.nslots = wp.segs.empty() ? 0u : domains * per_dom;
      dp.domains = domains;
      dp.per_dom = per_dom;
      dp.nlevels = wp.nlevels;
      dp.nsolo = (uint32_t)wp.solo.size();
      dp.nshares = (uint32_t)wp.shares.size();
      dp.level_base = (uint32_t)wide_level_base(wp);
      dp.members = dp.nslots;
      for (const PSeg& sg : wp.segs) dp.tiles += sg.ntiles;
      dp.ctr_words = wide_counter_words(wp);
      need = std::max(need, dp.ctr_words);
      continue;
    }
    const PPlan pl = build_plan(
        (uint32_t)b->count, kModeLevels[m], lv, tiles,
        [&](uint32_t t, int l) { return level_cls(b->hx[t], kModeAxis[m][l], kPersistPlaneClasses); },
        [&](uint32_t t) {
          const BXform& x = b->hx[t];
          double levels = 0;
          for (int l = 0; l < kModeLevels[m]; ++l) levels += tiles(t, l) ? 1.0 : 0.0;
          return (double)x.nrows * x.K * levels;
        },
        domains, per_dom, (double)b->pteam, 2, kSharesPerMember, (uint32_t)b->pspan);
    if (check_plan(pl) != 0) return GFFT_ERR_INVALID_ARG;  // cannot happen: build_plan's construction
    int st = upload_vec(&dp.seg, pl.segs);
    if (st == GFFT_OK) st = upload_vec(&dp.team, pl.teams);
    if (st == GFFT_OK) st = upload_vec(&dp.parts, pl.parts);
    if (st == GFFT_OK) st = upload_vec(&dp.local, pl.local);
    if (st == GFFT_OK) st = upload_vec(&dp.solo, pl.solo);
    if (st == GFFT_OK) st = upload_vec(&dp.team_ids, pl.team_ids);
    if (st == GFFT_OK) st = upload_vec(&dp.dom_off, pl.dom_off);
    if (st == GFFT_OK) st = upload_vec(&dp.dom_teams, pl.dom_teams);
    if (st != GFFT_OK) {
      free_plans(b);
      return st;
    }
    dp.nslots = pl.segs.empty() ? 0u : pl.domains * pl.per_dom;
    dp.domains = pl.domains;
    dp.per_dom = pl.per_dom;
    dp.nsolo = (uint32_t)pl.solo.size();
    dp.claim_words = pl.claim_words;
    for (const PSeg& sg : pl.segs) dp.tiles += sg.ntiles;
    for (const PTeam& tm : pl.teams)
      if (tm.k >= 2) {
        ++dp.teams_k2;
        dp.members += tm.k;
      }
    dp.ctr_words = counter_words(pl);
    need = std::max(need, dp.ctr_words);
  }
  if (need > b->ctr_words) {  // counters for the largest plan, zeroed
    if (b->d_ctr) (void)hipFree(b->d_ctr);
    b->d_ctr = nullptr;
    b->ctr_words = 0;
    if (hipMalloc(&b->d_ctr, sizeof(uint32_t) * need) != hipSuccess) {
      (void)hipGetLastError();
      free_plans(b);
      return GFFT_ERR_OUT_OF_MEMORY;
    }
    if (hipMemset(b->d_ctr, 0, sizeof(uint32_t) * need) != hipSuccess) return GFFT_ERR_HIP;
    b->ctr_words = need;
  }
  return GFFT_OK;
}

gfft::batch::PersistArgs persist_args(const gfft_batch_s* b, int mode) {
  using namespace gfft::batch;
  PersistArgs a{};
  const gfft_batch_s::DevPlan& dp = b->plan[mode];
  a.xf = b->d_xf;
  a.seg = dp.seg;
  a.team = dp.team;
  a.parts = dp.parts;
  a.local = dp.local;
  a.solo = dp.solo;
  a.team_ids = dp.team_ids;
  a.dom_off = dp.dom_off;
  a.dom_teams = dp.dom_teams;
  a.domains = dp.domains;
  a.per_dom = dp.per_dom;
  a.nsolo = dp.nsolo;
  a.nteams = dp.teams_k2;
  a.claim_words = dp.claim_words;
  a.ctr = b->d_ctr;
  a.status = b->d_ctr + gfft::kCtrStatus;
  a.host_abort = b->h_abort;
  const double ticks = b->wait_limit_s * (double)b->wall_khz * 1e3;
  a.wait_ticks = ticks >= 1.8e19 ? ~0ull : (uint64_t)ticks;
  // team plans: a member steals as soon as its own shares are done; global-
  // level plans: only after 50 us without completion (a list is unclaimed
  // only while its slot's workgroup has not arrived)
  a.steal_ticks = b->pteam == 0 ? (uint64_t)(50e-6 * (double)b->wall_khz * 1e3) : 0;
  for (int c = 0; c < kClasses; ++c)
    a.tw[c] = static_cast<const gfft::cx<float>*>(b->d_tw[c][kSetVariant[b->pset][c]]);
  a.wide = b->pteam == 0 ? 1u : 0u;
  a.nlevels = dp.nlevels;
  a.nshares = dp.nshares;
  a.level_base = dp.level_base;
  a.slots = dp.domains * dp.per_dom;
  a.wshare = dp.wshare;
  a.lvl_off = dp.lvl_off;
  a.own_off = dp.own_off;
  a.own = dp.own;
  return a;
}

uint32_t persist_grid(const gfft_batch_s* b, int mode) { return std::max<uint32_t>(1u, b->plan[mode].nslots); }

int launch_persistent(gfft_batch_s* b, int mode, hipStream_t s) {
  if (b->plan[mode].nslots == 0) return GFFT_OK;
  const gfft::batch::PersistArgs a = persist_args(b, mode);
  return gfft::batch::launch_persist(b->pset, persist_grid(b, mode), b->plds, a, s) == hipSuccess ? GFFT_OK
                                                                                                   : GFFT_ERR_HIP;
}

int build_persist_graph(gfft_batch_s* b, int mode, hipGraph_t* out) {
  hipGraph_t g = nullptr;
  if (hipGraphCreate(&g, 0) != hipSuccess) return GFFT_ERR_HIP;
  if (b->plan[mode].nslots > 0) {
    gfft::batch::PersistArgs a = persist_args(b, mode);
    void* params[] = {&a};
    hipKernelNodeParams p{};
    p.func = const_cast<void*>(gfft::batch::persist_kernel(b->pset));
    p.gridDim = dim3(persist_grid(b, mode));
    p.blockDim = dim3(gfft::batch::kBT);
    p.sharedMemBytes = (unsigned)b->plds;
    p.kernelParams = params;
    p.extra = nullptr;
    hipGraphNode_t node = nullptr;
    if (hipGraphAddKernelNode(&node, g, nullptr, 0, &p) != hipSuccess) {
      (void)hipGraphDestroy(g);
      return GFFT_ERR_HIP;
    }
  }
  *out = g;
  return GFFT_OK;
}

// Graph path: one kernel node per (class, level); each class is a chain.  A
// class on the plain variant is one child-graph node captured from its plan
// launches (its data pointers and scales are then part of the graph, so
// set_data / set_scales drop the graphs).
int build_graph(gfft_batch_s* b, int mode, hipGraph_t* out) {
  using namespace gfft;
  hipGraph_t g = nullptr;
  if (hipGraphCreate(&g, 0) != hipSuccess) return GFFT_ERR_HIP;
  for (int c = 0; c < kClasses; ++c) {
    if (b->n[c] == 0) continue;
    if (b->var[c] == kPlainVariant) {
      hipStream_t cs = b->ws[b->lane[c]];
      hipGraph_t child = nullptr;
      if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGraphDestroy(g);
        return GFFT_ERR_HIP;
      }
      const int st = launch_plain(b, c, mode, cs);
      const hipError_t ce = hipStreamEndCapture(cs, &child);
      hipGraphNode_t node = nullptr;
      if (st != GFFT_OK || ce != hipSuccess || !child ||
          hipGraphAddChildGraphNode(&node, g, nullptr, 0, child) != hipSuccess) {
        if (child) (void)hipGraphDestroy(child);
        (void)hipGraphDestroy(g);
        return st != GFFT_OK ? st : GFFT_ERR_HIP;
      }
      (void)hipGraphDestroy(child);
      continue;
    }
    hipGraphNode_t prev = nullptr;
    for (int l = 0; l < kModeLevels[mode]; ++l) {
      LevelArgs a = level_args(b, c, mode, l);
      if (a.total == 0) continue;
      void* params[] = {&a};
      hipKernelNodeParams p{};
      const VarInfo& vi = var(c, b->var[c]);
      p.func = const_cast<void*>(a.axis == kAxisPlane ? vi.plane_kernel : vi.kernel);
      if (!p.func) {
        (void)hipGraphDestroy(g);
        return GFFT_ERR_HIP;
      }
      p.gridDim = dim3(level_grid(b, c, a.axis, a.total));
      p.blockDim = dim3(kBT);
      p.sharedMemBytes = 0;
      p.kernelParams = params;
      p.extra = nullptr;
      hipGraphNode_t node = nullptr;
      if (hipGraphAddKernelNode(&node, g, prev ? &prev : nullptr, prev ? 1 : 0, &p) != hipSuccess) {
        (void)hipGraphDestroy(g);
        return GFFT_ERR_HIP;
      }
      prev = node;
    }
  }
  *out = g;
  return GFFT_OK;
}

int check_dev(const gfft_batch_s* b) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return GFFT_ERR_HIP;
  return dev == b->device ? GFFT_OK : GFFT_ERR_DEVICE_MISMATCH;
}

// Synchronous: earlier launches may still read the descriptors.
int upload_xf(gfft_batch_s* b) {
  if (hipDeviceSynchronize() != hipSuccess) return GFFT_ERR_HIP;
  return hipMemcpy(b->d_xf, b->hx.data(), sizeof(BXform) * b->count, hipMemcpyHostToDevice) == hipSuccess
             ? GFFT_OK
             : GFFT_ERR_HIP;
}

// Times `reps` eager executions of `mode` after one warm-up; ms per execution.
int time_exec(gfft_batch_s* b, int mode, int reps, hipStream_t s, hipEvent_t e0, hipEvent_t e1, float* ms) {
  auto launch = [&] {
    return b->sched == GFFT_BATCH_PERSISTENT ? launch_persistent(b, mode, s) : launch_eager(b, mode, s);
  };
  int st = launch();
  if (st != GFFT_OK) return st;
  if (hipEventRecord(e0, s) != hipSuccess) return GFFT_ERR_HIP;
  for (int r = 0; r < reps && st == GFFT_OK; ++r) st = launch();
  if (st != GFFT_OK) return st;
  float t = 0;
  if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
      hipEventElapsedTime(&t, e0, e1) != hipSuccess)
    return GFFT_ERR_HIP;
  *ms = t / reps;
  return GFFT_OK;
}

// Persistent scheduler: every variant set with each (team factor, span)
// pair, team factor 0 being the global-level plan.  Two passes over the
// candidates, each candidate's time the faster of its two (the passes
// interleave, so a slow moment of the device does not decide alone); a
// candidate must win by 1 %; a set whose kernel does not launch is skipped.
int tune_persistent(gfft_batch_s* b, int mode, int reps, hipStream_t s, hipEvent_t e0, hipEvent_t e1, float* best) {
  static const int kCandidates[][2] = {{0, 2}, {1, 2}, {2, 2}, {4, 2}, {1, 1}, {2, 1}};
  constexpr int kCand = (int)(sizeof(kCandidates) / sizeof(kCandidates[0]));
  float ms_of[gfft::batch::kSets][kCand];
  bool runs[gfft::batch::kSets];
  for (int set = 0; set < gfft::batch::kSets; ++set) {
    runs[set] = true;
    for (int i = 0; i < kCand; ++i) ms_of[set][i] = 1e30f;
  }
  int st = GFFT_OK;
  b->sched = GFFT_BATCH_PERSISTENT;
  for (int pass = 0; pass < 2 && st == GFFT_OK; ++pass)
    for (int set = 0; set < gfft::batch::kSets && st == GFFT_OK; ++set)
      for (int i = 0; i < kCand && runs[set]; ++i) {
        b->pset = set;
        b->pteam = kCandidates[i][0];
        b->pspan = kCandidates[i][1];
        if ((st = build_plans(b)) != GFFT_OK) break;
        float ms = 0;
        const int rt = time_exec(b, mode, reps, s, e0, e1, &ms);
        if (rt == GFFT_ERR_HIP) {  // this set's kernel cannot run here: skip it
          (void)hipGetLastError();
          if (hipDeviceSynchronize() != hipSuccess) return GFFT_ERR_HIP;
          runs[set] = false;
          break;
        }
        if ((st = rt) != GFFT_OK) break;
        ms_of[set][i] = std::min(ms_of[set][i], ms);
      }
  if (st != GFFT_OK) return st;
  int best_set = -1, best_i = 0;
  float best_ms = 1e30f;
  for (int set = 0; set < gfft::batch::kSets; ++set)
    for (int i = 0; i < kCand; ++i)
      if (runs[set] && ms_of[set][i] < 0.99f * best_ms) {
        best_ms = ms_of[set][i];
        best_set = set;
        best_i = i;
      }
  if (best_set < 0) return GFFT_ERR_HIP;
  b->pset = best_set;
  b->pteam = kCandidates[best_i][0];
  b->pspan = kCandidates[best_i][1];
  *best = best_ms;
  return build_plans(b);
}

// Level scheduler: every class's variants and the per-transform-plan choice,
// each class's level chain timed alone.
int tune_levels(gfft_batch_s* b, int mode, int reps, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  int st = GFFT_OK;
  for (int c = 0; c < gfft::batch::kClasses && st == GFFT_OK; ++c) {
    if (b->n[c] == 0) continue;
    int best = 0;
    float best_ms = 1e30f;
    for (int v = 0; v <= gfft::batch::kPlainVariant && st == GFFT_OK; ++v) {
      b->var[c] = v;
      if ((st = launch_chain(b, c, mode, s)) != GFFT_OK) break;  // warm-up
      if (hipEventRecord(e0, s) != hipSuccess) st = GFFT_ERR_HIP;
      for (int r = 0; r < reps && st == GFFT_OK; ++r) st = launch_chain(b, c, mode, s);
      float ms = 0;
      if (st == GFFT_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        st = GFFT_ERR_HIP;
      if (st == GFFT_OK && ms < best_ms) {
        best_ms = ms;
        best = v;
      }
    }
    b->var[c] = best;
  }
  return st;
}

}  // namespace

extern "C" {

int gfft_batch_query(int64_t nx, int64_t ny, int64_t nz, int prec) {
  if (prec != GFFT_F32) return GFFT_ERR_UNSUPPORTED_SIZE;
  if (nx < 1 || ny < 1 || nz < 1) return GFFT_ERR_INVALID_DIMS;
  const int c = gfft::batch::class_of(nz);
  if (c < 0) return GFFT_ERR_UNSUPPORTED_SIZE;
  if (ny > 1 && gfft::batch::class_of(ny) != c) return GFFT_ERR_UNSUPPORTED_SIZE;
  if (nx > 1 && gfft::batch::class_of(nx) != c) return GFFT_ERR_UNSUPPORTED_SIZE;
  return GFFT_OK;
}

int gfft_batch_create(gfft_batch* out, int count, const int64_t* dims, int prec) {
  using namespace gfft;
  using namespace gfft::batch;
  if (!out || count < 1 || !dims) return GFFT_ERR_INVALID_ARG;
  *out = nullptr;
  for (int i = 0; i < count; ++i) {
    const int st = gfft_batch_query(dims[3 * i], dims[3 * i + 1], dims[3 * i + 2], prec);
    if (st != GFFT_OK) return st;
  }
  gfft_batch_s* b = new (std::nothrow) gfft_batch_s;
  if (!b) return GFFT_ERR_OUT_OF_MEMORY;
  b->count = count;
  b->dims.assign(dims, dims + 3 * count);
  if (hipGetDevice(&b->device) != hipSuccess) {
    delete b;
    return GFFT_ERR_HIP;
  }
  // sort by class (stable: caller order within a class)
  std::vector<int> order(count);
  for (int i = 0; i < count; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int c2) { return class_of(dims[3 * a + 2]) < class_of(dims[3 * c2 + 2]); });
  b->pos.assign(count, 0);
  b->hx.resize(count);
  b->bytes.resize(count);
  double work[kClasses] = {};
  for (int j = 0; j < count; ++j) {
    const int i = order[j];
    b->pos[i] = j;
    const int64_t nx = dims[3 * i], ny = dims[3 * i + 1], nz = dims[3 * i + 2];
    BXform& x = b->hx[j];
    x.data = nullptr;
    x.K = (int32_t)(nz / 2 + 1);
    x.ny = (int32_t)ny;
    x.nrows = nx * ny;
    x.ncols_y = (uint32_t)(nx * x.K);
    x.ncols_x = (uint32_t)(ny * x.K);
    x.cls[0] = class_of(nz);
    x.cls[1] = ny > 1 ? class_of(ny) : -1;
    x.cls[2] = nx > 1 ? class_of(nx) : -1;
    x.scale_r2c = 1.0f;
    x.scale_c2r = 1.0f;
    x.plane = ny > 1 && x.cls[0] < kPlaneClasses ? 1 : 0;
    b->bytes[i] = (size_t)nx * ny * x.K * 8;
    b->n[x.cls[0]]++;
    work[x.cls[0]] += (double)b->bytes[i];
  }
  for (int c = 1; c < kClasses; ++c) b->off[c] = b->off[c - 1] + b->n[c - 1];
  // tile layout per class, variant, mode and level
  std::vector<uint32_t> prefix;
  for (int c = 0; c < kClasses; ++c)
    for (int v = 0; v < kVariants; ++v)
      for (int m = 0; m < 3; ++m) {
        b->prefix_off[c][v][m] = prefix.size();
        const int w = var(c, v).W;
        for (int l = 0; l < kModeLevels[m]; ++l) {
          const size_t base = prefix.size();
          prefix.resize(base + b->n[c] + 1, 0);
          for (int k = 0; k < b->n[c]; ++k)
            prefix[base + k + 1] = prefix[base + k] + tiles_of(b->hx[b->off[c] + k], kModeAxis[m][l], w);
          b->level_tiles[c][v][m][l] = prefix[base + b->n[c]];
        }
      }
  // class chains onto lanes: heaviest first, each to the least loaded lane
  {
    int cl[kClasses];
    for (int c = 0; c < kClasses; ++c) cl[c] = c;
    std::stable_sort(cl, cl + kClasses, [&](int a, int c2) { return work[a] > work[c2]; });
    double load[kLanes] = {};
    for (int k = 0; k < kClasses; ++k) {
      const int c = cl[k];
      if (b->n[c] == 0) continue;
      const int ln = (int)(std::min_element(load, load + kLanes) - load);
      b->lane[c] = ln;
      load[ln] += work[c];
    }
  }
  b->cus = device_cu_count();
  b->xccs = device_xcc_count();
  int st = b->cus > 0 ? GFFT_OK : GFFT_ERR_HIP;
  auto alloc = [&](void** p, size_t n) {
    if (st == GFFT_OK && hipMalloc(p, n) != hipSuccess) {
      (void)hipGetLastError();
      st = GFFT_ERR_OUT_OF_MEMORY;
    }
  };
  alloc((void**)&b->d_xf, sizeof(BXform) * count);
  alloc((void**)&b->d_prefix, sizeof(uint32_t) * std::max<size_t>(prefix.size(), 1));
  for (int c = 0; c < kClasses && st == GFFT_OK; ++c) {
    if (b->n[c] == 0) continue;
    for (int v = 0; v < kVariants && st == GFFT_OK; ++v) {
      const std::vector<unsigned char> tw = twiddle_table(var(c, v).sched, F32);
      alloc(&b->d_tw[c][v], tw.size());
      if (st == GFFT_OK && hipMemcpy(b->d_tw[c][v], tw.data(), tw.size(), hipMemcpyHostToDevice) != hipSuccess)
        st = GFFT_ERR_HIP;
      int occ = 0;
      if (st == GFFT_OK &&
          (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, var(c, v).kernel, kBT, 0) != hipSuccess || occ < 1)) {
        (void)hipGetLastError();  // tolerated: one workgroup per CU
        occ = 1;
      }
      b->occ[c][v] = occ;
      occ = 0;
      if (st == GFFT_OK && var(c, v).plane_kernel &&
          (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, var(c, v).plane_kernel, kBT, 0) != hipSuccess ||
           occ < 1)) {
        (void)hipGetLastError();
        occ = 1;
      }
      b->occ_plane[c][v] = occ;
    }
  }
  for (int i = 0; i < kLanes && st == GFFT_OK; ++i) {
    if (hipStreamCreateWithFlags(&b->ws[i], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_join[i], hipEventDisableTiming) != hipSuccess)
      st = GFFT_ERR_HIP;
  }
  if (st == GFFT_OK && hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess) st = GFFT_ERR_HIP;
  // single-transform plans of the plain variant, one per distinct geometry
  b->xplan.assign(count, nullptr);
  for (int j = 0; j < count && st == GFFT_OK; ++j) {
    const int i = order[j];
    const auto key = std::make_tuple(dims[3 * i], dims[3 * i + 1], dims[3 * i + 2]);
    auto it = b->plans.find(key);
    if (it == b->plans.end()) {
      std::unique_ptr<PlanCore> pc(new (std::nothrow) PlanCore);
      if (!pc) {
        st = GFFT_ERR_OUT_OF_MEMORY;
        break;
      }
      st = plan_host(*pc, dims[3 * i], dims[3 * i + 1], dims[3 * i + 2], 1, 0, F32);
      if (st == GFFT_OK) st = upload_twiddles(*pc);
      if (st != GFFT_OK) {
        free_twiddles(*pc);
        break;
      }
      (void)apply_tune_cache(*pc);
      it = b->plans.emplace(key, std::move(pc)).first;
    }
    b->xplan[j] = it->second.get();
  }
  if (st == GFFT_OK &&
      hipMemcpy(b->d_prefix, prefix.data(), sizeof(uint32_t) * prefix.size(), hipMemcpyHostToDevice) != hipSuccess)
    st = GFFT_ERR_HIP;
  if (st == GFFT_OK) st = upload_xf(b);
  // persistent execution: the host abort word, plans and counters
  if (st == GFFT_OK) {
    if (hipHostMalloc((void**)&b->h_abort, sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
      b->h_abort = nullptr;
      st = GFFT_ERR_OUT_OF_MEMORY;
    }
    if (st == GFFT_OK) *b->h_abort = 0;
    if (st == GFFT_OK &&
        (hipDeviceGetAttribute(&b->wall_khz, hipDeviceAttributeWallClockRate, b->device) != hipSuccess ||
         b->wall_khz <= 0))
      st = GFFT_ERR_HIP;
    if (st == GFFT_OK) st = build_plans(b);
  }
  if (st != GFFT_OK) {
    release(b);
    delete b;
    return st;
  }
  *out = b;
  return GFFT_OK;
}

int gfft_batch_set_data(gfft_batch b, void* const* ptrs) {
  if (!b || !ptrs) return GFFT_ERR_INVALID_ARG;
  for (int i = 0; i < b->count; ++i)  // whole float2 elements: 8 B aligned
    if (!ptrs[i] || (reinterpret_cast<uintptr_t>(ptrs[i]) & 7u) != 0) return GFFT_ERR_INVALID_ARG;
  for (int i = 0; i < b->count; ++i) b->hx[b->pos[i]].data = static_cast<float*>(ptrs[i]);
  const int st = upload_xf(b);  // synchronises the device first
  if (b->sched == GFFT_BATCH_LEVELS && any_plain(b)) drop_graphs(b);
  if (st == GFFT_OK) b->data_set = true;
  return st;
}

int gfft_batch_set_scales(gfft_batch b, const double* r2c_scale, const double* c2r_scale) {
  if (!b) return GFFT_ERR_INVALID_ARG;
  for (int i = 0; i < b->count; ++i) {
    b->hx[b->pos[i]].scale_r2c = r2c_scale ? (float)r2c_scale[i] : 1.0f;
    b->hx[b->pos[i]].scale_c2r = c2r_scale ? (float)c2r_scale[i] : 1.0f;
  }
  const int st = upload_xf(b);
  if (b->sched == GFFT_BATCH_LEVELS && any_plain(b)) drop_graphs(b);
  return st;
}

int gfft_batch_exec(gfft_batch b, int mode, gfft_stream_t stream) {
  if (!b || mode < 0 || mode > 2 || !b->data_set) return GFFT_ERR_INVALID_ARG;
  const int st = check_dev(b);
  if (st != GFFT_OK) return st;
  if (b->sched == GFFT_BATCH_PERSISTENT) return launch_persistent(b, mode, reinterpret_cast<hipStream_t>(stream));
  return launch_eager(b, mode, reinterpret_cast<hipStream_t>(stream));
}

int gfft_batch_exec_noalloc(gfft_batch b, int mode, gfft_stream_t stream) {
  if (!b || mode < 0 || mode > 2 || !b->data_set) return GFFT_ERR_INVALID_ARG;
  int st = check_dev(b);
  if (st != GFFT_OK) return st;
  const gfft_batch_s::DevPlan& dp = b->plan[mode];
  if (dp.nslots == 0) return GFFT_OK;
  if (dp.ctr_words > gfft::batch::kStaticCounterWords) return GFFT_ERR_SCRATCH_TOO_SMALL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  uint32_t* ctr = nullptr;
  if (gfft::batch::static_counters(&ctr, s) != hipSuccess) return GFFT_ERR_HIP;
  gfft::batch::PersistArgs a = persist_args(b, mode);
  a.ctr = ctr;  // the launch's counters; failures still reach the batch's status word
  return gfft::batch::launch_persist(b->pset, persist_grid(b, mode), b->plds, a, s) == hipSuccess ? GFFT_OK
                                                                                                   : GFFT_ERR_HIP;
}

int gfft_batch_exec_graph(gfft_batch b, int mode, gfft_stream_t stream) {
  if (!b || mode < 0 || mode > 2 || !b->data_set) return GFFT_ERR_INVALID_ARG;
  int st = check_dev(b);
  if (st != GFFT_OK) return st;
  if (!b->gexec[mode]) {
    hipGraph_t g = nullptr;
    st = b->sched == GFFT_BATCH_PERSISTENT ? build_persist_graph(b, mode, &g) : build_graph(b, mode, &g);
    if (st != GFFT_OK) return st;
    hipGraphExec_t ge = nullptr;
    if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(g);
      return GFFT_ERR_HIP;
    }
    b->graph[mode] = g;
    b->gexec[mode] = ge;
  }
  return hipGraphLaunch(b->gexec[mode], reinterpret_cast<hipStream_t>(stream)) == hipSuccess ? GFFT_OK
                                                                                            : GFFT_ERR_HIP;
}

int gfft_batch_set_variant(gfft_batch b, int64_t edge, int variant) {
  if (!b || variant < 0 || variant > gfft::batch::kPlainVariant) return GFFT_ERR_INVALID_ARG;
  const int c = gfft::batch::class_of(edge);
  if (c < 0) return GFFT_ERR_INVALID_ARG;
  if (hipDeviceSynchronize() != hipSuccess) return GFFT_ERR_HIP;
  b->var[c] = variant;
  drop_graphs(b);
  return GFFT_OK;
}

int gfft_batch_variant(gfft_batch b, int64_t edge, int* variant) {
  if (!b || !variant) return GFFT_ERR_INVALID_ARG;
  const int c = gfft::batch::class_of(edge);
  if (c < 0) return GFFT_ERR_INVALID_ARG;
  *variant = b->var[c];
  return GFFT_OK;
}

int gfft_batch_tune(gfft_batch b, int mode, int reps, gfft_stream_t stream) {
  if (!b || mode < 0 || mode > 2 || !b->data_set) return GFFT_ERR_INVALID_ARG;
  int st = check_dev(b);
  if (st != GFFT_OK) return st;
  if (reps < 1) reps = 1;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess) return GFFT_ERR_HIP;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return GFFT_ERR_HIP;
  }
  drop_graphs(b);
  // both schedulers tuned, then the faster kept (eager timings of `mode`)
  float ms_p = 1e30f, ms_l = 1e30f;
  st = tune_persistent(b, mode, reps, s, e0, e1, &ms_p);
  if (st == GFFT_OK) {
    b->sched = GFFT_BATCH_LEVELS;
    st = tune_levels(b, mode, reps, s, e0, e1);
  }
  if (st == GFFT_OK) st = time_exec(b, mode, reps, s, e0, e1, &ms_l);
  if (st == GFFT_OK) b->sched = ms_p <= ms_l ? GFFT_BATCH_PERSISTENT : GFFT_BATCH_LEVELS;
  b->tuned_ms[0] = ms_p;
  b->tuned_ms[1] = ms_l;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  drop_graphs(b);
  return st;
}

int gfft_batch_data_bytes(gfft_batch b, int index, size_t* bytes) {
  if (!b || !bytes || index < 0 || index >= b->count) return GFFT_ERR_INVALID_ARG;
  *bytes = b->bytes[index];
  return GFFT_OK;
}

int gfft_batch_tiles(gfft_batch b, int mode, uint32_t* tiles, uint32_t* launches) {
  if (!b || !tiles || mode < 0 || mode > 2) return GFFT_ERR_INVALID_ARG;
  if (b->sched == GFFT_BATCH_PERSISTENT) {
    *tiles = b->plan[mode].tiles;
    if (launches) *launches = b->plan[mode].nslots ? 1u : 0u;
    return GFFT_OK;
  }
  uint32_t t = 0, n = 0;
  for (int c = 0; c < gfft::batch::kClasses; ++c)
    for (int l = 0; l < kModeLevels[mode] && b->var[c] != gfft::batch::kPlainVariant; ++l) {
      const uint32_t k = b->level_tiles[c][b->var[c]][mode][l];
      t += k;
      n += k ? 1u : 0u;
    }
  *tiles = t;
  if (launches) *launches = n;
  return GFFT_OK;
}

int gfft_batch_set_scheduler(gfft_batch b, int scheduler) {
  if (!b || (scheduler != GFFT_BATCH_PERSISTENT && scheduler != GFFT_BATCH_LEVELS)) return GFFT_ERR_INVALID_ARG;
  if (scheduler != b->sched) {
    drop_graphs(b);
    b->sched = scheduler;
  }
  return GFFT_OK;
}

int gfft_batch_scheduler(gfft_batch b, int* scheduler) {
  if (!b || !scheduler) return GFFT_ERR_INVALID_ARG;
  *scheduler = b->sched;
  return GFFT_OK;
}

int gfft_batch_set_persist(gfft_batch b, int set, int team_factor) {
  if (!b || set < 0 || set >= gfft::batch::kSets || team_factor < 0 || team_factor > 64) return GFFT_ERR_INVALID_ARG;
  drop_graphs(b);
  b->pset = set;
  b->pteam = team_factor;
  return build_plans(b);
}

int gfft_batch_set_span(gfft_batch b, int parts) {
  if (!b || parts < 1 || parts > (int)gfft::kMaxDomains) return GFFT_ERR_INVALID_ARG;
  drop_graphs(b);
  b->pspan = parts;
  return build_plans(b);
}

int gfft_batch_span(gfft_batch b, int* parts) {
  if (!b || !parts) return GFFT_ERR_INVALID_ARG;
  *parts = b->pspan;
  return GFFT_OK;
}

int gfft_batch_persist(gfft_batch b, int* set, int* team_factor, int* variants) {
  if (!b || !set || !team_factor) return GFFT_ERR_INVALID_ARG;
  *set = b->pset;
  *team_factor = b->pteam;
  if (variants)
    for (int c = 0; c < gfft::batch::kClasses; ++c) variants[c] = gfft::batch::kSetVariant[b->pset][c];
  return GFFT_OK;
}

int gfft_batch_tuned_ms(gfft_batch b, float* persistent_ms, float* levels_ms) {
  if (!b || !persistent_ms || !levels_ms) return GFFT_ERR_INVALID_ARG;
  *persistent_ms = b->tuned_ms[0];
  *levels_ms = b->tuned_ms[1];
  return GFFT_OK;
}

int gfft_batch_plan_info(gfft_batch b, int mode, uint32_t* slots, uint32_t* teams, uint32_t* members) {
  if (!b || mode < 0 || mode > 2 || !slots || !teams || !members) return GFFT_ERR_INVALID_ARG;
  *slots = b->plan[mode].nslots;
  *teams = b->plan[mode].teams_k2;
  *members = b->plan[mode].members;
  return GFFT_OK;
}

int gfft_batch_set_wait_limit(gfft_batch b, double seconds) {
  if (!b || !(seconds >= 0.0)) return GFFT_ERR_INVALID_ARG;
  drop_graphs(b);  // the bound is a kernel argument of the captured launch
  b->wait_limit_s = seconds;
  return GFFT_OK;
}

int gfft_batch_set_abort(gfft_batch b, int abort) {
  if (!b) return GFFT_ERR_INVALID_ARG;
  __atomic_store_n(b->h_abort, abort ? 1u : 0u, __ATOMIC_SEQ_CST);
  return GFFT_OK;
}

int gfft_batch_status(gfft_batch b, int* status) {
  if (!b || !status) return GFFT_ERR_INVALID_ARG;
  if (hipDeviceSynchronize() != hipSuccess) return GFFT_ERR_HIP;
  uint32_t w = 0;
  if (hipMemcpy(&w, b->d_ctr + gfft::kCtrStatus, sizeof w, hipMemcpyDeviceToHost) != hipSuccess) return GFFT_ERR_HIP;
  const uint32_t zero = 0;
  if (w && hipMemcpy(b->d_ctr + gfft::kCtrStatus, &zero, sizeof zero, hipMemcpyHostToDevice) != hipSuccess)
    return GFFT_ERR_HIP;
  *status = (w & gfft::kFailAborted) ? GFFT_ERR_ABORTED : (w & gfft::kFailTimeout) ? GFFT_ERR_TIMEOUT : GFFT_OK;
  return GFFT_OK;
}

int gfft_batch_destroy(gfft_batch b) {
  if (!b) return GFFT_ERR_INVALID_ARG;
  release(b);
  delete b;
  return GFFT_OK;
}

}  // extern "C"
