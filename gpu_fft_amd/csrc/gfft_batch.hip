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
// Each tile runs the same register-engine tile code as the single-transform
// kernels, configured so that every size class uses exactly 256 threads.
#include <algorithm>
#include <new>
#include <vector>

#include "gfft_kernels.h"
#include "gfft_plan.h"
#include "gpufft/gpufft.h"

namespace gfft {
namespace {

constexpr int kBT = 256;  // threads per workgroup, every tile kind
constexpr int kClasses = 8;
constexpr int kMaxLevels = 6;
constexpr int kLanes = 4;  // internal streams of the eager path

// Size classes of the batched kernels: schedule and points per thread; rows
// tiles carry 256 / (N/P) pairs, column tiles 256 / (N/P) columns.
template <int C>
struct BCls;
template <>
struct BCls<0> {
  using SC = Sched<8, 4>;
  static constexpr int P = 8;
};
template <>
struct BCls<1> {
  using SC = Sched<4, 4, 3>;
  static constexpr int P = 12;
};
template <>
struct BCls<2> {
  using SC = Sched<8, 8>;
  static constexpr int P = 8;
};
template <>
struct BCls<3> {
  using SC = Sched<8, 4, 3>;
  static constexpr int P = 24;
};
template <>
struct BCls<4> {
  using SC = Sched<8, 4, 4>;
  static constexpr int P = 8;
};
template <>
struct BCls<5> {
  using SC = Sched<4, 4, 2, 5>;
  static constexpr int P = 20;
};
template <>
struct BCls<6> {
  using SC = Sched<8, 8, 3>;
  static constexpr int P = 24;
};
template <>
struct BCls<7> {
  using SC = Sched<8, 8, 4>;
  static constexpr int P = 8;
};

template <int C>
struct BCfg {
  using SC = typename BCls<C>::SC;
  static constexpr int N = SC::N;
  static constexpr int P = BCls<C>::P;
  static constexpr int TPT = N / P;
  static constexpr int W = kBT / TPT;  // pairs per rows tile = columns per cols tile
  static_assert(W * TPT == kBT, "tile must use all 256 threads");
  using R2C = RowsR2C<float, SC, P, W>;
  using C2R = RowsC2R<float, SC, P, W>;
  using FWD = ColsPass<float, false, SC, P, W>;
  using INV = ColsPass<float, true, SC, P, W>;
  static constexpr int lds_entries() {
    const int rows = W * R2C::LROW;
    const int cols = N * W;
    return rows > cols ? rows : cols;
  }
};

constexpr int kClassN[kClasses] = {32, 48, 64, 96, 128, 160, 192, 256};

int class_of(int64_t n) {
  for (int c = 0; c < kClasses; ++c)
    if (kClassN[c] == n) return c;
  return -1;
}

Schedule class_sched(int c) {
  switch (c) {
    case 0: return make_sched<BCls<0>::SC>();
    case 1: return make_sched<BCls<1>::SC>();
    case 2: return make_sched<BCls<2>::SC>();
    case 3: return make_sched<BCls<3>::SC>();
    case 4: return make_sched<BCls<4>::SC>();
    case 5: return make_sched<BCls<5>::SC>();
    case 6: return make_sched<BCls<6>::SC>();
    default: return make_sched<BCls<7>::SC>();
  }
}

int class_width(int c) {
  const int n = kClassN[c];
  const int P[kClasses] = {8, 12, 8, 24, 8, 20, 24, 8};
  return kBT / (n / P[c]);
}

// One transform of the batch.
struct BXform {
  float* data;       // in-place array (Nx, Ny, 2K)
  int32_t cls[3];    // size class of the z, y, x pass (-1: length-1 strided axis)
  int32_t K;
  int32_t ny;
  int64_t nrows;     // Nx * Ny
  uint32_t ncols_y;  // Nx * K
  uint32_t ncols_x;  // Ny * K
  float scale_r2c, scale_c2r;
};

// One level launch: one axis pass over every transform of one class.
struct LevelArgs {
  const BXform* xf;        // the class's transforms
  const uint32_t* prefix;  // [count + 1] tile prefix sums of this level
  const cx<float>* tw;     // the class's twiddle table
  int32_t count;
  int32_t axis;            // 0 = z rows, 1 = y columns, 2 = x columns
  int32_t inv;
  uint32_t total;          // tiles of this level = prefix[count]
};

__device__ __forceinline__ ColMap dmap(uint32_t inner, int64_t outer, int64_t rlo) {
  ColMap m;
  m.inner = inner;
  m.nsub = 0x7fffffff;
  m.outer = outer;
  m.rlo = rlo;
  m.rhi = 0;
  return m;
}

template <int C>
__device__ __forceinline__ void run_tile(const BXform& x, int axis, bool inv, uint32_t tile,
                                         const cx<float>* tw, cx<float>* smem) {
  using F = BCfg<C>;
  if (axis == 0) {
    RowGeom g;
    g.nrows = x.nrows;
    g.K = x.K;
    g.scale = inv ? x.scale_c2r : 1.0;
    if (inv)
      F::C2R::tile(x.data, g, tw, (int64_t)tile * F::W, smem);
    else
      F::R2C::tile(x.data, g, tw, (int64_t)tile * F::W, smem);
    return;
  }
  ColGeom g;
  if (axis == 1) {
    g.ncols = x.ncols_y;
    g.in = g.out = dmap((uint32_t)x.K, (int64_t)x.ny * x.K, x.K);
  } else {
    g.ncols = x.ncols_x;
    g.in = g.out = dmap(x.ncols_x, 0, (int64_t)x.ncols_x);
  }
  // the forward scale belongs to the last R2C pass (x, or y/z when x is length 1)
  const bool last_fwd = !inv && (axis == 2 || (axis == 1 && x.cls[2] < 0));
  g.scale = last_fwd ? x.scale_r2c : 1.0;
  cx<float>* d = reinterpret_cast<cx<float>*>(x.data);
  if (inv)
    F::INV::tile(d, d, g, tw, tile, smem);
  else
    F::FWD::tile(d, d, g, tw, tile, smem);
}

// Workgroups stride over the level's tiles; tile k belongs to the transform
// t with prefix[t] <= k < prefix[t + 1].
template <int C>
__global__ __launch_bounds__(kBT) void k_batch_level(LevelArgs a) {
  __shared__ cx<float> smem[BCfg<C>::lds_entries()];
  for (uint32_t k = blockIdx.x; k < a.total; k += gridDim.x) {
    int lo = 0, hi = a.count;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (a.prefix[mid] <= k)
        lo = mid;
      else
        hi = mid;
    }
    run_tile<C>(a.xf[lo], a.axis, a.inv != 0, k - a.prefix[lo], a.tw, smem);
    __syncthreads();  // the tile's LDS reads are done before the next tile's writes
  }
}

const void* level_kernel(int c) {
  switch (c) {
    case 0: return reinterpret_cast<const void*>(&k_batch_level<0>);
    case 1: return reinterpret_cast<const void*>(&k_batch_level<1>);
    case 2: return reinterpret_cast<const void*>(&k_batch_level<2>);
    case 3: return reinterpret_cast<const void*>(&k_batch_level<3>);
    case 4: return reinterpret_cast<const void*>(&k_batch_level<4>);
    case 5: return reinterpret_cast<const void*>(&k_batch_level<5>);
    case 6: return reinterpret_cast<const void*>(&k_batch_level<6>);
    default: return reinterpret_cast<const void*>(&k_batch_level<7>);
  }
}

hipError_t launch_level(int c, dim3 grid, const LevelArgs& a, hipStream_t s) {
  switch (c) {
    case 0: hipLaunchKernelGGL(k_batch_level<0>, grid, dim3(kBT), 0, s, a); break;
    case 1: hipLaunchKernelGGL(k_batch_level<1>, grid, dim3(kBT), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_batch_level<2>, grid, dim3(kBT), 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_batch_level<3>, grid, dim3(kBT), 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_batch_level<4>, grid, dim3(kBT), 0, s, a); break;
    case 5: hipLaunchKernelGGL(k_batch_level<5>, grid, dim3(kBT), 0, s, a); break;
    case 6: hipLaunchKernelGGL(k_batch_level<6>, grid, dim3(kBT), 0, s, a); break;
    default: hipLaunchKernelGGL(k_batch_level<7>, grid, dim3(kBT), 0, s, a); break;
  }
  return hipGetLastError();
}

}  // namespace
}  // namespace gfft

using gfft::BXform;

// Transforms are grouped by size class; group c occupies positions
// [off[c], off[c] + n[c]) of the device descriptor array.
struct gfft_batch_s {
  int count = 0;
  int device = 0;
  std::vector<int64_t> dims;   // 3 per transform, caller order
  std::vector<size_t> bytes;   // data bytes per transform, caller order
  std::vector<int> pos;        // caller index -> sorted position
  std::vector<BXform> hx;      // sorted by class
  int off[gfft::kClasses] = {}, n[gfft::kClasses] = {};
  BXform* d_xf = nullptr;
  uint32_t* d_prefix = nullptr;  // [class][mode][level][n_c + 1], packed
  size_t prefix_off[gfft::kClasses][3] = {};
  uint32_t level_tiles[gfft::kClasses][3][gfft::kMaxLevels] = {};
  void* d_tw[gfft::kClasses] = {};
  int occ[gfft::kClasses] = {};   // resident workgroups per CU of the class's kernel
  int lane[gfft::kClasses] = {};  // internal stream of the class's chain (eager path)
  int cus = 0;
  bool data_set = false;
  hipStream_t ws[gfft::kLanes] = {};
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_join[gfft::kLanes] = {};
  hipGraph_t graph[3] = {};
  hipGraphExec_t gexec[3] = {};
};

namespace {

const int kModeAxis[3][gfft::kMaxLevels] = {{0, 1, 2}, {2, 1, 0}, {0, 1, 2, 2, 1, 0}};
const int kModeInv[3][gfft::kMaxLevels] = {{0, 0, 0}, {1, 1, 1}, {0, 0, 0, 1, 1, 1}};
const int kModeLevels[3] = {3, 3, 6};

uint32_t tiles_of(const BXform& x, int axis) {
  const int c = x.cls[axis];
  if (c < 0) return 0;
  const int w = gfft::class_width(c);
  if (axis == 0) return (uint32_t)(((x.nrows + 1) / 2 + w - 1) / w);
  const uint32_t nc = axis == 1 ? x.ncols_y : x.ncols_x;
  return (nc + w - 1) / w;
}

void release(gfft_batch_s* b) {
  for (int m = 0; m < 3; ++m) {
    if (b->gexec[m]) (void)hipGraphExecDestroy(b->gexec[m]);
    if (b->graph[m]) (void)hipGraphDestroy(b->graph[m]);
  }
  for (int i = 0; i < gfft::kLanes; ++i) {
    if (b->ev_join[i]) (void)hipEventDestroy(b->ev_join[i]);
    if (b->ws[i]) (void)hipStreamDestroy(b->ws[i]);
  }
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->d_xf) (void)hipFree(b->d_xf);
  if (b->d_prefix) (void)hipFree(b->d_prefix);
  for (int c = 0; c < gfft::kClasses; ++c)
    if (b->d_tw[c]) (void)hipFree(b->d_tw[c]);
}

gfft::LevelArgs level_args(const gfft_batch_s* b, int c, int mode, int l) {
  gfft::LevelArgs a{};
  a.xf = b->d_xf + b->off[c];
  a.prefix = b->d_prefix + b->prefix_off[c][mode] + (size_t)l * (b->n[c] + 1);
  a.tw = static_cast<const gfft::cx<float>*>(b->d_tw[c]);
  a.count = b->n[c];
  a.axis = kModeAxis[mode][l];
  a.inv = kModeInv[mode][l];
  a.total = b->level_tiles[c][mode][l];
  return a;
}

uint32_t level_grid(const gfft_batch_s* b, int c, uint32_t total) {
  return std::min<uint32_t>(total, (uint32_t)(b->cus * b->occ[c]));
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
    for (int l = 0; l < kModeLevels[mode]; ++l) {
      const LevelArgs a = level_args(b, c, mode, l);
      if (a.total == 0) continue;
      if (launch_level(c, dim3(level_grid(b, c, a.total)), a, b->ws[b->lane[c]]) != hipSuccess)
        return GFFT_ERR_HIP;
    }
  }
  for (int i = 0; i < kLanes; ++i) {
    if (!used[i]) continue;
    if (hipEventRecord(b->ev_join[i], b->ws[i]) != hipSuccess) return GFFT_ERR_HIP;
    if (hipStreamWaitEvent(s, b->ev_join[i], 0) != hipSuccess) return GFFT_ERR_HIP;
  }
  return GFFT_OK;
}

// Graph path: one kernel node per (class, level); each class is a chain.
int build_graph(gfft_batch_s* b, int mode, hipGraph_t* out) {
  using namespace gfft;
  hipGraph_t g = nullptr;
  if (hipGraphCreate(&g, 0) != hipSuccess) return GFFT_ERR_HIP;
  for (int c = 0; c < kClasses; ++c) {
    if (b->n[c] == 0) continue;
    hipGraphNode_t prev = nullptr;
    for (int l = 0; l < kModeLevels[mode]; ++l) {
      LevelArgs a = level_args(b, c, mode, l);
      if (a.total == 0) continue;
      void* params[] = {&a};
      hipKernelNodeParams p{};
      p.func = const_cast<void*>(level_kernel(c));
      p.gridDim = dim3(level_grid(b, c, a.total));
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

}  // namespace

extern "C" {

int gfft_batch_query(int64_t nx, int64_t ny, int64_t nz, int prec) {
  if (prec != GFFT_F32) return GFFT_ERR_UNSUPPORTED_SIZE;
  if (nx < 1 || ny < 1 || nz < 1) return GFFT_ERR_INVALID_DIMS;
  const int c = gfft::class_of(nz);
  if (c < 0) return GFFT_ERR_UNSUPPORTED_SIZE;
  if (ny > 1 && gfft::class_of(ny) != c) return GFFT_ERR_UNSUPPORTED_SIZE;
  if (nx > 1 && gfft::class_of(nx) != c) return GFFT_ERR_UNSUPPORTED_SIZE;
  return GFFT_OK;
}

int gfft_batch_create(gfft_batch* out, int count, const int64_t* dims, int prec) {
  using namespace gfft;
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
    b->bytes[i] = (size_t)nx * ny * x.K * 8;
    b->n[x.cls[0]]++;
    work[x.cls[0]] += (double)b->bytes[i];
  }
  for (int c = 1; c < kClasses; ++c) b->off[c] = b->off[c - 1] + b->n[c - 1];
  // tile layout per class, mode and level
  std::vector<uint32_t> prefix;
  for (int c = 0; c < kClasses; ++c)
    for (int m = 0; m < 3; ++m) {
      b->prefix_off[c][m] = prefix.size();
      for (int l = 0; l < kModeLevels[m]; ++l) {
        const size_t base = prefix.size();
        prefix.resize(base + b->n[c] + 1, 0);
        for (int k = 0; k < b->n[c]; ++k)
          prefix[base + k + 1] = prefix[base + k] + tiles_of(b->hx[b->off[c] + k], kModeAxis[m][l]);
        b->level_tiles[c][m][l] = prefix[base + b->n[c]];
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
    const std::vector<unsigned char> tw = twiddle_table(class_sched(c), F32);
    alloc(&b->d_tw[c], tw.size());
    if (st == GFFT_OK && hipMemcpy(b->d_tw[c], tw.data(), tw.size(), hipMemcpyHostToDevice) != hipSuccess)
      st = GFFT_ERR_HIP;
    int occ = 0;
    if (st == GFFT_OK &&
        (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, level_kernel(c), kBT, 0) != hipSuccess || occ < 1))
      occ = 1;
    b->occ[c] = occ;
  }
  for (int i = 0; i < kLanes && st == GFFT_OK; ++i) {
    if (hipStreamCreateWithFlags(&b->ws[i], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_join[i], hipEventDisableTiming) != hipSuccess)
      st = GFFT_ERR_HIP;
  }
  if (st == GFFT_OK && hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess) st = GFFT_ERR_HIP;
  if (st == GFFT_OK &&
      hipMemcpy(b->d_prefix, prefix.data(), sizeof(uint32_t) * prefix.size(), hipMemcpyHostToDevice) != hipSuccess)
    st = GFFT_ERR_HIP;
  if (st == GFFT_OK) st = upload_xf(b);
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
  for (int i = 0; i < b->count; ++i)
    if (!ptrs[i]) return GFFT_ERR_INVALID_ARG;
  for (int i = 0; i < b->count; ++i) b->hx[b->pos[i]].data = static_cast<float*>(ptrs[i]);
  const int st = upload_xf(b);
  if (st == GFFT_OK) b->data_set = true;
  return st;
}

int gfft_batch_set_scales(gfft_batch b, const double* r2c_scale, const double* c2r_scale) {
  if (!b) return GFFT_ERR_INVALID_ARG;
  for (int i = 0; i < b->count; ++i) {
    b->hx[b->pos[i]].scale_r2c = r2c_scale ? (float)r2c_scale[i] : 1.0f;
    b->hx[b->pos[i]].scale_c2r = c2r_scale ? (float)c2r_scale[i] : 1.0f;
  }
  return upload_xf(b);
}

int gfft_batch_exec(gfft_batch b, int mode, gfft_stream_t stream) {
  if (!b || mode < 0 || mode > 2 || !b->data_set) return GFFT_ERR_INVALID_ARG;
  const int st = check_dev(b);
  if (st != GFFT_OK) return st;
  return launch_eager(b, mode, reinterpret_cast<hipStream_t>(stream));
}

int gfft_batch_exec_graph(gfft_batch b, int mode, gfft_stream_t stream) {
  if (!b || mode < 0 || mode > 2 || !b->data_set) return GFFT_ERR_INVALID_ARG;
  int st = check_dev(b);
  if (st != GFFT_OK) return st;
  if (!b->gexec[mode]) {
    hipGraph_t g = nullptr;
    st = build_graph(b, mode, &g);
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

int gfft_batch_data_bytes(gfft_batch b, int index, size_t* bytes) {
  if (!b || !bytes || index < 0 || index >= b->count) return GFFT_ERR_INVALID_ARG;
  *bytes = b->bytes[index];
  return GFFT_OK;
}

int gfft_batch_tiles(gfft_batch b, int mode, uint32_t* tiles, uint32_t* launches) {
  if (!b || !tiles || mode < 0 || mode > 2) return GFFT_ERR_INVALID_ARG;
  uint32_t t = 0, n = 0;
  for (int c = 0; c < gfft::kClasses; ++c)
    for (int l = 0; l < kModeLevels[mode]; ++l) {
      t += b->level_tiles[c][mode][l];
      n += b->level_tiles[c][mode][l] ? 1u : 0u;
    }
  *tiles = t;
  if (launches) *launches = n;
  return GFFT_OK;
}

int gfft_batch_destroy(gfft_batch b) {
  if (!b) return GFFT_ERR_INVALID_ARG;
  release(b);
  delete b;
  return GFFT_OK;
}

}  // extern "C"
