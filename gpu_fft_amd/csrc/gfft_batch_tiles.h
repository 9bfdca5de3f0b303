// Tile code shared by the batched kernels (gfft_batch.hip: one launch per
// size class and level; gfft_batch_persist.hip: one persistent launch for
// the whole batch).  Transforms are grouped by size class (all three edges
// of a transform belong to one class); each class has kVariants compiled
// variants (radix order; points per thread, which sets the pencils per
// 256-thread tile) and every tile runs the same register-engine tile code as
// the single-transform kernels.  For the classes up to 128 the z and y
// passes of a transform run fused, one x-plane per tile held in LDS (the
// plane kernels' code), so those transforms make two HBM round trips per
// direction instead of three.
#pragma once
#include "gfft_batch_sched.h"
#include "gfft_kernels.h"
#include "gfft_plane.h"

namespace gfft {
namespace batch {

constexpr int kBT = 256;  // threads per workgroup, every tile kind
constexpr int kClasses = 8;
constexpr int kVariants = 3;
constexpr int kPlainVariant = kVariants;  // per-transform plans
constexpr int kMaxLevels = 8;
constexpr int kAxisPlane = 3;    // level axis code of the fused z + y plane pass
constexpr int kPlaneClasses = 7;         // level scheduler: classes 0 .. 6 (32 .. 192) fuse z and y
constexpr int kPersistPlaneClasses = 5;  // persistent kernel: 0 .. 4 (a 160 or 192 plane needs
                                         // 105-150 KB of LDS, one workgroup per CU)
constexpr int kLanes = 4;  // internal streams of the eager path

// Variant V of size class C: radix order and points per thread P; with
// TPT = N/P threads per transform a tile carries W = 256/TPT row pairs (z)
// or columns (y, x).
template <int C, int V>
struct BVar;
#define GFFT_BVAR(C, V, P_, ...)     \
  template <>                        \
  struct BVar<C, V> {                \
    using SC = Sched<__VA_ARGS__>;   \
    static constexpr int P = P_;     \
  };
GFFT_BVAR(0, 0, 8, 8, 4)
GFFT_BVAR(0, 1, 8, 4, 8)
GFFT_BVAR(0, 2, 4, 8, 4)
GFFT_BVAR(1, 0, 12, 4, 4, 3)
GFFT_BVAR(1, 1, 12, 3, 4, 4)
GFFT_BVAR(1, 2, 6, 4, 4, 3)
GFFT_BVAR(2, 0, 8, 8, 8)
GFFT_BVAR(2, 1, 4, 4, 4, 4)
GFFT_BVAR(2, 2, 16, 8, 8)
GFFT_BVAR(3, 0, 24, 8, 4, 3)
GFFT_BVAR(3, 1, 12, 8, 4, 3)
GFFT_BVAR(3, 2, 24, 3, 8, 4)
GFFT_BVAR(4, 0, 8, 8, 4, 4)
GFFT_BVAR(4, 1, 8, 4, 4, 8)
GFFT_BVAR(4, 2, 16, 8, 4, 4)
GFFT_BVAR(5, 0, 20, 4, 4, 2, 5)
GFFT_BVAR(5, 1, 10, 8, 4, 5)
GFFT_BVAR(5, 2, 20, 8, 4, 5)
GFFT_BVAR(6, 0, 24, 8, 8, 3)
GFFT_BVAR(6, 1, 12, 8, 8, 3)
GFFT_BVAR(6, 2, 24, 3, 8, 8)
GFFT_BVAR(7, 0, 8, 8, 8, 4)
GFFT_BVAR(7, 1, 16, 8, 8, 4)
GFFT_BVAR(7, 2, 8, 4, 8, 8)
#undef GFFT_BVAR

constexpr int kClassN[kClasses] = {32, 48, 64, 96, 128, 160, 192, 256};

template <int C, int V>
struct BCfg {
  static constexpr int CLS = C, VAR = V;
  using SC = typename BVar<C, V>::SC;
  static constexpr int N = SC::N;
  static_assert(N == kClassN[C], "variant size");
  static constexpr int P = BVar<C, V>::P;
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

// Shape of the fused plane pass of a class (256 threads): z and y points per
// thread (the z exchange of a row pair stays inside one wavefront) and the
// kz columns per y step.
template <int C>
struct PlaneShape;
template <>
struct PlaneShape<0> {
  static constexpr int PZ = 4, PY = 4, CW = 32;
};
template <>
struct PlaneShape<1> {
  static constexpr int PZ = 6, PY = 3, CW = 16;
};
template <>
struct PlaneShape<2> {
  static constexpr int PZ = 8, PY = 4, CW = 16;
};
template <>
struct PlaneShape<3> {
  static constexpr int PZ = 12, PY = 6, CW = 16;
};
template <>
struct PlaneShape<4> {
  static constexpr int PZ = 8, PY = 8, CW = 16;
};
template <>
struct PlaneShape<5> {
  static constexpr int PZ = 10, PY = 10, CW = 16;
};
template <>
struct PlaneShape<6> {
  static constexpr int PZ = 12, PY = 12, CW = 16;
};

// The plane pass of class C with variant V's radix schedule (and twiddles).
template <int C, int V>
using BPlane = PlanePass<float, typename BVar<C, V>::SC, PlaneShape<C>::PZ, typename BVar<C, V>::SC,
                         PlaneShape<C>::PY, PlaneShape<C>::CW>;
// The same with LDS-only barriers in its y stage (pipelined plane loops).
template <int C, int V>
using BPlaneL = PlanePass<float, typename BVar<C, V>::SC, PlaneShape<C>::PZ, typename BVar<C, V>::SC,
                          PlaneShape<C>::PY, PlaneShape<C>::CW, 1>;
// One wavefront per plane (small planes): kz columns in groups of 64 / TPTY.
template <int C, int V>
using BPlaneW = PlanePass<float, typename BVar<C, V>::SC, PlaneShape<C>::PZ, typename BVar<C, V>::SC,
                          PlaneShape<C>::PY, 64 / (BVar<C, V>::SC::N / PlaneShape<C>::PY), 2>;

inline int class_of(int64_t n) {
  for (int c = 0; c < kClasses; ++c)
    if (kClassN[c] == n) return c;
  return -1;
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
  int32_t plane;     // 1: the z and y passes run fused per x-plane (level axis kAxisPlane)
};

// Size class of a level of transform x (-1: the level is empty for x);
// z and y run fused when x.plane is set and the class is below max_plane.
GFFT_HD inline int level_cls(const BXform& x, int axis, int max_plane = kPlaneClasses) {
  const bool fused = x.plane && x.cls[0] < max_plane;
  if (axis == kAxisPlane) return fused ? x.cls[0] : -1;
  if (fused && axis < 2) return -1;
  return x.cls[axis];
}

// One level launch: one axis pass over every transform of one class.
struct LevelArgs {
  const BXform* xf;        // the class's transforms
  const uint32_t* prefix;  // [count + 1] tile prefix sums of this level
  const cx<float>* tw;     // twiddle table of the class's variant
  int32_t count;
  int32_t axis;            // 0 = z rows, 1 = y columns, 2 = x columns, 3 = z + y per x-plane
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

template <typename F>
__device__ __forceinline__ void run_tile(const BXform& x, int axis, bool inv, uint32_t tile,
                                         const cx<float>* tw, cx<float>* smem) {
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

// x-plane `plane` of transform x through the fused plane pass PP (R2C: z
// then y; C2R: y then z, writing the Nz reals of every row).  smem holds
// PP::PLANE + PP::SCRATCH entries; ends with a barrier, so the next call may
// reuse them.
template <typename PP>
__device__ __forceinline__ void plane_tile(const BXform& x, bool inv, uint32_t plane, const cx<float>* tw,
                                           cx<float>* smem) {
  cx<float>* g = reinterpret_cast<cx<float>*>(x.data) + (size_t)plane * PP::PLANE;
  PP::copy_in(smem, g);
  __syncthreads();
  if (!inv) {
    PP::z_r2c(smem, tw);
    __syncthreads();
    // the forward scale belongs to this pass when x is length 1
    const float sc = x.cls[2] < 0 ? x.scale_r2c : 1.0f;
    PP::template y_pass<false>(smem, smem + PP::PLANE, tw, sc, sc != 1.0f);
    __syncthreads();
    for (int i = threadIdx.x; i < PP::PLANE; i += PP::NT) st_stream(g + i, smem[i]);
  } else {
    PP::template y_pass<true>(smem, smem + PP::PLANE, tw, 1.0f, false);
    __syncthreads();
    PP::z_c2r(smem, tw, x.scale_c2r, x.scale_c2r != 1.0f);
    __syncthreads();
    constexpr int H = PP::NZ / 2;
    for (int i = threadIdx.x; i < PP::NY * H; i += PP::NT) {
      const int y = i / H, m = i - y * H;
      st_stream(g + y * PP::K + m, smem[y * PP::K + m]);
    }
  }
  __syncthreads();
}

// Planes through a one-wavefront plane pass PP (PP::NT == 64, wavefront-level
// barriers): the calling wavefront runs the planes next(0), next(1), ...
// (next(i, &x, &plane) returns false past its last) in its own LDS region L
// of PP::PLANE + PP::SCRATCH entries, loading each next plane into registers
// before the current one's stages, so its loads are in flight meanwhile.
// R2C runs z then y and writes the half spectra back, C2R y then z and the
// Nz reals of every row; the scales are each plane's transform's (the
// forward one belongs to this pass when x is length 1).  Planes must be
// disjoint (loading the next before storing the current is then safe in
// place).
template <typename PP, typename Next>
__device__ __forceinline__ void wave_planes(Next next, bool inv, const cx<float>* tw, cx<float>* L) {
  static_assert(PP::NT == 64, "one plane per wavefront");
  constexpr int U = (PP::PLANE + 63) / 64;
  constexpr bool kFull = U * 64 == PP::PLANE;
  const int lane = (int)(threadIdx.x & 63u);
  cx<float> nxt[U];
  const BXform* nx = nullptr;
  uint32_t np = 0;
  bool more = next(0u, &nx, &np);
  if (more) {
    const cx<float>* g = reinterpret_cast<const cx<float>*>(nx->data) + (size_t)np * PP::PLANE;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = u * 64 + lane;
      if (kFull || i < PP::PLANE) nxt[u] = ld_stream(g + i);
    }
  }
  for (uint32_t k = 0; more; ++k) {
    const BXform* x = nx;
    cx<float>* g = reinterpret_cast<cx<float>*>(x->data) + (size_t)np * PP::PLANE;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = u * 64 + lane;
      if (kFull || i < PP::PLANE) L[i] = nxt[u];
    }
    wave_sync();
    more = next(k + 1, &nx, &np);
    if (more) {
      const cx<float>* gn = reinterpret_cast<const cx<float>*>(nx->data) + (size_t)np * PP::PLANE;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = u * 64 + lane;
        if (kFull || i < PP::PLANE) nxt[u] = ld_stream(gn + i);
      }
    }
    if (!inv) {
      PP::z_r2c(L, tw);
      wave_sync();
      const float sc = x->cls[2] < 0 ? x->scale_r2c : 1.0f;
      PP::template y_pass<false>(L, L + PP::PLANE, tw, sc, sc != 1.0f);
      wave_sync();
      for (int i = lane; i < PP::PLANE; i += 64) st_stream(g + i, L[i]);
    } else {
      PP::template y_pass<true>(L, L + PP::PLANE, tw, 1.0f, false);
      wave_sync();
      PP::z_c2r(L, tw, x->scale_c2r, x->scale_c2r != 1.0f);
      wave_sync();
      constexpr int H = PP::NZ / 2;
      for (int i = lane; i < PP::NY * H; i += 64) {
        const int y = i / H, m = i - y * H;
        st_stream(g + y * PP::K + m, L[y * PP::K + m]);
      }
    }
    wave_sync();  // the stores' LDS reads are done before the next plane lands
  }
}

// ---------------------------------------------------------------------------
// Persistent execution (gfft_batch_persist.hip): one launch runs the plan of
// a mode (gfft_batch_sched.h).  A variant set fixes one compiled variant per
// class; the set's kernel calls the tile code of those eight variants (its
// register count is the largest of them).
// ---------------------------------------------------------------------------
constexpr int kSets = 2;
// set 0: per class the variant with the fewest registers (kernel: 180
// VGPRs); set 1: variant 1 of every class (228).  Both run 2 waves per SIMD.
// All variants 0 would need 280 registers (with AGPRs), which does not
// launch.  (Set 0 held to 128 registers, 4 waves per SIMD, compiles with a
// few spills and ran the team levels 2-4x slower.)
constexpr int kSetVariant[kSets][kClasses] = {{2, 2, 1, 1, 0, 1, 1, 0}, {1, 1, 1, 1, 1, 1, 1, 1}};

struct PersistArgs {
  const BXform* xf;            // descriptors, batch order (= team order)
  const PSeg* seg;             // levels of every transform
  const PTeam* team;           // one per transform
  const PPart* parts;          // the teams' parts (a span team has several, on distinct domains)
  const PEntry* local;         // domains * per_dom local slots, domain-major
  const uint32_t* solo;        // solo queue (transforms run whole by one workgroup)
  const uint32_t* team_ids;    // transforms run by teams
  const uint32_t* dom_off;     // domains + 1: the teams of domain d are dom_teams[dom_off[d] ..
  const uint32_t* dom_teams;   //   dom_off[d + 1]), longest first (the helpers' lists)
  uint32_t domains, per_dom;   // L2 domains (XCDs) and local slots per domain
  uint32_t nsolo, nteams;
  uint32_t claim_words;
  uint32_t* ctr;               // counter words (kCtrShares + claim_words)
  uint32_t* status;            // sticky status word the failure bits are ORed into
  const uint32_t* host_abort;  // pinned host word: nonzero = every wait gives up
  uint64_t wait_ticks;         // bound of one wait, device wall-clock ticks
  uint64_t steal_ticks;        // a waiting member steals unclaimed shares after this long
  const cx<float>* tw[kClasses];  // twiddles of each class's variant in the set
  // global-level plans (gfft_batch_sched.h WPlan; team, local, solo and
  // team_ids unused): wide = 1
  uint32_t wide;
  uint32_t nlevels, nshares, slots;  // slots = domains * per_dom
  uint32_t level_base;               // counter word of level 0's hand-over block (kLevelWords per level)
  const WShare* wshare;              // level-major shares
  const uint32_t* lvl_off;           // nlevels + 1
  const uint32_t* own_off;           // nlevels * slots + 1
  const uint32_t* own;               // share indices per (level, slot)
};

// Host side of the persistent kernels.
const void* persist_kernel(int set);
// dynamic LDS bytes of set `set` for the classes marked present
size_t persist_lds_bytes(int set, const bool* present);
hipError_t launch_persist(int set, uint32_t grid, size_t lds, const PersistArgs& a, hipStream_t s);

// Counter words of the allocation-free entry point: a file-scope device
// array in gfft_batch_persist.hip, one per GPU, held until the process ends.
constexpr size_t kStaticCounterWords = size_t(1) << 16;  // 256 KiB
// Device address of the calling device's array (hipGetSymbolAddress once per
// device, cached); zeroes it with a fill kernel on `s` at the first call on a
// device.  Returns a HIP status.
hipError_t static_counters(uint32_t** out, hipStream_t s);

}  // namespace batch
}  // namespace gfft
