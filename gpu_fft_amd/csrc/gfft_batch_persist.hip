// Persistent batched kernel: every axis pass of every transform of a batch
// mode in one launch, eager or captured in a graph, following the plan of
// gfft_batch_sched.h.
//
// A workgroup reads its XCC id (its L2 domain) and takes the next local slot
// of that domain (arrival counter).  If the slot holds a team member, it runs
// the member's part of the team's levels; then it takes solo transforms from
// the shared queue (one ticket each) and runs each whole, level after level.
// A solo transform needs no synchronisation beyond the workgroup barrier
// between levels (one workgroup's stores are visible to its own later loads).
//
// Every workgroup that touches a team's array runs in the team's domain (or,
// for a span team, in one of its parts' domains), so within a part they all
// go through one L2, and a level's hand-over needs L2-level ordering only
// (the ordering the memory model uses for a workgroup whose waves are split
// over CUs of one XCD):
//   run a share: claim it (compare-and-swap on the share's word), run its
//             tiles, every thread waits for its stores to complete
//             (s_waitcnt vmcnt(0): they are in the L2), workgroup barrier,
//             one atomic store of "done" to the share's word;
//   wait:     the first wavefront reads the level's share words (atomic
//             loads, 64 per instruction) until all are done; barrier; every
//             wavefront runs an agent-scope acquire fence (buffer_inv sc1:
//             drops the CU's L1 lines, keeps the L2's dirty lines) before
//             it loads the next level.
// A span team's level also waits, once every share is done, until each part
// that ran shares (recorded in the done words: kShareDone + part) has been
// released: one workgroup of the part (the first to claim the part's release
// word) writes its XCD's L2 back (agent-scope release fence, then the flag),
// and the acquire after the wait drops stale lines of the other parts' L2s.
// Otherwise no L2 write-back happens inside the launch.  (A workgroup-scope
// L1 invalidate, buffer_inv sc0, is not enough outside threadgroup-split
// mode: with it, lines a member had read in an earlier level were now and
// then served stale from its L1.)
//
// While waiting, after steal_ticks, a member claims and runs any share still
// free (its owner has not started), so a team never depends on all of its
// members being resident.  Every wait is bounded: it gives up when the
// launch's wall-clock bound passes or when the host sets the batch's abort
// word (pinned host memory, system-scope loads), and marks the launch
// failed; a failed launch's workgroups stop computing and leave.  A
// workgroup that finds the solo queue empty helps the teams of its own
// domain (same L2): it claims free shares of each team's current level, runs
// them and marks them done as a member would, after an agent-scope acquire
// (the level before is complete in the L2), and leaves when those teams are
// complete.  The last workgroup to leave runs the teams of which no member
// arrived (whole, alone), moves the failure bits into the sticky status word
// and zeroes the other counter words, so the next launch (or graph replay)
// starts from zero without a memset.
//
// With a global-level plan (team factor 0, WPlan in gfft_batch_sched.h) every
// workgroup takes shares of every transform's level g in level g of the
// launch: it claims its slot's list of the level (one compare-and-swap), runs
// it, and publishes: its stores complete (in its XCD's L2), then it adds the
// count it ran to its domain's done word.  Each domain's first workgroup (its
// leader) sums the done words; at the level's share count it writes its XCD's
// L2 back once (agent-scope release) if the domain did work, flags that, waits
// for every working domain's flag and raises its domain's go flag, which the
// domain's other workgroups watch.  Every wavefront then acquires at agent
// scope (L1 and stale L2 lines dropped) before the next level.  A workgroup
// that has waited steal_ticks claims and runs any list still unclaimed.
// Waits are bounded and abortable as for the team plans.
#include <map>
#include <mutex>

#include "gfft_batch_tiles.h"

namespace gfft {
namespace batch {

// Counter words of gfft_batch_exec_noalloc (see static_counters): every
// call on one GPU shares this array, which keeps the last call's contents
// (all zero after a launch completes: its last workgroup resets them).
static __device__ __align__(256) uint32_t scratch[kStaticCounterWords];

namespace {

__global__ void k_fill_zero(uint32_t* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

template <int S, int C>
using SCfg = BCfg<C, kSetVariant[S][C]>;

// Tiles first, first + step, ... (count of them) of one level of one
// transform, software-pipelined: while one tile runs its LDS stages, the
// global loads of the next are in flight in a second register set (stage
// barriers wait for LDS traffic only; the twiddles are copied to LDS, after
// the tile area, so no global load in the stages waits behind the prefetch).
// Tiles of a level are disjoint, so loading the next before storing the
// current is safe in place.  The thread's lane indexes are made opaque at
// the top of every tile, so the per-element offsets are recomputed per tile
// instead of being held in registers across the loop (which doubled the
// register count).
template <typename F>
struct Seq {
  using SC = typename F::SC;
  static constexpr int P = F::P, W = F::W, N = SC::N;
  using R2C = RowsR2C<float, SC, P, W, true>;
  using C2R = RowsC2R<float, SC, P, W, true>;
  template <bool INV>
  using COLS = ColsPass<float, INV, SC, P, W, true>;
  static constexpr int kTw = twiddle_count<SC>();
  static constexpr bool kPlane = F::CLS < kPersistPlaneClasses;  // z + y fused levels
  // plane loops (in functions of their own, class_planes, so that they do
  // not add to the rows' and columns' registers): one wavefront per plane
  // for 32 and 48 (a 256-thread plane of those leaves most threads idle in
  // its stages), the workgroup on one plane, pipelined, for 64; 96 and 128
  // unpipelined (the next plane's 19 or 33 entries per thread spill into
  // accumulator registers, which takes the kernel to one wave per SIMD)
  static constexpr bool kPlaneWave = kPlane && F::CLS <= 1;
  static constexpr bool kPlanePipe = kPlane && F::CLS == 2;
  static constexpr int kWaves = kBT / 64;
  static constexpr int base_entries() {
    if constexpr (kPlaneWave) {
      using PP = BPlaneW<F::CLS, F::VAR>;
      constexpr int e = kWaves * (PP::PLANE + PP::SCRATCH);
      return F::lds_entries() > e ? F::lds_entries() : e;
    } else if constexpr (kPlane) {
      using PP = BPlane<F::CLS, F::VAR>;
      return F::lds_entries() > PP::PLANE + PP::SCRATCH ? F::lds_entries() : PP::PLANE + PP::SCRATCH;
    } else {
      return F::lds_entries();
    }
  }
  static constexpr int lds_entries() { return base_entries() + kTw; }

  __device__ static void rows_r2c(const BXform& x, uint32_t first, uint32_t step, uint32_t count,
                                  const cx<float>* stw, cx<float>* smem) {
    RowGeom g;
    g.nrows = x.nrows;
    g.K = x.K;
    g.scale = 1.0;
    const int gi0 = threadIdx.x / R2C::TPT, t0 = threadIdx.x % R2C::TPT;
    cx<float> cur[R2C::E::PV], nxt[R2C::E::PV];
    R2C::load(nxt, x.data, g, (int64_t)first * W + gi0, t0);
    for (uint32_t k = 0; k < count; ++k) {
      int gi = gi0, t = t0;
      __asm__ volatile("" : "+v"(gi), "+v"(t));
      RowLds<float, N> L{smem + gi * R2C::LROW};
#pragma unroll
      for (int i = 0; i < R2C::E::PV; ++i) cur[i] = nxt[i];
      if (k + 1 < count) R2C::load(nxt, x.data, g, (int64_t)(first + (k + 1) * step) * W + gi, t);
      R2C::E::run(cur, t, L, stw);
      R2C::finish(cur, x.data, g, (int64_t)(first + k * step) * W + gi, t, L);
    }
  }

  __device__ static void rows_c2r(const BXform& x, uint32_t first, uint32_t step, uint32_t count,
                                  const cx<float>* stw, cx<float>* smem) {
    RowGeom g;
    g.nrows = x.nrows;
    g.K = x.K;
    g.scale = x.scale_c2r;
    const int gi0 = threadIdx.x / C2R::TPT, t0 = threadIdx.x % C2R::TPT;
    cx<float> ca[C2R::NI], cb[C2R::NI], na[C2R::NI], nb[C2R::NI], v[C2R::E::PV];
    C2R::load(na, nb, x.data, g, (int64_t)first * W + gi0, t0);
    for (uint32_t k = 0; k < count; ++k) {
      int gi = gi0, t = t0;
      __asm__ volatile("" : "+v"(gi), "+v"(t));
      cx<float>* S = smem + gi * C2R::LROW;
      RowLds<float, N> L{S};
#pragma unroll
      for (int i = 0; i < C2R::NI; ++i) {
        ca[i] = na[i];
        cb[i] = nb[i];
      }
      if (k + 1 < count) C2R::load(na, nb, x.data, g, (int64_t)(first + (k + 1) * step) * W + gi, t);
      stage_barrier<C2R::BM>();  // the previous tile's stage reads of S are done
      C2R::build(v, ca, cb, S, t);
      C2R::E::run(v, t, L, stw);
      C2R::store(v, x.data, g, (int64_t)(first + k * step) * W + gi, t);
    }
  }

  // x-planes first, first + step, ... (count of them) through the fused plane
  // pass, software-pipelined like the rows and columns: the next plane's
  // loads go out into registers before the current plane's LDS stages, whose
  // barriers wait for LDS traffic only, and land in LDS once the current
  // plane's stores have read it.  Planes of a level are disjoint, so loading
  // the next before storing the current is safe in place.
  __device__ static void planes(const BXform& x, bool inv, uint32_t first, uint32_t step, uint32_t count,
                                const cx<float>* stw, cx<float>* smem) {
    using PP = BPlaneL<F::CLS, F::VAR>;
    static_assert(PP::NT == kBT, "one plane per 256-thread workgroup");
    constexpr int U = (PP::PLANE + kBT - 1) / kBT;
    constexpr bool kFull = U * kBT == PP::PLANE;
    cx<float>* base = reinterpret_cast<cx<float>*>(x.data);
    cx<float> nxt[U];
    {
      const cx<float>* g = base + (size_t)first * PP::PLANE;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = u * kBT + (int)threadIdx.x;
        if (kFull || i < PP::PLANE) nxt[u] = ld_stream(g + i);
      }
    }
    for (uint32_t k = 0; k < count; ++k) {
      cx<float>* g = base + (size_t)(first + k * step) * PP::PLANE;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = u * kBT + (int)threadIdx.x;
        if (kFull || i < PP::PLANE) smem[i] = nxt[u];
      }
      lds_barrier();
      if (k + 1 < count) {
        const cx<float>* gn = base + (size_t)(first + (k + 1) * step) * PP::PLANE;
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int i = u * kBT + (int)threadIdx.x;
          if (kFull || i < PP::PLANE) nxt[u] = ld_stream(gn + i);
        }
      }
      if (!inv) {
        PP::z_r2c(smem, stw);
        lds_barrier();
        // the forward scale belongs to this pass when x is length 1
        const float sc = x.cls[2] < 0 ? x.scale_r2c : 1.0f;
        PP::template y_pass<false>(smem, smem + PP::PLANE, stw, sc, sc != 1.0f);
        lds_barrier();
        for (int i = threadIdx.x; i < PP::PLANE; i += kBT) st_stream(g + i, smem[i]);
      } else {
        PP::template y_pass<true>(smem, smem + PP::PLANE, stw, 1.0f, false);
        lds_barrier();
        PP::z_c2r(smem, stw, x.scale_c2r, x.scale_c2r != 1.0f);
        lds_barrier();
        constexpr int H = PP::NZ / 2;
        for (int i = threadIdx.x; i < PP::NY * H; i += kBT) {
          const int y = i / H, m = i - y * H;
          st_stream(g + y * PP::K + m, smem[y * PP::K + m]);
        }
      }
      lds_barrier();  // the stores' LDS reads are done before the next plane lands
    }
  }

  // x-planes through the plane pass, one wavefront per plane: wavefront w
  // takes elements w, w + 4, ... of the tile list (first, first + step, ...)
  // (wave_planes in gfft_batch_tiles.h).
  __device__ static void planes_wave(const BXform& x, bool inv, uint32_t first, uint32_t step, uint32_t count,
                                     const cx<float>* stw, cx<float>* smem) {
    using PP = BPlaneW<F::CLS, F::VAR>;
    const uint32_t w = threadIdx.x >> 6;
    const BXform* xp = &x;
    wave_planes<PP>(
        [&](uint32_t i, const BXform** px, uint32_t* plane) {
          const uint32_t e = w + i * kWaves;
          *px = xp;
          *plane = first + e * step;
          return e < count;
        },
        inv, stw, smem + w * (PP::PLANE + PP::SCRATCH));
  }

  template <bool INV>
  __device__ static void cols(const BXform& x, int axis, uint32_t first, uint32_t step, uint32_t count,
                              const cx<float>* stw, cx<float>* smem) {
    using K_ = COLS<INV>;
    ColGeom g;
    if (axis == 1) {
      g.ncols = x.ncols_y;
      g.in = g.out = dmap((uint32_t)x.K, (int64_t)x.ny * x.K, x.K);
    } else {
      g.ncols = x.ncols_x;
      g.in = g.out = dmap(x.ncols_x, 0, (int64_t)x.ncols_x);
    }
    // the forward scale belongs to the last R2C pass (x, or y when x is length 1)
    const bool last_fwd = !INV && (axis == 2 || (axis == 1 && x.cls[2] < 0));
    g.scale = last_fwd ? x.scale_r2c : 1.0;
    cx<float>* d = reinterpret_cast<cx<float>*>(x.data);
    const int c0 = threadIdx.x % W, t0 = threadIdx.x / W;
    cx<float> cur[K_::E::PV], nxt[K_::E::PV];
    K_::load(nxt, d, g, first, c0, t0);
    for (uint32_t k = 0; k < count; ++k) {
      int c = c0, t = t0;
      __asm__ volatile("" : "+v"(c), "+v"(t));
      ColLds<float, W> L{smem, c};
#pragma unroll
      for (int i = 0; i < K_::E::PV; ++i) cur[i] = nxt[i];
      if (k + 1 < count) K_::load(nxt, d, g, first + (k + 1) * step, c, t);
      K_::E::run(cur, t, L, stw);
      K_::store(cur, d, g, first + k * step, c, t);
    }
  }

  // inlined into each set's class_tiles, so that each copy is compiled under
  // its own kernel's register limit
  __device__ __forceinline__ static void run(const BXform& x, int axis, int inv, uint32_t first, uint32_t step,
                                             uint32_t count, const cx<float>* tw, cx<float>* smem) {
    cx<float>* stw = smem + base_entries();
    for (int i = threadIdx.x; i < SC::twoff(SC::S); i += kBT) stw[i] = tw[i];
    __syncthreads();  // the table is complete (and the previous call's LDS reads are done)
    if (count > 0) {
      if (axis == 0) {
        if (inv)
          rows_c2r(x, first, step, count, stw, smem);
        else
          rows_r2c(x, first, step, count, stw, smem);
      } else if (inv) {
        cols<true>(x, axis, first, step, count, stw, smem);
      } else {
        cols<false>(x, axis, first, step, count, stw, smem);
      }
    }
    __syncthreads();  // every LDS access of this call is done before the next call's writes
  }

  // the fused plane level's tiles (plane classes only)
  __device__ __forceinline__ static void run_planes(const BXform& x, int inv, uint32_t first, uint32_t step,
                                                    uint32_t count, const cx<float>* tw, cx<float>* smem) {
    cx<float>* stw = smem + base_entries();
    for (int i = threadIdx.x; i < SC::twoff(SC::S); i += kBT) stw[i] = tw[i];
    __syncthreads();  // the table is complete (and the previous call's LDS reads are done)
    if (count > 0) {
      if constexpr (kPlaneWave) {
        planes_wave(x, inv != 0, first, step, count, stw, smem);
      } else if constexpr (kPlanePipe) {
        planes(x, inv != 0, first, step, count, stw, smem);
      } else if constexpr (kPlane) {
        for (uint32_t k = 0; k < count; ++k)
          plane_tile<BPlane<F::CLS, F::VAR>>(x, inv != 0, first + k * step, stw, smem);
      }
    }
    __syncthreads();  // every LDS access of this call is done before the next call's writes
  }
};

// One class's tile code as a function of its own (not inlined): its
// registers are allocated for it alone, so the kernel needs the largest
// class's count rather than what one function holding all eight would.
template <int S, int C>
__device__ __attribute__((noinline)) void class_tiles(const BXform& x, int axis, int inv, uint32_t first,
                                                      uint32_t step, uint32_t count, const cx<float>* tw,
                                                      cx<float>* smem) {
  Seq<SCfg<S, C>>::run(x, axis, inv, first, step, count, tw, smem);
}

// A plane class's fused plane levels, in a function of its own (see
// class_tiles).
template <int S, int C>
__device__ __attribute__((noinline)) void class_planes(const BXform& x, int inv, uint32_t first, uint32_t step,
                                                       uint32_t count, const cx<float>* tw, cx<float>* smem) {
  Seq<SCfg<S, C>>::run_planes(x, inv, first, step, count, tw, smem);
}

template <int S>
__device__ __forceinline__ void run_set_tiles(const BXform& x, const PSeg& sg, uint32_t first, uint32_t step,
                                              uint32_t count, const PersistArgs& a, cx<float>* smem) {
  static_assert(kPersistPlaneClasses == 5, "plane classes 0 .. 4");
  if (sg.axis == kAxisPlane) {
    switch (sg.cls) {
      case 0: class_planes<S, 0>(x, sg.inv, first, step, count, a.tw[0], smem); break;
      case 1: class_planes<S, 1>(x, sg.inv, first, step, count, a.tw[1], smem); break;
      case 2: class_planes<S, 2>(x, sg.inv, first, step, count, a.tw[2], smem); break;
      case 3: class_planes<S, 3>(x, sg.inv, first, step, count, a.tw[3], smem); break;
      default: class_planes<S, 4>(x, sg.inv, first, step, count, a.tw[4], smem); break;
    }
    return;
  }
  switch (sg.cls) {
    case 0: class_tiles<S, 0>(x, sg.axis, sg.inv, first, step, count, a.tw[0], smem); break;
    case 1: class_tiles<S, 1>(x, sg.axis, sg.inv, first, step, count, a.tw[1], smem); break;
    case 2: class_tiles<S, 2>(x, sg.axis, sg.inv, first, step, count, a.tw[2], smem); break;
    case 3: class_tiles<S, 3>(x, sg.axis, sg.inv, first, step, count, a.tw[3], smem); break;
    case 4: class_tiles<S, 4>(x, sg.axis, sg.inv, first, step, count, a.tw[4], smem); break;
    case 5: class_tiles<S, 5>(x, sg.axis, sg.inv, first, step, count, a.tw[5], smem); break;
    case 6: class_tiles<S, 6>(x, sg.axis, sg.inv, first, step, count, a.tw[6], smem); break;
    default: class_tiles<S, 7>(x, sg.axis, sg.inv, first, step, count, a.tw[7], smem); break;
  }
}

__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool launch_failed(const PersistArgs& a) { return ld_agent(a.ctr + kCtrFail) != 0u; }

// The XCC (L2 domain) this wavefront runs on.
__device__ __forceinline__ uint32_t xcc_id() {
  uint32_t v;
  __asm__ volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
  return v & 0xfu;
}

// This wavefront's global stores have completed (they are in the L2).
__device__ __forceinline__ void stores_complete() { __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Later loads of this wavefront see what the L2 holds (the L1 is dropped).
__device__ __forceinline__ void acquire_level() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

// Agent-scope release followed by a flag store: the fence writes this XCD's
// L2 back (buffer_wbl2 sc1); the explicit wait keeps the store behind the
// write-back's completion (a release store compiled to the write-back and the
// store with no wait between them, so the flag could be seen before the data
// reached memory).
__device__ __forceinline__ void release_store(uint32_t* p, uint32_t v) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// share m of level sg (tiles m, m + n, ... of n shares), then publish it as
// done by a workgroup of team part `part`
template <int S>
__device__ __forceinline__ void run_share(const PersistArgs& a, const BXform& x, const PSeg& sg, uint32_t n,
                                          uint32_t m, uint32_t part, cx<float>* smem) {
  run_set_tiles<S>(x, sg, m, n, share_tiles(sg.ntiles, n, m), a, smem);
  stores_complete();  // this thread's stores are in the L2
  __syncthreads();    // every thread's
  if (threadIdx.x == 0)
    __hip_atomic_store(a.ctr + kCtrShares + sg.claim0 + m, kShareDone + part, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool claim_share(const PersistArgs& a, const PSeg& sg, uint32_t m) {
  uint32_t expect = kShareFree;
  return __hip_atomic_compare_exchange_strong(a.ctr + kCtrShares + sg.claim0 + m, &expect, kShareClaimed,
                                              __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait-step outcomes, decided by the first wavefront
constexpr uint32_t kStepDone = 0xffffffffu, kStepFail = 0xfffffffeu, kStepSpin = 0xfffffffdu,
                   kStepRun = 0xfffffffcu;

// State of one workgroup's wait (first wavefront).
struct WaitState {
  uint64_t t0;       // wall clock at the start of the wait
  uint64_t checked;  // wall clock of the last look at the failure and abort words (0: none yet)
  uint32_t next;     // where the next steal attempt starts (the scan runs downwards)
  uint32_t nap;      // sleep length of the next spin (grows to kMaxNap)
  uint32_t phase;    // global-level plans: the leader's hand-over phase (wide_wait_step)
  uint64_t steal_at; // global-level plans: wall clock of the next steal attempt
};

constexpr uint32_t kMaxNap = 16;            // x 512 clocks: about 4 us between looks at most
constexpr uint32_t kMaxWideNap = 1;         // level hand-overs: looks at most ~0.2 us apart
constexpr uint64_t kCheckTicks = 1000;      // failure / abort words read at most every 10 us

// Bitwise OR of a value over the 64 lanes of the calling wavefront.
__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o);
  return v;
}

// First wavefront only (all 64 lanes): is level sg of a team of k members,
// q shares each and nparts parts complete for a workgroup of part `part`?
// Every share done and, for a span team, every part that ran shares of it
// released; this workgroup releases its own part when it is the first to
// claim the part's release word (its XCD's L2 written back, then the flag).
// *any_free: a share of the level was seen unclaimed.
__device__ __forceinline__ bool level_complete(const PersistArgs& a, const PSeg& sg, uint32_t k, uint32_t q,
                                               uint32_t nparts, uint32_t part, bool* any_free) {
  const uint32_t lane = threadIdx.x;
  const uint32_t ns = k * q;
  uint32_t* w = a.ctr + kCtrShares + sg.claim0;
  bool mine_done = true, mine_free = false;
  uint32_t mask = 0;
  for (uint32_t j = lane; j < ns; j += 64) {
    const uint32_t v = ld_agent(w + j);
    mine_done = mine_done && v >= kShareDone;
    mine_free = mine_free || v == kShareFree;
    if (v >= kShareDone) mask |= 1u << ((v - kShareDone) & 31u);
  }
  *any_free = __ballot(mine_free) != 0ull;
  if (!__all(mine_done)) return false;
  if (nparts < 2) return true;
  mask = wave_or(mask);
  uint32_t* rel = w + ns;  // the level's release words, one per part
  if ((mask >> part) & 1u) {
    uint32_t won = 0;
    if (lane == 0 && ld_agent(rel + part) == kRelNone) {
      uint32_t expect = kRelNone;
      won = __hip_atomic_compare_exchange_strong(rel + part, &expect, kRelClaimed, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)
                ? 1u
                : 0u;
    }
    if (lane == 0 && won) release_store(rel + part, kRelDone);  // the part's stores: from its L2 to memory
  }
  bool ok = true;
  if (lane < nparts) ok = ((mask >> lane) & 1u) == 0u || ld_agent(rel + lane) == kRelDone;
  return __all(ok);
}

// First wavefront only (all 64 lanes): one look at the level of team tm
// (this workgroup of part `part`).
// Returns kStepDone, kStepFail, kStepSpin, or the index (< k q) of a share
// this workgroup has just claimed (steal).  Every decision is made by lane 0
// and broadcast, so the wavefront's control flow stays uniform.  The
// launch-wide failure word is read at most every kCheckTicks, the host's
// abort word (pinned host memory, read across PCIe) only by watch_host
// workgroups (one per team) at that rate, which turn it into the failure
// word; the sleep between looks grows, so the waiting workgroups put little
// traffic on the memory system.
__device__ __forceinline__ uint32_t wait_step(const PersistArgs& a, const PSeg& sg, const PTeam& tm, uint32_t part,
                                              bool watch_host, WaitState* ws) {
  const uint32_t lane = threadIdx.x;
  const uint32_t k = tm.k * tm.q;  // shares of the level
  const uint32_t* w = a.ctr + kCtrShares + sg.claim0;
  bool any_free = false;
  if (level_complete(a, sg, tm.k, tm.q, tm.nparts, part, &any_free)) return kStepDone;
  uint32_t out = kStepSpin;
  if (lane == 0) {
    const uint64_t now = wall_clock64();
    const uint64_t el = now - ws->t0;
    if (now - ws->checked > kCheckTicks || el > a.wait_ticks) {
      ws->checked = now;
      if (launch_failed(a)) {
        out = kStepFail;
      } else if (watch_host && a.host_abort &&
                 __hip_atomic_load(a.host_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u) {
        __hip_atomic_fetch_or(a.ctr + kCtrFail, kFailAborted, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        out = kStepFail;
      } else if (el > a.wait_ticks) {
        __hip_atomic_fetch_or(a.ctr + kCtrFail, kFailTimeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        out = kStepFail;
      }
    }
    if (out == kStepSpin && any_free && el > a.steal_ticks) {
      // steal the first free share at or below the rotating start (owners
      // run their shares upwards, so a slow member's last ones come first)
      for (uint32_t n = 0; n < k; ++n) {
        const uint32_t j = ws->next >= n ? ws->next - n : ws->next + k - n;
        if (ld_agent(w + j) != kShareFree) continue;
        uint32_t expect = kShareFree;
        if (__hip_atomic_compare_exchange_strong(a.ctr + kCtrShares + sg.claim0 + j, &expect, kShareClaimed,
                                                 __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          out = j;
          ws->next = j > 0 ? j - 1 : k - 1;
          break;
        }
      }
    }
  }
  out = __shfl(out, 0);
  if (out == kStepSpin) {
    for (uint32_t i = 0; i < ws->nap; ++i) __builtin_amdgcn_s_sleep(8);
    ws->nap = ws->nap < kMaxNap ? 2 * ws->nap : kMaxNap;
  } else {
    ws->nap = 1;
  }
  return out;
}

// First wavefront only (all 64 lanes): one look for work among the teams of
// domain `dom` (the first 64 of its list) whose first member in the domain
// (of the team's part there) has arrived.  Lane i keeps team i's current
// level (the first not complete, see level_complete) in *lvl between looks.
// Returns kStepDone (every such team is complete), kStepFail, kStepSpin
// (nothing free now), or kStepRun with *team, *seg, *share and *part set: a
// share of a current level just claimed.
// Timeouts and the failure word are handled as in wait_step; the host's
// abort word is left to the teams' first members.
__device__ __forceinline__ uint32_t help_step(const PersistArgs& a, uint32_t dom, uint32_t* lvl, WaitState* ws,
                                              uint32_t* team, uint32_t* seg, uint32_t* share, uint32_t* part) {
  const uint32_t lane = threadIdx.x;
  const uint32_t t0 = a.dom_off[dom], nt = min(a.dom_off[dom + 1] - t0, 64u);
  const uint32_t arrived = ld_agent(a.ctr + kCtrArrive + dom);
  bool unfinished = false;
  for (uint32_t i = 0; i < nt; ++i) {
    const uint32_t t = a.dom_teams[t0 + i];
    const PTeam tm = a.team[t];
    uint32_t pi = 0;  // the team's part in this domain
    while (pi + 1 < tm.nparts && a.parts[tm.part0 + pi].dom != dom) ++pi;
    if (arrived <= a.parts[tm.part0 + pi].slot0) continue;  // no member yet: possibly an orphan, run whole at the end
    const uint32_t ns = tm.k * tm.q;
    uint32_t cur = __shfl(*lvl, (int)i);
    while (cur < tm.nseg) {
      const PSeg sg = a.seg[tm.seg0 + cur];
      uint32_t* w = a.ctr + kCtrShares + sg.claim0;
      bool any_free = false;
      if (level_complete(a, sg, tm.k, tm.q, tm.nparts, pi, &any_free)) {
        ++cur;
        continue;
      }
      if (!any_free) break;  // the level is under way (or handing over) with nothing free
      uint32_t free_j = kStepDone;
      for (uint32_t j = lane; j < ns; j += 64)
        if (free_j == kStepDone && ld_agent(w + j) == kShareFree) free_j = j;
      uint64_t m = __ballot(free_j != kStepDone);
      while (m) {
        const int b = __ffsll((long long)m) - 1;
        m &= m - 1;
        const uint32_t j = __shfl(free_j, b);
        uint32_t got = 0;
        if (lane == 0) {
          uint32_t expect = kShareFree;
          got = __hip_atomic_compare_exchange_strong(w + j, &expect, kShareClaimed, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                    ? 1u
                    : 0u;
        }
        if (__shfl(got, 0)) {
          if (lane == i) *lvl = cur;
          *team = t;
          *seg = tm.seg0 + cur;
          *share = j;
          *part = pi;
          ws->nap = 1;
          ws->t0 = wall_clock64();
          return kStepRun;
        }
      }
      break;  // the level is under way with nothing free
    }
    if (lane == i) *lvl = cur;
    unfinished = unfinished || cur < tm.nseg;
  }
  if (!unfinished) return kStepDone;
  uint32_t out = kStepSpin;
  if (lane == 0) {
    const uint64_t now = wall_clock64();
    const uint64_t el = now - ws->t0;
    if (now - ws->checked > kCheckTicks || el > a.wait_ticks) {
      ws->checked = now;
      if (launch_failed(a)) {
        out = kStepFail;
      } else if (el > a.wait_ticks) {
        __hip_atomic_fetch_or(a.ctr + kCtrFail, kFailTimeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        out = kStepFail;
      }
    }
  }
  out = __shfl(out, 0);
  if (out == kStepSpin) {
    for (uint32_t i = 0; i < ws->nap; ++i) __builtin_amdgcn_s_sleep(8);
    ws->nap = ws->nap < kMaxNap ? 2 * ws->nap : kMaxNap;
  }
  return out;
}

// A helper's loop; sh: the body's LDS words.  Returns true when the launch
// failed.  The per-class tile functions save no callee-saved registers only
// while no call of theirs may be a tail call (LLVM's no-CSR rule for local
// functions); a call whose arguments all point outside the caller's frame
// may be, so the helper passes local copies of the descriptors, as the other
// callers do (with references into the plan arrays the 96 and 128 functions
// saved registers into accumulator registers, and the kernel fell to one
// wave per SIMD).
template <int S>
__device__ __forceinline__ bool help_teams(const PersistArgs& a, uint32_t dom, cx<float>* smem,
                                                     uint32_t* sh) {
  uint32_t *s_step = sh, *s_team = sh + 3, *s_seg = sh + 4, *s_share = sh + 5, *s_part = sh + 7;
  const uint32_t tid = threadIdx.x;
  uint32_t lvl = 0;  // first wavefront: lane i, team i's current level
  WaitState ws{0, 0, 0, 1, 0, 0};
  if (tid < 64) ws.t0 = wall_clock64();
  for (;;) {
    if (tid < 64) {
      uint32_t ht = 0, hs = 0, hj = 0, hp = 0;
      const uint32_t st = help_step(a, dom, &lvl, &ws, &ht, &hs, &hj, &hp);
      if (tid == 0) {
        *s_step = st;
        *s_team = ht;
        *s_seg = hs;
        *s_share = hj;
        *s_part = hp;
      }
    }
    __syncthreads();
    const uint32_t st = *s_step, ht = *s_team, hs = *s_seg, hj = *s_share, hp = *s_part;
    __syncthreads();  // read before the next look writes them
    if (st == kStepDone) return false;
    if (st == kStepFail) return true;
    if (st == kStepRun) {
      acquire_level();  // the level before is complete in the L2
      const PTeam tm = a.team[ht];
      const BXform x = a.xf[ht];
      const PSeg sg = a.seg[hs];
      run_share<S>(a, x, sg, tm.k * tm.q, hj, hp, smem);
    }
  }
}

// every level of transform t, by this workgroup alone
template <int S>
__device__ void run_whole(const PersistArgs& a, uint32_t t, cx<float>* smem) {
  const PTeam tm = a.team[t];
  const BXform x = a.xf[t];
  for (uint32_t s = tm.seg0; s < tm.seg0 + tm.nseg; ++s) {
    const PSeg sg = a.seg[s];
    run_set_tiles<S>(x, sg, 0, 1, sg.ntiles, a, smem);
  }
}

// The kernel's body; sh: eight LDS words (step, last, claimed, team, member,
// job, domain, part).
template <int S>
__device__ __forceinline__ void persist_body(const PersistArgs& a, cx<float>* smem, uint32_t* sh) {
  uint32_t *s_step = sh, *s_last = sh + 1, *s_claimed = sh + 2, *s_team = sh + 3, *s_member = sh + 4,
           *s_job = sh + 5, *s_dom = sh + 6, *s_part = sh + 7;
  const uint32_t tid = threadIdx.x;
  bool failed = false;
  // the next local slot of this workgroup's domain
  if (tid == 0) {
    const uint32_t d = xcc_id();
    uint32_t team = kNoTeam, member = 0, part = 0;
    if (d < a.domains) {
      const uint32_t li = __hip_atomic_fetch_add(a.ctr + kCtrArrive + d, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (li < a.per_dom) {
        const PEntry e = a.local[d * a.per_dom + li];
        team = e.team;
        member = e.member;
        part = e.part;
      }
    }
    *s_team = team;
    *s_member = member;
    *s_dom = d;
    *s_part = part;
  }
  __syncthreads();
  const uint32_t team = *s_team, member = *s_member, dom = *s_dom, part = *s_part;
  if (team != kNoTeam) {  // a team member: per level the own share, then wait (stealing) for the level
    const PTeam tm = a.team[team];
    const BXform x = a.xf[team];
    const uint32_t ns = tm.k * tm.q;  // shares of a level
    for (uint32_t s = tm.seg0; s < tm.seg0 + tm.nseg && !failed; ++s) {
      const PSeg sg = a.seg[s];
      for (uint32_t i = 0; i < tm.q; ++i) {  // the own shares, in turn
        const uint32_t own = member + i * tm.k;
        if (tid == 0) *s_claimed = claim_share(a, sg, own) ? 1u : 0u;
        __syncthreads();
        const bool mine = *s_claimed != 0u;
        __syncthreads();  // s_claimed is read before the next claim writes it
        if (mine) run_share<S>(a, x, sg, ns, own, part, smem);
      }
      WaitState ws{0, 0, ns - 1 - member, 1, 0, 0};
      if (tid < 64) ws.t0 = wall_clock64();  // checked = 0: the first look reads the launch-wide words
      for (;;) {
        if (tid < 64) {
          const uint32_t st = wait_step(a, sg, tm, part, member == 0, &ws);
          if (tid == 0) *s_step = st;
        }
        __syncthreads();
        const uint32_t st = *s_step;
        __syncthreads();  // s_step is read before the next step writes it
        if (st == kStepDone) break;
        if (st == kStepFail) {
          failed = true;
          break;
        }
        if (st < ns) run_share<S>(a, x, sg, ns, st, part, smem);  // a stolen share
      }
      if (!failed) acquire_level();  // the level's stores, from the L2
    }
  }
  // solo transforms from the shared queue
  while (!failed) {
    if (tid == 0)
      *s_job = launch_failed(a) ? kNoTeam
                                : __hip_atomic_fetch_add(a.ctr + kCtrSolo, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const uint32_t j = *s_job;
    __syncthreads();  // s_job is read before the next ticket writes it
    if (j >= a.nsolo) break;
    run_whole<S>(a, a.solo[j], smem);
  }
  // then help the teams of this workgroup's domain
  if (!failed && dom < a.domains && a.dom_off[dom + 1] > a.dom_off[dom]) failed = help_teams<S>(a, dom, smem, sh);
  // leave; the last workgroup out runs the orphaned teams and resets the counters
  if (tid == 0) {
    const uint32_t prev = __hip_atomic_fetch_add(a.ctr + kCtrExited, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    *s_last = prev == gridDim.x - 1 ? 1u : 0u;
  }
  __syncthreads();
  if (*s_last) {
    if (!launch_failed(a))
      for (uint32_t i = 0; i < a.nteams; ++i) {
        const uint32_t t = a.team_ids[i];
        const PTeam tm = a.team[t];
        bool orphan = true;  // no member of any part arrived
        for (uint32_t i = 0; i < tm.nparts; ++i) {
          const PPart pt = a.parts[tm.part0 + i];
          orphan = orphan && ld_agent(a.ctr + kCtrArrive + pt.dom) <= pt.slot0;
        }
        if (orphan) run_whole<S>(a, t, smem);
      }
    if (tid == 0) {
      const uint32_t f = ld_agent(a.ctr + kCtrFail);
      if (f) __hip_atomic_fetch_or(a.status, f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const uint32_t words = kCtrShares + a.claim_words;
    for (uint32_t i = tid; i < words; i += kBT)
      if (a.ctr + i != a.status) __hip_atomic_store(a.ctr + i, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ---------------------------------------------------------------------------
// Global-level plans
// ---------------------------------------------------------------------------
// Sum of a value over the 64 lanes of the calling wavefront (every lane gets it).
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Run the list of slot `slot` in level g (its shares, each a set of tiles of
// one transform level); returns the number of shares.
template <int S>
__device__ __forceinline__ uint32_t run_list(const PersistArgs& a, uint32_t g, uint32_t slot, cx<float>* smem) {
  const uint32_t b0 = a.own_off[g * a.slots + slot], b1 = a.own_off[g * a.slots + slot + 1];
  for (uint32_t i = b0; i < b1; ++i) {
    const WShare w = a.wshare[a.own[i]];
    const PSeg sg = a.seg[w.seg];
    const BXform x = a.xf[sg.xf];
    run_set_tiles<S>(x, sg, w.first, w.step, w.count, a, smem);
  }
  return b1 - b0;
}

// Every thread's stores are in the L2; then thread 0 adds n to its domain's
// done word (its domain's leader writes the L2 back once the level is
// complete).  A workgroup without a domain releases itself (the L2 written
// back before the add).
__device__ __forceinline__ void publish(uint32_t* word, uint32_t n, bool own_release) {
  stores_complete();
  __syncthreads();
  if (threadIdx.x == 0 && n) {
    if (own_release)
      __hip_atomic_fetch_add(word, n, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    else
      __hip_atomic_fetch_add(word, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// First wavefront only: claim an unclaimed slot list of level g (a slot whose
// workgroup has not run it), scanning `windows` windows of 64 claim words (one
// load per lane each) from ws->next onward.  Returns the claimed slot, or
// kStepSpin.
__device__ __forceinline__ uint32_t wide_steal(const PersistArgs& a, uint32_t g, WaitState* ws, int windows) {
  const uint32_t lane = threadIdx.x;
  uint32_t* w = a.ctr + kCtrShares + (size_t)g * a.slots;
  const uint32_t n = a.slots;
  uint32_t out = kStepSpin;
  for (int win = 0; win < windows && out == kStepSpin; ++win) {
    const uint32_t base = ws->next;
    const uint32_t s = (base + lane) % n;
    const bool free_ = lane < n && ld_agent(w + s) == kShareFree &&
                       a.own_off[g * n + s + 1] > a.own_off[g * n + s];  // lists with work only
    uint64_t m = __ballot(free_);
    while (m && out == kStepSpin) {
      const uint32_t b = (uint32_t)(__ffsll((long long)m) - 1);
      m &= m - 1;
      uint32_t got = 0;
      if (lane == 0) {
        uint32_t expect = kShareFree;
        got = __hip_atomic_compare_exchange_strong(w + (base + b) % n, &expect, kShareClaimed, __ATOMIC_RELAXED,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                  ? 1u
                  : 0u;
      }
      if (__shfl(got, 0)) out = (base + b) % n;
    }
    ws->next = (base + 64) % n;
  }
  return out;
}

// First wavefront only: one look at level g's hand-over (block L of
// kLevelWords words).  The domain's leader (or a workgroup without a domain)
// runs the hand-over in phases kept in ws->phase: 0 = the done words sum to
// the level's share count; 1 = its domain's L2 written back if the domain did
// work, released[dom] set; 2 = every domain that did work has released, go[dom]
// set.  The other workgroups watch go[dom].  Returns kStepDone, kStepFail,
// kStepSpin or a slot whose list was just claimed (a steal, tried after a
// wait of steal_ticks without completion).
__device__ __forceinline__ uint32_t wide_wait_step(const PersistArgs& a, uint32_t g, uint32_t dom, bool leader,
                                                   WaitState* ws) {
  // only the leaders read the host's abort word (across PCIe; every waiting
  // workgroup reading it at each level boundary held each look up for tens of
  // microseconds); they turn it into the launch's failure word
  const bool watch_host = leader;
  const uint32_t lane = threadIdx.x;
  uint32_t* L = a.ctr + a.level_base + (size_t)g * kLevelWords;
  const uint32_t D = a.domains;
  bool done = false;
  if (leader) {
    if (ws->phase == 0) {
      uint32_t v = lane <= kMaxDomains ? ld_agent(L + kLevelDone + lane * kFlagStride) : 0u;
      if (wave_sum(v) >= a.lvl_off[g + 1] - a.lvl_off[g]) ws->phase = 1;
    }
    if (ws->phase == 1) {
      if (dom < D) {
        const uint32_t mine = ld_agent(L + kLevelDone + dom * kFlagStride);
        if (lane == 0 && mine)  // this domain's stores of the level: from its L2 to memory
          release_store(L + kLevelReleased + dom * kFlagStride, 1u);
      }
      ws->phase = 2;
    }
    if (ws->phase == 2) {
      bool ok = true;
      if (lane < D)
        ok = ld_agent(L + kLevelDone + lane * kFlagStride) == 0u ||
             ld_agent(L + kLevelReleased + lane * kFlagStride) != 0u;
      done = __all(ok);
      if (done && lane == 0 && dom < D)
        __hip_atomic_store(L + kLevelGo + dom * kFlagStride, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  } else {
    done = ld_agent(L + kLevelGo + dom * kFlagStride) != 0u;
  }
  if (done) return kStepDone;
  uint32_t out = kStepSpin;
  bool steal = false;
  if (lane == 0) {
    const uint64_t now = wall_clock64();
    const uint64_t el = now - ws->t0;
    if (now - ws->checked > kCheckTicks || el > a.wait_ticks) {
      ws->checked = now;
      if (launch_failed(a)) {
        out = kStepFail;
      } else if (watch_host && a.host_abort &&
                 __hip_atomic_load(a.host_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u) {
        __hip_atomic_fetch_or(a.ctr + kCtrFail, kFailAborted, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        out = kStepFail;
      } else if (el > a.wait_ticks) {
        __hip_atomic_fetch_or(a.ctr + kCtrFail, kFailTimeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        out = kStepFail;
      }
    }
    steal = out == kStepSpin && ws->phase == 0 && now >= ws->steal_at;
    if (steal) ws->steal_at = now + a.steal_ticks;  // a scan that finds nothing waits again
  }
  out = __shfl(out, 0);
  if (__shfl(steal ? 1u : 0u, 0)) out = wide_steal(a, g, ws, (int)((a.slots + 63) / 64));
  if (out == kStepSpin) {
    for (uint32_t i = 0; i < ws->nap; ++i) __builtin_amdgcn_s_sleep(8);
    ws->nap = ws->nap < kMaxWideNap ? 2 * ws->nap : kMaxWideNap;
  } else {
    ws->nap = 1;
  }
  return out;
}

// The kernel's body for a global-level plan; sh as for persist_body.
template <int S>
__device__ __forceinline__ void wide_body(const PersistArgs& a, cx<float>* smem, uint32_t* sh) {
  uint32_t *s_step = sh, *s_last = sh + 1, *s_claimed = sh + 2, *s_slot = sh + 3, *s_dom = sh + 4,
           *s_poll = sh + 5;
  const uint32_t tid = threadIdx.x;
  uint32_t* claim_w = a.ctr + kCtrShares;  // per level and slot: its list claimed
  if (tid == 0) {
    const uint32_t d = xcc_id();
    // no domain: no slot (only steals), releases its own work, leads itself
    uint32_t slot = a.slots, dom = kMaxDomains, poll = 1;
    if (d < a.domains) {
      const uint32_t li = __hip_atomic_fetch_add(a.ctr + kCtrArrive + d, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      dom = d;
      poll = li == 0 ? 1u : 0u;  // the domain's first workgroup leads its hand-overs
      if (li < a.per_dom) slot = d * a.per_dom + li;
    }
    *s_slot = slot;
    *s_dom = dom;
    *s_poll = poll;
  }
  __syncthreads();
  const uint32_t slot = *s_slot, dom = *s_dom;
  const bool leader = *s_poll != 0u;
  bool failed = false, solo_left = a.nsolo > 0;
  for (uint32_t g = 0; g < a.nlevels && !failed; ++g) {
    uint32_t* done = a.ctr + a.level_base + (size_t)g * kLevelWords + kLevelDone + dom * kFlagStride;
    const bool own_release = dom >= a.domains;
    uint32_t ran = 0;
    if (slot < a.slots) {
      if (tid == 0) {
        uint32_t expect = kShareFree;
        *s_claimed = __hip_atomic_compare_exchange_strong(claim_w + (size_t)g * a.slots + slot, &expect, kShareClaimed,
                                                          __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                         ? 1u
                         : 0u;
      }
      __syncthreads();
      const bool mine = *s_claimed != 0u;
      __syncthreads();  // s_claimed is read before the next claim writes it
      if (mine) ran = run_list<S>(a, g, slot, smem);
    }
    publish(done, ran, own_release);
    WaitState ws{0, 0, (uint32_t)(((uint64_t)(slot < a.slots ? slot : 0) + 1) % a.slots), 1, 0, 0};
    if (tid < 64) {
      ws.t0 = wall_clock64();
      ws.steal_at = ws.t0 + a.steal_ticks;
    }
    for (;;) {
      if (tid < 64) {
        const uint32_t st = wide_wait_step(a, g, dom, leader, &ws);
        if (tid == 0) *s_step = st;
      }
      __syncthreads();
      const uint32_t st = *s_step;
      __syncthreads();  // s_step is read before the next step writes it
      if (st == kStepDone) break;
      if (st == kStepFail) {
        failed = true;
        break;
      }
      if (st < a.slots) {  // an absent workgroup's list
        publish(done, run_list<S>(a, g, st, smem), own_release);
      } else if (!leader && solo_left) {
        // a solo transform while the level hands over (the leaders keep
        // their hand-over duties)
        if (tid == 0) *s_claimed = __hip_atomic_fetch_add(a.ctr + kCtrSolo, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const uint32_t j = *s_claimed;
        __syncthreads();
        if (j < a.nsolo)
          run_whole<S>(a, a.solo[j], smem);
        else
          solo_left = false;
      }
    }
    if (!failed) acquire_level();  // the level's stores, from memory
  }
  // the solo transforms still queued
  while (!failed && solo_left) {
    if (tid == 0)
      *s_claimed = launch_failed(a) ? kNoTeam
                                    : __hip_atomic_fetch_add(a.ctr + kCtrSolo, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const uint32_t j = *s_claimed;
    __syncthreads();
    if (j >= a.nsolo) break;
    run_whole<S>(a, a.solo[j], smem);
  }
  // leave; the last workgroup out moves the failure bits and resets the counters
  if (tid == 0) {
    const uint32_t prev = __hip_atomic_fetch_add(a.ctr + kCtrExited, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    *s_last = prev == gridDim.x - 1 ? 1u : 0u;
  }
  __syncthreads();
  if (*s_last) {
    if (tid == 0) {
      const uint32_t f = ld_agent(a.ctr + kCtrFail);
      if (f) __hip_atomic_fetch_or(a.status, f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const uint32_t words = a.level_base + a.nlevels * kLevelWords;
    for (uint32_t i = tid; i < words; i += kBT)
      if (a.ctr + i != a.status) __hip_atomic_store(a.ctr + i, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <int S>
__global__ __launch_bounds__(kBT) void k_batch_persist(PersistArgs a) {
  extern __shared__ cx<float> smem[];
  __shared__ uint32_t sh[8];
  if (a.wide)
    wide_body<S>(a, smem, sh);
  else
    persist_body<S>(a, smem, sh);
}

template <int S>
size_t lds_of_set(const bool* present) {
  const int e[kClasses] = {Seq<SCfg<S, 0>>::lds_entries(), Seq<SCfg<S, 1>>::lds_entries(),
                           Seq<SCfg<S, 2>>::lds_entries(), Seq<SCfg<S, 3>>::lds_entries(),
                           Seq<SCfg<S, 4>>::lds_entries(), Seq<SCfg<S, 5>>::lds_entries(),
                           Seq<SCfg<S, 6>>::lds_entries(), Seq<SCfg<S, 7>>::lds_entries()};
  int m = 1;
  for (int c = 0; c < kClasses; ++c)
    if (present[c] && e[c] > m) m = e[c];
  return (size_t)m * sizeof(cx<float>);
}

}  // namespace

const void* persist_kernel(int set) {
  return set == 0 ? reinterpret_cast<const void*>(&k_batch_persist<0>)
                  : reinterpret_cast<const void*>(&k_batch_persist<1>);
}

size_t persist_lds_bytes(int set, const bool* present) {
  return set == 0 ? lds_of_set<0>(present) : lds_of_set<1>(present);
}

hipError_t static_counters(uint32_t** out, hipStream_t s) {
  static std::mutex mu;
  static std::map<int, uint32_t*> ptr;  // per device: the array's device address
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(mu);
  auto it = ptr.find(dev);
  if (it == ptr.end()) {
    uint32_t* p = nullptr;
    e = hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    // first use on this device: clear it with a kernel (stream-ordered, and a
    // kernel node if the caller captures the stream)
    size_t n = kStaticCounterWords;
    void* args[] = {&p, &n};
    e = hipLaunchKernel(reinterpret_cast<const void*>(&k_fill_zero), dim3(64), dim3(256), args, 0, s);
    if (e != hipSuccess) return e;
    it = ptr.emplace(dev, p).first;
  }
  *out = it->second;
  return hipSuccess;
}

// The launch's own status (not the thread's last error, which an earlier
// tolerated failure elsewhere may have left).
hipError_t launch_persist(int set, uint32_t grid, size_t lds, const PersistArgs& a, hipStream_t s) {
  void* args[] = {const_cast<PersistArgs*>(&a)};
  return hipLaunchKernel(persist_kernel(set), dim3(grid), dim3(kBT), args, lds, s);
}

}  // namespace batch
}  // namespace gfft
