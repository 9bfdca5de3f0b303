// Axis-pass kernels of the 3D R2C/C2R engine (gfx950).
//
// z pass (contiguous rows).  Two real rows a, b of Nz points are transformed
// together as one complex sequence z = a + i b ("two for one"), then split:
//   A[k] = (Z[k] + conj Z[N-k]) / 2,   B[k] = (Z[k] - conj Z[N-k]) / (2i)
// for k < K = floor(N/2) + 1.  Works for even and odd Nz.  The inverse builds
// Z[n] = A[n] + i B[n] from the Hermitian extension of both half spectra
// (imaginary parts of bin 0 and, for even N, bin N/2 are dropped, the numpy
// irfft convention) and keeps Re -> row a, Im -> row b.  Rows are processed
// in place: a workgroup reads all of its rows before it writes any.
//
// Strided passes (y and x).  A workgroup owns CW adjacent columns (CW * 8 or
// 16 bytes = 128 B per row segment); lanes are column-fastest so each global
// access of a wave covers 64/CW rows x 128 B.  The FFT along the column runs
// in the register engine with LDS exchanges between stages.
#pragma once
#include "gfft_device.h"
#include "gfft_internal.h"

namespace gfft {

template <int N>
struct RowPad {
  // LDS slot of logical point n of one pair: one pad entry every 8 points
  // keeps the stride-R stores of the first stages free of bank conflicts.
  __device__ __forceinline__ static constexpr int idx(int n) { return n + (n >> 3); }
  static constexpr int lrow(int K) { return (N + (N >> 3) + 1) > 2 * K ? (N + (N >> 3) + 1) : 2 * K; }
};

template <typename T, int N>
struct RowLds {
  cx<T>* b;
  __device__ __forceinline__ cx<T>& at(int n) { return b[RowPad<N>::idx(n)]; }
};

template <typename T, int CW>
struct ColLds {
  cx<T>* b;
  int c;
  __device__ __forceinline__ cx<T>& at(int n) { return b[n * CW + c]; }
};

// ---------------------------------------------------------------------------
// z-axis R2C, register engine
// ---------------------------------------------------------------------------
// Rows tiles: thread gi*TPT + t works on row pair gi, so when TPT divides 64
// every LDS exchange of a pair stays inside one wavefront (barrier mode 2).
template <int N, int P, bool LB>
constexpr int rows_barrier() {
  return (64 % (N / P) == 0) ? 2 : (LB ? 1 : 0);
}

template <typename T, typename SC, int P, int G, bool LB = false>
struct RowsR2C {
  static constexpr int BM = rows_barrier<SC::N, P, LB>();
  using E = RegEngine<T, false, SC, P, BM>;
  using C = cx<T>;
  static constexpr int N = SC::N;
  static constexpr int TPT = E::TPT;
  static constexpr int NT = G * TPT;
  static_assert(NT <= 1024, "workgroup larger than 1024 threads");
  static constexpr int KC = N / 2 + 1;  // half-spectrum bins per row (= RowGeom::K)
  static constexpr int LROW = RowPad<N>::lrow(KC);

  // stage-0 inputs of pair `pair` (rows 2*pair and 2*pair + 1): z = a + i b
  __device__ __forceinline__ static void load(C (&v)[E::PV], const T* __restrict__ data, const RowGeom& g,
                                              int64_t pair, int t) {
    const int64_t ra = 2 * pair, rb = ra + 1;
    const bool va = ra < g.nrows, vb = rb < g.nrows;
    const T* A = data + ra * 2 * (int64_t)KC;
    const T* B = A + 2 * (int64_t)KC;
#pragma unroll
    for (int q = 0; q < E::NB0; ++q) {
      if (!E::live0(t, q)) continue;
#pragma unroll
      for (int r = 0; r < E::R0; ++r) {
        const int n = E::in_index(t, q, r);
        v[q * E::R0 + r] = mk<T>(va ? ld_stream(A + n) : (T)0, vb ? ld_stream(B + n) : (T)0);
      }
    }
  }

  // after E::run: Z goes through LDS, is split into the two half spectra and
  // stored over the pair's own rows
  __device__ __forceinline__ static void finish(C (&v)[E::PV], T* __restrict__ data, const RowGeom& g,
                                                int64_t pair, int t, RowLds<T, N>& L) {
    stage_barrier<BM>();
#pragma unroll
    for (int q = 0; q < E::NBL; ++q) {
      if (!E::liveL(t, q)) continue;
#pragma unroll
      for (int r = 0; r < E::RL; ++r) L.at(E::out_index(t, q, r)) = v[q * E::RL + r];
    }
    stage_barrier<BM>();
    const int64_t ra = 2 * pair, rb = ra + 1;
    const bool va = ra < g.nrows, vb = rb < g.nrows;
    C* Ac = reinterpret_cast<C*>(data) + ra * KC;
    C* Bc = Ac + KC;
    const T h = (T)(0.5 * g.scale);
    constexpr int NI = (KC + TPT - 1) / TPT;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int k = t + i * TPT;
      if ((NI * TPT == KC) || k < KC) {
        const C zk = L.at(k), zm = L.at(k == 0 ? 0 : N - k);
        const C xa = mk<T>((zk.x + zm.x) * h, (zk.y - zm.y) * h);
        const C xb = mk<T>((zk.y + zm.y) * h, (zm.x - zk.x) * h);
        if (va) st_stream(Ac + k, xa);
        if (vb) st_stream(Bc + k, xb);
      }
    }
  }

  __device__ __forceinline__ static void tile(T* __restrict__ data, RowGeom g, const C* __restrict__ tw,
                                              int64_t pair0, C* smem) {
    const int gi = threadIdx.x / TPT, t = threadIdx.x % TPT;
    RowLds<T, N> L{smem + gi * LROW};
    C v[E::PV];
    load(v, data, g, pair0 + gi, t);
    E::run(v, t, L, tw);
    finish(v, data, g, pair0 + gi, t, L);
  }
};

// ---------------------------------------------------------------------------
// z-axis C2R, register engine
// ---------------------------------------------------------------------------
template <typename T, typename SC, int P, int G, bool LB = false>
struct RowsC2R {
  static constexpr int BM = rows_barrier<SC::N, P, LB>();
  using E = RegEngine<T, true, SC, P, BM>;
  using C = cx<T>;
  static constexpr int N = SC::N;
  static constexpr int TPT = E::TPT;
  static constexpr int NT = G * TPT;
  static_assert(NT <= 1024, "workgroup larger than 1024 threads");
  static constexpr int KC = N / 2 + 1;
  static constexpr int NI = (KC + TPT - 1) / TPT;  // half-spectrum bins per thread and row
  static constexpr int LROW = RowPad<N>::lrow(KC);

  __device__ __forceinline__ static C hermitian_z(const C* S, int n) {
    const bool lo = 2 * n <= N;
    const int m = lo ? n : N - n;
    C xa = S[m], xb = S[KC + m];
    if (n == 0 || 2 * n == N) {
      xa.y = (T)0;
      xb.y = (T)0;
    }
    // lo: Z = A + iB ; hi: Z = conj(A) + i conj(B)
    return lo ? mk<T>(xa.x - xb.y, xa.y + xb.x) : mk<T>(xa.x + xb.y, xb.x - xa.y);
  }

  // both half spectra of the pair (all loads issued before any use)
  __device__ __forceinline__ static void load(C (&xa)[NI], C (&xb)[NI], const T* __restrict__ data,
                                              const RowGeom& g, int64_t pair, int t) {
    const int64_t ra = 2 * pair, rb = ra + 1;
    const bool va = ra < g.nrows, vb = rb < g.nrows;
    const C* Ac = reinterpret_cast<const C*>(data) + ra * KC;
    const C* Bc = Ac + KC;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int k = t + i * TPT;
      const bool ok = (NI * TPT == KC) || k < KC;
      xa[i] = (va && ok) ? ld_stream(Ac + k) : mk<T>(0, 0);
      xb[i] = (vb && ok) ? ld_stream(Bc + k) : mk<T>(0, 0);
    }
  }

  // LDS staging of the spectra, then the stage-0 inputs of the complex FFT
  __device__ __forceinline__ static void build(C (&v)[E::PV], const C (&xa)[NI], const C (&xb)[NI], C* S, int t) {
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int k = t + i * TPT;
      if ((NI * TPT == KC) || k < KC) {
        S[k] = xa[i];
        S[KC + k] = xb[i];
      }
    }
    stage_barrier<BM>();
#pragma unroll
    for (int q = 0; q < E::NB0; ++q) {
      if (!E::live0(t, q)) continue;
#pragma unroll
      for (int r = 0; r < E::R0; ++r) v[q * E::R0 + r] = hermitian_z(S, E::in_index(t, q, r));
    }
  }

  __device__ __forceinline__ static void store(C (&v)[E::PV], T* __restrict__ data, const RowGeom& g,
                                               int64_t pair, int t) {
    const int64_t ra = 2 * pair, rb = ra + 1;
    const bool va = ra < g.nrows, vb = rb < g.nrows;
    T* A = data + ra * 2 * (int64_t)KC;
    T* B = A + 2 * (int64_t)KC;
    const T sc = (T)g.scale;
#pragma unroll
    for (int q = 0; q < E::NBL; ++q) {
      if (!E::liveL(t, q)) continue;
#pragma unroll
      for (int r = 0; r < E::RL; ++r) {
        const int n = E::out_index(t, q, r);
        C z = v[q * E::RL + r];
        if (g.scale != 1.0) z = mk<T>(z.x * sc, z.y * sc);
        if (va) st_stream(A + n, z.x);
        if (vb) st_stream(B + n, z.y);
      }
    }
  }

  __device__ __forceinline__ static void tile(T* __restrict__ data, RowGeom g, const C* __restrict__ tw,
                                              int64_t pair0, C* smem) {
    const int gi = threadIdx.x / TPT, t = threadIdx.x % TPT;
    C* S = smem + gi * LROW;
    C xa[NI], xb[NI], v[E::PV];
    load(xa, xb, data, g, pair0 + gi, t);
    build(v, xa, xb, S, t);
    RowLds<T, N> L{S};
    E::run(v, t, L, tw);
    // every read of this workgroup's rows happened before the staging barrier
    store(v, data, g, pair0 + gi, t);
  }
};

// ---------------------------------------------------------------------------
// Strided C2C column pass, register engine
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t map_base(const ColMap& m, uint32_t c) {
  return (int64_t)(c / m.inner) * m.outer + (int64_t)(c % m.inner);
}
template <int N>
__device__ __forceinline__ int64_t map_row(const ColMap& m, int n) {
  if (m.nsub >= N) return (int64_t)n * m.rlo;  // uniform branch: single stride
  return (int64_t)(n % m.nsub) * m.rlo + (int64_t)(n / m.nsub) * m.rhi;
}

// LB selects LDS-only stage barriers (persistent, prefetching kernels).
template <typename T, bool INV, typename SC, int P, int CW, bool LB = false>
struct ColsPass {
  using E = RegEngine<T, INV, SC, P, LB>;
  using C = cx<T>;
  static constexpr int N = SC::N;
  static constexpr int TPT = E::TPT;
  static constexpr int NT = CW * TPT;
  static_assert(NT <= 1024, "workgroup larger than 1024 threads");

  // `in` and `out` may be the same array (in-place pass): no __restrict__
  __device__ __forceinline__ static void load(C (&v)[E::PV], const C* in, const ColGeom& g, uint32_t tile,
                                              int c, int t) {
    const uint32_t col = tile * CW + c;
    const bool valid = col < g.ncols;
    const C* p = in + (valid ? map_base(g.in, col) : 0);
#pragma unroll
    for (int q = 0; q < E::NB0; ++q) {
      if (!E::live0(t, q)) continue;
#pragma unroll
      for (int r = 0; r < E::R0; ++r)
        v[q * E::R0 + r] = valid ? ld_stream(p + map_row<N>(g.in, E::in_index(t, q, r))) : mk<T>(0, 0);
    }
  }

  __device__ __forceinline__ static void store(C (&v)[E::PV], C* out, const ColGeom& g, uint32_t tile, int c,
                                               int t) {
    const uint32_t col = tile * CW + c;
    if (col >= g.ncols) return;
    C* p = out + map_base(g.out, col);
    const T sc = (T)g.scale;
#pragma unroll
    for (int q = 0; q < E::NBL; ++q) {
      if (!E::liveL(t, q)) continue;
#pragma unroll
      for (int r = 0; r < E::RL; ++r) {
        C z = v[q * E::RL + r];
        if (g.scale != 1.0) z = mk<T>(z.x * sc, z.y * sc);
        st_stream(p + map_row<N>(g.out, E::out_index(t, q, r)), z);
      }
    }
  }

  // Straight-line forms for tiles whose CW columns all exist: no per-lane
  // guard and the row map's form decided once per tile, so no path skips a
  // load or a store.  The compiler's wait counting is then exact across a
  // persistent tile loop (it waits for the prefetched loads only, not for
  // the previous tile's stores).
  template <bool TWO>
  __device__ __forceinline__ static int64_t row_off(const ColMap& m, int n) {
    return TWO ? (int64_t)(n % m.nsub) * m.rlo + (int64_t)(n / m.nsub) * m.rhi : (int64_t)n * m.rlo;
  }
  template <bool TWO>
  __device__ __forceinline__ static void load_rows(C (&v)[E::PV], const C* p, const ColMap& m, int t) {
#pragma unroll
    for (int q = 0; q < E::NB0; ++q) {
      if (!E::live0(t, q)) continue;
#pragma unroll
      for (int r = 0; r < E::R0; ++r) v[q * E::R0 + r] = ld_stream(p + row_off<TWO>(m, E::in_index(t, q, r)));
    }
  }
  template <bool TWO>
  __device__ __forceinline__ static void store_rows(C (&v)[E::PV], C* p, const ColMap& m, int t, T sc,
                                                    bool scaled) {
#pragma unroll
    for (int q = 0; q < E::NBL; ++q) {
      if (!E::liveL(t, q)) continue;
#pragma unroll
      for (int r = 0; r < E::RL; ++r) {
        C z = v[q * E::RL + r];
        if (scaled) z = mk<T>(z.x * sc, z.y * sc);
        st_stream(p + row_off<TWO>(m, E::out_index(t, q, r)), z);
      }
    }
  }
  // single-stride row maps on both sides (nsub >= N), all columns present
  __device__ __forceinline__ static void load_full(C (&v)[E::PV], const C* in, const ColGeom& g, uint32_t tile,
                                                   int c, int t) {
    load_rows<false>(v, in + map_base(g.in, tile * CW + c), g.in, t);
  }
  __device__ __forceinline__ static void store_full(C (&v)[E::PV], C* out, const ColGeom& g, uint32_t tile,
                                                    int c, int t) {
    store_rows<false>(v, out + map_base(g.out, tile * CW + c), g.out, t, (T)g.scale, g.scale != 1.0);
  }

  __device__ __forceinline__ static void tile(const C* in, C* out, const ColGeom& g, const C* __restrict__ tw,
                                              uint32_t tile_id, C* smem) {
    const int c = threadIdx.x % CW, t = threadIdx.x / CW;
    C v[E::PV];
    load(v, in, g, tile_id, c, t);
    ColLds<T, CW> L{smem, c};
    E::run(v, t, L, tw);
    store(v, out, g, tile_id, c, t);
  }
};

// ---------------------------------------------------------------------------
// __global__ wrappers (one tile per workgroup, XCD-aware tile order)
// ---------------------------------------------------------------------------
template <typename T, typename SC, int P, int G>
__global__ __attribute__((amdgpu_flat_work_group_size(1, RowsR2C<T, SC, P, G>::NT))) void k_rows_r2c(T* data, RowGeom g, const cx<T>* tw) {
  using K_ = RowsR2C<T, SC, P, G>;
  __shared__ cx<T> smem[G * K_::LROW];
  const uint32_t t0 = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t tile = g.rev ? gridDim.x - 1 - t0 : t0;
  K_::tile(data, g, tw, (int64_t)tile * G, smem);
}

template <typename T, typename SC, int P, int G>
__global__ __attribute__((amdgpu_flat_work_group_size(1, RowsC2R<T, SC, P, G>::NT))) void k_rows_c2r(T* data, RowGeom g, const cx<T>* tw) {
  using K_ = RowsC2R<T, SC, P, G>;
  __shared__ cx<T> smem[G * K_::LROW];
  const uint32_t t0 = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t tile = g.rev ? gridDim.x - 1 - t0 : t0;
  K_::tile(data, g, tw, (int64_t)tile * G, smem);
}

// ORD 0: XCD-aware tile order (each XCD walks a contiguous range, so tiles
// that share straddled cache lines meet in one L2).  ORD 2: tile = workgroup
// id, so the tiles in flight form one contiguous range across all XCDs (for
// line-aligned tiles: the rows they read are then read densely).
template <typename T, bool INV, typename SC, int P, int CW, int ORD = 0>
__global__ __attribute__((amdgpu_flat_work_group_size(1, ColsPass<T, INV, SC, P, CW>::NT))) void k_cols(
    const cx<T>* in, cx<T>* out, ColGeom g, const cx<T>* tw) {
  using K_ = ColsPass<T, INV, SC, P, CW>;
  __shared__ cx<T> smem[SC::N * CW];
  const uint32_t t0 = ORD == 2 ? blockIdx.x : xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t tile = g.rev ? gridDim.x - 1 - t0 : t0;
  K_::tile(in, out, g, tw, tile, smem);
}

// Persistent, software-pipelined column pass.  Workgroup w owns the
// contiguous tile range [w*T/G, (w+1)*T/G) (XCD-aware order, so one XCD walks
// neighbouring tiles); while tile i runs its LDS stages, the global loads of
// tile i+1 are already in flight in a second register set.  Stage barriers
// wait for LDS traffic only, so the prefetch is not drained by them.
template <typename T, bool INV, typename SC, int P, int CW>
using ColsPipe = ColsPass<T, INV, SC, P, CW, true>;

// Tiles of a persistent workgroup: ORD 0 = the contiguous range
// [w*T/G, (w+1)*T/G); ORD 1 = round robin w, w+G, w+2G, ... (neighbouring
// tiles run concurrently on neighbouring workgroups of one XCD).
// With rev, tile t of the walk is tile ntiles-1-t: the pass starts with the
// tiles the previous pass wrote last (still in the memory-side cache).
struct TileWalk {
  uint32_t first, count, step, last;
  int32_t rev;
  __device__ __forceinline__ uint32_t at(uint32_t k) const {
    const uint32_t t = first + k * step;
    return rev ? last - t : t;
  }
};
template <int ORD>
__device__ __forceinline__ TileWalk tile_walk(uint32_t ntiles, int32_t rev) {
  const uint32_t G = gridDim.x;
  const uint32_t w = xcd_tile(blockIdx.x, G);
  TileWalk tw;
  tw.last = ntiles - 1;
  tw.rev = rev;
  if (ORD == 0) {
    tw.first = (uint32_t)(((uint64_t)w * ntiles) / G);
    tw.count = (uint32_t)(((uint64_t)(w + 1) * ntiles) / G) - tw.first;
    tw.step = 1;
  } else {
    tw.first = w;
    tw.count = w < ntiles ? (ntiles - w + G - 1) / G : 0;
    tw.step = G;
  }
  return tw;
}

// Twiddles of a persistent kernel live in LDS: a global twiddle load would
// make the in-order vector-memory counter wait for the prefetched tile too.
template <typename SC, typename C>
__device__ __forceinline__ void copy_twiddles(C* stw, const C* __restrict__ tw, int nthreads) {
  for (int i = threadIdx.x; i < SC::twoff(SC::S); i += nthreads) stw[i] = tw[i];
}
template <typename SC>
constexpr int twiddle_count() {
  return SC::twoff(SC::S) > 0 ? SC::twoff(SC::S) : 1;
}

// RA 1: the thread's column and row indexes are made opaque to the compiler
// at the top of every tile, so the per-element global and LDS offsets
// derived from them are recomputed per tile.  Otherwise (RA 0) they are
// hoisted out of the tile loop and held in registers for the whole kernel:
// cheaper for small tiles, but for 768/1024-point columns at 16-24 points
// per thread that exceeds the register budget (spills, or half the waves).
//
// FULL: every tile has CW existing columns (ncols % CW == 0) and both row
// maps have a single stride.  The tile loop then runs straight-line loads
// and stores, and its last iteration (no prefetch) is peeled off, so the
// wait before a tile's first use of its prefetched rows does not also wait
// for the previous tile's stores.
template <typename T, bool INV, typename SC, int P, int CW, int ORD, int RA = 0, bool FULL = false>
__global__ __attribute__((amdgpu_flat_work_group_size(1, ColsPipe<T, INV, SC, P, CW>::NT))) void k_cols_pipe(
    const cx<T>* in, cx<T>* out, ColGeom g, const cx<T>* tw, uint32_t ntiles) {
  using K_ = ColsPipe<T, INV, SC, P, CW>;
  __shared__ cx<T> smem[SC::N * CW];
  __shared__ cx<T> stw[twiddle_count<SC>()];
  const TileWalk walk = tile_walk<ORD>(ntiles, g.rev);
  if (walk.count == 0) return;  // uniform across the workgroup
  copy_twiddles<SC>(stw, tw, K_::NT);
  const int c0 = threadIdx.x % CW, t0 = threadIdx.x / CW;
  cx<T> cur[K_::E::PV], nxt[K_::E::PV];
  if constexpr (FULL) {
    K_::load_full(nxt, in, g, walk.at(0), c0, t0);
    for (uint32_t k = 0; k + 1 < walk.count; ++k) {
      int c = c0, t = t0;
      if constexpr (RA != 0) __asm__ volatile("" : "+v"(c), "+v"(t));
      ColLds<T, CW> L{smem, c};
#pragma unroll
      for (int i = 0; i < K_::E::PV; ++i) cur[i] = nxt[i];
      K_::load_full(nxt, in, g, walk.at(k + 1), c, t);
      K_::E::run(cur, t, L, stw);
      K_::store_full(cur, out, g, walk.at(k), c, t);
    }
    ColLds<T, CW> L{smem, c0};
    K_::E::run(nxt, t0, L, stw);
    K_::store_full(nxt, out, g, walk.at(walk.count - 1), c0, t0);
    return;
  }
  K_::load(nxt, in, g, walk.at(0), c0, t0);
  for (uint32_t k = 0; k < walk.count; ++k) {
    int c = c0, t = t0;
    if constexpr (RA != 0) __asm__ volatile("" : "+v"(c), "+v"(t));
    ColLds<T, CW> L{smem, c};
#pragma unroll
    for (int i = 0; i < K_::E::PV; ++i) cur[i] = nxt[i];
    if (k + 1 < walk.count) K_::load(nxt, in, g, walk.at(k + 1), c, t);
    // the first stage uses no twiddles; two barriers precede the first
    // twiddle read, so the LDS copy above is complete by then
    K_::E::run(cur, t, L, stw);
    K_::store(cur, out, g, walk.at(k), c, t);
  }
}

// Persistent, software-pipelined z passes (tile = G row pairs).
template <typename T, typename SC, int P, int G, int ORD>
__global__ __attribute__((amdgpu_flat_work_group_size(1, RowsR2C<T, SC, P, G, true>::NT))) void k_rows_r2c_pipe(
    T* data, RowGeom g, const cx<T>* tw, uint32_t ntiles) {
  using K_ = RowsR2C<T, SC, P, G, true>;
  __shared__ cx<T> smem[G * K_::LROW];
  __shared__ cx<T> stw[twiddle_count<SC>()];
  const TileWalk walk = tile_walk<ORD>(ntiles, g.rev);
  if (walk.count == 0) return;
  copy_twiddles<SC>(stw, tw, K_::NT);
  __syncthreads();  // the table is complete before any wave's first stage
  const int gi = threadIdx.x / K_::TPT, t = threadIdx.x % K_::TPT;
  RowLds<T, SC::N> L{smem + gi * K_::LROW};
  cx<T> cur[K_::E::PV], nxt[K_::E::PV];
  K_::load(nxt, data, g, (int64_t)walk.at(0) * G + gi, t);
  for (uint32_t k = 0; k < walk.count; ++k) {
#pragma unroll
    for (int i = 0; i < K_::E::PV; ++i) cur[i] = nxt[i];
    if (k + 1 < walk.count) K_::load(nxt, data, g, (int64_t)walk.at(k + 1) * G + gi, t);
    K_::E::run(cur, t, L, stw);
    K_::finish(cur, data, g, (int64_t)walk.at(k) * G + gi, t, L);
  }
}

template <typename T, typename SC, int P, int G, int ORD>
__global__ __attribute__((amdgpu_flat_work_group_size(1, RowsC2R<T, SC, P, G, true>::NT))) void k_rows_c2r_pipe(
    T* data, RowGeom g, const cx<T>* tw, uint32_t ntiles) {
  using K_ = RowsC2R<T, SC, P, G, true>;
  __shared__ cx<T> smem[G * K_::LROW];
  __shared__ cx<T> stw[twiddle_count<SC>()];
  const TileWalk walk = tile_walk<ORD>(ntiles, g.rev);
  if (walk.count == 0) return;
  copy_twiddles<SC>(stw, tw, K_::NT);
  __syncthreads();  // the table is complete before any wave's first stage
  const int gi = threadIdx.x / K_::TPT, t = threadIdx.x % K_::TPT;
  cx<T>* S = smem + gi * K_::LROW;
  RowLds<T, SC::N> L{S};
  cx<T> ca[K_::NI], cb[K_::NI], na[K_::NI], nb[K_::NI], v[K_::E::PV];
  K_::load(na, nb, data, g, (int64_t)walk.at(0) * G + gi, t);
  for (uint32_t k = 0; k < walk.count; ++k) {
#pragma unroll
    for (int i = 0; i < K_::NI; ++i) {
      ca[i] = na[i];
      cb[i] = nb[i];
    }
    if (k + 1 < walk.count) K_::load(na, nb, data, g, (int64_t)walk.at(k + 1) * G + gi, t);
    stage_barrier<K_::BM>();  // the previous tile's stage reads of S are done
    K_::build(v, ca, cb, S, t);
    K_::E::run(v, t, L, stw);
    K_::store(v, data, g, (int64_t)walk.at(k) * G + gi, t);
  }
}

int device_cu_count();

// Host launchers with the registry's signatures.  Each clears the thread's
// last HIP error before its launch, so the hipGetLastError() read after it
// reports that launch only, not a failure an earlier call left behind (the
// occupancy queries tolerate theirs, and so may the caller's own code).
inline void clear_launch_error() { (void)hipGetLastError(); }

template <typename T, typename SC, int P, int G>
hipError_t launch_rows_r2c(void* data, RowGeom g, const void* tw, hipStream_t s) {
  const int64_t pairs = (g.nrows + 1) / 2;
  const int64_t blocks = (pairs + G - 1) / G;
  if (blocks <= 0) return hipSuccess;
  clear_launch_error();
  hipLaunchKernelGGL((k_rows_r2c<T, SC, P, G>), dim3((uint32_t)blocks), dim3(RowsR2C<T, SC, P, G>::NT), 0, s,
                     static_cast<T*>(data), g, static_cast<const cx<T>*>(tw));
  return hipGetLastError();
}

template <typename T, typename SC, int P, int G>
hipError_t launch_rows_c2r(void* data, RowGeom g, const void* tw, hipStream_t s) {
  const int64_t pairs = (g.nrows + 1) / 2;
  const int64_t blocks = (pairs + G - 1) / G;
  if (blocks <= 0) return hipSuccess;
  clear_launch_error();
  hipLaunchKernelGGL((k_rows_c2r<T, SC, P, G>), dim3((uint32_t)blocks), dim3(RowsC2R<T, SC, P, G>::NT), 0, s,
                     static_cast<T*>(data), g, static_cast<const cx<T>*>(tw));
  return hipGetLastError();
}

template <typename T, bool INV, typename SC, int P, int CW, int ORD = 0>
hipError_t launch_cols(const void* in, void* out, ColGeom g, const void* tw, hipStream_t s) {
  const int64_t blocks = ((int64_t)g.ncols + CW - 1) / CW;
  if (blocks <= 0) return hipSuccess;
  clear_launch_error();
  hipLaunchKernelGGL((k_cols<T, INV, SC, P, CW, ORD>), dim3((uint32_t)blocks), dim3(ColsPass<T, INV, SC, P, CW>::NT),
                     0, s, static_cast<const cx<T>*>(in), static_cast<cx<T>*>(out), g, static_cast<const cx<T>*>(tw));
  return hipGetLastError();
}

inline int64_t persistent_blocks(int64_t ntiles, int occ) {
  int64_t blocks = (int64_t)device_cu_count() * occ;
  return blocks < ntiles ? blocks : ntiles;
}

// FL 1: the straight-line FULL kernel where the geometry allows it (else the
// general one); FL 0: always the general kernel.
template <typename T, bool INV, typename SC, int P, int CW, int OCC, int ORD, int RA = 0, int FL = 0>
hipError_t launch_cols_pipe(const void* in, void* out, ColGeom g, const void* tw, hipStream_t s) {
  const int64_t ntiles = ((int64_t)g.ncols + CW - 1) / CW;
  if (ntiles <= 0) return hipSuccess;
  const dim3 grid((uint32_t)persistent_blocks(ntiles, OCC)), block(ColsPipe<T, INV, SC, P, CW>::NT);
  clear_launch_error();
  if (FL && g.ncols % CW == 0 && g.in.nsub >= SC::N && g.out.nsub >= SC::N)
    hipLaunchKernelGGL((k_cols_pipe<T, INV, SC, P, CW, ORD, RA, true>), grid, block, 0, s,
                       static_cast<const cx<T>*>(in), static_cast<cx<T>*>(out), g, static_cast<const cx<T>*>(tw),
                       (uint32_t)ntiles);
  else
    hipLaunchKernelGGL((k_cols_pipe<T, INV, SC, P, CW, ORD, RA, false>), grid, block, 0, s,
                       static_cast<const cx<T>*>(in), static_cast<cx<T>*>(out), g, static_cast<const cx<T>*>(tw),
                       (uint32_t)ntiles);
  return hipGetLastError();
}

template <typename T, typename SC, int P, int G, int OCC, int ORD>
hipError_t launch_rows_r2c_pipe(void* data, RowGeom g, const void* tw, hipStream_t s) {
  const int64_t ntiles = ((g.nrows + 1) / 2 + G - 1) / G;
  if (ntiles <= 0) return hipSuccess;
  clear_launch_error();
  hipLaunchKernelGGL((k_rows_r2c_pipe<T, SC, P, G, ORD>), dim3((uint32_t)persistent_blocks(ntiles, OCC)),
                     dim3(RowsR2C<T, SC, P, G, true>::NT), 0, s, static_cast<T*>(data), g,
                     static_cast<const cx<T>*>(tw), (uint32_t)ntiles);
  return hipGetLastError();
}

template <typename T, typename SC, int P, int G, int OCC, int ORD>
hipError_t launch_rows_c2r_pipe(void* data, RowGeom g, const void* tw, hipStream_t s) {
  const int64_t ntiles = ((g.nrows + 1) / 2 + G - 1) / G;
  if (ntiles <= 0) return hipSuccess;
  clear_launch_error();
  hipLaunchKernelGGL((k_rows_c2r_pipe<T, SC, P, G, ORD>), dim3((uint32_t)persistent_blocks(ntiles, OCC)),
                     dim3(RowsC2R<T, SC, P, G, true>::NT), 0, s, static_cast<T*>(data), g,
                     static_cast<const cx<T>*>(tw), (uint32_t)ntiles);
  return hipGetLastError();
}

// Registry entry builders
template <typename SC>
constexpr Schedule make_sched() {
  Schedule s{};
  s.N = SC::N;
  s.S = SC::S;
  for (int i = 0; i < SC::S; ++i) s.R[i] = SC::R[i];
  return s;
}

template <typename T>
constexpr int prec_of() {
  return sizeof(T) == 4 ? F32 : F64;
}

template <typename T, typename SC, int P, int G>
KernelEntry entry_rows_r2c() {
  return KernelEntry{prec_of<T>(), ROWS_R2C, SC::N, make_sched<SC>(), P, G, RowsR2C<T, SC, P, G>::NT,
                     reinterpret_cast<void*>(&launch_rows_r2c<T, SC, P, G>)};
}
template <typename T, typename SC, int P, int G>
KernelEntry entry_rows_c2r() {
  return KernelEntry{prec_of<T>(), ROWS_C2R, SC::N, make_sched<SC>(), P, G, RowsC2R<T, SC, P, G>::NT,
                     reinterpret_cast<void*>(&launch_rows_c2r<T, SC, P, G>)};
}
template <typename T, bool INV, typename SC, int P, int CW, int ORD = 0>
KernelEntry entry_cols() {
  return KernelEntry{prec_of<T>(), INV ? COLS_INV : COLS_FWD, SC::N, make_sched<SC>(), P, CW,
                     ColsPass<T, INV, SC, P, CW>::NT,
                     reinterpret_cast<void*>(&launch_cols<T, INV, SC, P, CW, ORD>), 0, ORD};
}

template <typename T, bool INV, typename SC, int P, int CW, int OCC, int ORD, int RA = 0, int FL = 0>
KernelEntry entry_cols_pipe() {
  return KernelEntry{prec_of<T>(),
                     INV ? COLS_INV : COLS_FWD,
                     SC::N,
                     make_sched<SC>(),
                     P,
                     CW,
                     ColsPipe<T, INV, SC, P, CW>::NT,
                     reinterpret_cast<void*>(&launch_cols_pipe<T, INV, SC, P, CW, OCC, ORD, RA, FL>),
                     OCC,
                     ORD,
                     RA,
                     FL};
}

template <typename T, typename SC, int P, int G, int OCC, int ORD>
KernelEntry entry_rows_r2c_pipe() {
  return KernelEntry{prec_of<T>(),
                     ROWS_R2C,
                     SC::N,
                     make_sched<SC>(),
                     P,
                     G,
                     RowsR2C<T, SC, P, G, true>::NT,
                     reinterpret_cast<void*>(&launch_rows_r2c_pipe<T, SC, P, G, OCC, ORD>),
                     OCC,
                     ORD};
}

template <typename T, typename SC, int P, int G, int OCC, int ORD>
KernelEntry entry_rows_c2r_pipe() {
  return KernelEntry{prec_of<T>(),
                     ROWS_C2R,
                     SC::N,
                     make_sched<SC>(),
                     P,
                     G,
                     RowsC2R<T, SC, P, G, true>::NT,
                     reinterpret_cast<void*>(&launch_rows_c2r_pipe<T, SC, P, G, OCC, ORD>),
                     OCC,
                     ORD};
}

}  // namespace gfft
