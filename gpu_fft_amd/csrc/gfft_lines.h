// Strided C2C pass with line-aligned global accesses (gfx950).
//
// The in-place half-spectrum rows of the y pass are K = Nz/2 + 1 complex
// long (2056 B at Nz = 512 FP32), so the 128 B segment that a CW-column tile
// reads from each row starts inside a cache line and touches two lines; the
// neighbouring tile touches the other half of each.  Here a workgroup walks
// the tiles of a plane in order and reads each row's memory as aligned
// CW-element chunks (one 128 B line per row and chunk), keeping the part of
// a chunk that belongs to the next tile in registers:
//
//   row r of plane x starts sigma = (row address / element size) mod CW
//   elements into a chunk, so chunk m of the row covers the columns
//   kz in [CW*m - sigma, CW*m - sigma + CW).  Lane l of a CW-lane row group
//   holds position l of a chunk, i.e. column kz = CW*m + l - sigma: of tile m
//   when l >= sigma, of tile m - 1 when l < sigma.  In both cases its column
//   inside the tile is c = (l - sigma) mod CW.
//
// A thread's rows differ by multiples of CW (TPT and N/R are multiples of
// CW) and K is the same for every row, so sigma is one value per thread and
// plane, and the thread simply takes column c of the tile: no lane exchange
// is needed, only a per-lane choice between two chunks.  Tile j's inputs are
// chunk j (lanes l >= sigma) and chunk j + 1 (l < sigma); walking forward,
// chunk j + 1 is loaded while tile j - 1 runs and kept for tile j + 1.
// Stores mirror this: a whole chunk j is written once tile j is done, its
// first sigma positions from tile j - 1's outputs kept in registers.  A
// walk that starts or ends inside a plane (the workgroup's range boundary)
// writes only its own positions of the boundary chunk; the neighbouring
// workgroup writes the rest.  Positions outside [0, K) of a row (the
// neighbouring rows' columns in the first and last chunk) are masked with
// the buffer descriptor's range check (offset >= the plane's byte size:
// loads return zero, stores are dropped), so every access is branch-free.
#pragma once
#include "gfft_kernels.h"

namespace gfft {

typedef int gfft_v2i __attribute__((ext_vector_type(2)));
typedef int gfft_v4i __attribute__((ext_vector_type(4)));

// descriptor flags of a raw buffer on gfx9-family parts (stride 0,
// 32-bit data format, no swizzle)
constexpr int kBufFlags = 0x00020000;
// offset that fails the range check of every descriptor used here
constexpr uint32_t kOutOfRange = 0x80000000u;

template <typename T>
struct BufIO;
template <>
struct BufIO<float> {
  __device__ __forceinline__ static cx<float> ld(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(cx<float>, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
  }
  __device__ __forceinline__ static void st(__amdgpu_buffer_rsrc_t r, uint32_t off, cx<float> v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(gfft_v2i, v), r, (int)off, 0, 0);
  }
};
template <>
struct BufIO<double> {
  __device__ __forceinline__ static cx<double> ld(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(cx<double>, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
  }
  __device__ __forceinline__ static void st(__amdgpu_buffer_rsrc_t r, uint32_t off, cx<double> v) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(gfft_v4i, v), r, (int)off, 0, 0);
  }
};

// Geometry of a line-aligned pass (host-computed from a single-stride,
// in-place ColGeom): column c = (plane c / inner, kz = c % inner); point n
// of a column sits at plane * outer + n * rlo + kz.
struct LinesGeom {
  uint32_t ntiles;       // planes * tpp
  uint32_t tpp;          // tiles per plane, ceil(inner / CW)
  uint32_t inner;        // columns per plane
  uint32_t plane_bytes;  // ((N - 1) * rlo + inner) * element size, < 2^31
  uint32_t rlo_bytes;    // rlo * element size
  uint32_t rlo_mod;      // rlo mod CW
  uint32_t outer_mod;    // outer mod CW
  uint32_t base_mod;     // (data address / element size) mod CW
  int64_t outer;         // elements between planes
  double scale;          // applied to the outputs
  int32_t rev;           // 1: tiles run last to first
};

template <typename T, bool INV, typename SC, int P, int CW>
struct ColsLines {
  using E = RegEngine<T, INV, SC, P, 1>;
  using C = cx<T>;
  static constexpr int N = SC::N;
  static constexpr int TPT = E::TPT;
  static constexpr int NT = CW * TPT;
  static_assert(NT <= 1024, "workgroup larger than 1024 threads");
  static_assert((CW & (CW - 1)) == 0 && CW * (int)sizeof(C) == 128, "one chunk = one 128 B line");
  static_assert(TPT % CW == 0 && (N / E::R0) % CW == 0 && (N / E::RL) % CW == 0,
                "a thread's rows must be congruent mod CW");
  static_assert(E::EXACT(0) && E::EXACT(SC::S - 1), "every first / last stage slot must be live");

  // byte offset of position l of chunk m in a row (relative to the row
  // start), or an offset that fails the range check when the position
  // holds another row's column (kz outside [0, inner))
  __device__ __forceinline__ static uint32_t bias(const LinesGeom& g, int64_t m, int l, int sigma) {
    const int64_t kz = (int64_t)CW * m + l - sigma;
    return (kz >= 0 && kz < (int64_t)g.inner) ? (uint32_t)kz * (uint32_t)sizeof(C) : kOutOfRange;
  }

  __device__ __forceinline__ static void load_chunk(C (&v)[E::PV], __amdgpu_buffer_rsrc_t r, uint32_t b,
                                                    const LinesGeom& g, int t) {
#pragma unroll
    for (int q = 0; q < E::NB0; ++q)
#pragma unroll
      for (int k = 0; k < E::R0; ++k)
        v[q * E::R0 + k] = BufIO<T>::ld(r, (uint32_t)E::in_index(t, q, k) * g.rlo_bytes + b);
  }

  // store position l of a chunk: `sel` picks a (the current tile) or b
  __device__ __forceinline__ static void store_chunk(const C (&a)[E::PV], const C (&b)[E::PV], bool sel,
                                                     __amdgpu_buffer_rsrc_t r, uint32_t bo, const LinesGeom& g,
                                                     int t) {
    const T sc = (T)g.scale;
    const bool scaled = g.scale != 1.0;
#pragma unroll
    for (int q = 0; q < E::NBL; ++q)
#pragma unroll
      for (int k = 0; k < E::RL; ++k) {
        C z = sel ? a[q * E::RL + k] : b[q * E::RL + k];
        if (scaled) z = mk<T>(z.x * sc, z.y * sc);
        BufIO<T>::st(r, (uint32_t)E::out_index(t, q, k) * g.rlo_bytes + bo, z);
      }
  }

  __device__ __forceinline__ static __amdgpu_buffer_rsrc_t plane_rsrc(C* data, const LinesGeom& g, uint32_t plane) {
    return __builtin_amdgcn_make_buffer_rsrc(data + (int64_t)plane * g.outer, (short)0, (int)g.plane_bytes, kBufFlags);
  }
  // chunk offset of the thread's rows in a plane
  __device__ __forceinline__ static int sigma_of(const LinesGeom& g, uint32_t plane, int t) {
    return (int)((g.base_mod + (uint64_t)plane * g.outer_mod + (uint64_t)t * g.rlo_mod) & (CW - 1));
  }
};

// Persistent line-aligned column pass: workgroup w walks the contiguous tile
// range [w*T/G, (w+1)*T/G) (forward, or backward with g.rev), tile = (plane,
// CW columns); the loads of the next chunk are in flight while a tile runs
// its LDS stages (LDS-only barriers).
template <typename T, bool INV, typename SC, int P, int CW>
__global__ __attribute__((amdgpu_flat_work_group_size(1, ColsLines<T, INV, SC, P, CW>::NT))) void k_cols_lines(
    cx<T>* data, LinesGeom g, const cx<T>* tw) {
  using K_ = ColsLines<T, INV, SC, P, CW>;
  using E = typename K_::E;
  using C = cx<T>;
  __shared__ C smem[SC::N * CW];
  __shared__ C stw[twiddle_count<SC>()];
  const TileWalk walk = tile_walk<0>(g.ntiles, g.rev);
  if (walk.count == 0) return;  // uniform across the workgroup
  copy_twiddles<SC>(stw, tw, K_::NT);
  const int l = threadIdx.x % CW, t = threadIdx.x / CW;
  const bool fwd = g.rev == 0;

  uint32_t tile = walk.at(0);
  uint32_t plane = tile / g.tpp;
  int64_t j = tile % g.tpp;
  int sigma = K_::sigma_of(g, plane, t);
  __amdgpu_buffer_rsrc_t rs = K_::plane_rsrc(data, g, plane);
  C older[E::PV], newer[E::PV], carry[E::PV];
  // forward: older = chunk j, newer = chunk j + 1; backward: the reverse
  K_::load_chunk(older, rs, K_::bias(g, fwd ? j : j + 1, l, sigma), g, t);
  K_::load_chunk(newer, rs, K_::bias(g, fwd ? j + 1 : j, l, sigma), g, t);
  bool have_prev = false;
  for (uint32_t k = 0; k < walk.count; ++k) {
    C cur[E::PV];
    const bool take_newer = (l < sigma) == fwd;
#pragma unroll
    for (int i = 0; i < E::PV; ++i) cur[i] = take_newer ? newer[i] : older[i];
    const int col = (l - sigma) & (CW - 1);

    // chunks of the next tile: one new chunk inside a plane, two at a
    // plane change
    bool next_same = false;
    uint32_t plane2 = plane;
    int64_t j2 = j;
    int sigma2 = sigma;
    __amdgpu_buffer_rsrc_t rs2 = rs;
    if (k + 1 < walk.count) {
      const uint32_t tile2 = walk.at(k + 1);
      plane2 = tile2 / g.tpp;
      j2 = tile2 % g.tpp;
      next_same = plane2 == plane;
      if (next_same) {
#pragma unroll
        for (int i = 0; i < E::PV; ++i) older[i] = newer[i];
        K_::load_chunk(newer, rs, K_::bias(g, fwd ? j2 + 1 : j2, l, sigma), g, t);
      } else {
        sigma2 = K_::sigma_of(g, plane2, t);
        rs2 = K_::plane_rsrc(data, g, plane2);
        K_::load_chunk(older, rs2, K_::bias(g, fwd ? j2 : j2 + 1, l, sigma2), g, t);
        K_::load_chunk(newer, rs2, K_::bias(g, fwd ? j2 + 1 : j2, l, sigma2), g, t);
      }
    }

    // the first stage uses no twiddles; two barriers precede the first
    // twiddle read, so the LDS copy above is complete by then
    ColLds<T, CW> L{smem, col};
    E::run(cur, t, L, stw);

    // the whole chunk this tile completes: forward chunk j (positions
    // l < sigma from the previous tile), backward chunk j + 1 (positions
    // l >= sigma from the previous tile)
    {
      const bool mine = (l >= sigma) == fwd;
      const uint32_t b = (mine || have_prev) ? K_::bias(g, fwd ? j : j + 1, l, sigma) : kOutOfRange;
      K_::store_chunk(cur, carry, mine, rs, b, g, t);
    }
    // at the end of a run in this plane, this tile's positions of the next
    // chunk in walk order (forward chunk j + 1 at l < sigma, backward chunk j
    // at l >= sigma)
    if (!next_same) {
      const bool mine = (l < sigma) == fwd;
      const uint32_t b = mine ? K_::bias(g, fwd ? j + 1 : j, l, sigma) : kOutOfRange;
      K_::store_chunk(cur, cur, true, rs, b, g, t);
    }
#pragma unroll
    for (int i = 0; i < E::PV; ++i) carry[i] = cur[i];
    have_prev = next_same;
    plane = plane2;
    j = j2;
    sigma = sigma2;
    rs = rs2;
  }
}

// Eligible geometries: in place, one single-stride map for both sides, whole
// planes, a plane below 2 GiB.  Others run the persistent column kernel.
template <typename T, int N, int CW>
bool lines_geom(const void* in, const void* out, const ColGeom& g, LinesGeom* lg) {
  const int64_t es = (int64_t)sizeof(cx<T>);
  if (in != out || g.in.nsub < N || g.out.nsub < N || g.in.inner != g.out.inner || g.in.outer != g.out.outer ||
      g.in.rlo != g.out.rlo || g.in.inner == 0 || g.ncols % g.in.inner != 0 || g.in.rlo < (int64_t)g.in.inner)
    return false;
  if ((reinterpret_cast<uintptr_t>(in) % es) != 0) return false;
  const int64_t planes = g.ncols / g.in.inner;
  if (planes > 1 && g.in.outer < (int64_t)(N - 1) * g.in.rlo + g.in.inner) return false;
  const int64_t pb = ((int64_t)(N - 1) * g.in.rlo + g.in.inner) * es;
  const int64_t tpp = ((int64_t)g.in.inner + CW - 1) / CW;
  if (pb >= (int64_t)kOutOfRange || (int64_t)g.in.rlo * es >= (int64_t)kOutOfRange || planes * tpp >= (1ll << 32))
    return false;
  lg->ntiles = (uint32_t)(planes * tpp);
  lg->tpp = (uint32_t)tpp;
  lg->inner = g.in.inner;
  lg->plane_bytes = (uint32_t)pb;
  lg->rlo_bytes = (uint32_t)(g.in.rlo * es);
  lg->rlo_mod = (uint32_t)(g.in.rlo % CW);
  lg->outer_mod = (uint32_t)(g.in.outer % CW);
  lg->base_mod = (uint32_t)((reinterpret_cast<uintptr_t>(in) / es) % CW);
  lg->outer = g.in.outer;
  lg->scale = g.scale;
  lg->rev = g.rev;
  return true;
}

template <typename T, bool INV, typename SC, int P, int CW, int OCC>
hipError_t launch_cols_lines(const void* in, void* out, ColGeom g, const void* tw, hipStream_t s) {
  LinesGeom lg;
  if (!lines_geom<T, SC::N, CW>(in, out, g, &lg))
    return launch_cols_pipe<T, INV, SC, P, CW, OCC, 0>(in, out, g, tw, s);
  if (lg.ntiles == 0) return hipSuccess;
  hipLaunchKernelGGL((k_cols_lines<T, INV, SC, P, CW>), dim3((uint32_t)persistent_blocks(lg.ntiles, OCC)),
                     dim3(ColsLines<T, INV, SC, P, CW>::NT), 0, s, static_cast<cx<T>*>(out), lg,
                     static_cast<const cx<T>*>(tw));
  return hipGetLastError();
}

// straight = 2 marks the line-aligned kernel in the registry
template <typename T, bool INV, typename SC, int P, int CW, int OCC>
KernelEntry entry_cols_lines() {
  return KernelEntry{prec_of<T>(),
                     INV ? COLS_INV : COLS_FWD,
                     SC::N,
                     make_sched<SC>(),
                     P,
                     CW,
                     ColsLines<T, INV, SC, P, CW>::NT,
                     reinterpret_cast<void*>(&launch_cols_lines<T, INV, SC, P, CW, OCC>),
                     OCC,
                     0,
                     0,
                     2};
}

#define GFFT_COLS_LINES(T, SC, PC, CW, OCC) \
  entry_cols_lines<T, false, SC, PC, CW, OCC>(), entry_cols_lines<T, true, SC, PC, CW, OCC>()

}  // namespace gfft
