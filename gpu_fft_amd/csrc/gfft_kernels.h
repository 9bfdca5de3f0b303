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
template <typename T, typename SC, int P, int G>
struct RowsR2C {
  using E = RegEngine<T, false, SC, P>;
  using C = cx<T>;
  static constexpr int N = SC::N;
  static constexpr int TPT = E::TPT;
  static constexpr int NT = G * TPT;
  static constexpr int LROW = RowPad<N>::lrow(N / 2 + 1);

  __device__ __forceinline__ static void tile(T* __restrict__ data, RowGeom g, const C* __restrict__ tw,
                                              int64_t pair0, C* smem) {
    const int gi = threadIdx.x / TPT, t = threadIdx.x % TPT;
    const int K = g.K;
    const int64_t ra = 2 * (pair0 + gi), rb = ra + 1;
    const bool va = ra < g.nrows, vb = rb < g.nrows;
    const T* A = data + ra * 2 * (int64_t)K;
    const T* B = A + 2 * (int64_t)K;
    C v[P];
#pragma unroll
    for (int q = 0; q < P / E::R0; ++q)
#pragma unroll
      for (int r = 0; r < E::R0; ++r) {
        const int n = E::in_index(t, q, r);
        v[q * E::R0 + r] = mk<T>(va ? A[n] : (T)0, vb ? B[n] : (T)0);
      }
    RowLds<T, N> L{smem + gi * LROW};
    E::run(v, t, L, tw);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < P / E::RL; ++q)
#pragma unroll
      for (int r = 0; r < E::RL; ++r) L.at(E::out_index(t, q, r)) = v[q * E::RL + r];
    __syncthreads();
    C* Ac = reinterpret_cast<C*>(data) + ra * K;
    C* Bc = Ac + K;
    const T h = (T)(0.5 * g.scale);
    for (int k = t; k < K; k += TPT) {
      const C zk = L.at(k), zm = L.at(k == 0 ? 0 : N - k);
      const C xa = mk<T>((zk.x + zm.x) * h, (zk.y - zm.y) * h);
      const C xb = mk<T>((zk.y + zm.y) * h, (zm.x - zk.x) * h);
      if (va) Ac[k] = xa;
      if (vb) Bc[k] = xb;
    }
  }
};

// ---------------------------------------------------------------------------
// z-axis C2R, register engine
// ---------------------------------------------------------------------------
template <typename T, typename SC, int P, int G>
struct RowsC2R {
  using E = RegEngine<T, true, SC, P>;
  using C = cx<T>;
  static constexpr int N = SC::N;
  static constexpr int TPT = E::TPT;
  static constexpr int NT = G * TPT;
  static constexpr int LROW = RowPad<N>::lrow(N / 2 + 1);

  __device__ __forceinline__ static C hermitian_z(const C* S, int K, int n) {
    const bool lo = 2 * n <= N;
    const int m = lo ? n : N - n;
    C xa = S[m], xb = S[K + m];
    if (n == 0 || 2 * n == N) {
      xa.y = (T)0;
      xb.y = (T)0;
    }
    // lo: Z = A + iB ; hi: Z = conj(A) + i conj(B)
    return lo ? mk<T>(xa.x - xb.y, xa.y + xb.x) : mk<T>(xa.x + xb.y, xb.x - xa.y);
  }

  __device__ __forceinline__ static void tile(T* __restrict__ data, RowGeom g, const C* __restrict__ tw,
                                              int64_t pair0, C* smem) {
    const int gi = threadIdx.x / TPT, t = threadIdx.x % TPT;
    const int K = g.K;
    const int64_t ra = 2 * (pair0 + gi), rb = ra + 1;
    const bool va = ra < g.nrows, vb = rb < g.nrows;
    const C* Ac = reinterpret_cast<const C*>(data) + ra * K;
    const C* Bc = Ac + K;
    C* S = smem + gi * LROW;
    for (int k = t; k < K; k += TPT) {
      S[k] = va ? Ac[k] : mk<T>(0, 0);
      S[K + k] = vb ? Bc[k] : mk<T>(0, 0);
    }
    __syncthreads();
    C v[P];
#pragma unroll
    for (int q = 0; q < P / E::R0; ++q)
#pragma unroll
      for (int r = 0; r < E::R0; ++r) v[q * E::R0 + r] = hermitian_z(S, K, E::in_index(t, q, r));
    RowLds<T, N> L{S};
    E::run(v, t, L, tw);
    // every read of this workgroup's rows happened before the staging barrier
    T* A = data + ra * 2 * (int64_t)K;
    T* B = A + 2 * (int64_t)K;
    const T sc = (T)g.scale;
    if (g.scale != 1.0) {
#pragma unroll
      for (int i = 0; i < P; ++i) v[i] = mk<T>(v[i].x * sc, v[i].y * sc);
    }
#pragma unroll
    for (int q = 0; q < P / E::RL; ++q)
#pragma unroll
      for (int r = 0; r < E::RL; ++r) {
        const int n = E::out_index(t, q, r);
        const C z = v[q * E::RL + r];
        if (va) A[n] = z.x;
        if (vb) B[n] = z.y;
      }
  }
};

// ---------------------------------------------------------------------------
// Strided C2C column pass, register engine
// ---------------------------------------------------------------------------
template <typename T, bool INV, typename SC, int P, int CW>
struct ColsPass {
  using E = RegEngine<T, INV, SC, P>;
  using C = cx<T>;
  static constexpr int N = SC::N;
  static constexpr int TPT = E::TPT;
  static constexpr int NT = CW * TPT;

  __device__ __forceinline__ static void tile(C* __restrict__ data, ColGeom g, const C* __restrict__ tw,
                                              uint32_t tile_id, C* smem) {
    const int c = threadIdx.x % CW, t = threadIdx.x / CW;
    const uint32_t col = tile_id * CW + c;
    const bool valid = col < g.ncols;
    const int64_t base = valid ? (int64_t)(col / g.inner) * g.outer + (int64_t)(col % g.inner) : 0;
    C* p = data + base;
    C v[P];
#pragma unroll
    for (int q = 0; q < P / E::R0; ++q)
#pragma unroll
      for (int r = 0; r < E::R0; ++r)
        v[q * E::R0 + r] = valid ? p[(int64_t)E::in_index(t, q, r) * g.rstride] : mk<T>(0, 0);
    ColLds<T, CW> L{smem, c};
    E::run(v, t, L, tw);
    if (g.scale != 1.0) {
      const T sc = (T)g.scale;
#pragma unroll
      for (int i = 0; i < P; ++i) v[i] = mk<T>(v[i].x * sc, v[i].y * sc);
    }
    if (valid) {
#pragma unroll
      for (int q = 0; q < P / E::RL; ++q)
#pragma unroll
        for (int r = 0; r < E::RL; ++r) p[(int64_t)E::out_index(t, q, r) * g.rstride] = v[q * E::RL + r];
    }
  }
};

// ---------------------------------------------------------------------------
// __global__ wrappers (one tile per workgroup, XCD-aware tile order)
// ---------------------------------------------------------------------------
template <typename T, typename SC, int P, int G>
__global__ __attribute__((amdgpu_flat_work_group_size(1, RowsR2C<T, SC, P, G>::NT))) void k_rows_r2c(T* data, RowGeom g, const cx<T>* tw) {
  using K_ = RowsR2C<T, SC, P, G>;
  __shared__ cx<T> smem[G * K_::LROW];
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  K_::tile(data, g, tw, (int64_t)tile * G, smem);
}

template <typename T, typename SC, int P, int G>
__global__ __attribute__((amdgpu_flat_work_group_size(1, RowsC2R<T, SC, P, G>::NT))) void k_rows_c2r(T* data, RowGeom g, const cx<T>* tw) {
  using K_ = RowsC2R<T, SC, P, G>;
  __shared__ cx<T> smem[G * K_::LROW];
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  K_::tile(data, g, tw, (int64_t)tile * G, smem);
}

template <typename T, bool INV, typename SC, int P, int CW>
__global__ __attribute__((amdgpu_flat_work_group_size(1, ColsPass<T, INV, SC, P, CW>::NT))) void k_cols(cx<T>* data, ColGeom g, const cx<T>* tw) {
  using K_ = ColsPass<T, INV, SC, P, CW>;
  __shared__ cx<T> smem[SC::N * CW];
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  K_::tile(data, g, tw, tile, smem);
}

// Host launchers with the registry's signatures.
template <typename T, typename SC, int P, int G>
hipError_t launch_rows_r2c(void* data, RowGeom g, const void* tw, hipStream_t s) {
  const int64_t pairs = (g.nrows + 1) / 2;
  const int64_t blocks = (pairs + G - 1) / G;
  if (blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_rows_r2c<T, SC, P, G>), dim3((uint32_t)blocks), dim3(RowsR2C<T, SC, P, G>::NT), 0, s,
                     static_cast<T*>(data), g, static_cast<const cx<T>*>(tw));
  return hipGetLastError();
}

template <typename T, typename SC, int P, int G>
hipError_t launch_rows_c2r(void* data, RowGeom g, const void* tw, hipStream_t s) {
  const int64_t pairs = (g.nrows + 1) / 2;
  const int64_t blocks = (pairs + G - 1) / G;
  if (blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_rows_c2r<T, SC, P, G>), dim3((uint32_t)blocks), dim3(RowsC2R<T, SC, P, G>::NT), 0, s,
                     static_cast<T*>(data), g, static_cast<const cx<T>*>(tw));
  return hipGetLastError();
}

template <typename T, bool INV, typename SC, int P, int CW>
hipError_t launch_cols(void* data, ColGeom g, const void* tw, hipStream_t s) {
  const int64_t blocks = ((int64_t)g.ncols + CW - 1) / CW;
  if (blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_cols<T, INV, SC, P, CW>), dim3((uint32_t)blocks), dim3(ColsPass<T, INV, SC, P, CW>::NT), 0,
                     s, static_cast<cx<T>*>(data), g, static_cast<const cx<T>*>(tw));
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
template <typename T, bool INV, typename SC, int P, int CW>
KernelEntry entry_cols() {
  return KernelEntry{prec_of<T>(), INV ? COLS_INV : COLS_FWD, SC::N, make_sched<SC>(), P, CW,
                     ColsPass<T, INV, SC, P, CW>::NT, reinterpret_cast<void*>(&launch_cols<T, INV, SC, P, CW>)};
}

}  // namespace gfft
