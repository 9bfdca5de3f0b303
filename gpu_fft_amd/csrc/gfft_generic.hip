// Run-time-schedule kernels: the same Stockham recurrence and z-pass
// two-for-one split as the register engine, with every stage staged through
// LDS ping-pong buffers.  Used for any 13-smooth size that has no compiled
// register variant (the planner prefers compiled variants).
#include "gfft_kernels.h"

namespace gfft {

namespace {

template <typename T>
struct GenCfg;
template <>
struct GenCfg<float> {
  static constexpr int CW = 8;  // columns per tile (64 B segments)
};
template <>
struct GenCfg<double> {
  static constexpr int CW = 4;
};

constexpr int kGenThreads = 256;
constexpr size_t kMaxLds = 160 * 1024;

template <int R, bool INV, typename T>
__device__ __forceinline__ void gen_bfly(const cx<T>* in, cx<T>* out, int i, int N, int p, const cx<T>* tw, int ld) {
  cx<T> a[R];
#pragma unroll
  for (int r = 0; r < R; ++r) a[r] = in[(i + r * (N / R)) * ld];
  const int k = i % p;
  if (p > 1) {
    const cx<T>* w = tw + k * (R - 1);
#pragma unroll
    for (int r = 1; r < R; ++r) a[r] = twmul<INV>(a[r], w[r - 1]);
  }
  Dft<R, INV, T>::run(a);
  const int j = (i - k) * R + k;
#pragma unroll
  for (int r = 0; r < R; ++r) out[(j + r * p) * ld] = a[r];
}

template <bool INV, typename T>
__device__ __forceinline__ void gen_dispatch(int R, const cx<T>* in, cx<T>* out, int i, int N, int p,
                                             const cx<T>* tw, int ld) {
  switch (R) {
    case 2: gen_bfly<2, INV, T>(in, out, i, N, p, tw, ld); break;
    case 3: gen_bfly<3, INV, T>(in, out, i, N, p, tw, ld); break;
    case 4: gen_bfly<4, INV, T>(in, out, i, N, p, tw, ld); break;
    case 5: gen_bfly<5, INV, T>(in, out, i, N, p, tw, ld); break;
    case 7: gen_bfly<7, INV, T>(in, out, i, N, p, tw, ld); break;
    case 11: gen_bfly<11, INV, T>(in, out, i, N, p, tw, ld); break;
    case 13: gen_bfly<13, INV, T>(in, out, i, N, p, tw, ld); break;
    default: gen_bfly<8, INV, T>(in, out, i, N, p, tw, ld); break;
  }
}

// Runs all stages; data starts in buf[0]; returns the buffer holding the result.
// Rows layout: element (g, n) at g*N + n (ld 1).  Cols: (c, n) at n*CW + c.
template <bool INV, typename T, bool ROWS, int CW>
__device__ int gen_stages(cx<T>* buf0, cx<T>* buf1, int nset, const RtSched& sc, const cx<T>* tw) {
  cx<T>* b[2] = {buf0, buf1};
  int cur = 0;
  const int N = sc.N;
  for (int s = 0; s < sc.S; ++s) {
    const int R = sc.R[s], p = sc.p[s];
    const int nb = N / R;
    const cx<T>* in = b[cur];
    cx<T>* out = b[cur ^ 1];
    for (int f = threadIdx.x; f < nb * nset; f += blockDim.x) {
      int i, off, ld;
      if (ROWS) {
        const int g = f / nb;
        i = f - g * nb;
        off = g * N;
        ld = 1;
      } else {
        const int c = f % CW;
        i = f / CW;
        off = c;
        ld = CW;
      }
      gen_dispatch<INV, T>(R, in + off, out + off, i, N, p, tw + sc.twoff[s], ld);
    }
    __syncthreads();
    cur ^= 1;
  }
  return cur;
}

template <typename T>
__global__ __launch_bounds__(kGenThreads) void k_gen_rows_r2c(T* data, RowGeom g, RtSched sc, const cx<T>* tw,
                                                              int G) {
  extern __shared__ unsigned char raw[];
  cx<T>* buf0 = reinterpret_cast<cx<T>*>(raw);
  const int N = sc.N, K = g.K;
  cx<T>* buf1 = buf0 + G * N;
  const uint32_t t0 = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t pair0 = (int64_t)(g.rev ? gridDim.x - 1 - t0 : t0) * G;
  for (int f = threadIdx.x; f < G * N; f += blockDim.x) {
    const int gi = f / N, n = f - gi * N;
    const int64_t ra = 2 * (pair0 + gi), rb = ra + 1;
    const T a = ra < g.nrows ? data[ra * 2 * (int64_t)K + n] : (T)0;
    const T b = rb < g.nrows ? data[rb * 2 * (int64_t)K + n] : (T)0;
    buf0[f] = mk<T>(a, b);
  }
  __syncthreads();
  const int cur = gen_stages<false, T, true, 1>(buf0, buf1, G, sc, tw);
  const cx<T>* Z = cur ? buf1 : buf0;
  const T h = (T)(0.5 * g.scale);
  for (int f = threadIdx.x; f < G * K; f += blockDim.x) {
    const int gi = f / K, k = f - gi * K;
    const int64_t ra = 2 * (pair0 + gi), rb = ra + 1;
    const cx<T> zk = Z[gi * N + k], zm = Z[gi * N + (k == 0 ? 0 : N - k)];
    cx<T>* Ac = reinterpret_cast<cx<T>*>(data) + ra * K;
    if (ra < g.nrows) Ac[k] = mk<T>((zk.x + zm.x) * h, (zk.y - zm.y) * h);
    if (rb < g.nrows) Ac[K + k] = mk<T>((zk.y + zm.y) * h, (zm.x - zk.x) * h);
  }
}

template <typename T>
__global__ __launch_bounds__(kGenThreads) void k_gen_rows_c2r(T* data, RowGeom g, RtSched sc, const cx<T>* tw,
                                                              int G) {
  extern __shared__ unsigned char raw[];
  cx<T>* buf0 = reinterpret_cast<cx<T>*>(raw);
  const int N = sc.N, K = g.K;
  const int ls = N + 2;  // staging stride per pair (>= 2K)
  cx<T>* buf1 = buf0 + G * ls;
  const uint32_t t0 = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t pair0 = (int64_t)(g.rev ? gridDim.x - 1 - t0 : t0) * G;
  // stage both half spectra of every pair in buf1
  for (int f = threadIdx.x; f < G * K; f += blockDim.x) {
    const int gi = f / K, k = f - gi * K;
    const int64_t ra = 2 * (pair0 + gi), rb = ra + 1;
    const cx<T>* Ac = reinterpret_cast<const cx<T>*>(data) + ra * K;
    buf1[gi * ls + k] = ra < g.nrows ? Ac[k] : mk<T>(0, 0);
    buf1[gi * ls + K + k] = rb < g.nrows ? Ac[K + k] : mk<T>(0, 0);
  }
  __syncthreads();
  for (int f = threadIdx.x; f < G * N; f += blockDim.x) {
    const int gi = f / N, n = f - gi * N;
    const cx<T>* S = buf1 + gi * ls;
    const bool lo = 2 * n <= N;
    const int m = lo ? n : N - n;
    cx<T> xa = S[m], xb = S[K + m];
    if (n == 0 || 2 * n == N) {
      xa.y = (T)0;
      xb.y = (T)0;
    }
    buf0[gi * N + n] = lo ? mk<T>(xa.x - xb.y, xa.y + xb.x) : mk<T>(xa.x + xb.y, xb.x - xa.y);
  }
  __syncthreads();
  const int cur = gen_stages<true, T, true, 1>(buf0, buf1, G, sc, tw);
  const cx<T>* Z = cur ? buf1 : buf0;
  for (int f = threadIdx.x; f < G * N; f += blockDim.x) {
    const int gi = f / N, n = f - gi * N;
    const int64_t ra = 2 * (pair0 + gi), rb = ra + 1;
    const cx<T> z = Z[gi * N + n];
    if (ra < g.nrows) data[ra * 2 * (int64_t)K + n] = z.x * (T)g.scale;
    if (rb < g.nrows) data[rb * 2 * (int64_t)K + n] = z.y * (T)g.scale;
  }
}

__device__ __forceinline__ int64_t gen_base(const ColMap& m, uint32_t c) {
  return (int64_t)(c / m.inner) * m.outer + (int64_t)(c % m.inner);
}
__device__ __forceinline__ int64_t gen_row(const ColMap& m, int n, int N) {
  if (m.nsub >= N) return (int64_t)n * m.rlo;
  return (int64_t)(n % m.nsub) * m.rlo + (int64_t)(n / m.nsub) * m.rhi;
}

template <typename T, bool INV>
__global__ __launch_bounds__(kGenThreads) void k_gen_cols(const cx<T>* in, cx<T>* out, ColGeom g, RtSched sc,
                                                             const cx<T>* tw) {
  constexpr int CW = GenCfg<T>::CW;
  extern __shared__ unsigned char raw[];
  cx<T>* buf0 = reinterpret_cast<cx<T>*>(raw);
  const int N = sc.N;
  cx<T>* buf1 = buf0 + N * CW;
  const uint32_t t0 = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t tile = g.rev ? gridDim.x - 1 - t0 : t0;
  for (int f = threadIdx.x; f < N * CW; f += blockDim.x) {
    const int c = f % CW, n = f / CW;
    const uint32_t col = tile * CW + c;
    cx<T> v = mk<T>(0, 0);
    if (col < g.ncols) {
      v = in[gen_base(g.in, col) + gen_row(g.in, n, N)];
    }
    buf0[f] = v;
  }
  __syncthreads();
  const int cur = gen_stages<INV, T, false, CW>(buf0, buf1, CW, sc, tw);
  const cx<T>* Z = cur ? buf1 : buf0;
  for (int f = threadIdx.x; f < N * CW; f += blockDim.x) {
    const int c = f % CW, n = f / CW;
    const uint32_t col = tile * CW + c;
    if (col < g.ncols) {
      const cx<T> z = Z[f];
      out[gen_base(g.out, col) + gen_row(g.out, n, N)] = mk<T>(z.x * (T)g.scale, z.y * (T)g.scale);
    }
  }
}

RtSched to_dev(const RtSchedHost& h) {
  RtSched d{};
  d.N = h.N;
  d.S = h.S;
  for (int i = 0; i < kMaxStages && i < kMaxRadices; ++i) {
    d.R[i] = h.R[i];
    d.p[i] = h.p[i];
    d.twoff[i] = h.twoff[i];
  }
  return d;
}

int gen_pairs_per_block(int N) {
  int G = 2048 / N;
  if (G < 1) G = 1;
  if (G > 8) G = 8;
  return G;
}

template <typename T>
size_t rows_lds_bytes(int N, int G) {
  // two buffers of G pairs, each pair (N + 2) entries: the C2R kernel stages
  // 2K <= N + 2 half-spectrum bins per pair in the second buffer
  return (size_t)2 * G * (size_t)(N + 2) * sizeof(cx<T>);
}

template <typename T>
hipError_t gen_rows(int kind, void* data, RowGeom g, const RtSchedHost& h, const void* tw, hipStream_t s) {
  const int N = h.N;
  const int G = gen_pairs_per_block(N);
  const int64_t pairs = (g.nrows + 1) / 2;
  const int64_t blocks = (pairs + G - 1) / G;
  if (blocks <= 0) return hipSuccess;
  const size_t lds = rows_lds_bytes<T>(N, G);
  if (lds > kMaxLds) return hipErrorInvalidValue;
  RtSched sc = to_dev(h);
  if (kind == ROWS_R2C) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_gen_rows_r2c<T>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    clear_launch_error();
    hipLaunchKernelGGL(k_gen_rows_r2c<T>, dim3((uint32_t)blocks), dim3(kGenThreads), lds, s, static_cast<T*>(data),
                       g, sc, static_cast<const cx<T>*>(tw), G);
  } else {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_gen_rows_c2r<T>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    clear_launch_error();
    hipLaunchKernelGGL(k_gen_rows_c2r<T>, dim3((uint32_t)blocks), dim3(kGenThreads), lds, s, static_cast<T*>(data),
                       g, sc, static_cast<const cx<T>*>(tw), G);
  }
  return hipGetLastError();
}

template <typename T>
size_t cols_lds_bytes(int N) {
  return (size_t)2 * N * GenCfg<T>::CW * sizeof(cx<T>);
}

template <typename T>
hipError_t gen_cols(int kind, const void* in, void* out, ColGeom g, const RtSchedHost& h, const void* tw,
                    hipStream_t s) {
  constexpr int CW = GenCfg<T>::CW;
  const int64_t blocks = ((int64_t)g.ncols + CW - 1) / CW;
  if (blocks <= 0) return hipSuccess;
  const size_t lds = cols_lds_bytes<T>(h.N);
  if (lds > kMaxLds) return hipErrorInvalidValue;
  RtSched sc = to_dev(h);
  const void* fn = kind == COLS_INV ? reinterpret_cast<const void*>(&k_gen_cols<T, true>)
                                    : reinterpret_cast<const void*>(&k_gen_cols<T, false>);
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  clear_launch_error();
  if (kind == COLS_INV)
    hipLaunchKernelGGL((k_gen_cols<T, true>), dim3((uint32_t)blocks), dim3(kGenThreads), lds, s,
                       static_cast<const cx<T>*>(in), static_cast<cx<T>*>(out), g, sc, static_cast<const cx<T>*>(tw));
  else
    hipLaunchKernelGGL((k_gen_cols<T, false>), dim3((uint32_t)blocks), dim3(kGenThreads), lds, s,
                       static_cast<const cx<T>*>(in), static_cast<cx<T>*>(out), g, sc, static_cast<const cx<T>*>(tw));
  return hipGetLastError();
}

}  // namespace

hipError_t launch_generic_rows(int prec, int kind, void* data, RowGeom g, const RtSchedHost& sc, const void* tw,
                               hipStream_t s) {
  return prec == F32 ? gen_rows<float>(kind, data, g, sc, tw, s) : gen_rows<double>(kind, data, g, sc, tw, s);
}

hipError_t launch_generic_cols(int prec, int kind, const void* in, void* out, ColGeom g, const RtSchedHost& sc,
                               const void* tw,
                               hipStream_t s) {
  return prec == F32 ? gen_cols<float>(kind, in, out, g, sc, tw, s) : gen_cols<double>(kind, in, out, g, sc, tw, s);
}

size_t generic_rows_lds(int prec, int N, int K) {
  (void)K;
  const int G = gen_pairs_per_block(N);
  const size_t b = prec == F32 ? rows_lds_bytes<float>(N, G) : rows_lds_bytes<double>(N, G);
  return b <= kMaxLds ? b : 0;
}

size_t generic_cols_lds(int prec, int N) {
  const size_t b = prec == F32 ? cols_lds_bytes<float>(N) : cols_lds_bytes<double>(N);
  return b <= kMaxLds ? b : 0;
}

}  // namespace gfft
