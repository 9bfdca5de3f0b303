// Wavefront-shuffle variant of a column (x) pass, measured against the
// library's own x pass on the same data.
//
// The library's column kernels run every radix stage of an N-point column
// transform in registers and exchange between stages through LDS (Stockham,
// 16 columns per 256-thread tile, columns fastest across the lanes so that a
// wavefront's loads cover whole 128 B row segments).  This variant keeps the
// first factor in registers and does every later stage with cross-lane DPP
// moves, no LDS at all:
//   N = N1 * 16; lane t (0..15 of a 16-lane DPP row) of column c holds
//   x[t + 16 j], j < N1;
//   1. N1-point DFT over j in registers (k1 = 0..N1-1),
//   2. times W_N^(t k1),
//   3. 16-point DFT over the 16 lanes of the row for every k1: radix-2
//      decimation in frequency, partners at lane distance 8, 4, 2, 1 fetched
//      with DPP (xor 1 and 2: quad_perm; xor 4 and 8: a row or half-row
//      mirror followed by a second move),
//   4. lane t holds X[k1 + N1 * bitrev4(t)]; it stores those.
// A DPP row is 16 lanes, so a wavefront holds 4 columns and a 256-thread
// workgroup 16 (the same tile width as the library's kernels), but each load
// instruction now touches 16 rows x 4 columns (32 B per row segment) instead
// of 4 rows x 16 columns.
//
// A second form (k_cols_lane_lds) keeps the library's global access pattern
// and moves the tile between it and the lane layout through LDS.
//
// Output: one JSON line with the library's x-pass time (tuned plan), both
// variants' times and their relative L2 differences from the library's
// result, for the 128^3 and 256^3 FP32 grids.  Exit status 1 if a result
// differs by more than 1e-5 or a call fails.
//
// Build: make -C tests/shuffle_bench     Run: tests/shuffle_bench/col_shuffle [reps]
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gfft_device.h"
#include "gpufft/gpufft.h"

using gfft::cx;
using gfft::mk;

namespace {

template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}

// value of lane (t ^ H) of the calling lane's 16-lane row
template <int H>
__device__ __forceinline__ float xor_fetch(float v) {
  if constexpr (H == 1) return dpp<0xB1>(v);                 // quad_perm [1,0,3,2]
  if constexpr (H == 2) return dpp<0x4E>(v);                 // quad_perm [2,3,0,1]
  if constexpr (H == 4) return dpp<0x1B>(dpp<0x141>(v));     // half-row mirror (t^7), quad_perm [3,2,1,0] (^3)
  if constexpr (H == 8) return dpp<0x141>(dpp<0x140>(v));    // row mirror (t^15), half-row mirror (^7)
  return v;
}

// one radix-2 decimation-in-frequency stage across lanes (span H): lower
// lanes keep v + partner, upper lanes (partner - v) * w
template <int H, int N1>
__device__ __forceinline__ void lane_stage(cx<float> (&v)[N1], float sgn, cx<float> w) {
#pragma unroll
  for (int s = 0; s < N1; ++s) {
    const cx<float> p = mk<float>(xor_fetch<H>(v[s].x), xor_fetch<H>(v[s].y));
    const cx<float> d = mk<float>(fmaf(sgn, v[s].x, p.x), fmaf(sgn, v[s].y, p.y));
    v[s] = H == 1 ? d : gfft::twmul<false>(d, w);
  }
}

// 16-point DFT in registers, natural order: n = 4 m + l, radix 4 over m,
// twiddle W16^(l k1), radix 4 over l; X[k1 + 4 k2]
__device__ __forceinline__ void dft16(cx<float> (&a)[16], const cx<float>* tab, int tstep) {
  cx<float> b[4][4];
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    cx<float> q[4] = {a[l], a[4 + l], a[8 + l], a[12 + l]};
    gfft::Dft<4, false, float>::run(q);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) b[l][k1] = (l * k1) ? gfft::twmul<false>(q[k1], tab[l * k1 * tstep]) : q[k1];
  }
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    cx<float> q[4] = {b[0][k1], b[1][k1], b[2][k1], b[3][k1]};
    gfft::Dft<4, false, float>::run(q);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) a[k1 + 4 * k2] = q[k2];
  }
}

// The same lane FFT with the library's global access pattern: the tile is
// loaded and stored columns-fastest (each wavefront instruction covers 4 rows
// x 16 columns, 128 B row segments) and moved between that layout and the
// lane layout through LDS (two LDS round trips, as many as the library's
// three-stage Stockham schedules exchange through).
template <int N1>
__device__ __forceinline__ void lane_fft(cx<float> (&v)[N1], int t, const cx<float>* __restrict__ tab) {
  constexpr int N = N1 * 16;
  if constexpr (N1 == 8)
    gfft::Dft<8, false, float>::run(v);
  else
    dft16(v, tab, N / 16);
#pragma unroll
  for (int k1 = 1; k1 < N1; ++k1) v[k1] = gfft::twmul<false>(v[k1], tab[t * k1]);
  const cx<float> one = mk<float>(1.0f, 0.0f);
  const cx<float> w8 = (t & 8) ? tab[(t & 7) * (N / 16)] : one;
  const cx<float> w4 = (t & 4) ? tab[(t & 3) * (N / 8)] : one;
  const cx<float> w2 = (t & 2) ? tab[(t & 1) * (N / 4)] : one;
  lane_stage<8>(v, (t & 8) ? -1.0f : 1.0f, w8);
  lane_stage<4>(v, (t & 4) ? -1.0f : 1.0f, w4);
  lane_stage<2>(v, (t & 2) ? -1.0f : 1.0f, w2);
  lane_stage<1>(v, (t & 1) ? -1.0f : 1.0f, one);
}

template <int N1>
__global__ __launch_bounds__(256) void k_cols_lane_lds(cx<float>* data, int64_t stride,
                                                       const cx<float>* __restrict__ tab) {
  constexpr int N = N1 * 16, LP = 17;  // LDS row pitch (16 columns + 1)
  __shared__ cx<float> sm[N * LP];
  const int col = threadIdx.x & 15, r0 = threadIdx.x >> 4;  // columns-fastest I/O layout
  const int lane = threadIdx.x & 63, t = lane & 15, cl = (threadIdx.x >> 6) * 4 + (lane >> 4);
  cx<float>* d = data + (int64_t)blockIdx.x * 16;
  cx<float> v[N1];
#pragma unroll
  for (int j = 0; j < N1; ++j) v[j] = d[(int64_t)(r0 + 16 * j) * stride + col];
#pragma unroll
  for (int j = 0; j < N1; ++j) sm[(r0 + 16 * j) * LP + col] = v[j];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < N1; ++j) v[j] = sm[(t + 16 * j) * LP + cl];
  lane_fft<N1>(v, t, tab);
  const int k2 = ((t & 1) << 3) | ((t & 2) << 1) | ((t & 4) >> 1) | ((t & 8) >> 3);
  __syncthreads();  // every lane-layout read of sm is done
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) sm[(k1 + N1 * k2) * LP + cl] = v[k1];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < N1; ++j) d[(int64_t)(r0 + 16 * j) * stride + col] = sm[(r0 + 16 * j) * LP + col];
}

// tab[m] = W_N^m, m < N
template <int N1>
__global__ __launch_bounds__(256) void k_cols_lane(cx<float>* data, int64_t stride, const cx<float>* __restrict__ tab) {
  constexpr int N = N1 * 16;
  const int lane = threadIdx.x & 63, t = lane & 15;
  const int64_t c = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 6) * 4 + (lane >> 4);
  cx<float> v[N1];
#pragma unroll
  for (int j = 0; j < N1; ++j) v[j] = data[(int64_t)(t + 16 * j) * stride + c];
  lane_fft<N1>(v, t, tab);
  const int k2 = ((t & 1) << 3) | ((t & 2) << 1) | ((t & 4) >> 1) | ((t & 8) >> 3);  // bitrev4(t)
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) data[(int64_t)(k1 + N1 * k2) * stride + c] = v[k1];
}

#define CK(x)                                                              \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess) {                                                \
      std::printf("{\"error\": \"%s: %s\"}\n", #x, hipGetErrorString(e_)); \
      std::exit(1);                                                        \
    }                                                                      \
  } while (0)
#define CG(x)                                                                          \
  do {                                                                                 \
    int s_ = (x);                                                                      \
    if (s_ != GFFT_OK) {                                                               \
      std::printf("{\"error\": \"%s: %s\"}\n", #x, gfft_strerror(s_));                 \
      std::exit(1);                                                                    \
    }                                                                                  \
  } while (0)

struct Result {
  int n;
  double lib_us, lane_us, lds_us, rel_err, rel_err_lds;
};

template <int N1>
Result run_grid(int reps) {
  constexpr int N = N1 * 16;
  const int64_t K = N / 2 + 1, ncols = (int64_t)N * K, elems = (int64_t)N * ncols;
  if (ncols % 16 != 0) {
    std::printf("{\"error\": \"columns not a multiple of 16\"}\n");
    std::exit(1);
  }
  gfft_plan plan = nullptr;
  CG(gfft_plan_create(&plan, N, N, N, 1, 0, GFFT_F32));
  size_t bytes = 0;
  CG(gfft_plan_data_bytes(plan, &bytes));
  if (bytes < (size_t)elems * sizeof(cx<float>)) {
    std::printf("{\"error\": \"plan data smaller than the grid\"}\n");
    std::exit(1);
  }
  hipStream_t s;
  CK(hipStreamCreate(&s));
  cx<float> *a = nullptr, *b = nullptr, *c = nullptr, *tab = nullptr;
  CK(hipMalloc(&a, bytes));
  CK(hipMalloc(&b, bytes));
  CK(hipMalloc(&c, bytes));
  CK(hipMalloc(&tab, sizeof(cx<float>) * N));
  CG(gfft_plan_tune(plan, a, s, 3));
  std::vector<cx<float>> h((size_t)elems), ht(N);
  uint64_t z = 0x9E3779B97F4A7C15ull;
  for (auto& v : h) {
    z = z * 6364136223846793005ull + 1442695040888963407ull;
    v.x = (float)((z >> 40) * (1.0 / 16777216.0)) - 0.5f;
    z = z * 6364136223846793005ull + 1442695040888963407ull;
    v.y = (float)((z >> 40) * (1.0 / 16777216.0)) - 0.5f;
  }
  for (int m = 0; m < N; ++m) {
    const double ang = -2.0 * M_PI * (double)m / (double)N;
    ht[m].x = (float)std::cos(ang);
    ht[m].y = (float)std::sin(ang);
  }
  CK(hipMemcpy(tab, ht.data(), sizeof(cx<float>) * N, hipMemcpyHostToDevice));
  // correctness: the same input through both x passes
  CK(hipMemcpy(a, h.data(), sizeof(cx<float>) * elems, hipMemcpyHostToDevice));
  CK(hipMemcpy(b, h.data(), sizeof(cx<float>) * elems, hipMemcpyHostToDevice));
  CK(hipMemcpy(c, h.data(), sizeof(cx<float>) * elems, hipMemcpyHostToDevice));
  CG(gfft_exec_axis(plan, 2, 0, a, s));
  const dim3 grid((unsigned)(ncols / 16)), block(256);
  hipLaunchKernelGGL(k_cols_lane<N1>, grid, block, 0, s, b, ncols, tab);
  hipLaunchKernelGGL(k_cols_lane_lds<N1>, grid, block, 0, s, c, ncols, tab);
  CK(hipGetLastError());
  CK(hipStreamSynchronize(s));
  std::vector<cx<float>> ra((size_t)elems), rb((size_t)elems), rc((size_t)elems);
  CK(hipMemcpy(ra.data(), a, sizeof(cx<float>) * elems, hipMemcpyDeviceToHost));
  CK(hipMemcpy(rb.data(), b, sizeof(cx<float>) * elems, hipMemcpyDeviceToHost));
  CK(hipMemcpy(rc.data(), c, sizeof(cx<float>) * elems, hipMemcpyDeviceToHost));
  double num = 0, num_c = 0, den = 0;
  for (int64_t i = 0; i < elems; ++i) {
    const double dx = (double)ra[i].x - rb[i].x, dy = (double)ra[i].y - rb[i].y;
    const double ex = (double)ra[i].x - rc[i].x, ey = (double)ra[i].y - rc[i].y;
    num += dx * dx + dy * dy;
    num_c += ex * ex + ey * ey;
    den += (double)ra[i].x * ra[i].x + (double)ra[i].y * ra[i].y;
  }
  // timing: the two passes alternate in blocks on one stream
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  double lib_ms = 0, lane_ms = 0, lds_ms = 0;
  for (int round = 0; round < 4; ++round) {
    for (int w = 0; w < 3; ++w) CG(gfft_exec_axis(plan, 2, 0, a, s));
    CK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) CG(gfft_exec_axis(plan, 2, 0, a, s));
    CK(hipEventRecord(e1, s));
    CK(hipEventSynchronize(e1));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, e0, e1));
    lib_ms += ms;
    for (int w = 0; w < 3; ++w) hipLaunchKernelGGL(k_cols_lane<N1>, grid, block, 0, s, b, ncols, tab);
    CK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_cols_lane<N1>, grid, block, 0, s, b, ncols, tab);
    CK(hipEventRecord(e1, s));
    CK(hipEventSynchronize(e1));
    CK(hipGetLastError());
    CK(hipEventElapsedTime(&ms, e0, e1));
    lane_ms += ms;
    for (int w = 0; w < 3; ++w) hipLaunchKernelGGL(k_cols_lane_lds<N1>, grid, block, 0, s, c, ncols, tab);
    CK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_cols_lane_lds<N1>, grid, block, 0, s, c, ncols, tab);
    CK(hipEventRecord(e1, s));
    CK(hipEventSynchronize(e1));
    CK(hipGetLastError());
    CK(hipEventElapsedTime(&ms, e0, e1));
    lds_ms += ms;
  }
  Result res{N, 1e3 * lib_ms / (4.0 * reps), 1e3 * lane_ms / (4.0 * reps), 1e3 * lds_ms / (4.0 * reps),
             std::sqrt(num / den), std::sqrt(num_c / den)};
  CK(hipEventDestroy(e0));
  CK(hipEventDestroy(e1));
  CK(hipFree(a));
  CK(hipFree(b));
  CK(hipFree(c));
  CK(hipFree(tab));
  CK(hipStreamDestroy(s));
  CG(gfft_plan_destroy(plan));
  return res;
}

}  // namespace

int main(int argc, char** argv) {
  const int reps = argc > 1 ? std::atoi(argv[1]) : 50;
  if (reps < 1 || reps > 10000) return 1;
  const Result r[2] = {run_grid<8>(reps), run_grid<16>(reps)};
  bool ok = true;
  std::printf("{\"workload\": \"x pass (C2C along x, forward) of an N^3 FP32 R2C grid, in place\", \"reps\": %d, \"grids\": [",
              reps);
  for (int i = 0; i < 2; ++i) {
    const double bytes = 2.0 * (double)r[i].n * r[i].n * (r[i].n / 2 + 1) * 8.0;
    std::printf("%s{\"n\": %d, \"library_us\": %.2f, \"lane_shuffle_us\": %.2f, \"lane_shuffle_lds_io_us\": %.2f, "
                "\"library_GBps\": %.1f, \"lane_shuffle_GBps\": %.1f, \"lane_shuffle_lds_io_GBps\": %.1f, "
                "\"rel_l2_diff\": %.3g, \"rel_l2_diff_lds_io\": %.3g}",
                i ? ", " : "", r[i].n, r[i].lib_us, r[i].lane_us, r[i].lds_us, bytes / (r[i].lib_us * 1e3),
                bytes / (r[i].lane_us * 1e3), bytes / (r[i].lds_us * 1e3), r[i].rel_err, r[i].rel_err_lds);
    ok = ok && r[i].rel_err <= 1e-5 && r[i].rel_err_lds <= 1e-5;
  }
  std::printf("]}\n");
  return ok ? 0 : 1;
}
