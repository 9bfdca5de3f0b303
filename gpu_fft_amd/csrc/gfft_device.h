// Device-side building blocks of the gfx950 FFT engine.
//
//  * cx<T>            : POD complex, 8/16-byte aligned (one dwordx2 / dwordx4)
//  * Dft<R, INV>      : radix-2/3/4/5/7/8/11/13 butterflies, FMA form, in registers
//  * Sched<R...>      : a compile-time mixed-radix Stockham schedule
//  * RegEngine        : register-resident Stockham; each thread owns P points,
//                       stages exchange through LDS (one barrier pair each)
//  * RtSched/gen_*    : the same Stockham recurrence with a run-time schedule,
//                       LDS ping-pong, for sizes without a compiled schedule
//
// Stockham recurrence (all stages, both engines), for radix R and p = product
// of the radices already applied, butterfly i in [0, N/R), k = i mod p:
//    a_r = x[i + r N/R] * W_{pR}^{r k}        (r = 0..R-1, W_L = e^{-2 pi i/L})
//    A   = DFT_R(a)
//    y[(i - k) R + k + r p] = A_r
// The output of the last stage is in natural order.  Inverse transforms use
// the conjugate twiddles and butterflies (sign +1), unnormalised.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gfft {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> mk(T a, T b) {
  cx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }

// a * w (INV = false) or a * conj(w) (INV = true)
template <bool INV, typename T>
__device__ __forceinline__ cx<T> twmul(cx<T> a, cx<T> w) {
  if (INV) return mk<T>(fma(a.x, w.x, a.y * w.y), fma(a.y, w.x, -(a.x * w.y)));
  return mk<T>(fma(a.x, w.x, -(a.y * w.y)), fma(a.x, w.y, a.y * w.x));
}
// Global accesses of the passes.  GFFT_NONTEMPORAL=1 adds the nontemporal
// hint; measured on MI355X it halves the strided passes (row segments that
// straddle cache lines are then fetched once per tile instead of once per
// L2), so the default build uses plain accesses.
#ifndef GFFT_NONTEMPORAL
#define GFFT_NONTEMPORAL 0
#endif
template <typename T>
struct VecOf;
template <>
struct VecOf<float> {
  typedef float type __attribute__((ext_vector_type(2)));
};
template <>
struct VecOf<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};

template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
#if GFFT_NONTEMPORAL
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}
template <typename T>
__device__ __forceinline__ cx<T> ld_stream(const cx<T>* p) {
#if GFFT_NONTEMPORAL
  const typename VecOf<T>::type v = __builtin_nontemporal_load(reinterpret_cast<const typename VecOf<T>::type*>(p));
  return mk<T>(v.x, v.y);
#else
  return *p;
#endif
}
template <typename T>
__device__ __forceinline__ void st_stream(T* p, T v) {
#if GFFT_NONTEMPORAL
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}
template <typename T>
__device__ __forceinline__ void st_stream(cx<T>* p, cx<T> v) {
#if GFFT_NONTEMPORAL
  typename VecOf<T>::type w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<typename VecOf<T>::type*>(p));
#else
  *p = v;
#endif
}

// multiply by W_4 of the transform direction: -i forward, +i inverse
template <bool INV, typename T>
__device__ __forceinline__ cx<T> rot4(cx<T> a) {
  return INV ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
}
// multiply by the real scalar s and then by W_4 of the direction
template <bool INV, typename T>
__device__ __forceinline__ cx<T> rot4s(cx<T> a, T s) {
  return INV ? mk<T>(-a.y * s, a.x * s) : mk<T>(a.y * s, -a.x * s);
}

// ---------------------------------------------------------------------------
// Butterflies (unnormalised DFT of R points, in place)
// ---------------------------------------------------------------------------
template <int R, bool INV, typename T>
struct Dft;

template <bool INV, typename T>
struct Dft<2, INV, T> {
  __device__ __forceinline__ static void run(cx<T>* a) {
    cx<T> t = a[0] - a[1];
    a[0] = a[0] + a[1];
    a[1] = t;
  }
};

template <bool INV, typename T>
struct Dft<3, INV, T> {
  __device__ __forceinline__ static void run(cx<T>* a) {
    const T s3 = (T)0.86602540378443864676372317075293618;  // sin(2 pi/3)
    cx<T> b = a[1] + a[2];
    cx<T> d = a[1] - a[2];
    cx<T> y0 = a[0] + b;
    cx<T> t = mk<T>(fma((T)-0.5, b.x, a[0].x), fma((T)-0.5, b.y, a[0].y));
    cx<T> u = rot4s<INV, T>(d, s3);  // -i s3 d (forward)
    a[0] = y0;
    a[1] = t + u;
    a[2] = t - u;
  }
};

template <bool INV, typename T>
struct Dft<4, INV, T> {
  __device__ __forceinline__ static void run(cx<T>* a) {
    cx<T> t0 = a[0] + a[2], t1 = a[0] - a[2];
    cx<T> t2 = a[1] + a[3], t3 = rot4<INV>(a[1] - a[3]);
    a[0] = t0 + t2;
    a[2] = t0 - t2;
    a[1] = t1 + t3;
    a[3] = t1 - t3;
  }
};

template <bool INV, typename T>
struct Dft<5, INV, T> {
  __device__ __forceinline__ static void run(cx<T>* a) {
    const T c1 = (T)0.30901699437494742410229341718281906;   // cos(2 pi/5)
    const T c2 = (T)-0.80901699437494742410229341718281906;  // cos(4 pi/5)
    const T s1 = (T)0.95105651629515357211643933337938214;   // sin(2 pi/5)
    const T s2 = (T)0.58778525229247312916870595463907277;   // sin(4 pi/5)
    cx<T> b1 = a[1] + a[4], d1 = a[1] - a[4];
    cx<T> b2 = a[2] + a[3], d2 = a[2] - a[3];
    cx<T> y0 = a[0] + b1 + b2;
    cx<T> t1 = mk<T>(fma(c2, b2.x, fma(c1, b1.x, a[0].x)), fma(c2, b2.y, fma(c1, b1.y, a[0].y)));
    cx<T> t2 = mk<T>(fma(c1, b2.x, fma(c2, b1.x, a[0].x)), fma(c1, b2.y, fma(c2, b1.y, a[0].y)));
    cx<T> u1 = mk<T>(fma(s1, d1.x, s2 * d2.x), fma(s1, d1.y, s2 * d2.y));
    cx<T> u2 = mk<T>(fma(s2, d1.x, -(s1 * d2.x)), fma(s2, d1.y, -(s1 * d2.y)));
    u1 = rot4<INV>(u1);  // -i u1 forward
    u2 = rot4<INV>(u2);
    a[0] = y0;
    a[1] = t1 + u1;
    a[4] = t1 - u1;
    a[2] = t2 + u2;
    a[3] = t2 - u2;
  }
};

// cos(2 pi n / R) and sin(2 pi n / R) for n = 1..(R-1)/2 of the odd primes
// 7, 11, 13 (20 significant digits, long double).
template <int R>
struct PrimeTab;
template <>
struct PrimeTab<7> {
  __device__ __forceinline__ static constexpr double c(int n) {
    return n == 1 ? 0.62348980185873355790 : n == 2 ? -0.22252093395631433606 : -0.90096886790241908068;
  }
  __device__ __forceinline__ static constexpr double s(int n) {
    return n == 1 ? 0.78183148246802978688 : n == 2 ? 0.97492791218182362259 : 0.43388373911755821505;
  }
};
template <>
struct PrimeTab<11> {
  __device__ __forceinline__ static constexpr double c(int n) {
    return n == 1   ? 0.84125353283118118090
           : n == 2 ? 0.41541501300188646601
           : n == 3 ? -0.14231483827328507435
           : n == 4 ? -0.65486073394528499680
                    : -0.95949297361449735851;
  }
  __device__ __forceinline__ static constexpr double s(int n) {
    return n == 1   ? 0.54064081745559756340
           : n == 2 ? 0.90963199535451835293
           : n == 3 ? 0.98982144188093274190
           : n == 4 ? 0.75574957435425834209
                    : 0.28173255684142980458;
  }
};
template <>
struct PrimeTab<13> {
  __device__ __forceinline__ static constexpr double c(int n) {
    return n == 1   ? 0.88545602565320990465
           : n == 2 ? 0.56806474673115583350
           : n == 3 ? 0.12053668025532310943
           : n == 4 ? -0.35460488704253555553
           : n == 5 ? -0.74851074817110103613
                    : -0.97094181742605200014;
  }
  __device__ __forceinline__ static constexpr double s(int n) {
    return n == 1   ? 0.46472317204376852899
           : n == 2 ? 0.82298386589365637317
           : n == 3 ? 0.99270887409805398600
           : n == 4 ? 0.93501624268541485018
           : n == 5 ? 0.66312265824079527296
                    : 0.23931566428755787686;
  }
};

// Odd prime R in the symmetric pair form (the radix-5 butterfly above is the
// R = 5 case): with b_j = a_j + a_{R-j}, d_j = a_j - a_{R-j} (j = 1..m,
// m = (R-1)/2),
//   X_0 = a_0 + sum_j b_j,
//   X_k, X_{R-k} = t_k +- W4 u_k,  t_k = a_0 + sum_j cos(2 pi jk/R) b_j,
//                                  u_k = sum_j sin(2 pi jk/R) d_j,
// W4 = -i forward, +i inverse.  (m+1)^2 real multiply-adds per component
// instead of R^2.
template <int R, bool INV, typename T>
struct DftOdd {
  static constexpr int M = (R - 1) / 2;
  // cos / sin of 2 pi n / R for any n, folded into 1..m
  __device__ __forceinline__ static constexpr T cs(int n) {
    const int r = n % R;
    return (T)PrimeTab<R>::c(r <= M ? r : R - r);
  }
  __device__ __forceinline__ static constexpr T sn(int n) {
    const int r = n % R;
    return r <= M ? (T)PrimeTab<R>::s(r) : (T)-PrimeTab<R>::s(R - r);
  }
  __device__ __forceinline__ static void run(cx<T>* a) {
    cx<T> b[M], d[M];
#pragma unroll
    for (int j = 1; j <= M; ++j) {
      b[j - 1] = a[j] + a[R - j];
      d[j - 1] = a[j] - a[R - j];
    }
    const cx<T> a0 = a[0];
    cx<T> y0 = a0;
#pragma unroll
    for (int j = 0; j < M; ++j) y0 = y0 + b[j];
#pragma unroll
    for (int k = 1; k <= M; ++k) {
      cx<T> t = a0;
      cx<T> u = mk<T>((T)0, (T)0);
#pragma unroll
      for (int j = 1; j <= M; ++j) {
        const T c = cs(j * k), s = sn(j * k);
        t = mk<T>(fma(c, b[j - 1].x, t.x), fma(c, b[j - 1].y, t.y));
        u = mk<T>(fma(s, d[j - 1].x, u.x), fma(s, d[j - 1].y, u.y));
      }
      u = rot4<INV>(u);
      a[k] = t + u;
      a[R - k] = t - u;
    }
    a[0] = y0;
  }
};

template <bool INV, typename T>
struct Dft<7, INV, T> : DftOdd<7, INV, T> {};
template <bool INV, typename T>
struct Dft<11, INV, T> : DftOdd<11, INV, T> {};
template <bool INV, typename T>
struct Dft<13, INV, T> : DftOdd<13, INV, T> {};

template <bool INV, typename T>
struct Dft<8, INV, T> {
  __device__ __forceinline__ static void run(cx<T>* a) {
    const T h = (T)0.70710678118654752440084436210484904;  // 1/sqrt(2)
    cx<T> t0 = a[0] + a[4], t1 = a[0] - a[4];
    cx<T> t2 = a[2] + a[6], t3 = rot4<INV>(a[2] - a[6]);
    cx<T> t4 = a[1] + a[5], t5 = a[1] - a[5];
    cx<T> t6 = a[3] + a[7], t7 = rot4<INV>(a[3] - a[7]);
    cx<T> e0 = t0 + t2, e2 = t0 - t2, e1 = t1 + t3, e3 = t1 - t3;
    cx<T> o0 = t4 + t6, o2 = t4 - t6, o1 = t5 + t7, o3 = t5 - t7;
    // W8^1, W8^2, W8^3 of the direction applied to o1, o2, o3
    cx<T> w1, w3;
    if (INV) {
      w1 = mk<T>((o1.x - o1.y) * h, (o1.x + o1.y) * h);
      w3 = mk<T>(-(o3.x + o3.y) * h, (o3.x - o3.y) * h);
    } else {
      w1 = mk<T>((o1.x + o1.y) * h, (o1.y - o1.x) * h);
      w3 = mk<T>((o3.y - o3.x) * h, -(o3.x + o3.y) * h);
    }
    cx<T> w2 = rot4<INV>(o2);
    a[0] = e0 + o0;
    a[4] = e0 - o0;
    a[1] = e1 + w1;
    a[5] = e1 - w1;
    a[2] = e2 + w2;
    a[6] = e2 - w2;
    a[3] = e3 + w3;
    a[7] = e3 - w3;
  }
};

// ---------------------------------------------------------------------------
// Compile-time schedules
// ---------------------------------------------------------------------------
template <int... Rs>
struct Sched {
  static constexpr int S = sizeof...(Rs);
  static constexpr int R[S] = {Rs...};
  static constexpr int N = (Rs * ...);
  static constexpr int p(int s) {
    int v = 1;
    for (int i = 0; i < s; ++i) v *= R[i];
    return v;
  }
  // offset (in complex entries) of stage s's twiddles in the plan table:
  // stage s >= 1 stores p(s) * (R[s]-1) entries, k-major.
  static constexpr int twoff(int s) {
    int o = 0;
    for (int i = 1; i < s; ++i) o += p(i) * (R[i] - 1);
    return o;
  }
};

// Workgroup barrier for LDS traffic only: waits for this wave's LDS
// operations (lgkmcnt(0)), not for its outstanding global loads, so register
// prefetches of the next tile stay in flight across the stage exchanges.
// The empty asm statements keep the compiler from moving memory accesses
// across the barrier.  s_waitcnt immediate (gfx9 encoding): vmcnt = 63,
// expcnt = 7, lgkmcnt = 0.
__device__ __forceinline__ void lds_barrier() {
  __asm__ volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xC07F);
  __builtin_amdgcn_s_barrier();
  __asm__ volatile("" ::: "memory");
}

// Exchange between the threads of one wavefront only: LDS operations of a
// wave complete in program order, so waiting for this wave's own LDS traffic
// (lgkmcnt(0)) and keeping the compiler from moving memory accesses across
// the point is enough; other waves are not waited for.
__device__ __forceinline__ void wave_sync() {
  __asm__ volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xC07F);
  __builtin_amdgcn_wave_barrier();
  __asm__ volatile("" ::: "memory");
}

// Barrier between the stages of the register engine.  BM 0: workgroup
// barrier; 1: workgroup barrier on LDS traffic only (prefetched global loads
// stay in flight); 2: wave-local (every thread sharing the LDS region is in
// the calling thread's wavefront).
template <int BM>
__device__ __forceinline__ void stage_barrier() {
  if constexpr (BM == 2)
    wave_sync();
  else if constexpr (BM == 1)
    lds_barrier();
  else
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Register-resident engine.  TPT = N/P threads share one transform.  At
// stage s (radix R) thread t owns butterflies i = t + TPT*q for q < NB(s) =
// ceil((N/R) / TPT); when TPT*NB(s) exceeds N/R the surplus slots are idle
// (guarded), so any radix order works with any P dividing N.  v[q*R + r] is
// input/output r of butterfly q; PV register slots cover the widest stage.
// When every radix divides P, NB(s) = P/R, PV = P and no guard is emitted.
// After run(), v[q*RL + r] holds X[t + TPT*q + r*N/RL] (valid slots).
// LB selects LDS-only barriers between stages (see lds_barrier).
// ---------------------------------------------------------------------------
template <typename T, bool INV, typename SC, int P, int LB = 0>
struct RegEngine {
  using C = cx<T>;
  static constexpr int N = SC::N;
  static constexpr int TPT = N / P;
  static_assert(N % P == 0, "P must divide N");

  static constexpr int NB(int s) { return (N / SC::R[s] + TPT - 1) / TPT; }
  static constexpr bool EXACT(int s) { return NB(s) * TPT == N / SC::R[s]; }
  static constexpr int pv() {
    int m = 0;
    for (int s = 0; s < SC::S; ++s) m = NB(s) * SC::R[s] > m ? NB(s) * SC::R[s] : m;
    return m;
  }
  static constexpr int PV = pv();
  static constexpr int R0 = SC::R[0];
  static constexpr int RL = SC::R[SC::S - 1];
  static constexpr int NB0 = NB(0);
  static constexpr int NBL = NB(SC::S - 1);

  template <int s>
  __device__ __forceinline__ static bool live(int t, int q) {
    return EXACT(s) || t + TPT * q < N / SC::R[s];
  }
  // slot (t, q) of the first / last stage holds a real butterfly
  __device__ __forceinline__ static bool live0(int t, int q) { return live<0>(t, q); }
  __device__ __forceinline__ static bool liveL(int t, int q) { return live<SC::S - 1>(t, q); }

  template <int s>
  __device__ __forceinline__ static void bfly(C (&v)[PV], int t, const C* __restrict__ tw) {
    constexpr int R = SC::R[s];
    constexpr int p = SC::p(s);
#pragma unroll
    for (int q = 0; q < NB(s); ++q) {
      if (!live<s>(t, q)) continue;
      if constexpr (p > 1) {
        const int i = t + TPT * q;
        const int k = i % p;
        const C* w = tw + SC::twoff(s) + k * (R - 1);
#pragma unroll
        for (int r = 1; r < R; ++r) v[q * R + r] = twmul<INV>(v[q * R + r], w[r - 1]);
      }
      Dft<R, INV, T>::run(&v[q * R]);
    }
  }

  template <int s, typename L>
  __device__ __forceinline__ static void put(C (&v)[PV], int t, L& lds) {
    constexpr int R = SC::R[s];
    constexpr int p = SC::p(s);
#pragma unroll
    for (int q = 0; q < NB(s); ++q) {
      if (!live<s>(t, q)) continue;
      const int i = t + TPT * q;
      const int k = i % p;
      const int j = (i - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) lds.at(j + r * p) = v[q * R + r];
    }
  }

  template <int s, typename L>
  __device__ __forceinline__ static void get(C (&v)[PV], int t, L& lds) {
    constexpr int R = SC::R[s];
#pragma unroll
    for (int q = 0; q < NB(s); ++q) {
      if (!live<s>(t, q)) continue;
      const int i = t + TPT * q;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q * R + r] = lds.at(i + r * (N / R));
    }
  }

  template <int s, typename L>
  __device__ __forceinline__ static void run_from(C (&v)[PV], int t, L& lds, const C* __restrict__ tw) {
    bfly<s>(v, t, tw);
    if constexpr (s + 1 < SC::S) {
      stage_barrier<LB>();
      put<s>(v, t, lds);
      stage_barrier<LB>();
      get<s + 1>(v, t, lds);
      run_from<s + 1>(v, t, lds, tw);
    }
  }

  // v holds stage-0 inputs: v[q*R0 + r] = x[t + TPT*q + r*N/R0] (live slots)
  template <typename L>
  __device__ __forceinline__ static void run(C (&v)[PV], int t, L& lds, const C* __restrict__ tw) {
    run_from<0>(v, t, lds, tw);
  }

  // element index of stage-0 input / last-stage output slot (q, r)
  __device__ __forceinline__ static constexpr int in_index(int t, int q, int r) {
    return t + TPT * q + r * (N / R0);
  }
  __device__ __forceinline__ static constexpr int out_index(int t, int q, int r) {
    return t + TPT * q + r * (N / RL);
  }
};

// ---------------------------------------------------------------------------
// Run-time schedule (any product of 2, 3, 4, 5, 7, 8, 11, 13)
// ---------------------------------------------------------------------------
constexpr int kMaxStages = 16;
struct RtSched {
  int N;
  int S;
  int R[kMaxStages];
  int p[kMaxStages];
  int twoff[kMaxStages];
};

// XCD-aware block remap: hardware hands consecutive workgroup ids to the 8
// XCDs round-robin; remap so that each XCD walks a contiguous range of tiles
// (neighbouring tiles share straddled cache lines in that XCD's L2).
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nblocks) {
  constexpr uint32_t X = 8;
  const uint32_t xcd = b % X, j = b / X;
  const uint32_t base = nblocks / X, rem = nblocks % X;
  return (xcd < rem) ? xcd * (base + 1) + j : rem * (base + 1) + (xcd - rem) * base + j;
}

}  // namespace gfft
