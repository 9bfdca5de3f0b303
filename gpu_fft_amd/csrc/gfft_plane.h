// Fused z + y passes for x-planes that fit in LDS (small grids).
//
// One workgroup owns one x-plane: Ny rows of Nz reals (pitch 2K) = Ny x K
// complex half spectra, contiguous in global memory.  R2C:
//   0. the plane is copied into LDS (one coalesced read),
//   1. z R2C of each row pair in LDS (two-for-one, the register engine's
//      exchange runs inside the pair's own two rows),
//   2. y C2C of each kz column in LDS (exchange through a column scratch),
//   3. the plane is copied back (one coalesced write).
// C2R runs 0, y C2C inverse, z C2R, and writes the Nz reals of every row.
// The arithmetic (schedules, twiddles, split / Hermitian build, scale) is
// that of the separate z and y passes, so the results agree to rounding (the
// compiler may contract multiply-adds differently in the two kernels); the
// plane makes one HBM round trip instead of two.
#pragma once
#include "gfft_kernels.h"

namespace gfft {

// Exchange slots of a z transform inside its pair's own rows (N entries,
// no room for padding): point n lives at n ^ ((n >> 3) & 7), a permutation
// inside each block of 8 that spreads the stride-8 accesses of the first
// stages over the LDS banks (N is a multiple of 8 for every plane size).
template <typename T>
struct FlatLds {
  cx<T>* b;
  __device__ __forceinline__ cx<T>& at(int n) { return b[n ^ ((n >> 3) & 7)]; }
};

// y exchange inside the plane's own column: entry n of
// the column starting at b, s entries apart
template <typename T>
struct StrideLds {
  cx<T>* b;
  int s;
  __device__ __forceinline__ cx<T>& at(int n) { return b[n * s]; }
};

// The y exchange runs in the columns themselves (the engine's barrier before
// its first exchange write follows the column loads, and one more barrier
// precedes the output writes), so no Ny x CW column scratch is needed: only a
// tail of max(Ny, K) entries follows the plane.  Lanes past the last column
// exchange there, and an odd Ny's last z row runs its exchange into it.  The
// LDS a scratch would take goes to wider column groups (more threads).
// BMY: the y stage's barriers (0: workgroup barriers; 1: LDS-only barriers,
// for loops that keep the next plane's global loads in flight across them;
// 2: one wavefront per plane, NT = 64: lane ids, wavefront-level barriers,
// so each wavefront of a workgroup can run a plane of its own).
template <typename T, typename SCZ, int PZ, typename SCY, int PY, int CW, int BMY = 0>
struct PlanePass {
  using C = cx<T>;
  static constexpr int NZ = SCZ::N, NY = SCY::N, K = NZ / 2 + 1;
  using EZF = RegEngine<T, false, SCZ, PZ, 2>;
  using EZI = RegEngine<T, true, SCZ, PZ, 2>;
  using EYF = RegEngine<T, false, SCY, PY, BMY>;
  using EYI = RegEngine<T, true, SCY, PY, BMY>;
  static constexpr int TPTZ = NZ / PZ, TPTY = NY / PY;
  static constexpr int NT = CW * TPTY;
  static constexpr int GZ = NT / TPTZ;  // row pairs in flight
  static constexpr int NPAIRS = (NY + 1) / 2;
  static constexpr int PLANE = NY * K;  // complex entries of one x-plane
  static constexpr int SCRATCH = NY > K ? NY : K;  // the exchange tail
  static constexpr size_t LDS_BYTES = (size_t)(PLANE + SCRATCH) * sizeof(C);
  static_assert(64 % TPTZ == 0, "the z exchange must stay inside one wavefront");
  static_assert(NT % TPTZ == 0 && NT <= 1024, "thread count");
  // the z exchange of an odd Ny's last (single-row) pair runs past the plane
  // end by K - 2 (even Nz) entries, into the column scratch
  static_assert(SCRATCH >= K, "column scratch");
  static_assert(LDS_BYTES <= 160 * 1024, "plane does not fit in LDS");
  static_assert(NZ % 8 == 0, "FlatLds permutes inside blocks of 8 points");
  static_assert(BMY != 2 || NT == 64, "one wavefront per plane");

  // the thread's index in the plane's thread group
  __device__ __forceinline__ static int lid() { return BMY == 2 ? (int)(threadIdx.x & 63u) : (int)threadIdx.x; }

  // all of a thread's loads are issued before its LDS stores (U in flight)
  __device__ __forceinline__ static void copy_in(C* plane, const C* src) {
    constexpr int U = (PLANE + NT - 1) / NT < 16 ? (PLANE + NT - 1) / NT : 16;
    for (int i0 = 0; i0 < PLANE; i0 += U * NT) {
      C r[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = i0 + u * NT + lid();
        if (i < PLANE) r[u] = ld_stream(src + i);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = i0 + u * NT + lid();
        if (i < PLANE) plane[i] = r[u];
      }
    }
  }

  // z R2C of every row pair of the LDS plane, in place
  __device__ __forceinline__ static void z_r2c(C* plane, const C* __restrict__ tw) {
    const int gi = lid() / TPTZ, t = lid() % TPTZ;
    constexpr int NI = (K + TPTZ - 1) / TPTZ;
    for (int pair = gi; pair < NPAIRS; pair += GZ) {
      const int ra = 2 * pair;
      const bool vb = ra + 1 < NY;
      T* A = reinterpret_cast<T*>(plane) + (size_t)ra * 2 * K;
      const T* B = A + 2 * K;
      C v[EZF::PV];
#pragma unroll
      for (int q = 0; q < EZF::NB0; ++q) {
        if (!EZF::live0(t, q)) continue;
#pragma unroll
        for (int r = 0; r < EZF::R0; ++r) {
          const int n = EZF::in_index(t, q, r);
          v[q * EZF::R0 + r] = mk<T>(A[n], vb ? B[n] : (T)0);
        }
      }
      FlatLds<T> L{reinterpret_cast<C*>(A)};
      EZF::run(v, t, L, tw);
      stage_barrier<2>();
#pragma unroll
      for (int q = 0; q < EZF::NBL; ++q) {
        if (!EZF::liveL(t, q)) continue;
#pragma unroll
        for (int r = 0; r < EZF::RL; ++r) L.at(EZF::out_index(t, q, r)) = v[q * EZF::RL + r];
      }
      stage_barrier<2>();
      const T h = (T)0.5;
      C xa[NI], xb[NI];
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const int k = t + i * TPTZ;
        if ((NI * TPTZ == K) || k < K) {
          const C zk = L.at(k), zm = L.at(k == 0 ? 0 : NZ - k);
          xa[i] = mk<T>((zk.x + zm.x) * h, (zk.y - zm.y) * h);
          xb[i] = mk<T>((zk.y + zm.y) * h, (zm.x - zk.x) * h);
        }
      }
      stage_barrier<2>();
      C* Ac = reinterpret_cast<C*>(A);
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const int k = t + i * TPTZ;
        if ((NI * TPTZ == K) || k < K) {
          Ac[k] = xa[i];
          if (vb) Ac[K + k] = xb[i];
        }
      }
    }
  }

  // z C2R of every row pair of the LDS plane, in place (reals, pitch 2K)
  __device__ __forceinline__ static void z_c2r(C* plane, const C* __restrict__ tw, T sc, bool scaled) {
    using RC = RowsC2R<T, SCZ, PZ, 1>;
    const int gi = lid() / TPTZ, t = lid() % TPTZ;
    for (int pair = gi; pair < NPAIRS; pair += GZ) {
      const int ra = 2 * pair;
      const bool vb = ra + 1 < NY;
      C* S = plane + (size_t)ra * K;  // half spectra of rows ra, ra + 1
      C v[EZI::PV];
#pragma unroll
      for (int q = 0; q < EZI::NB0; ++q) {
        if (!EZI::live0(t, q)) continue;
#pragma unroll
        for (int r = 0; r < EZI::R0; ++r) {
          const int n = EZI::in_index(t, q, r);
          if (vb) {
            v[q * EZI::R0 + r] = RC::hermitian_z(S, n);
          } else {  // a single row: its partner spectrum is zero
            const bool lo = 2 * n <= NZ;
            C xa = S[lo ? n : NZ - n];
            if (n == 0 || 2 * n == NZ) xa.y = (T)0;
            v[q * EZI::R0 + r] = lo ? xa : mk<T>(xa.x, -xa.y);
          }
        }
      }
      FlatLds<T> L{S};
      EZI::run(v, t, L, tw);
      stage_barrier<2>();  // every read of the pair's exchange region is done
      T* A = reinterpret_cast<T*>(S);
      T* B = A + 2 * K;
#pragma unroll
      for (int q = 0; q < EZI::NBL; ++q) {
        if (!EZI::liveL(t, q)) continue;
#pragma unroll
        for (int r = 0; r < EZI::RL; ++r) {
          const int n = EZI::out_index(t, q, r);
          C z = v[q * EZI::RL + r];
          if (scaled) z = mk<T>(z.x * sc, z.y * sc);
          A[n] = z.x;
          if (vb) B[n] = z.y;
        }
      }
    }
  }

  // y C2C of every kz column of the LDS plane, in place.  The DC and Nyquist
  // columns (kz = 0 and NZ / 2) are transformed together, as one complex
  // column, so NZ / 2 columns take the rounds of CW instead of NZ / 2 + 1:
  //   R2C: both columns are real after the z pass (its two-for-one split
  //        leaves their imaginary parts exactly zero), so column 0 takes
  //        c = x0 + i xN; after the transform the two spectra are the
  //        Hermitian and anti-Hermitian parts of C: X0[k] = (C[k] +
  //        conj(C[-k])) / 2, XN[k] = (C[k] - conj(C[-k])) / 2i;
  //   C2R: the z pass reads only the real parts of the two columns (numpy's
  //        irfft convention), which are the inverse transforms of their
  //        Hermitian parts h0, hN; column 0 takes h0 + i hN, whose inverse
  //        transform has x0 as its real part and xN as its imaginary part.
  template <bool INV>
  __device__ __forceinline__ static void y_pass(C* plane, C* scratch, const C* __restrict__ tw, T sc, bool scaled) {
    using E = RegEngine<T, INV, SCY, PY, BMY>;
    constexpr int H = NZ / 2;  // the Nyquist column; columns 1 .. H - 1 and the packed column 0
    const T half = (T)0.5;
    if constexpr (INV) {
      // Hermitian parts of columns 0 and H, packed into column 0 (pairs k, -k)
      for (int p = lid(); p <= NY / 2; p += NT) {
        const int k1 = p, k2 = (NY - p) % NY;
        const C a1 = plane[k1 * K], a2 = plane[k2 * K], b1 = plane[k1 * K + H], b2 = plane[k2 * K + H];
        const C h0 = mk<T>((a1.x + a2.x) * half, (a1.y - a2.y) * half);  // h0[k2] = conj(h0[k1])
        const C hn = mk<T>((b1.x + b2.x) * half, (b1.y - b2.y) * half);
        plane[k1 * K] = mk<T>(h0.x - hn.y, h0.y + hn.x);
        if (k2 != k1) plane[k2 * K] = mk<T>(h0.x + hn.y, hn.x - h0.y);
      }
      stage_barrier<BMY>();
    }
    const int c = lid() % CW, t = lid() / CW;
    for (int k0 = 0; k0 < H; k0 += CW) {
      const int k = k0 + c;
      const bool valid = k < H;
      C v[E::PV];
#pragma unroll
      for (int q = 0; q < E::NB0; ++q) {
        if (!E::live0(t, q)) continue;
#pragma unroll
        for (int r = 0; r < E::R0; ++r) {
          const int n = E::in_index(t, q, r);
          if (!valid)
            v[q * E::R0 + r] = mk<T>(0, 0);
          else if (!INV && k == 0)
            v[q * E::R0 + r] = mk<T>(plane[n * K].x, plane[n * K + H].x);
          else
            v[q * E::R0 + r] = plane[n * K + k];
        }
      }
      StrideLds<T> L{valid ? plane + k : scratch, valid ? K : 1};
      E::run(v, t, L, tw);
      stage_barrier<BMY>();  // every exchange read of the columns is done
      if (valid) {
#pragma unroll
        for (int q = 0; q < E::NBL; ++q) {
          if (!E::liveL(t, q)) continue;
#pragma unroll
          for (int r = 0; r < E::RL; ++r) {
            const int n = E::out_index(t, q, r);
            C z = v[q * E::RL + r];
            if (k == 0) {
              if (INV) {  // x0 and xN, each real
                plane[n * K] = mk<T>(z.x, (T)0);
                plane[n * K + H] = mk<T>(z.y, (T)0);
              } else {  // C, split below
                plane[n * K] = z;
              }
              continue;
            }
            if (scaled) z = mk<T>(z.x * sc, z.y * sc);
            plane[n * K + k] = z;
          }
        }
      }
    }
    if constexpr (!INV) {
      stage_barrier<BMY>();  // the packed column's transform is complete
      const T f = scaled ? sc * half : half;
      for (int p = lid(); p <= NY / 2; p += NT) {
        const int k1 = p, k2 = (NY - p) % NY;
        const C c1 = plane[k1 * K], c2 = plane[k2 * K];
        const C x0 = mk<T>((c1.x + c2.x) * f, (c1.y - c2.y) * f);  // X0[k2] = conj(X0[k1])
        const C xn = mk<T>((c1.y + c2.y) * f, (c2.x - c1.x) * f);  // XN[k2] = conj(XN[k1])
        plane[k1 * K] = x0;
        plane[k1 * K + H] = xn;
        if (k2 != k1) {
          plane[k2 * K] = mk<T>(x0.x, -x0.y);
          plane[k2 * K + H] = mk<T>(xn.x, -xn.y);
        }
      }
    }
  }
};

template <typename T, typename SCZ, int PZ, typename SCY, int PY, int CW>
__global__ __attribute__((amdgpu_flat_work_group_size(1, PlanePass<T, SCZ, PZ, SCY, PY, CW>::NT))) void
k_plane_r2c(T* data, const cx<T>* twz, const cx<T>* twy, double yscale) {
  using K_ = PlanePass<T, SCZ, PZ, SCY, PY, CW>;
  __shared__ cx<T> smem[K_::PLANE + K_::SCRATCH];
  cx<T>* g = reinterpret_cast<cx<T>*>(data) + (size_t)blockIdx.x * K_::PLANE;
  K_::copy_in(smem, g);
  __syncthreads();
  K_::z_r2c(smem, twz);
  __syncthreads();
  K_::template y_pass<false>(smem, smem + K_::PLANE, twy, (T)yscale, yscale != 1.0);
  __syncthreads();
  for (int i = threadIdx.x; i < K_::PLANE; i += K_::NT) st_stream(g + i, smem[i]);
}

template <typename T, typename SCZ, int PZ, typename SCY, int PY, int CW>
__global__ __attribute__((amdgpu_flat_work_group_size(1, PlanePass<T, SCZ, PZ, SCY, PY, CW>::NT))) void
k_plane_c2r(T* data, const cx<T>* twz, const cx<T>* twy, double zscale) {
  using K_ = PlanePass<T, SCZ, PZ, SCY, PY, CW>;
  __shared__ cx<T> smem[K_::PLANE + K_::SCRATCH];
  cx<T>* g = reinterpret_cast<cx<T>*>(data) + (size_t)blockIdx.x * K_::PLANE;
  K_::copy_in(smem, g);
  __syncthreads();
  K_::template y_pass<true>(smem, smem + K_::PLANE, twy, (T)1, false);
  __syncthreads();
  K_::z_c2r(smem, twz, (T)zscale, zscale != 1.0);
  __syncthreads();
  // the Nz reals of every row, as Nz / 2 pairs (the two pad reals of a row
  // are unspecified after a C2R, as in the reference; here they keep the x
  // pass's output)
  constexpr int H = K_::NZ / 2;
  for (int i = threadIdx.x; i < K_::NY * H; i += K_::NT) {
    const int y = i / H, m = i - y * H;
    st_stream(g + y * K_::K + m, smem[y * K_::K + m]);
  }
}

template <typename T, typename SCZ, int PZ, typename SCY, int PY, int CW, bool INV>
hipError_t launch_plane(void* data, int64_t nplanes, double scale, const void* twz, const void* twy,
                        hipStream_t s) {
  using K_ = PlanePass<T, SCZ, PZ, SCY, PY, CW>;
  if (nplanes <= 0) return hipSuccess;
  clear_launch_error();
  if (INV)
    hipLaunchKernelGGL((k_plane_c2r<T, SCZ, PZ, SCY, PY, CW>), dim3((uint32_t)nplanes), dim3(K_::NT), 0, s,
                       static_cast<T*>(data), static_cast<const cx<T>*>(twz), static_cast<const cx<T>*>(twy),
                       scale);
  else
    hipLaunchKernelGGL((k_plane_r2c<T, SCZ, PZ, SCY, PY, CW>), dim3((uint32_t)nplanes), dim3(K_::NT), 0, s,
                       static_cast<T*>(data), static_cast<const cx<T>*>(twz), static_cast<const cx<T>*>(twy),
                       scale);
  return hipGetLastError();
}

template <typename T, typename SCZ, int PZ, typename SCY, int PY, int CW>
PlaneEntry entry_plane() {
  return PlaneEntry{prec_of<T>(),
                    SCZ::N,
                    SCY::N,
                    make_sched<SCZ>(),
                    make_sched<SCY>(),
                    PlanePass<T, SCZ, PZ, SCY, PY, CW>::NT,
                    reinterpret_cast<void*>(&launch_plane<T, SCZ, PZ, SCY, PY, CW, false>),
                    reinterpret_cast<void*>(&launch_plane<T, SCZ, PZ, SCY, PY, CW, true>)};
}

}  // namespace gfft
