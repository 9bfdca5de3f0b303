// C++ counterpart of the reference harness GPU_FFT_TEST_SCRIPT/test.cu:86-178,
// written against the drop-in header include/GPU_FFT/GPU_FFT.h (p = 1).
//
//   gpu_fft_test Nx [Ny Nz] [f32|f64]
//
// Fills f = 8 (sin x sin 2y sin 3z + sin 4x sin 5y sin 6z) with zero pads
// (test.cu:118-136), runs GPU_FFT_R2C, normalises by 1/N with a kernel
// (test.cu:53-65,149), prints every bin with |re| or |im| > 0.1 in the
// reference's format (test.cu:67-84), runs GPU_FFT_C2R and prints the
// max abs round-trip error.  Unlike test.cu it accepts non-cubic grids, checks
// only z < Nz (test.cu:166 also reads pad slots), and exits non-zero unless
// exactly the 8 expected bins appear and the error is below a bound.
#include <GPU_FFT/GPU_FFT.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

template <typename T2>
__global__ void normalize(T2* data, int64 n_complex, double inv_n) {
  const int64 i = (int64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_complex) {
    T2 v = data[i];
    v.x *= inv_n;
    v.y *= inv_n;
    data[i] = v;
  }
}

template <typename T1, typename T2>
int run(int64 Nx, int64 Ny, int64 Nz, double tol) {
  const int64 K = Nz / 2 + 1;
  const int64 nreal = Nx * Ny * 2 * K;
  const double pi = M_PI, dx = 2 * pi / Nx, dy = 2 * pi / Ny, dz = 2 * pi / Nz;
  std::vector<T1> in(nreal, T1(0)), out(nreal);
  for (int64 i = 0; i < Nx; ++i)
    for (int64 j = 0; j < Ny; ++j)
      for (int64 k = 0; k < Nz; ++k)
        in[(i * Ny + j) * 2 * K + k] =
            (T1)(8 * (std::sin(1.0 * i * dx) * std::sin(2.0 * j * dy) * std::sin(3.0 * k * dz) +
                      std::sin(4.0 * i * dx) * std::sin(5.0 * j * dy) * std::sin(6.0 * k * dz)));
  T2* d = nullptr;
  d = TRANSITIONS::__Memory_allocation_gpu__(d, (size_t)(Nx * Ny * K));
  TRANSITIONS::__Memory_copy_cpu_to_gpu__((T2*)in.data(), d, (size_t)(Nx * Ny * K));
  MPI_Comm comm{};
  GPU_FFT::INIT_GPU_FFT<T1, T2>(Nx, Ny, Nz, 1, 0, comm);
  GPU_FFT::GPU_FFT_R2C<T1, T2>((T1*)d);
  TRANSITIONS::__Device_synchronize__();
  const int64 nc = Nx * Ny * K;
  hipLaunchKernelGGL(normalize<T2>, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, 0, d, nc,
                     1.0 / (double)(Nx * Ny * Nz));
  TRANSITIONS::__Device_synchronize__();
  std::vector<T2> spec(nc);
  TRANSITIONS::__Memory_copy_gpu_to_cpu__(d, spec.data(), (size_t)nc);
  int npeaks = 0, bad_peaks = 0;
  for (int64 x = 0; x < Nx; ++x)
    for (int64 y = 0; y < Ny; ++y)
      for (int64 z = 0; z < K; ++z) {
        const T2 v = spec[(x * Ny + y) * K + z];
        if (std::fabs(v.x) > 0.1 || std::fabs(v.y) > 0.1) {
          std::printf("\n rank = %d , data at (%lld,%lld,%lld)  ------>  %f,%f", 0, x, y, z, (double)v.x, (double)v.y);
          ++npeaks;
          // expected: +-i at (+-1,+-2,3) and (+-4,+-5,6), sign sx*sy
          const int64 ax = x <= Nx / 2 ? x : x - Nx, ay = y <= Ny / 2 ? y : y - Ny;
          const bool m1 = std::llabs(ax) == 1 && std::llabs(ay) == 2 && z == 3;
          const bool m2 = std::llabs(ax) == 4 && std::llabs(ay) == 5 && z == 6;
          const double want = (ax > 0) == (ay > 0) ? 1.0 : -1.0;
          if (!(m1 || m2) || std::fabs(v.x) > tol * 1e3 || std::fabs(v.y - want) > tol * 1e3) ++bad_peaks;
        }
      }
  GPU_FFT::GPU_FFT_C2R<T1, T2>(d);
  TRANSITIONS::__Device_synchronize__();
  TRANSITIONS::__Memory_copy_gpu_to_cpu__((T2*)d, (T2*)out.data(), (size_t)nc);
  double max_err = 0;
  for (int64 i = 0; i < Nx * Ny; ++i)
    for (int64 k = 0; k < Nz; ++k) {
      const double e = std::fabs((double)in[i * 2 * K + k] - (double)out[i * 2 * K + k]);
      if (e > max_err) max_err = e;
    }
  std::printf("\n max error = %g\n", max_err);
  GPU_FFT::FINALIZE_GPU_FFT<T1, T2>();
  (void)hipFree(d);
  const bool ok = npeaks == 8 && bad_peaks == 0 && max_err < tol;
  std::printf("%s: %d peaks (%d unexpected), max error %g (bound %g)\n", ok ? "PASS" : "FAIL", npeaks, bad_peaks,
              max_err, tol);
  return ok ? 0 : 1;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s Nx [Ny Nz] [f32|f64]\n", argv[0]);
    return 2;
  }
  const int64 Nx = std::atoll(argv[1]);
  const int64 Ny = argc > 3 ? std::atoll(argv[2]) : Nx;
  const int64 Nz = argc > 3 ? std::atoll(argv[3]) : Nx;
  const bool f32 = argc > 2 && std::strcmp(argv[argc - 1], "f32") == 0;
  std::printf("\n Nx, Ny, Nz = %lld,%lld,%lld (%s)", Nx, Ny, Nz, f32 ? "f32" : "f64");
  try {
    return f32 ? run<float, float2>(Nx, Ny, Nz, 2e-5) : run<double, double2>(Nx, Ny, Nz, 3e-14);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 3;
  }
}
