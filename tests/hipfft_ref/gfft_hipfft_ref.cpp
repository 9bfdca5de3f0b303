// TEST-ONLY comparator: the reference's own arithmetic at p = 1.
//
// GPU_FFT's transform is a sequence of hipFFT executions with the plan
// geometry of GPU_FFT/src/initialize_GPU_FFT/GPU_FFT_init.cu:59-91:
//   R2C/C2R : rank 2, n = {Ny, Nz}, NULL embed (packed, in-place padded
//             rows), batch Nx/p                          (GPU_FFT_init.cu:61-64,90-91)
//   C2C     : rank 1, n = {Nx}, istride = ostride = (Ny/p)*K,
//             idist = odist = 1, batch (Ny/p)*K          (GPU_FFT_init.cu:67-73,89)
// and the call order of FFT_R2C.cu:45-69 / FFT_C2R.cu:45-69.  At p = 1 the
// slab and chunk transposes are identity permutations and the exchange is a
// full-array device copy (FFT_R2C.cu:62), so the sequence below computes what
// the reference computes: R2C in place, copy to the work buffer, C2C buffer ->
// data (forward); C2C data -> buffer, copy back, C2R in place (inverse).
// Built by __graft_entry__.build(); never linked into libgpufft.so.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>
#include <stdint.h>

namespace {

struct Plans {
  hipfftHandle r2c = nullptr, c2r = nullptr, c2c = nullptr;
  void* buffer = nullptr;
  ~Plans() {
    if (r2c) hipfftDestroy(r2c);
    if (c2r) hipfftDestroy(c2r);
    if (c2c) hipfftDestroy(c2c);
    if (buffer) (void)hipFree(buffer);
  }
};

int make(Plans& p, int64_t nx, int64_t ny, int64_t nz, int f64) {
  const int64_t K = nz / 2 + 1;
  int n2[2] = {(int)ny, (int)nz};
  int n1[1] = {(int)nx};
  int emb[1] = {(int)nx};
  size_t ws = 0;
  const size_t bytes = (size_t)nx * ny * K * (f64 ? 16 : 8);
  if (hipMalloc(&p.buffer, bytes) != hipSuccess) return 1000;
  if (hipfftCreate(&p.r2c) || hipfftCreate(&p.c2r) || hipfftCreate(&p.c2c)) return 1001;
  int st = hipfftMakePlanMany(p.c2c, 1, n1, emb, (int)(ny * K), 1, emb, (int)(ny * K), 1,
                              f64 ? HIPFFT_Z2Z : HIPFFT_C2C, (int)(ny * K), &ws);
  if (st) return st;
  st = hipfftMakePlanMany(p.r2c, 2, n2, nullptr, 1, 1, nullptr, 1, 1, f64 ? HIPFFT_D2Z : HIPFFT_R2C, (int)nx, &ws);
  if (st) return st;
  st = hipfftMakePlanMany(p.c2r, 2, n2, nullptr, 1, 1, nullptr, 1, 1, f64 ? HIPFFT_Z2D : HIPFFT_C2R, (int)nx, &ws);
  return st;
}

}  // namespace

extern "C" int gref_r2c(int64_t nx, int64_t ny, int64_t nz, int f64, void* data) {
  Plans p;
  int st = make(p, nx, ny, nz, f64);
  if (st) return st;
  const size_t bytes = (size_t)nx * ny * (nz / 2 + 1) * (f64 ? 16 : 8);
  if (f64) {
    st = hipfftExecD2Z(p.r2c, (hipfftDoubleReal*)data, (hipfftDoubleComplex*)data);
    if (!st && hipMemcpy(p.buffer, data, bytes, hipMemcpyDeviceToDevice) != hipSuccess) st = 1002;
    if (!st) st = hipfftExecZ2Z(p.c2c, (hipfftDoubleComplex*)p.buffer, (hipfftDoubleComplex*)data, HIPFFT_FORWARD);
  } else {
    st = hipfftExecR2C(p.r2c, (hipfftReal*)data, (hipfftComplex*)data);
    if (!st && hipMemcpy(p.buffer, data, bytes, hipMemcpyDeviceToDevice) != hipSuccess) st = 1002;
    if (!st) st = hipfftExecC2C(p.c2c, (hipfftComplex*)p.buffer, (hipfftComplex*)data, HIPFFT_FORWARD);
  }
  if (!st && hipDeviceSynchronize() != hipSuccess) st = 1003;
  return st;
}

extern "C" int gref_c2r(int64_t nx, int64_t ny, int64_t nz, int f64, void* data) {
  Plans p;
  int st = make(p, nx, ny, nz, f64);
  if (st) return st;
  const size_t bytes = (size_t)nx * ny * (nz / 2 + 1) * (f64 ? 16 : 8);
  if (f64) {
    st = hipfftExecZ2Z(p.c2c, (hipfftDoubleComplex*)data, (hipfftDoubleComplex*)p.buffer, HIPFFT_BACKWARD);
    if (!st && hipMemcpy(data, p.buffer, bytes, hipMemcpyDeviceToDevice) != hipSuccess) st = 1002;
    if (!st) st = hipfftExecZ2D(p.c2r, (hipfftDoubleComplex*)data, (hipfftDoubleReal*)data);
  } else {
    st = hipfftExecC2C(p.c2c, (hipfftComplex*)data, (hipfftComplex*)p.buffer, HIPFFT_BACKWARD);
    if (!st && hipMemcpy(data, p.buffer, bytes, hipMemcpyDeviceToDevice) != hipSuccess) st = 1002;
    if (!st) st = hipfftExecC2R(p.c2r, (hipfftComplex*)data, (hipfftReal*)data);
  }
  if (!st && hipDeviceSynchronize() != hipSuccess) st = 1003;
  return st;
}
