// GPU_FFT.h -- source-level drop-in for the reference's public API
// (GPU_FFT/src/GPU_FFT/GPU_FFT.h:62-127 and transitions/transitions.h),
// implemented header-only on the C ABI of libgpufft.so (gpufft/gpufft.h).
//
// Kept from the reference:
//   * TRANSITIONS::T1_f/T2_f/T1_d/T2_d (float, float2, double, double2;
//     transitions.h:77-80), the blocking memory helpers
//     __Memory_allocation_gpu__ / __Memory_copy_cpu_to_gpu__ /
//     __Memory_copy_gpu_to_cpu__ / __Memory_copy_gpu_to_gpu__ and
//     __Device_synchronize__ (transitions.h:149-208), FFT_FORWARD/FFT_INVERSE.
//   * int64 (FFT_defination.h:44).
//   * GPU_FFT::INIT_GPU_FFT<T1,T2>(Nx,Ny,Nz,procs,rank,comm),
//     GPU_FFT::GPU_FFT_R2C<T1,T2>(T1*), GPU_FFT::GPU_FFT_C2R<T1,T2>(T2*)
//     (GPU_FFT.h:114-121) with the same layouts, in place, unnormalised, on
//     the null stream, returning while the last pass may still run
//     (FFT_R2C.cu:69).  One process-wide plan per precision, as the
//     reference's globals (GPU_FFT.cu:43-87); the dimension and rank globals
//     __Nx__, __Ny__, __Nz__, __procs__, __rank__ are mirrored.
// Changed on purpose:
//   * errors throw GPU_FFT::Error (the reference ignores every status);
//   * a second INIT replaces the plan instead of leaking it, and
//     FINALIZE_GPU_FFT<T1,T2>() releases it;
//   * procs > 1 is accepted by INIT (local share planned) but R2C/C2R throw:
//     the inter-rank exchange (FFT_R2C.cu:51-63) is not part of this library;
//   * MPI is optional: <mpi.h> is used when present (or when
//     GPU_FFT_WITH_MPI is defined), otherwise MPI_Comm is a placeholder
//     type and the communicator argument is ignored, as it is at p = 1.
//     INIT_GPU_FFT_COMM<T1,T2> keeps the reference's bookkeeping (rank,
//     procs, communicator; communications_setup.cu:56-71) and
//     __get_mpi_datatype__ maps T2_f/T2_d to MPI_C_COMPLEX /
//     MPI_C_DOUBLE_COMPLEX (communications_setup.cu:43-53) when <mpi.h> is
//     in use.
#ifndef GPU_FFT_DROPIN_H_
#define GPU_FFT_DROPIN_H_

// The reference defines this macro in its public headers (GPU_FFT.h:40-46,
// transitions.h:39-45) and its callers use it (GPU_FFT_TEST_SCRIPT/test.cu:41).
#ifndef inline_qualifier
#if defined(_MSC_VER)
#define inline_qualifier __inline
#else
#define inline_qualifier inline
#endif
#endif

#include <hip/hip_runtime.h>

#include <cmath>
#include <initializer_list>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "gpufft/gpufft.h"

#if defined(GPU_FFT_WITH_MPI) || (!defined(GPU_FFT_NO_MPI) && defined(__has_include) && __has_include(<mpi.h>))
#include <mpi.h>
#define GPU_FFT_HAVE_MPI 1
#else
typedef int MPI_Comm;
#define GPU_FFT_HAVE_MPI 0
#endif

typedef long long int64;

namespace TRANSITIONS {

typedef float T1_f;
typedef float2 T2_f;
typedef double T1_d;
typedef double2 T2_d;

constexpr int FFT_FORWARD = -1;  // HIPFFT_FORWARD
constexpr int FFT_INVERSE = 1;   // HIPFFT_BACKWARD

inline void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

inline void __Device_synchronize__() { check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

template <typename T>
inline T* __Memory_allocation_gpu__(T* data_pointer, size_t count) {
  check_hip(hipMalloc(&data_pointer, sizeof(T) * count), "hipMalloc");
  return data_pointer;
}

template <typename T>
inline void __Memory_copy_cpu_to_gpu__(T* data_cpu, T* data_gpu, size_t count) {
  check_hip(hipMemcpy(data_gpu, data_cpu, sizeof(T) * count, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

template <typename T>
inline void __Memory_copy_gpu_to_cpu__(T* data_gpu, T* data_cpu, size_t count) {
  check_hip(hipMemcpy(data_cpu, data_gpu, sizeof(T) * count, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

template <typename T>
inline void __Memory_copy_gpu_to_gpu__(T* data_src, T* data_dest, size_t count) {
  check_hip(hipMemcpy(data_dest, data_src, sizeof(T) * count, hipMemcpyDeviceToDevice), "hipMemcpy D2D");
}

}  // namespace TRANSITIONS

namespace GPU_FFT {

class Error : public std::runtime_error {
 public:
  Error(int status, const char* what)
      : std::runtime_error(std::string(what) + ": " + gfft_strerror(status)), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

inline void check(int status, const char* what) {
  if (status != GFFT_OK) throw Error(status, what);
}

inline int64 __Nx__ = 0;
inline int64 __Ny__ = 0;
inline int64 __Nz__ = 0;
inline int __procs__ = 1;
inline int __rank__ = 0;
inline MPI_Comm __MPI_COMMUNICATOR__{};

#if GPU_FFT_HAVE_MPI
// Request slots for the 2(p-1) point-to-point transfers of one exchange
// (communications_setup.cu:67). Owned here (the reference leaks them); this
// library posts no transfers itself, so they stay unused.
inline std::vector<MPI_Request> __requests__;

// MPI type of one complex element (communications_setup.cu:43-53).
template <typename T>
MPI_Datatype __get_mpi_datatype__(T a);
template <>
inline MPI_Datatype __get_mpi_datatype__(float2) { return MPI_C_COMPLEX; }
template <>
inline MPI_Datatype __get_mpi_datatype__(double2) { return MPI_C_DOUBLE_COMPLEX; }
#endif

// Communication bookkeeping, as the reference's INIT_GPU_FFT_COMM
// (communications_setup.cu:56-71; declared at GPU_FFT.h:111-112): records
// the communicator, rank and rank count. INIT_GPU_FFT calls it, as the
// reference's INIT does (GPU_FFT_init.cu:57).
template <typename T1, typename T2>
void INIT_GPU_FFT_COMM(int procs, int rank, MPI_Comm& MPI_COMMUNICATOR_INPUT) {
  if (procs < 1 || rank < 0 || rank >= procs) throw Error(GFFT_ERR_INVALID_ARG, "INIT_GPU_FFT_COMM");
  __MPI_COMMUNICATOR__ = MPI_COMMUNICATOR_INPUT;
  __rank__ = rank;
  __procs__ = procs;
#if GPU_FFT_HAVE_MPI
  __requests__.assign(2 * static_cast<size_t>(procs - 1), MPI_REQUEST_NULL);
#endif
}

namespace detail {

template <typename T1, typename T2>
struct Traits;
template <>
struct Traits<float, float2> {
  static constexpr int prec = GFFT_F32;
};
template <>
struct Traits<double, double2> {
  static constexpr int prec = GFFT_F64;
};

// One plan per precision, like the reference's plan globals.
template <typename T1>
inline gfft_plan& plan_slot() {
  static gfft_plan p = nullptr;
  return p;
}

}  // namespace detail

template <typename T1, typename T2>
void INIT_GPU_FFT(int64 Nx_in, int64 Ny_in, int64 Nz_in, int procs, int rank, MPI_Comm& MPI_COMMUNICATOR_INPUT) {
  gfft_plan& slot = detail::plan_slot<T1>();
  gfft_plan p = nullptr;
  check(gfft_plan_create(&p, Nx_in, Ny_in, Nz_in, procs, rank, detail::Traits<T1, T2>::prec), "INIT_GPU_FFT");
  if (slot) gfft_plan_destroy(slot);
  slot = p;
  __Nx__ = Nx_in;
  __Ny__ = Ny_in;
  __Nz__ = Nz_in;
  INIT_GPU_FFT_COMM<T1, T2>(procs, rank, MPI_COMMUNICATOR_INPUT);
}

template <typename T1, typename T2>
void GPU_FFT_R2C(T1* input_data) {
  gfft_plan p = detail::plan_slot<T1>();
  if (!p) throw Error(GFFT_ERR_INVALID_ARG, "GPU_FFT_R2C before INIT_GPU_FFT");
  check(gfft_exec_r2c(p, static_cast<void*>(input_data), nullptr), "GPU_FFT_R2C");
}

template <typename T1, typename T2>
void GPU_FFT_C2R(T2* input_data) {
  gfft_plan p = detail::plan_slot<T1>();
  if (!p) throw Error(GFFT_ERR_INVALID_ARG, "GPU_FFT_C2R before INIT_GPU_FFT");
  check(gfft_exec_c2r(p, static_cast<void*>(input_data), nullptr), "GPU_FFT_C2R");
}

// Not in the reference (which never frees): releases the precision's plan.
template <typename T1, typename T2>
void FINALIZE_GPU_FFT() {
  gfft_plan& slot = detail::plan_slot<T1>();
  if (slot) gfft_plan_destroy(slot);
  slot = nullptr;
}

}  // namespace GPU_FFT

#endif  // GPU_FFT_DROPIN_H_
