// Shared definitions for the MI355X (gfx950) DMRG sweep engine.
//
// Scalar types: real FP64 (`double`) and complex128 (`cdouble`, two doubles,
// real part first -- the memory layout of hipDoubleComplex / std::complex).
// All matrices are column-major, as in the reference's rocBLAS calls
// (gpu-rocm/dmrg2-gpu/src/dmrg2_gpu_impl.h).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace dmrg {

struct cdouble {
  double x, y;
  cdouble() = default;
  __host__ __device__ cdouble(double r) : x(r), y(0.0) {}
  __host__ __device__ cdouble(double r, double i) : x(r), y(i) {}
};

__host__ __device__ inline cdouble operator+(cdouble a, cdouble b) { return {a.x + b.x, a.y + b.y}; }
__host__ __device__ inline cdouble operator-(cdouble a, cdouble b) { return {a.x - b.x, a.y - b.y}; }
__host__ __device__ inline cdouble operator*(cdouble a, cdouble b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__host__ __device__ inline cdouble operator*(double s, cdouble a) { return {s * a.x, s * a.y}; }
__host__ __device__ inline cdouble operator*(cdouble a, double s) { return {s * a.x, s * a.y}; }
__host__ __device__ inline cdouble& operator+=(cdouble& a, cdouble b) { a.x += b.x; a.y += b.y; return a; }
__host__ __device__ inline cdouble& operator-=(cdouble& a, cdouble b) { a.x -= b.x; a.y -= b.y; return a; }
__host__ __device__ inline cdouble operator-(cdouble a) { return {-a.x, -a.y}; }

__host__ __device__ inline double conj(double a) { return a; }
__host__ __device__ inline cdouble conj(cdouble a) { return {a.x, -a.y}; }
__host__ __device__ inline double real(double a) { return a; }
__host__ __device__ inline double real(cdouble a) { return a.x; }
__host__ __device__ inline double imag(double) { return 0.0; }
__host__ __device__ inline double imag(cdouble a) { return a.y; }
__host__ __device__ inline double abs2(double a) { return a * a; }
__host__ __device__ inline double abs2(cdouble a) { return a.x * a.x + a.y * a.y; }

template <class T> struct is_cplx { static constexpr bool value = false; };
template <> struct is_cplx<cdouble> { static constexpr bool value = true; };

template <class T> __host__ __device__ inline T make_scalar(double r, double i);
template <> __host__ __device__ inline double make_scalar<double>(double r, double) { return r; }
template <> __host__ __device__ inline cdouble make_scalar<cdouble>(double r, double i) { return {r, i}; }

// Transposition operators for GEMM operands.
enum Op : int { OP_N = 0, OP_T = 1, OP_C = 2, OP_H = 3 };  // C: conj only, H: conj-transpose
__host__ __device__ constexpr bool op_trans(int op) { return op == OP_T || op == OP_H; }
__host__ __device__ constexpr bool op_conj(int op) { return op == OP_C || op == OP_H; }

// Status codes shared with the C-ABI (include/dmrg_amd.h).
enum Status : int {
  ST_OK = 0,
  ST_HIP = 1,
  ST_SOLVER_NO_CONV = 2,
  ST_BAD_ARG = 3,
  ST_NONFINITE = 4,
  ST_STATE = 5,
  ST_OOM = 6,
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define DMRG_HIP(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      throw ::dmrg::Error(::dmrg::ST_HIP, std::string(#call) + ": " +               \
                                          hipGetErrorString(e_) + " at " +          \
                                          __FILE__ + ":" + std::to_string(__LINE__)); \
  } while (0)

#define DMRG_REQUIRE(cond, msg)                                                     \
  do {                                                                              \
    if (!(cond)) throw ::dmrg::Error(::dmrg::ST_BAD_ARG, std::string(msg));         \
  } while (0)

// Check the launch of the kernel just queued.
#define DMRG_LAUNCH_CHECK() DMRG_HIP(hipGetLastError())

inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace dmrg
