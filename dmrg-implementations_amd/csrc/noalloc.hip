// Left-environment update that allocates nothing (dmrg_op_env_left_noalloc).
//
// The kernels are those of dmrg_op_env_left (env_left in contract.hip: GEMM,
// MPO-bond row kernel, GEMM).  Their two intermediates X and Y, D*d*ci*co
// scalars each, live in the file-scope device array `scratch` below instead
// of context memory, so a call makes no hipMalloc/hipFree and needs no
// context (creating one allocates).
//
// `scratch` names a host-side stand-in in host code, not a device address.
// Its address on a device is looked up once per device with
// hipGetSymbolAddress, with that device current, and only that pointer is
// handed to kernels.  Each GPU holds its own copy of the array until the
// process ends.  Calls on one GPU share it: calls that overlap on one GPU
// (two streams, or two threads) race.  Nothing needs clearing between calls,
// since X and Y are written in full before they are read.
#include <mutex>

#include "../../include/dmrg_amd.h"
#include "contract.hpp"

namespace dmrg {
namespace {

constexpr size_t kScratchBytes = (size_t)DMRG_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};  // device address of `scratch` per device (null: not looked up)

// Device address of `scratch` on the current device; looked up on the first
// call per device.
void* scratch_ptr() {
  int dev = -1;
  DMRG_HIP(hipGetDevice(&dev));
  DMRG_REQUIRE(dev >= 0 && dev < kMaxDevices, "noalloc: device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    DMRG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    DMRG_REQUIRE(p != nullptr, "noalloc: no device address for the scratch array");
    g_ptr[dev] = p;
  }
  return g_ptr[dev];
}

template <class T>
void env_left_noalloc(hipStream_t st, const T* L, const T* A, const T* W, T* Lnew, int ci, int co,
                      int d, int D) {
  const size_t per = (size_t)D * d * ci * co;  // scalars of X, and of Y
  // Y starts at the first 256-byte boundary after X
  const size_t xbytes = (per * sizeof(T) + 255) & ~(size_t)255;
  if (xbytes + per * sizeof(T) > kScratchBytes)
    throw Error(ST_BAD_ARG, "dmrg_op_env_left_noalloc: the intermediates (2 x D*d*ci*co scalars, " +
                                std::to_string(xbytes + per * sizeof(T)) + " bytes) exceed the " +
                                std::to_string(kScratchBytes) + "-byte scratch");
  unsigned char* base = static_cast<unsigned char*>(scratch_ptr());
  ContractScratch<T> sc;
  sc.X = reinterpret_cast<T*>(base);
  sc.Y = reinterpret_cast<T*>(base + xbytes);
  sc.cap = per;
  env_left<T>(st, L, A, W, Lnew, ci, co, d, D, sc);
}

}  // namespace

void env_left_noalloc_op(hipStream_t st, int dtype, int ci, int co, int d, int D, const void* L,
                         const void* A, const void* W, void* Lnew) {
  DMRG_REQUIRE(ci >= 1 && co >= 1 && d >= 1 && D >= 1, "dmrg_op_env_left_noalloc: bad shape");
  if (dtype == DMRG_DTYPE_F64)
    env_left_noalloc<double>(st, (const double*)L, (const double*)A, (const double*)W,
                             (double*)Lnew, ci, co, d, D);
  else if (dtype == DMRG_DTYPE_C128)
    env_left_noalloc<cdouble>(st, (const cdouble*)L, (const cdouble*)A, (const cdouble*)W,
                              (cdouble*)Lnew, ci, co, d, D);
  else
    throw Error(ST_BAD_ARG, "dmrg_op_env_left_noalloc: bad dtype");
}

}  // namespace dmrg
