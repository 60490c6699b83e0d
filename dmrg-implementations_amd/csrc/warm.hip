// Warm start of the truncation SVD (see warm.hpp).
#include "warm.hpp"
#include "contract.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace dmrg {
namespace {

// Row and column of flat index e of a column-major matrix with n rows, in
// 32-bit arithmetic (every matrix these kernels index has at most 2048^2
// entries; a 64-bit division is a long software sequence on the GPU).
__device__ inline int mod32(size_t e, int n) { return (int)((unsigned)e % (unsigned)n); }
__device__ inline int div32(size_t e, int n) { return (int)((unsigned)e / (unsigned)n); }

__device__ inline double absv(double a) { return fabs(a); }
__device__ inline double absv(cdouble a) { return sqrt(a.x * a.x + a.y * a.y); }

// Gershgorin bounds of the Hermitian P = Z^H Z, one wave per column j:
// part[3 blk + 0] = min_j (P_jj - r_j), part[3 blk + 1] = max_j (P_jj + r_j),
// r_j = sum_{i != j} |P_ij|, and part[3 blk + 2] = max |P - I| over the
// workgroup's columns (NaN as +inf, and as -inf / +inf in the bounds).  Every
// eigenvalue of P lies in [min, max].  Grid of kPolarParts workgroups.
template <class T>
__global__ __launch_bounds__(256) void k_gersh(const T* __restrict__ P, int n,
                                               double* __restrict__ part) {
  __shared__ double red[3][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double gmin = INFINITY, gmax = -INFINITY, dev = 0.0;
  for (int j = blockIdx.x * 4 + wave; j < n; j += gridDim.x * 4) {
    double off = 0.0, diag = 0.0, dv = 0.0;
    for (int i = lane; i < n; i += 64) {
      const T p = P[i + (size_t)j * n];
      if (i == j) {
        diag = real(p);
        dv = fmax(dv, absv(p - T(1)));
      } else {
        const double v = absv(p);
        off += v;
        dv = fmax(dv, v);
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      off += __shfl_xor(off, o);
      diag += __shfl_xor(diag, o);
      dv = fmax(dv, __shfl_xor(dv, o));
    }
    if (!(off == off) || !(diag == diag) || !(dv == dv)) {
      gmin = -INFINITY;
      gmax = INFINITY;
      dev = INFINITY;
    } else {
      gmin = fmin(gmin, diag - off);
      gmax = fmax(gmax, diag + off);
      dev = fmax(dev, dv);
    }
  }
  if (lane == 0) {
    red[0][wave] = gmin;
    red[1][wave] = gmax;
    red[2][wave] = dev;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      gmin = fmin(gmin, red[0][w]);
      gmax = fmax(gmax, red[1][w]);
      dev = fmax(dev, red[2][w]);
    }
    part[3 * blockIdx.x] = gmin;
    part[3 * blockIdx.x + 1] = gmax;
    part[3 * blockIdx.x + 2] = dev;
  }
}

// part[blk] = max |P - I| over the workgroup's entries (NaN as +inf; part may
// be null); P <- (a/2)(3I - a^2 P).  Grid of kPolarParts workgroups.
template <class T>
__global__ __launch_bounds__(256) void k_ns_prep(T* __restrict__ P, int n, double a,
                                                 double* __restrict__ part) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  double mx = 0.0;
  const size_t tot = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + t; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    const int r = mod32(e, n), c = div32(e, n);
    const T one = (r == c) ? T(1) : T(0);
    const T p = P[e];
    double v = absv(p - one);
    if (!(v == v)) v = INFINITY;
    mx = fmax(mx, v);
    P[e] = (3.0 * one - (a * a) * p) * (0.5 * a);
  }
  if (!part) return;
  red[t] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) red[t] = fmax(red[t], red[t + o]);
    __syncthreads();
  }
  if (t == 0) part[blockIdx.x] = red[0];
}

// Steps of the scaled iteration from the lower bound l until the bound
// reaches 1 to rounding (the iterate is then unitary in exact arithmetic).
int polar_steps(double l) {
  int k = 0;
  while (l < 1.0 - 1e-15 && k < kPolarMaxIter) {
    const double a = std::sqrt(3.0 / (1.0 + l + l * l));
    l = std::min(1.0, 0.5 * a * l * (3.0 - a * a * l * l));
    ++k;
  }
  return k;
}

}  // namespace

template <class T>
void overlap_left_step(hipStream_t st, const T* Tin, int r, int c, const T* Aold, int r2,
                       const T* Anew, int c2, int d, T* Tout, T* tmp) {
  const bool cx = is_cplx<T>::value;
  for (int s = 0; s < d; ++s) {
    // tmp (r x c2) = Tin (r x c) * Anew_s (c x c2, ld c*d)
    gemm<T>(st, OP_N, OP_N, r, c2, c, T(1), Tin, r, Anew + (size_t)s * c, c * d, T(0), tmp, r,
            gemm_pick_tile(r, c2, c, cx));
    // Tout (r2 x c2) += Aold_s^H (r2 x r) * tmp
    gemm<T>(st, OP_H, OP_N, r2, c2, r, T(1), Aold + (size_t)s * r, r * d, tmp, r,
            s == 0 ? T(0) : T(1), Tout, r2, gemm_pick_tile(r2, c2, r, cx));
  }
}

template <class T>
void overlap_right_step(hipStream_t st, const T* Tin, int r, int c, const T* Bold, int r2,
                        const T* Bnew, int c2, int d, T* Tout, T* tmp) {
  const bool cx = is_cplx<T>::value;
  for (int s = 0; s < d; ++s) {
    // tmp (r x c2) = Tin (r x c) * Bnew_s^T, Bnew_s (c2 x c) at Bnew + s*c2, ld c2*d
    gemm<T>(st, OP_N, OP_T, r, c2, c, T(1), Tin, r, Bnew + (size_t)s * c2, c2 * d, T(0), tmp, r,
            gemm_pick_tile(r, c2, c, cx));
    // Tout (r2 x c2) += conj(Bold_s) (r2 x r, at Bold + s*r2, ld r2*d) * tmp
    gemm<T>(st, OP_C, OP_N, r2, c2, r, T(1), Bold + (size_t)s * r2, r2 * d, tmp, r,
            s == 0 ? T(0) : T(1), Tout, r2, gemm_pick_tile(r2, c2, r, cx));
  }
}

template <class T>
bool polar_unitary(hipStream_t st, const T* Tin, int n, T* Z0, T* Z1, T* P, double* dev,
                   double* hp, double start_tol, T** Zout) {
  const bool cx = is_cplx<T>::value;
  copy_dev<T>(st, Z0, Tin, (size_t)n * n);
  T* Z = Z0;
  T* Zn = Z1;
  // dev[it*kPolarParts + blk]: workgroup maxima of |Z_it^H Z_it - I|; the host
  // reduces them after the first step and after steps 9, 11, 13, ... from the
  // fixed bound (which needs 13), or K-2, K, K+2, ... from a Gershgorin bound
  // that needs K steps to reach 1 (steps past convergence keep Z unitary
  // to rounding, the iteration being stable at its fixed point, and by then
  // a = 1: the bound l reaches 1 no later than the singular values).  The
  // first step also bounds the spectrum of Z^H Z by Gershgorin's discs: when
  // they leave out 0, sqrt(min) replaces the fixed start l = kPolarL0 (an
  // overlap of two bases that differ little is close to unitary, and its
  // iteration then needs 3-6 steps instead of ~13).
  int seen = 0;
  double l = kPolarL0;  // lower bound of the smallest singular value of Z
  int first_check = 9;
  for (int it = 0; it < kPolarMaxIter; ++it) {
    gemm_herm<T>(st, OP_H, OP_N, n, n, T(1), Z, n, Z, n, P, n, gemm_pick_tile(n, n, n, cx));
    if (it == 0) {
      double* const gp = hp + 2 * (size_t)kPolarParts;  // past slots 0 and 1 of the read-back
      hipLaunchKernelGGL(k_gersh<T>, dim3(kPolarParts), dim3(256), 0, st, (const T*)P, n, dev);
      DMRG_LAUNCH_CHECK();
      DMRG_HIP(hipMemcpyAsync(gp, dev, 3 * (size_t)kPolarParts * sizeof(double),
                              hipMemcpyDeviceToHost, st));
      DMRG_HIP(hipStreamSynchronize(st));
      double gmin = INFINITY, h = 0.0;
      for (int b = 0; b < kPolarParts; ++b) {
        gmin = std::min(gmin, gp[3 * b]);
        const double v = gp[3 * b + 2];
        h = (v == v) ? std::max(h, v) : INFINITY;
      }
      if (!(h == h) || h > start_tol) return false;
      if (h <= 1e-14) {
        *Zout = Z;
        return true;
      }
      if (gmin == gmin && gmin > kPolarL0 * kPolarL0) {
        l = std::min(1.0, std::sqrt(gmin));
        first_check = std::max(1, polar_steps(l) - 2);
      }
      seen = 1;
    }
    const double a = l < 1.0 ? std::sqrt(3.0 / (1.0 + l + l * l)) : 1.0;
    l = std::min(1.0, 0.5 * a * l * (3.0 - a * a * l * l));
    hipLaunchKernelGGL(k_ns_prep<T>, dim3(kPolarParts), dim3(256), 0, st, P, n, a,
                       it == 0 ? nullptr : dev + (size_t)it * kPolarParts);
    DMRG_LAUNCH_CHECK();
    if (it > 0 && ((it >= first_check && (it - first_check) % 2 == 0) || it == kPolarMaxIter - 1)) {
      DMRG_HIP(hipMemcpyAsync(hp + (size_t)seen * kPolarParts, dev + (size_t)seen * kPolarParts,
                              (size_t)(it + 1 - seen) * kPolarParts * sizeof(double),
                              hipMemcpyDeviceToHost, st));
      DMRG_HIP(hipStreamSynchronize(st));
      for (int j = seen; j <= it; ++j) {
        double h = 0.0;
        for (int b = 0; b < kPolarParts; ++b) {
          const double v = hp[(size_t)j * kPolarParts + b];
          h = (v == v) ? std::max(h, v) : INFINITY;
        }
        // a zero singular value never converges: give up when the start is
        // too far or the deviation is still O(1) after 24 steps (from
        // s = 1e-5 the scaled iteration needs about 24)
        if (!(h == h) || (j >= 24 && h > 0.5)) return false;
        if (h <= 1e-14) {
          *Zout = Z;
          return true;
        }
      }
      seen = it + 1;
    }
    gemm<T>(st, OP_N, OP_N, n, n, n, T(1), Z, n, P, n, T(0), Zn, n, gemm_pick_tile(n, n, n, cx));
    std::swap(Z, Zn);
  }
  return false;
}

template <class T>
void warm_basis_left(hipStream_t st, const T* Z, int cL, int d, const T* X, T* out) {
  const int N = cL * d;
  for (int s = 0; s < d; ++s)
    gemm<T>(st, OP_H, OP_N, cL, N, cL, T(1), Z, cL, X + (size_t)s * cL, N, T(0),
            out + (size_t)s * cL, N, gemm_pick_tile(cL, N, cL, is_cplx<T>::value));
}

template <class T>
void warm_basis_right(hipStream_t st, const T* Z, int cR, int d, const T* X, T* out) {
  const int N = d * cR;
  // rows i = (s, j) -> s + N*j (two-level, n1 = d); k = b' -> d*b'; n = b -> d*b
  const View2 vx{{d, 1, N}, {1 << 30, d, 0}};
  const View2 vz{{1 << 30, 1, 0}, {1 << 30, cR, 0}};  // Z(b', b) at b' + cR b
  gemm_tensor<T>(st, d * N, cR, cR, T(1), X, vx, false, false, Z, vz, true, true, T(0), out, vx,
                 gemm_pick_tile(d * N, cR, cR, is_cplx<T>::value));
}

#define DMRG_WARM_INST(T)                                                                       \
  template void overlap_left_step<T>(hipStream_t, const T*, int, int, const T*, int, const T*, \
                                     int, int, T*, T*);                                         \
  template void overlap_right_step<T>(hipStream_t, const T*, int, int, const T*, int, const T*, \
                                      int, int, T*, T*);                                        \
  template bool polar_unitary<T>(hipStream_t, const T*, int, T*, T*, T*, double*, double*,      \
                                 double, T**);                                                  \
  template void warm_basis_left<T>(hipStream_t, const T*, int, int, const T*, T*);               \
  template void warm_basis_right<T>(hipStream_t, const T*, int, int, const T*, T*);
DMRG_WARM_INST(double)
DMRG_WARM_INST(cdouble)
#undef DMRG_WARM_INST

}  // namespace dmrg
