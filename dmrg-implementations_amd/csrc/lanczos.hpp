// Lanczos building blocks (replaces the rocBLAS dot/axpy/gemv/nrm2/scal
// sequence and the rocsolver_dsteqr calls of the reference's
// lanczos_eigensolver, gpu-rocm/dmrg2-gpu/src/dmrg2_gpu_impl.h:970-1160, and
// the scalar kernels of gpu-rocm/common/scalar_traits.h:422-489).
//
// Reductions carry no atomics or fences: every workgroup of a vector kernel
// writes one partial sum, and the next kernel in the stream reduces the
// partials (each of its workgroups, in the same fixed order), so results are
// deterministic and bit-identical across workgroups.
#pragma once
#include "common.hpp"

namespace dmrg {

struct LzDev {
  // device scalars / small arrays
  double* alpha = nullptr;     // [kmax]
  double* beta = nullptr;      // [kmax]
  double* y = nullptr;         // [kmax] Ritz coefficients
  double* scal = nullptr;      // [8]: 0 norm, 1 inv, 2 prev_energy, 3 energy, 4 check value
  int* status = nullptr;       // [4]: 0 stop flag, 1 niter
  void* part = nullptr;        // partial sums (bytes), lz_part_doubles(kmax, n) doubles
  void* cvec = nullptr;        // [kmax] projection coefficients (T)
  int kmax = 0;
  size_t part_bytes = 0;
  void alloc(int kmax_, size_t part_bytes_, size_t elem_bytes);
  void release();
};

// Number of workgroups used by the vector kernels for length n.
int lz_blocks(size_t n);
// Size of LzDev::part, in doubles, for Krylov dimension kmax and length n.
size_t lz_part_doubles(int kmax, size_t n);

// Norm partials of x (consumed by lz_fill_if_small / lz_scale).
template <class T> void lz_norm(hipStream_t st, LzDev& d, const T* x, size_t n);
// If |x| (from the last lz_norm) < thr, overwrite x with the splitmix64
// stream `seed` (U[-1,1)).
template <class T> void lz_fill_if_small(hipStream_t st, LzDev& d, T* x, size_t n, uint64_t seed,
                                         double thr);
// nrm = norm from the last lz_norm / lz_orth_norm / lz_ritz; scal[0] = nrm,
// scal[1] = 1/nrm (0 when nrm <= 1e-300), *beta_out = nrm if given;
// dst = src / nrm if dst is given.
template <class T> void lz_scale(hipStream_t st, LzDev& d, T* dst, const T* src, size_t n,
                                 double* beta_out = nullptr);
// beta[it] = |w| (after lz_orth_norm); vnext = w / beta[it] if vnext is given.
template <class T> void lz_next(hipStream_t st, LzDev& d, T* vnext, const T* w, size_t n, int it);
// Partials of Re <v, w>; lz_sub_dots (or lz_alpha_value) turns them into alpha[it].
template <class T> void lz_alpha(hipStream_t st, LzDev& d, const T* v, const T* w, size_t n, int it);
// alpha[it] = Re <v, w> from the partials of the last lz_alpha (length n).
void lz_alpha_value(hipStream_t st, LzDev& d, size_t n, int it);
// alpha[it] = <v_it, w>;  w -= alpha[it] v_it + beta[it-1] v_{it-1};
// partials of c_j = <v_j, w>, j <= it
template <class T> void lz_sub_dots(hipStream_t st, LzDev& d, T* w, const T* V, size_t n, int it);
// w -= sum_j c_j v_j;  norm partials of w (lz_next gives beta[it])
template <class T> void lz_orth_norm(hipStream_t st, LzDev& d, T* w, const T* V, size_t n, int it);
// prev_energy := 1e30, status := 0 (start of a Lanczos call).
void lz_reset(hipStream_t st, LzDev& d);
// Convergence check on T(ncheck) (every third iteration).
// Converged when |E - E_prev| < max(eig_tol, eig_rel |E|).
void lz_check(hipStream_t st, LzDev& d, int ncheck, double beta_tol, double eig_tol,
              double eig_rel = 0.0);
// Lowest eigenpair of T(niter) -> scal[3], y.
void lz_final(hipStream_t st, LzDev& d, int niter);
// out = V y (niter columns) and its norm partials (lz_scale normalises)
template <class T> void lz_ritz(hipStream_t st, LzDev& d, T* out, const T* V, size_t n, int niter);

// Fill `x` with entries scale*U[-1,1) from the splitmix64 stream `seed`,
// starting at stream position `offset` (complex: two outputs per entry).
template <class T> void fill_random(hipStream_t st, T* x, size_t n, uint64_t seed, uint64_t offset,
                                    double scale);

}  // namespace dmrg
