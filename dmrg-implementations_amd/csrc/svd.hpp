// FP64 / complex128 SVD for the two-site truncation (replaces the reference's
// rocsolver_gesvdj / gesvd call, gpu-rocm/dmrg2-gpu/src/dmrg2_gpu_impl.h:1193-1202,
// gpu-rocm/common/scalar_traits.h:178-188 and :375-391).
//
// Algorithm: block one-sided Jacobi (Hestenes).  The working matrix W (p x q,
// q <= p; the input is conjugate-transposed first when it is wide) is split
// into column blocks of 8.  Each round pairs all blocks (round-robin); for
// every block pair a 16x16 Gram matrix is accumulated over row slices,
// diagonalised by cyclic two-sided Jacobi in LDS, and the resulting unitary
// is applied to the pair's 16 columns of W and of the accumulated V.  Sweeps
// repeat until every pair satisfies |w_i^H w_j| <= tol * |w_i| |w_j|.
// One-sided Jacobi keeps high relative accuracy for small singular values,
// which the reference's absolute 1e-14 truncation threshold relies on.
#pragma once
#include "common.hpp"

namespace dmrg {

template <class T>
struct SvdWork {
  // Device workspace sized for p, q <= cap.
  T* W = nullptr;        // cap * capq
  T* V = nullptr;        // capq * capq
  T* gram = nullptr;     // partial Gram matrices
  double* sigma = nullptr;
  int* perm = nullptr;
  unsigned long long* off = nullptr;   // max relative off-diagonal of a sweep
  int* newk = nullptr;
  int cap = 0, capq = 0;
  size_t gram_elems = 0;
  void alloc(int cap_rows);
  void release();
};

struct SvdResult {
  int rows_p, cols_q;  // working matrix shape
  bool transposed;     // working matrix is A^H
  int sweeps;
  double off;          // final relative off-diagonal measure
};

// Factor A (m x n, column-major, lda) as U S V^H, singular values sorted
// descending into work.sigma/perm.  Synchronises the stream once per sweep.
template <class T>
SvdResult svd_jacobi(hipStream_t st, SvdWork<T>& w, const T* A, int m, int n, int lda,
                     int max_sweeps = 40);

// Truncation rank on device: first i < k with sigma_i < max(cutoff, rel*sigma_0),
// else k, at least 1 (scalar_traits.h:497-505).  Returns it on the host.
template <class T>
int svd_rank(hipStream_t st, SvdWork<T>& w, const SvdResult& r, int k, double cutoff,
             double rel_cutoff);

// Write the truncated factors.  left (m x nk, ldl) gets U or U*S; right
// (nk x n, ldr) gets V^H or S*V^H, depending on scale_left.
template <class T>
void svd_write(hipStream_t st, SvdWork<T>& w, const SvdResult& r, int m, int n, int nk,
               bool scale_left, T* left, int ldl, T* right, int ldr);

// Copy the first `count` sorted singular values to a device buffer.
template <class T>
void svd_sigma(hipStream_t st, SvdWork<T>& w, const SvdResult& r, int count, double* out);

}  // namespace dmrg
