// FP64 / complex128 truncated SVD for the two-site (and single-site) split
// (replaces the reference's rocsolver_gesvdj / gesvd call and its truncation
// kernels: gpu-rocm/dmrg2-gpu/src/dmrg2_gpu_impl.h:1166-1343,
// gpu-rocm/common/scalar_traits.h:178-188, :375-391 and :497-547).
//
// Method: eigen-decomposition of the reduced density matrix on the side whose
// factor must be orthonormal.
//   * 'U' side (the reference's direction 'R': A[site] = U, A[site+1] = S Vh):
//       rho = theta theta^H (m x m); U_k = top-k eigenvectors of rho;
//       S Vh = U_k^H theta (one GEMM).
//   * 'V' side (direction 'L': A[site] = U S, A[site+1] = Vh):
//       rho = theta^H theta (n x n); V_k = top-k eigenvectors;
//       U S = theta V_k (one GEMM), Vh = V_k^H.
// The orthonormal factor comes straight from the eigen-solver (orthonormal to
// machine precision whatever the spectrum); the other factor is a product,
// never a division by sigma.  This is the numerical method of the reference's
// own real SVD: rocsolver_dgesvdj diagonalises A^T A (or A A^T) by Jacobi and
// takes the singular values from a QR factorisation of A V (rocSOLVER 7.2
// header, rocsolver-functions.h, GESVDJ).  Here sigma_j is the diagonal of
// that QR's R, obtained from the Gram matrix of the product factor by a
// first-order Schur complement (k_qr_sigma), so the reference's absolute
// cutoff (first sigma < 1e-14, dmrg2_gpu_impl.h:1276 and
// scalar_traits.h:497-505) sees the same deflated values.
//
// Eigen-solver: two-sided block Jacobi.  rho (zero-padded to Np, a multiple
// of 32) is split into 16-wide blocks; each round pairs all blocks
// (round-robin) and one workgroup per pair diagonalises its 32x32 Hermitian
// sub-matrix in LDS by cyclic Jacobi, producing a 32x32 unitary Q_P; a second
// kernel applies the block-diagonal transform rho <- P^H rho P, V <- V P with
// MFMA tiles.  Every sweep starts by sorting the diagonal (largest first),
// which groups the graded DMRG spectrum and cuts the sweep count.  A sweep
// ends with a device reduction of max |rho_ij| (i != j); the solver stops
// when it is below tol * |rho|_F.
#pragma once
#include "common.hpp"

namespace dmrg {

template <class T>
struct SvdWork {
  T* A[2] = {nullptr, nullptr};  // density matrix, ping-pong (capN^2 each)
  T* V[2] = {nullptr, nullptr};  // accumulated eigenvectors
  T* Y = nullptr;                // staging for the gathered / product factor
  T* W = nullptr;                // work matrix of the first-order steps
  T* Q = nullptr;                // per-pair 32x32 rotations of one round
  int* qflag = nullptr;          // per pair: 1 if its Q is not the identity
  int* perm = nullptr;           // sort permutation
  int* rank = nullptr;           // inverse of perm (position of each index)
  double* gdiag = nullptr;       // squared norms of the k product vectors
  double* red = nullptr;         // reduction partials / small results (1024 doubles)
  double* sigma = nullptr;       // singular values in output order
  double* resid = nullptr;       // first-pass QR residuals
  double* scal = nullptr;        // [0] |rho|_F, [1] discarded weight, [2] trace(rho)
  int* newk = nullptr;
  double* host = nullptr;  // pinned host staging of every read-back (2 cap + 2*256 + 16 doubles)
  int* host_i = nullptr;   // pinned host staging of the kept rank
  int cap = 0;  // largest supported rho dimension (multiple of 32)
  void alloc(int cap_n);
  void release();
};

struct SvdParams {
  int max_sweeps = 40;  // outer block-Jacobi sweeps
  double tol = 1e-14;   // stop when max |rho_ij| <= tol * |rho|_F
  // svd_split only: converge just the coupling between the k leading
  // eigenvectors (k = min(m, n, kmax)) and the rest, i.e. the kept subspace,
  // not the individual vectors inside it.  The truncated product left*right
  // is the same (to tol); left is an orthonormal basis of the kept subspace
  // rather than the singular vectors, and sigma_out holds the norms of the
  // product vectors after projecting out the earlier ones.  Falls back to
  // full convergence when the cutoff keeps fewer than k vectors.
  bool subspace = false;
};

struct SvdResult {
  int N;             // dimension of the density matrix that was diagonalised
  int sweeps;        // outer Jacobi sweeps (all schedules)
  int bip_sweeps;    // of which kept-vs-rest (bipartite) sweeps
  int oa_steps;      // first-order all-pairs steps before the sweeps
  double off;        // final max |rho_ij| / |rho|_F
  int nk;            // kept rank
  double discarded;  // sum of the discarded sigma^2 (trace - kept)
};

// Truncated split of A (m x n, column-major, lda).
//   k = min(m, n, kmax); nk = first j < k with sigma_j < max(cutoff,
//   rel_cutoff * sigma_0), else k; at least 1.
//   scale_left == false ('U' side): left (m x nk, ldl) = U, right (nk x n,
//     leading dimension nk) = S Vh.
//   scale_left == true ('V' side): left = U S, right = Vh.
//   sigma_out (optional, device): the nk kept singular values.
//   warm (optional, device, N x N, ld N; N = m for the 'U' side, n for 'V'):
//     an orthonormal basis of that side to start the Jacobi from (rho is
//     first rotated into it).  Any orthonormal basis gives the same result;
//     one close to the eigenbasis needs fewer sweeps.
//   basis_out (optional, device, N x N, ld N): receives the full sorted
//     eigenbasis, the warm start of a later split at the same bond.
// Synchronises the stream once per Jacobi sweep and once for the rank.
template <class T>
SvdResult svd_split(hipStream_t st, SvdWork<T>& w, const T* A, int m, int n, int lda, int kmax,
                    double cutoff, double rel_cutoff, bool scale_left, T* left, int ldl, T* right,
                    double* sigma_out, const SvdParams& prm, const T* warm = nullptr,
                    T* basis_out = nullptr);

// Work space of rsvd_split: sketch, bases and the projected matrix, each
// cap_len x cap_r, plus Gram-Schmidt scratch.
template <class T>
struct RsvdWork {
  T *Om = nullptr, *Y = nullptr, *Z = nullptr, *B = nullptr, *Wb = nullptr, *C = nullptr;
  double* nrm = nullptr;
  int len = 0, r = 0;
  void alloc(int cap_len, int cap_r);
  void release();
};

// Randomised truncated split (the reference's RSVD option,
// gpu-rocm/dmrg2-gpu/src/dmrg2_gpu_impl.h:1186-1266, Halko-Martinsson-Tropp):
// a sketch of r = min(k + oversample, m, n) random combinations of the side
// whose factor is orthonormal (uniform [-1, 1) entries from the splitmix64
// stream `seed`), `power` re-orthonormalised subspace iterations, an
// orthonormal basis Q by block Gram-Schmidt, and the exact split (svd_split,
// every pair converged) of the r-dimensional projection; the orthonormal
// factor is mapped back through Q.  Same outputs and truncation rule as
// svd_split (no sigma_out, no warm start).  The result approximates the
// exact split to within the sketch error: exact when the rank of A is at
// most r, otherwise the kept space is off by roughly
// (sigma_r / sigma_k)^(2 power + 1).
template <class T>
SvdResult rsvd_split(hipStream_t st, SvdWork<T>& w, RsvdWork<T>& rw, const T* A, int m, int n,
                     int lda, int kmax, double cutoff, double rel_cutoff, bool scale_left, T* left,
                     int ldl, T* right, const SvdParams& prm, int oversample, int power,
                     uint64_t seed);

// Eigen-decomposition of a Hermitian matrix H (N x N, ldh): eigenvalues
// descending into evals (device, N), eigenvectors into evecs (N x N, lde).
template <class T>
int heig(hipStream_t st, SvdWork<T>& w, const T* H, int N, int ldh, double* evals, T* evecs,
         int lde, const SvdParams& prm);

}  // namespace dmrg
