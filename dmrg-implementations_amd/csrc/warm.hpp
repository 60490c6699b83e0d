// Warm start of the truncation SVD from the previous sweep (two-site engine).
//
// At bond i of a left-to-right sweep the density matrix rho = theta theta^H
// lives in the basis (a, s) of the left block (sites 0..i-1, bond index a)
// times site i.  The previous left-to-right sweep diagonalised rho at the same
// bond in the basis of the left block as it was then.  The overlap
// T[a_old, a_new] = <a_old|a_new> of the two left blocks follows from the
// saved left-canonical tensors by the transfer recursion
//     T_j = sum_s A_old[j]_s^H T_{j-1} A_new[j]_s,   T_{-1} = [1],
// and (P^H (x) I_d) X_old, with P the unitary polar factor of T, is an exact
// orthonormal basis of the new (a, s) space that is close to the new
// eigenbasis once the sweep has converged.  Right-to-left sweeps mirror this
// with the right blocks, T_j = sum_s conj(B_old[j]_s) T_{j+1} B_new[j]_s^T.
// The warm basis only changes how many Jacobi sweeps the split needs; the
// split itself is exact for any orthonormal start.
#pragma once
#include "common.hpp"
#include "gemm.hpp"

namespace dmrg {

// Tout (r2 x c2) = sum_s Aold_s^H Tin Anew_s; Tin r x c, Aold (r, d, r2),
// Anew (c, d, c2); tmp holds r * c2.
template <class T>
void overlap_left_step(hipStream_t st, const T* Tin, int r, int c, const T* Aold, int r2,
                       const T* Anew, int c2, int d, T* Tout, T* tmp);

// Tout (r2 x c2) = sum_s conj(Bold_s) Tin Bnew_s^T; Tin r x c, Bold (r2, d, r),
// Bnew (c2, d, c); tmp holds r * c2.
template <class T>
void overlap_right_step(hipStream_t st, const T* Tin, int r, int c, const T* Bold, int r2,
                        const T* Bnew, int c2, int d, T* Tout, T* tmp);

constexpr int kPolarMaxIter = 40;
constexpr int kPolarParts = 256;  // workgroups of the deviation reduction
constexpr double kPolarL0 = 3e-4;  // initial lower bound of the smallest singular value

// Unitary polar factor of the n x n matrix Tin by the scaled Newton-Schulz
// iteration Z <- (a/2) Z (3I - a^2 Z^H Z) of Chen & Chow (2014), which maps
// every singular value s by the same odd cubic.  The scale a follows a lower
// bound l of the smallest singular value, a = sqrt(3/(1 + l + l^2)) and
// l <- (a/2) l (3 - a^2 l^2), from l = kPolarL0 (a -> 1, the plain
// iteration, as l -> 1).  A singular value below the bound still grows by
// about (3/2) a per step, so the bound is a guess that costs steps when it is
// wrong, never convergence.  Returns false (and leaves *Zout undefined) when
// max|T^H T - I| > start_tol, when the iteration stalls (a singular value
// near 0) or when it does not reach max|Z^H Z - I| <= 1e-14 in kPolarMaxIter
// steps.  When Gershgorin's discs of Z^H Z at the first step leave out 0,
// the square root of their lower end replaces kPolarL0 as the start bound.
// Z0, Z1, P are n x n work buffers; dev holds kPolarMaxIter * kPolarParts
// device doubles and hp as many pinned host doubles (the read-backs of the
// deviation and of the bounds).  Synchronises the stream after the first
// step, then every second step from two steps before the bound reaches 1
// (from step 9 with the fixed bound).
template <class T>
bool polar_unitary(hipStream_t st, const T* Tin, int n, T* Z0, T* Z1, T* P, double* dev,
                   double* hp, double start_tol, T** Zout);

// Left side: out[(a + cL s), j] = sum_a' conj(Z[a', a]) X[(a' + cL s), j],
// N = cL d columns (ld N).
template <class T>
void warm_basis_left(hipStream_t st, const T* Z, int cL, int d, const T* X, T* out);

// Right side: out[(s + d b), j] = sum_b' conj(Z[b', b]) X[(s + d b'), j],
// N = d cR columns (ld N).
template <class T>
void warm_basis_right(hipStream_t st, const T* Z, int cR, int d, const T* X, T* out);

}  // namespace dmrg
