// Hand-written MFMA GEMM for gfx950.
//
//   C[M x N] = alpha * op(A) * op(B) + beta * C
//
// Replaces the rocBLAS dgemm / zgemm / gemm_batched calls of the reference
// (gpu-rocm/common/scalar_traits.h:46-199, :204-406).  FP64 runs on
// v_mfma_f64_16x16x4_f64; complex128 is four real MFMAs per k-step on separate
// real/imaginary LDS planes; FP32 runs on v_mfma_f32_16x16x4_f32.
//
// Two addressing modes:
//  * plain: column-major with leading dimensions (BLAS semantics);
//  * two-level ("tensor") addressing, where a row or column index i of an
//    operand is split as i = i1 + n1*i2 and lives at i1*s1 + i2*s2.  This lets
//    one launch contract a pair of tensor legs against another pair without a
//    permutation pass (used for the MPO-bond contraction of H_eff and of the
//    environment updates).
#pragma once
#include "common.hpp"

namespace dmrg {

struct GemmTile {
  int bm, bn;  // workgroup tile: 32x32 or 64x64
  int ks = 1;  // 2: eight waves, each pair splitting every k-tile in halves
};

// Index map of one operand dimension: i -> (i % n1) * s1 + (i / n1) * s2.
struct Dim2 {
  int n1;
  long long s1, s2;
};
// Operand view: element (r, c) at row(r) + col(c).
struct View2 {
  Dim2 row, col;
};
inline View2 view_colmajor(long long ld) {  // plain column-major matrix
  return View2{{1 << 30, 1, 0}, {1 << 30, ld, 0}};
}

template <class T>
void gemm(hipStream_t st, int opA, int opB, int M, int N, int K, T alpha, const T* A, int lda,
          const T* B, int ldb, T beta, T* C, int ldc, GemmTile tile = {64, 64});

// Strided batch: batch b uses A + b*sA, B + b*sB, C + b*sC.
// C = alpha op(A) op(B) for a product known to be Hermitian (N x N; e.g.
// X^H X, or X^H M X with M Hermitian): only the square tiles on and above the
// diagonal are computed, each also writing its conjugate transpose below.
// The result's two triangles are then exact conjugates off the diagonal tiles.
template <class T>
void gemm_herm(hipStream_t st, int opA, int opB, int N, int K, T alpha, const T* A, int lda,
               const T* B, int ldb, T* C, int ldc, GemmTile tile = {64, 64});

template <class T>
void gemm_strided(hipStream_t st, int opA, int opB, int M, int N, int K, T alpha, const T* A,
                  int lda, long long sA, const T* B, int ldb, long long sB, T beta, T* C,
                  int ldc, long long sC, int batch, GemmTile tile = {64, 64});

// Two-level-addressed GEMM.  A is viewed as (M x K), B as (K x N), C as
// (M x N); conjA / conjB conjugate the operands.  `a_kfast` / `b_kfast` say
// which operand index is memory-contiguous (k or m/n) to pick coalesced loads.
template <class T>
void gemm_tensor(hipStream_t st, int M, int N, int K, T alpha, const T* A, View2 va, bool conjA,
                 bool a_kfast, const T* B, View2 vb, bool conjB, bool b_kfast, T beta, T* C,
                 View2 vc, GemmTile tile = {64, 64});

// C (M x N) = A B for tiny inner and outer sizes (K <= 32, N <= 32): the
// MPO-bond step of H_eff and of the environment updates, where a tiled GEMM
// wastes most of its MFMA width and the work is one pass over A and C.  A and
// C use two-level addressing, B is column-major K x N.  Returns false (and
// does nothing) when the sizes are out of range.
template <class T>
bool gemm_rows_small(hipStream_t st, int M, int N, int K, const T* A, View2 va, const T* B, T* C,
                     View2 vc);

// The two-site H_eff's first two steps in one launch (d = 2, real scalars,
// 2 <= D <= 8): Y[a' + cL ss' + cL d^2 b + cL d^2 cR n] =
//   sum_{w, ss} X[(w + D a'), (ss + d^2 b)] WW[(w + D ss) + D d^2 (ss' + d^2 n)]
// with X = L^T theta (L: cL x (D cL), theta: cL x (d^2 cR), both ld cL)
// formed tile by tile on the MFMA and contracted with WW in LDS, so X never
// reaches HBM; the same sums in the same order as the GEMM followed by
// gemm_rows_small.  Returns false (and does nothing) outside that range.
// (Instantiated for TB = T.)
template <class T, class TB = T>
bool heff2_x_mpo(hipStream_t st, int cL, int cR, int d, int D, const T* Lenv, const TB* theta,
                 const T* WW, T* Y);

// gemm_nn_splitk with FP32 operands and an FP64 result C = s1 s2 (A B)
// (s1, s2 device scalars, each taken as 1 when not positive); the partial
// products and their sum stay FP32.
bool gemm_nn_splitk_to_f64(hipStream_t st, int M, int N, int K, const float* A, int lda,
                           const float* B, int ldb, double* C, int ldc, float* ws,
                           size_t ws_elems, const double* s1, const double* s2);

// FP32 C (M x N, ldc) = op(A) op(B) with split-BF16 products: each operand as
// three bf16 terms, the six leading cross products on the bf16 MFMA with FP32
// accumulation (FP32-class accuracy).  op: OP_N or OP_T (conjugation is a
// no-op for real data); alpha = 1, beta = 0.
void gemm_f32_split_bf16(hipStream_t st, int opA, int opB, int M, int N, int K, const float* A,
                         int lda, const float* B, int ldb, float* C, int ldc);

// C = A B (both op N, alpha 1, beta 0) split along K into slices computed
// into the workspace ws (ws_elems elements) and summed in a fixed order:
// for K-heavy products whose 64x64 grid cannot fill the chip.  Returns false
// (and does nothing) when splitting would not help or ws is too small.
template <class T>
bool gemm_nn_splitk(hipStream_t st, int M, int N, int K, const T* A, int lda, const T* B, int ldb,
                    T* C, int ldc, T* ws, size_t ws_elems);

// Pick a tile so that the grid covers the chip.
GemmTile gemm_pick_tile(int M, int N, int K, bool complex_);

// Process-wide tile autotuner of gemm / gemm_herm: when on, the first call of
// each shape (scalar type, ops, M, N, K) times the 32x32 and 64x64 tiles,
// each with four waves or eight (k-split), on the caller's stream and keeps
// the fastest for the rest of the process
// (the tile argument is then only the fallback of calls made during graph
// capture).  gemm_autotune_entries() counts the shapes tuned so far.
void gemm_autotune(bool on);
int gemm_autotune_entries();

}  // namespace dmrg
