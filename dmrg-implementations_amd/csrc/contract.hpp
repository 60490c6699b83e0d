// Tensor-network contractions of the sweep, written as MFMA GEMMs.
//
// Layouts (column-major, as in the reference GPU engines):
//   MPS     A[a + s*cL + b*cL*d]                       (cL, d, cR)
//   theta   th[a + s1*cL + s2*cL*d + b*cL*d*d]          (cL, d, d, cR)
//   env     L[a + w*chi + a2*chi*D]   (ket, mpo, bra)   (chi, D, chi)
//   MPO     W[w + s*D + sp*D*d + wp*D*d*d]  (s ket, sp bra)
// (gpu-rocm/dmrg2-gpu/src/dmrg2_gpu_impl.h:582-934, test_dmrg2_gpu.cpp:22-34)
//
// Every contraction is three launches: a big GEMM against an environment, an
// MPO-bond GEMM with two-level addressing (K = D*d or D*d^2), and a big GEMM
// against the other environment / MPS.  No permutation passes.
#pragma once
#include "common.hpp"
#include "gemm.hpp"

namespace dmrg {

// Scratch for one contraction stream: X and Y each hold D*d^2*chi^2 entries.
template <class T>
struct ContractScratch {
  T* X = nullptr;
  T* Y = nullptr;
  size_t cap = 0;
  bool split_bf16 = false;  // FP32 only: the H_eff GEMMs as split-BF16 products
};

// Two-site H_eff: out = H_eff(theta); WW is the fused two-site MPO as a
// (D*d^2) x (D*d^2) matrix WW[(w + D*(s1 + d*s2)) + D*d^2*((s1p + d*s2p) + d^2*n)].
template <class T>
void heff2_apply(hipStream_t st, const T* Lenv, const T* Renv, const T* WW, const T* theta, T* out,
                 int cL, int cR, int d, int D, ContractScratch<T>& sc);

// The FP32 two-site H_eff on FP64 vectors: out = sL sR H_eff(FP32 L, R, WW)
// theta, with L and R scaled to unit max-abs by sL and sR (device scalars).
// theta is rounded to FP32 into thf; where the fused first step and the
// split-K out-step apply, the split-K sum writes the widened, rescaled result
// directly, otherwise the FP32 result goes through outf (n elements each)
// and a conversion.  Same values either way.
void heff2_apply_f32_io64(hipStream_t st, const float* Lenv, const float* Renv, const float* WW,
                          const double* theta, double* out, int cL, int cR, int d, int D,
                          ContractScratch<float>& sc, float* thf, float* outf, const double* sL,
                          const double* sR);

// Single-site H_eff (dmrg_gpu_impl.h:538-705); W is the raw MPO tensor.
template <class T>
void heff1_apply(hipStream_t st, const T* Lenv, const T* Renv, const T* W, const T* theta, T* out,
                 int cL, int cR, int d, int D, ContractScratch<T>& sc);

// Lnew (co, D, co) from L (ci, D, ci), A (ci, d, co), W (raw).
template <class T>
void env_left(hipStream_t st, const T* Lenv, const T* A, const T* W, T* Lnew, int ci, int co, int d,
              int D, ContractScratch<T>& sc);

// Rnew (co, D, co) from R (ci, D, ci), A (co, d, ci), Wr (right-env MPO matrix
// Wr[(s + d*wp) + d*D*(w + D*sp)] = W[w, s, sp, wp]).
template <class T>
void env_right(hipStream_t st, const T* Renv, const T* A, const T* Wr, T* Rnew, int ci, int co,
               int d, int D, ContractScratch<T>& sc);

// Host builders of the derived MPO matrices (from the raw layout).
template <class T>
void build_ww2(const T* W1, const T* W2, int d, int D, T* WW);
template <class T>
void build_wright(const T* W, int d, int D, T* Wr);

// Elementwise helpers.
void convert_f64_f32(hipStream_t st, const double* in, float* out, size_t n);
void convert_f32_f64(hipStream_t st, const float* in, double* out, size_t n);
// out[0] = max |in[i]| (device scalar).
void absmax_f64(hipStream_t st, const double* in, size_t n, double* out);
// out = in / s (s = *scale, or 1 when *scale == 0).  The FP32 matvec keeps its
// environments at unit scale: those of an unnormalised start state can exceed
// the FP32 range (the right environments of a random L=128, chi=512 MPS reach
// ~1e67).
void convert_f64_f32_scaled(hipStream_t st, const double* in, float* out, size_t n,
                            const double* scale);
// out = in * s1 * s2 (each 1 when 0): undoes the two environment scales.
void convert_f32_f64_scaled(hipStream_t st, const float* in, double* out, size_t n,
                            const double* s1, const double* s2);
template <class T>
void set_boundary_env(hipStream_t st, T* env, int D, int w);  // (1, D, 1) unit vector e_w

// FP32 environments of the mixed modes: an environment is a float array e
// with a device double s, the true environment being s * e with max |e| = 1
// (a random start state's right environments reach ~1e67, past the FP32
// range, so the magnitude lives in s).
// out[0] = max |in[i]| as a double (0 for n = 0).
void absmax_f32(hipStream_t st, const float* in, size_t n, double* out);
// e /= m and *s_out = *s_in * m, m = *maxv (1 when 0): renormalises a freshly
// contracted FP32 environment and carries the factor in its scale.
void rescale_env_f32(hipStream_t st, float* e, size_t n, const double* maxv, const double* s_in,
                     double* s_out);
// s[0..n) = v
void fill_f64(hipStream_t st, double* s, size_t n, double v);

// dst(i, j) = src(i, j) for i < rows, j < cols (column-major, leading
// dimensions ldd and lds) on the device, by a copy kernel: every copy of the
// sweep is a kernel of its own rather than a runtime copy, so a stream's
// work is kernels only.
template <class T>
void copy_dev(hipStream_t st, T* dst, size_t ldd, const T* src, size_t lds, size_t rows,
              size_t cols);
template <class T>
inline void copy_dev(hipStream_t st, T* dst, const T* src, size_t n) {
  copy_dev<T>(st, dst, n, src, n, n, 1);
}

// env_left on the library's own fixed device scratch (noalloc.hip): no
// allocation, no context.  dtype is a DMRG_DTYPE_* code.  Throws BAD_ARG
// when the two intermediates exceed DMRG_NOALLOC_SCRATCH_BYTES.
void env_left_noalloc_op(hipStream_t st, int dtype, int ci, int co, int d, int D, const void* L,
                         const void* A, const void* W, void* Lnew);

}  // namespace dmrg
