/*
 * dmrg_amd.h -- C-ABI of the MI355X (gfx950) DMRG sweep engine.
 *
 * Drop-in for the single-GPU sweep path of the reference
 * (gpu-rocm/dmrg2-gpu DMRG2GPU<Scalar> and gpu-rocm/dmrg-gpu DMRGGPU<Scalar>):
 * an MPO and an initial MPS go in, the ground-state energy and the truncated
 * MPS come out.  Plain pointers and sizes only; no exception crosses the ABI.
 *
 * Layouts (all column-major, exactly the reference's device layouts):
 *   MPO site tensor  W[w + s*D + sp*D*d + wp*D*d*d]   s = ket, sp = bra,
 *                    left boundary row w = 0, right boundary column wp = D-1
 *                    (gpu-rocm/dmrg2-gpu/src/test_dmrg2_gpu.cpp:22-34)
 *   MPS site tensor  A[a + s*cL + b*cL*d]                 shape (cL, d, cR)
 *   Complex scalars are interleaved (re, im) doubles (hipDoubleComplex).
 *
 * Ownership: the caller owns every host buffer; the handle owns all device
 * memory.  No pointer is retained past the call that receives it.
 * Threading: one handle per host thread; a handle is not re-entrant.
 */
#ifndef DMRG_AMD_H
#define DMRG_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference throws std::runtime_error from HIP_CHECK /
 * ROCBLAS_CHECK (gpu-rocm/common/hip_check.h:13-31) and on dsteqr failure
 * (gpu-rocm/dmrg2-gpu/src/dmrg2_gpu_impl.h:1120-1122); here each failure is a
 * return code with the message in dmrg_last_error(). */
enum {
  DMRG_OK = 0,
  DMRG_ERR_HIP = 1,             /* HIP runtime / launch failure          */
  DMRG_ERR_SOLVER_NO_CONV = 2,  /* SVD or eigensolver did not converge   */
  DMRG_ERR_BAD_ARG = 3,         /* invalid argument or shape             */
  DMRG_ERR_NONFINITE = 4,       /* NaN / Inf met in the sweep            */
  DMRG_ERR_STATE = 5,           /* call out of order (e.g. run before set_mpo) */
  DMRG_ERR_OOM = 6              /* device allocation failed              */
};

enum { DMRG_DTYPE_F64 = 0, DMRG_DTYPE_C128 = 1 };
enum { DMRG_SINGLE_SITE = 1, DMRG_TWO_SITE = 2 };
/* Arithmetic of the H_eff matvec and of the environment updates.  Lanczos
 * (tridiagonal solve, reorthogonalisation) and the truncation SVD are FP64 in
 * every mode.  FP32: the matvec GEMMs, and (option env_precision = 1) the
 * environment updates, on FP32 MFMA.  BF16X3: the same with the matvec's two
 * GEMMs as split-BF16 products (each operand hi+mid+lo in bf16, the six
 * leading cross products on the bf16 MFMA, FP32 accumulation).  Both mixed
 * modes are real two-site only. */
enum { DMRG_PREC_FP64 = 0, DMRG_PREC_FP32 = 1, DMRG_PREC_BF16X3 = 2 };

typedef struct dmrg_options {
  int krylov_max;        /* Lanczos Krylov dimension cap; reference: min(100, n)
                            (dmrg2_gpu_impl.h:138, :974).  <= 128.            */
  double lanczos_tol;    /* beta breakdown threshold, reference 1e-12 (:975) */
  double eig_tol;        /* Ritz-value convergence threshold, 1e-12 (:976)   */
  double svd_cutoff;     /* absolute singular-value cutoff, 1e-14 (:1276)    */
  double svd_rel_cutoff; /* relative cutoff (sigma_i < rel*sigma_0); 0 = off */
  int precision;         /* DMRG_PREC_*                                       */
  int profile;           /* 1: time every H_eff matvec with HIP events (stats
                            heff_ms); 0 (default): no events, heff_ms = 0     */
  int device;            /* HIP device ordinal                                */
  int svd_max_sweeps;    /* Jacobi sweep cap of the SVD eigen-solver          */
  double svd_tol;        /* Jacobi stop: max |rho_ij| <= svd_tol * |rho|_F    */
  int svd_warm;          /* two-site: start each split from the previous sweep's
                            eigenbasis at that bond (1, default) or from I (0) */
  double fp32_eig_rel;   /* FP32 matvec mode only: Lanczos also converges when
                            |dE| < fp32_eig_rel*|E| (FP32 rounding of H_eff
                            bounds how far the Ritz value settles); 1e-9 */
  int svd_subspace;      /* 1 (default): a split's Jacobi stops once the kept
                            chi vectors decouple from the discarded ones; the
                            truncated state is the same, the kept tensor is an
                            orthonormal basis of the kept space rather than the
                            singular vectors.  0: converge every pair.       */
  double svd_tol_mixed;  /* Jacobi stop in the FP32 / BF16X3 modes (1e-13): the
                            two-site vectors there carry FP32 rounding (~1e-7),
                            which couples kept and discarded directions near
                            the cut at ~1e-13 |rho|_F; converging further would
                            resolve that rounding.                            */
  int env_precision;     /* FP32 / BF16X3 modes only.  0 (default): FP64
                            environments, converted to FP32 per bond; the
                            energy is the FP64 Rayleigh quotient of the last
                            bond.  1: the left and right environment updates
                            run on FP32 MFMA and the environments stay FP32
                            (each with an FP64 scale); the energy of every
                            sweep is the FP64 <psi|H|psi>/<psi|psi> of the
                            state.  FP32 environments accumulate rounding
                            along the chain: |dE|/|E| 2.4e-10 against FP64 at
                            L=128, chi=512 (2e-12 with FP64 environments).    */
  int svd_rsvd;          /* 1: randomised two-site split (the reference's
                            DMRG_GPU_OPT_RSVD, dmrg2_gpu_impl.h:1186-1266) on the
                            bonds where the reference uses it (it keeps fewer
                            than min(m, n) - rsvd_oversample values and
                            m > 2 k); 0 (default): exact split everywhere.   */
  int rsvd_oversample;   /* extra sketch columns, 10 as in the reference      */
  int rsvd_power;        /* re-orthonormalised subspace iterations (0: the
                            reference's single pass)                          */
  int graph;             /* 1: each bond's Lanczos iterations between two
                            convergence checks run as a captured HIP graph,
                            cached per bond and bond dimensions and replayed
                            (the reference's per-bond graph cache,
                            dmrg2_gpu_impl.h:615-637); ignored with profile. */
  int autotune;          /* 1: turn on the process-wide GEMM tile autotuner:
                            the first call of each GEMM shape whose default
                            tile has one wave per output sub-tile times the
                            32x32 and 64x64 MFMA tiles and keeps the faster
                            for every later call of that shape in the
                            process.  Both sum over k in the same order, so
                            results do not depend on the choice.
                            No reference counterpart (rocBLAS picks its own
                            kernels).                                         */
} dmrg_options;

typedef struct dmrg_stats {
  double lanczos_ms;     /* host-measured phase times (sum over bonds)      */
  double heff_ms;        /* HIP-event time of all H_eff matvecs on the stream */
  double svd_ms;
  double env_ms;
  double sweep_s_last;   /* wall seconds of the last completed sweep        */
  double heff_flop;      /* algorithmic H_eff flops (F_heff per matvec)     */
  double env_flop;
  long long matvecs;
  long long lanczos_calls;
  long long svd_calls;
  long long svd_sweeps;
  int sweeps_done;
  double max_discarded;  /* largest truncated weight sum(sigma_i^2) of a bond */
  double env_build_s;    /* wall seconds of the last environment build      */
  double sweeps_s_total; /* wall seconds of all sweeps (environment build excluded) */
  long long svd_warm_starts; /* splits that started from a previous sweep's basis */
  long long svd_bip_sweeps;  /* Jacobi sweeps that rotated only kept-vs-discarded pairs */
  long long svd_oa_steps;    /* first-order all-pairs steps before the Jacobi sweeps */
  double trunc_sum_last;     /* truncation error of the last sweep: the discarded
                                weight sum(sigma_i^2) summed over its splits   */
  long long rsvd_calls;      /* splits that took the randomised path            */
  long long graph_launches;  /* Lanczos stretches replayed as graphs            */
  long long graph_captures;  /* graphs captured (first use per bond and shape)  */
  long long tuned_shapes;    /* GEMM shapes timed by the autotuner (process-wide) */
} dmrg_stats;

typedef struct dmrg_engine* dmrg_handle;

/* Fill `opt` with the reference's hard-coded defaults: Krylov cap 100
 * (gpu-rocm/dmrg2-gpu/src/dmrg2_gpu_impl.h:138), Lanczos tolerances 1e-12
 * (:975-976), SVD cutoff 1e-14 (:1276); the other fields are this engine's. */
void dmrg_default_options(dmrg_options* opt);

/* Replaces DMRG2GPU(int L, int d, int chi_max, int D_mpo, double tol)
 * (gpu-rocm/dmrg2-gpu/src/dmrg2_gpu.h:30; impl :51-227) and
 * DMRGGPU(...) (gpu-rocm/dmrg-gpu/src/dmrg_gpu.h:30).  Bond dimensions start
 * at min(d^min(i, L-i), chi_max) (dmrg2_gpu_impl.h:70-76).  opt may be NULL. */
int dmrg_create(int L, int d, int chi_max, int D, double tol, int dtype, int variant,
                const dmrg_options* opt, dmrg_handle* out);
/* Replaces ~DMRG2GPU (dmrg2_gpu_impl.h:266-343).  NULL is ignored. */
void dmrg_destroy(dmrg_handle h);

/* Replaces set_mpo(const std::vector<Scalar*>&) (dmrg2_gpu.h:38; impl :435-576):
 * W[i] points to the D*d*d*D scalars of site i. */
int dmrg_set_mpo(dmrg_handle h, const void* const* W);

/* Replaces initialize_mps_random(double scale) (dmrg2_gpu_impl.h:382-393).
 * Entries scale*U[-1,1) from the counter-based splitmix64 stream `seed`
 * (one stream over all sites, Fortran order, complex: re then im); the
 * reference's unseeded rand() is not reproducible. */
int dmrg_init_mps_random(dmrg_handle h, uint64_t seed, double scale);
/* Replaces initialize_mps_product() (dmrg2_gpu_impl.h:396-413). */
int dmrg_init_mps_product(dmrg_handle h);
/* Replaces initialize_mps_neel() (dmrg2_gpu_impl.h:415-429). */
int dmrg_init_mps_neel(dmrg_handle h);
/* Set one site tensor (cL, d, cR); the bond dimensions of the neighbouring
 * bonds become cL and cR.  Sites may be set one at a time: dmrg_run and
 * dmrg_expectation return DMRG_ERR_STATE while a site's tensor disagrees with
 * its neighbour's on a shared bond.  No reference counterpart (the reference only has
 * the initialize_mps_* generators); used to feed identical states to the
 * engine and to a CPU path, and to load the reference's .bin/.npz files. */
int dmrg_set_mps(dmrg_handle h, int site, const void* A, int cL, int cR);

/* Replaces double run(int n_sweeps) (dmrg2_gpu_impl.h:1426-1457): build the
 * environments, then alternate L->R and R->L sweeps until |E - E_prev| < tol
 * after the second sweep.  energy may be NULL.  When `quiet` is 0 it prints
 * the reference's stdout contract ("Final energy: %.12f", "Total wall time:
 * %.3f s", "  env_build_sec: %.3f  timer_scope: sweep_only"). */
int dmrg_run(dmrg_handle h, int n_sweeps, int quiet, double* energy);

/* Batched solve of independent chains (no reference counterpart; the
 * reference runs chains one binary invocation at a time).  hs[0..n) are
 * distinct engines (BAD_ARG if one is listed twice) created and set up as
 * usual (any mix of L, chi, d, D, variant and precision); each runs
 * dmrg_run(n_sweeps, quiet).  `threads` host threads feed them to the GPU
 * concurrently, longest chain first, so the latency-bound phases of different
 * chains overlap; each thread owns one stream, on which the chains it runs do
 * all their work (the engines' own streams are idle meanwhile).
 * energies[i] (and seconds[i], may be NULL) get chain i's result.  The first
 * error stops the remaining chains and is returned. */
int dmrg_run_batch(dmrg_handle* hs, int n, int n_sweeps, int threads, double* energies,
                   double* seconds);

/* Segment-parallel two-site sweep on one GPU: replaces PDMRGGPU::run
 * (gpu-rocm/pdmrg-gpu/src/pdmrg_gpu_impl.h:2585-2753; class at
 * pdmrg_gpu.h:13-60).  `warmup` serial full sweeps, then up to `outer`
 * iterations of `local` staggered passes: the chain's `segments` segments
 * (lengths differing by at most one, at least 2 sites each) sweep their own
 * bonds concurrently, even segments left to right and odd ones right to left
 * and then the reverse, each pass followed by the boundary merges
 * theta = psi_L diag(V) psi_R with V = 1/max(S, 1e-12) of the boundary's last
 * split; stops when the energy changes by less than tol.  Then `polish`
 * serial full sweeps (two or more at large chi: the merges' energy is not
 * monotone there, and the polish sweeps restore the serial result).  Two-site
 * FP64 / complex128 only (BAD_ARG otherwise).  The energy is the last
 * merge's, or the last polish sweep's. */
int dmrg_run_pdmrg(dmrg_handle h, int segments, int outer, int local, int warmup, int polish,
                   int quiet, double* energy);

/* Replaces get_energy() (dmrg2_gpu.h:40). */
int dmrg_get_energy(dmrg_handle h, double* energy);
/* Energies at the end of each completed sweep: the count in *n, the first
 * min(cap, *n) of them in out (which may be NULL when cap <= 0).  No
 * reference counterpart: run keeps only the last one (dmrg2_gpu_impl.h:1440-1446). */
int dmrg_get_sweep_energies(dmrg_handle h, double* out, int cap, int* n);
/* Replaces chi_L / chi_R (dmrg2_gpu.h:42-43): dims[0..L] of all bonds. */
int dmrg_get_bond_dims(dmrg_handle h, int* dims);
/* Replaces get_mps (dmrg2_gpu_impl.h:1463-1472) for one site: copies
 * cL*d*cR scalars to `out`. */
int dmrg_get_mps(dmrg_handle h, int site, void* out);
/* Copy environment `idx` (left: L[idx], sites 0..idx-1 contracted; right:
 * R[idx], sites idx..L-1) of shape (chi, D, chi), layout (ket, mpo, bra), to
 * `out` (NULL: only report chi).  The reference keeps d_L_envs_/d_R_envs_
 * private (dmrg2_gpu.h:60-61); exposed here for checks and the CPU baseline. */
int dmrg_get_env(dmrg_handle h, int left, int idx, void* out, int* chi);
/* FP64 <psi|H|psi> / <psi|psi> of the current MPS by a full environment
 * contraction.  No reference counterpart; used to check a returned state. */
int dmrg_expectation(dmrg_handle h, double* energy, double* norm2);
/* Per-phase timings and counters; no reference counterpart beyond the
 * DMRG_GPU_PROFILE phase timers (gpu-rocm/common/gpu_opts.h:57-93). */
int dmrg_get_stats(dmrg_handle h, dmrg_stats* st);
/* Zero those counters (no reference counterpart). */
int dmrg_reset_stats(dmrg_handle h);

/* Message of the last failure on this thread ("" if none).  No reference
 * counterpart: the reference prints the message and throws
 * (gpu-rocm/common/hip_check.h:13-21). */
const char* dmrg_last_error(void);

/* ------------------------------------------------------------------------
 * Operator-level entry points on device buffers (for parity tests and the
 * roofline bench).  Each replaces one library-call sequence of the reference
 * and runs on the given HIP stream (0 = default stream); `ctx` owns scratch
 * and names the device, which each call makes current on its thread.
 * ---------------------------------------------------------------------- */
typedef struct dmrg_ctx* dmrg_ctx_handle;
/* Create / destroy an op-level context.  No reference counterpart: the
 * reference keeps its workspaces inside the engine object
 * (gpu-rocm/dmrg2-gpu/src/dmrg2_gpu.h:95-140). */
int dmrg_ctx_create(int device, dmrg_ctx_handle* out);
void dmrg_ctx_destroy(dmrg_ctx_handle c);

/* Two-site H_eff matvec: apply_heff_two_site (dmrg2_gpu_impl.h:606-780).
 * L (cL,D,cL), R (cR,D,cR), W1/W2 raw MPO tensors of the two sites (device),
 * theta/out (cL,d,d,cR).  dtype: DMRG_DTYPE_*. */
int dmrg_op_heff2(dmrg_ctx_handle c, void* stream, int dtype, int cL, int cR, int d, int D,
                  const void* L, const void* R, const void* W1, const void* W2,
                  const void* theta, void* out);
/* The two-site matvec of the mixed modes: L, R, theta, out in FP32 (layouts
 * as dmrg_op_heff2), W1/W2 the raw FP64 MPO tensors (device), fused on the
 * host and rounded to FP32; prec DMRG_PREC_FP32 (FP32 MFMA) or
 * DMRG_PREC_BF16X3 (split-BF16 products).  The reference has no FP32 path:
 * this is the mixed modes' substitute for apply_heff_two_site
 * (dmrg2_gpu_impl.h:606-780). */
int dmrg_op_heff2_f32(dmrg_ctx_handle c, void* stream, int prec, int cL, int cR, int d, int D,
                      const float* L, const float* R, const double* W1, const double* W2,
                      const float* theta, float* out);
/* Single-site matvec: DMRGGPU::apply_heff (dmrg_gpu_impl.h:538-705). */
int dmrg_op_heff1(dmrg_ctx_handle c, void* stream, int dtype, int cL, int cR, int d, int D,
                  const void* L, const void* R, const void* W, const void* theta, void* out);
/* update_left_env (dmrg2_gpu_impl.h:786-860): Lnew (co,D,co). */
int dmrg_op_env_left(dmrg_ctx_handle c, void* stream, int dtype, int ci, int co, int d, int D,
                     const void* L, const void* A, const void* W, void* Lnew);
/* Size of the fixed device scratch behind dmrg_op_env_left_noalloc: 16 MiB
 * per GPU, held from the first call on that GPU until the process ends. */
#define DMRG_NOALLOC_SCRATCH_BYTES (16u << 20)
/* dmrg_op_env_left without allocation.  No reference counterpart: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * kernels and result as dmrg_op_env_left, but no context: the two
 * intermediates (D*d*ci*co scalars each) live in the library's fixed device
 * scratch, so it returns BAD_ARG when 2*D*d*ci*co*sizeof(scalar) (each
 * rounded up to 256 bytes) exceeds DMRG_NOALLOC_SCRATCH_BYTES (e.g. FP64
 * ci = co = 256, d = 2, D = 5 fits).  Runs on the calling thread's current
 * device.  All calls on one GPU share that scratch: calls that overlap on
 * one GPU (two streams or two threads) race; calls in order on one stream
 * do not. */
int dmrg_op_env_left_noalloc(void* stream, int dtype, int ci, int co, int d, int D,
                             const void* L, const void* A, const void* W, void* Lnew);
/* update_right_env (dmrg2_gpu_impl.h:866-934): A (co,d,ci), R (ci,D,ci). */
int dmrg_op_env_right(dmrg_ctx_handle c, void* stream, int dtype, int ci, int co, int d, int D,
                      const void* R, const void* A, const void* W, void* Rnew);
/* Truncated SVD split of svd_split (dmrg2_gpu_impl.h:1166-1343): M (m x n),
 * keep k <= min(m,n) values, new rank by the cutoffs; left (m x nk) = U or
 * U*S, right (nk x n) = V^H or S*V^H (scale_left selects); sigma gets the nk
 * kept values (device, may be NULL).  *nk_out on the host. */
int dmrg_op_svd_split(dmrg_ctx_handle c, void* stream, int dtype, int m, int n, int k,
                      double cutoff, double rel_cutoff, int scale_left, const void* M,
                      void* left, void* right, double* sigma, int* nk_out);
/* Randomised truncated split (the reference's RSVD path,
 * dmrg2_gpu_impl.h:1186-1266): a sketch of r = min(k + oversample, m, n)
 * uniform random combinations (splitmix64 stream `seed`), `power`
 * re-orthonormalised subspace iterations, a block Gram-Schmidt basis Q and
 * the exact split of the projected r-dimensional matrix.  Outputs as
 * dmrg_op_svd_split (without sigma); exact when rank(M) <= r. */
int dmrg_op_rsvd_split(dmrg_ctx_handle c, void* stream, int dtype, int m, int n, int k,
                       int oversample, int power, uint64_t seed, double cutoff,
                       double rel_cutoff, int scale_left, const void* M, void* left,
                       void* right, int* nk_out);
/* Hermitian eigen-decomposition H = V diag(evals) V^H of a device matrix
 * H (N x N, column-major, N <= 2048) by the block-Jacobi solver behind
 * dmrg_op_svd_split; evals descending (device, N), evecs (device, N x N).
 * The reference reaches the same solver class through rocsolver_dgesvdj
 * (scalar_traits.h:166-188), whose V is the Jacobi eigenbasis of A^T A.
 * *sweeps_out (host, may be NULL) gets the Jacobi sweep count. */
int dmrg_op_heig(dmrg_ctx_handle c, void* stream, int dtype, int N, const void* H, double* evals,
                 void* evecs, int* sweeps_out);
/* Lowest eigenpair of the symmetric tridiagonal (alpha[0..k), beta[0..k-1))
 * on the device -- replaces rocsolver_dsteqr (dmrg2_gpu_impl.h:1110-1125).
 * Host arrays in, host results out. */
int dmrg_op_tridiag_lowest(dmrg_ctx_handle c, int k, const double* alpha, const double* beta,
                           double* lam, double* y);
/* One Lanczos solve on the two-site H_eff (lanczos_eigensolver,
 * dmrg2_gpu_impl.h:970-1160).  theta is overwritten with the Ritz vector. */
int dmrg_op_lanczos2(dmrg_ctx_handle c, void* stream, int dtype, int cL, int cR, int d, int D,
                     const void* L, const void* R, const void* W1, const void* W2, void* theta,
                     int krylov_max, int site, double* energy, int* niter);

/* C = alpha*op(A)*op(B) + beta*C on device buffers (column-major; op: 0 N,
 * 1 T, 2 conj, 3 conj-transpose).  Replaces the Traits::gemm dispatch of
 * gpu-rocm/common/scalar_traits.h:46-58 (rocblas_dgemm / rocblas_zgemm).
 * alpha/beta point to one scalar of the dtype (host memory). */
int dmrg_op_gemm(dmrg_ctx_handle c, void* stream, int dtype, int opA, int opB, int M, int N,
                 int K, const void* alpha, const void* A, int lda, const void* B, int ldb,
                 const void* beta, void* Cm, int ldc);
/* FP32 C = op(A) op(B) (op 0 N, 1 T) as the mixed modes run it: prec
 * DMRG_PREC_FP32 (FP32 MFMA) or DMRG_PREC_BF16X3 (split-BF16 products).  The
 * reference has no FP32 GEMM; this is the matvec GEMM of the mixed modes
 * (their substitute for the rocblas_dgemm calls of apply_heff_two_site,
 * dmrg2_gpu_impl.h:686-755). */
int dmrg_op_gemm_f32(dmrg_ctx_handle c, void* stream, int prec, int opA, int opB, int M, int N,
                     int K, const float* A, int lda, const float* B, int ldb, float* Cm, int ldc);

/* Device-memory helpers so callers need no other HIP binding (no reference
 * counterpart: the reference's callers never hold device memory). */
int dmrg_dev_malloc(void** p, size_t bytes);
int dmrg_dev_free(void* p);
int dmrg_memcpy_h2d(void* dst, const void* src, size_t bytes);
int dmrg_memcpy_d2h(void* dst, const void* src, size_t bytes);
/* Device-to-device copy (a copy kernel of known size; the traffic
 * calibration of the profiling runs uses it).  No reference counterpart. */
int dmrg_memcpy_d2d(void* dst, const void* src, size_t bytes);
/* Wait for all device work; visible HIP devices (0 without a GPU).  No
 * reference counterpart. */
int dmrg_dev_synchronize(void);
int dmrg_device_count(int* n);

#ifdef __cplusplus
}
#endif
#endif /* DMRG_AMD_H */
