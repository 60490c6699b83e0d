// DMRG sweep engine (host side, C++), the MI355X counterpart of the
// reference's DMRG2GPU<Scalar> (gpu-rocm/dmrg2-gpu/src/dmrg2_gpu.h:23-201) and
// DMRGGPU<Scalar> (gpu-rocm/dmrg-gpu/src/dmrg_gpu.h:23-219).
#pragma once
#include <chrono>
#include <functional>
#include <vector>

#include "common.hpp"
#include "contract.hpp"
#include "lanczos.hpp"
#include "svd.hpp"
#include "warm.hpp"

namespace dmrg {

struct Options {
  int krylov_max = 100;
  double lanczos_tol = 1e-12;
  double eig_tol = 1e-12;
  double svd_cutoff = 1e-14;
  double svd_rel_cutoff = 0.0;
  int precision = 0;  // 0 fp64, 1 fp32 matvec
  int profile = 0;
  int device = 0;
  int svd_max_sweeps = 40;
  double svd_tol = 1e-14;
  int svd_warm = 1;  // two-site: start each split's Jacobi from the previous sweep's basis
  // FP32 matvec mode: the Lanczos Ritz value cannot settle below the FP32
  // rounding of H_eff, so the check also accepts |dE| < fp32_eig_rel |E|.
  double fp32_eig_rel = 1e-9;
  // Truncation: converge only the kept/discarded coupling of each split
  // (same truncated state; the kept vectors are a basis of the kept space).
  int svd_subspace = 1;
  // Jacobi stop of the split in the FP32 / BF16X3 matvec modes.
  double svd_tol_mixed = 1e-13;
};

struct Stats {
  double lanczos_ms = 0, heff_ms = 0, svd_ms = 0, env_ms = 0, sweep_s_last = 0;
  double heff_flop = 0, env_flop = 0;
  long long matvecs = 0, lanczos_calls = 0, svd_calls = 0, svd_sweeps = 0;
  int sweeps_done = 0;
  double max_discarded = 0;
  double env_build_s = 0, sweeps_s_total = 0;
  long long svd_warm_starts = 0;
  long long svd_bip_sweeps = 0, svd_oa_steps = 0;
};

// Algorithmic flop counts (real arithmetic; complex = 4x).
inline double flops_heff2(int cL, int cR, int d, int D, bool cplx) {
  double dd = (double)d * d;
  double f = 2.0 * D * dd * cL * cR * ((double)cL + cR) + 2.0 * cL * cR * (D * dd) * (D * dd);
  return cplx ? 4 * f : f;
}
inline double flops_heff1(int cL, int cR, int d, int D, bool cplx) {
  double f = 2.0 * D * d * cL * (double)cR * ((double)cL + cR) + 2.0 * cL * (double)cR * (D * d) * (D * d);
  return cplx ? 4 * f : f;
}
inline double flops_env(int ci, int co, int d, int D, bool cplx) {
  double f = 2.0 * D * d * ci * (double)co * ((double)ci + co) + 2.0 * ci * (double)co * (D * d) * (D * d);
  return cplx ? 4 * f : f;
}

// One Lanczos solve (reference semantics, dmrg2_gpu_impl.h:970-1160) with a
// caller-supplied matvec.  theta (length n) is replaced by the Ritz vector.
template <class T>
struct LanczosRunner {
  LzDev* lz = nullptr;
  T* V = nullptr;      // kmax * n
  T* work = nullptr;   // n
  int* h_status = nullptr;   // pinned [4]
  double* h_scal = nullptr;  // pinned [8]
  std::vector<hipEvent_t> ev;  // 2 per matvec
  bool timed = false;          // record the events (option profile): H_eff time in the stats
  double beta_tol = 1e-12, eig_tol = 1e-12, eig_rel = 0.0;
  int last_niter = 0, last_matvecs = 0;
  double last_heff_ms = 0;
  double run(hipStream_t st, T* theta, size_t n, int kmax, int site,
             const std::function<void(const T*, T*)>& matvec);
};

template <class T>
class Engine {
 public:
  Engine(int L, int d, int chi, int D, double tol, int variant, const Options& opt);
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  void set_mpo(const T* const* W);
  void init_random(uint64_t seed, double scale);
  void init_product();
  void init_neel();
  void set_mps(int site, const T* A, int cL, int cR);
  double run(int n_sweeps, bool quiet);
  void get_mps(int site, T* out) const;
  void expectation(double* e, double* nrm2);
  int get_env(bool left, int idx, T* out);

  int L, d, chi, D;
  double tol;
  int variant;
  Options opt;
  std::vector<int> bond;
  double energy = 0.0;
  std::vector<double> sweep_E;
  Stats stats;

 private:
  // Shape (cL, cR) of the tensor each site holds.  set_mps changes one site's
  // bonds, so run and expectation check every site against bond[] first.
  std::vector<int> tl_, tr_;
  void sync_shapes();
  void check_shapes(const char* who) const;
  void alloc_all();
  void free_all();
  void build_envs();
  void wait_env();
  void env_left_at(int site);
  void env_right_at(int site);  // builds R[site] from R[site+1] and A[site]
  void flush_env_timing();
  std::vector<hipEvent_t> env_ev;  // option profile: event pairs around env updates
  size_t env_ev_n = 0;
  bool env_timing = false;
  double optimize_bond(int site, char dir, bool rq);
  double optimize_site(int site, char dir, bool split);
  double sweep2_lr();
  double sweep2_rl();
  double sweep1_lr();
  double sweep1_rl();
  double lanczos_two(int site, bool rq);
  double lanczos_one(int site);
  void matvec2(int site, const T* in, T* out, bool fp64);
  void split_two(int site, char dir);
  // warm start of the truncation SVD (two-site; see warm.hpp)
  void warm_alloc();
  void warm_free();
  void warm_begin();
  const T* warm_basis(int site, char dir);
  void warm_after_split(int site, char dir);
  void split_one(int site, char dir);
  double rayleigh2(int site);

  hipStream_t s_main = nullptr, s_env = nullptr;
  hipEvent_t ev_canon = nullptr, ev_env = nullptr;
  bool env_pending = false;
  bool mpo_set = false;

  std::vector<T*> mps;
  T* mps_spare = nullptr;
  std::vector<T*> Lenv, Renv, Wd, Wrd, WWd;
  std::vector<float*> WWf;
  float *Lf = nullptr, *Rf = nullptr, *thf = nullptr, *outf = nullptr;
  double* envScale = nullptr;  // max |L|, max |R| of the current bond (FP32 mode)
  T *theta = nullptr, *work = nullptr, *V = nullptr;
  ContractScratch<T> sc_main, sc_env;
  ContractScratch<float> sc_f;
  LzDev lz;
  SvdWork<T> svdw;
  SvdParams svd_params() const {
    SvdParams p;
    p.max_sweeps = opt.svd_max_sweeps;
    p.tol = opt.precision != 0 ? opt.svd_tol_mixed : opt.svd_tol;
    p.subspace = opt.svd_subspace != 0;
    return p;
  }
  LanczosRunner<T> runner;
  bool warm_on = false;
  std::vector<T*> wX[2];           // per bond: last sorted eigenbasis (N x N)
  std::vector<int> wXN[2], wXB[2];  // its N and block bond dimension (0: none)
  std::vector<T*> wA[2];           // per site: canonical tensor of the last sweep
  std::vector<int> wAL[2], wAR[2];  // its bond dimensions (0: none)
  T* wT[2] = {nullptr, nullptr};   // running block overlap (ping-pong)
  T *wZ0 = nullptr, *wZ1 = nullptr, *wP = nullptr, *wTmp = nullptr, *wXw = nullptr;
  double* wDev = nullptr;
  double* wHost = nullptr;  // pinned read-back of the polar iteration's deviation
  int wTr = 0, wTc = 0, wCur = 0;
  int* h_status = nullptr;
  double* h_scal = nullptr;
  T* h_aux = nullptr;  // pinned [8]: Rayleigh quotient and expectation read-backs
  T *ex_a = nullptr, *ex_b = nullptr, *ex_id = nullptr;  // expectation(): envs, identity MPO
  size_t nmax = 0;
  int kmax = 0;
};

}  // namespace dmrg
