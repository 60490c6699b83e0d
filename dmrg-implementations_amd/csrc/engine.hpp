// DMRG sweep engine (host side, C++), the MI355X counterpart of the
// reference's DMRG2GPU<Scalar> (gpu-rocm/dmrg2-gpu/src/dmrg2_gpu.h:23-201) and
// DMRGGPU<Scalar> (gpu-rocm/dmrg-gpu/src/dmrg_gpu.h:23-219).
#pragma once
#include <chrono>
#include <functional>
#include <unordered_map>
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
  // FP32 / BF16X3 modes: 1 keeps the environments in FP32 (scaled) and runs
  // their updates on FP32 MFMA, the sweep energy then coming from an FP64
  // contraction of the state; 0 (default) keeps FP64 environments, converted
  // per bond.  FP32 environments accumulate FP32 rounding along the chain
  // (~1e-6 of H_eff against ~1e-7 from the matvec alone), which measured
  // |dE|/|E| = 2.4e-10 against FP64 at Heisenberg L=128, chi=512 after 10
  // sweeps, against 2e-12 with FP64 environments.
  int env_precision = 0;
  // Randomised split (the reference's RSVD option): used for a two-site
  // split when it keeps fewer than min(m, n) - oversample values and
  // m > 2 k, as in the reference; power = subspace iterations.
  int svd_rsvd = 0;
  int rsvd_oversample = 10;
  int rsvd_power = 0;
  // Replay each bond's Lanczos iterations between two convergence checks as
  // a captured HIP graph (cached per bond and bond dimensions): one launch
  // instead of ~8 per iteration.  Off with profile (events are not captured).
  int graph = 0;
  // Turn on the process-wide GEMM tile autotuner (gemm_autotune) at creation.
  int autotune = 0;
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
  long long rsvd_calls = 0;
  long long graph_launches = 0, graph_captures = 0;
  long long tuned_shapes = 0;  // GEMM shapes the autotuner has timed (process-wide)
  double trunc_sum_last = 0;  // discarded weight summed over the last sweep's splits
  double trunc_sum_cur = 0;   // running sum of the sweep in progress
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
  // graph_key != 0 (and not timed): each stretch of iterations between two
  // convergence checks runs as a HIP graph captured on first use and cached
  // under (graph_key, stretch); the key must change whenever anything the
  // stretch's launches read as an argument changes (site, bond dimensions).
  double run(hipStream_t st, T* theta, size_t n, int kmax, int site,
             const std::function<void(const T*, T*)>& matvec, uint64_t graph_key = 0);
  struct GraphEntry {
    hipGraphExec_t exec = nullptr;
    size_t n = 0;
    int a = 0, b = 0;
  };
  std::unordered_map<uint64_t, GraphEntry> graphs;
  long long graph_launches = 0, graph_captures = 0;
  void clear_graphs();
};

// Per-segment work space of the segment-parallel sweep (pdmrg.hip): its own
// stream, two-site vectors, Krylov basis, contraction scratch, split and
// Lanczos state, so segments run concurrently on one device.
template <class T>
struct SegWork {
  hipStream_t st = nullptr;
  T *theta = nullptr, *work = nullptr, *V = nullptr, *tmp = nullptr;
  double* sigma = nullptr;  // singular values of a boundary split
  ContractScratch<T> sc;
  SvdWork<T> svd;
  LzDev lz;
  LanczosRunner<T> runner;
  int* h_status = nullptr;
  double* h_scal = nullptr;
  Stats stats;
  double energy = 0.0;
  void alloc(int chi, int d, int D, int kmax, const Options& opt);
  void release();
};

// Options of the segment-parallel run (the reference's PDMRGGPU::run
// arguments, gpu-rocm/pdmrg-gpu/src/pdmrg_gpu.h:40-41).
struct PdmrgParams {
  int segments = 2;
  int outer = 10;   // outer iterations (stop when the energy changes by < tol)
  int local = 2;    // staggered segment sweeps + boundary merges per outer iteration
  int warmup = 1;   // serial full sweeps before the parallel phase
  int polish = 1;   // serial full sweeps after it
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
  // Run on the caller's stream(s) until restore_streams(): both of the
  // engine's own streams are drained first, and `main` may equal `env` (then
  // the environment updates queue behind the bond work).  The batch runner
  // lends each worker thread's stream to the chains it runs, so a hardware
  // queue carries whole chains rather than every chain's main or env stream.
  void borrow_streams(hipStream_t main, hipStream_t env);
  void restore_streams();
  // Segment-parallel sweep (PDMRG, the reference's pdmrg-gpu,
  // gpu-rocm/pdmrg-gpu/src/pdmrg_gpu_impl.h:2407-2753), two-site FP64.
  double run_pdmrg(const PdmrgParams& p, bool quiet);
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
  // FP32 environments (mixed modes with env_precision = 1): Lfe/Rfe hold
  // the unit-max float arrays, envS the scales (envS[i] of Lfe[i],
  // envS[L + 1 + i] of Rfe[i]); Af stages one site tensor in FP32.
  bool env32 = false;
  std::vector<float*> Lfe, Rfe, Wf, Wrf;
  double* envS = nullptr;
  double* envMax = nullptr;
  float* Af = nullptr;
  ContractScratch<float> sc_envf;
  void env_left_at32(int site);
  void env_right_at32(int site);
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
  // Graph-cache key of a bond's Lanczos launches (0: no graphs): every
  // argument they take is fixed by the site and its bond dimensions.
  uint64_t graph_key(int site, int cL, int cR) const {
    if (!opt.graph || opt.profile) return 0;
    return 1 + (uint64_t)site + (uint64_t)(L + 1) * ((uint64_t)cL + (uint64_t)(chi + 1) * cR);
  }

  // segment-parallel run (pdmrg.hip)
  std::vector<SegWork<T>*> segw;
  void seg_free();
  double seg_bond(SegWork<T>& w, int site, char dir, const double* vb);
  void seg_sweep(SegWork<T>& w, int first, int last, char dir);
  std::vector<double*> pd_V;  // per boundary: V = 1/clip(S) on the device
  std::vector<int> pd_Vn;     // its length (the boundary bond dimension)

  hipStream_t s_main = nullptr, s_env = nullptr;
  hipStream_t own_main = nullptr, own_env = nullptr;  // while borrowing: the engine's streams
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
  RsvdWork<T> rsvdw;
  uint64_t rsvd_seq = 0;  // RSVD calls so far: the sketch of call j uses stream 0x5EED0000 + j
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
