// Segment-parallel two-site sweep on one GPU (PDMRG): the reference's
// pdmrg-gpu (gpu-rocm/pdmrg-gpu/src/pdmrg_gpu.h:13-60,
// pdmrg_gpu_impl.h:2315-2753), after Stoudenmire and White.  The chain is cut
// into P segments of consecutive sites; each segment sweeps its own bonds on
// its own stream and host thread, concurrently with the others, and the
// boundary bonds between segments are then optimised from
// theta = psi_L diag(V) psi_R with V = 1/S of the boundary's last split.
//
// Staggering (pdmrg_gpu_impl.h:2674-2696): even segments sweep left to right
// while odd ones sweep right to left, so every even boundary then has fresh
// environments on both sides and is merged; the directions then swap and the
// odd boundaries are merged.  Segments touch disjoint environments: segment
// k's left-to-right pass writes L[first+1 .. last] and reads R[.. last+1],
// which the right-to-left pass of segment k+1 (writing R[last+2 ..]) leaves
// alone.
#include <algorithm>
#include <cmath>
#include <exception>
#include <thread>

#include "engine.hpp"

namespace dmrg {
namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// dst(i, j) = src(i, j) * v[i]  (rows x cols, ld rows); inv: / max(v[i], reg)
template <class T>
__global__ void k_scale_rows(T* __restrict__ dst, const T* __restrict__ src, int rows, int cols,
                             const double* __restrict__ v, int inv, double reg) {
  const size_t tot = (size_t)rows * cols;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % rows);
    const double f = inv ? 1.0 / fmax(v[i], reg) : v[i];
    dst[e] = src[e] * f;
  }
}
// V[i] = 1 / max(s[i], reg) for i < n
__global__ void k_inv_clip(double* __restrict__ V, const double* __restrict__ s, int n, double reg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) V[i] = 1.0 / fmax(s[i], reg);
}

template <class T>
void scale_rows(hipStream_t st, T* dst, const T* src, int rows, int cols, const double* v, bool inv,
                double reg) {
  const size_t tot = (size_t)rows * cols;
  if (!tot) return;
  hipLaunchKernelGGL(k_scale_rows<T>, dim3(std::min<size_t>(cdiv(tot, 256), 8192)), dim3(256), 0,
                     st, dst, src, rows, cols, v, inv ? 1 : 0, reg);
  DMRG_LAUNCH_CHECK();
}

// Clip of the singular values in V = 1/S (pdmrg_gpu_impl.h:2499-2505).
constexpr double kVReg = 1e-12;

void add_stats(Stats& a, const Stats& b) {
  a.lanczos_ms += b.lanczos_ms;
  a.heff_ms += b.heff_ms;
  a.svd_ms += b.svd_ms;
  a.env_ms += b.env_ms;
  a.heff_flop += b.heff_flop;
  a.env_flop += b.env_flop;
  a.matvecs += b.matvecs;
  a.lanczos_calls += b.lanczos_calls;
  a.svd_calls += b.svd_calls;
  a.svd_sweeps += b.svd_sweeps;
  a.max_discarded = std::max(a.max_discarded, b.max_discarded);
  a.trunc_sum_cur += b.trunc_sum_cur;
}

// Run f(i) for i < n on n host threads (one per segment or boundary), each
// on the engine's device; the first exception is rethrown after all join.
template <class F>
void parallel_for(int n, int device, F&& f) {
  std::vector<std::thread> th;
  std::vector<std::exception_ptr> ex(n);
  for (int i = 0; i < n; ++i)
    th.emplace_back([&, i] {
      try {
        DMRG_HIP(hipSetDevice(device));
        f(i);
      } catch (...) {
        ex[i] = std::current_exception();
      }
    });
  for (auto& t : th) t.join();
  for (auto& e : ex)
    if (e) std::rethrow_exception(e);
}

}  // namespace

template <class T>
void SegWork<T>::alloc(int chi, int d, int D, int kmax, const Options& opt) {
  DMRG_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  const size_t n = (size_t)chi * d * d * chi;
  auto dm = [](void** p, size_t b) { DMRG_HIP(hipMalloc(p, b)); };
  dm((void**)&theta, n * sizeof(T));
  dm((void**)&work, n * sizeof(T));
  dm((void**)&tmp, (size_t)chi * d * chi * sizeof(T));
  dm((void**)&V, (size_t)kmax * n * sizeof(T));
  dm((void**)&sigma, (size_t)chi * d * sizeof(double));
  sc.cap = (size_t)D * d * d * chi * chi;
  dm((void**)&sc.X, sc.cap * sizeof(T));
  dm((void**)&sc.Y, sc.cap * sizeof(T));
  svd.alloc(chi * d);
  lz.alloc(std::max(kmax, 1), lz_part_doubles(std::max(kmax, 1), n) * sizeof(double), sizeof(T));
  DMRG_HIP(hipHostMalloc((void**)&h_status, 4 * sizeof(int), hipHostMallocDefault));
  DMRG_HIP(hipHostMalloc((void**)&h_scal, 8 * sizeof(double), hipHostMallocDefault));
  runner.lz = &lz;
  runner.V = V;
  runner.work = work;
  runner.h_status = h_status;
  runner.h_scal = h_scal;
  runner.beta_tol = opt.lanczos_tol;
  runner.eig_tol = opt.eig_tol;
  runner.eig_rel = 0.0;
}

template <class T>
void SegWork<T>::release() {
  if (st) (void)hipStreamSynchronize(st);
  for (void* p : {(void*)theta, (void*)work, (void*)tmp, (void*)V, (void*)sigma, (void*)sc.X,
                  (void*)sc.Y})
    if (p) (void)hipFree(p);
  theta = work = tmp = V = nullptr;
  sigma = nullptr;
  sc.X = sc.Y = nullptr;
  svd.release();
  lz.release();
  for (auto e : runner.ev) (void)hipEventDestroy(e);
  runner.ev.clear();
  runner.clear_graphs();
  if (h_status) (void)hipHostFree(h_status);
  if (h_scal) (void)hipHostFree(h_scal);
  h_status = nullptr;
  h_scal = nullptr;
  if (st) (void)hipStreamDestroy(st);
  st = nullptr;
}

template <class T>
void Engine<T>::seg_free() {
  for (auto* w : segw) {
    w->release();
    delete w;
  }
  segw.clear();
  for (auto p : pd_V)
    if (p) (void)hipFree(p);
  pd_V.clear();
  pd_Vn.clear();
}

// Optimise bond (site, site+1) on segment work space w: theta from the two
// site tensors (with diag(vb) between them at a boundary), Lanczos, split.
template <class T>
double Engine<T>::seg_bond(SegWork<T>& w, int site, char dir, const double* vb) {
  const int cL = bond[site], cm = bond[site + 1], cR = bond[site + 2];
  const T* B = mps[site + 1];
  if (vb) {  // psi_R with its rows scaled by V (pdmrg_gpu_impl.h:2356-2390)
    scale_rows<T>(w.st, w.tmp, mps[site + 1], cm, d * cR, vb, false, 0.0);
    B = w.tmp;
  }
  gemm<T>(w.st, OP_N, OP_N, cL * d, d * cR, cm, T(1), mps[site], cL * d, B, cm, T(0), w.theta,
          cL * d, gemm_pick_tile(cL * d, d * cR, cm, is_cplx<T>::value));
  const size_t n = (size_t)cL * d * d * cR;
  double t0 = now_s();
  const double E = w.runner.run(w.st, w.theta, n, opt.krylov_max, site, [&](const T* in, T* out) {
    heff2_apply<T>(w.st, Lenv[site], Renv[site + 2], WWd[site], in, out, cL, cR, d, D, w.sc);
  });
  w.stats.lanczos_ms += 1e3 * (now_s() - t0);
  w.stats.matvecs += w.runner.last_matvecs;
  w.stats.lanczos_calls += 1;
  w.stats.heff_flop += w.runner.last_matvecs * flops_heff2(cL, cR, d, D, is_cplx<T>::value);
  // split; a boundary needs the true singular values (V = 1/S), so every
  // pair is converged there
  const int m = cL * d, nn = d * cR;
  const int k = std::min(std::min(m, nn), chi);
  SvdParams prm = svd_params();
  if (vb) prm.subspace = false;
  t0 = now_s();
  SvdResult r = svd_split<T>(w.st, w.svd, w.theta, m, nn, m, k, opt.svd_cutoff, opt.svd_rel_cutoff,
                             dir == 'L', mps[site], m, mps[site + 1], vb ? w.sigma : nullptr, prm);
  bond[site + 1] = r.nk;
  w.stats.svd_ms += 1e3 * (now_s() - t0);
  w.stats.svd_calls += 1;
  w.stats.svd_sweeps += r.sweeps;
  w.stats.max_discarded = std::max(w.stats.max_discarded, r.discarded);
  w.stats.trunc_sum_cur += r.discarded;
  return E;
}

// Sweep the bonds first .. last-1 of one segment on its work space.
template <class T>
void Engine<T>::seg_sweep(SegWork<T>& w, int first, int last, char dir) {
  if (dir == 'R') {
    for (int site = first; site < last; ++site) {
      w.energy = seg_bond(w, site, 'R', nullptr);
      env_left<T>(w.st, Lenv[site], mps[site], Wd[site], Lenv[site + 1], bond[site],
                  bond[site + 1], d, D, w.sc);
      w.stats.env_flop += flops_env(bond[site], bond[site + 1], d, D, is_cplx<T>::value);
    }
  } else {
    for (int site = last - 1; site >= first; --site) {
      w.energy = seg_bond(w, site, 'L', nullptr);
      env_right<T>(w.st, Renv[site + 2], mps[site + 1], Wrd[site + 1], Renv[site + 1],
                   bond[site + 2], bond[site + 1], d, D, w.sc);
      w.stats.env_flop += flops_env(bond[site + 2], bond[site + 1], d, D, is_cplx<T>::value);
    }
  }
  DMRG_HIP(hipStreamSynchronize(w.st));
}

template <class T>
double Engine<T>::run_pdmrg(const PdmrgParams& p, bool quiet) {
  if (!mpo_set) throw Error(ST_STATE, "run_pdmrg: set_mpo has not been called");
  DMRG_REQUIRE(variant == 2 && opt.precision == 0,
               "run_pdmrg: the segment-parallel sweep is two-site FP64 / complex128");
  DMRG_REQUIRE(p.segments >= 2 && L >= 2 * p.segments,
               "run_pdmrg: need at least 2 segments of at least 2 sites");
  DMRG_REQUIRE(p.outer >= 0 && p.local >= 1 && p.warmup >= 0 && p.polish >= 0,
               "run_pdmrg: negative sweep counts");
  check_shapes("run_pdmrg");
  auto resync = [this] { sync_shapes(); };
  struct Resync {
    decltype(resync)& f;
    ~Resync() { f(); }
  } resync_at_exit{resync};
  DMRG_HIP(hipSetDevice(opt.device));
  const int P = p.segments;
  if ((int)segw.size() != P) {
    seg_free();
    for (int i = 0; i < P; ++i) {
      segw.push_back(new SegWork<T>());
      segw.back()->alloc(chi, d, D, kmax, opt);
    }
  }
  for (auto* w : segw) w->stats = Stats();
  double te = now_s();
  build_envs();
  const double t_start = now_s();
  stats.env_build_s = t_start - te;
  // warm-up: serial full two-site sweeps
  for (int sw = 0; sw < p.warmup; ++sw) {
    sweep2_lr();
    energy = sweep2_rl();
  }
  DMRG_HIP(hipStreamSynchronize(s_main));
  DMRG_HIP(hipStreamSynchronize(s_env));
  // partition (pdmrg_gpu_impl.h:582-601): lengths differ by at most one
  std::vector<int> first(P), last(P);
  for (int k = 0, pos = 0; k < P; ++k) {
    const int len = L / P + (k < L % P ? 1 : 0);
    first[k] = pos;
    last[k] = pos + len - 1;
    pos += len;
  }
  const int NB = P - 1;
  // V = ones at every boundary (pdmrg_gpu_impl.h:607-617)
  for (auto q : pd_V)
    if (q) (void)hipFree(q);
  pd_V.assign(NB, nullptr);
  pd_Vn.assign(NB, 0);
  for (int b = 0; b < NB; ++b) {
    DMRG_HIP(hipMalloc((void**)&pd_V[b], (size_t)chi * d * sizeof(double)));
    pd_Vn[b] = bond[last[b] + 1];
    fill_f64(s_main, pd_V[b], (size_t)pd_Vn[b], 1.0);
  }
  DMRG_HIP(hipStreamSynchronize(s_main));
  // merge one boundary on work space w (pdmrg_gpu_impl.h:2407-2578)
  auto merge = [&](int b, SegWork<T>& w) {
    const int site = last[b];
    DMRG_REQUIRE(pd_Vn[b] == bond[site + 1], "run_pdmrg: boundary V out of step with its bond");
    w.energy = seg_bond(w, site, 'R', pd_V[b]);
    const int nk = bond[site + 1], cR = bond[site + 2];
    hipLaunchKernelGGL(k_inv_clip, dim3(cdiv(nk, 256)), dim3(256), 0, w.st, pd_V[b],
                       (const double*)w.sigma, nk, kVReg);
    DMRG_LAUNCH_CHECK();
    pd_Vn[b] = nk;
    // left environment from U, right environment from the canonical Vh =
    // diag(1/S) (S Vh) (the site keeps S Vh)
    env_left<T>(w.st, Lenv[site], mps[site], Wd[site], Lenv[site + 1], bond[site], nk, d, D, w.sc);
    scale_rows<T>(w.st, w.tmp, mps[site + 1], nk, d * cR, w.sigma, true, kVReg);
    env_right<T>(w.st, Renv[site + 2], w.tmp, Wrd[site + 1], Renv[site + 1], cR, nk, d, D, w.sc);
    DMRG_HIP(hipStreamSynchronize(w.st));
  };
  auto merge_parity = [&](int parity) {
    std::vector<int> bs;
    for (int b = parity; b < NB; b += 2) bs.push_back(b);
    if (bs.empty()) return;
    parallel_for((int)bs.size(), opt.device, [&](int i) { merge(bs[i], *segw[i % P]); });
    energy = segw[(bs.size() - 1) % P]->energy;
  };
  double e_prev = energy;
  int outer_done = 0;
  for (int outer = 0; outer < p.outer; ++outer) {
    ++outer_done;
    double e_local = energy;
    for (int loc = 0; loc < p.local; ++loc) {
      parallel_for(P, opt.device,
                   [&](int k) { seg_sweep(*segw[k], first[k], last[k], k % 2 == 0 ? 'R' : 'L'); });
      merge_parity(0);
      parallel_for(P, opt.device,
                   [&](int k) { seg_sweep(*segw[k], first[k], last[k], k % 2 == 0 ? 'L' : 'R'); });
      merge_parity(1);
      if (!std::isfinite(energy)) throw Error(ST_NONFINITE, "run_pdmrg: non-finite energy");
      const double dl = std::fabs(energy - e_local);
      e_local = energy;
      if (dl < tol && loc > 0) break;
    }
    const double dE = std::fabs(energy - e_prev);
    e_prev = energy;
    if (dE < tol) break;
  }
  // polish: serial full sweeps from freshly built environments
  if (p.polish > 0) {
    build_envs();
    for (int sw = 0; sw < p.polish; ++sw) {
      sweep2_lr();
      energy = sweep2_rl();
    }
  }
  DMRG_HIP(hipStreamSynchronize(s_main));
  DMRG_HIP(hipStreamSynchronize(s_env));
  for (auto* w : segw) add_stats(stats, w->stats);
  const double total = now_s() - t_start;
  stats.sweeps_s_total += total;
  sweep_E.push_back(energy);
  if (!quiet) {
    printf("Converged after %d outer iterations of %d segments\n", outer_done, P);
    printf("Final energy: %.12f\n", energy);
    printf("Total wall time: %.3f s\n", total);
    printf("  env_build_sec: %.3f  timer_scope: sweep_only\n", stats.env_build_s);
    fflush(stdout);
  }
  return energy;
}

template struct SegWork<double>;
template struct SegWork<cdouble>;
template double Engine<double>::run_pdmrg(const PdmrgParams&, bool);
template double Engine<cdouble>::run_pdmrg(const PdmrgParams&, bool);
template void Engine<double>::seg_free();
template void Engine<cdouble>::seg_free();

}  // namespace dmrg
