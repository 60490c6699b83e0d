// DMRG sweep engine: sweep driver, Lanczos driver, truncation split.
// Follows gpu-rocm/dmrg2-gpu/src/dmrg2_gpu_impl.h (two-site) and
// gpu-rocm/dmrg-gpu/src/dmrg_gpu_impl.h (single-site) step by step; the
// library calls there are replaced by the kernels of gemm/contract/lanczos/svd.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace dmrg {

namespace {
template <class F>
struct OnExit {
  F f;
  ~OnExit() { f(); }
};

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
template <class T> bool finite_v(T v) { return std::isfinite(real(v)) && std::isfinite(imag(v)); }
}  // namespace

// ---------------------------------------------------------------------------
// Lanczos driver (dmrg2_gpu_impl.h:970-1160)
// ---------------------------------------------------------------------------
template <class T>
double LanczosRunner<T>::run(hipStream_t st, T* theta, size_t n, int kmax, int site,
                             const std::function<void(const T*, T*)>& matvec,
                             uint64_t graph_key) {
  const int max_iter = (int)std::min<size_t>((size_t)kmax, n);
  DMRG_REQUIRE(max_iter >= 1, "lanczos: empty problem");
  if (timed && (int)ev.size() < 2 * max_iter) {
    size_t old = ev.size();
    ev.resize(2 * max_iter);
    for (size_t i = old; i < ev.size(); ++i) DMRG_HIP(hipEventCreate(&ev[i]));
  }
  const bool use_graph = graph_key != 0 && !timed;
  int nmv = 0;
  // Iterations [a, b): the prologue (v0 = theta/|theta|; a zero-norm start
  // replaced by the seeded stream 42+site) before the first, the convergence
  // check after iteration b-1 when `check`.
  auto stretch = [&](int a, int b, bool check) {
    if (a == 0) {
      lz_reset(st, *lz);
      lz_norm<T>(st, *lz, theta, n);
      lz_fill_if_small<T>(st, *lz, theta, n, (uint64_t)(42 + site), 1e-14);
      lz_norm<T>(st, *lz, theta, n);
      lz_scale<T>(st, *lz, V, theta, n);
    }
    for (int it = a; it < b; ++it) {
      T* vi = V + (size_t)it * n;
      if (timed) DMRG_HIP(hipEventRecord(ev[2 * it], st));
      matvec(vi, work);
      if (timed) DMRG_HIP(hipEventRecord(ev[2 * it + 1], st));
      lz_alpha<T>(st, *lz, vi, work, n, it);
      lz_sub_dots<T>(st, *lz, work, V, n, it);
      lz_orth_norm<T>(st, *lz, work, V, n, it);
      lz_next<T>(st, *lz, it + 1 < max_iter ? V + (size_t)(it + 1) * n : nullptr, work, n, it);
    }
    if (check) {
      lz_check(st, *lz, b, beta_tol, eig_tol, eig_rel);
      DMRG_HIP(hipMemcpyAsync(h_status, lz->status, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    }
  };
  int niter = max_iter;
  int a = 0, chunk = 0;
  while (a < max_iter) {
    // the reference checks after iterations 6, 9, 12, ... (it >= 4, it % 3 == 0)
    int b = a;
    while (b < max_iter && !(b >= 4 && b % 3 == 0)) ++b;
    const bool check = b < max_iter;
    if (check) ++b;  // iteration b-1 is the checked one
    if (use_graph) {
      const uint64_t key = graph_key * 64 + (uint64_t)chunk;
      auto itg = graphs.find(key);
      if (itg != graphs.end() && (itg->second.n != n || itg->second.a != a || itg->second.b != b)) {
        (void)hipGraphExecDestroy(itg->second.exec);
        graphs.erase(itg);
        itg = graphs.end();
      }
      if (itg == graphs.end()) {
        hipGraph_t g = nullptr;
        DMRG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
          stretch(a, b, check);
        } catch (...) {
          (void)hipStreamEndCapture(st, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        DMRG_HIP(hipStreamEndCapture(st, &g));
        GraphEntry ge;
        hipError_t e = hipGraphInstantiate(&ge.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        DMRG_HIP(e);
        ge.n = n;
        ge.a = a;
        ge.b = b;
        itg = graphs.emplace(key, ge).first;
        ++graph_captures;
      }
      DMRG_HIP(hipGraphLaunch(itg->second.exec, st));
      ++graph_launches;
    } else {
      stretch(a, b, check);
    }
    nmv += b - a;
    if (check) {
      DMRG_HIP(hipStreamSynchronize(st));
      if (h_status[0]) {
        niter = h_status[1];
        break;
      }
    }
    a = b;
    ++chunk;
  }
  lz_final(st, *lz, niter);
  lz_ritz<T>(st, *lz, theta, V, n, niter);
  lz_scale<T>(st, *lz, theta, theta, n);
  DMRG_HIP(hipMemcpyAsync(h_scal, lz->scal, 8 * sizeof(double), hipMemcpyDeviceToHost, st));
  DMRG_HIP(hipStreamSynchronize(st));
  double ms = 0.0;
  for (int j = 0; timed && j < nmv; ++j) {
    float t = 0.f;
    DMRG_HIP(hipEventElapsedTime(&t, ev[2 * j], ev[2 * j + 1]));
    ms += t;
  }
  last_heff_ms = ms;
  last_niter = niter;
  last_matvecs = nmv;
  const double E = h_scal[3];
  if (!std::isfinite(E) || !(h_scal[0] > 0.0))
    throw Error(ST_NONFINITE, "lanczos: non-finite energy or zero Ritz vector at site " +
                                  std::to_string(site));
  return E;
}

template <class T>
void LanczosRunner<T>::clear_graphs() {
  for (auto& kv : graphs)
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
  graphs.clear();
}

// ---------------------------------------------------------------------------
// Construction (dmrg2_gpu_impl.h:51-227)
// ---------------------------------------------------------------------------
namespace {
void dmalloc(void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(ST_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
}
}  // namespace

template <class T>
Engine<T>::Engine(int L_, int d_, int chi_, int D_, double tol_, int variant_, const Options& o)
    : L(L_), d(d_), chi(chi_), D(D_), tol(tol_), variant(variant_), opt(o) {
  DMRG_REQUIRE(L >= 2, "L must be >= 2");
  DMRG_REQUIRE(d >= 1 && d <= 64, "bad physical dimension");
  DMRG_REQUIRE(chi >= 1 && D >= 1, "bad chi_max / D");
  DMRG_REQUIRE(variant == 1 || variant == 2, "variant must be 1 or 2");
  DMRG_REQUIRE(opt.krylov_max >= 1 && opt.krylov_max <= 128, "krylov_max must be in [1, 128]");
  DMRG_REQUIRE(opt.precision == 0 ||
                   ((opt.precision == 1 || opt.precision == 2) && !is_cplx<T>::value && variant == 2),
               "FP32 matvec mode is available for the real two-site engine");
  DMRG_REQUIRE((long long)chi * d <= 2048, "chi_max * d must be <= 2048");
  DMRG_HIP(hipSetDevice(opt.device));
  if (opt.autotune) gemm_autotune(true);
  bond.assign(L + 1, 1);
  for (int i = 1; i < L; ++i) {
    int e = std::min(i, L - i);
    double ex = std::pow((double)d, (double)e);
    bond[i] = ex > chi ? chi : (int)ex;
  }
  sync_shapes();
  try {
    alloc_all();
  } catch (...) {
    free_all();
    throw;
  }
}

template <class T>
Engine<T>::~Engine() {
  free_all();
}

template <class T>
void Engine<T>::alloc_all() {
  DMRG_HIP(hipStreamCreateWithFlags(&s_main, hipStreamNonBlocking));
  DMRG_HIP(hipStreamCreateWithFlags(&s_env, hipStreamNonBlocking));
  DMRG_HIP(hipEventCreateWithFlags(&ev_canon, hipEventDisableTiming));
  DMRG_HIP(hipEventCreateWithFlags(&ev_env, hipEventDisableTiming));
  const size_t site_sz = (size_t)chi * d * chi;
  const size_t env_sz = (size_t)chi * D * chi;
  const int dv = (variant == 2) ? d * d : d;
  nmax = (size_t)chi * dv * chi;
  kmax = std::min<size_t>((size_t)opt.krylov_max, nmax);
  mps.assign(L, nullptr);
  for (int i = 0; i < L; ++i) dmalloc((void**)&mps[i], site_sz * sizeof(T));
  dmalloc((void**)&mps_spare, site_sz * sizeof(T));
  Lenv.assign(L + 1, nullptr);
  Renv.assign(L + 1, nullptr);
  for (int i = 0; i <= L; ++i) {
    size_t sz = (i == 0 || i == L) ? (size_t)D : env_sz;
    dmalloc((void**)&Lenv[i], sz * sizeof(T));
    dmalloc((void**)&Renv[i], sz * sizeof(T));
  }
  Wd.assign(L, nullptr);
  Wrd.assign(L, nullptr);
  const size_t wsz = (size_t)D * d * d * D;
  for (int i = 0; i < L; ++i) {
    dmalloc((void**)&Wd[i], wsz * sizeof(T));
    dmalloc((void**)&Wrd[i], wsz * sizeof(T));
  }
  if (variant == 2) {
    const size_t K = (size_t)D * d * d;
    WWd.assign(L - 1, nullptr);
    for (int i = 0; i < L - 1; ++i) dmalloc((void**)&WWd[i], K * K * sizeof(T));
    if (opt.precision != 0) {
      WWf.assign(L - 1, nullptr);
      for (int i = 0; i < L - 1; ++i) dmalloc((void**)&WWf[i], K * K * sizeof(float));
      dmalloc((void**)&Lf, env_sz * sizeof(float));
      dmalloc((void**)&envScale, 2 * sizeof(double));
      dmalloc((void**)&Rf, env_sz * sizeof(float));
      dmalloc((void**)&thf, nmax * sizeof(float));
      dmalloc((void**)&outf, nmax * sizeof(float));
      sc_f.cap = (size_t)D * d * d * chi * chi;
      sc_f.split_bf16 = opt.precision == 2;
      dmalloc((void**)&sc_f.X, sc_f.cap * sizeof(float));
      dmalloc((void**)&sc_f.Y, sc_f.cap * sizeof(float));
      if (opt.env_precision != 0) {
        env32 = true;
        Lfe.assign(L + 1, nullptr);
        Rfe.assign(L + 1, nullptr);
        for (int i = 0; i <= L; ++i) {
          size_t sz = (i == 0 || i == L) ? (size_t)D : env_sz;
          dmalloc((void**)&Lfe[i], sz * sizeof(float));
          dmalloc((void**)&Rfe[i], sz * sizeof(float));
        }
        Wf.assign(L, nullptr);
        Wrf.assign(L, nullptr);
        const size_t wsz = (size_t)D * d * d * D;
        for (int i = 0; i < L; ++i) {
          dmalloc((void**)&Wf[i], wsz * sizeof(float));
          dmalloc((void**)&Wrf[i], wsz * sizeof(float));
        }
        dmalloc((void**)&envS, 2 * (size_t)(L + 1) * sizeof(double));
        dmalloc((void**)&envMax, sizeof(double));
        dmalloc((void**)&Af, site_sz * sizeof(float));
        sc_envf.cap = (size_t)D * d * chi * chi;
        dmalloc((void**)&sc_envf.X, sc_envf.cap * sizeof(float));
        dmalloc((void**)&sc_envf.Y, sc_envf.cap * sizeof(float));
      }
    }
  }
  dmalloc((void**)&theta, nmax * sizeof(T));
  dmalloc((void**)&work, nmax * sizeof(T));
  dmalloc((void**)&V, (size_t)kmax * nmax * sizeof(T));
  sc_main.cap = (size_t)D * dv * chi * chi;
  dmalloc((void**)&sc_main.X, sc_main.cap * sizeof(T));
  dmalloc((void**)&sc_main.Y, sc_main.cap * sizeof(T));
  sc_env.cap = (size_t)D * d * chi * chi;
  dmalloc((void**)&sc_env.X, sc_env.cap * sizeof(T));
  dmalloc((void**)&sc_env.Y, sc_env.cap * sizeof(T));
  lz.alloc(std::max(kmax, 1), lz_part_doubles(std::max(kmax, 1), nmax) * sizeof(double), sizeof(T));
  svdw.alloc(chi * d);
  if (opt.svd_rsvd) {
    DMRG_REQUIRE(opt.rsvd_oversample >= 0 && opt.rsvd_power >= 0 && opt.rsvd_power <= 8,
                 "rsvd: oversample >= 0 and power in [0, 8]");
    rsvdw.alloc(chi * d, std::min(chi * d, chi + opt.rsvd_oversample));
  }
  if (variant == 2 && opt.svd_warm) warm_alloc();
  DMRG_HIP(hipHostMalloc((void**)&h_status, 4 * sizeof(int), hipHostMallocDefault));
  DMRG_HIP(hipHostMalloc((void**)&h_scal, 8 * sizeof(double), hipHostMallocDefault));
  DMRG_HIP(hipHostMalloc((void**)&h_aux, 8 * sizeof(T), hipHostMallocDefault));
  // <psi|H|psi> contraction: two environment buffers and the identity MPO
  dmalloc((void**)&ex_a, env_sz * sizeof(T));
  dmalloc((void**)&ex_b, env_sz * sizeof(T));
  dmalloc((void**)&ex_id, (size_t)d * d * sizeof(T));
  {
    std::vector<T> idm((size_t)d * d, T(0));
    for (int s = 0; s < d; ++s) idm[s + (size_t)d * s] = T(1);
    DMRG_HIP(hipMemcpy(ex_id, idm.data(), idm.size() * sizeof(T), hipMemcpyHostToDevice));
  }
  if (opt.profile) {
    // start/end pair per environment update of one sweep (2 (L-1) updates)
    env_ev.resize(4 * (size_t)L);
    for (auto& e : env_ev) DMRG_HIP(hipEventCreate(&e));
  }
  runner.lz = &lz;
  runner.timed = opt.profile != 0;
  runner.V = V;
  runner.work = work;
  runner.h_status = h_status;
  runner.h_scal = h_scal;
  runner.beta_tol = opt.lanczos_tol;
  runner.eig_tol = opt.eig_tol;
  runner.eig_rel = opt.precision != 0 ? opt.fp32_eig_rel : 0.0;
}

template <class T>
void Engine<T>::borrow_streams(hipStream_t main, hipStream_t env) {
  DMRG_REQUIRE(main && env, "borrow_streams: NULL stream");
  DMRG_REQUIRE(!own_main, "borrow_streams: already borrowing");
  DMRG_HIP(hipStreamSynchronize(s_main));
  DMRG_HIP(hipStreamSynchronize(s_env));
  own_main = s_main;
  own_env = s_env;
  s_main = main;
  s_env = env;
}

template <class T>
void Engine<T>::restore_streams() {
  if (!own_main) return;
  const hipError_t e1 = hipStreamSynchronize(s_main);
  const hipError_t e2 = hipStreamSynchronize(s_env);
  s_main = own_main;
  s_env = own_env;
  own_main = own_env = nullptr;
  DMRG_HIP(e1);
  DMRG_HIP(e2);
}

template <class T>
void Engine<T>::free_all() {
  if (own_main) {  // still borrowing: drain the lent streams, destroy only our own
    (void)hipStreamSynchronize(s_main);
    (void)hipStreamSynchronize(s_env);
    s_main = own_main;
    s_env = own_env;
    own_main = own_env = nullptr;
  }
  seg_free();
  if (s_main) (void)hipStreamSynchronize(s_main);
  if (s_env) (void)hipStreamSynchronize(s_env);
  auto f = [](void* p) { if (p) (void)hipFree(p); };
  for (auto p : mps) f(p);
  f(mps_spare);
  for (auto p : Lenv) f(p);
  for (auto p : Renv) f(p);
  for (auto p : Wd) f(p);
  for (auto p : Wrd) f(p);
  for (auto p : WWd) f(p);
  for (auto p : WWf) f(p);
  f(Lf); f(Rf); f(thf); f(outf); f(envScale);
  envScale = nullptr;
  f(sc_f.X); f(sc_f.Y);
  for (auto p : Lfe) f(p);
  for (auto p : Rfe) f(p);
  for (auto p : Wf) f(p);
  for (auto p : Wrf) f(p);
  Lfe.clear(); Rfe.clear(); Wf.clear(); Wrf.clear();
  f(envS); f(envMax); f(Af); f(sc_envf.X); f(sc_envf.Y);
  envS = envMax = nullptr;
  Af = nullptr;
  sc_envf.X = sc_envf.Y = nullptr;
  env32 = false;
  f(theta); f(work); f(V);
  f(sc_main.X); f(sc_main.Y); f(sc_env.X); f(sc_env.Y);
  mps.clear(); Lenv.clear(); Renv.clear(); Wd.clear(); Wrd.clear(); WWd.clear(); WWf.clear();
  mps_spare = nullptr; Lf = Rf = thf = outf = nullptr; theta = work = V = nullptr;
  sc_f.X = sc_f.Y = nullptr; sc_main.X = sc_main.Y = sc_env.X = sc_env.Y = nullptr;
  lz.release();
  svdw.release();
  rsvdw.release();
  warm_free();
  for (auto e : runner.ev) (void)hipEventDestroy(e);
  runner.ev.clear();
  runner.clear_graphs();
  if (h_status) (void)hipHostFree(h_status);
  if (h_scal) (void)hipHostFree(h_scal);
  if (h_aux) (void)hipHostFree(h_aux);
  h_status = nullptr;
  h_scal = nullptr;
  h_aux = nullptr;
  f(ex_a); f(ex_b); f(ex_id);
  ex_a = ex_b = ex_id = nullptr;
  for (auto e : env_ev) (void)hipEventDestroy(e);
  env_ev.clear();
  if (ev_canon) (void)hipEventDestroy(ev_canon);
  if (ev_env) (void)hipEventDestroy(ev_env);
  if (s_main) (void)hipStreamDestroy(s_main);
  if (s_env) (void)hipStreamDestroy(s_env);
  ev_canon = ev_env = nullptr;
  s_main = s_env = nullptr;
}

// ---------------------------------------------------------------------------
// MPO / MPS setup (dmrg2_gpu_impl.h:382-576)
// ---------------------------------------------------------------------------
template <class T>
void Engine<T>::set_mpo(const T* const* W) {
  DMRG_REQUIRE(W != nullptr, "set_mpo: null");
  mpo_set = false;  // until every site is in: a rejected call leaves no half-set MPO usable
  const size_t wsz = (size_t)D * d * d * D;
  std::vector<T> wr(wsz);
  for (int i = 0; i < L; ++i) {
    DMRG_REQUIRE(W[i] != nullptr, "set_mpo: null site tensor");
    for (size_t k = 0; k < wsz; ++k)
      if (!finite_v(W[i][k])) throw Error(ST_NONFINITE, "set_mpo: non-finite MPO entry");
    DMRG_HIP(hipMemcpy(Wd[i], W[i], wsz * sizeof(T), hipMemcpyHostToDevice));
    build_wright<T>(W[i], d, D, wr.data());
    DMRG_HIP(hipMemcpy(Wrd[i], wr.data(), wsz * sizeof(T), hipMemcpyHostToDevice));
    if (env32) {
      std::vector<float> wf(wsz);
      for (size_t k = 0; k < wsz; ++k) wf[k] = (float)real(W[i][k]);
      DMRG_HIP(hipMemcpy(Wf[i], wf.data(), wsz * sizeof(float), hipMemcpyHostToDevice));
      for (size_t k = 0; k < wsz; ++k) wf[k] = (float)real(wr[k]);
      DMRG_HIP(hipMemcpy(Wrf[i], wf.data(), wsz * sizeof(float), hipMemcpyHostToDevice));
    }
  }
  if (variant == 2) {
    const size_t K = (size_t)D * d * d;
    std::vector<T> ww(K * K);
    std::vector<float> wwf(K * K);
    for (int i = 0; i < L - 1; ++i) {
      build_ww2<T>(W[i], W[i + 1], d, D, ww.data());
      DMRG_HIP(hipMemcpy(WWd[i], ww.data(), K * K * sizeof(T), hipMemcpyHostToDevice));
      if (opt.precision != 0) {
        for (size_t k = 0; k < K * K; ++k) wwf[k] = (float)real(ww[k]);
        DMRG_HIP(hipMemcpy(WWf[i], wwf.data(), K * K * sizeof(float), hipMemcpyHostToDevice));
      }
    }
  }
  mpo_set = true;
}

template <class T>
void Engine<T>::init_random(uint64_t seed, double scale) {
  uint64_t off = 0;
  for (int i = 0; i < L; ++i) {
    size_t n = (size_t)bond[i] * d * bond[i + 1];
    fill_random<T>(s_main, mps[i], n, seed, off, scale);
    off += n;
  }
  DMRG_HIP(hipStreamSynchronize(s_main));
  sync_shapes();
}

template <class T>
void Engine<T>::init_product() {
  for (int i = 1; i < L; ++i) bond[i] = 1;
  std::vector<T> h(d, T(0));
  h[0] = T(1);
  for (int i = 0; i < L; ++i) DMRG_HIP(hipMemcpy(mps[i], h.data(), d * sizeof(T), hipMemcpyHostToDevice));
  sync_shapes();
}

template <class T>
void Engine<T>::init_neel() {
  for (int i = 0; i < L; ++i) {
    int cL = bond[i], cR = bond[i + 1];
    std::vector<T> h((size_t)cL * d * cR, T(0));
    int s = i % 2;
    for (int a = 0; a < std::min(cL, cR); ++a) h[a + (size_t)s * cL + (size_t)a * cL * d] = T(1);
    DMRG_HIP(hipMemcpy(mps[i], h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  }
  sync_shapes();
}

template <class T>
void Engine<T>::sync_shapes() {
  tl_.assign(bond.begin(), bond.end() - 1);
  tr_.assign(bond.begin() + 1, bond.end());
}

template <class T>
void Engine<T>::check_shapes(const char* who) const {
  for (int i = 0; i < L; ++i)
    if (tl_[i] != bond[i] || tr_[i] != bond[i + 1])
      throw Error(ST_STATE, std::string(who) + ": site " + std::to_string(i) + " holds a " +
                                std::to_string(tl_[i]) + " x d x " + std::to_string(tr_[i]) +
                                " tensor but its bonds are " + std::to_string(bond[i]) + " and " +
                                std::to_string(bond[i + 1]) +
                                " (set_mps shapes disagree between neighbouring sites)");
}

template <class T>
void Engine<T>::set_mps(int site, const T* A, int cL, int cR) {
  DMRG_REQUIRE(site >= 0 && site < L, "set_mps: bad site");
  DMRG_REQUIRE(cL >= 1 && cR >= 1 && cL <= chi && cR <= chi, "set_mps: bond dims out of range");
  DMRG_REQUIRE(site != 0 || cL == 1, "set_mps: first site needs cL = 1");
  DMRG_REQUIRE(site != L - 1 || cR == 1, "set_mps: last site needs cR = 1");
  size_t n = (size_t)cL * d * cR;
  for (size_t k = 0; k < n; ++k)
    if (!finite_v(A[k])) throw Error(ST_NONFINITE, "set_mps: non-finite entry");
  bond[site] = cL;
  bond[site + 1] = cR;
  tl_[site] = cL;
  tr_[site] = cR;
  DMRG_HIP(hipMemcpy(mps[site], A, n * sizeof(T), hipMemcpyHostToDevice));
}

template <class T>
void Engine<T>::get_mps(int site, T* out) const {
  DMRG_REQUIRE(site >= 0 && site < L, "get_mps: bad site");
  size_t n = (size_t)bond[site] * d * bond[site + 1];
  DMRG_HIP(hipStreamSynchronize(s_main));
  DMRG_HIP(hipStreamSynchronize(s_env));
  DMRG_HIP(hipMemcpy(out, mps[site], n * sizeof(T), hipMemcpyDeviceToHost));
}

// ---------------------------------------------------------------------------
// Environments (dmrg2_gpu_impl.h:786-964)
// ---------------------------------------------------------------------------
template <class T>
void Engine<T>::env_left_at(int site) {
  const bool timed = env_timing && env_ev_n + 2 <= env_ev.size();
  if (timed) DMRG_HIP(hipEventRecord(env_ev[env_ev_n], s_env));
  env_left<T>(s_env, Lenv[site], mps[site], Wd[site], Lenv[site + 1], bond[site], bond[site + 1], d,
              D, sc_env);
  if (timed) {
    DMRG_HIP(hipEventRecord(env_ev[env_ev_n + 1], s_env));
    env_ev_n += 2;
  }
  stats.env_flop += flops_env(bond[site], bond[site + 1], d, D, is_cplx<T>::value);
}

template <class T>
void Engine<T>::env_right_at(int site) {
  const bool timed = env_timing && env_ev_n + 2 <= env_ev.size();
  if (timed) DMRG_HIP(hipEventRecord(env_ev[env_ev_n], s_env));
  env_right<T>(s_env, Renv[site + 1], mps[site], Wrd[site], Renv[site], bond[site + 1], bond[site], d,
               D, sc_env);
  if (timed) {
    DMRG_HIP(hipEventRecord(env_ev[env_ev_n + 1], s_env));
    env_ev_n += 2;
  }
  stats.env_flop += flops_env(bond[site + 1], bond[site], d, D, is_cplx<T>::value);
}

// FP32 environment updates (mixed modes): the same contractions on FP32 MFMA
// from the FP32 copy of the site tensor, then renormalised to unit max-abs
// with the factor carried into the new scale.
template <class T>
void Engine<T>::env_left_at32(int site) {
  if constexpr (!is_cplx<T>::value) {
    const bool timed = env_timing && env_ev_n + 2 <= env_ev.size();
    if (timed) DMRG_HIP(hipEventRecord(env_ev[env_ev_n], s_env));
    const int ci = bond[site], co = bond[site + 1];
    convert_f64_f32(s_env, mps[site], Af, (size_t)ci * d * co);
    env_left<float>(s_env, Lfe[site], Af, Wf[site], Lfe[site + 1], ci, co, d, D, sc_envf);
    const size_t n = (size_t)co * D * co;
    absmax_f32(s_env, Lfe[site + 1], n, envMax);
    rescale_env_f32(s_env, Lfe[site + 1], n, envMax, envS + site, envS + site + 1);
    if (timed) {
      DMRG_HIP(hipEventRecord(env_ev[env_ev_n + 1], s_env));
      env_ev_n += 2;
    }
    stats.env_flop += flops_env(ci, co, d, D, false);
  }
}

template <class T>
void Engine<T>::env_right_at32(int site) {
  if constexpr (!is_cplx<T>::value) {
    const bool timed = env_timing && env_ev_n + 2 <= env_ev.size();
    if (timed) DMRG_HIP(hipEventRecord(env_ev[env_ev_n], s_env));
    const int ci = bond[site + 1], co = bond[site];
    convert_f64_f32(s_env, mps[site], Af, (size_t)co * d * ci);
    env_right<float>(s_env, Rfe[site + 1], Af, Wrf[site], Rfe[site], ci, co, d, D, sc_envf);
    const size_t n = (size_t)co * D * co;
    double* const sR = envS + (L + 1);
    absmax_f32(s_env, Rfe[site], n, envMax);
    rescale_env_f32(s_env, Rfe[site], n, envMax, sR + site + 1, sR + site);
    if (timed) {
      DMRG_HIP(hipEventRecord(env_ev[env_ev_n + 1], s_env));
      env_ev_n += 2;
    }
    stats.env_flop += flops_env(ci, co, d, D, false);
  }
}

// Adds the device time of the environment updates recorded since the last
// call (option profile) to stats.env_ms; the streams must be synchronised.
template <class T>
void Engine<T>::flush_env_timing() {
  for (size_t j = 0; j + 1 < env_ev_n; j += 2) {
    float t = 0.f;
    DMRG_HIP(hipEventElapsedTime(&t, env_ev[j], env_ev[j + 1]));
    stats.env_ms += t;
  }
  env_ev_n = 0;
}

template <class T>
void Engine<T>::build_envs() {
  if (env32) {
    if constexpr (!is_cplx<T>::value) {
      set_boundary_env<float>(s_env, Lfe[0], D, 0);
      set_boundary_env<float>(s_env, Rfe[L], D, D - 1);
      fill_f64(s_env, envS, 2 * (size_t)(L + 1), 1.0);
      for (int i = L - 1; i >= 0; --i) env_right_at32(i);
    }
  } else {
    set_boundary_env<T>(s_env, Lenv[0], D, 0);
    set_boundary_env<T>(s_env, Renv[L], D, D - 1);
    for (int i = L - 1; i >= 0; --i) env_right_at(i);
  }
  DMRG_HIP(hipStreamSynchronize(s_env));
  env_pending = false;
}

template <class T>
void Engine<T>::wait_env() {
  if (env_pending) {
    DMRG_HIP(hipStreamWaitEvent(s_main, ev_env, 0));
    env_pending = false;
  }
}

// ---------------------------------------------------------------------------
// Two-site optimisation (dmrg2_gpu_impl.h:582-1359)
// ---------------------------------------------------------------------------
template <class T>
void Engine<T>::matvec2(int site, const T* in, T* out, bool fp64) {
  const int cL = bond[site], cR = bond[site + 2];
  if (fp64 || opt.precision == 0) {
    heff2_apply<T>(s_main, Lenv[site], Renv[site + 2], WWd[site], in, out, cL, cR, d, D, sc_main);
  } else {
    if constexpr (!is_cplx<T>::value) {
      if (env32)
        heff2_apply_f32_io64(s_main, Lfe[site], Rfe[site + 2], WWf[site], in, out, cL, cR, d, D,
                             sc_f, thf, outf, envS + site, envS + (L + 1) + site + 2);
      else
        heff2_apply_f32_io64(s_main, Lf, Rf, WWf[site], in, out, cL, cR, d, D, sc_f, thf, outf,
                             envScale, envScale + 1);
    }
  }
}

template <class T>
double Engine<T>::lanczos_two(int site, bool rq) {
  const int cL = bond[site], cR = bond[site + 2];
  const size_t n = (size_t)cL * d * d * cR;
  if (opt.precision != 0) {
    if constexpr (!is_cplx<T>::value) {
      const size_t nl = (size_t)cL * D * cL, nr = (size_t)cR * D * cR;
      if (!env32) {
        absmax_f64(s_main, Lenv[site], nl, envScale);
        absmax_f64(s_main, Renv[site + 2], nr, envScale + 1);
        convert_f64_f32_scaled(s_main, Lenv[site], Lf, nl, envScale);
        convert_f64_f32_scaled(s_main, Renv[site + 2], Rf, nr, envScale + 1);
      }
    }
  }
  double t0 = now_s();
  double E = runner.run(s_main, theta, n, opt.krylov_max, site,
                        [&](const T* in, T* out) { matvec2(site, in, out, false); },
                        graph_key(site, cL, cR));
  stats.lanczos_ms += 1e3 * (now_s() - t0);
  stats.heff_ms += runner.last_heff_ms;
  stats.matvecs += runner.last_matvecs;
  stats.lanczos_calls += 1;
  stats.heff_flop += runner.last_matvecs * flops_heff2(cL, cR, d, D, is_cplx<T>::value);
  stats.graph_launches = runner.graph_launches;
  stats.graph_captures = runner.graph_captures;
  // FP64 environments: the FP64 Rayleigh quotient at the sweep's last bond;
  // FP32 environments: run() takes the energy from an FP64 contraction
  if (rq && opt.precision != 0 && !env32) E = rayleigh2(site);
  return E;
}

// FP64 Rayleigh quotient <theta|H|theta> of the (normalised) Ritz vector.
template <class T>
double Engine<T>::rayleigh2(int site) {
  const int cL = bond[site], cR = bond[site + 2];
  const size_t n = (size_t)cL * d * d * cR;
  matvec2(site, theta, work, true);
  lz_alpha<T>(s_main, lz, theta, work, n, 0);
  lz_alpha_value(s_main, lz, n, 0);
  double* const ha = reinterpret_cast<double*>(h_aux);
  DMRG_HIP(hipMemcpyAsync(ha, lz.alpha, sizeof(double), hipMemcpyDeviceToHost, s_main));
  DMRG_HIP(hipStreamSynchronize(s_main));
  return ha[0];
}

template <class T>
void Engine<T>::split_two(int site, char dir) {
  const int cL = bond[site], cR = bond[site + 2];
  const int m = cL * d, n = d * cR;
  const int full_k = std::min(m, n), k = std::min(full_k, chi);
  double t0 = now_s();
  const int wi = dir == 'R' ? 0 : 1;
  // the reference's RSVD condition (dmrg2_gpu_impl.h:1186-1188)
  const bool rsvd = opt.svd_rsvd && full_k > k + opt.rsvd_oversample && m > 2 * k;
  SvdResult r;
  if (rsvd) {
    r = rsvd_split<T>(s_main, svdw, rsvdw, theta, m, n, m, k, opt.svd_cutoff, opt.svd_rel_cutoff,
                      dir == 'L', mps[site], m, mps[site + 1], svd_params(), opt.rsvd_oversample,
                      opt.rsvd_power, 0x5EED0000ull + rsvd_seq++);
    stats.rsvd_calls += 1;
    if (warm_on) wXN[wi][site] = 0;  // no stored basis for this bond
  } else {
    const T* warm = warm_on ? warm_basis(site, dir) : nullptr;
    T* keep = warm_on ? wX[wi][site] : nullptr;
    r = svd_split<T>(s_main, svdw, theta, m, n, m, k, opt.svd_cutoff, opt.svd_rel_cutoff,
                     dir == 'L', mps[site], m, mps[site + 1], nullptr, svd_params(), warm, keep);
    if (warm_on) {
      wXN[wi][site] = dir == 'R' ? m : n;
      wXB[wi][site] = dir == 'R' ? cL : cR;
    }
    stats.svd_warm_starts += warm ? 1 : 0;
  }
  bond[site + 1] = r.nk;
  DMRG_HIP(hipEventRecord(ev_canon, s_main));
  if (warm_on) warm_after_split(site, dir);
  stats.svd_ms += 1e3 * (now_s() - t0);
  stats.svd_calls += 1;
  stats.svd_sweeps += r.sweeps;
  stats.svd_bip_sweeps += r.bip_sweeps;
  stats.svd_oa_steps += r.oa_steps;
  stats.max_discarded = std::max(stats.max_discarded, r.discarded);
  stats.trunc_sum_cur += r.discarded;
}

// ---------------------------------------------------------------------------
// Warm start of the truncation SVD (see warm.hpp)
// ---------------------------------------------------------------------------
template <class T>
void Engine<T>::warm_alloc() {
  const size_t capN = (size_t)chi * d;
  const size_t asz = (size_t)chi * d * chi;
  for (int w = 0; w < 2; ++w) {
    wX[w].assign(L - 1, nullptr);
    wXN[w].assign(L - 1, 0);
    wXB[w].assign(L - 1, 0);
    for (int i = 0; i < L - 1; ++i) dmalloc((void**)&wX[w][i], capN * capN * sizeof(T));
    wA[w].assign(L, nullptr);
    wAL[w].assign(L, 0);
    wAR[w].assign(L, 0);
    for (int i = 0; i < L; ++i) dmalloc((void**)&wA[w][i], asz * sizeof(T));
  }
  const size_t tsz = (size_t)chi * chi;
  for (T** p : {&wT[0], &wT[1], &wZ0, &wZ1, &wP, &wTmp}) dmalloc((void**)p, tsz * sizeof(T));
  dmalloc((void**)&wXw, capN * capN * sizeof(T));
  dmalloc((void**)&wDev, (size_t)kPolarMaxIter * kPolarParts * sizeof(double));
  DMRG_HIP(hipHostMalloc((void**)&wHost, (size_t)kPolarMaxIter * kPolarParts * sizeof(double),
                         hipHostMallocDefault));
  warm_on = true;
}

template <class T>
void Engine<T>::warm_free() {
  auto f = [](void* p) { if (p) (void)hipFree(p); };
  for (int w = 0; w < 2; ++w) {
    for (auto p : wX[w]) f(p);
    for (auto p : wA[w]) f(p);
    wX[w].clear();
    wA[w].clear();
  }
  for (T** p : {&wT[0], &wT[1], &wZ0, &wZ1, &wP, &wTmp, &wXw}) {
    f(*p);
    *p = nullptr;
  }
  f(wDev);
  wDev = nullptr;
  if (wHost) (void)hipHostFree(wHost);
  wHost = nullptr;
  warm_on = false;
}

// Start of a half-sweep: the empty block overlaps itself, T = [1].
template <class T>
void Engine<T>::warm_begin() {
  if (!warm_on) return;
  set_boundary_env<T>(s_main, wT[wCur], 1, 0);  // the 1 x 1 matrix [1], written on the device
  wTr = wTc = 1;
}

template <class T>
const T* Engine<T>::warm_basis(int site, char dir) {
  const int wi = dir == 'R' ? 0 : 1;
  const int cb = dir == 'R' ? bond[site] : bond[site + 2];
  const int N = cb * d;
  if (wXN[wi][site] != N || wXB[wi][site] != cb || wTr != cb || wTc != cb) return nullptr;
  T* Z = nullptr;
  // any T with singular values in (0, 1] has a polar factor; the basis it
  // gives is exact, and close to the new eigenbasis where the block changed
  // little, so accept every convergent start
  if (!polar_unitary<T>(s_main, wT[wCur], cb, wZ0, wZ1, wP, wDev, wHost, 1.01, &Z)) return nullptr;
  if (dir == 'R') warm_basis_left<T>(s_main, Z, cb, d, wX[wi][site], wXw);
  else warm_basis_right<T>(s_main, Z, cb, d, wX[wi][site], wXw);
  return wXw;
}

// After a split: extend the block overlap by the new canonical tensor and
// keep that tensor for the next sweep.
template <class T>
void Engine<T>::warm_after_split(int site, char dir) {
  if (dir == 'R') {
    const int cL = bond[site], nk = bond[site + 1];
    const T* Anew = mps[site];
    if (wTr > 0 && wAL[0][site] == wTr && wTc == cL) {
      overlap_left_step<T>(s_main, wT[wCur], wTr, wTc, wA[0][site], wAR[0][site], Anew, nk, d,
                           wT[1 - wCur], wTmp);
      wCur = 1 - wCur;
      wTr = wAR[0][site];
      wTc = nk;
    } else {
      wTr = wTc = 0;
    }
    copy_dev<T>(s_main, wA[0][site], Anew, (size_t)cL * d * nk);
    wAL[0][site] = cL;
    wAR[0][site] = nk;
  } else {
    const int nk = bond[site + 1], cR = bond[site + 2];
    const T* Bnew = mps[site + 1];
    if (wTr > 0 && wAR[1][site + 1] == wTr && wTc == cR) {
      overlap_right_step<T>(s_main, wT[wCur], wTr, wTc, wA[1][site + 1], wAL[1][site + 1], Bnew,
                            nk, d, wT[1 - wCur], wTmp);
      wCur = 1 - wCur;
      wTr = wAL[1][site + 1];
      wTc = nk;
    } else {
      wTr = wTc = 0;
    }
    copy_dev<T>(s_main, wA[1][site + 1], Bnew, (size_t)nk * d * cR);
    wAL[1][site + 1] = nk;
    wAR[1][site + 1] = cR;
  }
}

template <class T>
double Engine<T>::optimize_bond(int site, char dir, bool rq) {
  wait_env();
  const int cL = bond[site], cm = bond[site + 1], cR = bond[site + 2];
  gemm<T>(s_main, OP_N, OP_N, cL * d, d * cR, cm, T(1), mps[site], cL * d, mps[site + 1], cm, T(0),
          theta, cL * d, gemm_pick_tile(cL * d, d * cR, cm, is_cplx<T>::value));
  double E = lanczos_two(site, rq);
  split_two(site, dir);
  return E;
}

template <class T>
double Engine<T>::sweep2_lr() {
  double E = 0.0;
  warm_begin();
  for (int site = 0; site < L - 1; ++site) {
    E = optimize_bond(site, 'R', false);
    DMRG_HIP(hipStreamWaitEvent(s_env, ev_canon, 0));
    if (env32) env_left_at32(site);
    else env_left_at(site);
    DMRG_HIP(hipEventRecord(ev_env, s_env));
    env_pending = true;
  }
  wait_env();
  return E;
}

template <class T>
double Engine<T>::sweep2_rl() {
  double E = 0.0;
  warm_begin();
  for (int site = L - 2; site >= 0; --site) {
    E = optimize_bond(site, 'L', site == 0);
    DMRG_HIP(hipStreamWaitEvent(s_env, ev_canon, 0));
    if (env32) env_right_at32(site + 1);
    else env_right_at(site + 1);
    DMRG_HIP(hipEventRecord(ev_env, s_env));
    env_pending = true;
  }
  wait_env();
  return E;
}

// ---------------------------------------------------------------------------
// Single-site optimisation (dmrg_gpu_impl.h:895-1445)
// ---------------------------------------------------------------------------
template <class T>
double Engine<T>::lanczos_one(int site) {
  const int cL = bond[site], cR = bond[site + 1];
  const size_t n = (size_t)cL * d * cR;
  double t0 = now_s();
  double E = runner.run(s_main, theta, n, opt.krylov_max, site, [&](const T* in, T* out) {
    heff1_apply<T>(s_main, Lenv[site], Renv[site + 1], Wd[site], in, out, cL, cR, d, D, sc_main);
  }, graph_key(site, cL, cR));
  stats.lanczos_ms += 1e3 * (now_s() - t0);
  stats.heff_ms += runner.last_heff_ms;
  stats.matvecs += runner.last_matvecs;
  stats.lanczos_calls += 1;
  stats.heff_flop += runner.last_matvecs * flops_heff1(cL, cR, d, D, is_cplx<T>::value);
  stats.graph_launches = runner.graph_launches;
  stats.graph_captures = runner.graph_captures;
  return E;
}

template <class T>
void Engine<T>::split_one(int site, char dir) {
  const int cL = bond[site], cR = bond[site + 1];
  double t0 = now_s();
  if (dir == 'R') {
    const int m = cL * d, n = cR;
    const int k = std::min(std::min(m, n), chi);
    // A[site] = U; S*Vh (nk x cR) -> work; A[site+1] <- (S Vh) A[site+1]
    SvdResult r = svd_split<T>(s_main, svdw, theta, m, n, m, k, opt.svd_cutoff,
                               opt.svd_rel_cutoff, false, mps[site], m, work, nullptr, svd_params());
    const int nk = r.nk;
    stats.max_discarded = std::max(stats.max_discarded, r.discarded);
    stats.trunc_sum_cur += r.discarded;
    DMRG_HIP(hipEventRecord(ev_canon, s_main));
    const int cR2 = bond[site + 2];
    gemm<T>(s_main, OP_N, OP_N, nk, d * cR2, cR, T(1), work, nk, mps[site + 1], cR, T(0), mps_spare,
            nk, gemm_pick_tile(nk, d * cR2, cR, is_cplx<T>::value));
    std::swap(mps[site + 1], mps_spare);
    bond[site + 1] = nk;
    stats.svd_sweeps += r.sweeps;
  } else {
    const int m = cL, n = d * cR;
    const int k = std::min(std::min(m, n), chi);
    // A[site] = Vh; U*S (cL x nk) -> work; A[site-1] <- A[site-1] (U S)
    SvdResult r = svd_split<T>(s_main, svdw, theta, m, n, m, k, opt.svd_cutoff,
                               opt.svd_rel_cutoff, true, work, m, mps[site], nullptr, svd_params());
    const int nk = r.nk;
    stats.max_discarded = std::max(stats.max_discarded, r.discarded);
    stats.trunc_sum_cur += r.discarded;
    DMRG_HIP(hipEventRecord(ev_canon, s_main));
    const int cL0 = bond[site - 1];
    gemm<T>(s_main, OP_N, OP_N, cL0 * d, nk, cL, T(1), mps[site - 1], cL0 * d, work, cL, T(0),
            mps_spare, cL0 * d, gemm_pick_tile(cL0 * d, nk, cL, is_cplx<T>::value));
    std::swap(mps[site - 1], mps_spare);
    bond[site] = nk;
    stats.svd_sweeps += r.sweeps;
  }
  stats.svd_ms += 1e3 * (now_s() - t0);
  stats.svd_calls += 1;
}

template <class T>
double Engine<T>::optimize_site(int site, char dir, bool split) {
  wait_env();
  const size_t n = (size_t)bond[site] * d * bond[site + 1];
  copy_dev<T>(s_main, theta, mps[site], n);
  double E = lanczos_one(site);
  if (split) {
    split_one(site, dir);
  } else {
    copy_dev<T>(s_main, mps[site], theta, n);
  }
  return E;
}

template <class T>
double Engine<T>::sweep1_lr() {
  for (int site = 0; site < L - 1; ++site) {
    optimize_site(site, 'R', true);
    DMRG_HIP(hipStreamWaitEvent(s_env, ev_canon, 0));
    env_left_at(site);
    DMRG_HIP(hipEventRecord(ev_env, s_env));
    env_pending = true;
  }
  return optimize_site(L - 1, 'R', false);
}

template <class T>
double Engine<T>::sweep1_rl() {
  for (int site = L - 1; site >= 1; --site) {
    optimize_site(site, 'L', true);
    DMRG_HIP(hipStreamWaitEvent(s_env, ev_canon, 0));
    env_right_at(site);
    DMRG_HIP(hipEventRecord(ev_env, s_env));
    env_pending = true;
  }
  return optimize_site(0, 'L', false);
}

// ---------------------------------------------------------------------------
// run (dmrg2_gpu_impl.h:1426-1457)
// ---------------------------------------------------------------------------
template <class T>
double Engine<T>::run(int n_sweeps, bool quiet) {
  if (!mpo_set) throw Error(ST_STATE, "run: set_mpo has not been called");
  DMRG_REQUIRE(n_sweeps >= 0, "run: negative sweep count");
  check_shapes("run");
  // the sweeps keep every tensor consistent with bond[], however they end
  auto resync = [this] { sync_shapes(); };
  OnExit<decltype(resync)> resync_at_exit{resync};
  DMRG_HIP(hipSetDevice(opt.device));  // the calling thread may not be the creating one
  double te = now_s();
  build_envs();
  double t_start = now_s();
  double env_time = t_start - te;
  stats.env_build_s = env_time;
  if (!quiet) printf("  Environment build: %.3f s\n", env_time);
  double e_prev = 0.0;
  env_timing = !env_ev.empty();  // sweep updates only; the build above is env_build_s
  env_ev_n = 0;
  struct TimingOff {
    Engine* e;
    ~TimingOff() { e->env_timing = false; }
  } timing_off{this};
  for (int sweep = 0; sweep < n_sweeps; ++sweep) {
    double ts = now_s();
    stats.trunc_sum_cur = 0.0;
    if (variant == 2) {
      sweep2_lr();
      energy = sweep2_rl();
    } else {
      sweep1_lr();
      energy = sweep1_rl();
    }
    if (env32) {
      // FP32 environments: the sweep energy is the FP64 <psi|H|psi>/<psi|psi>
      // of the state (the last bond of the R->L half-sweep keeps every
      // singular value, so this is the Rayleigh quotient of its Ritz vector)
      double e = 0.0, n2 = 0.0;
      expectation(&e, &n2);
      energy = e;
    }
    DMRG_HIP(hipStreamSynchronize(s_main));
    DMRG_HIP(hipStreamSynchronize(s_env));
    stats.sweep_s_last = now_s() - ts;
    stats.trunc_sum_last = stats.trunc_sum_cur;
    flush_env_timing();
    stats.sweeps_s_total += stats.sweep_s_last;
    stats.sweeps_done += 1;
    sweep_E.push_back(energy);
    if (!std::isfinite(energy)) throw Error(ST_NONFINITE, "run: non-finite sweep energy");
    double dE = std::fabs(energy - e_prev);
    if (dE < tol && sweep > 0) break;
    e_prev = energy;
  }
  double total = now_s() - t_start;
  if (!quiet) {
    printf("Final energy: %.12f\n", energy);
    printf("Total wall time: %.3f s\n", total);
    printf("  env_build_sec: %.3f  timer_scope: sweep_only\n", env_time);
    fflush(stdout);
  }
  return energy;
}

// ---------------------------------------------------------------------------
// <psi|H|psi> / <psi|psi> by full left contraction (FP64).
// ---------------------------------------------------------------------------
template <class T>
void Engine<T>::expectation(double* e, double* nrm2) {
  if (!mpo_set) throw Error(ST_STATE, "expectation: set_mpo has not been called");
  check_shapes("expectation");
  DMRG_HIP(hipSetDevice(opt.device));
  DMRG_HIP(hipStreamSynchronize(s_env));
  // the environments hold sweep state: contract in the two buffers set aside
  // at creation (ex_a, ex_b) with the identity MPO ex_id for the norm
  T* ea = ex_a;
  T* eb = ex_b;
  const T* wid = ex_id;
  T* hv = h_aux;
  for (int pass = 0; pass < 2; ++pass) {
    const int Dp = pass == 0 ? D : 1;
    set_boundary_env<T>(s_main, ea, Dp, 0);
    for (int i = 0; i < L; ++i) {
      env_left<T>(s_main, ea, mps[i], pass == 0 ? Wd[i] : wid, eb, bond[i], bond[i + 1], d, Dp,
                  sc_main);
      std::swap(ea, eb);
    }
    DMRG_HIP(hipMemcpyAsync(hv + pass, ea + (Dp - 1), sizeof(T), hipMemcpyDeviceToHost, s_main));
    DMRG_HIP(hipStreamSynchronize(s_main));
  }
  *nrm2 = real(hv[1]);
  *e = real(hv[0]) / real(hv[1]);
}

template <class T>
int Engine<T>::get_env(bool left, int idx, T* out) {
  DMRG_REQUIRE(idx >= 0 && idx <= L, "get_env: bad index");
  const int c = bond[idx];
  if (out) {
    DMRG_HIP(hipStreamSynchronize(s_main));
    DMRG_HIP(hipStreamSynchronize(s_env));
    const size_t n = (size_t)c * D * c;
    if (env32) {
      // FP32 environment times its scale
      std::vector<float> hf(n);
      double sc = 1.0;
      DMRG_HIP(hipMemcpy(hf.data(), left ? Lfe[idx] : Rfe[idx], n * sizeof(float),
                         hipMemcpyDeviceToHost));
      DMRG_HIP(hipMemcpy(&sc, envS + (left ? 0 : L + 1) + idx, sizeof(double),
                         hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; ++i) out[i] = T((double)hf[i] * sc);
    } else {
      DMRG_HIP(hipMemcpy(out, left ? Lenv[idx] : Renv[idx], n * sizeof(T), hipMemcpyDeviceToHost));
    }
  }
  return c;
}

template class Engine<double>;
template class Engine<cdouble>;
template struct LanczosRunner<double>;
template struct LanczosRunner<cdouble>;

}  // namespace dmrg
