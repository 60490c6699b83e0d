// extern "C" boundary (include/dmrg_amd.h).  No exception crosses it: every
// entry point returns a status code and records the message per thread.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/dmrg_amd.h"
#include "engine.hpp"

using namespace dmrg;

namespace {
thread_local std::string g_err;

int fail(int code, const std::string& m) {
  g_err = m;
  return code;
}

template <class F>
int guard(F&& f) {
  try {
    g_err.clear();
    f();
    return DMRG_OK;
  } catch (const Error& e) {
    return fail(e.code, e.what());
  } catch (const std::bad_alloc&) {
    return fail(DMRG_ERR_OOM, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(DMRG_ERR_HIP, e.what());
  } catch (...) {
    return fail(DMRG_ERR_HIP, "unknown error");
  }
}

Options to_opts(const dmrg_options* o) {
  Options r;
  if (o) {
    r.krylov_max = o->krylov_max;
    r.lanczos_tol = o->lanczos_tol;
    r.eig_tol = o->eig_tol;
    r.svd_cutoff = o->svd_cutoff;
    r.svd_rel_cutoff = o->svd_rel_cutoff;
    r.precision = o->precision;
    r.profile = o->profile;
    r.device = o->device;
    r.svd_max_sweeps = o->svd_max_sweeps;
    r.svd_tol = o->svd_tol;
    r.svd_warm = o->svd_warm;
    r.fp32_eig_rel = o->fp32_eig_rel;
    r.svd_subspace = o->svd_subspace;
    r.svd_tol_mixed = o->svd_tol_mixed;
    r.env_precision = o->env_precision;
    r.svd_rsvd = o->svd_rsvd;
    r.rsvd_oversample = o->rsvd_oversample;
    r.rsvd_power = o->rsvd_power;
    r.graph = o->graph;
    r.autotune = o->autotune;
  }
  return r;
}
}  // namespace

struct dmrg_engine {
  int dtype = 0;
  std::unique_ptr<Engine<double>> r;
  std::unique_ptr<Engine<cdouble>> c;
  template <class F>
  void with(F&& f) {
    if (r) f(*r);
    else if (c) f(*c);
    else throw Error(ST_STATE, "engine not initialised");
  }
};

extern "C" {

void dmrg_default_options(dmrg_options* o) {
  if (!o) return;
  Options d;
  o->krylov_max = d.krylov_max;
  o->lanczos_tol = d.lanczos_tol;
  o->eig_tol = d.eig_tol;
  o->svd_cutoff = d.svd_cutoff;
  o->svd_rel_cutoff = d.svd_rel_cutoff;
  o->precision = d.precision;
  o->profile = d.profile;
  o->device = d.device;
  o->svd_max_sweeps = d.svd_max_sweeps;
  o->svd_tol = d.svd_tol;
  o->svd_warm = d.svd_warm;
  o->fp32_eig_rel = d.fp32_eig_rel;
  o->svd_subspace = d.svd_subspace;
  o->svd_tol_mixed = d.svd_tol_mixed;
  o->env_precision = d.env_precision;
  o->svd_rsvd = d.svd_rsvd;
  o->rsvd_oversample = d.rsvd_oversample;
  o->rsvd_power = d.rsvd_power;
  o->graph = d.graph;
  o->autotune = d.autotune;
}

const char* dmrg_last_error(void) { return g_err.c_str(); }

int dmrg_create(int L, int d, int chi_max, int D, double tol, int dtype, int variant,
                const dmrg_options* opt, dmrg_handle* out) {
  if (!out) return fail(DMRG_ERR_BAD_ARG, "dmrg_create: out is NULL");
  *out = nullptr;
  return guard([&] {
    DMRG_REQUIRE(dtype == DMRG_DTYPE_F64 || dtype == DMRG_DTYPE_C128, "dmrg_create: bad dtype");
    Options o = opt ? to_opts(opt) : Options();
    auto h = std::make_unique<dmrg_engine>();
    h->dtype = dtype;
    if (dtype == DMRG_DTYPE_F64) h->r = std::make_unique<Engine<double>>(L, d, chi_max, D, tol, variant, o);
    else h->c = std::make_unique<Engine<cdouble>>(L, d, chi_max, D, tol, variant, o);
    *out = h.release();
  });
}

void dmrg_destroy(dmrg_handle h) { delete h; }

int dmrg_run_batch(dmrg_handle* hs, int n, int n_sweeps, int threads, double* energies,
                   double* seconds) {
  if (!hs || !energies || n < 0) return fail(DMRG_ERR_BAD_ARG, "dmrg_run_batch: bad argument");
  for (int i = 0; i < n; ++i)
    if (!hs[i]) return fail(DMRG_ERR_BAD_ARG, "dmrg_run_batch: NULL engine");
  return guard([&] {
    {
      // one engine listed twice would run on two threads at once
      std::vector<dmrg_handle> sorted(hs, hs + n);
      std::sort(sorted.begin(), sorted.end());
      DMRG_REQUIRE(std::adjacent_find(sorted.begin(), sorted.end()) == sorted.end(),
                   "dmrg_run_batch: the same engine is listed twice");
    }
    // longest first: estimated cost L * chi^3 * d^3 per sweep
    std::vector<int> order(n);
    std::vector<double> cost(n);
    for (int i = 0; i < n; ++i) {
      order[i] = i;
      hs[i]->with([&](auto& e) {
        cost[i] = (double)e.L * e.chi * (double)e.chi * e.chi * e.d * e.d * e.d;
      });
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::atomic<int> next{0};
    std::atomic<bool> stop{false};
    std::mutex mu;
    int err_code = 0;
    std::string err_msg;
    const int nt = std::max(1, std::min(threads, n));
    // One stream per worker, created here in worker order, carries both the
    // bond work and the environment updates of the chains that worker runs:
    // the runtime spreads streams over its hardware queues in creation order,
    // so the workers share the queues evenly, and a chain never waits behind
    // another chain's kernels on a queue that it shares with it more than
    // the worker count forces.
    struct Streams {
      std::vector<hipStream_t> s;
      ~Streams() {
        for (hipStream_t x : s) (void)hipStreamDestroy(x);
      }
    } ws;
    for (int t = 0; t < nt; ++t) {
      hipStream_t x = nullptr;
      DMRG_HIP(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
      ws.s.push_back(x);
    }
    auto worker = [&](int t) {
      hipStream_t st = ws.s[t];
      for (;;) {
        if (stop.load()) return;
        const int j = next.fetch_add(1);
        if (j >= n) return;
        const int i = order[j];
        try {
          const auto t0 = std::chrono::steady_clock::now();
          double E = 0.0;
          hs[i]->with([&](auto& e) {
            struct Lend {
              decltype(e)& eng;
              ~Lend() {
                try { eng.restore_streams(); } catch (...) {}
              }
            };
            e.borrow_streams(st, st);
            Lend lend{e};
            E = e.run(n_sweeps, true);
            e.restore_streams();
          });
          energies[i] = E;
          if (seconds)
            seconds[i] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        } catch (const Error& e) {
          std::lock_guard<std::mutex> lk(mu);
          if (!err_code) { err_code = e.code; err_msg = "chain " + std::to_string(i) + ": " + e.what(); }
          stop = true;
        } catch (const std::exception& e) {
          std::lock_guard<std::mutex> lk(mu);
          if (!err_code) { err_code = ST_HIP; err_msg = "chain " + std::to_string(i) + ": " + e.what(); }
          stop = true;
        } catch (...) {
          // an exception escaping a worker thread would end the process
          std::lock_guard<std::mutex> lk(mu);
          if (!err_code) { err_code = ST_HIP; err_msg = "chain " + std::to_string(i) + ": unknown error"; }
          stop = true;
        }
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(worker, t);
    worker(0);
    for (auto& th : pool) th.join();
    if (err_code) throw Error(err_code, err_msg);
  });
}

int dmrg_set_mpo(dmrg_handle h, const void* const* W) {
  if (!h || !W) return fail(DMRG_ERR_BAD_ARG, "dmrg_set_mpo: NULL argument");
  return guard([&] {
    if (h->r) h->r->set_mpo(reinterpret_cast<const double* const*>(W));
    else h->c->set_mpo(reinterpret_cast<const cdouble* const*>(W));
  });
}

int dmrg_init_mps_random(dmrg_handle h, uint64_t seed, double scale) {
  if (!h) return fail(DMRG_ERR_BAD_ARG, "NULL handle");
  return guard([&] { h->with([&](auto& e) { e.init_random(seed, scale); }); });
}
int dmrg_init_mps_product(dmrg_handle h) {
  if (!h) return fail(DMRG_ERR_BAD_ARG, "NULL handle");
  return guard([&] { h->with([&](auto& e) { e.init_product(); }); });
}
int dmrg_init_mps_neel(dmrg_handle h) {
  if (!h) return fail(DMRG_ERR_BAD_ARG, "NULL handle");
  return guard([&] { h->with([&](auto& e) { e.init_neel(); }); });
}
int dmrg_set_mps(dmrg_handle h, int site, const void* A, int cL, int cR) {
  if (!h || !A) return fail(DMRG_ERR_BAD_ARG, "dmrg_set_mps: NULL argument");
  return guard([&] {
    if (h->r) h->r->set_mps(site, (const double*)A, cL, cR);
    else h->c->set_mps(site, (const cdouble*)A, cL, cR);
  });
}

int dmrg_run(dmrg_handle h, int n_sweeps, int quiet, double* energy) {
  if (!h) return fail(DMRG_ERR_BAD_ARG, "NULL handle");
  return guard([&] {
    double e = 0.0;
    h->with([&](auto& eng) { e = eng.run(n_sweeps, quiet != 0); });
    if (energy) *energy = e;
  });
}

int dmrg_run_pdmrg(dmrg_handle h, int segments, int outer, int local, int warmup, int polish,
                   int quiet, double* energy) {
  if (!h) return fail(DMRG_ERR_BAD_ARG, "NULL handle");
  return guard([&] {
    PdmrgParams p;
    p.segments = segments;
    p.outer = outer;
    p.local = local;
    p.warmup = warmup;
    p.polish = polish;
    double e = 0.0;
    h->with([&](auto& eng) { e = eng.run_pdmrg(p, quiet != 0); });
    if (energy) *energy = e;
  });
}

int dmrg_get_energy(dmrg_handle h, double* energy) {
  if (!h || !energy) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] { h->with([&](auto& e) { *energy = e.energy; }); });
}

int dmrg_get_sweep_energies(dmrg_handle h, double* out, int cap, int* n) {
  if (!h || !n || (cap > 0 && !out)) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] {
    h->with([&](auto& e) {
      *n = (int)e.sweep_E.size();
      for (int i = 0; i < std::min(cap, *n); ++i) out[i] = e.sweep_E[i];
    });
  });
}

int dmrg_get_bond_dims(dmrg_handle h, int* dims) {
  if (!h || !dims) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] { h->with([&](auto& e) { for (int i = 0; i <= e.L; ++i) dims[i] = e.bond[i]; }); });
}

int dmrg_get_mps(dmrg_handle h, int site, void* out) {
  if (!h || !out) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] {
    if (h->r) h->r->get_mps(site, (double*)out);
    else h->c->get_mps(site, (cdouble*)out);
  });
}

int dmrg_get_env(dmrg_handle h, int left, int idx, void* out, int* chi) {
  if (!h || !chi) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] {
    if (h->r) *chi = h->r->get_env(left != 0, idx, (double*)out);
    else *chi = h->c->get_env(left != 0, idx, (cdouble*)out);
  });
}

int dmrg_expectation(dmrg_handle h, double* energy, double* norm2) {
  if (!h || !energy || !norm2) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] { h->with([&](auto& e) { e.expectation(energy, norm2); }); });
}

int dmrg_get_stats(dmrg_handle h, dmrg_stats* st) {
  if (!h || !st) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] {
    h->with([&](auto& e) {
      const Stats& s = e.stats;
      st->lanczos_ms = s.lanczos_ms;
      st->heff_ms = s.heff_ms;
      st->svd_ms = s.svd_ms;
      st->env_ms = s.env_ms;
      st->sweep_s_last = s.sweep_s_last;
      st->heff_flop = s.heff_flop;
      st->env_flop = s.env_flop;
      st->matvecs = s.matvecs;
      st->lanczos_calls = s.lanczos_calls;
      st->svd_calls = s.svd_calls;
      st->svd_sweeps = s.svd_sweeps;
      st->sweeps_done = s.sweeps_done;
      st->max_discarded = s.max_discarded;
      st->env_build_s = s.env_build_s;
      st->sweeps_s_total = s.sweeps_s_total;
      st->svd_warm_starts = s.svd_warm_starts;
      st->svd_bip_sweeps = s.svd_bip_sweeps;
      st->svd_oa_steps = s.svd_oa_steps;
      st->trunc_sum_last = s.trunc_sum_last;
      st->rsvd_calls = s.rsvd_calls;
      st->graph_launches = s.graph_launches;
      st->graph_captures = s.graph_captures;
      st->tuned_shapes = gemm_autotune_entries();
    });
  });
}

int dmrg_reset_stats(dmrg_handle h) {
  if (!h) return fail(DMRG_ERR_BAD_ARG, "NULL handle");
  return guard([&] { h->with([&](auto& e) { e.stats = Stats(); }); });
}

// ---------------------------------------------------------------------------
// Operator-level entry points
// ---------------------------------------------------------------------------
}  // extern "C"

namespace {
template <class T>
struct CtxT {
  ContractScratch<T> sc;
  SvdWork<T> svd;
  RsvdWork<T> rsvd;
  LzDev lz;
  T* buf = nullptr;  // derived MPO matrices
  size_t buf_cap = 0;
  T* V = nullptr;
  T* work = nullptr;
  size_t V_cap = 0, work_cap = 0;
  LanczosRunner<T> runner;
  ~CtxT() {
    if (sc.X) (void)hipFree(sc.X);
    if (sc.Y) (void)hipFree(sc.Y);
    if (buf) (void)hipFree(buf);
    if (V) (void)hipFree(V);
    if (work) (void)hipFree(work);
    for (auto e : runner.ev) (void)hipEventDestroy(e);
    svd.release();
    rsvd.release();
    lz.release();
  }
  void need_sc(size_t n) {
    if (n <= sc.cap) return;
    if (sc.X) DMRG_HIP(hipFree(sc.X));
    if (sc.Y) DMRG_HIP(hipFree(sc.Y));
    sc.X = sc.Y = nullptr;
    sc.cap = 0;  // a failed allocation below leaves no stale capacity
    DMRG_HIP(hipMalloc(&sc.X, n * sizeof(T)));
    DMRG_HIP(hipMalloc(&sc.Y, n * sizeof(T)));
    sc.cap = n;
  }
  T* need_buf(size_t n) {
    if (n > buf_cap) {
      if (buf) DMRG_HIP(hipFree(buf));
      buf = nullptr;
      buf_cap = 0;
      DMRG_HIP(hipMalloc(&buf, n * sizeof(T)));
      buf_cap = n;
    }
    return buf;
  }
  void need_svd(int cap) {
    if (cap > svd.cap) svd.alloc(cap);
  }
  void need_lz(int kmax, size_t n) {
    size_t pb = lz_part_doubles(kmax, n) * sizeof(double);
    if (lz.kmax < kmax || lz.part_bytes < pb) lz.alloc(std::max(kmax, 128), std::max(pb, lz.part_bytes), sizeof(T));
  }
};
// Scratch of the FP32 ops: contraction intermediates and the fused MPO.
struct CtxF32 {
  ContractScratch<float> sc;
  float* buf = nullptr;
  size_t buf_cap = 0;
  ~CtxF32() {
    if (sc.X) (void)hipFree(sc.X);
    if (sc.Y) (void)hipFree(sc.Y);
    if (buf) (void)hipFree(buf);
  }
  void need_sc(size_t n) {
    if (n <= sc.cap) return;
    if (sc.X) DMRG_HIP(hipFree(sc.X));
    if (sc.Y) DMRG_HIP(hipFree(sc.Y));
    sc.X = sc.Y = nullptr;
    sc.cap = 0;
    DMRG_HIP(hipMalloc(&sc.X, n * sizeof(float)));
    DMRG_HIP(hipMalloc(&sc.Y, n * sizeof(float)));
    sc.cap = n;
  }
  float* need_buf(size_t n) {
    if (n > buf_cap) {
      if (buf) DMRG_HIP(hipFree(buf));
      buf = nullptr;
      buf_cap = 0;
      DMRG_HIP(hipMalloc(&buf, n * sizeof(float)));
      buf_cap = n;
    }
    return buf;
  }
};
}  // namespace

struct dmrg_ctx {
  int device = 0;
  CtxT<double> r;
  CtxT<cdouble> c;
  CtxF32 f;  // the FP32 matvec of the mixed modes
  int* h_status = nullptr;
  double* h_scal = nullptr;
  ~dmrg_ctx() {
    if (h_status) (void)hipHostFree(h_status);
    if (h_scal) (void)hipHostFree(h_scal);
  }
};

namespace {
template <class T> CtxT<T>& ctx_of(dmrg_ctx* c);
template <> CtxT<double>& ctx_of<double>(dmrg_ctx* c) { return c->r; }
template <> CtxT<cdouble>& ctx_of<cdouble>(dmrg_ctx* c) { return c->c; }

template <class F>
void by_dtype(int dtype, F&& f) {
  if (dtype == DMRG_DTYPE_F64) f(double{});
  else if (dtype == DMRG_DTYPE_C128) f(cdouble{});
  else throw Error(ST_BAD_ARG, "bad dtype");
}

// Every op runs on its context's device, whichever thread calls it.
void use(const dmrg_ctx* c) { DMRG_HIP(hipSetDevice(c->device)); }

template <class T>
std::vector<T> d2h(const void* p, size_t n) {
  std::vector<T> h(n);
  DMRG_HIP(hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost));
  return h;
}
}  // namespace

extern "C" {

int dmrg_ctx_create(int device, dmrg_ctx_handle* out) {
  if (!out) return fail(DMRG_ERR_BAD_ARG, "NULL out");
  *out = nullptr;
  return guard([&] {
    DMRG_HIP(hipSetDevice(device));
    auto c = std::make_unique<dmrg_ctx>();
    c->device = device;
    DMRG_HIP(hipHostMalloc((void**)&c->h_status, 4 * sizeof(int), hipHostMallocDefault));
    DMRG_HIP(hipHostMalloc((void**)&c->h_scal, 8 * sizeof(double), hipHostMallocDefault));
    *out = c.release();
  });
}

void dmrg_ctx_destroy(dmrg_ctx_handle c) {
  if (c) (void)hipDeviceSynchronize();
  delete c;
}

int dmrg_op_heff2(dmrg_ctx_handle c, void* stream, int dtype, int cL, int cR, int d, int D,
                  const void* L, const void* R, const void* W1, const void* W2, const void* theta,
                  void* out) {
  if (!c || !L || !R || !W1 || !W2 || !theta || !out) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] {
    use(c);
    DMRG_REQUIRE(cL >= 1 && cR >= 1 && d >= 1 && D >= 1, "bad shape");
    by_dtype(dtype, [&](auto tag) {
      using T = decltype(tag);
      auto& x = ctx_of<T>(c);
      hipStream_t st = (hipStream_t)stream;
      // the caller's pending writes of the MPO and the previous call's reads
      // of the context's MPO buffer are ordered on the stream: drain it first
      DMRG_HIP(hipStreamSynchronize((hipStream_t)stream));
      const size_t wsz = (size_t)D * d * d * D, K = (size_t)D * d * d;
      auto h1 = d2h<T>(W1, wsz), h2 = d2h<T>(W2, wsz);
      std::vector<T> ww(K * K);
      build_ww2<T>(h1.data(), h2.data(), d, D, ww.data());
      T* dww = x.need_buf(K * K);
      DMRG_HIP(hipMemcpy(dww, ww.data(), K * K * sizeof(T), hipMemcpyHostToDevice));
      x.need_sc((size_t)D * d * d * cL * cR);
      heff2_apply<T>(st, (const T*)L, (const T*)R, dww, (const T*)theta, (T*)out, cL, cR, d, D, x.sc);
    });
  });
}

int dmrg_op_heff2_f32(dmrg_ctx_handle c, void* stream, int prec, int cL, int cR, int d, int D,
                      const float* L, const float* R, const double* W1, const double* W2,
                      const float* theta, float* out) {
  if (!c || !L || !R || !W1 || !W2 || !theta || !out) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] {
    use(c);
    DMRG_REQUIRE(cL >= 1 && cR >= 1 && d >= 1 && D >= 1, "bad shape");
    DMRG_REQUIRE(prec == DMRG_PREC_FP32 || prec == DMRG_PREC_BF16X3, "prec must be FP32 or BF16X3");
    hipStream_t st = (hipStream_t)stream;
    // the caller's pending writes of the MPO and the previous call's reads
    // of the context's MPO buffer are ordered on the stream: drain it first
    DMRG_HIP(hipStreamSynchronize(st));
    auto& x = c->f;
    const size_t wsz = (size_t)D * d * d * D, K = (size_t)D * d * d;
    auto h1 = d2h<double>(W1, wsz), h2 = d2h<double>(W2, wsz);
    std::vector<double> ww(K * K);
    build_ww2<double>(h1.data(), h2.data(), d, D, ww.data());
    std::vector<float> wwf(K * K);
    for (size_t i = 0; i < K * K; ++i) wwf[i] = (float)ww[i];
    float* dww = x.need_buf(K * K);
    DMRG_HIP(hipMemcpy(dww, wwf.data(), K * K * sizeof(float), hipMemcpyHostToDevice));
    x.need_sc((size_t)D * d * d * cL * cR);
    x.sc.split_bf16 = prec == DMRG_PREC_BF16X3;
    heff2_apply<float>(st, L, R, dww, theta, out, cL, cR, d, D, x.sc);
  });
}

int dmrg_op_heff1(dmrg_ctx_handle c, void* stream, int dtype, int cL, int cR, int d, int D,
                  const void* L, const void* R, const void* W, const void* theta, void* out) {
  if (!c || !L || !R || !W || !theta || !out) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] {
    use(c);
    DMRG_REQUIRE(cL >= 1 && cR >= 1 && d >= 1 && D >= 1, "bad shape");
    by_dtype(dtype, [&](auto tag) {
      using T = decltype(tag);
      auto& x = ctx_of<T>(c);
      x.need_sc((size_t)D * d * cL * cR);
      heff1_apply<T>((hipStream_t)stream, (const T*)L, (const T*)R, (const T*)W, (const T*)theta,
                     (T*)out, cL, cR, d, D, x.sc);
    });
  });
}

int dmrg_op_env_left(dmrg_ctx_handle c, void* stream, int dtype, int ci, int co, int d, int D,
                     const void* L, const void* A, const void* W, void* Lnew) {
  if (!c || !L || !A || !W || !Lnew) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] {
    use(c);
    DMRG_REQUIRE(ci >= 1 && co >= 1 && d >= 1 && D >= 1, "bad shape");
    by_dtype(dtype, [&](auto tag) {
      using T = decltype(tag);
      auto& x = ctx_of<T>(c);
      x.need_sc((size_t)D * d * ci * co);
      env_left<T>((hipStream_t)stream, (const T*)L, (const T*)A, (const T*)W, (T*)Lnew, ci, co, d, D, x.sc);
    });
  });
}

int dmrg_op_env_left_noalloc(void* stream, int dtype, int ci, int co, int d, int D, const void* L,
                             const void* A, const void* W, void* Lnew) {
  if (!L || !A || !W || !Lnew) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] {
    env_left_noalloc_op((hipStream_t)stream, dtype, ci, co, d, D, L, A, W, Lnew);
  });
}

int dmrg_op_env_right(dmrg_ctx_handle c, void* stream, int dtype, int ci, int co, int d, int D,
                      const void* R, const void* A, const void* W, void* Rnew) {
  if (!c || !R || !A || !W || !Rnew) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] {
    use(c);
    DMRG_REQUIRE(ci >= 1 && co >= 1 && d >= 1 && D >= 1, "bad shape");
    by_dtype(dtype, [&](auto tag) {
      using T = decltype(tag);
      auto& x = ctx_of<T>(c);
      // the caller's pending writes of the MPO and the previous call's reads
      // of the context's MPO buffer are ordered on the stream: drain it first
      DMRG_HIP(hipStreamSynchronize((hipStream_t)stream));
      const size_t wsz = (size_t)D * d * d * D;
      auto hw = d2h<T>(W, wsz);
      std::vector<T> wr(wsz);
      build_wright<T>(hw.data(), d, D, wr.data());
      T* dwr = x.need_buf(wsz);
      DMRG_HIP(hipMemcpy(dwr, wr.data(), wsz * sizeof(T), hipMemcpyHostToDevice));
      x.need_sc((size_t)D * d * ci * co);
      env_right<T>((hipStream_t)stream, (const T*)R, (const T*)A, dwr, (T*)Rnew, ci, co, d, D, x.sc);
    });
  });
}

int dmrg_op_svd_split(dmrg_ctx_handle c, void* stream, int dtype, int m, int n, int k,
                      double cutoff, double rel_cutoff, int scale_left, const void* M, void* left,
                      void* right, double* sigma, int* nk_out) {
  if (!c || !M || !left || !right || !nk_out) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] {
    use(c);
    DMRG_REQUIRE(m >= 1 && n >= 1 && k >= 1 && k <= std::min(m, n), "bad shape");
    by_dtype(dtype, [&](auto tag) {
      using T = decltype(tag);
      auto& x = ctx_of<T>(c);
      hipStream_t st = (hipStream_t)stream;
      // the staging buffers hold products up to max(m, n) on a side (U_k^H M is k x n)
      x.need_svd(std::max(m, n));
      SvdResult r = svd_split<T>(st, x.svd, (const T*)M, m, n, m, k, cutoff, rel_cutoff,
                                 scale_left != 0, (T*)left, m, (T*)right, sigma, SvdParams{});
      DMRG_HIP(hipStreamSynchronize(st));
      *nk_out = r.nk;
    });
  });
}

int dmrg_op_rsvd_split(dmrg_ctx_handle c, void* stream, int dtype, int m, int n, int k,
                       int oversample, int power, uint64_t seed, double cutoff,
                       double rel_cutoff, int scale_left, const void* M, void* left,
                       void* right, int* nk_out) {
  if (!c || !M || !left || !right || !nk_out) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] {
    use(c);
    DMRG_REQUIRE(m >= 1 && n >= 1 && k >= 1 && k <= std::min(m, n), "bad shape");
    DMRG_REQUIRE(oversample >= 0 && power >= 0 && power <= 8, "bad oversample / power");
    by_dtype(dtype, [&](auto tag) {
      using T = decltype(tag);
      auto& x = ctx_of<T>(c);
      hipStream_t st = (hipStream_t)stream;
      const int r = std::min(k + oversample, std::min(m, n));
      x.need_svd(std::max(std::max(m, n), r));
      if (x.rsvd.len < std::max(m, n) || x.rsvd.r < r) x.rsvd.alloc(std::max(m, n), r);
      SvdResult res = rsvd_split<T>(st, x.svd, x.rsvd, (const T*)M, m, n, m, k, cutoff,
                                    rel_cutoff, scale_left != 0, (T*)left, m, (T*)right,
                                    SvdParams{}, oversample, power, seed);
      DMRG_HIP(hipStreamSynchronize(st));
      *nk_out = res.nk;
    });
  });
}

int dmrg_op_heig(dmrg_ctx_handle c, void* stream, int dtype, int N, const void* H, double* evals,
                 void* evecs, int* sweeps_out) {
  if (!c || !H || !evals || !evecs) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] {
    use(c);
    DMRG_REQUIRE(N >= 1 && N <= 2048, "bad shape");
    by_dtype(dtype, [&](auto tag) {
      using T = decltype(tag);
      auto& x = ctx_of<T>(c);
      hipStream_t st = (hipStream_t)stream;
      x.need_svd(N);
      int sw = heig<T>(st, x.svd, (const T*)H, N, N, evals, (T*)evecs, N, SvdParams{});
      DMRG_HIP(hipStreamSynchronize(st));
      if (sweeps_out) *sweeps_out = sw;
    });
  });
}

int dmrg_op_tridiag_lowest(dmrg_ctx_handle c, int k, const double* alpha, const double* beta,
                           double* lam, double* y) {
  if (!c || !alpha || !lam || !y || (k > 1 && !beta)) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] {
    use(c);
    DMRG_REQUIRE(k >= 1 && k <= 128, "k must be in [1, 128]");
    auto& x = c->r;
    x.need_lz(128, 1);
    DMRG_HIP(hipMemcpy(x.lz.alpha, alpha, k * sizeof(double), hipMemcpyHostToDevice));
    if (k > 1) DMRG_HIP(hipMemcpy(x.lz.beta, beta, (k - 1) * sizeof(double), hipMemcpyHostToDevice));
    lz_final(0, x.lz, k);
    DMRG_HIP(hipDeviceSynchronize());
    double sc[8];
    DMRG_HIP(hipMemcpy(sc, x.lz.scal, sizeof(sc), hipMemcpyDeviceToHost));
    DMRG_HIP(hipMemcpy(y, x.lz.y, k * sizeof(double), hipMemcpyDeviceToHost));
    *lam = sc[3];
  });
}

int dmrg_op_lanczos2(dmrg_ctx_handle c, void* stream, int dtype, int cL, int cR, int d, int D,
                     const void* L, const void* R, const void* W1, const void* W2, void* theta,
                     int krylov_max, int site, double* energy, int* niter) {
  if (!c || !L || !R || !W1 || !W2 || !theta || !energy) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] {
    use(c);
    DMRG_REQUIRE(cL >= 1 && cR >= 1 && d >= 1 && D >= 1, "bad shape");
    DMRG_REQUIRE(krylov_max >= 1 && krylov_max <= 128, "krylov_max must be in [1, 128]");
    by_dtype(dtype, [&](auto tag) {
      using T = decltype(tag);
      auto& x = ctx_of<T>(c);
      hipStream_t st = (hipStream_t)stream;
      const size_t n = (size_t)cL * d * d * cR;
      // the caller's pending writes of the MPO and the previous call's reads
      // of the context's MPO buffer are ordered on the stream: drain it first
      DMRG_HIP(hipStreamSynchronize((hipStream_t)stream));
      const size_t wsz = (size_t)D * d * d * D, K = (size_t)D * d * d;
      auto h1 = d2h<T>(W1, wsz), h2 = d2h<T>(W2, wsz);
      std::vector<T> ww(K * K);
      build_ww2<T>(h1.data(), h2.data(), d, D, ww.data());
      T* dww = x.need_buf(K * K);
      DMRG_HIP(hipMemcpy(dww, ww.data(), K * K * sizeof(T), hipMemcpyHostToDevice));
      x.need_sc((size_t)D * d * d * cL * cR);
      const int kmax = (int)std::min<size_t>(krylov_max, n);
      if ((size_t)kmax * n > x.V_cap) {
        if (x.V) DMRG_HIP(hipFree(x.V));
        x.V = nullptr;
        x.V_cap = 0;
        DMRG_HIP(hipMalloc(&x.V, (size_t)kmax * n * sizeof(T)));
        x.V_cap = (size_t)kmax * n;
      }
      if (n > x.work_cap) {
        if (x.work) DMRG_HIP(hipFree(x.work));
        x.work = nullptr;
        x.work_cap = 0;
        DMRG_HIP(hipMalloc(&x.work, n * sizeof(T)));
        x.work_cap = n;
      }
      x.need_lz(kmax, n);
      x.runner.lz = &x.lz;
      x.runner.V = x.V;
      x.runner.work = x.work;
      x.runner.h_status = c->h_status;
      x.runner.h_scal = c->h_scal;
      double E = x.runner.run(st, (T*)theta, n, kmax, site, [&](const T* in, T* o) {
        heff2_apply<T>(st, (const T*)L, (const T*)R, dww, in, o, cL, cR, d, D, x.sc);
      });
      *energy = E;
      if (niter) *niter = x.runner.last_niter;
    });
  });
}

int dmrg_op_gemm(dmrg_ctx_handle c, void* stream, int dtype, int opA, int opB, int M, int N,
                 int K, const void* alpha, const void* A, int lda, const void* B, int ldb,
                 const void* beta, void* Cm, int ldc) {
  if (!c || !alpha || !beta || !A || !B || !Cm) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] {
    use(c);
    DMRG_REQUIRE(M >= 1 && N >= 1 && K >= 1, "bad shape");
    DMRG_REQUIRE(ldc >= M, "ldc < M");
    DMRG_REQUIRE(lda >= (op_trans(opA) ? K : M) && ldb >= (op_trans(opB) ? N : K), "bad ld");
    by_dtype(dtype, [&](auto tag) {
      using T = decltype(tag);
      gemm<T>((hipStream_t)stream, opA, opB, M, N, K, *(const T*)alpha, (const T*)A, lda,
              (const T*)B, ldb, *(const T*)beta, (T*)Cm, ldc,
              gemm_pick_tile(M, N, K, is_cplx<T>::value));
    });
  });
}

int dmrg_op_gemm_f32(dmrg_ctx_handle c, void* stream, int prec, int opA, int opB, int M, int N,
                     int K, const float* A, int lda, const float* B, int ldb, float* Cm, int ldc) {
  if (!c || !A || !B || !Cm) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] {
    use(c);
    DMRG_REQUIRE(M >= 1 && N >= 1 && K >= 1, "bad shape");
    DMRG_REQUIRE((opA == OP_N || opA == OP_T) && (opB == OP_N || opB == OP_T), "op must be N or T");
    DMRG_REQUIRE(ldc >= M, "ldc < M");
    DMRG_REQUIRE(lda >= (op_trans(opA) ? K : M) && ldb >= (op_trans(opB) ? N : K), "bad ld");
    DMRG_REQUIRE(prec == DMRG_PREC_FP32 || prec == DMRG_PREC_BF16X3, "prec must be FP32 or BF16X3");
    if (prec == DMRG_PREC_BF16X3)
      gemm_f32_split_bf16((hipStream_t)stream, opA, opB, M, N, K, A, lda, B, ldb, Cm, ldc);
    else
      gemm<float>((hipStream_t)stream, opA, opB, M, N, K, 1.0f, A, lda, B, ldb, 0.0f, Cm, ldc,
                  gemm_pick_tile(M, N, K, false));
  });
}

int dmrg_dev_malloc(void** p, size_t bytes) {
  if (!p) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] { DMRG_HIP(hipMalloc(p, bytes)); });
}
int dmrg_dev_free(void* p) {
  return guard([&] { if (p) DMRG_HIP(hipFree(p)); });
}
int dmrg_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  return guard([&] { DMRG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}
int dmrg_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  return guard([&] { DMRG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}
int dmrg_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  return guard([&] { DMRG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice)); });
}
int dmrg_dev_synchronize(void) {
  return guard([&] { DMRG_HIP(hipDeviceSynchronize()); });
}
int dmrg_device_count(int* n) {
  if (!n) return fail(DMRG_ERR_BAD_ARG, "NULL argument");
  return guard([&] {
    hipError_t e = hipGetDeviceCount(n);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      *n = 0;
    }
  });
}

}  // extern "C"
