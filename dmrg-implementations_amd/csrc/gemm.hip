// MFMA GEMM kernels for gfx950 (see gemm.hpp).
//
// Workgroup: 256 threads = 4 wave64s in a 2x2 grid.  Each wave owns a
// (BM/2) x (BN/2) sub-tile made of 16x16 MFMA blocks.  Operand tiles of depth
// BK (16 or 32) are staged through LDS (double buffered, one-element row
// padding), with the global loads of the next tile (PF=1) or of the next two
// tiles (PF=2) issued before the MFMAs of the current one.  The MFMA is issued with the B fragment in the "A" slot so
// that each lane's four accumulator registers land on four columns of C with
// consecutive lanes on consecutive rows: epilogue stores of a column-major C
// are 128-byte runs.
#include "gemm.hpp"

#include <type_traits>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <unordered_map>

namespace dmrg {
namespace {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <class T> struct GT;
// G1: accumulator register r of lane l holds output row (l/16) + 4r of the
// 16x16 block (f64 MFMA, group size 1); otherwise row 4(l/16) + r (f32 MFMA,
// group size 4).  Column = l % 16 in both cases.
template <> struct GT<double> { using E = double; using V4 = d4_t; static constexpr int P = 1; static constexpr bool G1 = true; };
template <> struct GT<cdouble> { using E = double; using V4 = d4_t; static constexpr int P = 2; static constexpr bool G1 = true; };
template <> struct GT<float> { using E = float; using V4 = f4_t; static constexpr int P = 1; static constexpr bool G1 = false; };

__device__ inline d4_t mfma(double a, double b, d4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline f4_t mfma(float a, float b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <class T> __device__ inline void split(T v, typename GT<T>::E* out);
template <> __device__ inline void split<double>(double v, double* o) { o[0] = v; }
template <> __device__ inline void split<float>(float v, float* o) { o[0] = v; }
template <> __device__ inline void split<cdouble>(cdouble v, double* o) { o[0] = v.x; o[1] = v.y; }

template <class T> __device__ inline T conj_v(T v) { return conj(v); }
template <> __device__ inline float conj_v<float>(float v) { return v; }

// Plain column-major operand; TR: stored transposed.
template <bool TR> struct AddrPlain {
  long long ld;
  __device__ long long operator()(int r, int c) const {
    return TR ? (long long)c + (long long)r * ld : (long long)r + (long long)c * ld;
  }
};
// Two-level operand.
struct AddrTensor {
  View2 v;
  __device__ static long long dim(const Dim2& d, int i) {
    return (long long)(i % d.n1) * d.s1 + (long long)(i / d.n1) * d.s2;
  }
  __device__ long long operator()(int r, int c) const { return dim(v.row, r) + dim(v.col, c); }
};

// KS = 2: eight waves, two per output sub-tile, each multiplying one half of
// every k-tile; the halves are summed through LDS at the end (twice the waves
// per SIMD to hide latency when a grid has about one workgroup per CU).
template <class T, bool CA, bool CB, bool KFA, bool KFB, int BM, int BN, int BK, int PF, int KS,
          class AA, class AB, class AC>
__global__ __launch_bounds__(256 * KS) void k_gemm(int M, int N, int K, T alpha,
                                                   const T* __restrict__ A, AA aa, long long sA,
                                                   const T* __restrict__ B, AB ab, long long sB,
                                                   T beta, T* __restrict__ C, AC ac, long long sC,
                                                   int herm) {
  using E = typename GT<T>::E;
  using V4 = typename GT<T>::V4;
  constexpr int P = GT<T>::P;
  constexpr int NT = 256 * KS;
  constexpr int PADM = BM + 1, PADN = BN + 1;
  constexpr int FM = BM / 32, FN = BN / 32;  // 16x16 blocks per wave
  constexpr int NA = BM * BK / NT, NB = BN * BK / NT;
  static_assert(KS == 1 || KS == 2, "k split");
  static_assert(NA >= 1 && NB >= 1 && BK % (4 * KS) == 0, "tile too small for the k split");

  __shared__ E As[2][P][BK][PADM];
  __shared__ E Bs[2][P][BK][PADN];
  __shared__ E Red[KS > 1 ? 4 * P * FM * FN * 4 * 64 : 1];

  // Hermitian product (M == N, BM == BN, beta == 0): only the tiles on and
  // above the diagonal are computed; each writes its conjugate mirror too.
  if (herm && blockIdx.x > blockIdx.y) return;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = (wave >> 1) & 1, wk = wave >> 2;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  A += blockIdx.z * sA;
  B += blockIdx.z * sB;
  C += blockIdx.z * sC;

  static_assert(PF >= 1 && PF <= 4, "prefetch depth");
  T ra[PF][NA], rb[PF][NB];

  auto idxA = [&](int r, int& m, int& k) {
    if (!KFA) { m = t % BM; k = t / BM + r * (NT / BM); }
    else      { k = t % BK; m = t / BK + r * (NT / BK); }
  };
  auto idxB = [&](int r, int& n, int& k) {
    if (!KFB) { n = t % BN; k = t / BN + r * (NT / BN); }
    else      { k = t % BK; n = t / BK + r * (NT / BK); }
  };
  // register stage `s` <- operand tiles at depth k0; a tile inside all three
  // bounds loads without per-element checks (one branch for the whole stage)
  const bool inner_mn = m0 + BM <= M && n0 + BN <= N;
  auto load = [&](int s, int k0) {
    if (inner_mn && k0 + BK <= K) {
#pragma unroll
      for (int r = 0; r < NA; ++r) {
        int m, k;
        idxA(r, m, k);
        const T v = A[aa(m0 + m, k0 + k)];
        ra[s][r] = CA ? conj_v(v) : v;
      }
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        int n, k;
        idxB(r, n, k);
        const T v = B[ab(k0 + k, n0 + n)];
        rb[s][r] = CB ? conj_v(v) : v;
      }
      return;
    }
#pragma unroll
    for (int r = 0; r < NA; ++r) {
      int m, k;
      idxA(r, m, k);
      int gm = m0 + m, gk = k0 + k;
      T v = T(0);
      if (gm < M && gk < K) v = A[aa(gm, gk)];
      ra[s][r] = CA ? conj_v(v) : v;
    }
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      int n, k;
      idxB(r, n, k);
      int gn = n0 + n, gk = k0 + k;
      T v = T(0);
      if (gn < N && gk < K) v = B[ab(gk, gn)];
      rb[s][r] = CB ? conj_v(v) : v;
    }
  };
  // LDS buffer `buf` <- register stage `s`
  auto store = [&](int s, int buf) {
#pragma unroll
    for (int r = 0; r < NA; ++r) {
      int m, k;
      idxA(r, m, k);
      E e[P];
      split<T>(ra[s][r], e);
#pragma unroll
      for (int p = 0; p < P; ++p) As[buf][p][k][m] = e[p];
    }
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      int n, k;
      idxB(r, n, k);
      E e[P];
      split<T>(rb[s][r], e);
#pragma unroll
      for (int p = 0; p < P; ++p) Bs[buf][p][k][n] = e[p];
    }
  };

  V4 acc[P][FM][FN];
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) acc[p][i][j] = V4{0, 0, 0, 0};

  const int nk = (K + BK - 1) / BK;
  const int li = lane & 15, lk = lane >> 4;
  auto compute = [&](int buf) {
#pragma unroll
    for (int kk = wk * (BK / KS); kk < (wk + 1) * (BK / KS); kk += 4) {
      E af[P][FM], bf[P][FN];
#pragma unroll
      for (int p = 0; p < P; ++p) {
#pragma unroll
        for (int i = 0; i < FM; ++i) af[p][i] = As[buf][p][kk + lk][wm * (BM / 2) + i * 16 + li];
#pragma unroll
        for (int j = 0; j < FN; ++j) bf[p][j] = Bs[buf][p][kk + lk][wn * (BN / 2) + j * 16 + li];
      }
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          if constexpr (P == 1) {
            acc[0][i][j] = mfma(bf[0][j], af[0][i], acc[0][i][j]);
          } else {
            acc[0][i][j] = mfma(bf[0][j], af[0][i], acc[0][i][j]);
            acc[0][i][j] = mfma(-bf[1][j], af[1][i], acc[0][i][j]);
            acc[1][i][j] = mfma(bf[1][j], af[0][i], acc[1][i][j]);
            acc[1][i][j] = mfma(bf[0][j], af[1][i], acc[1][i][j]);
          }
        }
    }
  };
  // PF register stages: tile j travels in stage j % PF, then LDS buffer
  // j & 1.  While tile kt is multiplied, tiles kt+1 .. kt+PF are in flight
  // (kt+PF issued into the stage that tile kt just left); tile kt+1 is then
  // staged into the other LDS buffer.  The loop is unrolled by PF so that
  // stage indices are compile-time constants.
  load(0, 0);
  store(0, 0);
#pragma unroll
  for (int s = 1; s < PF; ++s)
    if (s < nk) load(s, s * BK);
  __syncthreads();
  for (int kt = 0; kt < nk; kt += PF) {
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int tile = kt + u;
      if (tile >= nk) break;
      const int buf = (PF % 2 == 0) ? (u & 1) : (tile & 1);  // kt is even when PF is
      if (tile + PF < nk) load(u, (tile + PF) * BK);
      compute(buf);
      if (tile + 1 < nk) store((u + 1) % PF, buf ^ 1);
      __syncthreads();
    }
  }

  if constexpr (KS > 1) {
    // the second k-half's partial sums join the first's (the loop ended at
    // a barrier, so no wave reads As/Bs any more)
    E* red = Red + (size_t)(wave & 3) * (P * FM * FN * 4 * 64);
    if (wk == 1) {
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) red[(((p * FM + i) * FN + j) * 4 + r) * 64 + lane] = acc[p][i][j][r];
    }
    __syncthreads();
    if (wk == 1) return;
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[p][i][j][r] += red[(((p * FM + i) * FN + j) * 4 + r) * 64 + lane];
  }
  // Epilogue: lane -> row li of C; register r -> column lk + 4r (f64) or
  // 4*lk + r (f32) of each 16x16 block.
  const bool use_beta = !(real(beta) == 0.0 && imag(beta) == 0.0);
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int gm = m0 + wm * (BM / 2) + i * 16 + li;
        int gn = n0 + wn * (BN / 2) + j * 16 + (GT<T>::G1 ? lk + 4 * r : 4 * lk + r);
        if (gm < M && gn < N) {
          T v;
          if constexpr (P == 1) v = T(acc[0][i][j][r]);
          else v = make_scalar<T>(acc[0][i][j][r], acc[1][i][j][r]);
          v = alpha * v;
          long long idx = ac(gm, gn);
          if (use_beta) v = v + beta * C[idx];
          C[idx] = v;
          if (herm && blockIdx.x != blockIdx.y) C[ac(gn, gm)] = conj_v(v);
        }
      }
}

template <class T, bool CA, bool CB, bool KFA, bool KFB, class AA, class AB, class AC>
void launch(hipStream_t st, GemmTile tl, int M, int N, int K, T alpha, const T* A, AA aa,
            long long sA, const T* B, AB ab, long long sB, T beta, T* C, AC ac, long long sC,
            int batch, int herm = 0) {
#define DMRG_GEMM_LAUNCH(BM_, BN_, BK_, PF_)                                                \
  if (tl.ks == 2)                                                                          \
    hipLaunchKernelGGL((k_gemm<T, CA, CB, KFA, KFB, BM_, BN_, BK_, PF_, 2, AA, AB, AC>),    \
                       dim3(cdiv(M, BM_), cdiv(N, BN_), batch), dim3(512), 0, st, M, N, K,  \
                       alpha, A, aa, sA, B, ab, sB, beta, C, ac, sC, herm);                 \
  else                                                                                     \
    hipLaunchKernelGGL((k_gemm<T, CA, CB, KFA, KFB, BM_, BN_, BK_, PF_, 1, AA, AB, AC>),    \
                       dim3(cdiv(M, BM_), cdiv(N, BN_), batch), dim3(256), 0, st, M, N, K,  \
                       alpha, A, aa, sA, B, ab, sB, beta, C, ac, sC, herm)
  // 32x32 tiles serve grids that cannot fill the chip with 64x64 ones; a deep
  // k-step and two register stages then hide the load latency that the low
  // tile count per CU leaves exposed.
  // 64x64 tiles: two register stages as well (measured on the H_eff shapes:
  // FP64 +12-15 % with BK=16, FP32 +5-26 % with BK=32).  Complex: BK=16 and
  // two stages for both tiles (+4-13 % on the C3 shapes; four MFMAs per k-step).
  if (tl.bm == 32 && tl.bn == 32) {
    if constexpr (is_cplx<T>::value) DMRG_GEMM_LAUNCH(32, 32, 16, 2);
    else DMRG_GEMM_LAUNCH(32, 32, 32, 2);
  } else {
    if constexpr (is_cplx<T>::value) DMRG_GEMM_LAUNCH(64, 64, 16, 2);
    else if constexpr (std::is_same<T, float>::value) DMRG_GEMM_LAUNCH(64, 64, 32, 2);
    else DMRG_GEMM_LAUNCH(64, 64, 16, 2);
  }
#undef DMRG_GEMM_LAUNCH
  DMRG_LAUNCH_CHECK();
}

template <class T, bool TA, bool TB, bool CA, bool CB>
void plain4(hipStream_t st, GemmTile tl, int M, int N, int K, T alpha, const T* A, int lda,
            long long sA, const T* B, int ldb, long long sB, T beta, T* C, int ldc, long long sC,
            int batch, int herm) {
  launch<T, CA, CB, TA, !TB>(st, tl, M, N, K, alpha, A, AddrPlain<TA>{lda}, sA, B,
                             AddrPlain<TB>{ldb}, sB, beta, C, AddrPlain<false>{ldc}, sC, batch,
                             herm);
}

template <class T>
void dispatch(hipStream_t st, int opA, int opB, GemmTile tl, int M, int N, int K, T alpha,
              const T* A, int lda, long long sA, const T* B, int ldb, long long sB, T beta, T* C,
              int ldc, long long sC, int batch, int herm) {
  const bool ta = op_trans(opA), tb = op_trans(opB);
  bool ca = op_conj(opA), cb = op_conj(opB);
  if (!is_cplx<T>::value) ca = cb = false;
#define DMRG_P4(TA_, TB_, CA_, CB_) \
  plain4<T, TA_, TB_, CA_, CB_>(st, tl, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, \
                                batch, herm)
  if (!ca && !cb) {
    if (!ta && !tb) DMRG_P4(false, false, false, false);
    else if (!ta && tb) DMRG_P4(false, true, false, false);
    else if (ta && !tb) DMRG_P4(true, false, false, false);
    else DMRG_P4(true, true, false, false);
  } else {
    if constexpr (is_cplx<T>::value) {
      int code = (ta ? 8 : 0) | (tb ? 4 : 0) | (ca ? 2 : 0) | (cb ? 1 : 0);
      switch (code) {
        case 1: DMRG_P4(false, false, false, true); break;
        case 2: DMRG_P4(false, false, true, false); break;
        case 3: DMRG_P4(false, false, true, true); break;
        case 5: DMRG_P4(false, true, false, true); break;
        case 6: DMRG_P4(false, true, true, false); break;
        case 7: DMRG_P4(false, true, true, true); break;
        case 9: DMRG_P4(true, false, false, true); break;
        case 10: DMRG_P4(true, false, true, false); break;
        case 11: DMRG_P4(true, false, true, true); break;
        case 13: DMRG_P4(true, true, false, true); break;
        case 14: DMRG_P4(true, true, true, false); break;
        default: DMRG_P4(true, true, true, true); break;
      }
    }
  }
#undef DMRG_P4
}

// ---------------------------------------------------------------------------
// Tile autotuner.  When enabled (gemm_autotune), the first non-batched call
// of each (scalar type, ops, M, N, K, Hermitian) shape whose caller tile has
// one wave per output sub-tile times the 32x32 and 64x64 tiles on the
// caller's stream, with the device otherwise idle (one warm-up each, then
// four interleaved rounds of one launch each into a scratch output, keeping
// each tile's fastest), and keeps the fastest for the rest of the process --
// the caller's tile unless the other is at least 3 % faster; later calls of
// that shape use it.  Both tiles sum over k in the same order, so tuning
// never changes a result: runs are bit-reproducible across processes.  A call made while its stream is being captured
// into a graph uses the caller's tile and records nothing.
// ---------------------------------------------------------------------------
struct TuneKey {
  int dt, opA, opB, M, N, K, herm;
  bool operator==(const TuneKey& o) const {
    return dt == o.dt && opA == o.opA && opB == o.opB && M == o.M && N == o.N && K == o.K &&
           herm == o.herm;
  }
};
struct TuneKeyHash {
  size_t operator()(const TuneKey& k) const {
    uint64_t h = 1469598103934665603ull;
    for (int v : {k.dt, k.opA, k.opB, k.M, k.N, k.K, k.herm}) {
      h ^= (uint64_t)(uint32_t)v;
      h *= 1099511628211ull;
    }
    return (size_t)h;
  }
};
std::mutex g_tune_mu;
std::unordered_map<TuneKey, GemmTile, TuneKeyHash> g_tune;
std::atomic<int> g_tune_on{0};
// DMRG_GEMM_AUTOTUNE_LOG=1 prints each tuning decision (shape, candidate
// times, choice) to stderr.
const bool g_tune_log = [] {
  const char* v = std::getenv("DMRG_GEMM_AUTOTUNE_LOG");
  return v && v[0] == '1';
}();
constexpr int kTuneDevices = 64;
void* g_tune_buf[kTuneDevices] = {};
size_t g_tune_cap[kTuneDevices] = {};

template <class T> constexpr int dt_code() { return is_cplx<T>::value ? 1 : (std::is_same<T, float>::value ? 2 : 0); }

template <class T>
GemmTile tuned_tile(hipStream_t st, int opA, int opB, GemmTile tl, int M, int N, int K,
                    const T* A, int lda, const T* B, int ldb, int herm) {
  // Only tiles that sum over k in the caller's order are candidates, so the
  // choice never changes a result: with one wave per output sub-tile every
  // tile size accumulates k in the same sequence; an eight-wave (k-split)
  // tile's order depends on its k-tile, so a k-split caller tile stands.
  if (tl.ks != 1) return tl;
  const TuneKey key{dt_code<T>(), opA, opB, M, N, K, herm};
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(key);
  if (it != g_tune.end()) return it->second;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  DMRG_HIP(hipStreamIsCapturing(st, &cs));
  if (cs != hipStreamCaptureStatusNone) return tl;
  int dev = 0;
  DMRG_HIP(hipGetDevice(&dev));
  DMRG_REQUIRE(dev >= 0 && dev < kTuneDevices, "gemm autotune: device ordinal out of range");
  const size_t bytes = (size_t)M * N * sizeof(T);
  if (g_tune_cap[dev] < bytes) {
    if (g_tune_buf[dev]) DMRG_HIP(hipFree(g_tune_buf[dev]));
    g_tune_buf[dev] = nullptr;
    g_tune_cap[dev] = 0;
    DMRG_HIP(hipMalloc(&g_tune_buf[dev], bytes));
    g_tune_cap[dev] = bytes;
  }
  T* Cs = static_cast<T*>(g_tune_buf[dev]);
  hipEvent_t e0, e1;
  DMRG_HIP(hipEventCreate(&e0));
  DMRG_HIP(hipEventCreate(&e1));
  constexpr int NC = 2, ROUNDS = 4;
  const GemmTile cand[NC] = {{32, 32, 1}, {64, 64, 1}};
  float t_min[NC];
  GemmTile best = tl;
  float best_ms = 1e30f;
  try {
    // Time on an idle device: other streams' work (the environment updates
    // that overlap a bond) would otherwise bias whichever candidate it meets.
    DMRG_HIP(hipDeviceSynchronize());
    for (const GemmTile& c : cand)  // warm-up: code and operands into the caches
      dispatch<T>(st, opA, opB, c, M, N, K, T(1), A, lda, 0, B, ldb, 0, T(0), Cs, M, 0, 1, herm);
    // Rounds interleave the candidates, and each keeps its fastest launch,
    // so work that other streams run meanwhile biases no candidate.
    for (int r = 0; r < ROUNDS; ++r)
      for (int c = 0; c < NC; ++c) {
        DMRG_HIP(hipEventRecord(e0, st));
        dispatch<T>(st, opA, opB, cand[c], M, N, K, T(1), A, lda, 0, B, ldb, 0, T(0), Cs, M, 0, 1,
                    herm);
        DMRG_HIP(hipEventRecord(e1, st));
        DMRG_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        DMRG_HIP(hipEventElapsedTime(&ms, e0, e1));
        t_min[c] = r == 0 ? ms : std::min(t_min[c], ms);
      }
    // the caller's tile stands unless another is at least 3 % faster
    for (int c = 0; c < NC; ++c)
      if (cand[c].bm == tl.bm && cand[c].bn == tl.bn && cand[c].ks == tl.ks) best_ms = t_min[c];
    for (int c = 0; c < NC; ++c)
      if (t_min[c] < 0.97f * best_ms) {
        best_ms = t_min[c];
        best = cand[c];
      }
    if (g_tune_log)
      std::fprintf(stderr,
                   "[gemm autotune] dt %d op %d%d%s M %d N %d K %d: 32x32 %.1f, 64x64 %.1f us -> "
                   "%dx%d\n",
                   key.dt, opA, opB, herm ? " herm" : "", M, N, K, 1e3f * t_min[0],
                   1e3f * t_min[1], best.bm, best.bn);
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  g_tune.emplace(key, best);
  return best;
}

template <class T>
void plain(hipStream_t st, int opA, int opB, GemmTile tl, int M, int N, int K, T alpha, const T* A,
           int lda, long long sA, const T* B, int ldb, long long sB, T beta, T* C, int ldc,
           long long sC, int batch, int herm = 0) {
  DMRG_REQUIRE(M >= 0 && N >= 0 && K >= 0 && batch >= 0, "gemm: negative size");
  DMRG_REQUIRE(opA >= 0 && opA <= 3 && opB >= 0 && opB <= 3, "gemm: bad op");
  if (M == 0 || N == 0 || batch == 0) return;
  if (batch == 1 && g_tune_on.load(std::memory_order_relaxed))
    tl = tuned_tile<T>(st, opA, opB, tl, M, N, K, A, lda, B, ldb, herm);
  dispatch<T>(st, opA, opB, tl, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch,
              herm);
}

}  // namespace

void gemm_autotune(bool on) { g_tune_on.store(on ? 1 : 0); }
int gemm_autotune_entries() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return (int)g_tune.size();
}

GemmTile gemm_pick_tile(int M, int N, int K, bool complex_) {
  (void)K;
  const long long t64 = (long long)cdiv(M, 64) * cdiv(N, 64);
  const long long t32 = (long long)cdiv(M, 32) * cdiv(N, 32);
  // 64x64 once there are about four tiles per CU (256 CUs); complex too
  // (measured at 1280^3: 32x32 with two stages 53.8 TF/s, 64x64 48.3).
  // At most one 32x32 tile per CU: eight waves splitting every k-tile
  // (512^3: 24.5 TF/s against 21.3 with four waves).
  (void)complex_;
  if (t64 >= 1024) return {64, 64, 1};
  return {32, 32, t32 <= 256 ? 2 : 1};
}

namespace {
// One thread per row r: C(r, :) = A(r, :) B with B (K x N, column-major),
// K <= KM and N <= 32.  B sits in LDS zero-padded to KM rows, so the unrolled
// k-loop needs no bounds; the operand column offsets are tabulated once per
// workgroup and each thread does one div/mod pair per operand row.
template <class T, int KM>
__global__ __launch_bounds__(256) void k_rows_small(int M, int N, int K, const T* __restrict__ A,
                                                    View2 va, const T* __restrict__ B,
                                                    T* __restrict__ C, View2 vc) {
  __shared__ T Bs[KM * 32];
  __shared__ long long ca[KM], cc[32];
  for (int e = threadIdx.x; e < KM * 32; e += 256) {
    const int k = e % KM, j = e / KM;
    Bs[e] = (k < K && j < N) ? B[k + (size_t)K * j] : T(0);
  }
  for (int k = threadIdx.x; k < KM; k += 256) ca[k] = k < K ? AddrTensor::dim(va.col, k) : 0;
  for (int j = threadIdx.x; j < N; j += 256) cc[j] = AddrTensor::dim(vc.col, j);
  __syncthreads();
  for (int r = blockIdx.x * 256 + threadIdx.x; r < M; r += gridDim.x * 256) {
    const long long ra = AddrTensor::dim(va.row, r), rc = AddrTensor::dim(vc.row, r);
    T x[KM];
#pragma unroll
    for (int k = 0; k < KM; ++k) x[k] = k < K ? A[ra + ca[k]] : T(0);
    for (int j = 0; j < N; ++j) {
      T y = T(0);
#pragma unroll
      for (int k = 0; k < KM; ++k) y = y + x[k] * Bs[k + KM * j];
      C[rc + cc[j]] = y;
    }
  }
}
}  // namespace

namespace {
// ---------------------------------------------------------------------------
// Split-BF16 FP32 GEMM: C = A B with every FP32 operand written as
// hi + mid + lo, three bf16 numbers (round to nearest even, then the exact FP32
// residual, rounded again), so together they carry 24 significand bits.  The
// six products hi*hi, hi*mid, mid*hi, hi*lo, lo*hi, mid*mid (the ones above
// 2^-24 relative) run on v_mfma_f32_16x16x32_bf16 with FP32 accumulation; the
// result has FP32-class accuracy.  hi*hi goes to its own accumulator.
//
// MFMA operand layout (gfx950, checked with random integer inputs): lane l
// holds row l%16 of the left operand and column l%16 of the right one, over
// the eight consecutive k = 8*(l/16) .. +7; result register r is row
// 4*(l/16)+r, column l%16.  B is passed in the left slot, so the result is
// C^T per lane: lane l -> row l%16 of C, register r -> column 4*(l/16)+r.
//
// A(m, k) = A[k + m*lda] (KFA) or A[m + k*lda]; B(k, n) = B[k + n*ldb] (KFB)
// or B[n + k*ldb]; C column-major.  64x64 workgroup tile, 2x2 waves of 32x32,
// BK = 32, LDS rows of 40 bf16 (80 B: 16-byte aligned, no bank conflicts for
// the 16 rows of one 128-bit read).
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
constexpr int BF_BK = 32, BF_LD = 40;

// x = hi + mid + lo: the float -> bf16 casts compile to gfx950's
// v_cvt_pk_bf16_f32 (round to nearest even, two values per instruction), the
// bf16 -> float ones to a shift.
__device__ inline void split3(float x, unsigned short& h, unsigned short& m, unsigned short& l) {
  const __bf16 bh = (__bf16)x;
  const float r1 = x - (float)bh;
  const __bf16 bm = (__bf16)r1;
  const __bf16 bl = (__bf16)(r1 - (float)bm);
  memcpy(&h, &bh, 2);
  memcpy(&m, &bm, 2);
  memcpy(&l, &bl, 2);
}

template <bool KFA, bool KFB>
__global__ __launch_bounds__(256) void k_gemm_bf3(int M, int N, int K, const float* __restrict__ A,
                                                  long long lda, const float* __restrict__ B,
                                                  long long ldb, float* __restrict__ C,
                                                  long long ldc) {
  constexpr int BM = 64, BN = 64, NL = BM * BF_BK / 256;  // 8 elements of A and of B per thread
  __shared__ __attribute__((aligned(16))) unsigned short As[2][3][BM][BF_LD];
  __shared__ __attribute__((aligned(16))) unsigned short Bs[2][3][BN][BF_LD];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  float ra[NL], rb[NL];
  // k-fast operand: thread t holds the 8 consecutive k 8 (t%4) .. +7 of row
  // t/4 (two 16-byte loads; one 16-byte LDS store per bf16 plane).  m-fast
  // operand: row t%64 at k = t/64 + 4r, stored element by element.
  auto load_op = [&](bool kfast, const float* __restrict__ X, long long ld, int rows0, int nrows,
                     int k0, float* r) {
    if (kfast) {
      const int row = t >> 2, kq = (t & 3) * 8;
      const int g = rows0 + row, gk = k0 + kq;
      const float* src = X + gk + (long long)g * ld;
      if (g < nrows && gk + 8 <= K && (ld & 3) == 0 && ((uintptr_t)X & 15) == 0) {
        const float4 a = reinterpret_cast<const float4*>(src)[0];
        const float4 b = reinterpret_cast<const float4*>(src)[1];
        r[0] = a.x; r[1] = a.y; r[2] = a.z; r[3] = a.w;
        r[4] = b.x; r[5] = b.y; r[6] = b.z; r[7] = b.w;
      } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) r[q] = (g < nrows && gk + q < K) ? src[q] : 0.f;
      }
    } else {
#pragma unroll
      for (int q = 0; q < NL; ++q) {
        const int g = rows0 + (t & 63), gk = k0 + (t >> 6) + 4 * q;
        r[q] = (g < nrows && gk < K) ? X[g + (long long)gk * ld] : 0.f;
      }
    }
  };
  auto load = [&](int k0) {
    load_op(KFA, A, lda, m0, M, k0, ra);
    load_op(KFB, B, ldb, n0, N, k0, rb);
  };
  auto store_op = [&](bool kfast, unsigned short (*S)[BM][BF_LD], const float* r) {
    if (kfast) {
      const int row = t >> 2, kq = (t & 3) * 8;
      unsigned short h[8], mi[8], lo[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) split3(r[q], h[q], mi[q], lo[q]);
      uint4 v;
      memcpy(&v, h, 16);
      *reinterpret_cast<uint4*>(&S[0][row][kq]) = v;
      memcpy(&v, mi, 16);
      *reinterpret_cast<uint4*>(&S[1][row][kq]) = v;
      memcpy(&v, lo, 16);
      *reinterpret_cast<uint4*>(&S[2][row][kq]) = v;
    } else {
#pragma unroll
      for (int q = 0; q < NL; ++q) {
        const int row = t & 63, k = (t >> 6) + 4 * q;
        unsigned short h, mi, lo;
        split3(r[q], h, mi, lo);
        S[0][row][k] = h;
        S[1][row][k] = mi;
        S[2][row][k] = lo;
      }
    }
  };
  auto store = [&](int buf) {
    store_op(KFA, As[buf], ra);
    store_op(KFB, Bs[buf], rb);
  };
  f4_t acc_hh[2][2], acc_x[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc_hh[i][j] = acc_x[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};
  const int li = lane & 15, lg = lane >> 4;
  const int nk = (K + BF_BK - 1) / BF_BK;
  load(0);
  store(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) load((kt + 1) * BF_BK);
    bf8_t af[2][3], bfr[2][3];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        af[i][p] = *reinterpret_cast<const bf8_t*>(&As[buf][p][wm * 32 + i * 16 + li][lg * 8]);
        bfr[i][p] = *reinterpret_cast<const bf8_t*>(&Bs[buf][p][wn * 32 + i * 16 + li][lg * 8]);
      }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        acc_hh[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[j][0], af[i][0], acc_hh[i][j], 0, 0, 0);
        f4_t x = acc_x[i][j];
        x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[j][1], af[i][1], x, 0, 0, 0);
        x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[j][0], af[i][2], x, 0, 0, 0);
        x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[j][2], af[i][0], x, 0, 0, 0);
        x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[j][0], af[i][1], x, 0, 0, 0);
        x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[j][1], af[i][0], x, 0, 0, 0);
        acc_x[i][j] = x;
      }
    if (kt + 1 < nk) store(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm * 32 + i * 16 + li;
        const int gn = n0 + wn * 32 + j * 16 + lg * 4 + r;
        if (gm < M && gn < N) C[gm + (long long)gn * ldc] = acc_hh[i][j][r] + acc_x[i][j][r];
      }
}
}  // namespace

void gemm_f32_split_bf16(hipStream_t st, int opA, int opB, int M, int N, int K, const float* A,
                         int lda, const float* B, int ldb, float* C, int ldc) {
  DMRG_REQUIRE(M >= 0 && N >= 0 && K >= 0, "gemm_bf16x3: negative size");
  DMRG_REQUIRE(opA >= 0 && opA <= 3 && opB >= 0 && opB <= 3, "gemm_bf16x3: bad op");
  if (M == 0 || N == 0) return;
  // op(A) (m, k): transposed storage is k-fast; op(B) (k, n): plain storage is k-fast
  const bool kfa = op_trans(opA), kfb = !op_trans(opB);
  const dim3 grid(cdiv(M, 64), cdiv(N, 64));
#define DMRG_BF3(KA_, KB_) \
  hipLaunchKernelGGL((k_gemm_bf3<KA_, KB_>), grid, dim3(256), 0, st, M, N, K, A, (long long)lda, B, \
                     (long long)ldb, C, (long long)ldc)
  if (kfa && kfb) DMRG_BF3(true, true);
  else if (kfa) DMRG_BF3(true, false);
  else if (kfb) DMRG_BF3(false, true);
  else DMRG_BF3(false, false);
#undef DMRG_BF3
  DMRG_LAUNCH_CHECK();
}

namespace {
// C[i] = sum_{z<S} W[i + z*stride] for i < M*N over an M x N block (ld ldc);
// fixed summation order (deterministic).
// FP32 partials into an FP64 result (TO = double, T = float): the sum is
// formed in FP32 as for an FP32 result, then widened and multiplied by
// s1 * s2 (each taken as 1 when not positive), the factor that undoes the
// unit-max scaling of the FP32 operands.
template <class T, class TO>
__global__ void k_splitk_sum(const T* __restrict__ W, int S, long long stride, int M, int N,
                             TO* __restrict__ C, int ldc, const double* __restrict__ s1,
                             const double* __restrict__ s2) {
  double f = 1.0;
  if constexpr (!std::is_same<T, TO>::value) {
    const double a = *s1, b = *s2;
    f = (a > 0.0 ? a : 1.0) * (b > 0.0 ? b : 1.0);
  }
  const long long tot = (long long)M * N;
  const bool packed = ldc == M;  // C contiguous: no row/column split per element
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    T v = W[e];
    for (int z = 1; z < S; ++z) v = v + W[e + z * stride];
    const long long o = packed ? e : (e % M) + (e / M) * (long long)ldc;
    if constexpr (std::is_same<T, TO>::value) C[o] = v;
    else C[o] = (TO)v * f;
  }
}

template <class T, class TO>
bool splitk_impl(hipStream_t st, int M, int N, int K, const T* A, int lda, const T* B, int ldb,
                 TO* C, int ldc, T* ws, size_t ws_elems, const double* s1, const double* s2) {
  // S equal slices (S | K) in one batched launch, while the 64x64 grid
  // leaves CUs idle and each slice keeps K/S >= 512.  FP32 (BK = 32 tiles)
  // stops at two slices: measured 80 TF/s against 74 for four on the C4
  // out-step (2048 x 512 x 2560); FP64 (BK = 16) is best at four.
  const int smax = std::is_same<T, float>::value ? 2 : 8;
  const long long t64 = (long long)cdiv(M, 64) * cdiv(N, 64);
  int S = 1;
  while (S < smax && t64 * S * 2 <= 1024 && K % (S * 2) == 0 && K / (S * 2) >= 512 &&
         (size_t)(S * 2) * M * N <= ws_elems)
    S *= 2;
  if (S == 1) return false;
  const int Kc = K / S;
  // slice z: A(:, z Kc ..) (A + z Kc lda), B(z Kc .., :) (B + z Kc) -> ws + z M N
  // eight-wave tiles: +11 % (FP64) and +17 % (FP32) on the C4 out-step slices
  gemm_strided<T>(st, OP_N, OP_N, M, N, Kc, T(1), A, lda, (long long)Kc * lda, B, ldb,
                  (long long)Kc, T(0), ws, M, (long long)M * N, S, GemmTile{64, 64, 2});
  const long long tot = (long long)M * N;
  hipLaunchKernelGGL((k_splitk_sum<T, TO>),
                     dim3((unsigned)std::min<long long>(cdiv(tot, 256), 4096)), dim3(256), 0, st,
                     (const T*)ws, S, (long long)M * N, M, N, C, ldc, s1, s2);
  DMRG_LAUNCH_CHECK();
  return true;
}
}  // namespace

template <class T>
bool gemm_nn_splitk(hipStream_t st, int M, int N, int K, const T* A, int lda, const T* B, int ldb,
                    T* C, int ldc, T* ws, size_t ws_elems) {
  return splitk_impl<T, T>(st, M, N, K, A, lda, B, ldb, C, ldc, ws, ws_elems, nullptr, nullptr);
}

bool gemm_nn_splitk_to_f64(hipStream_t st, int M, int N, int K, const float* A, int lda,
                           const float* B, int ldb, double* C, int ldc, float* ws,
                           size_t ws_elems, const double* s1, const double* s2) {
  return splitk_impl<float, double>(st, M, N, K, A, lda, B, ldb, C, ldc, ws, ws_elems, s1, s2);
}

namespace {
// First step of the two-site H_eff with the MPO-bond step applied in the
// epilogue (d = 2, real scalars).  The workgroup computes the tile of
// X = L^T theta with rows (w + D a') for 16 consecutive a' (all D values of w)
// and columns (ss + 4 b) for 16 consecutive b (all four ss): wave j owns the
// columns 16 j .. 16 j + 15 and all D row blocks, one 16x16 MFMA block each
// per 4-deep k step, with the k-loop of k_gemm (two register stages, two LDS
// buffers).  The finished tile goes to LDS (over the operand buffers, which
// the k-loop no longer reads), and thread (a', b) contracts its 4D values
// with WW into the 4D outputs Y[a' + cL ss' + 4 cL b + 4 cL cR n], summing
// over c = w + D ss in the order of k_rows_small.  X never reaches HBM.
// theta may be stored in another precision (TB): each element is converted
// to T on load.
template <class T, class TB, int D, int BK>
__global__ __launch_bounds__(256) void k_heff2_x_mpo(int cL, int cR, const T* __restrict__ Lenv,
                                                     const TB* __restrict__ theta,
                                                     const T* __restrict__ WW,
                                                     T* __restrict__ Y) {
  using V4 = typename GT<T>::V4;
  constexpr int DD = 4, BM = 16 * D, BN = 16 * DD, K2 = D * DD;
  constexpr int NA = BM * BK / 256, NB = BN * BK / 256;
  static_assert(NA * 256 == BM * BK && NB * 256 == BN * BK, "tile");
  constexpr int OPS = 2 * BK * (BM + 1) + 2 * BK * (BN + 1), XS = BM * (BN + 1);
  __shared__ T smem[OPS > XS ? OPS : XS];
  __shared__ T Ws[K2 * K2];
  T(*As)[BK][BM + 1] = reinterpret_cast<T(*)[BK][BM + 1]>(smem);
  T(*Bs)[BK][BN + 1] = reinterpret_cast<T(*)[BK][BN + 1]>(smem + 2 * BK * (BM + 1));
  T(*Xs)[BN + 1] = reinterpret_cast<T(*)[BN + 1]>(smem);
  const int t = threadIdx.x, lane = t & 63, wn = t >> 6;
  const int M = D * cL, N = DD * cR, K = cL;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  for (int e = t; e < K2 * K2; e += 256) Ws[e] = WW[e];
  // A(m, k) = Lenv[k + m cL], B(k, n) = theta[k + n cL]: both k-fast
  T ra[2][NA], rb[2][NB];
  const bool inner = m0 + BM <= M && n0 + BN <= N;
  auto load = [&](int s, int k0) {
    const bool full = inner && k0 + BK <= K;
#pragma unroll
    for (int r = 0; r < NA; ++r) {
      const int k = t % BK, m = t / BK + r * (256 / BK);
      const int gm = m0 + m, gk = k0 + k;
      ra[s][r] = (full || (gm < M && gk < K)) ? Lenv[gk + (size_t)gm * cL] : T(0);
    }
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      const int k = t % BK, n = t / BK + r * (256 / BK);
      const int gn = n0 + n, gk = k0 + k;
      rb[s][r] = (full || (gn < N && gk < K)) ? (T)theta[gk + (size_t)gn * cL] : T(0);
    }
  };
  auto store = [&](int s, int buf) {
#pragma unroll
    for (int r = 0; r < NA; ++r) As[buf][t % BK][t / BK + r * (256 / BK)] = ra[s][r];
#pragma unroll
    for (int r = 0; r < NB; ++r) Bs[buf][t % BK][t / BK + r * (256 / BK)] = rb[s][r];
  };
  V4 acc[D];
#pragma unroll
  for (int i = 0; i < D; ++i) acc[i] = V4{0, 0, 0, 0};
  const int li = lane & 15, lk = lane >> 4;
  auto compute = [&](int buf) {
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T af[D];
#pragma unroll
      for (int i = 0; i < D; ++i) af[i] = As[buf][kk + lk][i * 16 + li];
      const T bf = Bs[buf][kk + lk][wn * 16 + li];
#pragma unroll
      for (int i = 0; i < D; ++i) acc[i] = mfma(bf, af[i], acc[i]);
    }
  };
  const int nk = (K + BK - 1) / BK;
  load(0, 0);
  store(0, 0);
  if (nk > 1) load(1, BK);
  __syncthreads();
  for (int kt = 0; kt < nk; kt += 2) {
    if (kt + 2 < nk) load(0, (kt + 2) * BK);
    compute(0);
    if (kt + 1 < nk) store(1, 1);
    __syncthreads();
    if (kt + 1 >= nk) break;
    if (kt + 3 < nk) load(1, (kt + 3) * BK);
    compute(1);
    if (kt + 2 < nk) store(0, 0);
    __syncthreads();
  }
  // every exit of the loop passed a barrier: the operand buffers are free
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r)
      Xs[i * 16 + li][wn * 16 + (GT<T>::G1 ? lk + 4 * r : 4 * lk + r)] = acc[i][r];
  __syncthreads();
  const int al = t & 15, bl = t >> 4;
  const int a = blockIdx.x * 16 + al, b = blockIdx.y * 16 + bl;
  if (a >= cL || b >= cR) return;
  T x[K2];
#pragma unroll
  for (int w = 0; w < D; ++w)
#pragma unroll
    for (int ss = 0; ss < DD; ++ss) x[w + D * ss] = Xs[w + D * al][ss + DD * bl];
  T* const yb = Y + a + (size_t)cL * DD * b;
  const size_t sn = (size_t)cL * DD * cR;
#pragma unroll
  for (int n = 0; n < D; ++n)
#pragma unroll
    for (int sp = 0; sp < DD; ++sp) {
      const int c2 = sp + DD * n;
      T y = T(0);
#pragma unroll
      for (int c = 0; c < K2; ++c) y = y + x[c] * Ws[c + K2 * c2];
      yb[(size_t)cL * sp + sn * n] = y;
    }
}
}  // namespace

template <class T, class TB>
bool heff2_x_mpo(hipStream_t st, int cL, int cR, int d, int D, const T* Lenv, const TB* theta,
                 const T* WW, T* Y) {
  if constexpr (is_cplx<T>::value || is_cplx<TB>::value) {
    return false;
  } else {
    if (d != 2 || D < 2 || D > 8 || cL < 1 || cR < 1) return false;
    const dim3 grid(cdiv(cL, 16), cdiv(cR, 16));
    switch (D) {
#define DMRG_XM(D_) \
  case D_: hipLaunchKernelGGL((k_heff2_x_mpo<T, TB, D_, 16>), grid, dim3(256), 0, st, cL, cR, Lenv, theta, WW, Y); break
      DMRG_XM(2); DMRG_XM(3); DMRG_XM(4); DMRG_XM(5); DMRG_XM(6); DMRG_XM(7); DMRG_XM(8);
#undef DMRG_XM
      default: return false;
    }
    DMRG_LAUNCH_CHECK();
    return true;
  }
}

template <class T>
bool gemm_rows_small(hipStream_t st, int M, int N, int K, const T* A, View2 va, const T* B, T* C,
                     View2 vc) {
  if (K > 32 || N > 32 || K < 1) return false;
  if (M == 0 || N == 0) return true;
  const unsigned grid = (unsigned)std::min<long long>(cdiv(M, 256), 8192);
  if (K <= 16)
    hipLaunchKernelGGL((k_rows_small<T, 16>), dim3(grid), dim3(256), 0, st, M, N, K, A, va, B, C, vc);
  else if (K <= 24)
    hipLaunchKernelGGL((k_rows_small<T, 24>), dim3(grid), dim3(256), 0, st, M, N, K, A, va, B, C, vc);
  else
    hipLaunchKernelGGL((k_rows_small<T, 32>), dim3(grid), dim3(256), 0, st, M, N, K, A, va, B, C, vc);
  DMRG_LAUNCH_CHECK();
  return true;
}

template <class T>
void gemm(hipStream_t st, int opA, int opB, int M, int N, int K, T alpha, const T* A, int lda,
          const T* B, int ldb, T beta, T* C, int ldc, GemmTile tile) {
  plain<T>(st, opA, opB, tile, M, N, K, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, 1);
}

template <class T>
void gemm_herm(hipStream_t st, int opA, int opB, int N, int K, T alpha, const T* A, int lda,
               const T* B, int ldb, T* C, int ldc, GemmTile tile) {
  DMRG_REQUIRE(tile.bm == tile.bn, "gemm_herm: square tiles only");
  plain<T>(st, opA, opB, tile, N, N, K, alpha, A, lda, 0, B, ldb, 0, T(0), C, ldc, 0, 1, 1);
}

template <class T>
void gemm_strided(hipStream_t st, int opA, int opB, int M, int N, int K, T alpha, const T* A,
                  int lda, long long sA, const T* B, int ldb, long long sB, T beta, T* C, int ldc,
                  long long sC, int batch, GemmTile tile) {
  plain<T>(st, opA, opB, tile, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
}

template <class T>
void gemm_tensor(hipStream_t st, int M, int N, int K, T alpha, const T* A, View2 va, bool conjA,
                 bool a_kfast, const T* B, View2 vb, bool conjB, bool b_kfast, T beta, T* C,
                 View2 vc, GemmTile tile) {
  DMRG_REQUIRE(M >= 0 && N >= 0 && K >= 0, "gemm_tensor: negative size");
  if (M == 0 || N == 0) return;
  if (!is_cplx<T>::value) conjA = conjB = false;
  AddrTensor aa{va}, ab{vb}, ac{vc};
#define DMRG_GT(CA_, CB_, KA_, KB_) \
  launch<T, CA_, CB_, KA_, KB_>(st, tile, M, N, K, alpha, A, aa, 0, B, ab, 0, beta, C, ac, 0, 1)
  int code = (conjA ? 8 : 0) | (conjB ? 4 : 0) | (a_kfast ? 2 : 0) | (b_kfast ? 1 : 0);
  if constexpr (is_cplx<T>::value) {
    switch (code) {
      case 0: DMRG_GT(false, false, false, false); break;
      case 1: DMRG_GT(false, false, false, true); break;
      case 2: DMRG_GT(false, false, true, false); break;
      case 3: DMRG_GT(false, false, true, true); break;
      case 4: DMRG_GT(false, true, false, false); break;
      case 5: DMRG_GT(false, true, false, true); break;
      case 6: DMRG_GT(false, true, true, false); break;
      case 7: DMRG_GT(false, true, true, true); break;
      case 8: DMRG_GT(true, false, false, false); break;
      case 9: DMRG_GT(true, false, false, true); break;
      case 10: DMRG_GT(true, false, true, false); break;
      case 11: DMRG_GT(true, false, true, true); break;
      case 12: DMRG_GT(true, true, false, false); break;
      case 13: DMRG_GT(true, true, false, true); break;
      case 14: DMRG_GT(true, true, true, false); break;
      default: DMRG_GT(true, true, true, true); break;
    }
  } else {
    switch (code & 3) {
      case 0: DMRG_GT(false, false, false, false); break;
      case 1: DMRG_GT(false, false, false, true); break;
      case 2: DMRG_GT(false, false, true, false); break;
      default: DMRG_GT(false, false, true, true); break;
    }
  }
#undef DMRG_GT
}

#define DMRG_INST(T)                                                                         \
  template void gemm<T>(hipStream_t, int, int, int, int, int, T, const T*, int, const T*, int, \
                        T, T*, int, GemmTile);                                                 \
  template void gemm_herm<T>(hipStream_t, int, int, int, int, T, const T*, int, const T*, int,  \
                             T*, int, GemmTile);                                                \
  template void gemm_strided<T>(hipStream_t, int, int, int, int, int, T, const T*, int,        \
                                long long, const T*, int, long long, T, T*, int, long long,    \
                                int, GemmTile);                                                \
  template void gemm_tensor<T>(hipStream_t, int, int, int, T, const T*, View2, bool, bool,     \
                               const T*, View2, bool, bool, T, T*, View2, GemmTile);           \
  template bool gemm_rows_small<T>(hipStream_t, int, int, int, const T*, View2, const T*, T*,  \
                                   View2);                                                      \
  template bool gemm_nn_splitk<T>(hipStream_t, int, int, int, const T*, int, const T*, int, T*, \
                                  int, T*, size_t);                                             \
  template bool heff2_x_mpo<T, T>(hipStream_t, int, int, int, int, const T*, const T*, const T*, \
                                  T*);
DMRG_INST(double)
DMRG_INST(cdouble)
DMRG_INST(float)
#undef DMRG_INST


}  // namespace dmrg
