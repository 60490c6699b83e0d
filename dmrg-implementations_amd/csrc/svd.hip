// Density-matrix truncated SVD and block-Jacobi Hermitian eigen-solver
// (see svd.hpp).
#include "svd.hpp"
#include "contract.hpp"
#include "lanczos.hpp"
#include "gemm.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace dmrg {
namespace {

// Row and column of flat index e of a column-major matrix with n rows, in
// 32-bit arithmetic (every matrix these kernels index has at most 2048^2
// entries; a 64-bit division is a long software sequence on the GPU).
__device__ inline int mod32(size_t e, int n) { return (int)((unsigned)e % (unsigned)n); }
__device__ inline int div32(size_t e, int n) { return (int)((unsigned)e / (unsigned)n); }

constexpr int JB = 16;       // Jacobi block width
constexpr int JP = 2 * JB;   // order of one pair sub-problem
constexpr int LDP = JP + 1;  // LDS column stride (padding)
constexpr int NT = 256;      // threads of the pair kernels
constexpr int SORT_MAX = 2048;

typedef double d4_t __attribute__((ext_vector_type(4)));

// Round-robin tournament: block pair k of round r among nb blocks (nb even).
__device__ __host__ inline void rr_pair(int nb, int r, int k, int& a, int& b) {
  const int n1 = nb - 1;
  if (k == 0) { a = r; b = n1; }
  else { a = (r + k) % n1; b = (r - k + n1) % n1; }
}
// Block pair k of round r under schedule ka.  ka == 0: round robin over all
// nb blocks (nb-1 rounds cover every pair).  ka > 0: the first ka blocks
// (the kept group after the sort) against the other nb-ka ("bipartite"; only
// couplings across the two groups are rotated): with a = min(ka, nb-ka) on the
// smaller side S and b = max on the larger side G, pair k < a is
// (S_k, G_{(k+r) mod b}), so b rounds meet every S-G pair once; pairs k >= a
// join the G blocks left over in round r two by two and are not rotated
// (pair_is_idle).  Every round is a perfect matching of the nb blocks.
__device__ __host__ inline void sched_pair(int nb, int ka, int r, int k, int& I, int& J) {
  if (ka == 0) { rr_pair(nb, r, k, I, J); return; }
  const int kb = nb - ka;
  const int a = ka < kb ? ka : kb, b = nb - a;
  const int s0 = ka < kb ? 0 : ka, g0 = ka < kb ? ka : 0;  // first block of S and of G
  if (k < a) {
    I = s0 + k;
    J = g0 + (k + r) % b;
  } else {
    const int q = k - a;
    I = g0 + (a + 2 * q + r) % b;
    J = g0 + (a + 2 * q + 1 + r) % b;
  }
}
__device__ __host__ inline bool pair_is_idle(int nb, int ka, int k) {
  if (ka == 0) return false;
  const int kb = nb - ka;
  return k >= (ka < kb ? ka : kb);
}
// Rounds of one sweep under schedule ka.
inline int sched_rounds(int nb, int ka) {
  if (ka == 0) return nb - 1;
  return std::max(ka, nb - ka);
}
// Global row/column of local index i (0..JP-1) of the pair (I, J).
__device__ inline int gidx(int I, int J, int i) { return i < JB ? I * JB + i : J * JB + (i - JB); }

__device__ inline double absv(double a) { return fabs(a); }
__device__ inline double absv(cdouble a) { return sqrt(a.x * a.x + a.y * a.y); }

template <class T> struct NP { static constexpr int P = 1; };
template <> struct NP<cdouble> { static constexpr int P = 2; };

__device__ inline void put(double* pl, int idx, int plane_stride, double v) { pl[idx] = v; }
__device__ inline void put(double* pl, int idx, int plane_stride, cdouble v) {
  pl[idx] = v.x;
  pl[idx + plane_stride] = v.y;
}
template <class T> __device__ inline T get(const double* pl, int idx, int plane_stride);
template <> __device__ inline double get<double>(const double* pl, int idx, int) { return pl[idx]; }
template <> __device__ inline cdouble get<cdouble>(const double* pl, int idx, int ps) {
  return cdouble(pl[idx], pl[idx + ps]);
}

// ---------------------------------------------------------------------------
// Set-up: symmetrise the N x N block of rho (lower := conj(upper), real
// diagonal), zero the padding up to Np, V = I.
template <class T>
__global__ void k_prep(T* __restrict__ A, T* __restrict__ V, int N, int Np, int set_v) {
  const size_t tot = (size_t)Np * Np;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int r = mod32(e, Np), c = div32(e, Np);
    if (r >= N || c >= N) A[e] = T(0);
    else if (r > c) A[e] = conj(A[c + (size_t)r * Np]);
    else if (r == c) A[e] = T(real(A[e]));
    if (set_v) V[e] = (r == c) ? T(1) : T(0);
  }
}

// V (Np x Np) = the caller's N x N basis X (ld N), identity on the padding.
template <class T>
__global__ void k_set_basis(T* __restrict__ V, const T* __restrict__ X, int N, int Np) {
  const size_t tot = (size_t)Np * Np;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int r = mod32(e, Np), c = div32(e, Np);
    V[e] = (r < N && c < N) ? X[r + (size_t)c * N] : ((r == c) ? T(1) : T(0));
  }
}

// Partial sums of |rho|_F^2 and trace(rho): red[blk] and red[NRED + blk].
constexpr int NRED = 256;  // max workgroups of k_norms_part
template <class T>
__global__ __launch_bounds__(256) void k_norms_part(const T* __restrict__ A, int N, int Np,
                                                    double* __restrict__ red) {
  __shared__ double s1[256], s2[256];
  const int t = threadIdx.x;
  double f = 0.0, tr = 0.0;
  const size_t tot = (size_t)Np * Np;
  for (size_t e = blockIdx.x * (size_t)256 + t; e < tot; e += (size_t)gridDim.x * 256)
    f += abs2(A[e]);
  for (int i = blockIdx.x * 256 + t; i < N; i += gridDim.x * 256) tr += real(A[i + (size_t)i * Np]);
  s1[t] = f;
  s2[t] = tr;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) { s1[t] += s1[t + o]; s2[t] += s2[t + o]; }
    __syncthreads();
  }
  if (t == 0) { red[blockIdx.x] = s1[0]; red[NRED + blockIdx.x] = s2[0]; }
}
// scal[0] = |rho|_F, scal[2] = trace(rho) from nb partials (one workgroup,
// fixed order).
__global__ __launch_bounds__(256) void k_norms_fin(const double* __restrict__ red, int nb,
                                                   double* scal) {
  __shared__ double s1[256], s2[256];
  const int t = threadIdx.x;
  s1[t] = t < nb ? red[t] : 0.0;
  s2[t] = t < nb ? red[NRED + t] : 0.0;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) { s1[t] += s1[t + o]; s2[t] += s2[t + o]; }
    __syncthreads();
  }
  if (t == 0) { scal[0] = sqrt(s1[0]); scal[2] = s2[0]; }
}

// rank[i] = position of diagonal entry i in the stable descending order of
// k_sort (ties by index, NaN as -inf); padding i >= N gets rank i.
// keys[0] = largest entry, keys[1] = entry of rank keep-1.
// Sixteen lanes per entry (all with the same i): lane q counts the entries
// j = q, q + 16, ...; the group sums by shuffles.  Grid: cdiv(Np, 16).
template <class T>
__global__ __launch_bounds__(256) void k_rank_diag(const T* __restrict__ A, int N, int Np,
                                                   int keep, int* __restrict__ rank,
                                                   double* __restrict__ keys) {
  __shared__ double dg[SORT_MAX];
  const int t = threadIdx.x;
  for (int i = t; i < N; i += 256) {
    double v = real(A[i + (size_t)i * Np]);
    dg[i] = (v == v) ? v : -INFINITY;
  }
  __syncthreads();
  const int q = t & 15;
  for (int i = blockIdx.x * 16 + (t >> 4); i < Np; i += gridDim.x * 16) {
    int r = 0;
    double di = 0.0;
    if (i < N) {
      di = dg[i];
#pragma unroll 8
      for (int j = q; j < N; j += 16) {
        const double dj = dg[j];
        r += (dj > di) || (dj == di && j < i);
      }
    }
    r += __shfl_xor(r, 1);
    r += __shfl_xor(r, 2);
    r += __shfl_xor(r, 4);
    r += __shfl_xor(r, 8);
    if (q == 0) {
      if (i >= N) {
        rank[i] = i;
      } else {
        rank[i] = r;
        if (r == 0) keys[0] = di;
        if (r == keep - 1) keys[1] = di;
      }
    }
  }
}

// Stable descending sort of n <= 2048 keys -> perm (and the sorted keys).
// Key i is the real diagonal A(i, i) (ld Np) when A is given, else keys[i];
// positions i >= n_real (zero padding of the Jacobi matrix) get -inf, so the
// padding always sorts to the end and real indices keep positions < n_real.
// By ranks: the position of key i is the number of keys before it in the
// order (larger, or equal with a smaller index; NaN as -inf), and
// perm[position] = i.  Sixteen lanes per key count j = q, q + 16, ... and sum
// by shuffles; grid cdiv(n, 16) workgroups of 256 threads.
template <class T>
__global__ __launch_bounds__(256) void k_sort(const T* __restrict__ A, int Np,
                                              const double* __restrict__ keys, int n, int n_real,
                                              int* __restrict__ perm, double* __restrict__ keys_out) {
  __shared__ double key[SORT_MAX];
  const int t = threadIdx.x;
  for (int i = t; i < n; i += 256) {
    double v = -INFINITY;
    if (i < n_real) v = A ? real(A[i + (size_t)i * Np]) : keys[i];
    key[i] = (v == v) ? v : -INFINITY;  // keeps every output index below n
  }
  __syncthreads();
  const int q = t & 15;
  const int i = blockIdx.x * 16 + (t >> 4);
  if (i >= n) return;  // whole groups of sixteen lanes leave together
  const double ki = key[i];
  int r = 0;
#pragma unroll 8
  for (int j = q; j < n; j += 16) {
    const double kj = key[j];
    r += (kj > ki) || (kj == ki && j < i);
  }
  r += __shfl_xor(r, 1);
  r += __shfl_xor(r, 2);
  r += __shfl_xor(r, 4);
  r += __shfl_xor(r, 8);
  if (q == 0) {
    perm[r] = i;
    if (keys_out) keys_out[r] = ki;
  }
}

// B(i, j) = A(p_i, p_j); W(:, j) = V(:, p_j).
template <class T>
__global__ void k_permute(const T* __restrict__ A, const T* __restrict__ V, T* __restrict__ B,
                          T* __restrict__ W, const int* __restrict__ perm, int Np) {
  const size_t tot = (size_t)Np * Np;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int r = mod32(e, Np), c = div32(e, Np);
    const int pc = perm[c];
    B[e] = A[perm[r] + (size_t)pc * Np];
    W[e] = V[r + (size_t)pc * Np];
  }
}

// ---------------------------------------------------------------------------
// Kernel A: one workgroup per block pair of round `round`.  Loads the JP x JP
// Hermitian sub-matrix, runs up to `inner_max` cyclic Jacobi sweeps on it in
// LDS (skipping pairs with |a_ab| <= rot_tol * |rho|_F), writes the rotated
// sub-matrix back and the accumulated unitary to Qout (JP x JP, ld JP).
//
// Rotation of the pair (a, b): with beta = a_ab = |beta| e^{i phi},
// tau = (a_bb - a_aa) / (2 |beta|), t = sign(tau) / (|tau| + sqrt(1 + tau^2)),
// c = 1/sqrt(1+t^2), s = t c, J = [[c, s e^{i phi}], [-s e^{-i phi}, c]];
// J^H [[a_aa, beta], [conj(beta), a_bb]] J = diag(a_aa - t|beta|, a_bb + t|beta|)
// (the symmetric Schur decomposition with a phase, Golub & Van Loan 8.4).
// Evaluated as u = |a_bb - a_aa|, r = sqrt(u^2 + 4|beta|^2),
// y = 1/sqrt(2 r (u + r)): c = (u + r) y, s e^{i phi} = 2 sign(tau) beta y,
// t|beta| = 4 sign(tau) |beta|^2 r y^2 (t = 2 sign(tau) |beta| / (u + r) and
// 1/(u + r) = 2 r y^2).  Every term is a sum of non-negative numbers, so there
// is no cancellation, also for |beta| << u, and c^2 + |s|^2 = 1 to the
// accuracy of y.
//
// Inner schedule: the circle method over the JP local indices.  Round g pairs
// positions (p, JP-1-p), p < HP = JP/2; the index at each position advances
// by adv() from round to round (position == index at round 0), so the pair p
// of round g+1 joins the indices at positions adv(p) and adv(JP-1-p) of
// round g, which lie in two different pairs of round g.
//
// NTA threads (4 waves) per sub-problem, one barrier per round.  Threads
// t < NBLK each own one unordered pair of pairs {k1 < k2} of the round and
// replace the 2x2 block rho(k1 rows, k2 cols) by J1^H X J2 (and the mirror
// block by its conjugate transpose, which keeps the sub-matrix exactly
// Hermitian).  The element that pairs the next round's pair p is in exactly
// one such block (a fixed block for each p, whatever the round): its owner
// computes the rotation of round g+1 from the updated value still in
// registers, writes it into the other half of the double-buffered (c, s)
// table, zeroes the element and its mirror and moves the two diagonal entries.
// No other thread touches those entries in that round.  Threads QT0 and up
// (waves 2-3) apply the round's rotations to the columns of Q.  A pair with |a_ab| <= thr gets (c, s) = (1, 0), which leaves every
// entry bit-for-bit unchanged, so the updates need no flags.  A sub-problem
// whose off-diagonal is already below 1e-6 |rho|_F gets one inner sweep (the
// next outer sweep re-checks it), others up to inner_max; a sweep without a
// rotation ends the loop.
constexpr int NTA = 256;
constexpr int HP = JP / 2;                  // pairs per inner round
constexpr int NBLK = HP * (HP - 1) / 2;     // off-diagonal 2x2 blocks per round
constexpr int QT0 = 128;                    // first thread of the Q updates
constexpr int NQP = JP * HP / (NTA - QT0);  // Q column-pair updates per Q thread
static_assert(NBLK <= QT0 && (JP * HP) % (NTA - QT0) == 0 && (NTA - QT0) % JP == 0,
              "thread mapping of the pair kernel");

__device__ inline int adv_pos(int i) { return i == 0 ? 0 : (i == JP - 1 ? 1 : i + 1); }
__device__ inline int pair_of(int pos) { return pos < HP ? pos : JP - 1 - pos; }

// 1/sqrt(x) for positive normal x: the hardware estimate refined by one
// third-order step, e = 1 - x y^2, y <- y (1 + e/2 + 3 e^2/8).
__device__ inline double rsqrt_refined(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}

// Rotation of one pair (formulas above); false (and the identity) when
// |beta|^2 <= thr2.
template <class T>
__device__ inline bool pair_rotation(double al, double ga, T be, double thr2, double& c, T& sp,
                                     double& tb) {
  const double m2 = abs2(be);
  c = 1.0;
  sp = T(0);
  tb = 0.0;
  if (!(m2 > thr2)) return false;
  const double dl = ga - al, u = fabs(dl), sg = dl >= 0.0 ? 1.0 : -1.0;
  const double r2 = fma(dl, dl, 4.0 * m2);
  if (r2 > 0x1p-900 && r2 < 0x1p+900) {
    const double r = r2 * rsqrt_refined(r2);
    const double y = rsqrt_refined(2.0 * r * (u + r));
    c = (u + r) * y;
    sp = be * (2.0 * sg * y);
    tb = 4.0 * sg * m2 * r * y * y;
  } else {  // far out of the usual range: library square root and division
    const double w = 1.0 / (u + sqrt(r2));
    tb = 2.0 * sg * m2 * w;
    c = 1.0 / sqrt(1.0 + 4.0 * m2 * w * w);
    sp = be * (2.0 * sg * c * w);
  }
  return true;
}

template <class T>
__global__ __launch_bounds__(NTA) void k_pair_eig(T* __restrict__ A, int ld, int nb, int ka,
                                                  int round, T* __restrict__ Qout,
                                                  int* __restrict__ qflag,
                                                  const double* __restrict__ scal, double rot_tol,
                                                  int inner_max) {
  __shared__ T As[JP * LDP];  // As[c * LDP + r] = sub(r, c)
  __shared__ T Qs[JP * LDP];
  __shared__ double rc[2][HP];
  __shared__ T rs[2][HP];
  __shared__ int s_rot[2];  // a rotation happened in an even / odd inner sweep
  const int k = blockIdx.x, t = threadIdx.x;
  if (pair_is_idle(nb, ka, k)) {
    if (t == 0) qflag[k] = 0;
    return;
  }
  int I, J;
  sched_pair(nb, ka, round, k, I, J);
  const double fro = scal[0];
  const double thr = rot_tol * fro;
  const double thr2 = thr * thr;
  const double near2 = 1e-12 * fro * fro;
  int big = 0, far = 0;
  for (int e = t; e < JP * JP; e += NTA) {
    const int r = e % JP, c = e / JP;
    const T v = A[(size_t)gidx(I, J, r) + (size_t)gidx(I, J, c) * ld];
    As[c * LDP + r] = v;
    Qs[c * LDP + r] = (r == c) ? T(1) : T(0);
    if (r != c) {
      const double a2v = abs2(v);
      big |= a2v > thr2;
      far |= a2v > near2;
    }
  }
  if (t < 2) s_rot[t] = 0;
  if (!__syncthreads_or(big)) {
    if (t == 0) qflag[k] = 0;
    return;
  }
  const int sweeps = __syncthreads_or(far) ? inner_max : 1;
  constexpr int RPS = JP - 1;  // rounds per inner sweep
  const int total = sweeps * RPS;

  // Block threads: rows from pair kr, columns from pair kc (k1 < k2 of the
  // thread's pair of pairs).  The owner of the next-round pair `own` orients
  // its block so that the element of that pair is the (0, 0) entry: rows
  // (r0, r1) with r0 the index of the pair's first position, columns
  // (c0, c1) with c0 the second; a pair taken in the order (b, a) has the
  // rotation [[c, -conj(s)], [s, c]] (fr, fc).  r0.. c1 are the indices at
  // those positions, advanced with the schedule.
  int kr = 0, kc = 0, own = -1;
  bool fr = false, fc = false;
  if (t < NBLK) {
    int k1 = 0, rem = t;
    while (rem >= HP - 1 - k1) { rem -= HP - 1 - k1; ++k1; }
    kr = k1;
    kc = k1 + 1 + rem;
    const int k2 = kc;
    for (int p = 0; p < HP; ++p) {
      const int u = adv_pos(p), v = adv_pos(JP - 1 - p);
      const int pu = pair_of(u), pv = pair_of(v);
      if ((pu == k1 && pv == k2) || (pu == k2 && pv == k1)) {
        own = p;
        kr = pu;
        kc = pv;
        fr = u != pu;
        fc = v != pv;
      }
    }
  }
  int r0 = fr ? JP - 1 - kr : kr, r1 = fr ? kr : JP - 1 - kr;
  int c0 = fc ? JP - 1 - kc : kc, c1 = fc ? kc : JP - 1 - kc;
  // Q threads: row qrow of Q, column pairs kq (indices aq[q], bq[q]).
  const int tq = t - QT0;
  const int qrow = tq >= 0 ? tq % JP : 0;
  int aq[NQP], bq[NQP];
#pragma unroll
  for (int q = 0; q < NQP; ++q) {
    const int kk = (tq >= 0 ? tq / JP : 0) + ((NTA - QT0) / JP) * q;
    aq[q] = kk;
    bq[q] = JP - 1 - kk;
  }

  // Rotations of round 0.
  if (t < HP) {
    const int a = t, b = JP - 1 - t;
    const double al = real(As[a * LDP + a]), ga = real(As[b * LDP + b]);
    double c, tb;
    T sp;
    if (pair_rotation(al, ga, As[b * LDP + a], thr2, c, sp, tb)) {
      As[a * LDP + a] = T(al - tb);
      As[b * LDP + b] = T(ga + tb);
      As[b * LDP + a] = T(0);
      As[a * LDP + b] = T(0);
      s_rot[0] = 1;
    }
    rc[0][t] = c;
    rs[0][t] = sp;
  }
  __syncthreads();

  int any = 0, rin = 0, sw = 0;  // round within the sweep, sweep
  for (int g = 0; g < total; ++g) {
    const int cb = g & 1;
    if (t < NBLK) {
      // every load of the round first
      const double cr = rc[cb][kr], cc = rc[cb][kc];
      T sr = rs[cb][kr], sc = rs[cb][kc];
      const T x00 = As[c0 * LDP + r0], x01 = As[c1 * LDP + r0];
      const T x10 = As[c0 * LDP + r1], x11 = As[c1 * LDP + r1];
      const double al = real(As[r0 * LDP + r0]), ga = real(As[c0 * LDP + c0]);
      if (fr) sr = -conj(sr);
      if (fc) sc = -conj(sc);
      // O = Jr^H X Jc
      const T y00 = cc * x00 - conj(sc) * x01, y01 = sc * x00 + cc * x01;
      const T y10 = cc * x10 - conj(sc) * x11, y11 = sc * x10 + cc * x11;
      T o00 = cr * y00 - sr * y10;
      const T o01 = cr * y01 - sr * y11;
      const T o10 = conj(sr) * y00 + cr * y10, o11 = conj(sr) * y01 + cr * y11;
      // Rotation of the next round from o00, evaluated by every block lane
      // on safe stand-in values where it is not needed (no divergence); the
      // owner keeps it.
      const bool live = own >= 0 && g + 1 < total;
      const double m2 = abs2(o00);
      const bool rot = live && m2 > thr2;
      double c = 1.0, tb = 0.0;
      T sp = T(0);
      {
        const double dl = rot ? ga - al : 0.0, m2s = rot ? m2 : 1.0;
        const double u = fabs(dl), sg = dl >= 0.0 ? 1.0 : -1.0;
        const double r2 = fma(dl, dl, 4.0 * m2s);
        double cw, yw, tw;
        if (r2 > 0x1p-900 && r2 < 0x1p+900) {
          const double r = r2 * rsqrt_refined(r2);
          const double y = rsqrt_refined(2.0 * r * (u + r));
          cw = (u + r) * y;
          yw = 2.0 * sg * y;
          tw = 4.0 * sg * m2s * r * y * y;
        } else {  // far out of the usual range: library square root and division
          const double w = 1.0 / (u + sqrt(r2));
          tw = 2.0 * sg * m2s * w;
          cw = 1.0 / sqrt(1.0 + 4.0 * m2s * w * w);
          yw = 2.0 * sg * cw * w;
        }
        if (rot) {
          c = cw;
          sp = o00 * yw;
          tb = tw;
          o00 = T(0);
        }
      }
      if (live) {
        As[r0 * LDP + r0] = T(al - tb);
        As[c0 * LDP + c0] = T(ga + tb);
        rc[cb ^ 1][own] = c;
        rs[cb ^ 1][own] = sp;
        if (rot) s_rot[(sw + (rin + 1 == RPS)) & 1] = 1;
      }
      As[c0 * LDP + r0] = o00;
      As[c1 * LDP + r0] = o01;
      As[c0 * LDP + r1] = o10;
      As[c1 * LDP + r1] = o11;
      As[r0 * LDP + c0] = conj(o00);
      As[r0 * LDP + c1] = conj(o01);
      As[r1 * LDP + c0] = conj(o10);
      As[r1 * LDP + c1] = conj(o11);
    } else if (tq >= 0) {
      // Q <- Q J (all loads first: the stores go to other entries)
      T qa[NQP], qb[NQP], sq[NQP];
      double cq[NQP];
#pragma unroll
      for (int q = 0; q < NQP; ++q) {
        const int kk = tq / JP + ((NTA - QT0) / JP) * q;
        cq[q] = rc[cb][kk];
        sq[q] = rs[cb][kk];
        qa[q] = Qs[aq[q] * LDP + qrow];
        qb[q] = Qs[bq[q] * LDP + qrow];
      }
#pragma unroll
      for (int q = 0; q < NQP; ++q) {
        Qs[aq[q] * LDP + qrow] = cq[q] * qa[q] - conj(sq[q]) * qb[q];
        Qs[bq[q] * LDP + qrow] = sq[q] * qa[q] + cq[q] * qb[q];
      }
    }
    // The flag of sweep sw-1 is free again once every thread has read it at
    // the end of that sweep (one barrier ago at least).
    if (t == 0 && rin == 1 && sw > 0) s_rot[(sw - 1) & 1] = 0;
    r0 = adv_pos(r0);
    r1 = adv_pos(r1);
    c0 = adv_pos(c0);
    c1 = adv_pos(c1);
#pragma unroll
    for (int q = 0; q < NQP; ++q) { aq[q] = adv_pos(aq[q]); bq[q] = adv_pos(bq[q]); }
    __syncthreads();
    if (++rin == RPS) {
      // sweep sw ends: its flag, and the next sweep's (set only by the
      // rotations of that sweep's first round, already applied)
      const int rot = s_rot[sw & 1] | s_rot[(sw + 1) & 1];
      any |= rot;
      rin = 0;
      ++sw;
      if (!rot) break;
    }
  }
  if (t == 0) qflag[k] = any;
  if (!any) return;
  for (int e = t; e < JP * JP; e += NTA) {
    const int r = e % JP, c = e / JP;
    A[(size_t)gidx(I, J, r) + (size_t)gidx(I, J, c) * ld] = As[c * LDP + r];
    Qout[(size_t)k * JP * JP + e] = Qs[c * LDP + r];
  }
}
// ---------------------------------------------------------------------------
// Kernel B: apply the round's block-diagonal unitary P (pair blocks Q_P):
//   off-diagonal pair blocks  rho_PQ <- Q_P^H rho_PQ Q_Q  (and the mirror),
//   eigenvector tiles         V(rows, P) <- V(rows, P) Q_P.
// The diagonal pair blocks were written by k_pair_eig.  JP x JP x JP products
// on v_mfma_f64_16x16x4_f64 (complex: four real MFMAs on re/im planes); wave w
// owns the 16x16 output tiles w, w+4, ...  The MFMA is issued with the right operand
// in the first slot, so lane l's accumulator register r holds output
// (row l%16, column l/16 + 4r) of its 16x16 tile.
constexpr int TD = JP / 16;        // 16x16 output tiles per dimension
constexpr int TPW = TD * TD / 4;   // tiles per wave (4 waves)
static_assert(TPW >= 1 && (TD * TD) % 4 == 0, "tile mapping");

template <class T>
__device__ inline void tile_mm(const double* __restrict__ Ls, bool ltrans, const double* __restrict__ Rs,
                               d4_t (&acc)[NP<T>::P][TPW], int wave, int lane) {
  constexpr int P = NP<T>::P;
  constexpr int PS = JP * LDP;
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int u = 0; u < TPW; ++u) acc[p][u] = d4_t{0.0, 0.0, 0.0, 0.0};
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll 4
  for (int kk = 0; kk < JP; kk += 4) {
    const int kx = kk + lk;
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      const int id = wave + 4 * u, it = id % TD, jt = id / TD;
      const int i = it * 16 + li, j = jt * 16 + li;
      // right operand R(kx, j) stored R[j * LDP + kx]
      const double br = Rs[j * LDP + kx];
      // left operand L(i, kx) = X[kx * LDP + i], or conj(X[i * LDP + kx]) when transposed
      const int ia = ltrans ? (i * LDP + kx) : (kx * LDP + i);
      const double ar = Ls[ia];
      acc[0][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(br, ar, acc[0][u], 0, 0, 0);
      if (P == 2) {
        const double bi = Rs[PS + j * LDP + kx];
        double ai = Ls[PS + ia];
        if (ltrans) ai = -ai;
        // re += ar*br - ai*bi ; im += ar*bi + ai*br
        acc[0][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(-bi, ai, acc[0][u], 0, 0, 0);
        acc[P - 1][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(bi, ar, acc[P - 1][u], 0, 0, 0);
        acc[P - 1][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(br, ai, acc[P - 1][u], 0, 0, 0);
      }
    }
  }
}

// Write accumulators (tile layout above) into LDS planes Xs (col-major, LDP).
template <class T>
__device__ inline void acc_to_lds(double* __restrict__ Xs, d4_t (&acc)[NP<T>::P][TPW], int wave,
                                  int lane) {
  constexpr int P = NP<T>::P;
  constexpr int PS = JP * LDP;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    const int id = wave + 4 * u, it = id % TD, jt = id / TD;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = it * 16 + li, j = jt * 16 + lk + 4 * r;
#pragma unroll
      for (int p = 0; p < P; ++p) Xs[p * PS + j * LDP + i] = acc[p][u][r];
    }
  }
}

template <class T>
__global__ __launch_bounds__(NT) void k_pair_apply(T* __restrict__ A, T* __restrict__ V, int ld,
                                                   int nb, int ka, int round,
                                                   const T* __restrict__ Q,
                                                   const int* __restrict__ qflag) {
  constexpr int P = NP<T>::P;
  constexpr int PS = JP * LDP;
  __shared__ double Xs[P * PS];
  __shared__ double Rs[P * PS];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int np = nb / 2;
  const int noff = np * (np - 1) / 2;
  int bid = blockIdx.x;
  if (bid < noff) {
    // decode the pair block (p1, p2), p1 < p2
    int p1 = 0;
    while (bid >= np - 1 - p1) { bid -= np - 1 - p1; ++p1; }
    const int p2 = p1 + 1 + bid;
    const int f1 = qflag[p1], f2 = qflag[p2];
    if (!(f1 | f2)) return;
    int I1, J1, I2, J2;
    sched_pair(nb, ka, round, p1, I1, J1);
    sched_pair(nb, ka, round, p2, I2, J2);
    for (int e = t; e < JP * JP; e += NT) {
      const int r = e % JP, c = e / JP;
      put(Xs, c * LDP + r, PS, A[(size_t)gidx(I1, J1, r) + (size_t)gidx(I2, J2, c) * ld]);
      if (f2) put(Rs, c * LDP + r, PS, Q[(size_t)p2 * JP * JP + e]);
    }
    __syncthreads();
    d4_t acc[P][TPW];
    if (f2) {
      tile_mm<T>(Xs, false, Rs, acc, wave, lane);  // X Q2
      __syncthreads();
      acc_to_lds<T>(Xs, acc, wave, lane);
    }
    if (f1) {
      for (int e = t; e < JP * JP; e += NT) {
        const int r = e % JP, c = e / JP;
        put(Rs, c * LDP + r, PS, Q[(size_t)p1 * JP * JP + e]);
      }
      __syncthreads();
      tile_mm<T>(Rs, true, Xs, acc, wave, lane);  // Q1^H (X Q2)
      __syncthreads();
      acc_to_lds<T>(Xs, acc, wave, lane);
    }
    __syncthreads();
    for (int e = t; e < JP * JP; e += NT) {
      const int r = e % JP, c = e / JP;
      A[(size_t)gidx(I1, J1, r) + (size_t)gidx(I2, J2, c) * ld] = get<T>(Xs, c * LDP + r, PS);
      // mirror: rho(g2(r), g1(c)) = conj(Z(c, r))
      A[(size_t)gidx(I2, J2, r) + (size_t)gidx(I1, J1, c) * ld] = conj(get<T>(Xs, r * LDP + c, PS));
    }
  } else {
    bid -= noff;
    const int pp = bid % np, rt = bid / np;
    if (!qflag[pp]) return;
    int I, J;
    sched_pair(nb, ka, round, pp, I, J);
    const size_t r0 = (size_t)rt * JP;
    for (int e = t; e < JP * JP; e += NT) {
      const int r = e % JP, c = e / JP;
      put(Xs, c * LDP + r, PS, V[r0 + r + (size_t)gidx(I, J, c) * ld]);
      put(Rs, c * LDP + r, PS, Q[(size_t)pp * JP * JP + e]);
    }
    __syncthreads();
    d4_t acc[P][TPW];
    tile_mm<T>(Xs, false, Rs, acc, wave, lane);
    __syncthreads();
    acc_to_lds<T>(Xs, acc, wave, lane);
    __syncthreads();
    for (int e = t; e < JP * JP; e += NT) {
      const int r = e % JP, c = e / JP;
      V[r0 + r + (size_t)gidx(I, J, c) * ld] = get<T>(Xs, c * LDP + r, PS);
    }
  }
}


// part[blk] = max |A(i,j)|, i != j, over the workgroup's columns j of the
// N x N block (NaN as +inf); grid of NRED, columns blk, blk + NRED, ...  With
// rank given, only pairs with exactly one of rank[i], rank[j] < keep (the
// side of every index is tabulated in LDS first).
template <class T>
__global__ __launch_bounds__(256) void k_offmax(const T* __restrict__ A, int N, int Np,
                                                const int* __restrict__ rank, int keep,
                                                double* __restrict__ part) {
  __shared__ double red[256];
  __shared__ unsigned char side[SORT_MAX];
  const int t = threadIdx.x;
  if (rank) {
    for (int i = t; i < N; i += 256) side[i] = rank[i] < keep ? 1 : 0;
    __syncthreads();
  }
  double mx = 0.0;
  for (int c = blockIdx.x; c < N; c += gridDim.x) {
    const T* col = A + (size_t)c * Np;
    const unsigned char sc = rank ? side[c] : 0;
    for (int r = t; r < N; r += 256) {
      if (r == c || (rank && side[r] == sc)) continue;
      double v = absv(col[r]);
      if (!(v == v)) v = INFINITY;
      mx = fmax(mx, v);
    }
  }
  red[t] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) red[t] = fmax(red[t], red[t + o]);
    __syncthreads();
  }
  if (t == 0) part[blockIdx.x] = red[0];
}

// out(i, j) = V(i, perm[j]) (conj-transposed into out(j, i) with `herm`).
template <class T>
__global__ void k_gather(const T* __restrict__ V, int ldv, const int* __restrict__ perm, int rows,
                         int cols, int herm, T* __restrict__ out, int ldo) {
  const size_t tot = (size_t)rows * cols;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    if (!herm) {
      const int i = mod32(e, rows), j = div32(e, rows);
      out[i + (size_t)j * ldo] = V[i + (size_t)perm[j] * ldv];
    } else {
      const int j = mod32(e, cols), i = div32(e, cols);
      out[j + (size_t)i * ldo] = conj(V[i + (size_t)perm[j] * ldv]);
    }
  }
}

// sigma_j of the QR step.  G = Y^H Y is the Gram matrix of the product
// factor's k vectors (ordered by eigenvalue); sigma_j^2 is the residual of
// vector j after projecting out the earlier ones (R_jj^2 of G = R^H R), taken
// to first order, sigma_j^2 = G_jj - sum_{i<j} |G_ij|^2 / G_ii.  A vector that
// is numerically parallel to earlier ones -- the image of a null-space
// eigenvector contaminated by a kept direction -- then gets sigma ~ 0, as the
// |R_jj| of the QR factorisation of A V in rocsolver_dgesvdj does.
// Two passes keep the first-order sum from counting a shared direction twice:
// pass 0 computes every residual with all earlier vectors; pass 1 subtracts
// only "clean" pivots i (pass-0 residual >= G_ii / 2), so parallel junk
// vectors never act as pivots.
template <class T>
__global__ __launch_bounds__(256) void k_qr_sigma(const T* __restrict__ G, int ldg, int pass,
                                                  const double* __restrict__ r0,
                                                  double* __restrict__ out,
                                                  double* __restrict__ gdiag) {
  __shared__ double red[256];
  const int j = blockIdx.x, t = threadIdx.x;
  double s = 0.0;
  for (int i = t; i < j; i += 256) {
    const double gii = real(G[i + (size_t)i * ldg]);
    if (gii > 0.0 && (pass == 0 || r0[i] >= 0.5 * gii)) s += abs2(G[i + (size_t)j * ldg]) / gii;
  }
  red[t] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) red[t] += red[t + o];
    __syncthreads();
  }
  if (t == 0) {
    const double gjj = real(G[j + (size_t)j * ldg]);
    const double r = gjj - red[0];
    out[j] = pass == 0 ? r : sqrt(fmax(r, 0.0));
    if (gdiag) gdiag[j] = gjj;
  }
}

// Clean pivots of the QR step, normalised: vector i (row i of Y when by_row,
// else column i) scaled by 1/sqrt(G_ii) when its pass-0 residual keeps at
// least half of its norm (r0[i] >= G_ii / 2), zero otherwise.
template <class T>
__global__ void k_clean_scale(const T* __restrict__ Y, int ldy, int k, int len, int by_row,
                              const T* __restrict__ G, int ldg, const double* __restrict__ r0,
                              T* __restrict__ out, int ldo) {
  const size_t tot = (size_t)k * len;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    int i, c;
    if (by_row) { i = mod32(e, k); c = div32(e, k); }
    else { c = mod32(e, len); i = div32(e, len); }
    const double gii = real(G[i + (size_t)i * ldg]);
    const bool clean = gii > 0.0 && r0[i] >= 0.5 * gii;
    const double f = clean ? 1.0 / sqrt(gii) : 0.0;
    if (by_row) out[i + (size_t)c * ldo] = Y[i + (size_t)c * ldy] * f;
    else out[c + (size_t)i * ldo] = Y[c + (size_t)i * ldy] * f;
  }
}

// Keep only the entries of C (k x k) that couple a target vector to an
// earlier pivot: C[j, i], i < j (keep_lower) or C[i, j], i < j.
template <class T>
__global__ void k_mask_tri(T* __restrict__ C, int k, int keep_lower) {
  const size_t tot = (size_t)k * k;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int r = mod32(e, k), c = div32(e, k);
    const bool keep = keep_lower ? (r > c) : (r < c);
    if (!keep) C[e] = T(0);
  }
}

// sigma[j] = |row j| (by_row: Y is k x len, ld ldy) or |column j|.
template <class T>
__global__ __launch_bounds__(256) void k_vecnorm(const T* __restrict__ Y, int ldy, int len,
                                                 int by_row, double* __restrict__ sigma) {
  __shared__ double red[256];
  const int j = blockIdx.x, t = threadIdx.x;
  double s = 0.0;
  for (int i = t; i < len; i += 256)
    s += abs2(by_row ? Y[j + (size_t)i * ldy] : Y[i + (size_t)j * ldy]);
  red[t] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) red[t] += red[t + o];
    __syncthreads();
  }
  if (t == 0) sigma[j] = sqrt(red[0]);
}

// Truncation rank (scalar_traits.h:497-505 semantics, plus the optional
// relative cutoff) and discarded weight trace - sum_{j<nk} |product vector j|^2
// (the weight the kept orthonormal vectors carry, exact for any orthonormal
// basis of the kept space).
__global__ __launch_bounds__(256) void k_rank(const double* __restrict__ sigma,
                                              const double* __restrict__ gdiag, int k,
                                              double cutoff, double rel, double* scal, int* newk) {
  __shared__ int first[256];
  __shared__ double part[256];
  const int t = threadIdx.x;
  const double thr = fmax(cutoff, rel * sigma[0]);
  int f = k;  // first index failing the cutoff
  for (int i = t; i < k; i += 256) {
    const double s = sigma[i];
    if (s < thr || s != s) { f = i; break; }
  }
  first[t] = f;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) first[t] = min(first[t], first[t + o]);
    __syncthreads();
  }
  const int nk = first[0] == 0 ? 1 : first[0];
  double kept = 0.0;
  for (int i = t; i < nk; i += 256) kept += gdiag[i];
  part[t] = kept;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) part[t] += part[t + o];
    __syncthreads();
  }
  if (t == 0) {
    scal[1] = fmax(scal[2] - part[0], 0.0);
    *newk = nk;
  }
}

// out(j, c) = Y(j, c) for j < nk (ld nk) from Y (ld ldy).
template <class T>
__global__ void k_copy_rows(const T* __restrict__ Y, int ldy, int nk, int n, T* __restrict__ out) {
  const size_t tot = (size_t)nk * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int j = mod32(e, nk), c = div32(e, nk);
    out[e] = Y[j + (size_t)c * ldy];
  }
}

// First-order eigenvector correction of a nearly diagonal Hermitian A
// (Np x Np, real part of the diagonal = lambda): the generator E with
// E_ij = A_ij / (lambda_j - lambda_i) for pairs the perturbation resolves
// (|lambda_j - lambda_i| > sep |A_ij|), 0 otherwise (left to Jacobi) and on
// the diagonal.  E is exactly anti-Hermitian (A is Hermitian and the test is
// symmetric), so I + E is unitary up to O(|E|^2).  One workgroup per column.
template <class T>
__global__ __launch_bounds__(256) void k_oa_q(const T* __restrict__ A, T* __restrict__ Q, int N,
                                              int Np, double sep) {
  const int c = blockIdx.x;
  const double lc = real(A[c + (size_t)c * Np]);
  for (int r = threadIdx.x; r < Np; r += 256) {
    const size_t e = r + (size_t)c * Np;
    T q = T(0);
    if (r != c && r < N && c < N) {
      const T a = A[e];
      const double g = lc - real(A[r + (size_t)r * Np]);
      if (fabs(g) > sep * sqrt(abs2(a))) q = a * (1.0 / g);
    }
    Q[e] = q;
  }
}

// First Newton-Schulz factor from F = E^2 (Q_0 = I + E, Q_0^H Q_0 = I - F):
// F <- (3I - (I - F))/2 = I + F/2, and part[blk] = sum of |F|^2 =
// |Q_0^H Q_0 - I|^2 over the workgroup's entries (grid NRED).
template <class T>
__global__ __launch_bounds__(256) void k_ns_first(T* __restrict__ F, int n, double* __restrict__ part) {
  __shared__ double red[256];
  const size_t tot = (size_t)n * n;
  double s2 = 0.0;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int r = mod32(e, n), c = div32(e, n);
    const T f = F[e];
    s2 += abs2(f);
    F[e] = f * 0.5 + T(r == c ? 1.0 : 0.0);
  }
  red[threadIdx.x] = s2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// P <- (3I - P)/2 (one Newton-Schulz step's polynomial factor); with part
// given, part[blk] = sum of |P - I|^2 over the workgroup's entries (grid NRED).
template <class T>
__global__ __launch_bounds__(256) void k_ns_half(T* __restrict__ P, int n, double* __restrict__ part) {
  __shared__ double red[256];
  const size_t tot = (size_t)n * n;
  double s2 = 0.0;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int r = mod32(e, n), c = div32(e, n);
    const T one = (r == c) ? T(1) : T(0);
    const T p = P[e];
    s2 += abs2(p - one);
    P[e] = (3.0 * one - p) * 0.5;
  }
  if (!part) return;
  red[threadIdx.x] = s2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// X_ii += b.
template <class T>
__global__ void k_add_diag(T* __restrict__ X, int n, double b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) X[i + (size_t)i * n] = X[i + (size_t)i * n] + T(b);
}

inline int round_up(int a, int b) { return (a + b - 1) / b * b; }
constexpr int MGS_B = 32;
// Modified Gram-Schmidt, twice, on the block (the fallback of k_block_orth
// for an ill-conditioned block): for each column j, the norm and the inner
// products with the later columns come from one pass over the rows, the
// later columns are projected, and column j is normalised; a column below
// `dep` times its norm before the orthogonalisation becomes zero.
template <class T>
__device__ void mgs_block(T* __restrict__ Y, int len, int ldy, int b,
                          const double* __restrict__ nrm0, double dep, T (*red)[8],
                          double* nred) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (int pass = 0; pass < 2; ++pass) {
    for (int j = 0; j < b; ++j) {
      T* yj = Y + (size_t)j * ldy;
      // |y_j|^2 and <y_j, y_c> for c > j (unnormalised y_j)
      double nn = 0.0;
      T dot[MGS_B];
#pragma unroll
      for (int c = 0; c < MGS_B; ++c) dot[c] = T(0);
      for (int i = t; i < len; i += 256) {
        const T v = yj[i];
        nn += abs2(v);
        const T cv = conj(v);
#pragma unroll
        for (int c = 0; c < MGS_B; ++c)
          if (c > j && c < b) dot[c] += cv * Y[i + (size_t)c * ldy];
      }
      for (int o = 32; o > 0; o >>= 1) nn += __shfl_xor(nn, o);
#pragma unroll
      for (int c = 0; c < MGS_B; ++c) {
        if (c > j && c < b) {
          double re = real(dot[c]), im = imag(dot[c]);
          for (int o = 32; o > 0; o >>= 1) {
            re += __shfl_xor(re, o);
            im += __shfl_xor(im, o);
          }
          if (lane == 0) red[c][wv] = make_scalar<T>(re, im);
        }
      }
      if (lane == 0) nred[wv] = nn;
      __syncthreads();
      const double n2 = nred[0] + nred[1] + nred[2] + nred[3];
      const double nrm = sqrt(n2);
      // pass 0 against the norm before any orthogonalisation, pass 1
      // against the unit norm that pass 0 left
      const bool keep = n2 > 0.0 && nrm > dep * (pass == 0 ? nrm0[j] : 1.0);
      const double inv = keep ? 1.0 / nrm : 0.0;
      // y_c -= (<y_j, y_c> / |y_j|^2) y_j, then y_j /= |y_j| (or 0)
      for (int c = j + 1; c < b; ++c) {
        const T d = red[c][0] + red[c][1] + red[c][2] + red[c][3];
        const T coef = keep ? d * (inv * inv) : T(0);
        T* yc = Y + (size_t)c * ldy;
        for (int i = t; i < len; i += 256) yc[i] -= coef * yj[i];
      }
      __syncthreads();  // every column update has read y_j
      for (int i = t; i < len; i += 256) yj[i] = yj[i] * inv;
      __syncthreads();
    }
  }
}

// Orthonormalisation of the b <= MGS_B columns of one block of Y (len rows,
// ld ldy) by one workgroup, as Cholesky QR twice: the Gram matrix
// G = Y_b^H Y_b is accumulated in LDS over chunks of rows, factored
// G = R^H R in LDS, and every row y of the block is replaced by y R^-1
// (forward substitution, one thread per row).  Column j is numerically
// dependent when its squared norm after the earlier columns' projection,
// G_jj - sum_i |R_ij|^2, is not above (dep * ref_j)^2 and 1e-14 G_jj (the
// Gram matrix's own rounding), ref_j being its norm before any
// orthogonalisation (nrm0, device) in the first pass and 1 in the second; a
// dependent column is set to zero (the caller's sketch then simply carries
// fewer directions).  One pass over the block's rows per Gram matrix instead
// of MGS's b passes: the block of a sketch, projected against the earlier
// blocks, is well conditioned, so the squared condition number that Cholesky
// QR costs is harmless, and the second pass restores orthogonality to
// rounding.  A block whose first factorisation finds a column with less than
// 1e-4 of its norm left after the earlier columns' projection (condition
// number above ~1e4, where Cholesky QR would lose directions that MGS keeps)
// is left untouched by it and orthonormalised by MGS instead.
template <class T>
__global__ __launch_bounds__(256) void k_block_orth(T* __restrict__ Y, int len, int ldy, int b,
                                                    const double* __restrict__ nrm0,
                                                    double dep) {
  constexpr int CH = 32;                 // rows per chunk of the Gram accumulation
  __shared__ T Ys[CH][MGS_B + 1];        // chunk of the block
  __shared__ T G[MGS_B][MGS_B + 1];      // Gram matrix, then R (upper triangle)
  __shared__ double Gd0[MGS_B];          // G_jj before the factorisation
  __shared__ int kept[MGS_B];
  __shared__ int ill;
  const int t = threadIdx.x;
  // (p, q) entries of the upper triangle owned by thread t: e = t, t + 256, ...
  constexpr int NE = MGS_B * (MGS_B + 1) / 2, NU = (NE + 255) / 256;
  int pu[NU], qu[NU];
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    int p = 0, rem = t + 256 * u;  // row-major upper triangle: e -> (p, q), p <= q
    while (p < MGS_B && rem >= MGS_B - p) { rem -= MGS_B - p; ++p; }
    pu[u] = p;
    qu[u] = p + rem;
  }
  for (int pass = 0; pass < 2; ++pass) {
    T acc[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) acc[u] = T(0);
    for (int r0 = 0; r0 < len; r0 += CH) {
      for (int e = t; e < CH * MGS_B; e += 256) {
        const int i = e % CH, c = e / CH;
        Ys[i][c] = (r0 + i < len && c < b) ? Y[r0 + i + (size_t)c * ldy] : T(0);
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        if (t + 256 * u < NE) {
          const int p = pu[u], q = qu[u];
          T sacc = acc[u];
          for (int i = 0; i < CH; ++i) sacc += conj(Ys[i][p]) * Ys[i][q];
          acc[u] = sacc;
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      if (t + 256 * u < NE) {
        const int p = pu[u], q = qu[u];
        G[p][q] = acc[u];
        if (p == q) Gd0[p] = real(acc[u]);
      }
    }
    __syncthreads();
    // right-looking Cholesky with dependent columns dropped
    if (t == 0) ill = 0;
    for (int j = 0; j < b; ++j) {
      if (t == 0) {
        const double d = real(G[j][j]);
        const double ref = pass == 0 ? nrm0[j] : 1.0;
        const double tiny = fmax(dep * dep * ref * ref, 1e-14 * Gd0[j]);
        kept[j] = (d == d) && d > tiny;
        if (pass == 0 && !(d > 1e-8 * Gd0[j]) && Gd0[j] > 0.0) ill = 1;
        G[j][j] = T(kept[j] ? sqrt(d) : 0.0);
      }
      __syncthreads();
      const bool kj = kept[j];
      const double inv = kj ? 1.0 / real(G[j][j]) : 0.0;
      for (int q = j + 1 + t; q < b; q += 256) G[j][q] = kj ? G[j][q] * inv : T(0);
      __syncthreads();
      if (kj) {
        const int m = b - j - 1;  // trailing (p, q), j < p <= q < b
        for (int e = t; e < m * m; e += 256) {
          const int p = j + 1 + e % m, q = j + 1 + e / m;
          if (p <= q) G[p][q] -= conj(G[j][p]) * G[j][q];
        }
      }
      __syncthreads();
    }
    if (ill) {  // (set by thread 0 before the last barrier of the loop)
      __shared__ T red[MGS_B][8];
      __shared__ double nred[8];
      mgs_block<T>(Y, len, ldy, b, nrm0, dep, red, nred);
      return;
    }
    // rows: y <- y R^-1 (dependent columns become zero)
    for (int i = t; i < len; i += 256) {
      T x[MGS_B];
#pragma unroll
      for (int c = 0; c < MGS_B; ++c) x[c] = c < b ? Y[i + (size_t)c * ldy] : T(0);
#pragma unroll
      for (int c = 0; c < MGS_B; ++c) {
        if (c < b) {
          T v = x[c];
#pragma unroll
          for (int k = 0; k < c; ++k) v -= x[k] * G[k][c];
          x[c] = kept[c] ? v * (1.0 / real(G[c][c])) : T(0);
        }
      }
#pragma unroll
      for (int c = 0; c < MGS_B; ++c)
        if (c < b) Y[i + (size_t)c * ldy] = x[c];
    }
    __syncthreads();
  }
}

inline unsigned grid_for(size_t tot) { return (unsigned)std::min<size_t>(cdiv(tot, 256), 8192); }

// Block-Jacobi eigen-decomposition of the Hermitian matrix already stored in
// w.A[0] (N x N, ld Np, padding arbitrary).  Leaves the diagonalised matrix in
// w.A[cur], the eigenvectors in w.V[cur] and the descending order in w.perm.
// keep > 0: converged once the `keep` largest diagonal entries decouple from
// the rest (max |A_ij| over those pairs <= tol |A|_F); couplings inside either
// group are left as they are.  When the keep-th largest diagonal entry comes
// within a factor 1e4 of max(cut_abs^2, cut_rel^2 * largest) -- the cutoff
// may then drop kept vectors, which needs them resolved -- the iteration goes
// on to full convergence.
struct JacobiCount {
  int sweeps = 0, bip = 0, oa = 0;
};

template <class T>
int jacobi(hipStream_t st, SvdWork<T>& w, int N, int Np, const SvdParams& prm, JacobiCount& cnt,
           double& off_rel, bool v_given = false, int keep = 0, double cut_abs = 0.0,
           double cut_rel = 0.0) {
  int& sweeps = cnt.sweeps;
  const int nb = Np / JB;
  const int np = nb / 2;
  const size_t tot = (size_t)Np * Np;
  hipLaunchKernelGGL(k_prep<T>, dim3(grid_for(tot)), dim3(256), 0, st, w.A[0], w.V[0], N, Np,
                     v_given ? 0 : 1);
  {
    const int nbr = (int)std::min<size_t>(NRED, std::max<size_t>(1, cdiv(tot, 4096)));
    hipLaunchKernelGGL(k_norms_part<T>, dim3(nbr), dim3(256), 0, st, w.A[0], N, Np, w.red);
    hipLaunchKernelGGL(k_norms_fin, dim3(1), dim3(256), 0, st, (const double*)w.red, nbr, w.scal);
  }
  DMRG_LAUNCH_CHECK();
  double* const hs = w.host;  // pinned: the copy engine writes it directly
  DMRG_HIP(hipMemcpyAsync(hs, w.scal, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
  DMRG_HIP(hipStreamSynchronize(st));
  const double fro = hs[0];
  if (!(fro == fro) || std::isinf(fro)) throw Error(ST_NONFINITE, "svd: non-finite matrix entries");
  const double rot_tol = 0.125 * prm.tol;
  const int inner_max = (nb == 2) ? 64 : 2;
  const int noff = np * (np - 1) / 2, nvt = (Np / JP) * np;
  int cur = 0;
  sweeps = 0;
  off_rel = 0.0;
  if (fro == 0.0) {
    hipLaunchKernelGGL(k_sort<T>, dim3(cdiv(Np, 16)), dim3(256), 0, st, w.A[cur], Np, nullptr, Np, N,
                       w.perm, nullptr);
    DMRG_LAUNCH_CHECK();
    return cur;
  }
  double key_hi = 0.0, key_keep = 0.0;  // largest and keep-th largest diagonal entry
  double* const offpart = w.red + 2 * NRED;  // per-workgroup maxima of k_offmax
  auto offmax = [&](const T* M) {
    if (keep > 0)
      hipLaunchKernelGGL(k_rank_diag<T>, dim3(cdiv(Np, 16)), dim3(256), 0, st, M, N, Np, keep,
                         w.rank, w.red);
    hipLaunchKernelGGL(k_offmax<T>, dim3(NRED), dim3(256), 0, st, M, N, Np,
                       keep > 0 ? (const int*)w.rank : nullptr, keep, offpart);
    DMRG_LAUNCH_CHECK();
    double* const hp = w.host;         // NRED workgroup maxima
    double* const kk = w.host + NRED;  // largest and keep-th largest diagonal entry
    DMRG_HIP(hipMemcpyAsync(hp, offpart, NRED * sizeof(double), hipMemcpyDeviceToHost, st));
    if (keep > 0)
      DMRG_HIP(hipMemcpyAsync(kk, w.red, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    DMRG_HIP(hipStreamSynchronize(st));
    if (keep > 0) {
      key_hi = kk[0];
      key_keep = kk[1];
    }
    double o = 0.0;
    for (int b = 0; b < NRED; ++b) o = std::max(o, hp[b]);
    if (std::isinf(o) || !(o == o)) throw Error(ST_NONFINITE, "svd: non-finite matrix entries");
    return o;
  };
  // subspace convergence reached: accept it, or switch to full convergence
  // when the cutoff reaches into the kept group (returns the full off-norm).
  auto accept_keep = [&](const T* M, double o) {
    if (keep == 0) return o;
    const double floor = 1e4 * std::max(cut_abs * cut_abs, cut_rel * cut_rel * key_hi);
    if (key_keep >= floor) return o;
    keep = 0;
    return offmax(M);
  };
  bool done = false;
  double last_off = INFINITY;  // latest convergence measure (criterion in force)
  if (v_given) {
    // Started from a basis near the eigenbasis: up to two first-order
    // (all-pairs, GEMM-only) corrections before any Jacobi sweep.  A pair is
    // corrected when its gap exceeds sep |A_ij| (the first-order step then
    // leaves O(|A_ij|/sep)); the subspace criterion tolerates the rougher sep.
    // Steps go on while each one cuts the measure at least tenfold.  Pairs
    // near the cut whose gap is only 30-100 times their coupling stall the
    // step at sep = 1e2; the subspace criterion then retries with sep = 30
    // (first-order error |A_ij|/30 per step) before handing over to Jacobi.
    const GemmTile tl = gemm_pick_tile(Np, Np, Np, is_cplx<T>::value);
    double sep = keep > 0 ? 1e2 : 1e4;
    const double gate = keep > 0 ? 1e-4 : 1e-6;
    const int max_steps = keep > 0 ? 5 : 2;
    for (int it = 0; it <= max_steps && !done; ++it) {
      double o = offmax(w.A[0]);
      if (o <= prm.tol * fro) o = accept_keep(w.A[0], o);
      const double prev = last_off;
      last_off = o;
      off_rel = o / fro;
      if (o <= prm.tol * fro) { done = true; break; }
      if (it == max_steps || o > gate * fro) break;  // not perturbative: Jacobi
      if (it > 0 && o > 0.1 * prev) {                // stalled (near-degenerate pairs)
        if (keep > 0 && sep > 30.0) sep = 30.0;
        else break;                                  // Jacobi
      }
      hipLaunchKernelGGL(k_oa_q<T>, dim3(Np), dim3(256), 0, st, (const T*)w.A[0], w.Y, N, Np, sep);
      // The unitary factor Q2 of I + E by Newton-Schulz, Q <- Q (3I - Q^H Q)/2
      // from Q_0 = I + E.  Every eigenvalue x of Q^H Q - I goes to
      // -x^2 (3 - x)/4, so the Frobenius deviation delta = |Q^H Q - I|_F obeys
      // delta' <= delta^2 (3 + delta)/4.  E is anti-Hermitian, so
      // Q_0^H Q_0 = I - F with F = E^2: the first step's factor is I + F/2
      // (one product less), and delta_0 = |F|_F is read with it.  The
      // recursion then fixes the step count: the step taken once the bound is
      // <= 1e-8 is the last (it leaves O(1e-16)).  From a start that four steps
      // do not certify, the reading before step 4 decides (<= 1e-8 to take
      // it, else Jacobi takes over with A and V unchanged).
      T* freeb[4] = {w.A[1], w.V[1], w.W, w.Y};  // Y holds E
      double* const nspart = w.red + 2 * NRED;  // free while no offmax is pending
      T* Z = w.Y;
      T* P = freeb[0];
      T* Zn = freeb[1];
      gemm_herm<T>(st, OP_N, OP_N, Np, Np, T(1), w.Y, Np, w.Y, Np, P, Np, tl);  // F = E^2
      hipLaunchKernelGGL(k_ns_first<T>, dim3(NRED), dim3(256), 0, st, P, Np, nspart);
      hipLaunchKernelGGL(k_add_diag<T>, dim3(cdiv(Np, 256)), dim3(256), 0, st, w.Y, Np, 1.0);
      DMRG_LAUNCH_CHECK();
      gemm<T>(st, OP_N, OP_N, Np, Np, Np, T(1), Z, Np, P, Np, T(0), Zn, Np, tl);
      std::swap(Z, Zn);
      double* const dp = w.host;
      DMRG_HIP(hipMemcpyAsync(dp, nspart, NRED * sizeof(double), hipMemcpyDeviceToHost, st));
      DMRG_HIP(hipStreamSynchronize(st));
      double d2 = 0.0;
      for (int b = 0; b < NRED; ++b) d2 += dp[b];
      if (!(d2 == d2)) break;  // non-finite generator: Jacobi
      double dl = std::sqrt(d2);  // bound on the deviation before step `last`
      int last = 0;
      while (last < 3 && dl > 1e-8) {
        dl = dl * dl * (3.0 + dl) * 0.25;
        ++last;
      }
      const bool certified = dl <= 1e-8;
      if (!certified) last = 3;
      bool unitary = certified;
      for (int ns = 1; ns <= last; ++ns) {
        const bool check = !certified && ns == 3;
        gemm_herm<T>(st, OP_H, OP_N, Np, Np, T(1), Z, Np, Z, Np, P, Np, tl);
        hipLaunchKernelGGL(k_ns_half<T>, dim3(NRED), dim3(256), 0, st, P, Np,
                           check ? nspart : nullptr);
        gemm<T>(st, OP_N, OP_N, Np, Np, Np, T(1), Z, Np, P, Np, T(0), Zn, Np, tl);
        std::swap(Z, Zn);
        if (check) {
          DMRG_HIP(hipMemcpyAsync(dp, nspart, NRED * sizeof(double), hipMemcpyDeviceToHost, st));
          DMRG_HIP(hipStreamSynchronize(st));
          double c2 = 0.0;
          for (int b = 0; b < NRED; ++b) c2 += dp[b];
          unitary = c2 <= 1e-16;  // |Q^H Q - I|_F <= 1e-8 before this step (false for NaN)
        }
      }
      if (!unitary) break;  // leave it to Jacobi (A, V unchanged)
      T* const Q2 = Z;
      // The three other work buffers: X1 = A Q2, X2 = Q2^H X1 (the new A),
      // X3 = V Q2 (the new V).
      T* xs[3];
      int nx = 0;
      for (T* bptr : freeb)
        if (bptr != Q2) xs[nx++] = bptr;
      gemm<T>(st, OP_N, OP_N, Np, Np, Np, T(1), w.A[0], Np, Q2, Np, T(0), xs[0], Np, tl);
      gemm_herm<T>(st, OP_H, OP_N, Np, Np, T(1), Q2, Np, xs[0], Np, xs[1], Np, tl);
      gemm<T>(st, OP_N, OP_N, Np, Np, Np, T(1), w.V[0], Np, Q2, Np, T(0), xs[2], Np, tl);
      T* const oldA = w.A[0];
      T* const oldV = w.V[0];
      w.A[0] = xs[1];
      w.V[0] = xs[2];
      w.A[1] = oldA;
      w.V[1] = oldV;
      w.W = Q2;
      w.Y = xs[0];
      hipLaunchKernelGGL(k_prep<T>, dim3(grid_for(tot)), dim3(256), 0, st, w.A[0], w.V[0], N, Np, 0);
      DMRG_LAUNCH_CHECK();
      ++cnt.oa;
    }
  }
  bool bip_ok = true;  // bipartite sweeps allowed (revoked once one stalls)
  int n_bip = 0;
  for (int sw = 0; sw < prm.max_sweeps && !done; ++sw) {
    // sort the diagonal, permute rho and V into the other buffer
    hipLaunchKernelGGL(k_sort<T>, dim3(cdiv(Np, 16)), dim3(256), 0, st, w.A[cur], Np, nullptr, Np, N,
                       w.perm, nullptr);
    hipLaunchKernelGGL(k_permute<T>, dim3(grid_for(tot)), dim3(256), 0, st, w.A[cur], w.V[cur],
                       w.A[1 - cur], w.V[1 - cur], w.perm, Np);
    cur = 1 - cur;
    // From a warm start already close to the split, rotate only the couplings
    // between the kept group (the first keep/JB blocks after the sort) and the
    // rest: max(ka, nb-ka) rounds instead of nb-1.  From a far start that
    // schedule need not converge (the kept group itself is still mixing), and
    // near-degenerate rows at the cut can trade places between sweeps, which
    // turns couplings inside a group into cross couplings; so bipartite sweeps
    // run only while each cuts the measure tenfold (at most 4 per split), and
    // the full round robin finishes the job.
    const int ka = (bip_ok && keep > 0 && keep % JB == 0 && keep < Np && v_given &&
                    last_off <= 1e-4 * fro) ? keep / JB : 0;
    const double off_before = last_off;
    const int rounds = sched_rounds(nb, ka);
    for (int r = 0; r < rounds; ++r) {
      hipLaunchKernelGGL(k_pair_eig<T>, dim3(np), dim3(NTA), 0, st, w.A[cur], Np, nb, ka, r, w.Q,
                         w.qflag, (const double*)w.scal, rot_tol, inner_max);
      hipLaunchKernelGGL(k_pair_apply<T>, dim3(noff + nvt), dim3(NT), 0, st, w.A[cur], w.V[cur], Np,
                         nb, ka, r, (const T*)w.Q, (const int*)w.qflag);
    }
    double off = offmax(w.A[cur]);
    if (off <= prm.tol * fro) off = accept_keep(w.A[cur], off);
    last_off = off;
    if (ka > 0) ++cnt.bip;
    if (ka > 0 && (++n_bip >= 4 || off > 0.1 * off_before)) bip_ok = false;
    sweeps = sw + 1;
    off_rel = off / fro;
    if (off <= prm.tol * fro) { done = true; break; }
  }
  if (!done) throw Error(ST_SOLVER_NO_CONV, "svd: Jacobi eigen-solver did not converge");
  hipLaunchKernelGGL(k_sort<T>, dim3(cdiv(Np, 16)), dim3(256), 0, st, w.A[cur], Np, nullptr, Np, N, w.perm,
                     nullptr);
  DMRG_LAUNCH_CHECK();
  return cur;
}

}  // namespace

template <class T>
void SvdWork<T>::alloc(int cap_n) {
  release();
  cap = round_up(std::max(cap_n, 1), JP);
  DMRG_REQUIRE(cap <= SORT_MAX, "svd: dimension above 2048");
  const size_t mat = (size_t)cap * cap * sizeof(T);
  for (int i = 0; i < 2; ++i) {
    DMRG_HIP(hipMalloc(&A[i], mat));
    DMRG_HIP(hipMalloc(&V[i], mat));
  }
  DMRG_HIP(hipMalloc(&Y, mat));
  DMRG_HIP(hipMalloc(&W, mat));
  DMRG_HIP(hipMalloc(&Q, (size_t)(cap / JP) * JP * JP * sizeof(T)));
  DMRG_HIP(hipMalloc(&qflag, (cap / JP) * sizeof(int)));
  DMRG_HIP(hipMalloc(&perm, cap * sizeof(int)));
  DMRG_HIP(hipMalloc(&rank, cap * sizeof(int)));
  DMRG_HIP(hipMalloc(&gdiag, cap * sizeof(double)));
  DMRG_HIP(hipMalloc(&red, 4 * NRED * sizeof(double)));
  DMRG_HIP(hipMalloc(&sigma, cap * sizeof(double)));
  DMRG_HIP(hipMalloc(&resid, cap * sizeof(double)));
  DMRG_HIP(hipMalloc(&scal, 4 * sizeof(double)));
  DMRG_HIP(hipMalloc(&newk, sizeof(int)));
  // every device-to-host read of the solver lands in pinned memory, so no
  // copy goes through the runtime's pageable staging path
  DMRG_HIP(hipHostMalloc((void**)&host, (2 * (size_t)cap + 2 * NRED + 16) * sizeof(double),
                         hipHostMallocDefault));
  DMRG_HIP(hipHostMalloc((void**)&host_i, 4 * sizeof(int), hipHostMallocDefault));
}

template <class T>
void SvdWork<T>::release() {
  for (int i = 0; i < 2; ++i) {
    if (A[i]) (void)hipFree(A[i]);
    if (V[i]) (void)hipFree(V[i]);
    A[i] = V[i] = nullptr;
  }
  void* ps[] = {Y, W, Q, qflag, perm, rank, gdiag, red, sigma, resid, scal, newk};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  if (host) (void)hipHostFree(host);
  if (host_i) (void)hipHostFree(host_i);
  host = nullptr;
  host_i = nullptr;
  Y = nullptr;
  W = nullptr;
  Q = nullptr;
  qflag = nullptr;
  perm = nullptr;
  rank = nullptr;
  gdiag = nullptr;
  red = nullptr;
  sigma = nullptr;
  resid = nullptr;
  scal = nullptr;
  newk = nullptr;
  cap = 0;
}

template <class T>
SvdResult svd_split(hipStream_t st, SvdWork<T>& w, const T* A, int m, int n, int lda, int kmax,
                    double cutoff, double rel_cutoff, bool scale_left, T* left, int ldl, T* right,
                    double* sigma_out, const SvdParams& prm, const T* warm, T* basis_out) {
  DMRG_REQUIRE(m > 0 && n > 0 && lda >= m && kmax >= 1, "svd: bad shape");
  DMRG_REQUIRE(left && right && ldl >= m, "svd: bad output");
  const int N = scale_left ? n : m;
  const int Np = round_up(N, JP);
  DMRG_REQUIRE(Np <= w.cap, "svd: matrix exceeds workspace");
  const int k = std::min(std::min(m, n), kmax);
  // density matrix on the side of the orthonormal factor; from the caller's
  // orthonormal basis X (warm start, V = X) directly in that basis:
  // X^H theta theta^H X = Y Y^H with Y = X^H theta ('R'), X^H theta^H theta X =
  // Y^H Y with Y = theta X ('L'): one GEMM and one Hermitian product instead
  // of three products
  if (warm) {
    hipLaunchKernelGGL(k_set_basis<T>, dim3(grid_for((size_t)Np * Np)), dim3(256), 0, st, w.V[0],
                       warm, N, Np);
    if (!scale_left) {  // N = m; Y (m x n, ld m) in w.Y
      gemm<T>(st, OP_H, OP_N, m, n, m, T(1), w.V[0], Np, A, lda, T(0), w.Y, m,
              gemm_pick_tile(m, n, m, is_cplx<T>::value));
      gemm_herm<T>(st, OP_N, OP_H, m, n, T(1), w.Y, m, w.Y, m, w.A[0], Np,
                   gemm_pick_tile(m, m, n, is_cplx<T>::value));
    } else {  // N = n; Y (m x n, ld m) in w.Y
      gemm<T>(st, OP_N, OP_N, m, n, n, T(1), A, lda, w.V[0], Np, T(0), w.Y, m,
              gemm_pick_tile(m, n, n, is_cplx<T>::value));
      gemm_herm<T>(st, OP_H, OP_N, n, m, T(1), w.Y, m, w.Y, m, w.A[0], Np,
                   gemm_pick_tile(n, n, m, is_cplx<T>::value));
    }
  } else if (!scale_left) {
    gemm_herm<T>(st, OP_N, OP_H, m, n, T(1), A, lda, A, lda, w.A[0], Np,
                 gemm_pick_tile(m, m, n, is_cplx<T>::value));
  } else {
    gemm_herm<T>(st, OP_H, OP_N, n, m, T(1), A, lda, A, lda, w.A[0], Np,
                 gemm_pick_tile(n, n, m, is_cplx<T>::value));
  }
  SvdResult res;
  res.N = N;
  const int keep = (prm.subspace && k < N) ? k : 0;
  JacobiCount jc;
  const int cur = jacobi<T>(st, w, N, Np, prm, jc, res.off, warm != nullptr, keep, cutoff,
                            rel_cutoff);
  res.sweeps = jc.sweeps;
  res.bip_sweeps = jc.bip;
  res.oa_steps = jc.oa;
  if (basis_out)
    hipLaunchKernelGGL(k_gather<T>, dim3(grid_for((size_t)N * N)), dim3(256), 0, st,
                       (const T*)w.V[cur], Np, (const int*)w.perm, N, N, 0, basis_out, N);
  if (!scale_left) {
    // U_k -> left; S Vh = U_k^H theta -> Y (k x n); G = Y Y^H
    hipLaunchKernelGGL(k_gather<T>, dim3(grid_for((size_t)m * k)), dim3(256), 0, st,
                       (const T*)w.V[cur], Np, (const int*)w.perm, m, k, 0, left, ldl);
    gemm<T>(st, OP_H, OP_N, k, n, m, T(1), left, ldl, A, lda, T(0), w.Y, k,
            gemm_pick_tile(k, n, m, is_cplx<T>::value));
    gemm_herm<T>(st, OP_N, OP_H, k, n, T(1), w.Y, k, w.Y, k, w.A[1 - cur], k,
                 gemm_pick_tile(k, k, n, is_cplx<T>::value));
  } else {
    // V_k -> Y (n x k); U S = theta V_k -> left; G = left^H left
    hipLaunchKernelGGL(k_gather<T>, dim3(grid_for((size_t)n * k)), dim3(256), 0, st,
                       (const T*)w.V[cur], Np, (const int*)w.perm, n, k, 0, w.Y, n);
    gemm<T>(st, OP_N, OP_N, m, k, n, T(1), A, lda, w.Y, n, T(0), left, ldl,
            gemm_pick_tile(m, k, n, is_cplx<T>::value));
    gemm_herm<T>(st, OP_H, OP_N, k, m, T(1), left, ldl, left, ldl, w.A[1 - cur], k,
                 gemm_pick_tile(k, k, m, is_cplx<T>::value));
  }
  // sigma_j: residual of product vector j after projecting out the earlier
  // clean pivots, computed on the vectors themselves (absolute accuracy
  // ~eps |theta|, like |R_jj| of a QR), pivots chosen from the Gram matrix.
  hipLaunchKernelGGL(k_qr_sigma<T>, dim3(k), dim3(256), 0, st, (const T*)w.A[1 - cur], k, 0,
                     (const double*)nullptr, w.resid, w.gdiag);
  {
    const bool byr = !scale_left;            // 'U' side: vectors are the rows of Y (k x n)
    const T* Yv = byr ? w.Y : left;
    const int ldy = byr ? k : ldl;
    const int len = byr ? n : m;
    T* Qc = w.V[1 - cur];
    T* C = w.A[1 - cur];
    T* R = w.A[cur];
    const int ldq = byr ? k : m;
    hipLaunchKernelGGL(k_clean_scale<T>, dim3(grid_for((size_t)k * len)), dim3(256), 0, st, Yv,
                       ldy, k, len, byr ? 1 : 0, (const T*)w.A[1 - cur], k,
                       (const double*)w.resid, Qc, ldq);
    if (byr) {
      gemm<T>(st, OP_N, OP_H, k, k, len, T(1), Yv, ldy, Qc, ldq, T(0), C, k,
              gemm_pick_tile(k, k, len, is_cplx<T>::value));
      hipLaunchKernelGGL(k_mask_tri<T>, dim3(grid_for((size_t)k * k)), dim3(256), 0, st, C, k, 1);
      copy_dev<T>(st, R, (size_t)k, Yv, (size_t)ldy, (size_t)k, (size_t)len);
      gemm<T>(st, OP_N, OP_N, k, len, k, T(-1), C, k, Qc, ldq, T(1), R, k,
              gemm_pick_tile(k, len, k, is_cplx<T>::value));
      hipLaunchKernelGGL(k_vecnorm<T>, dim3(k), dim3(256), 0, st, (const T*)R, k, len, 1, w.sigma);
    } else {
      gemm<T>(st, OP_H, OP_N, k, k, len, T(1), Qc, ldq, Yv, ldy, T(0), C, k,
              gemm_pick_tile(k, k, len, is_cplx<T>::value));
      hipLaunchKernelGGL(k_mask_tri<T>, dim3(grid_for((size_t)k * k)), dim3(256), 0, st, C, k, 0);
      copy_dev<T>(st, R, (size_t)m, Yv, (size_t)ldy, (size_t)m, (size_t)k);
      gemm<T>(st, OP_N, OP_N, len, k, k, T(-1), Qc, ldq, C, k, T(1), R, m,
              gemm_pick_tile(len, k, k, is_cplx<T>::value));
      hipLaunchKernelGGL(k_vecnorm<T>, dim3(k), dim3(256), 0, st, (const T*)R, m, len, 0, w.sigma);
    }
  }
  hipLaunchKernelGGL(k_rank, dim3(1), dim3(256), 0, st, (const double*)w.sigma,
                     (const double*)w.gdiag, k, cutoff, rel_cutoff, w.scal, w.newk);
  DMRG_LAUNCH_CHECK();
  DMRG_HIP(hipMemcpyAsync(w.host_i, w.newk, sizeof(int), hipMemcpyDeviceToHost, st));
  DMRG_HIP(hipMemcpyAsync(w.host, w.scal + 1, sizeof(double), hipMemcpyDeviceToHost, st));
  DMRG_HIP(hipStreamSynchronize(st));
  const int nk = w.host_i[0];
  const double disc = w.host[0];
  if (keep > 0 && nk < k) {
    // the cutoff cuts inside the kept group, whose vectors are not resolved:
    // redo the split with every pair converged
    SvdParams full = prm;
    full.subspace = false;
    SvdResult r2 = svd_split<T>(st, w, A, m, n, lda, kmax, cutoff, rel_cutoff, scale_left, left,
                                ldl, right, sigma_out, full, warm, basis_out);
    r2.sweeps += res.sweeps;
    r2.bip_sweeps += res.bip_sweeps;
    r2.oa_steps += res.oa_steps;
    return r2;
  }
  res.nk = nk;
  res.discarded = disc;
  if (!scale_left)
    hipLaunchKernelGGL(k_copy_rows<T>, dim3(grid_for((size_t)nk * n)), dim3(256), 0, st,
                       (const T*)w.Y, k, nk, n, right);
  else
    hipLaunchKernelGGL(k_gather<T>, dim3(grid_for((size_t)n * nk)), dim3(256), 0, st,
                       (const T*)w.V[cur], Np, (const int*)w.perm, n, nk, 1, right, nk);
  if (sigma_out)
    copy_dev<double>(st, sigma_out, (const double*)w.sigma, (size_t)nk);
  DMRG_LAUNCH_CHECK();
  return res;
}

template <class T>
int heig(hipStream_t st, SvdWork<T>& w, const T* H, int N, int ldh, double* evals, T* evecs,
         int lde, const SvdParams& prm) {
  DMRG_REQUIRE(N > 0 && ldh >= N && lde >= N, "heig: bad shape");
  const int Np = round_up(N, JP);
  DMRG_REQUIRE(Np <= w.cap, "heig: matrix exceeds workspace");
  copy_dev<T>(st, w.A[0], (size_t)Np, H, (size_t)ldh, (size_t)N, (size_t)N);
  JacobiCount jc;
  double off = 0.0;
  const int cur = jacobi<T>(st, w, N, Np, prm, jc, off);
  const int sweeps = jc.sweeps;
  hipLaunchKernelGGL(k_sort<T>, dim3(cdiv(N, 16)), dim3(256), 0, st, (const T*)w.A[cur], Np, nullptr, N, N,
                     w.perm, evals);
  hipLaunchKernelGGL(k_gather<T>, dim3(grid_for((size_t)N * N)), dim3(256), 0, st,
                     (const T*)w.V[cur], Np, (const int*)w.perm, N, N, 0, evecs, lde);
  DMRG_LAUNCH_CHECK();
  return sweeps;
}

// Orthonormal basis of the columns of Y (len x r, ld ldy), in place, by
// block classical Gram-Schmidt with reorthogonalisation (each block of
// MGS_B columns projected twice against the finished ones by GEMMs, then
// Cholesky QR twice, or MGS twice for an ill-conditioned block, inside the
// block).  Numerically dependent columns become zero.
// C (>= r * MGS_B scalars) and nrm (>= r doubles) are device work space.
template <class T>
void orth_columns(hipStream_t st, T* Y, int len, int r, int ldy, T* C, double* nrm) {
  const bool cx = is_cplx<T>::value;
  hipLaunchKernelGGL(k_vecnorm<T>, dim3(r), dim3(256), 0, st, (const T*)Y, ldy, len, 0, nrm);
  DMRG_LAUNCH_CHECK();
  for (int j0 = 0; j0 < r; j0 += MGS_B) {
    const int b = std::min(MGS_B, r - j0);
    T* Yj = Y + (size_t)j0 * ldy;
    for (int rep = 0; rep < 2 && j0 > 0; ++rep) {
      // C (j0 x b) = Q^H Y_j;  Y_j -= Q C
      gemm<T>(st, OP_H, OP_N, j0, b, len, T(1), Y, ldy, Yj, ldy, T(0), C, j0,
              gemm_pick_tile(j0, b, len, cx));
      gemm<T>(st, OP_N, OP_N, len, b, j0, T(-1), Y, ldy, C, j0, T(1), Yj, ldy,
              gemm_pick_tile(len, b, j0, cx));
    }
    hipLaunchKernelGGL(k_block_orth<T>, dim3(1), dim3(256), 0, st, Yj, len, ldy, b,
                       (const double*)(nrm + j0), 1e-13);
    DMRG_LAUNCH_CHECK();
  }
}

template <class T>
void RsvdWork<T>::alloc(int cap_len, int cap_r) {
  release();
  const size_t big = (size_t)cap_len * cap_r;
  DMRG_HIP(hipMalloc(&Om, big * sizeof(T)));
  DMRG_HIP(hipMalloc(&Y, big * sizeof(T)));
  DMRG_HIP(hipMalloc(&Z, big * sizeof(T)));
  DMRG_HIP(hipMalloc(&B, big * sizeof(T)));
  DMRG_HIP(hipMalloc(&Wb, big * sizeof(T)));
  DMRG_HIP(hipMalloc(&C, (size_t)cap_r * MGS_B * sizeof(T)));
  DMRG_HIP(hipMalloc(&nrm, (size_t)cap_r * sizeof(double)));
  len = cap_len;
  r = cap_r;
}

template <class T>
void RsvdWork<T>::release() {
  for (T** p : {&Om, &Y, &Z, &B, &Wb, &C}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  if (nrm) (void)hipFree(nrm);
  nrm = nullptr;
  len = r = 0;
}

template <class T>
SvdResult rsvd_split(hipStream_t st, SvdWork<T>& w, RsvdWork<T>& rw, const T* A, int m, int n,
                     int lda, int kmax, double cutoff, double rel_cutoff, bool scale_left, T* left,
                     int ldl, T* right, const SvdParams& prm, int oversample, int power,
                     uint64_t seed) {
  DMRG_REQUIRE(m > 0 && n > 0 && lda >= m && kmax >= 1 && oversample >= 0 && power >= 0,
               "rsvd: bad shape");
  const bool cx = is_cplx<T>::value;
  const int k = std::min(std::min(m, n), kmax);
  const int r = std::min(k + oversample, std::min(m, n));
  const int len_s = scale_left ? n : m;  // side of the orthonormal factor
  const int len_o = scale_left ? m : n;  // the other side
  DMRG_REQUIRE(std::max(m, n) <= rw.len && r <= rw.r, "rsvd: work space too small");
  auto tile = [&](int M, int N, int K) { return gemm_pick_tile(M, N, K, cx); };
  // Range sketch of the orthonormal side: Y = A Om ('U' side) or A^H Om
  fill_random<T>(st, rw.Om, (size_t)len_o * r, seed, 0, 1.0);
  auto apply = [&](bool herm, const T* X, T* out) {  // out = A X or A^H X
    if (!herm)
      gemm<T>(st, OP_N, OP_N, m, r, n, T(1), A, lda, X, n, T(0), out, m, tile(m, r, n));
    else
      gemm<T>(st, OP_H, OP_N, n, r, m, T(1), A, lda, X, m, T(0), out, n, tile(n, r, m));
  };
  apply(scale_left, rw.Om, rw.Y);
  orth_columns<T>(st, rw.Y, len_s, r, len_s, rw.C, rw.nrm);
  for (int q = 0; q < power; ++q) {  // subspace (power) iterations, re-orthonormalised
    apply(!scale_left, rw.Y, rw.Z);
    orth_columns<T>(st, rw.Z, len_o, r, len_o, rw.C, rw.nrm);
    apply(scale_left, rw.Z, rw.Y);
    orth_columns<T>(st, rw.Y, len_s, r, len_s, rw.C, rw.nrm);
  }
  // Project and split the small matrix exactly, then map its orthonormal
  // factor back: U = Q U_B ('U' side) or V^H = V_B^H Q^H ('V' side).
  SvdParams p = prm;
  p.subspace = false;
  SvdResult res;
  if (!scale_left) {
    // B (r x n) = Q^H A; B = U_B (S Vh): right gets S Vh directly
    gemm<T>(st, OP_H, OP_N, r, n, m, T(1), rw.Y, m, A, lda, T(0), rw.B, r, tile(r, n, m));
    res = svd_split<T>(st, w, rw.B, r, n, r, k, cutoff, rel_cutoff, false, rw.Wb, r, right,
                       nullptr, p);
    gemm<T>(st, OP_N, OP_N, m, res.nk, r, T(1), rw.Y, m, rw.Wb, r, T(0), left, ldl,
            tile(m, res.nk, r));
  } else {
    // B (m x r) = A Q; B = (U S) V_B^H: left gets U S directly
    gemm<T>(st, OP_N, OP_N, m, r, n, T(1), A, lda, rw.Y, n, T(0), rw.B, m, tile(m, r, n));
    res = svd_split<T>(st, w, rw.B, m, r, m, k, cutoff, rel_cutoff, true, left, ldl, rw.Wb,
                       nullptr, p);
    // right (nk x n) = V_B^H (nk x r) Q^H (r x n)
    gemm<T>(st, OP_N, OP_H, res.nk, n, r, T(1), rw.Wb, res.nk, rw.Y, n, T(0), right, res.nk,
            tile(res.nk, n, r));
  }
  return res;
}

template struct SvdWork<double>;
template struct SvdWork<cdouble>;
template SvdResult svd_split<double>(hipStream_t, SvdWork<double>&, const double*, int, int, int,
                                     int, double, double, bool, double*, int, double*, double*,
                                     const SvdParams&, const double*, double*);
template SvdResult svd_split<cdouble>(hipStream_t, SvdWork<cdouble>&, const cdouble*, int, int,
                                      int, int, double, double, bool, cdouble*, int, cdouble*,
                                      double*, const SvdParams&, const cdouble*, cdouble*);
template struct RsvdWork<double>;
template struct RsvdWork<cdouble>;
template SvdResult rsvd_split<double>(hipStream_t, SvdWork<double>&, RsvdWork<double>&,
                                      const double*, int, int, int, int, double, double, bool,
                                      double*, int, double*, const SvdParams&, int, int, uint64_t);
template SvdResult rsvd_split<cdouble>(hipStream_t, SvdWork<cdouble>&, RsvdWork<cdouble>&,
                                       const cdouble*, int, int, int, int, double, double, bool,
                                       cdouble*, int, cdouble*, const SvdParams&, int, int,
                                       uint64_t);
template int heig<double>(hipStream_t, SvdWork<double>&, const double*, int, int, double*, double*,
                          int, const SvdParams&);
template int heig<cdouble>(hipStream_t, SvdWork<cdouble>&, const cdouble*, int, int, double*,
                           cdouble*, int, const SvdParams&);

}  // namespace dmrg
