// Block one-sided Jacobi SVD (see svd.hpp).
#include "svd.hpp"

#include <cmath>
#include <cstring>
#include <algorithm>
#include <vector>

namespace dmrg {
namespace {

constexpr int BW = 8;          // column block width
constexpr int PW = 2 * BW;     // columns per block pair
constexpr int ROWS = 256;      // rows per slice
constexpr int INNER_SWEEPS = 3;   // inner sweeps per block pair and round

// Round-robin tournament: block pair k of round r among nb blocks (nb even).
__device__ __host__ inline void rr_pair(int nb, int r, int k, int& a, int& b) {
  int n1 = nb - 1;
  if (k == 0) { a = r; b = n1; }
  else { a = (r + k) % n1; b = (r - k + n1) % n1; }
}
__device__ inline int pcol(int bi, int bj, int c) { return c < BW ? bi * BW + c : bj * BW + (c - BW); }

__device__ inline double absv(double a) { return fabs(a); }
__device__ inline double absv(cdouble a) { return sqrt(a.x * a.x + a.y * a.y); }
__device__ inline double unit_phase(double g, double) { return g >= 0.0 ? 1.0 : -1.0; }
__device__ inline cdouble unit_phase(cdouble g, double ag) { return cdouble(g.x / ag, g.y / ag); }

// W = A (m >= n) or A^H (m < n), zero-padded to qp columns; V = I.
template <class T>
__global__ void k_init(const T* __restrict__ A, int m, int n, int lda, int trans, T* __restrict__ W,
                       int p, int q, int qp, T* __restrict__ V) {
  size_t totW = (size_t)p * qp, totV = (size_t)qp * qp;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < totW + totV;
       i += (size_t)gridDim.x * blockDim.x) {
    if (i < totW) {
      int r = (int)(i % p), c = (int)(i / p);
      T v = T(0);
      if (c < q) v = trans ? conj(A[c + (size_t)r * lda]) : A[r + (size_t)c * lda];
      W[i] = v;
    } else {
      size_t j = i - totW;
      int r = (int)(j % qp), c = (int)(j / qp);
      V[j] = (r == c) ? T(1) : T(0);
    }
  }
}

// Partial 16x16 Gram of a block pair over one row slice: G[u][v] = w_u^H w_v.
template <class T>
__global__ __launch_bounds__(256) void k_gram(const T* __restrict__ W, int p, int nb, int round,
                                              int nsplit, int rows, T* __restrict__ gram) {
  const int k = blockIdx.x, s = blockIdx.y, t = threadIdx.x;
  int bi, bj;
  rr_pair(nb, round, k, bi, bj);
  __shared__ T tile[64][PW + 1];
  const int u = t % PW, v = t / PW;
  const int r0 = s * rows, r1 = min(p, r0 + rows);
  T acc = T(0);
  for (int rb = r0; rb < r1; rb += 64) {
    for (int e = t; e < 64 * PW; e += 256) {
      int rr = e % 64, c = e / 64, gr = rb + rr;
      tile[rr][c] = (gr < r1) ? W[gr + (size_t)pcol(bi, bj, c) * p] : T(0);
    }
    __syncthreads();
    const int nr = min(64, r1 - rb);
    for (int rr = 0; rr < nr; ++rr) acc += conj(tile[rr][u]) * tile[rr][v];
    __syncthreads();
  }
  gram[((size_t)k * nsplit + s) * (PW * PW) + u * PW + v] = acc;
}

struct Rot {
  double c, s;
  int p, q;
};

// Sum the Gram slices of a pair, diagonalise it by cyclic Jacobi (one wave),
// and apply the accumulated unitary to this slice's rows of W and V.
template <class T>
__global__ __launch_bounds__(64) void k_rotate(T* __restrict__ W, int p, T* __restrict__ V, int qp,
                                               int nb, int round, int nsplit, int rows,
                                               const T* __restrict__ gram, double tol,
                                               unsigned long long* __restrict__ off_max) {
  const int k = blockIdx.x, s = blockIdx.y, t = threadIdx.x;
  int bi, bj;
  rr_pair(nb, round, k, bi, bj);
  __shared__ T G[PW][PW + 1];
  __shared__ T Q[PW][PW + 1];
  __shared__ Rot rot[BW];
  __shared__ T ph[BW];
  __shared__ double red[64];
  __shared__ int any;

  for (int e = t; e < PW * PW; e += 64) {
    int u = e / PW, v = e % PW;
    T acc = T(0);
    for (int j = 0; j < nsplit; ++j) acc += gram[((size_t)k * nsplit + j) * (PW * PW) + u * PW + v];
    G[u][v] = acc;
    Q[u][v] = (u == v) ? T(1) : T(0);
  }
  __syncthreads();
  // Hermitian symmetrise: keep the upper triangle, real diagonal.
  for (int e = t; e < PW * PW; e += 64) {
    int u = e / PW, v = e % PW;
    if (u > v) G[u][v] = conj(G[v][u]);
    if (u == v) G[u][u] = T(real(G[u][u]));
  }
  __syncthreads();
  // Relative off-diagonal measure of this pair.
  double mx = 0.0;
  for (int e = t; e < PW * PW; e += 64) {
    int u = e / PW, v = e % PW;
    if (u < v) {
      double a = real(G[u][u]), b = real(G[v][v]);
      if (a > 0.0 && b > 0.0) mx = fmax(mx, absv(G[u][v]) / sqrt(a * b));
    }
  }
  red[t] = mx;
  __syncthreads();
  for (int w = 32; w > 0; w >>= 1) {
    if (t < w) red[t] = fmax(red[t], red[t + w]);
    __syncthreads();
  }
  const double off = red[0];
  if (s == 0 && t == 0) atomicMax(off_max, (unsigned long long)__double_as_longlong(off));
  if (!(off > tol)) return;  // already orthogonal (NaN falls through to rotate)

  // Cyclic Jacobi on the 16x16 Hermitian G, accumulating Q (G <- Q^H G Q).
  // A round rotates 8 disjoint index pairs; the two-sided update then splits
  // into 64 independent 2x2 blocks G[(pa,qa),(pb,qb)] <- Ja^H G Jb (one per
  // lane) plus the column rotation of Q, so a round needs two barriers.
  const double tin = tol * 0.25;
  for (int sw = 0; sw < INNER_SWEEPS; ++sw) {
    if (t == 0) any = 0;
    __syncthreads();
    for (int r = 0; r < PW - 1; ++r) {
      if (t < BW) {
        int a, b;
        rr_pair(PW, r, t, a, b);
        double al = real(G[a][a]), be = real(G[b][b]);
        T g = G[a][b];
        double ag = absv(g);
        Rot R{1.0, 0.0, a, b};
        T e = T(1);
        if (al > 0.0 && be > 0.0 && ag > tin * sqrt(al * be)) {
          double zeta = (be - al) / (2.0 * ag);
          double tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          double c = 1.0 / sqrt(1.0 + tt * tt);
          R.c = c;
          R.s = c * tt;
          e = unit_phase(g, ag);
          any = 1;
        }
        rot[t] = R;
        ph[t] = e;
      }
      __syncthreads();
      {
        // 2x2 block (ka, kb) of G: columns by Jb, then rows by Ja^H, where
        // J = [[c, s], [-s conj(e), c conj(e)]] on the (p, q) pair.
        const int ka = t / BW, kb = t % BW;
        const Rot Ra = rot[ka], Rb = rot[kb];
        if (Ra.s != 0.0 || Rb.s != 0.0) {
          const T ea = ph[ka], eb = conj(ph[kb]);
          T g00 = G[Ra.p][Rb.p], g01 = G[Ra.p][Rb.q], g10 = G[Ra.q][Rb.p], g11 = G[Ra.q][Rb.q];
          T x00 = Rb.c * g00 - Rb.s * (eb * g01), x01 = Rb.s * g00 + Rb.c * (eb * g01);
          T x10 = Rb.c * g10 - Rb.s * (eb * g11), x11 = Rb.s * g10 + Rb.c * (eb * g11);
          G[Ra.p][Rb.p] = Ra.c * x00 - Ra.s * (ea * x10);
          G[Ra.p][Rb.q] = Ra.c * x01 - Ra.s * (ea * x11);
          G[Ra.q][Rb.p] = Ra.s * x00 + Ra.c * (ea * x10);
          G[Ra.q][Rb.q] = Ra.s * x01 + Ra.c * (ea * x11);
        }
        // Q <- Q J (column rotations), 128 (row, pair) tasks over 64 lanes.
        for (int task = t; task < PW * BW; task += 64) {
          const int i = task / BW, kk = task % BW;
          const Rot R = rot[kk];
          if (R.s == 0.0) continue;
          const T ce = conj(ph[kk]);
          T xa = Q[i][R.p], xb = Q[i][R.q];
          Q[i][R.p] = R.c * xa - R.s * (ce * xb);
          Q[i][R.q] = R.s * xa + R.c * (ce * xb);
        }
      }
      __syncthreads();
    }
    const int more = any;
    __syncthreads();
    if (!more) break;
  }

  // Apply Q to the pair's columns of W (rows of this slice) and of V.
  const int r0 = s * rows;
  for (int pass = 0; pass < 2; ++pass) {
    T* M = pass ? V : W;
    const int ld = pass ? qp : p;
    const int r1 = min(ld, r0 + rows);
    for (int r = r0 + t; r < r1; r += 64) {
      T x[PW];
#pragma unroll
      for (int c = 0; c < PW; ++c) x[c] = M[r + (size_t)pcol(bi, bj, c) * ld];
#pragma unroll
      for (int c = 0; c < PW; ++c) {
        T acc = T(0);
#pragma unroll
        for (int u = 0; u < PW; ++u) acc += x[u] * Q[u][c];
        M[r + (size_t)pcol(bi, bj, c) * ld] = acc;
      }
    }
  }
}

template <class T>
__global__ void k_zero_u64(unsigned long long* p) { *p = 0ull; }

// sigma_c = |w_c| for c < q.
template <class T>
__global__ __launch_bounds__(256) void k_colnorm(const T* __restrict__ W, int p, double* sigma) {
  const int c = blockIdx.x, t = threadIdx.x;
  __shared__ double red[256];
  double acc = 0.0;
  for (int r = t; r < p; r += 256) acc += abs2(W[r + (size_t)c * p]);
  red[t] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) sigma[c] = sqrt(red[0]);
}

// Stable descending sort of q <= 2048 values (bitonic on (sigma, index)).
__global__ __launch_bounds__(1024) void k_sort(const double* sigma, int q, int* perm) {
  __shared__ double key[2048];
  __shared__ int idx[2048];
  const int t = threadIdx.x;
  for (int i = t; i < 2048; i += 1024) {
    key[i] = (i < q) ? sigma[i] : -1.0;
    idx[i] = i;
  }
  __syncthreads();
  for (int size = 2; size <= 2048; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = t; i < 2048; i += 1024) {
        int j = i ^ stride;
        if (j > i) {
          bool desc = ((i & size) == 0);
          // "a before b" when larger key, ties broken by smaller index.
          bool ib = (key[i] > key[j]) || (key[i] == key[j] && idx[i] < idx[j]);
          if (desc ? !ib : ib) {
            double tk = key[i]; key[i] = key[j]; key[j] = tk;
            int ti = idx[i]; idx[i] = idx[j]; idx[j] = ti;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = t; i < q; i += 1024) perm[i] = idx[i];
}

__global__ void k_rank(const double* sigma, const int* perm, int k, double cutoff, double rel,
                       int* newk) {
  double thr = fmax(cutoff, rel * sigma[perm[0]]);
  int nk = k;
  for (int i = 0; i < k; ++i) {
    double sv = sigma[perm[i]];
    if (sv < thr || sv != sv) { nk = i; break; }
  }
  if (nk == 0) nk = 1;
  *newk = nk;
}

// left (m x nk): U_x[i, j] (* sigma_j if scale_left)
template <class T>
__global__ void k_write_left(const T* __restrict__ W, int p, const T* __restrict__ V, int qp,
                             const double* __restrict__ sigma, const int* __restrict__ perm,
                             int trans, int m, int nk, int scale_left, T* __restrict__ out,
                             int ldo) {
  size_t tot = (size_t)m * nk;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    int i = (int)(e % m), j = (int)(e / m);
    int c = perm[j];
    double sg = sigma[c];
    T u;
    if (!trans) u = (sg > 0.0) ? W[i + (size_t)c * p] * (1.0 / sg) : T(0);
    else u = V[i + (size_t)c * qp];
    if (scale_left) u = u * sg;
    out[i + (size_t)j * ldo] = u;
  }
}

// right (nk x n): conj(V_x[col, j]) (* sigma_j unless scale_left)
template <class T>
__global__ void k_write_right(const T* __restrict__ W, int p, const T* __restrict__ V, int qp,
                              const double* __restrict__ sigma, const int* __restrict__ perm,
                              int trans, int n, int nk, int scale_left, T* __restrict__ out,
                              int ldo) {
  size_t tot = (size_t)nk * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    int j = (int)(e % nk), col = (int)(e / nk);
    int c = perm[j];
    double sg = sigma[c];
    T v;
    if (!trans) v = V[col + (size_t)c * qp];
    else v = (sg > 0.0) ? W[col + (size_t)c * p] * (1.0 / sg) : T(0);
    v = conj(v);
    if (!scale_left) v = v * sg;
    out[j + (size_t)col * ldo] = v;
  }
}

__global__ void k_gather_sigma(const double* sigma, const int* perm, int count, double* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = sigma[perm[i]];
}

inline int round_up(int a, int b) { return (a + b - 1) / b * b; }

}  // namespace

template <class T>
void SvdWork<T>::alloc(int cap_rows) {
  release();
  cap = cap_rows;
  capq = round_up(cap_rows, PW);
  int npairs = capq / PW;
  int nsplit = (capq + ROWS - 1) / ROWS;
  gram_elems = (size_t)npairs * nsplit * PW * PW;
  DMRG_HIP(hipMalloc(&W, (size_t)cap * capq * sizeof(T)));
  DMRG_HIP(hipMalloc(&V, (size_t)capq * capq * sizeof(T)));
  DMRG_HIP(hipMalloc(&gram, gram_elems * sizeof(T)));
  DMRG_HIP(hipMalloc(&sigma, capq * sizeof(double)));
  DMRG_HIP(hipMalloc(&perm, capq * sizeof(int)));
  DMRG_HIP(hipMalloc(&off, sizeof(unsigned long long)));
  DMRG_HIP(hipMalloc(&newk, sizeof(int)));
}

template <class T>
void SvdWork<T>::release() {
  if (W) (void)hipFree(W);
  if (V) (void)hipFree(V);
  if (gram) (void)hipFree(gram);
  if (sigma) (void)hipFree(sigma);
  if (perm) (void)hipFree(perm);
  if (off) (void)hipFree(off);
  if (newk) (void)hipFree(newk);
  W = V = gram = nullptr;
  sigma = nullptr;
  perm = nullptr;
  off = nullptr;
  newk = nullptr;
}

template <class T>
SvdResult svd_jacobi(hipStream_t st, SvdWork<T>& w, const T* A, int m, int n, int lda,
                     int max_sweeps) {
  DMRG_REQUIRE(m > 0 && n > 0 && lda >= m, "svd: bad shape");
  DMRG_REQUIRE(std::max(m, n) <= w.cap, "svd: matrix exceeds workspace");
  DMRG_REQUIRE(std::min(m, n) <= 2048, "svd: more than 2048 singular values");
  SvdResult res;
  res.transposed = (m < n);
  res.rows_p = res.transposed ? n : m;
  res.cols_q = res.transposed ? m : n;
  const int p = res.rows_p, q = res.cols_q;
  const int qp = round_up(q, PW);
  const int nb = qp / BW;
  const int npairs = nb / 2;
  const int nsplit = (std::max(p, qp) + ROWS - 1) / ROWS;
  DMRG_REQUIRE((size_t)npairs * nsplit * PW * PW <= w.gram_elems, "svd: gram workspace");

  size_t tot = (size_t)p * qp + (size_t)qp * qp;
  hipLaunchKernelGGL(k_init<T>, dim3(std::min<size_t>(cdiv(tot, 256), 4096)), dim3(256), 0, st, A,
                     m, n, lda, res.transposed ? 1 : 0, w.W, p, q, qp, w.V);
  DMRG_LAUNCH_CHECK();

  const double eps = 1.1102230246251565e-16;
  const double tol = std::max(1e-15, std::sqrt((double)p) * eps);
  res.sweeps = 0;
  res.off = 0.0;
  bool done = false;
  for (int sw = 0; sw < max_sweeps; ++sw) {
    hipLaunchKernelGGL(k_zero_u64<T>, dim3(1), dim3(1), 0, st, w.off);
    for (int r = 0; r < nb - 1; ++r) {
      hipLaunchKernelGGL(k_gram<T>, dim3(npairs, nsplit), dim3(256), 0, st, w.W, p, nb, r, nsplit,
                         ROWS, w.gram);
      hipLaunchKernelGGL(k_rotate<T>, dim3(npairs, nsplit), dim3(64), 0, st, w.W, p, w.V, qp, nb, r,
                         nsplit, ROWS, w.gram, tol, w.off);
    }
    DMRG_LAUNCH_CHECK();
    unsigned long long bits = 0;
    DMRG_HIP(hipMemcpyAsync(&bits, w.off, sizeof(bits), hipMemcpyDeviceToHost, st));
    DMRG_HIP(hipStreamSynchronize(st));
    double off;
    std::memcpy(&off, &bits, sizeof(off));
    res.sweeps = sw + 1;
    res.off = off;
    if (!(off == off)) throw Error(ST_NONFINITE, "svd: non-finite Gram entries");
    if (off <= tol) { done = true; break; }
  }
  if (!done) throw Error(ST_SOLVER_NO_CONV, "svd: Jacobi did not converge");
  hipLaunchKernelGGL(k_colnorm<T>, dim3(q), dim3(256), 0, st, w.W, p, w.sigma);
  hipLaunchKernelGGL(k_sort, dim3(1), dim3(1024), 0, st, w.sigma, q, w.perm);
  DMRG_LAUNCH_CHECK();
  return res;
}

template <class T>
int svd_rank(hipStream_t st, SvdWork<T>& w, const SvdResult& r, int k, double cutoff,
             double rel_cutoff) {
  DMRG_REQUIRE(k >= 1 && k <= r.cols_q, "svd_rank: bad k");
  hipLaunchKernelGGL(k_rank, dim3(1), dim3(1), 0, st, w.sigma, w.perm, k, cutoff, rel_cutoff,
                     w.newk);
  DMRG_LAUNCH_CHECK();
  int nk = 0;
  DMRG_HIP(hipMemcpyAsync(&nk, w.newk, sizeof(int), hipMemcpyDeviceToHost, st));
  DMRG_HIP(hipStreamSynchronize(st));
  return nk;
}

template <class T>
void svd_write(hipStream_t st, SvdWork<T>& w, const SvdResult& r, int m, int n, int nk,
               bool scale_left, T* left, int ldl, T* right, int ldr) {
  const int p = r.rows_p, qp = round_up(r.cols_q, PW);
  int tr = r.transposed ? 1 : 0;
  if (left) {
    size_t tot = (size_t)m * nk;
    hipLaunchKernelGGL(k_write_left<T>, dim3(std::min<size_t>(cdiv(tot, 256), 8192)), dim3(256), 0,
                       st, w.W, p, w.V, qp, w.sigma, w.perm, tr, m, nk, scale_left ? 1 : 0, left,
                       ldl);
  }
  if (right) {
    size_t tot = (size_t)nk * n;
    hipLaunchKernelGGL(k_write_right<T>, dim3(std::min<size_t>(cdiv(tot, 256), 8192)), dim3(256),
                       0, st, w.W, p, w.V, qp, w.sigma, w.perm, tr, n, nk, scale_left ? 1 : 0,
                       right, ldr);
  }
  DMRG_LAUNCH_CHECK();
}

template <class T>
void svd_sigma(hipStream_t st, SvdWork<T>& w, const SvdResult& r, int count, double* out) {
  DMRG_REQUIRE(count <= r.cols_q, "svd_sigma: count");
  hipLaunchKernelGGL(k_gather_sigma, dim3(cdiv(count, 256)), dim3(256), 0, st, w.sigma, w.perm,
                     count, out);
  DMRG_LAUNCH_CHECK();
}

template struct SvdWork<double>;
template struct SvdWork<cdouble>;
template SvdResult svd_jacobi<double>(hipStream_t, SvdWork<double>&, const double*, int, int, int,
                                      int);
template SvdResult svd_jacobi<cdouble>(hipStream_t, SvdWork<cdouble>&, const cdouble*, int, int,
                                       int, int);
template int svd_rank<double>(hipStream_t, SvdWork<double>&, const SvdResult&, int, double, double);
template int svd_rank<cdouble>(hipStream_t, SvdWork<cdouble>&, const SvdResult&, int, double,
                               double);
template void svd_write<double>(hipStream_t, SvdWork<double>&, const SvdResult&, int, int, int,
                                bool, double*, int, double*, int);
template void svd_write<cdouble>(hipStream_t, SvdWork<cdouble>&, const SvdResult&, int, int, int,
                                 bool, cdouble*, int, cdouble*, int);
template void svd_sigma<double>(hipStream_t, SvdWork<double>&, const SvdResult&, int, double*);
template void svd_sigma<cdouble>(hipStream_t, SvdWork<cdouble>&, const SvdResult&, int, double*);

}  // namespace dmrg
