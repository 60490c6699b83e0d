// Lanczos vector kernels, tridiagonal eigensolver and the device generator.
#include "lanczos.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>

namespace dmrg {
namespace {

constexpr int NT = 256;

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

template <class T> struct Acc;
template <> struct Acc<double> {
  static constexpr int N = 1;
  __device__ static void put(double v, double* o) { o[0] = v; }
  __device__ static double get(const double* o) { return o[0]; }
};
template <> struct Acc<cdouble> {
  static constexpr int N = 2;
  __device__ static void put(cdouble v, double* o) { o[0] = v.x; o[1] = v.y; }
  __device__ static cdouble get(const double* o) { return cdouble(o[0], o[1]); }
};

__device__ inline void chunk_range(size_t n, size_t& b, size_t& e) {
  size_t chunk = (n + gridDim.x - 1) / gridDim.x;
  chunk = (chunk + NT - 1) / NT * NT;
  b = blockIdx.x * chunk;
  e = b + chunk < n ? b + chunk : n;
}

// Reductions across workgroups carry no fence or atomic: each producing
// kernel writes one partial per workgroup, and the kernel that consumes the
// sum re-reduces the partials itself (every workgroup, same fixed order, so
// all workgroups see bit-identical values).  The kernel boundary makes the
// partials visible.  A device-scope release inside a kernel would write back
// the L2 of every XCD once per wave.

// Sum of part[0..nb) over the workgroup; the result is returned to every thread.
__device__ inline double sum_parts(const double* __restrict__ part, int nb, double* sh) {
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += NT) s += part[b];
  s = wave_sum(s);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < NT / 64; ++i) r += sh[i];
  __syncthreads();  // sh is free again on return
  return r;
}

// out[j] = sum_b part[j*nb + b] for j < cnt (one wave per j); ends with a barrier.
__device__ inline void sum_parts_multi(const double* __restrict__ part, int nb, int cnt, double* out) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int j = w; j < cnt; j += NT / 64) {
    double s = 0.0;
    for (int b = lane; b < nb; b += 64) s += part[(size_t)j * nb + b];
    s = wave_sum(s);
    if (lane == 0) out[j] = s;
  }
  __syncthreads();
}

// Partial sums of G values per thread -> part[(j0 + jj) * nb + blockIdx.x], jj < cnt.
template <int G>
__device__ inline void write_parts(const double (&v)[G], int cnt, double* part, int j0, int nb,
                                   double* sh /*[G * NT/64]*/) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int jj = 0; jj < G; ++jj) {
    const double s = wave_sum(v[jj]);
    if (lane == 0) sh[jj * (NT / 64) + w] = s;
  }
  __syncthreads();
  if ((int)threadIdx.x < cnt) {
    double r = 0.0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) r += sh[threadIdx.x * (NT / 64) + i];
    part[(size_t)(j0 + threadIdx.x) * nb + blockIdx.x] = r;
  }
  __syncthreads();
}

// partN[blk] = sum |x_i|^2 over the workgroup's chunk.
template <class T>
__global__ __launch_bounds__(NT) void k_norm(const T* __restrict__ x, size_t n,
                                            double* __restrict__ partN) {
  __shared__ double sh[NT / 64];
  size_t b, e;
  chunk_range(n, b, e);
  // four independent accumulators keep four loads in flight per thread
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  size_t i = b + threadIdx.x;
  for (; i + 3 * NT < e; i += 4 * NT) {
    a0 += abs2(x[i]);
    a1 += abs2(x[i + NT]);
    a2 += abs2(x[i + 2 * NT]);
    a3 += abs2(x[i + 3 * NT]);
  }
  for (; i < e; i += NT) a0 += abs2(x[i]);
  const double v[1] = {(a0 + a1) + (a2 + a3)};
  write_parts<1>(v, 1, partN, 0, gridDim.x, sh);
}

__device__ inline double u_pm1(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1ull) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
}
template <class T> __device__ inline T rand_entry(uint64_t seed, uint64_t idx, double scale);
template <> __device__ inline double rand_entry<double>(uint64_t seed, uint64_t idx, double scale) {
  return u_pm1(seed, idx) * scale;
}
template <> __device__ inline cdouble rand_entry<cdouble>(uint64_t seed, uint64_t idx, double scale) {
  return cdouble(u_pm1(seed, 2 * idx) * scale, u_pm1(seed, 2 * idx + 1) * scale);
}

// Fill x from the stream; with partN given, only when sqrt(sum partN) < thr.
template <class T>
__global__ __launch_bounds__(NT) void k_fill_random(T* x, size_t n, uint64_t seed, uint64_t offset,
                                                   double scale, const double* partN, int nbp,
                                                   double thr) {
  __shared__ double sh[NT / 64];
  if (partN && !(sqrt(sum_parts(partN, nbp, sh)) < thr)) return;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    x[i] = rand_entry<T>(seed, offset + i, scale);
}

// nrm = sqrt(sum partN); scal[0] = nrm, scal[1] = 1/nrm (0 if nrm <= 1e-300),
// *beta_out = nrm (if given); dst = src / nrm (if dst given).
template <class T>
__global__ __launch_bounds__(NT) void k_scale(T* dst, const T* src, size_t n,
                                             const double* __restrict__ partN, int nbp,
                                             double* scal, double* beta_out) {
  __shared__ double sh[NT / 64];
  const double nrm = sqrt(sum_parts(partN, nbp, sh));
  const double s = (nrm > 1e-300) ? 1.0 / nrm : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[0] = nrm;
    scal[1] = s;
    if (beta_out) *beta_out = nrm;
  }
  if (!dst) return;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i] * s;
}

// partA[blk] = Re <v, w> over the workgroup's chunk.
template <class T>
__global__ __launch_bounds__(NT) void k_alpha(const T* __restrict__ v, const T* __restrict__ w,
                                             size_t n, double* __restrict__ partA) {
  __shared__ double sh[NT / 64];
  size_t b, e;
  chunk_range(n, b, e);
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  size_t i = b + threadIdx.x;
  for (; i + 3 * NT < e; i += 4 * NT) {
    a0 += real(conj(v[i]) * w[i]);
    a1 += real(conj(v[i + NT]) * w[i + NT]);
    a2 += real(conj(v[i + 2 * NT]) * w[i + 2 * NT]);
    a3 += real(conj(v[i + 3 * NT]) * w[i + 3 * NT]);
  }
  for (; i < e; i += NT) a0 += real(conj(v[i]) * w[i]);
  const double vv[1] = {(a0 + a1) + (a2 + a3)};
  write_parts<1>(vv, 1, partA, 0, gridDim.x, sh);
}

// out[it] = sum partA (one workgroup).
__global__ __launch_bounds__(NT) void k_alpha_value(const double* __restrict__ partA, int nbp,
                                                    double* alpha, int it) {
  __shared__ double sh[NT / 64];
  const double a = sum_parts(partA, nbp, sh);
  if (threadIdx.x == 0) alpha[it] = a;
}

// alpha[it] = sum partA;  w -= alpha v_it + beta[it-1] v_{it-1};
// partC[(j*NC + c)*nb + blk] = component c of <v_j, w> over the chunk, j <= it.
// Dots are taken GD vectors at a time; the first group rides on the update pass.
template <class T>
__global__ __launch_bounds__(NT) void k_sub_dots(T* __restrict__ w, const T* __restrict__ V,
                                                size_t n, int it, const double* __restrict__ partA,
                                                double* alpha, const double* beta,
                                                double* __restrict__ partC) {
  constexpr int NC = Acc<T>::N;
  constexpr int GD = 8;
  __shared__ double sh[GD * NC * (NT / 64)];
  const int nb = gridDim.x;
  const double a = sum_parts(partA, nb, sh);
  if (blockIdx.x == 0 && threadIdx.x == 0) alpha[it] = a;
  size_t b, e;
  chunk_range(n, b, e);
  const double bm = it > 0 ? beta[it - 1] : 0.0;
  const T* vi = V + (size_t)it * n;
  const T* vm = it > 0 ? V + (size_t)(it - 1) * n : vi;
  for (int j0 = 0; j0 <= it; j0 += GD) {
    const int cnt = min(GD, it + 1 - j0);
    T acc[GD];
#pragma unroll
    for (int jj = 0; jj < GD; ++jj) acc[jj] = T(0);
    for (size_t i = b + threadIdx.x; i < e; i += NT) {
      T x = w[i];
      if (j0 == 0) {
        x = x - a * vi[i];
        if (it > 0) x = x - bm * vm[i];
        w[i] = x;
      }
#pragma unroll
      for (int jj = 0; jj < GD; ++jj)
        if (jj < cnt) acc[jj] += conj(V[(size_t)(j0 + jj) * n + i]) * x;
    }
    double o[GD * NC];
#pragma unroll
    for (int jj = 0; jj < GD; ++jj) Acc<T>::put(acc[jj], o + jj * NC);
    write_parts<GD * NC>(o, cnt * NC, partC, j0 * NC, nb, sh);
  }
}

// cvec[g] = sum_b partC[g*nb + b] (doubles; one workgroup per g, fixed order).
__global__ __launch_bounds__(NT) void k_coeffs(const double* __restrict__ partC, int nb,
                                               double* __restrict__ cvec) {
  __shared__ double sh[NT / 64];
  const double s = sum_parts(partC + (size_t)blockIdx.x * nb, nb, sh);
  if (threadIdx.x == 0) cvec[blockIdx.x] = s;
}

// w -= sum_j c_j v_j (c from cvec, or re-reduced from partC when cvec is null);
// partN[blk] = |w|^2 over the chunk.
template <class T>
__global__ __launch_bounds__(NT) void k_orth_norm(T* __restrict__ w, const T* __restrict__ V,
                                                 size_t n, int it, const double* __restrict__ partC,
                                                 const T* cvec, double* __restrict__ partN) {
  constexpr int NC = Acc<T>::N;
  __shared__ double sh[NT / 64];
  __shared__ double cs[128 * NC];
  const int nb = gridDim.x;
  if (cvec) {
    for (int j = threadIdx.x; j <= it; j += NT) Acc<T>::put(cvec[j], cs + j * NC);
    __syncthreads();
  } else {
    sum_parts_multi(partC, nb, (it + 1) * NC, cs);
  }
  size_t b, e;
  chunk_range(n, b, e);
  double a0 = 0.0, a1 = 0.0;
  size_t i = b + threadIdx.x;
  for (; i + NT < e; i += 2 * NT) {
    T x = w[i], x2 = w[i + NT];
    for (int j = 0; j <= it; ++j) {
      const T c = Acc<T>::get(cs + j * NC);
      const T* vj = V + (size_t)j * n;
      x = x - c * vj[i];
      x2 = x2 - c * vj[i + NT];
    }
    w[i] = x;
    w[i + NT] = x2;
    a0 += abs2(x);
    a1 += abs2(x2);
  }
  for (; i < e; i += NT) {
    T x = w[i];
    for (int j = 0; j <= it; ++j) x = x - Acc<T>::get(cs + j * NC) * V[(size_t)j * n + i];
    w[i] = x;
    a0 += abs2(x);
  }
  const double v[1] = {a0 + a1};
  write_parts<1>(v, 1, partN, 0, nb, sh);
}

// out = sum_j y_j v_j; partN[blk] = |out|^2 over the chunk.
template <class T>
__global__ __launch_bounds__(NT) void k_ritz(T* __restrict__ out, const T* __restrict__ V, size_t n,
                                            int niter, const double* y, double* __restrict__ partN) {
  __shared__ double sh[NT / 64];
  __shared__ double ys[128];
  for (int j = threadIdx.x; j < niter; j += NT) ys[j] = y[j];
  __syncthreads();
  size_t b, e;
  chunk_range(n, b, e);
  double acc = 0.0;
  for (size_t i = b + threadIdx.x; i < e; i += NT) {
    T x = T(0);
    for (int j = 0; j < niter; ++j) x = x + ys[j] * V[(size_t)j * n + i];
    out[i] = x;
    acc += abs2(x);
  }
  const double v[1] = {acc};
  write_parts<1>(v, 1, partN, 0, gridDim.x, sh);
}

// ---------------------------------------------------------------------------
// Symmetric tridiagonal T(k): diagonal a[0..k), off-diagonal b[0..k-1).
// ---------------------------------------------------------------------------

__device__ inline int sturm_count(const double* a, const double* b2, int k, double x, double pivmin) {
  double q = a[0] - x;
  int c = q < 0.0;
  for (int i = 1; i < k; ++i) {
    if (fabs(q) < pivmin) q = -pivmin;
    q = (a[i] - x) - b2[i - 1] / q;
    c += q < 0.0;
  }
  return c;
}

// Lowest eigenvalue by 64-way multisection (one wave).
__device__ double tri_lowest(const double* a, const double* b, int k, double* b2) {
  // One wave: every lane holds the same bracket [lo, hi] (the same arithmetic
  // on the same data), evaluates the Sturm count at one of 64 interior points
  // per round, and a ballot gives the first point with a count >= 1.
  const int t = threadIdx.x;
  double lo = a[0], hi = a[0], mb = 0.0;
  for (int i = 0; i < k; ++i) {
    double r = (i > 0 ? fabs(b[i - 1]) : 0.0) + (i < k - 1 ? fabs(b[i]) : 0.0);
    lo = fmin(lo, a[i] - r);
    hi = fmax(hi, a[i] + r);
    if (i < k - 1) mb = fmax(mb, b[i] * b[i]);
  }
  {
    const double w = fmax(fabs(lo), fabs(hi));
    lo -= 2.0 * DBL_EPSILON * w + DBL_MIN;
    hi += 2.0 * DBL_EPSILON * w + DBL_MIN;
  }
  const double pm = DBL_MIN * fmax(1.0, mb);
  for (int i = t; i < k - 1; i += 64) b2[i] = b[i] * b[i];
  __syncthreads();
  for (int round = 0; round < 24; ++round) {
    const double w = fmax(fabs(lo), fabs(hi));
    if (hi - lo <= 2.0 * DBL_EPSILON * w + 4.0 * DBL_MIN) break;
    const double x = lo + (hi - lo) * (double)(t + 1) / 65.0;
    const unsigned long long m = __ballot(sturm_count(a, b2, k, x, pm) >= 1);
    const int l = m ? __ffsll((long long)m) - 1 : 64;
    const double nlo = (l == 0) ? lo : lo + (hi - lo) * (double)l / 65.0;
    const double nhi = (l == 64) ? hi : lo + (hi - lo) * (double)(l + 1) / 65.0;
    lo = nlo;
    hi = nhi;
  }
  return 0.5 * (lo + hi);
}

// Eigenvector of T for eigenvalue lam by inverse iteration (thread 0).
// Work arrays (5 x 128 doubles + 128 ints) live in LDS, passed by the caller.
__device__ void tri_vector(const double* a, const double* b, int k, double lam, double* y,
                           double* d, double* du, double* dl, double* du2, double* x, int* piv) {
  if (k == 1) { y[0] = 1.0; return; }
  double tnorm = 0.0;
  for (int i = 0; i < k; ++i)
    tnorm = fmax(tnorm, fabs(a[i]) + (i > 0 ? fabs(b[i - 1]) : 0.0) + (i < k - 1 ? fabs(b[i]) : 0.0));
  const double tiny = fmax(DBL_EPSILON * tnorm, DBL_MIN);
  for (int i = 0; i < k; ++i) {
    d[i] = a[i] - lam;
    if (i < k - 1) { du[i] = b[i]; dl[i] = b[i]; }
    du2[i] = 0.0;
  }
  for (int i = 0; i < k - 1; ++i) {
    if (fabs(d[i]) >= fabs(dl[i])) {
      if (fabs(d[i]) < tiny) d[i] = d[i] >= 0.0 ? tiny : -tiny;
      double f = dl[i] / d[i];
      dl[i] = f;
      d[i + 1] -= f * du[i];
      piv[i] = i;
    } else {
      double f = d[i] / dl[i];
      d[i] = dl[i];
      dl[i] = f;
      double tmp = du[i];
      du[i] = d[i + 1];
      d[i + 1] = tmp - f * d[i + 1];
      if (i < k - 2) {
        du2[i] = du[i + 1];
        du[i + 1] = -f * du[i + 1];
      }
      piv[i] = i + 1;
    }
  }
  if (fabs(d[k - 1]) < tiny) d[k - 1] = d[k - 1] >= 0.0 ? tiny : -tiny;
  for (int i = 0; i < k; ++i) x[i] = 1.0 + 0.25 * (double)i / (double)k;
  for (int iter = 0; iter < 3; ++iter) {
    for (int i = 0; i < k - 1; ++i) {
      if (piv[i] == i) {
        x[i + 1] -= dl[i] * x[i];
      } else {
        double tmp = x[i];
        x[i] = x[i + 1];
        x[i + 1] = tmp - dl[i] * x[i];
      }
    }
    x[k - 1] /= d[k - 1];
    if (k > 1) x[k - 2] = (x[k - 2] - du[k - 2] * x[k - 1]) / d[k - 2];
    for (int i = k - 3; i >= 0; --i) x[i] = (x[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) / d[i];
    double s = 0.0, mx = 0.0;
    for (int i = 0; i < k; ++i) s += x[i] * x[i], mx = fmax(mx, fabs(x[i]));
    // rescale first to avoid overflow of the next solve
    double inv = 1.0 / sqrt(s);
    for (int i = 0; i < k; ++i) x[i] *= inv;
    (void)mx;
  }
  // deterministic sign: largest-magnitude component positive
  int im = 0;
  for (int i = 1; i < k; ++i)
    if (fabs(x[i]) > fabs(x[im])) im = i;
  double sg = x[im] < 0.0 ? -1.0 : 1.0;
  for (int i = 0; i < k; ++i) y[i] = sg * x[i];
}

__global__ __launch_bounds__(64) void k_tri(const double* alpha, const double* beta, int k, int mode,
                                           double beta_tol, double eig_tol, double eig_rel,
                                           double* scal, int* status, double* y) {
  __shared__ double b2[128];
  __shared__ double as[128], bs[128];  // the tridiagonal, read once from HBM
  __shared__ int bidx;
  __shared__ double wd[128], wdu[128], wdl[128], wdu2[128], wx[128];
  __shared__ int wpiv[128];
  const int t = threadIdx.x;
  for (int i = t; i < k; i += 64) {
    as[i] = alpha[i];
    bs[i] = beta[i];
  }
  __syncthreads();
  alpha = as;
  beta = bs;
  if (mode == 0) {  // check
    if (t == 0) {
      bidx = -1;
      for (int j = 0; j < k; ++j)
        if (beta[j] < beta_tol) { bidx = j; break; }
    }
    __syncthreads();
    if (bidx >= 0) {
      if (t == 0) { status[0] = 1; status[1] = bidx + 1; }
      return;
    }
    double lam = tri_lowest(alpha, beta, k, b2);
    if (t == 0) {
      scal[4] = lam;
      double prev = scal[2];
      if (fabs(lam - prev) < fmax(eig_tol, eig_rel * fabs(lam))) { status[0] = 1; status[1] = k; }
      else { status[0] = 0; scal[2] = lam; }
    }
  } else {  // final
    double lam = (k == 1) ? alpha[0] : tri_lowest(alpha, beta, k, b2);
    if (t == 0) {
      scal[3] = lam;
      tri_vector(alpha, beta, k, lam, y, wd, wdu, wdl, wdu2, wx, wpiv);
    }
  }
}

__global__ void k_lz_reset(double* scal, int* status) {
  scal[2] = 1e30;
  status[0] = 0;
  status[1] = 0;
}

}  // namespace

int lz_blocks(size_t n) {
  // one workgroup per 1024 elements up to 1024 workgroups (C4, n = 2^20: the
  // projections of iteration 9 take 28 us against 35 with 512 workgroups)
  size_t nb = (n + 1023) / 1024;
  if (nb < 1) nb = 1;
  if (nb > 1024) nb = 1024;
  return (int)nb;
}

size_t lz_part_doubles(int kmax, size_t n) { return (size_t)(2 * kmax + 2) * lz_blocks(n); }

void LzDev::alloc(int kmax_, size_t part_bytes_, size_t elem_bytes) {
  release();
  kmax = kmax_;
  part_bytes = part_bytes_;
  DMRG_HIP(hipMalloc(&alpha, kmax * sizeof(double)));
  DMRG_HIP(hipMalloc(&beta, kmax * sizeof(double)));
  DMRG_HIP(hipMalloc(&y, kmax * sizeof(double)));
  DMRG_HIP(hipMalloc(&scal, 8 * sizeof(double)));
  DMRG_HIP(hipMalloc(&status, 4 * sizeof(int)));
  DMRG_HIP(hipMalloc(&part, part_bytes));
  DMRG_HIP(hipMalloc(&cvec, kmax * elem_bytes));
  DMRG_HIP(hipMemset(alpha, 0, kmax * sizeof(double)));
  DMRG_HIP(hipMemset(beta, 0, kmax * sizeof(double)));
}

void LzDev::release() {
  for (void* p : {(void*)alpha, (void*)beta, (void*)y, (void*)scal, (void*)status, part, cvec})
    if (p) (void)hipFree(p);
  alpha = beta = y = scal = nullptr;
  status = nullptr;
  part = cvec = nullptr;
}

namespace {
// Partial-sum regions of LzDev::part (doubles): [0, nb) alpha, [nb, 2nb) norms,
// [2nb, ...) projection coefficients, (j*NC + c)*nb + blk.
inline double* partA(LzDev& d) { return (double*)d.part; }
inline double* partN(LzDev& d, int nb) { return (double*)d.part + nb; }
inline double* partC(LzDev& d, int nb) { return (double*)d.part + 2 * (size_t)nb; }
inline void need_parts(const LzDev& d, size_t n, int cnt) {
  const size_t need = (size_t)(cnt + 2) * lz_blocks(n) * sizeof(double);
  DMRG_REQUIRE(need <= d.part_bytes, "lanczos: partial buffer too small");
}
}  // namespace

template <class T>
void lz_norm(hipStream_t st, LzDev& d, const T* x, size_t n) {
  const int nb = lz_blocks(n);
  need_parts(d, n, 0);
  hipLaunchKernelGGL(k_norm<T>, dim3(nb), dim3(NT), 0, st, x, n, partN(d, nb));
  DMRG_LAUNCH_CHECK();
}

template <class T>
void lz_fill_if_small(hipStream_t st, LzDev& d, T* x, size_t n, uint64_t seed, double thr) {
  const int nb = lz_blocks(n);
  hipLaunchKernelGGL(k_fill_random<T>, dim3(std::min<size_t>(cdiv(n, NT), 2048)), dim3(NT), 0, st,
                     x, n, seed, 0ull, 1.0, (const double*)partN(d, nb), nb, thr);
  DMRG_LAUNCH_CHECK();
}

template <class T>
void fill_random(hipStream_t st, T* x, size_t n, uint64_t seed, uint64_t offset, double scale) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_random<T>, dim3(std::min<size_t>(cdiv(n, NT), 2048)), dim3(NT), 0, st,
                     x, n, seed, offset, scale, (const double*)nullptr, 0, 0.0);
  DMRG_LAUNCH_CHECK();
}

template <class T>
void lz_scale(hipStream_t st, LzDev& d, T* dst, const T* src, size_t n, double* beta_out) {
  const int nb = lz_blocks(n);
  hipLaunchKernelGGL(k_scale<T>, dim3(nb), dim3(NT), 0, st, dst, src, n,
                     (const double*)partN(d, nb), nb, d.scal, beta_out);
  DMRG_LAUNCH_CHECK();
}

template <class T>
void lz_next(hipStream_t st, LzDev& d, T* vnext, const T* w, size_t n, int it) {
  lz_scale<T>(st, d, vnext, w, n, d.beta + it);
}

template <class T>
void lz_alpha(hipStream_t st, LzDev& d, const T* v, const T* w, size_t n, int it) {
  (void)it;
  const int nb = lz_blocks(n);
  hipLaunchKernelGGL(k_alpha<T>, dim3(nb), dim3(NT), 0, st, v, w, n, partA(d));
  DMRG_LAUNCH_CHECK();
}

void lz_alpha_value(hipStream_t st, LzDev& d, size_t n, int it) {
  hipLaunchKernelGGL(k_alpha_value, dim3(1), dim3(NT), 0, st, (const double*)partA(d),
                     lz_blocks(n), d.alpha, it);
  DMRG_LAUNCH_CHECK();
}

template <class T>
void lz_sub_dots(hipStream_t st, LzDev& d, T* w, const T* V, size_t n, int it) {
  DMRG_REQUIRE(it < d.kmax && it < 128, "lanczos: Krylov dimension exceeds 128");
  need_parts(d, n, (it + 1) * Acc<T>::N);
  const int nb = lz_blocks(n);
  hipLaunchKernelGGL(k_sub_dots<T>, dim3(nb), dim3(NT), 0, st, w, V, n, it,
                     (const double*)partA(d), d.alpha, (const double*)d.beta, partC(d, nb));
  DMRG_LAUNCH_CHECK();
}

template <class T>
void lz_orth_norm(hipStream_t st, LzDev& d, T* w, const T* V, size_t n, int it) {
  const int nb = lz_blocks(n);
  const int cnt = (it + 1) * Acc<T>::N;
  // every workgroup re-reduces the coefficients while that stays small; past
  // that one workgroup per coefficient reduces them first (C4, n = 2^20, 1024
  // workgroups: 19 us against 24 at iteration 9, 30 against 43 at 19)
  const T* cv = nullptr;
  if ((size_t)cnt * nb > 1024) {
    hipLaunchKernelGGL(k_coeffs, dim3(cnt), dim3(NT), 0, st, (const double*)partC(d, nb), nb,
                       (double*)d.cvec);
    cv = (const T*)d.cvec;
  }
  hipLaunchKernelGGL(k_orth_norm<T>, dim3(nb), dim3(NT), 0, st, w, V, n, it,
                     (const double*)partC(d, nb), cv, partN(d, nb));
  DMRG_LAUNCH_CHECK();
}

void lz_check(hipStream_t st, LzDev& d, int ncheck, double beta_tol, double eig_tol,
              double eig_rel) {
  DMRG_REQUIRE(ncheck <= 128, "lanczos: tridiagonal larger than 128");
  hipLaunchKernelGGL(k_tri, dim3(1), dim3(64), 0, st, d.alpha, d.beta, ncheck, 0, beta_tol, eig_tol,
                     eig_rel, d.scal, d.status, d.y);
  DMRG_LAUNCH_CHECK();
}

void lz_final(hipStream_t st, LzDev& d, int niter) {
  DMRG_REQUIRE(niter >= 1 && niter <= 128, "lanczos: bad Krylov dimension");
  hipLaunchKernelGGL(k_tri, dim3(1), dim3(64), 0, st, d.alpha, d.beta, niter, 1, 0.0, 0.0, 0.0,
                     d.scal, d.status, d.y);
  DMRG_LAUNCH_CHECK();
}

void lz_reset(hipStream_t st, LzDev& d) {
  hipLaunchKernelGGL(k_lz_reset, dim3(1), dim3(1), 0, st, d.scal, d.status);
  DMRG_LAUNCH_CHECK();
}

template <class T>
void lz_ritz(hipStream_t st, LzDev& d, T* out, const T* V, size_t n, int niter) {
  DMRG_REQUIRE(niter <= 128, "lanczos: Krylov dimension exceeds 128");
  const int nb = lz_blocks(n);
  hipLaunchKernelGGL(k_ritz<T>, dim3(nb), dim3(NT), 0, st, out, V, n, niter, (const double*)d.y,
                     partN(d, nb));
  DMRG_LAUNCH_CHECK();
}

#define DMRG_INST(T)                                                                        \
  template void lz_norm<T>(hipStream_t, LzDev&, const T*, size_t);                          \
  template void lz_fill_if_small<T>(hipStream_t, LzDev&, T*, size_t, uint64_t, double);     \
  template void lz_scale<T>(hipStream_t, LzDev&, T*, const T*, size_t, double*);            \
  template void lz_next<T>(hipStream_t, LzDev&, T*, const T*, size_t, int);                 \
  template void lz_alpha<T>(hipStream_t, LzDev&, const T*, const T*, size_t, int);          \
  template void lz_sub_dots<T>(hipStream_t, LzDev&, T*, const T*, size_t, int);             \
  template void lz_orth_norm<T>(hipStream_t, LzDev&, T*, const T*, size_t, int);            \
  template void lz_ritz<T>(hipStream_t, LzDev&, T*, const T*, size_t, int);                 \
  template void fill_random<T>(hipStream_t, T*, size_t, uint64_t, uint64_t, double);
DMRG_INST(double)
DMRG_INST(cdouble)
#undef DMRG_INST

}  // namespace dmrg
