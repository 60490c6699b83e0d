// Contractions of the sweep as MFMA GEMMs (see contract.hpp).
#include "contract.hpp"

#include <type_traits>

#include <cstring>

namespace dmrg {

namespace {
inline GemmTile pick(int M, int N, int K) {
  return gemm_pick_tile(M, N, K, false);
}
template <class T> inline GemmTile pickT(int M, int N, int K) {
  return gemm_pick_tile(M, N, K, is_cplx<T>::value);
}
}  // namespace

// Y = X W for the MPO-bond step (W column-major K x N).
template <class T>
void mpo_step(hipStream_t st, int M, int N, int K, const T* X, View2 vx, const T* W, T* Y,
              View2 vy) {
  if (gemm_rows_small<T>(st, M, N, K, X, vx, W, Y, vy)) return;
  gemm_tensor<T>(st, M, N, K, T(1), X, vx, false, false, W, view_colmajor(K), false, true, T(0), Y,
                 vy, pickT<T>(M, N, K));
}

template <class T>
void heff2_apply(hipStream_t st, const T* Lenv, const T* Renv, const T* WW, const T* theta, T* out,
                 int cL, int cR, int d, int D, ContractScratch<T>& sc) {
  const int dd = d * d;
  const size_t need = (size_t)D * dd * cL * cR;
  DMRG_REQUIRE(need <= sc.cap, "heff2: scratch too small");
  const T one(1), zero(0);
  // X[(w + D a'), (ss + dd b)] = sum_a L[a, w, a'] theta[a, ss, b], then
  // Y[a' + cL ss' + cL dd b + cL dd cR n] = sum_{w,ss} X * WW[(w,ss),(ss',n)]:
  // one fused launch where it applies (d = 2, real, not split-BF16)
  bool fused = false;
  if constexpr (!is_cplx<T>::value) {
    if (!sc.split_bf16) fused = heff2_x_mpo<T>(st, cL, cR, d, D, Lenv, theta, WW, sc.Y);
  }
  if (!fused) {
    if constexpr (std::is_same<T, float>::value) {
      if (sc.split_bf16) {
        gemm_f32_split_bf16(st, OP_T, OP_N, D * cL, dd * cR, cL, Lenv, cL, theta, cL, sc.X, D * cL);
      } else {
        gemm<T>(st, OP_T, OP_N, D * cL, dd * cR, cL, one, Lenv, cL, theta, cL, zero, sc.X, D * cL,
                pickT<T>(D * cL, dd * cR, cL));
      }
    } else {
      gemm<T>(st, OP_T, OP_N, D * cL, dd * cR, cL, one, Lenv, cL, theta, cL, zero, sc.X, D * cL,
              pickT<T>(D * cL, dd * cR, cL));
    }
    View2 vx{{cL, (long long)D, (long long)D * cL * dd}, {D, 1, (long long)D * cL}};
    View2 vy{{cL, 1, (long long)cL * dd}, {dd, (long long)cL, (long long)cL * dd * cR}};
    mpo_step<T>(st, cL * cR, dd * D, D * dd, sc.X, vx, WW, sc.Y, vy);
  }
  // out[(a' + cL ss'), b'] = sum_{(b, n)} Y * R[(b + cR n), b']
  bool done = false;
  if constexpr (std::is_same<T, float>::value) {
    if (sc.split_bf16) {
      gemm_f32_split_bf16(st, OP_N, OP_N, cL * dd, cR, D * cR, sc.Y, cL * dd, Renv, cR * D, out,
                          cL * dd);
      done = true;
    }
  }
  // X is free again here: it serves as the split-K workspace
  if (!done)
    done = gemm_nn_splitk<T>(st, cL * dd, cR, D * cR, sc.Y, cL * dd, Renv, cR * D, out, cL * dd,
                             sc.X, sc.cap);
  if (!done)
    gemm<T>(st, OP_N, OP_N, cL * dd, cR, D * cR, one, sc.Y, cL * dd, Renv, cR * D, zero, out,
            cL * dd, pickT<T>(cL * dd, cR, D * cR));
}

void heff2_apply_f32_io64(hipStream_t st, const float* Lenv, const float* Renv, const float* WW,
                          const double* theta, double* out, int cL, int cR, int d, int D,
                          ContractScratch<float>& sc, float* thf, float* outf, const double* sL,
                          const double* sR) {
  const int dd = d * d;
  const size_t n = (size_t)cL * dd * cR;
  const size_t need = (size_t)D * dd * cL * cR;
  DMRG_REQUIRE(need <= sc.cap, "heff2: scratch too small");
  // theta is converted once: every row tile of the first step reads it, and
  // reading it as FP64 there doubles that L2 traffic (measured: the C4
  // matvec 14 us slower than the separate conversion)
  convert_f64_f32(st, theta, thf, n);
  if (!sc.split_bf16 && heff2_x_mpo<float>(st, cL, cR, d, D, Lenv, thf, WW, sc.Y)) {
    // X is free: it serves as the split-K workspace
    if (gemm_nn_splitk_to_f64(st, cL * dd, cR, D * cR, sc.Y, cL * dd, Renv, cR * D, out, cL * dd,
                              sc.X, sc.cap, sL, sR))
      return;
    gemm<float>(st, OP_N, OP_N, cL * dd, cR, D * cR, 1.f, sc.Y, cL * dd, Renv, cR * D, 0.f, outf,
                cL * dd, pickT<float>(cL * dd, cR, D * cR));
    convert_f32_f64_scaled(st, outf, out, n, sL, sR);
    return;
  }
  heff2_apply<float>(st, Lenv, Renv, WW, thf, outf, cL, cR, d, D, sc);
  convert_f32_f64_scaled(st, outf, out, n, sL, sR);
}

template <class T>
void heff1_apply(hipStream_t st, const T* Lenv, const T* Renv, const T* W, const T* theta, T* out,
                 int cL, int cR, int d, int D, ContractScratch<T>& sc) {
  const size_t need = (size_t)D * d * cL * cR;
  DMRG_REQUIRE(need <= sc.cap, "heff1: scratch too small");
  const T one(1), zero(0);
  gemm<T>(st, OP_T, OP_N, D * cL, d * cR, cL, one, Lenv, cL, theta, cL, zero, sc.X, D * cL,
          pickT<T>(D * cL, d * cR, cL));
  View2 vx{{cL, (long long)D, (long long)D * cL * d}, {D, 1, (long long)D * cL}};
  View2 vy{{cL, 1, (long long)cL * d}, {d, (long long)cL, (long long)cL * d * cR}};
  mpo_step<T>(st, cL * cR, d * D, D * d, sc.X, vx, W, sc.Y, vy);
  gemm<T>(st, OP_N, OP_N, cL * d, cR, D * cR, one, sc.Y, cL * d, Renv, cR * D, zero, out, cL * d,
          pickT<T>(cL * d, cR, D * cR));
}

template <class T>
void env_left(hipStream_t st, const T* Lenv, const T* A, const T* W, T* Lnew, int ci, int co, int d,
              int D, ContractScratch<T>& sc) {
  const size_t need = (size_t)D * d * ci * co;
  DMRG_REQUIRE(need <= sc.cap, "env_left: scratch too small");
  const T one(1), zero(0);
  gemm<T>(st, OP_T, OP_N, D * ci, d * co, ci, one, Lenv, ci, A, ci, zero, sc.X, D * ci,
          pickT<T>(D * ci, d * co, ci));
  View2 vx{{ci, (long long)D, (long long)D * ci * d}, {D, 1, (long long)D * ci}};
  View2 vy{{ci, 1, (long long)ci * d}, {d, (long long)ci, (long long)ci * d * co}};
  mpo_step<T>(st, ci * co, d * D, D * d, sc.X, vx, W, sc.Y, vy);
  gemm<T>(st, OP_T, OP_C, co * D, co, ci * d, one, sc.Y, ci * d, A, ci * d, zero, Lnew, co * D,
          pickT<T>(co * D, co, ci * d));
}

template <class T>
void env_right(hipStream_t st, const T* Renv, const T* A, const T* Wr, T* Rnew, int ci, int co,
               int d, int D, ContractScratch<T>& sc) {
  const size_t need = (size_t)D * d * ci * co;
  DMRG_REQUIRE(need <= sc.cap, "env_right: scratch too small");
  const T one(1), zero(0);
  // X[(a + co s), (w' + D b')] = sum_b A[a, s, b] R[b, w', b']
  gemm<T>(st, OP_N, OP_N, co * d, D * ci, ci, one, A, co * d, Renv, ci, zero, sc.X, co * d,
          pickT<T>(co * d, D * ci, ci));
  // Y[a + co w + co D sp + co D d b'] = sum_{s, w'} X * Wr[(s, w'), (w, sp)]
  View2 vx{{co, 1, (long long)co * d * D}, {d, (long long)co, (long long)co * d}};
  View2 vy{{co, 1, (long long)co * D * d}, {D, (long long)co, (long long)co * D}};
  mpo_step<T>(st, co * ci, D * d, d * D, sc.X, vx, Wr, sc.Y, vy);
  // Rnew[(a + co w), a'] = sum_{(sp, b')} Y * conj(A[a', sp, b'])
  gemm<T>(st, OP_N, OP_H, co * D, co, d * ci, one, sc.Y, co * D, A, co, zero, Rnew, co * D,
          pickT<T>(co * D, co, d * ci));
}

template <class T>
void build_ww2(const T* W1, const T* W2, int d, int D, T* WW) {
  const int dd = d * d, K = D * dd;
  for (int i = 0; i < K * K; ++i) WW[i] = T(0);
  auto w = [&](const T* W, int a, int s, int sp, int b) {
    return W[a + (size_t)D * (s + (size_t)d * (sp + (size_t)d * b))];
  };
  for (int a = 0; a < D; ++a)
    for (int n = 0; n < D; ++n)
      for (int s1 = 0; s1 < d; ++s1)
        for (int s2 = 0; s2 < d; ++s2)
          for (int s1p = 0; s1p < d; ++s1p)
            for (int s2p = 0; s2p < d; ++s2p) {
              T v(0);
              for (int m = 0; m < D; ++m) v = v + w(W1, a, s1, s1p, m) * w(W2, m, s2, s2p, n);
              int row = a + D * (s1 + d * s2);
              int col = (s1p + d * s2p) + dd * n;
              WW[row + (size_t)K * col] = v;
            }
}

template <class T>
void build_wright(const T* W, int d, int D, T* Wr) {
  for (int w = 0; w < D; ++w)
    for (int s = 0; s < d; ++s)
      for (int sp = 0; sp < d; ++sp)
        for (int wp = 0; wp < D; ++wp)
          Wr[(s + d * wp) + (size_t)d * D * (w + D * sp)] =
              W[w + (size_t)D * (s + (size_t)d * (sp + (size_t)d * wp))];
}

namespace {
__global__ void k_f64_f32(const double* in, float* out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = (float)in[i];
}
__global__ void k_f32_f64(const float* in, double* out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = (double)in[i];
}
__device__ inline double unit_scale(double s) { return s > 0.0 ? s : 1.0; }
__global__ void k_zero_d(double* x) { *x = 0.0; }
__global__ __launch_bounds__(256) void k_absmax(const double* in, size_t n, unsigned long long* out) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  // four independent maxima keep four loads in flight per thread
  double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
  const size_t st = (size_t)gridDim.x * blockDim.x;
  size_t i = blockIdx.x * (size_t)blockDim.x + t;
  for (; i + 3 * st < n; i += 4 * st) {
    m0 = fmax(m0, fabs(in[i]));
    m1 = fmax(m1, fabs(in[i + st]));
    m2 = fmax(m2, fabs(in[i + 2 * st]));
    m3 = fmax(m3, fabs(in[i + 3 * st]));
  }
  for (; i < n; i += st) m0 = fmax(m0, fabs(in[i]));
  red[t] = fmax(fmax(m0, m1), fmax(m2, m3));
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) red[t] = fmax(red[t], red[t + o]);
    __syncthreads();
  }
  if (t == 0) {
    unsigned long long bits;
    double v = red[0];
    memcpy(&bits, &v, sizeof(v));
    atomicMax(out, bits);
  }
}
__global__ void k_f64_f32_scaled(const double* in, float* out, size_t n, const double* s) {
  const double inv = 1.0 / unit_scale(*s);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = (float)(in[i] * inv);
}
__global__ void k_f32_f64_scaled(const float* in, double* out, size_t n, const double* s1,
                                 const double* s2) {
  const double f = unit_scale(*s1) * unit_scale(*s2);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = (double)in[i] * f;
}
__global__ __launch_bounds__(256) void k_absmax_f32(const float* in, size_t n,
                                                   unsigned long long* out) {
  __shared__ float red[256];
  const int t = threadIdx.x;
  float m = 0.f;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + t; i < n; i += (size_t)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf(in[i]));
  red[t] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) red[t] = fmaxf(red[t], red[t + o]);
    __syncthreads();
  }
  if (t == 0) {
    unsigned long long bits;
    double v = (double)red[0];
    memcpy(&bits, &v, sizeof(v));
    atomicMax(out, bits);  // non-negative doubles order as their bit patterns
  }
}
__global__ void k_rescale_f32(float* e, size_t n, const double* maxv, const double* s_in,
                              double* s_out) {
  const double m = unit_scale(*maxv);
  const double inv = 1.0 / m;  // in double: m may be an FP32 denormal
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    e[i] = (float)((double)e[i] * inv);
  if (blockIdx.x == 0 && threadIdx.x == 0) *s_out = unit_scale(*s_in) * m;
}
__global__ void k_fill_d(double* s, size_t n, double v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    s[i] = v;
}
// Grid (x: row blocks of 256, y: columns, strided by gridDim.y).
template <class T>
__global__ void k_copy2d(T* __restrict__ dst, size_t ldd, const T* __restrict__ src, size_t lds,
                         size_t rows, size_t cols) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= rows) return;
  for (size_t j = blockIdx.y; j < cols; j += gridDim.y) dst[i + j * ldd] = src[i + j * lds];
}
template <class T>
__global__ void k_unit(T* env, int D, int w) {
  int i = threadIdx.x;
  if (i < D) env[i] = (i == w) ? T(1) : T(0);
}
}  // namespace

void convert_f64_f32(hipStream_t st, const double* in, float* out, size_t n) {
  if (!n) return;
  hipLaunchKernelGGL(k_f64_f32, dim3(std::min<size_t>(cdiv(n, 256), 4096)), dim3(256), 0, st, in,
                     out, n);
  DMRG_LAUNCH_CHECK();
}
void convert_f32_f64(hipStream_t st, const float* in, double* out, size_t n) {
  if (!n) return;
  hipLaunchKernelGGL(k_f32_f64, dim3(std::min<size_t>(cdiv(n, 256), 4096)), dim3(256), 0, st, in,
                     out, n);
  DMRG_LAUNCH_CHECK();
}
void absmax_f64(hipStream_t st, const double* in, size_t n, double* out) {
  hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(1), 0, st, out);
  if (n)
    // at most 256 workgroups: each ends in one atomic on the same word
    hipLaunchKernelGGL(k_absmax, dim3(std::min<size_t>(cdiv(n, 256), 256)), dim3(256), 0, st, in,
                       n, reinterpret_cast<unsigned long long*>(out));
  DMRG_LAUNCH_CHECK();
}
void convert_f64_f32_scaled(hipStream_t st, const double* in, float* out, size_t n,
                            const double* scale) {
  if (!n) return;
  hipLaunchKernelGGL(k_f64_f32_scaled, dim3(std::min<size_t>(cdiv(n, 256), 4096)), dim3(256), 0,
                     st, in, out, n, scale);
  DMRG_LAUNCH_CHECK();
}
void convert_f32_f64_scaled(hipStream_t st, const float* in, double* out, size_t n,
                            const double* s1, const double* s2) {
  if (!n) return;
  hipLaunchKernelGGL(k_f32_f64_scaled, dim3(std::min<size_t>(cdiv(n, 256), 4096)), dim3(256), 0,
                     st, in, out, n, s1, s2);
  DMRG_LAUNCH_CHECK();
}
void absmax_f32(hipStream_t st, const float* in, size_t n, double* out) {
  hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(1), 0, st, out);
  if (n)
    hipLaunchKernelGGL(k_absmax_f32, dim3(std::min<size_t>(cdiv(n, 256), 1024)), dim3(256), 0, st,
                       in, n, reinterpret_cast<unsigned long long*>(out));
  DMRG_LAUNCH_CHECK();
}
void rescale_env_f32(hipStream_t st, float* e, size_t n, const double* maxv, const double* s_in,
                     double* s_out) {
  // one block also covers n = 0 (the scale is still written)
  hipLaunchKernelGGL(k_rescale_f32, dim3(std::max<unsigned>(1, std::min<size_t>(cdiv(n, 256), 4096))),
                     dim3(256), 0, st, e, n, maxv, s_in, s_out);
  DMRG_LAUNCH_CHECK();
}
template <class T>
void copy_dev(hipStream_t st, T* dst, size_t ldd, const T* src, size_t lds, size_t rows,
              size_t cols) {
  const size_t tot = rows * cols;
  if (!tot) return;
  const unsigned gx = cdiv(rows, 256);
  const unsigned gy = (unsigned)std::min<size_t>(cols, std::max<size_t>(1, 8192 / gx));
  hipLaunchKernelGGL(k_copy2d<T>, dim3(gx, gy), dim3(256), 0, st, dst, ldd, src, lds, rows, cols);
  DMRG_LAUNCH_CHECK();
}
template void copy_dev<double>(hipStream_t, double*, size_t, const double*, size_t, size_t, size_t);
template void copy_dev<cdouble>(hipStream_t, cdouble*, size_t, const cdouble*, size_t, size_t,
                                size_t);
template void copy_dev<float>(hipStream_t, float*, size_t, const float*, size_t, size_t, size_t);

void fill_f64(hipStream_t st, double* s, size_t n, double v) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_d, dim3(std::min<size_t>(cdiv(n, 256), 1024)), dim3(256), 0, st, s, n,
                     v);
  DMRG_LAUNCH_CHECK();
}
template <class T>
void set_boundary_env(hipStream_t st, T* env, int D, int w) {
  hipLaunchKernelGGL(k_unit<T>, dim3(1), dim3(64), 0, st, env, D, w);
  DMRG_LAUNCH_CHECK();
}

#define DMRG_INST(T)                                                                          \
  template void heff2_apply<T>(hipStream_t, const T*, const T*, const T*, const T*, T*, int,  \
                               int, int, int, ContractScratch<T>&);                           \
  template void heff1_apply<T>(hipStream_t, const T*, const T*, const T*, const T*, T*, int,  \
                               int, int, int, ContractScratch<T>&);                           \
  template void env_left<T>(hipStream_t, const T*, const T*, const T*, T*, int, int, int, int, \
                            ContractScratch<T>&);                                             \
  template void env_right<T>(hipStream_t, const T*, const T*, const T*, T*, int, int, int,    \
                             int, ContractScratch<T>&);                                       \
  template void build_ww2<T>(const T*, const T*, int, int, T*);                               \
  template void build_wright<T>(const T*, int, int, T*);
DMRG_INST(double)
DMRG_INST(cdouble)
DMRG_INST(float)
#undef DMRG_INST
template void set_boundary_env<double>(hipStream_t, double*, int, int);
template void set_boundary_env<cdouble>(hipStream_t, cdouble*, int, int);
template void set_boundary_env<float>(hipStream_t, float*, int, int);

}  // namespace dmrg
