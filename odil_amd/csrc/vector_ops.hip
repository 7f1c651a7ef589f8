// Vector kernels for the Krylov and Newton drivers: deterministic dot
// products and axpy forms whose scalar is a ratio of device values, so a
// whole solver iteration can run without host round trips (and be captured
// in a HIP graph).
#include "common.h"

namespace odil {

constexpr int kMaxVec = 4;

template <class T>
struct VP {
  T* p[kMaxVec];
};
template <class T>
struct CVP {
  const T* p[kMaxVec];
};

constexpr int kDotBlocks = 1024;

// Stage 1: per-block partial sums of sum_k x_k . y_k in FP64.
template <class T>
__global__ void __launch_bounds__(256)
    k_dot_partial(CVP<T> x, CVP<T> y, int nvar, size_t N,
                  double* __restrict__ partial) {
  __shared__ double red[4];
  double s = 0;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (int k = 0; k < nvar; ++k) {
    const T* a = x.p[k];
    const T* b = y.p[k];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += stride)
      s += (double)a[i] * (double)b[i];
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Stage 2: out[0] = sum of partials (fixed order), optionally also copies
// the previous value out[0] into out[1] first.
__global__ void __launch_bounds__(256)
    k_dot_final(const double* __restrict__ partial, int nb,
                double* __restrict__ out, int keep_prev) {
  __shared__ double red[4];
  double s = 0;
  for (int b = threadIdx.x; b < nb; b += 256) s += partial[b];
  s = block_sum(s, red);
  if (threadIdx.x == 0) {
    if (keep_prev) out[1] = out[0];
    out[0] = s;
  }
}

// y_k += sign * (num / den) * x_k  (den == 0 -> no update).
template <class T>
__global__ void __launch_bounds__(256)
    k_axpy_ratio(VP<T> y, CVP<T> x, int nvar, size_t N,
                 const double* __restrict__ num,
                 const double* __restrict__ den, double sign) {
  const double d = *den;
  const T a = d != 0.0 ? T(sign * (*num) / d) : T(0);
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (int k = 0; k < nvar; ++k)
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += stride)
      y.p[k][i] += a * x.p[k][i];
}

// x_k += (num / den) p_k and r_k -= (num / den) q_k in one pass: the two
// conjugate-gradient updates sharing one step length (den == 0 -> none).
// JZ: also y_k = omega r_k / dg_k with the updated r (k_jacobi_zero's
// arithmetic), the first sweep of the preconditioner cycle that follows.
template <class T, bool JZ>
__global__ void __launch_bounds__(256)
    k_axpy2_ratio(VP<T> x, CVP<T> p, VP<T> r, CVP<T> q, int nvar, size_t N,
                  const double* __restrict__ num,
                  const double* __restrict__ den, CVP<T> dg, VP<T> y,
                  T omega) {
  const double d = *den;
  const T a = d != 0.0 ? T(*num / d) : T(0);
  const T an = d != 0.0 ? T(-*num / d) : T(0);
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (int k = 0; k < nvar; ++k)
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += stride) {
      x.p[k][i] += a * p.p[k][i];
      const T rv = r.p[k][i] + an * q.p[k][i];
      r.p[k][i] = rv;
      if (JZ) y.p[k][i] = omega * rv / dg.p[k][i];
    }
}

// Flexible-CG pass after the preconditioner: z = zw converted to the
// Krylov precision T, per-block FP64 partial sums of r.z (partial[b]) and
// r.zo (partial[nb + b]) with zo the previous z, then zo = z. Same grid and
// summation order as k_dot_partial.
// COPY = false: the sums only (the caller keeps z and the previous z in
// two buffers used in turn).
template <class Tw, class T, bool COPY>
__global__ void __launch_bounds__(256)
    k_fcg_z(CVP<Tw> zw, CVP<T> r, VP<T> zo, int nvar, size_t N,
            double* __restrict__ partial) {
  __shared__ double red[4];
  double s = 0, t = 0;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (int k = 0; k < nvar; ++k) {
    const Tw* a = zw.p[k];
    const T* b = r.p[k];
    T* c = zo.p[k];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += stride) {
      const T z = (T)a[i];
      const double rv = (double)b[i];
      s += rv * (double)z;
      t += rv * (double)c[i];
      if (COPY) c[i] = z;
    }
  }
  s = block_sum(s, red);
  t = block_sum(t, red);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = s;
    partial[gridDim.x + blockIdx.x] = t;
  }
}

// Stage 2 of k_fcg_z: *rz_prev = *rz, *rz = sum, *rzo = sum (fixed order).
__global__ void __launch_bounds__(256)
    k_fcg_final(const double* __restrict__ partial, int nb,
                double* __restrict__ rz, double* __restrict__ rz_prev,
                double* __restrict__ rzo) {
  __shared__ double red[4];
  double s = 0, t = 0;
  for (int b = threadIdx.x; b < nb; b += 256) {
    s += partial[b];
    t += partial[nb + b];
  }
  s = block_sum(s, red);
  t = block_sum(t, red);
  if (threadIdx.x == 0) {
    *rz_prev = *rz;
    *rz = s;
    *rzo = t;
  }
}

// p_k = z_k + ((num - sub) / den) * p_k  (sub optional; den == 0 -> p = z).
template <class T>
__global__ void __launch_bounds__(256)
    k_xpay_ratio(VP<T> p, CVP<T> z, int nvar, size_t N,
                 const double* __restrict__ num,
                 const double* __restrict__ sub,
                 const double* __restrict__ den) {
  const double d = *den;
  const double n = *num - (sub ? *sub : 0.0);
  const T b = d != 0.0 ? T(n / d) : T(0);
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (int k = 0; k < nvar; ++k)
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += stride)
      p.p[k][i] = z.p[k][i] + b * p.p[k][i];
}

// One fused Adam step (AdamNativeOptimizer.run, optimizer.py:303-314):
//   m += (g - m)(1 - b1); v += (g^2 - v)(1 - b2); x -= alpha m / (sqrt v + eps)
// with alpha = lr sqrt(1 - b2^t) / (1 - b1^t) computed by the caller.
template <class T>
__global__ void __launch_bounds__(256)
    k_adam(T* __restrict__ x, const T* __restrict__ g, T* __restrict__ m,
           T* __restrict__ v, size_t N, T alpha, T b1, T b2, T eps) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += stride) {
    const T gi = g[i];
    T mi = m[i], vi = v[i];
    mi += (gi - mi) * (T(1) - b1);
    vi += (gi * gi - vi) * (T(1) - b2);
    m[i] = mi;
    v[i] = vi;
    x[i] -= mi * alpha / (sqrt(vi) + eps);
  }
}

// out = x + a y (host scalar a; out may alias x or y).
template <class T>
__global__ void __launch_bounds__(256)
    k_waxpy(T* out, const T* x, const T* y, size_t N, double a) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const T aa = T(a);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += stride)
    out[i] = x[i] + aa * y[i];
}

template <class T>
__global__ void __launch_bounds__(256)
    k_axpby(T* out, const T* x, const T* y, size_t N, double a, double b) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const T aa = T(a), bb = T(b);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += stride)
    out[i] = aa * x[i] + bb * y[i];
}

// One weighted-Jacobi sweep from a zero iterate, y_k = omega f_k / d_k
// (x + omega (f - A x) / d at x = 0, same rounding), optionally writing
// fw_k = f_k converted to the y precision: the first pre-smoothing sweep of
// a V-cycle level without the fill of the iterate, the operator product or,
// for an FP32 preconditioner, a separate conversion pass of its FP64 input.
template <class Tf, class Ty>
__global__ void __launch_bounds__(256)
    k_jacobi_zero(CVP<Tf> f, CVP<Ty> d, VP<Ty> y, VP<Ty> fw, int nvar,
                  size_t N, Ty omega) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const bool copy = fw.p[0] != nullptr;
  for (int k = 0; k < nvar; ++k)
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += stride) {
      const Ty fv = (Ty)f.p[k][i];
      if (copy) fw.p[k][i] = fv;
      y.p[k][i] = omega * fv / d.p[k][i];
    }
}

// One diagonal block of a right-looking blocked Cholesky factorisation of a
// dense SPD matrix (row-major, leading dimension lda): the block
// A[k0:k0+nb, k0:k0+nb] (already updated by the previous panels) is
// factored, A_kk = L L^T, L written back (the block's strict upper part
// zeroed), and W = L^{-1} (lower) written to w (nb x nb, row-major, leading
// dimension kCholNb). The panel and trailing updates around it are GEMMs.
// nb <= kCholNb.
constexpr int kCholNb = 64;

// Double-precision value of another lane (lane index uniform).
__device__ inline double readlane_f64(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// The factorisation and inversion run in one wavefront, in two phases.
// Factorisation: lane i keeps row i of L in registers (the column loop is
// unrolled so every register index is a constant); per column j the pivot
// comes from lane j by readlane, column j reaches every lane through LDS
// (broadcast reads) and each lane updates its own row. Inversion: L goes to
// LDS and lane c solves L w = e_c by forward substitution for column c of
// W = L^{-1}. No workgroup-wide barrier on either 64-step critical path.
__global__ void __launch_bounds__(64)
    k_chol_block_wave(double* __restrict__ A, int lda, int k0, int nb,
                      double* __restrict__ w) {
  __shared__ double col[kCholNb];
  __shared__ double Ls[kCholNb][kCholNb + 1];
  const int i = threadIdx.x;
  double Lr[kCholNb];
  const double* __restrict__ arow =
      A + (size_t)(k0 + (i < nb ? i : 0)) * lda + k0;
#pragma unroll
  for (int c = 0; c < kCholNb; ++c)
    Lr[c] = (i < nb && c <= i && c < nb) ? arow[c] : 0.0;
  // Columns j >= nb (a short last block) change nothing: lij = 0.
#pragma unroll
  for (int j = 0; j < kCholNb; ++j) {
    const bool live = j < nb;
    const double d = live ? sqrt(readlane_f64(Lr[j], j)) : 1.0;
    const double lij = (live && i > j) ? Lr[j] / d : 0.0;
    Lr[j] = (live && i > j) ? lij : ((live && i == j) ? d : Lr[j]);
    col[i] = lij;
    __syncthreads();
    // L[i][l] -= L[i][j] L[l][j] for l > j (entries right of the diagonal
    // are never read and are stored as zero).
#pragma unroll
    for (int l = j + 1; l < kCholNb; ++l) Lr[l] = fma(-lij, col[l], Lr[l]);
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < kCholNb; ++c) Ls[i][c] = (c <= i && i < nb) ? Lr[c] : 0.0;
  __syncthreads();
  // Column c of W: w_r = (delta_rc - sum_{k<r} L[r][k] w_k) / L[r][r]; rows
  // and columns past nb stay zero.
  const int cc = i;
  double wc[kCholNb];
#pragma unroll
  for (int r = 0; r < kCholNb; ++r) {
    double sacc = r == cc ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < r; ++k) sacc = fma(-Ls[r][k], wc[k], sacc);
    wc[r] = r < nb ? sacc / Ls[r][r] : 0.0;
  }
  if (i < nb) {
    double* __restrict__ orow = A + (size_t)(k0 + i) * lda + k0;
#pragma unroll
    for (int c = 0; c < kCholNb; ++c)
      if (c < nb) orow[c] = c <= i ? Lr[c] : 0.0;
#pragma unroll
    for (int r = 0; r < kCholNb; ++r)
      if (r < nb) w[r * kCholNb + cc] = wc[r];
  }
}

}  // namespace odil

using namespace odil;

namespace {
template <class T>
VP<T> mkv(void* const* a, int n) {
  VP<T> v{};
  for (int k = 0; k < n; ++k) v.p[k] = static_cast<T*>(a[k]);
  return v;
}
template <class T>
CVP<T> mkc(const void* const* a, int n) {
  CVP<T> v{};
  for (int k = 0; k < n; ++k) v.p[k] = static_cast<const T*>(a[k]);
  return v;
}
unsigned grid_for(size_t N) {
  size_t b = (N + 255) / 256;
  return (unsigned)(b < 2048 ? (b ? b : 1) : 2048);
}
}  // namespace

extern "C" {

// Number of FP64 partials `partial` must hold for odil_dot.
int odil_dot_partials(void) { return kDotBlocks; }

// out[0] = sum_k x_k . y_k (FP64). keep_prev != 0 moves the old out[0] to
// out[1] first (the previous r.z of conjugate gradients).
int odil_dot(int dtype, int nvar, long long N, const void* const* x,
             const void* const* y, double* partial, double* out,
             int keep_prev, void* stream) {
  if (nvar < 1 || nvar > kMaxVec || N < 0) {
    set_error("dot: bad nvar/N");
    return kErrArg;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    hipLaunchKernelGGL(k_dot_partial<double>, dim3(kDotBlocks), dim3(256), 0,
                       s, mkc<double>(x, nvar), mkc<double>(y, nvar), nvar,
                       (size_t)N, partial);
  else if (dtype == kF32)
    hipLaunchKernelGGL(k_dot_partial<float>, dim3(kDotBlocks), dim3(256), 0, s,
                       mkc<float>(x, nvar), mkc<float>(y, nvar), nvar,
                       (size_t)N, partial);
  else {
    set_error("dot: bad dtype");
    return kErrArg;
  }
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, s, partial,
                     kDotBlocks, out, keep_prev);
  return check_launch("dot");
}

int odil_axpy_ratio(int dtype, int nvar, long long N, void* const* y,
                    const void* const* x, const double* num,
                    const double* den, double sign, void* stream) {
  if (nvar < 1 || nvar > kMaxVec) {
    set_error("axpy_ratio: bad nvar");
    return kErrArg;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    hipLaunchKernelGGL(k_axpy_ratio<double>, dim3(grid_for(N)), dim3(256), 0,
                       s, mkv<double>(y, nvar), mkc<double>(x, nvar), nvar,
                       (size_t)N, num, den, sign);
  else if (dtype == kF32)
    hipLaunchKernelGGL(k_axpy_ratio<float>, dim3(grid_for(N)), dim3(256), 0,
                       s, mkv<float>(y, nvar), mkc<float>(x, nvar), nvar,
                       (size_t)N, num, den, sign);
  else {
    set_error("axpy_ratio: bad dtype");
    return kErrArg;
  }
  return check_launch("axpy_ratio");
}

int odil_axpy2_ratio(int dtype, int nvar, long long N, void* const* x,
                     const void* const* p, void* const* r,
                     const void* const* q, const double* num,
                     const double* den, void* stream) {
  if (nvar < 1 || nvar > kMaxVec || N < 0 || !num || !den) {
    set_error("axpy2_ratio: bad arguments");
    return kErrArg;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    hipLaunchKernelGGL((k_axpy2_ratio<double, false>), dim3(grid_for(N)),
                       dim3(256), 0, s, mkv<double>(x, nvar),
                       mkc<double>(p, nvar), mkv<double>(r, nvar),
                       mkc<double>(q, nvar), nvar, (size_t)N, num, den,
                       CVP<double>{}, VP<double>{}, 0.0);
  else if (dtype == kF32)
    hipLaunchKernelGGL((k_axpy2_ratio<float, false>), dim3(grid_for(N)),
                       dim3(256), 0, s, mkv<float>(x, nvar),
                       mkc<float>(p, nvar), mkv<float>(r, nvar),
                       mkc<float>(q, nvar), nvar, (size_t)N, num, den,
                       CVP<float>{}, VP<float>{}, 0.0f);
  else {
    set_error("axpy2_ratio: bad dtype");
    return kErrArg;
  }
  return check_launch("axpy2_ratio");
}

int odil_axpy2_jacobi_zero(int dtype, int nvar, long long N, void* const* x,
                           const void* const* p, void* const* r,
                           const void* const* q, const double* num,
                           const double* den, const void* const* d,
                           void* const* y, double omega, void* stream) {
  if (nvar < 1 || nvar > kMaxVec || N < 0 || !num || !den || !d || !y) {
    set_error("axpy2_jacobi_zero: bad arguments");
    return kErrArg;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    hipLaunchKernelGGL((k_axpy2_ratio<double, true>), dim3(grid_for(N)),
                       dim3(256), 0, s, mkv<double>(x, nvar),
                       mkc<double>(p, nvar), mkv<double>(r, nvar),
                       mkc<double>(q, nvar), nvar, (size_t)N, num, den,
                       mkc<double>(d, nvar), mkv<double>(y, nvar), omega);
  else if (dtype == kF32)
    hipLaunchKernelGGL((k_axpy2_ratio<float, true>), dim3(grid_for(N)),
                       dim3(256), 0, s, mkv<float>(x, nvar),
                       mkc<float>(p, nvar), mkv<float>(r, nvar),
                       mkc<float>(q, nvar), nvar, (size_t)N, num, den,
                       mkc<float>(d, nvar), mkv<float>(y, nvar),
                       (float)omega);
  else {
    set_error("axpy2_jacobi_zero: bad dtype");
    return kErrArg;
  }
  return check_launch("axpy2_jacobi_zero");
}

int odil_fcg_z(int dtype_w, int dtype, int nvar, long long N,
               const void* const* zw, const void* const* r, void* const* zo,
               double* partial, double* rz, double* rz_prev, double* rzo,
               void* stream) {
  if (nvar < 1 || nvar > kMaxVec || N < 0 || !partial || !rz || !rz_prev ||
      !rzo) {
    set_error("fcg_z: bad arguments");
    return kErrArg;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
#define ODIL_FCG_Z(TW, T)                                                    \
  hipLaunchKernelGGL((k_fcg_z<TW, T, true>), dim3(kDotBlocks), dim3(256), 0, \
                     s,                                                      \
                     mkc<TW>(zw, nvar), mkc<T>(r, nvar), mkv<T>(zo, nvar),   \
                     nvar, (size_t)N, partial)
  if (dtype_w == kF64 && dtype == kF64)
    ODIL_FCG_Z(double, double);
  else if (dtype_w == kF32 && dtype == kF64)
    ODIL_FCG_Z(float, double);
  else if (dtype_w == kF32 && dtype == kF32)
    ODIL_FCG_Z(float, float);
  else {
    set_error("fcg_z: unsupported dtypes %d, %d", dtype_w, dtype);
    return kErrArg;
  }
#undef ODIL_FCG_Z
  hipLaunchKernelGGL(k_fcg_final, dim3(1), dim3(256), 0, s, partial,
                     kDotBlocks, rz, rz_prev, rzo);
  return check_launch("fcg_z");
}

int odil_fcg_dots(int dtype, int nvar, long long N, const void* const* z,
                  const void* const* r, const void* const* zo,
                  double* partial, double* rz, double* rz_prev, double* rzo,
                  void* stream) {
  if (nvar < 1 || nvar > kMaxVec || N < 0 || !z || !r || !zo || !partial ||
      !rz || !rz_prev || !rzo) {
    set_error("fcg_dots: bad arguments");
    return kErrArg;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  // zo is only read (COPY = false).
  void* const* zo_rw = const_cast<void* const*>(zo);
  if (dtype == kF64)
    hipLaunchKernelGGL((k_fcg_z<double, double, false>), dim3(kDotBlocks),
                       dim3(256), 0, s, mkc<double>(z, nvar),
                       mkc<double>(r, nvar), mkv<double>(zo_rw, nvar), nvar,
                       (size_t)N, partial);
  else if (dtype == kF32)
    hipLaunchKernelGGL((k_fcg_z<float, float, false>), dim3(kDotBlocks),
                       dim3(256), 0, s, mkc<float>(z, nvar),
                       mkc<float>(r, nvar), mkv<float>(zo_rw, nvar), nvar,
                       (size_t)N, partial);
  else {
    set_error("fcg_dots: bad dtype %d", dtype);
    return kErrArg;
  }
  hipLaunchKernelGGL(k_fcg_final, dim3(1), dim3(256), 0, s, partial,
                     kDotBlocks, rz, rz_prev, rzo);
  return check_launch("fcg_dots");
}

// p = z + ((num - sub) / den) p; `sub` may be NULL (conjugate gradients:
// num = r.z; flexible CG: sub = r.z_old).
int odil_xpay_ratio(int dtype, int nvar, long long N, void* const* p,
                    const void* const* z, const double* num,
                    const double* sub, const double* den, void* stream) {
  if (nvar < 1 || nvar > kMaxVec) {
    set_error("xpay_ratio: bad nvar");
    return kErrArg;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    hipLaunchKernelGGL(k_xpay_ratio<double>, dim3(grid_for(N)), dim3(256), 0,
                       s, mkv<double>(p, nvar), mkc<double>(z, nvar), nvar,
                       (size_t)N, num, sub, den);
  else if (dtype == kF32)
    hipLaunchKernelGGL(k_xpay_ratio<float>, dim3(grid_for(N)), dim3(256), 0,
                       s, mkv<float>(p, nvar), mkc<float>(z, nvar), nvar,
                       (size_t)N, num, sub, den);
  else {
    set_error("xpay_ratio: bad dtype");
    return kErrArg;
  }
  return check_launch("xpay_ratio");
}

// One Adam step on N values in place (x, m, v updated; g read).
int odil_adam_step(int dtype, long long N, void* x, const void* g, void* m,
                   void* v, double alpha, double beta1, double beta2,
                   double eps, void* stream) {
  if (N < 0 || !x || !g || !m || !v) {
    set_error("adam_step: bad arguments");
    return kErrArg;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    hipLaunchKernelGGL(k_adam<double>, dim3(grid_for(N)), dim3(256), 0, s,
                       static_cast<double*>(x), static_cast<const double*>(g),
                       static_cast<double*>(m), static_cast<double*>(v),
                       (size_t)N, alpha, beta1, beta2, eps);
  else if (dtype == kF32)
    hipLaunchKernelGGL(k_adam<float>, dim3(grid_for(N)), dim3(256), 0, s,
                       static_cast<float*>(x), static_cast<const float*>(g),
                       static_cast<float*>(m), static_cast<float*>(v),
                       (size_t)N, (float)alpha, (float)beta1, (float)beta2,
                       (float)eps);
  else {
    set_error("adam_step: bad dtype %d", dtype);
    return kErrArg;
  }
  return check_launch("adam_step");
}

// y_k = omega f_k / d_k (k < nvar, N values each); fw_k = f_k in the y
// precision when fw is not NULL. dtype pairs (f, y): (FP64, FP64),
// (FP32, FP32), (FP64, FP32).
int odil_jacobi_zero(int dtype_f, int dtype_y, int nvar, long long N,
                     const void* const* f, const void* const* d,
                     void* const* y, void* const* fw, double omega,
                     void* stream) {
  if (nvar < 1 || nvar > kMaxVec || N < 0 || !f || !d || !y) {
    set_error("jacobi_zero: bad arguments");
    return kErrArg;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const VP<double> fw64 = fw ? mkv<double>(fw, nvar) : VP<double>{};
  const VP<float> fw32 = fw ? mkv<float>(fw, nvar) : VP<float>{};
  if (dtype_f == kF64 && dtype_y == kF64)
    hipLaunchKernelGGL((k_jacobi_zero<double, double>), dim3(grid_for(N)),
                       dim3(256), 0, s, mkc<double>(f, nvar),
                       mkc<double>(d, nvar), mkv<double>(y, nvar), fw64, nvar,
                       (size_t)N, omega);
  else if (dtype_f == kF32 && dtype_y == kF32)
    hipLaunchKernelGGL((k_jacobi_zero<float, float>), dim3(grid_for(N)),
                       dim3(256), 0, s, mkc<float>(f, nvar),
                       mkc<float>(d, nvar), mkv<float>(y, nvar), fw32, nvar,
                       (size_t)N, (float)omega);
  else if (dtype_f == kF64 && dtype_y == kF32)
    hipLaunchKernelGGL((k_jacobi_zero<double, float>), dim3(grid_for(N)),
                       dim3(256), 0, s, mkc<double>(f, nvar),
                       mkc<float>(d, nvar), mkv<float>(y, nvar), fw32, nvar,
                       (size_t)N, (float)omega);
  else {
    set_error("jacobi_zero: unsupported dtypes %d -> %d", dtype_f, dtype_y);
    return kErrArg;
  }
  return check_launch("jacobi_zero");
}

// Block size of odil_chol_block.
int odil_chol_block_size(void) { return kCholNb; }

// Factor and invert the diagonal block [k0, k0 + nb) of the dense SPD
// matrix A (FP64, row-major, n x n, leading dimension lda) in place; W gets
// the inverse of its Cholesky factor (kCholNb x kCholNb, row-major).
int odil_chol_block(double* A, int lda, int k0, int nb, double* W,
                    void* stream) {
  if (!A || !W || nb < 1 || nb > kCholNb || k0 < 0 || lda < k0 + nb) {
    set_error("chol_block: bad arguments");
    return kErrArg;
  }
  hipLaunchKernelGGL(k_chol_block_wave, dim3(1), dim3(64), 0,
                     static_cast<hipStream_t>(stream), A, lda, k0, nb, W);
  return check_launch("chol_block");
}

// out = x + a y on N values.
int odil_waxpy(int dtype, long long N, void* out, const void* x,
               const void* y, double a, void* stream) {
  if (N < 0 || !out || !x || !y) {
    set_error("waxpy: bad arguments");
    return kErrArg;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    hipLaunchKernelGGL(k_waxpy<double>, dim3(grid_for(N)), dim3(256), 0, s,
                       static_cast<double*>(out),
                       static_cast<const double*>(x),
                       static_cast<const double*>(y), (size_t)N, a);
  else if (dtype == kF32)
    hipLaunchKernelGGL(k_waxpy<float>, dim3(grid_for(N)), dim3(256), 0, s,
                       static_cast<float*>(out), static_cast<const float*>(x),
                       static_cast<const float*>(y), (size_t)N, a);
  else {
    set_error("waxpy: bad dtype %d", dtype);
    return kErrArg;
  }
  return check_launch("waxpy");
}

// out = a x + b y on N values.
int odil_axpby(int dtype, long long N, void* out, const void* x,
               const void* y, double a, double b, void* stream) {
  if (N < 0 || !out || !x || !y) {
    set_error("axpby: bad arguments");
    return kErrArg;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    hipLaunchKernelGGL(k_axpby<double>, dim3(grid_for(N)), dim3(256), 0, s,
                       static_cast<double*>(out),
                       static_cast<const double*>(x),
                       static_cast<const double*>(y), (size_t)N, a, b);
  else if (dtype == kF32)
    hipLaunchKernelGGL(k_axpby<float>, dim3(grid_for(N)), dim3(256), 0, s,
                       static_cast<float*>(out), static_cast<const float*>(x),
                       static_cast<const float*>(y), (size_t)N, a, b);
  else {
    set_error("axpby: bad dtype %d", dtype);
    return kErrArg;
  }
  return check_launch("axpby");
}

}  // extern "C"
