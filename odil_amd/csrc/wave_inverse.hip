// Inverse wave problem: recover the wave speed c(x) from sensor traces.
//
// Loss (the Problem._eval_loss sum of mean squares, util_op.py:628-654):
//   mean(f_wave^2) + mean(f_data^2) + k_reg^2 mean((dc/dx)^2)
// with f_wave the space-time residual of u_tt = c(x)^2 u_xx (WaveOp with a
// speed field; c = 1 is wave.py:67-122) and f_data = k_data (u - d) on the
// sensor columns (the reference's imposed-point equation form,
// ns_steady.py:396-411). One entry point returns the loss split and the
// gradients with respect to u and c:
//   grad_u: fused tile adjoint of the wave term (k_loss_grad) plus the
//           data term (k_winv_data);
//   grad_c: per-column sums over t of f_wave * df_wave/dc, in two stages
//           (k_winv_cgrad partials over 64-row chunks, then k_winv_final,
//           which also adds the regularisation and reduces the loss
//           partials in fixed order).
// A forward march (k_wave_march) solves f_wave = 0 row by row for the
// synthetic sensor data.
#include "wave_op.h"

namespace odil {

constexpr int kCgradRows = 64;  // rows per c-gradient partial

// Data misfit: grad_u += scale k_d f_d, per-block sums of f_d^2, where
// f_d = k_d (u - data[t][s]) on sensor columns (sensor_of[ix] = s >= 0).
template <class T>
__global__ void __launch_bounds__(256)
    k_winv_data(const T* __restrict__ u, T* __restrict__ gu,
                const int* __restrict__ sensor_of, const T* __restrict__ data,
                int ns, int nt, int nx, T kd, T scale,
                double* __restrict__ partial) {
  __shared__ double red[4];
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const size_t c = (size_t)b * 256 + threadIdx.x;
  double s2 = 0;
  if (c < (size_t)nt * nx) {
    const int it = (int)(c / nx), ix = (int)(c - (size_t)it * nx);
    const int s = sensor_of[ix];
    if (s >= 0) {
      const T fd = kd * (u[c] - data[(size_t)it * ns + s]);
      gu[c] += scale * kd * fd;
      s2 = (double)fd * (double)fd;
    }
  }
  s2 = block_sum(s2, red);
  if (threadIdx.x == 0) partial[b] = s2;
}

// pc[by][ix] = sum over rows t in [64 by, 64 by + 64), t >= 1, of
// f_wave(t, x) * df_wave/dc(x), df/dc = -2 c(x) u_xx(t - 1, x).
// Workgroup: 64 columns x 4 row lanes.
template <class T>
__global__ void __launch_bounds__(256)
    k_winv_cgrad(const T* __restrict__ u, WaveParams P, int nt, int nx,
                 double* __restrict__ pc) {
  __shared__ double sh[4][64];
  const int lx = threadIdx.x & 63, lr = threadIdx.x >> 6;
  const int ix = blockIdx.x * 64 + lx;
  const int t0 = blockIdx.y * kCgradRows;
  double acc = 0;
  if (ix < nx) {
    const T* cc = static_cast<const T*>(P.c);
    const T cx = cc[ix];
    const T dx = T(P.dx);
    const T* uf[1] = {u};
    for (int t = t0 + lr; t < min(t0 + kCgradRows, nt); t += 4) {
      if (t < 1) continue;
      T f[1];
      GlobalAcc<T, 1> acc1{uf, nt, nx, t, ix};
      WaveOp<T>::template residual<false>(acc1, t, ix, nt, nx, P, f);
      const T* row = u + (size_t)(t - 1) * nx;
      const T utm = row[ix];
      const T rxm = ix > 0 ? row[ix - 1]
                           : ghost(QUADH, row[ix + 1], utm,
                                   T(P.left_u[t - 1]));
      const T rxp = ix < nx - 1 ? row[ix + 1]
                                : ghost(QUADH, row[ix - 1], utm,
                                        T(P.right_u[t - 1]));
      const T uxx = (rxm - T(2) * utm + rxp) / (dx * dx);
      acc += (double)f[0] * (double)(T(-2) * cx * uxx);
    }
  }
  sh[lr][lx] = acc;
  __syncthreads();
  if (lr == 0 && ix < nx)
    pc[(size_t)blockIdx.y * nx + ix] =
        sh[0][lx] + sh[1][lx] + sh[2][lx] + sh[3][lx];
}

// One workgroup: grad_c = scale_w * sum_b pc[b][x] + regularisation
// gradient; split = [sum(pw)/N, sum(pd)/N, k_reg^2 mean((dc/dx)^2)], all
// sums in fixed order.
template <class T>
__global__ void __launch_bounds__(256)
    k_winv_final(const double* __restrict__ pw, int nbw,
                 const double* __restrict__ pd, int nbd,
                 const double* __restrict__ pc, int nby,
                 const T* __restrict__ c, T* __restrict__ gc, int nx,
                 double dx, double scale_w, double kreg2, double ncell,
                 double* __restrict__ split) {
  __shared__ double red[4];
  const double wr = kreg2 / ((nx - 1) * dx * dx);  // reg = wr sum dc^2
  double sr = 0;
  for (int ix = threadIdx.x; ix < nx; ix += 256) {
    double g = 0;
    for (int b = 0; b < nby; ++b) g += pc[(size_t)b * nx + ix];
    g *= scale_w;
    const double ci = (double)c[ix];
    if (ix > 0) g += 2 * wr * (ci - (double)c[ix - 1]);
    if (ix < nx - 1) {
      const double d = (double)c[ix + 1] - ci;
      g -= 2 * wr * d;
      sr += d * d;
    }
    gc[ix] = (T)g;
  }
  sr = block_sum(sr, red);
  double sw = 0, sd = 0;
  for (int b = threadIdx.x; b < nbw; b += 256) sw += pw[b];
  sw = block_sum(sw, red);
  for (int b = threadIdx.x; b < nbd; b += 256) sd += pd[b];
  sd = block_sum(sd, red);
  if (threadIdx.x == 0) {
    split[0] = sw / ncell;
    split[1] = sd / ncell;
    split[2] = wr * sr;
  }
}

// Solution of f_wave = 0 row by row (explicit):
//   u_0 = u0 + dt/2 ut0, u_1 = u_0 + dt ut0 + dt^2 c^2 u_xx(u_0),
//   u_t = 2 u_{t-1} - u_{t-2} + dt^2 c^2 u_xx(u_{t-1}).
// One workgroup; threads stride over x; a barrier separates rows.
template <class T>
__global__ void __launch_bounds__(1024)
    k_wave_march(WaveParams P, int nt, int nx, T* __restrict__ u) {
  const T* cc = static_cast<const T*>(P.c);
  const T dt = T(P.dt), dx = T(P.dx);
  for (int ix = threadIdx.x; ix < nx; ix += blockDim.x)
    u[ix] = T(P.init_u[ix] + 0.5 * P.dt * P.init_ut[ix]);
  __syncthreads();
  for (int t = 1; t < nt; ++t) {
    const T* row = u + (size_t)(t - 1) * nx;
    for (int ix = threadIdx.x; ix < nx; ix += blockDim.x) {
      const T utm = row[ix];
      const T rxm = ix > 0 ? row[ix - 1]
                           : ghost(QUADH, row[ix + 1], utm,
                                   T(P.left_u[t - 1]));
      const T rxp = ix < nx - 1 ? row[ix + 1]
                                : ghost(QUADH, row[ix - 1], utm,
                                        T(P.right_u[t - 1]));
      const T c2 = cc ? cc[ix] * cc[ix] : T(1);
      const T lap = dt * dt * c2 * (rxm - T(2) * utm + rxp) / (dx * dx);
      T v;
      if (t == 1)
        v = utm + dt * T(P.init_ut[ix]) + lap;
      else
        v = T(2) * utm - u[(size_t)(t - 2) * nx + ix] + lap;
      u[(size_t)t * nx + ix] = v;
    }
    __syncthreads();
  }
}

}  // namespace odil

using namespace odil;

namespace {
bool winv_bad(int nt, int nx, const double* a, const double* b,
              const double* c, const double* d, const void* speed) {
  if (nt < 3 || nx < 3 || !a || !b || !c || !d || !speed) {
    set_error("wave_inverse: need nt >= 3, nx >= 3, boundary arrays and c");
    return true;
  }
  return false;
}

int loss_blocks(int dtype, int nt, int nx) {
  return grad_nblocks(dtype == kF32, nt, nx);
}

int data_blocks(int nt, int nx) {
  return (int)(((size_t)nt * nx + 255) / 256);
}

template <class T>
int winv_run(int nt, int nx, const WaveParams& P, const int* sensor_of,
             const void* data, int ns, double kd, double kreg,
             const void* u, void* gu, void* gc, double* work, double* split,
             hipStream_t s) {
  const double ncell = (double)nt * nx;
  const int nbw = loss_blocks(sizeof(T) == 4 ? kF32 : kF64, nt, nx);
  const int nbd = data_blocks(nt, nx);
  const int nby = ceil_div(nt, kCgradRows);
  double* pw = work;
  double* pd = pw + nbw;
  double* pc = pd + nbd;
  const void* fields[1] = {u};
  void* grads[1] = {gu};
  const T scale = T(2.0 / ncell);
  int st;
  if (sizeof(T) == 4)
    st = launch_loss_grad_t<T, WaveOp<T>, kGradTx32, kGradTy32>(
        fields, grads, pw, nt, nx, scale, P, s, nullptr);
  else
    st = launch_loss_grad_t<T, WaveOp<T>, kGradTx64, kGradTy64>(
        fields, grads, pw, nt, nx, scale, P, s, nullptr);
  if (st) return st;
  hipLaunchKernelGGL(k_winv_data<T>, dim3(nbd), dim3(256), 0, s,
                     static_cast<const T*>(u), static_cast<T*>(gu),
                     sensor_of, static_cast<const T*>(data), ns, nt, nx,
                     T(kd), scale, pd);
  hipLaunchKernelGGL(k_winv_cgrad<T>, dim3(ceil_div(nx, 64), nby),
                     dim3(256), 0, s, static_cast<const T*>(u), P, nt, nx,
                     pc);
  hipLaunchKernelGGL(k_winv_final<T>, dim3(1), dim3(256), 0, s, pw, nbw, pd,
                     nbd, pc, nby, static_cast<const T*>(P.c),
                     static_cast<T*>(gc), nx, P.dx, 2.0 / ncell, kreg * kreg,
                     ncell, split);
  return check_launch("wave_inverse_loss_grad");
}
}  // namespace

extern "C" {

// Doubles of workspace odil_wave_inverse_loss_grad needs.
int odil_wave_inverse_work(int dtype, int nt, int nx) {
  if (nt < 1 || nx < 1) return 0;
  return loss_blocks(dtype, nt, nx) + data_blocks(nt, nx) +
         ceil_div(nt, kCgradRows) * nx;
}

int odil_wave_inverse_loss_grad(int dtype, int nt, int nx, double dt,
                                double dx, const double* init_u,
                                const double* init_ut, const double* left_u,
                                const double* right_u, const void* c,
                                const int* sensor_of, const void* data,
                                int nsensor, double k_data, double k_reg,
                                const void* u, void* grad_u, void* grad_c,
                                double* work, double* split, void* stream) {
  if (winv_bad(nt, nx, init_u, init_ut, left_u, right_u, c)) return kErrArg;
  if (!sensor_of || !data || nsensor < 1 || !u || !grad_u || !grad_c ||
      !work || !split) {
    set_error("wave_inverse_loss_grad: missing arrays");
    return kErrArg;
  }
  WaveParams P{dt, dx, init_u, init_ut, left_u, right_u, c};
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    return winv_run<double>(nt, nx, P, sensor_of, data, nsensor, k_data,
                            k_reg, u, grad_u, grad_c, work, split, s);
  if (dtype == kF32)
    return winv_run<float>(nt, nx, P, sensor_of, data, nsensor, k_data, k_reg,
                           u, grad_u, grad_c, work, split, s);
  set_error("wave_inverse_loss_grad: bad dtype %d", dtype);
  return kErrArg;
}

int odil_wave_march(int dtype, int nt, int nx, double dt, double dx,
                    const double* init_u, const double* init_ut,
                    const double* left_u, const double* right_u,
                    const void* c, void* u, void* stream) {
  if (nt < 2 || nx < 2 || !init_u || !init_ut || !left_u || !right_u || !u) {
    set_error("wave_march: bad arguments");
    return kErrArg;
  }
  WaveParams P{dt, dx, init_u, init_ut, left_u, right_u, c};
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int threads = nx < 1024 ? ((nx + 63) / 64) * 64 : 1024;
  if (dtype == kF64)
    hipLaunchKernelGGL(k_wave_march<double>, dim3(1), dim3(threads), 0, s, P,
                       nt, nx, static_cast<double*>(u));
  else if (dtype == kF32)
    hipLaunchKernelGGL(k_wave_march<float>, dim3(1), dim3(threads), 0, s, P,
                       nt, nx, static_cast<float*>(u));
  else {
    set_error("wave_march: bad dtype %d", dtype);
    return kErrArg;
  }
  return check_launch("wave_march");
}

}  // extern "C"
