// Linear advection p_t + a . grad p = 0 on a (t, x, y) space-time grid with
// constant velocity a = (ax, ay), periodic in x and y (reference:
// advection_2d.py:103-223, problem='direct', periodic=1, central=0,
// implicit=0). First-order upwind on sign(a); the row it = 0 uses the
// initial field p_init(x, y) as the previous time level (:200-202).
//
//   f(t,x,y) = (p - p_prev)/dt + ax Dx p + ay Dy p
//
// The operator is linear in p, so the loss gradient is (2/N) J^T f; the
// adjoint kernel gathers f from the 4 cells whose stencil contains c.
#include "common.h"

namespace odil {

struct AdvParams {
  int nt, nx, ny;
  double dt, dx, dy, ax, ay;
};

template <class T>
__device__ inline T adv_f(const T* __restrict__ p,
                          const T* __restrict__ pinit, const AdvParams& P,
                          int it, int ix, int iy) {
  const size_t plane = (size_t)P.nx * P.ny;
  const size_t c = it * plane + (size_t)ix * P.ny + iy;
  const T pc = p[c];
  const T pt = it == 0 ? pinit[(size_t)ix * P.ny + iy] : p[c - plane];
  const T dt = T(P.dt), dx = T(P.dx), dy = T(P.dy);
  T fx, fy;
  if (P.ax > 0) {
    const int xm = ix == 0 ? P.nx - 1 : ix - 1;
    fx = (pc - p[it * plane + (size_t)xm * P.ny + iy]) / dx;
  } else {
    const int xp = ix == P.nx - 1 ? 0 : ix + 1;
    fx = (p[it * plane + (size_t)xp * P.ny + iy] - pc) / dx;
  }
  if (P.ay > 0) {
    const int ym = iy == 0 ? P.ny - 1 : iy - 1;
    fy = (pc - p[it * plane + (size_t)ix * P.ny + ym]) / dy;
  } else {
    const int yp = iy == P.ny - 1 ? 0 : iy + 1;
    fy = (p[it * plane + (size_t)ix * P.ny + yp] - pc) / dy;
  }
  return (pc - pt) / dt + T(P.ax) * fx + T(P.ay) * fy;
}

template <class T>
__global__ void __launch_bounds__(256)
    k_adv_residual(const T* __restrict__ p, const T* __restrict__ pinit,
                   T* __restrict__ res, double* __restrict__ partial,
                   AdvParams P) {
  __shared__ double red[4];
  const size_t N = (size_t)P.nt * P.nx * P.ny;
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const size_t c = (size_t)b * 256 + threadIdx.x;
  double s = 0;
  if (c < N) {
    const int iy = (int)(c % P.ny);
    const size_t r = c / P.ny;
    const int ix = (int)(r % P.nx);
    const int it = (int)(r / P.nx);
    const T f = adv_f(p, pinit, P, it, ix, iy);
    if (res) res[c] = f;
    s = (double)f * (double)f;
  }
  if (partial) {
    s = block_sum(s, red);
    if (threadIdx.x == 0) partial[b] = s;
  }
}

// grad(c) = scale * [ f(c) (1/dt + |ax|/dx + |ay|/dy)
//                     - f(c + t)/dt - |ax|/dx f(c + ex) - |ay|/dy f(c + ey) ]
// with ex = +x if ax > 0 else -x (periodic), ey likewise, f(c + t) only for
// it < nt - 1.
template <class T>
__global__ void __launch_bounds__(256)
    k_adv_loss_grad(const T* __restrict__ p, const T* __restrict__ pinit,
                    T* __restrict__ grad, double* __restrict__ partial,
                    AdvParams P, T scale) {
  __shared__ double red[4];
  const size_t N = (size_t)P.nt * P.nx * P.ny;
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const size_t c = (size_t)b * 256 + threadIdx.x;
  double s = 0;
  if (c < N) {
    const int iy = (int)(c % P.ny);
    const size_t r = c / P.ny;
    const int ix = (int)(r % P.nx);
    const int it = (int)(r / P.nx);
    const T dt = T(P.dt), dx = T(P.dx), dy = T(P.dy);
    const T ax = T(fabs(P.ax)), ay = T(fabs(P.ay));
    const T f = adv_f(p, pinit, P, it, ix, iy);
    s = (double)f * (double)f;
    T g = f * (T(1) / dt + ax / dx + ay / dy);
    if (it < P.nt - 1) g -= adv_f(p, pinit, P, it + 1, ix, iy) / dt;
    const int jx = P.ax > 0 ? (ix == P.nx - 1 ? 0 : ix + 1)
                            : (ix == 0 ? P.nx - 1 : ix - 1);
    const int jy = P.ay > 0 ? (iy == P.ny - 1 ? 0 : iy + 1)
                            : (iy == 0 ? P.ny - 1 : iy - 1);
    if (P.ax != 0) g -= ax / dx * adv_f(p, pinit, P, it, jx, iy);
    if (P.ay != 0) g -= ay / dy * adv_f(p, pinit, P, it, ix, jy);
    grad[c] = g * scale;
  }
  if (partial) {
    s = block_sum(s, red);
    if (threadIdx.x == 0) partial[b] = s;
  }
}

}  // namespace odil

using namespace odil;

namespace {
bool adv_bad(int nt, int nx, int ny, const void* p, const void* pinit) {
  if (nt < 1 || nx < 2 || ny < 2 || !p || !pinit) {
    set_error("advection: need nt >= 1, nx, ny >= 2, p and p_init");
    return true;
  }
  return false;
}
}  // namespace

extern "C" {

int odil_advection_nblocks(int nt, int nx, int ny) {
  return ceil_div(nt * nx * ny, 256);
}

// p_init has the dtype of the field, shape (nx, ny).
int odil_advection_residual(int dtype, int nt, int nx, int ny, double dt,
                            double dx, double dy, double ax, double ay,
                            const void* p_init, const void* p, void* res,
                            double* partial, void* stream) {
  if (adv_bad(nt, nx, ny, p, p_init)) return kErrArg;
  AdvParams P{nt, nx, ny, dt, dx, dy, ax, ay};
  dim3 grid(ceil_div(nt * nx * ny, 256));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    hipLaunchKernelGGL(k_adv_residual<double>, grid, dim3(256), 0, s,
                       static_cast<const double*>(p),
                       static_cast<const double*>(p_init),
                       static_cast<double*>(res), partial, P);
  else if (dtype == kF32)
    hipLaunchKernelGGL(k_adv_residual<float>, grid, dim3(256), 0, s,
                       static_cast<const float*>(p),
                       static_cast<const float*>(p_init),
                       static_cast<float*>(res), partial, P);
  else {
    set_error("advection_residual: bad dtype %d", dtype);
    return kErrArg;
  }
  return check_launch("advection_residual");
}

int odil_advection_loss_grad(int dtype, int nt, int nx, int ny, double dt,
                             double dx, double dy, double ax, double ay,
                             const void* p_init, const void* p, void* grad,
                             double* partial, void* stream) {
  if (adv_bad(nt, nx, ny, p, p_init) || !grad) return kErrArg;
  AdvParams P{nt, nx, ny, dt, dx, dy, ax, ay};
  dim3 grid(ceil_div(nt * nx * ny, 256));
  const double scale = 2.0 / ((double)nt * nx * ny);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    hipLaunchKernelGGL(k_adv_loss_grad<double>, grid, dim3(256), 0, s,
                       static_cast<const double*>(p),
                       static_cast<const double*>(p_init),
                       static_cast<double*>(grad), partial, P, scale);
  else if (dtype == kF32)
    hipLaunchKernelGGL(k_adv_loss_grad<float>, grid, dim3(256), 0, s,
                       static_cast<const float*>(p),
                       static_cast<const float*>(p_init),
                       static_cast<float*>(grad), partial, P, (float)scale);
  else {
    set_error("advection_loss_grad: bad dtype %d", dtype);
    return kErrArg;
  }
  return check_launch("advection_loss_grad");
}

}  // extern "C"
