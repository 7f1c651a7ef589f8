// 1D wave equation u_tt = u_xx on a (t, x) space-time grid: residual, stencil
// Jacobian and fused loss adjoint (reference: wave.py:67-122).
//
// Axis 0 is t (nt cells), axis 1 is x (nx cells, contiguous). Row it = 0
// imposes u = u0 + dt/2 ut0; row it = 1 uses ut0 for the backward time
// difference; the x-neighbours at t-1 next to the walls come from quadratic
// extrapolation through the boundary value at t-1 (periodic index in t).
#include "stencil2d.h"

namespace odil {

struct WaveParams {
  double dt, dx;
  const double* init_u;   // [nx]
  const double* init_ut;  // [nx]
  const double* left_u;   // [nt], value at x = lower
  const double* right_u;  // [nt], value at x = upper
};

template <class T>
struct WaveOp {
  using Params = WaveParams;
  static constexpr int NF = 1, NE = 1, NS = 5, R = 2, RJ = 2;
  // Shifts (t, x): centre, t-1, t-2, (t-1, x-1), (t-1, x+1).
  static constexpr int SX[NS] = {0, -1, -2, -1, -1};
  static constexpr int SY[NS] = {0, 0, 0, -1, 1};
  static constexpr int NJ = 5;
  static constexpr int JE[NJ] = {0, 0, 0, 0, 0};
  static constexpr int JV[NJ] = {0, 0, 0, 0, 0};
  static constexpr int JS[NJ] = {0, 1, 2, 3, 4};

  template <class Acc>
  __device__ static inline void compute(const Acc& q, int it, int ix, int nt,
                                        int nx, const Params& P, T* f,
                                        T (*J)[NF][NS], bool want_jac) {
    const T dt = T(P.dt), dx = T(P.dx);
    const T u = q(0, 0, 0);
    if (it == 0) {
      f[0] = u - T(P.init_u[ix] + 0.5 * P.dt * P.init_ut[ix]);
      if (want_jac) {
        for (int s = 0; s < NS; ++s) J[0][0][s] = 0;
        J[0][0][0] = 1;
      }
      return;
    }
    const int itm = it - 1;  // previous time row (it >= 1 here)
    const T utm = q(0, -1, 0);
    const T rxm = q(0, -1, -1), rxp = q(0, -1, 1);
    const T lprev = T(P.left_u[itm]), rprev = T(P.right_u[itm]);
    const T uxm = (ix == 0) ? ghost(QUADH, rxp, utm, lprev) : rxm;
    const T uxp = (ix == nx - 1) ? ghost(QUADH, rxm, utm, rprev) : rxp;
    const T d_now = (u - utm) / dt;
    T d_prev;
    if (it == 1)
      d_prev = T(P.init_ut[ix]);
    else
      d_prev = (utm - q(0, -2, 0)) / dt;
    const T u_tt = (d_now - d_prev) / dt;
    const T u_xx = (uxm - T(2) * utm + uxp) / (dx * dx);
    f[0] = u_tt - u_xx;
    if (!want_jac) return;
    const T idt2 = T(1) / (dt * dt), idx2 = T(1) / (dx * dx);
    T c0 = idt2;
    T ctm = -idt2 + T(2) * idx2;
    T ctmm = 0;
    if (it >= 2) {
      ctm -= idt2;
      ctmm = idt2;
    }
    T cxm = -idx2, cxp = -idx2;
    if (ix == 0) {  // uxm is a halo: fold onto uxp and utm
      cxp += cxm * ghost_dfar<T>(QUADH);
      ctm += cxm * ghost_dnear<T>(QUADH);
      cxm = 0;
    }
    if (ix == nx - 1) {
      cxm += cxp * ghost_dfar<T>(QUADH);
      ctm += cxp * ghost_dnear<T>(QUADH);
      cxp = 0;
    }
    J[0][0][0] = c0;
    J[0][0][1] = ctm;
    J[0][0][2] = ctmm;
    J[0][0][3] = cxm;
    J[0][0][4] = cxp;
  }

  template <bool INT = false, class Acc>
  __device__ static inline void residual(const Acc& q, int i, int j, int n0,
                                         int n1, const Params& P, T* f) {
    compute(q, i, j, n0, n1, P, f, nullptr, false);
  }
  template <bool INT = false, class Acc>
  __device__ static inline void eval(const Acc& q, int i, int j, int n0,
                                     int n1, const Params& P, T* f,
                                     T (*J)[NF][NS]) {
    compute(q, i, j, n0, n1, P, f, J, true);
  }
};

}  // namespace odil

using namespace odil;

namespace {
bool wave_bad(int nt, int nx, const double* a, const double* b,
              const double* c, const double* d) {
  if (nt < 3 || nx < 2 || !a || !b || !c || !d) {
    set_error("wave: need nt >= 3, nx >= 2 and all boundary arrays");
    return true;
  }
  return false;
}
}  // namespace

extern "C" {

// Boundary arrays are FP64 device pointers: init_u, init_ut [nx],
// left_u, right_u [nt].
int odil_wave_residual(int dtype, int nt, int nx, double dt, double dx,
                       const double* init_u, const double* init_ut,
                       const double* left_u, const double* right_u,
                       const void* const* fields, void* const* res,
                       double* partial, void* stream) {
  if (wave_bad(nt, nx, init_u, init_ut, left_u, right_u)) return kErrArg;
  WaveParams P{dt, dx, init_u, init_ut, left_u, right_u};
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    return launch_residual<double, WaveOp<double>>(fields, res, partial, nt,
                                                   nx, P, s, nullptr);
  if (dtype == kF32)
    return launch_residual<float, WaveOp<float>>(fields, res, partial, nt, nx,
                                                 P, s, nullptr);
  set_error("wave_residual: bad dtype %d", dtype);
  return kErrArg;
}

int odil_wave_jac(int dtype, int nt, int nx, double dt, double dx,
                  const double* init_u, const double* init_ut,
                  const double* left_u, const double* right_u,
                  const void* const* fields, void* coef, void* const* res,
                  void* stream) {
  if (wave_bad(nt, nx, init_u, init_ut, left_u, right_u) || !coef)
    return kErrArg;
  WaveParams P{dt, dx, init_u, init_ut, left_u, right_u};
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    return launch_jac<double, WaveOp<double>>(fields, coef, res, nt, nx, P, s);
  if (dtype == kF32)
    return launch_jac<float, WaveOp<float>>(fields, coef, res, nt, nx, P, s);
  set_error("wave_jac: bad dtype %d", dtype);
  return kErrArg;
}

int odil_wave_jac_entries(int* ent, int* shifts) {
  for (int q = 0; q < WaveOp<double>::NJ; ++q) {
    ent[3 * q] = WaveOp<double>::JE[q];
    ent[3 * q + 1] = WaveOp<double>::JV[q];
    ent[3 * q + 2] = WaveOp<double>::JS[q];
  }
  for (int s = 0; s < WaveOp<double>::NS; ++s) {
    shifts[2 * s] = WaveOp<double>::SX[s];
    shifts[2 * s + 1] = WaveOp<double>::SY[s];
  }
  return WaveOp<double>::NJ;
}

int odil_wave_loss_grad(int dtype, int nt, int nx, double dt, double dx,
                        const double* init_u, const double* init_ut,
                        const double* left_u, const double* right_u,
                        const void* const* fields, void* const* grad,
                        double* partial, int* nblocks, void* stream) {
  if (wave_bad(nt, nx, init_u, init_ut, left_u, right_u) || !grad)
    return kErrArg;
  WaveParams P{dt, dx, init_u, init_ut, left_u, right_u};
  hipStream_t s = static_cast<hipStream_t>(stream);
  const double scale = 2.0 / ((double)nt * nx);
  if (dtype == kF64)
    return launch_loss_grad_t<double, WaveOp<double>, 8, 64>(
        fields, grad, partial, nt, nx, scale, P, s, nblocks);
  if (dtype == kF32)
    return launch_loss_grad_t<float, WaveOp<float>, 16, 64>(
        fields, grad, partial, nt, nx, (float)scale, P, s, nblocks);
  set_error("wave_loss_grad: bad dtype %d", dtype);
  return kErrArg;
}

int odil_wave_nblocks(int kind, int dtype, int nt, int nx) {
  if (kind == 0) return ceil_div(nx, 64) * ceil_div(nt, 4);
  return ceil_div(nx, 64) * ceil_div(nt, dtype == kF32 ? 16 : 8);
}

int odil_wave_normal_apply(int dtype, int nt, int nx, double dt, double dx,
                           const double* init_u, const double* init_ut,
                           const double* left_u, const double* right_u,
                           const void* const* state, const void* const* x,
                           const void* const* b, const void* const* d,
                           void* const* y, double omega, int mode,
                           void* stream) {
  if (wave_bad(nt, nx, init_u, init_ut, left_u, right_u)) return kErrArg;
  WaveParams P{dt, dx, init_u, init_ut, left_u, right_u};
  return dispatch_normal_apply<WaveOp<double>, WaveOp<float>>(
      dtype, state, x, b, d, y, nt, nx, omega, mode, P,
      static_cast<hipStream_t>(stream));
}

int odil_wave_jtf(int dtype, int nt, int nx, double dt, double dx,
                  const double* init_u, const double* init_ut,
                  const double* left_u, const double* right_u,
                  const void* const* fields, void* const* out, double scale,
                  double* partial, int* nblocks, void* stream) {
  if (wave_bad(nt, nx, init_u, init_ut, left_u, right_u) || !out)
    return kErrArg;
  WaveParams P{dt, dx, init_u, init_ut, left_u, right_u};
  return dispatch_jtf<WaveOp<double>, WaveOp<float>>(
      dtype, fields, out, partial, nt, nx, scale, P,
      static_cast<hipStream_t>(stream), nblocks);
}

}  // extern "C"
