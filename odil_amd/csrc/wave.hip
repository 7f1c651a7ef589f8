// 1D wave equation u_tt = u_xx on a (t, x) space-time grid: residual, stencil
// Jacobian and fused loss adjoint (reference: wave.py:67-122); the operator
// is WaveOp in wave_op.h.
#include "wave_op.h"


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
  WaveParams P{dt, dx, init_u, init_ut, left_u, right_u, nullptr};
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
  WaveParams P{dt, dx, init_u, init_ut, left_u, right_u, nullptr};
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    return launch_jac<double, WaveOp<double>>(fields, coef, res, nt, nx, P, s);
  if (dtype == kF32)
    return launch_jac<float, WaveOp<float>>(fields, coef, res, nt, nx, P, s);
  set_error("wave_jac: bad dtype %d", dtype);
  return kErrArg;
}

int odil_wave_jac_entries(int* ent, int* shifts) {
  if (ent)
    for (int q = 0; q < WaveOp<double>::NJ; ++q) {
      ent[3 * q] = WaveOp<double>::JE[q];
      ent[3 * q + 1] = WaveOp<double>::JV[q];
      ent[3 * q + 2] = WaveOp<double>::JS[q];
    }
  if (shifts)
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
  WaveParams P{dt, dx, init_u, init_ut, left_u, right_u, nullptr};
  hipStream_t s = static_cast<hipStream_t>(stream);
  const double scale = 2.0 / ((double)nt * nx);
  if (dtype == kF64)
    return launch_loss_grad_t<double, WaveOp<double>, kGradTx64, kGradTy64>(
        fields, grad, partial, nt, nx, scale, P, s, nblocks);
  if (dtype == kF32)
    return launch_loss_grad_t<float, WaveOp<float>, kGradTx32, kGradTy32>(
        fields, grad, partial, nt, nx, (float)scale, P, s, nblocks);
  set_error("wave_loss_grad: bad dtype %d", dtype);
  return kErrArg;
}

int odil_wave_nblocks(int kind, int dtype, int nt, int nx) {
  if (kind == 0) return residual_nblocks(nt, nx, WaveOp<float>::R);
  return grad_nblocks(dtype == kF32, nt, nx);
}

int odil_wave_normal_apply(int dtype, int nt, int nx, double dt, double dx,
                           const double* init_u, const double* init_ut,
                           const double* left_u, const double* right_u,
                           const void* const* state, const void* const* x,
                           const void* const* b, const void* const* d,
                           void* const* y, double omega, int mode,
                           void* stream) {
  if (wave_bad(nt, nx, init_u, init_ut, left_u, right_u)) return kErrArg;
  WaveParams P{dt, dx, init_u, init_ut, left_u, right_u, nullptr};
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
  WaveParams P{dt, dx, init_u, init_ut, left_u, right_u, nullptr};
  return dispatch_jtf<WaveOp<double>, WaveOp<float>>(
      dtype, fields, out, partial, nt, nx, scale, P,
      static_cast<hipStream_t>(stream), nblocks);
}

}  // extern "C"
