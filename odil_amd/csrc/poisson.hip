// Laplace-type operators on a 2D cell-centred grid: residual, stencil
// Jacobian and fused loss adjoint.
//
//  Poisson : f = lap(u) - src, Dirichlet u = 0 on the walls with halos from
//            quadratic extrapolation through the wall value (the no-slip
//            halo rule of ns_steady.py:219-225 applied to a scalar).
//  Laplace : (lap(u), lap(v), 0) with linear-extrapolation halos
//            (operator_laplace, util_op.py:1038-1079).
#include "poisson_op.h"

namespace odil {

struct LaplaceParams {
  double dx, dy;
};

template <class T>
struct LaplaceOp {
  using Params = LaplaceParams;
  static constexpr int NF = 2, NE = 3, NS = 5, R = 1, RJ = 1;
  static constexpr int SX[NS] = {0, -1, 1, 0, 0};
  static constexpr int SY[NS] = {0, 0, 0, -1, 1};
  static constexpr int NJ = 10;
  static constexpr int JE[NJ] = {0, 0, 0, 0, 0, 1, 1, 1, 1, 1};
  static constexpr int JV[NJ] = {0, 0, 0, 0, 0, 1, 1, 1, 1, 1};
  static constexpr int JS[NJ] = {0, 1, 2, 3, 4, 0, 1, 2, 3, 4};

  template <class Acc>
  __device__ static inline void go(const Acc& q, int i, int j, int n0,
                                   int n1, const Params& P, T* f,
                                   T (*J)[NF][NS]) {
    for (int k = 0; k < 2; ++k) {
      T c[5];
      lap5<T>(q(k, 0, 0), q(k, -1, 0), q(k, 1, 0), q(k, 0, -1), q(k, 0, 1),
              i, j, n0, n1, LINEAR, T(P.dx), T(P.dy), &f[k], J ? c : nullptr);
      if (J)
        for (int s = 0; s < 5; ++s) {
          J[k][k][s] = c[s];
          J[k][1 - k][s] = 0;
          J[2][k][s] = 0;
        }
    }
    f[2] = 0;
  }
  template <bool INT = false, class Acc>
  __device__ static inline void residual(const Acc& q, int i, int j, int n0,
                                         int n1, const Params& P, T* f) {
    go(q, i, j, n0, n1, P, f, nullptr);
  }
  template <bool INT = false, class Acc>
  __device__ static inline void eval(const Acc& q, int i, int j, int n0,
                                     int n1, const Params& P, T* f,
                                     T (*J)[NF][NS]) {
    go(q, i, j, n0, n1, P, f, J);
  }
};

template <class Op>
int entries_table(int* ent, int* shifts) {
  if (ent)
    for (int q = 0; q < Op::NJ; ++q) {
      ent[3 * q] = Op::JE[q];
      ent[3 * q + 1] = Op::JV[q];
      ent[3 * q + 2] = Op::JS[q];
    }
  if (shifts)
    for (int s = 0; s < Op::NS; ++s) {
      shifts[2 * s] = Op::SX[s];
      shifts[2 * s + 1] = Op::SY[s];
    }
  return Op::NJ;
}

}  // namespace odil

using namespace odil;

extern "C" {

// Poisson: `src` has the dtype of the fields.
int odil_poisson_residual(int dtype, int n0, int n1, double dx, double dy,
                          const void* src, const void* const* fields,
                          void* const* res, double* partial, void* stream) {
  if (n0 < 2 || n1 < 2 || !src || !fields) {
    set_error("poisson_residual: need n >= 2, src and fields");
    return kErrArg;
  }
  PoissonParams P{dx, dy, src};
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    return launch_residual<double, PoissonOp<double>>(fields, res, partial, n0,
                                                      n1, P, s, nullptr);
  if (dtype == kF32)
    return launch_residual<float, PoissonOp<float>>(fields, res, partial, n0,
                                                    n1, P, s, nullptr);
  set_error("poisson_residual: bad dtype %d", dtype);
  return kErrArg;
}

int odil_poisson_jac(int dtype, int n0, int n1, double dx, double dy,
                     const void* src, const void* const* fields, void* coef,
                     void* const* res, void* stream) {
  if (n0 < 2 || n1 < 2 || !src || !fields || !coef) {
    set_error("poisson_jac: bad arguments");
    return kErrArg;
  }
  PoissonParams P{dx, dy, src};
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    return launch_jac<double, PoissonOp<double>>(fields, coef, res, n0, n1, P,
                                                 s);
  if (dtype == kF32)
    return launch_jac<float, PoissonOp<float>>(fields, coef, res, n0, n1, P,
                                               s);
  set_error("poisson_jac: bad dtype %d", dtype);
  return kErrArg;
}

int odil_poisson_jac_entries(int* ent, int* shifts) {
  return entries_table<PoissonOp<double>>(ent, shifts);
}

int odil_poisson_loss_grad(int dtype, int n0, int n1, double dx, double dy,
                           const void* src, const void* const* fields,
                           void* const* grad, double* partial, int* nblocks,
                           void* stream) {
  if (n0 < 2 || n1 < 2 || !src || !fields || !grad) {
    set_error("poisson_loss_grad: bad arguments");
    return kErrArg;
  }
  PoissonParams P{dx, dy, src};
  hipStream_t s = static_cast<hipStream_t>(stream);
  const double scale = 2.0 / ((double)n0 * n1);
  if (dtype == kF64)
    return launch_loss_grad_t<double, PoissonOp<double>, kGradTx64, kGradTy64>(
        fields, grad, partial, n0, n1, scale, P, s, nblocks);
  if (dtype == kF32)
    return launch_loss_grad_t<float, PoissonOp<float>, kGradTx32, kGradTy32>(
        fields, grad, partial, n0, n1, (float)scale, P, s, nblocks);
  set_error("poisson_loss_grad: bad dtype %d", dtype);
  return kErrArg;
}

int odil_poisson_nblocks(int kind, int dtype, int n0, int n1) {
  if (kind == 0) return residual_nblocks(n0, n1, PoissonOp<float>::R);
  return grad_nblocks(dtype == kF32, n0, n1);
}

int odil_laplace_residual(int dtype, int n0, int n1, double dx, double dy,
                          const void* const* fields, void* const* res,
                          double* partial, void* stream) {
  if (n0 < 2 || n1 < 2 || !fields) {
    set_error("laplace_residual: bad arguments");
    return kErrArg;
  }
  LaplaceParams P{dx, dy};
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    return launch_residual<double, LaplaceOp<double>>(fields, res, partial, n0,
                                                      n1, P, s, nullptr);
  if (dtype == kF32)
    return launch_residual<float, LaplaceOp<float>>(fields, res, partial, n0,
                                                    n1, P, s, nullptr);
  set_error("laplace_residual: bad dtype %d", dtype);
  return kErrArg;
}

int odil_laplace_jac(int dtype, int n0, int n1, double dx, double dy,
                     const void* const* fields, void* coef, void* const* res,
                     void* stream) {
  if (n0 < 2 || n1 < 2 || !fields || !coef) {
    set_error("laplace_jac: bad arguments");
    return kErrArg;
  }
  LaplaceParams P{dx, dy};
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    return launch_jac<double, LaplaceOp<double>>(fields, coef, res, n0, n1, P,
                                                 s);
  if (dtype == kF32)
    return launch_jac<float, LaplaceOp<float>>(fields, coef, res, n0, n1, P,
                                               s);
  set_error("laplace_jac: bad dtype %d", dtype);
  return kErrArg;
}

int odil_laplace_jac_entries(int* ent, int* shifts) {
  return entries_table<LaplaceOp<double>>(ent, shifts);
}

int odil_laplace_loss_grad(int dtype, int n0, int n1, double dx, double dy,
                           const void* const* fields, void* const* grad,
                           double* partial, int* nblocks, void* stream) {
  if (n0 < 2 || n1 < 2 || !fields || !grad) {
    set_error("laplace_loss_grad: bad arguments");
    return kErrArg;
  }
  LaplaceParams P{dx, dy};
  hipStream_t s = static_cast<hipStream_t>(stream);
  const double scale = 2.0 / ((double)n0 * n1);
  if (dtype == kF64)
    return launch_loss_grad_t<double, LaplaceOp<double>, kGradTx64, kGradTy64>(
        fields, grad, partial, n0, n1, scale, P, s, nblocks);
  if (dtype == kF32)
    return launch_loss_grad_t<float, LaplaceOp<float>, kGradTx32, kGradTy32>(
        fields, grad, partial, n0, n1, (float)scale, P, s, nblocks);
  set_error("laplace_loss_grad: bad dtype %d", dtype);
  return kErrArg;
}

int odil_laplace_nblocks(int kind, int dtype, int n0, int n1) {
  if (kind == 0) return residual_nblocks(n0, n1, LaplaceOp<float>::R);
  return grad_nblocks(dtype == kF32, n0, n1);
}

int odil_poisson_normal_apply(int dtype, int n0, int n1, double dx, double dy,
                              const void* src, const void* const* state,
                              const void* const* x, const void* const* b,
                              const void* const* d, void* const* y,
                              double omega, int mode, void* stream) {
  if (n0 < 2 || n1 < 2 || !src) {
    set_error("poisson_normal_apply: bad arguments");
    return kErrArg;
  }
  PoissonParams P{dx, dy, src};
  return dispatch_normal_apply<PoissonOp<double>, PoissonOp<float>>(
      dtype, state, x, b, d, y, n0, n1, omega, mode, P,
      static_cast<hipStream_t>(stream));
}

int odil_poisson_jtf(int dtype, int n0, int n1, double dx, double dy,
                     const void* src, const void* const* fields,
                     void* const* out, double scale, double* partial,
                     int* nblocks, void* stream) {
  if (n0 < 2 || n1 < 2 || !src || !fields || !out) {
    set_error("poisson_jtf: bad arguments");
    return kErrArg;
  }
  PoissonParams P{dx, dy, src};
  return dispatch_jtf<PoissonOp<double>, PoissonOp<float>>(
      dtype, fields, out, partial, n0, n1, scale, P,
      static_cast<hipStream_t>(stream), nblocks);
}

int odil_laplace_normal_apply(int dtype, int n0, int n1, double dx, double dy,
                              const void* const* state, const void* const* x,
                              const void* const* b, const void* const* d,
                              void* const* y, double omega, int mode,
                              void* stream) {
  if (n0 < 2 || n1 < 2) {
    set_error("laplace_normal_apply: bad arguments");
    return kErrArg;
  }
  LaplaceParams P{dx, dy};
  return dispatch_normal_apply<LaplaceOp<double>, LaplaceOp<float>>(
      dtype, state, x, b, d, y, n0, n1, omega, mode, P,
      static_cast<hipStream_t>(stream));
}

int odil_laplace_jtf(int dtype, int n0, int n1, double dx, double dy,
                     const void* const* fields, void* const* out,
                     double scale, double* partial, int* nblocks,
                     void* stream) {
  if (n0 < 2 || n1 < 2 || !fields || !out) {
    set_error("laplace_jtf: bad arguments");
    return kErrArg;
  }
  LaplaceParams P{dx, dy};
  return dispatch_jtf<LaplaceOp<double>, LaplaceOp<float>>(
      dtype, fields, out, partial, n0, n1, scale, P,
      static_cast<hipStream_t>(stream), nblocks);
}

}  // extern "C"
