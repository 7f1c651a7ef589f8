// Inverse advection problem on a (t, x, y) space-time grid (reference:
// advection_2d.py:103-250, problem='inverse', central=0, implicit=0):
// unknown velocity (u, v) and scalar p, equations in the reference's order
//   fe    = (p - p_ex) alpha_t m(t)      alpha_t = alpha 0.5^(epoch/anneal_half)
//                                        (annealed by the caller), m = 1, or
//                                        with onlyfinal 1 only at t = 0, nt-1
//   fp    = (p - p_prev)/dt + u Dx p + v Dy p   first-order upwind on the sign
//           of the frozen velocity at the cell; p_prev = p_ex at t = 0
//   fsu, fsv = kreg lap_xy(u), kreg lap_xy(v)               (if kreg != 0)
//   ftu, ftv = ktreg (u - u(t-1))/dt, 0 at t = 0             (if ktreg != 0)
//   fdiv  = divfree ((u(x+1) - u(x-1))/2dx + (v(y+1) - v(y-1))/2dy)
//                                                           (if divfree != 0)
// x and y periodic (periodic = 1) or zero outside the grid (periodic = 0,
// apply_bc_zero); t neighbours are used only inside the grid.
//
// Loss gradient (the tape gradient with the frozen upwind direction held):
// grad_X(c) = (2/N) sum_e sum_c' f_e(c') d f_e(c') / d X(c), gathered from
// the equations at c and its t/x/y neighbours:
//   p: A m(c) fe(c) + fp(c)(1/dt + |u|/dx + |v|/dy)(c) - fp(c + t)/dt
//      - u(c + x) fp(c + x)/dx [u(c + x) > 0] + u(c - x) fp(c - x)/dx
//      [u(c - x) <= 0]   (and the same along y with v)
//   u: fp(c) Dx p(c) + kreg lap_xy(fsu)(c) + ktreg (ftu(c) - ftu(c + t))/dt
//      + divfree (fdiv(c - x) - fdiv(c + x)) / 2dx
//   v: likewise along y.
#include "common.h"

namespace odil {

constexpr int kAdvInvMaxEq = 7;

struct AdvInvParams {
  int nt, nx, ny;
  int onlyfinal, periodic;
  int ek, et, ed;  // equation index of fsu (fsv = ek + 1), ftu, fdiv or -1
  int ne;
  double dt, dx, dy, alpha, kreg, ktreg, divfree;
};

template <class T>
struct AdvInvFields {
  const T* u;
  const T* v;
  const T* p;
  const T* pex;
};

// Field value at (it, ix + di, iy + dj): wrapped in x, y (periodic) or zero
// outside the grid.
template <class T>
__device__ inline T adv_at(const T* a, const AdvInvParams& P, int it, int ix,
                           int iy) {
  if (P.periodic) {
    ix = ix < 0 ? ix + P.nx : (ix >= P.nx ? ix - P.nx : ix);
    iy = iy < 0 ? iy + P.ny : (iy >= P.ny ? iy - P.ny : iy);
  } else if (ix < 0 || ix >= P.nx || iy < 0 || iy >= P.ny) {
    return T(0);
  }
  return a[((size_t)it * P.nx + ix) * P.ny + iy];
}

// Does the cell (ix, iy) (possibly outside in x, y) hold equations?
__device__ inline bool adv_has(const AdvInvParams& P, int ix, int iy) {
  return P.periodic || (ix >= 0 && ix < P.nx && iy >= 0 && iy < P.ny);
}

template <class T>
__device__ inline void adv_wrap(const AdvInvParams& P, int& ix, int& iy) {
  ix = ix < 0 ? ix + P.nx : (ix >= P.nx ? ix - P.nx : ix);
  iy = iy < 0 ? iy + P.ny : (iy >= P.ny ? iy - P.ny : iy);
}

// Upwind differences of p at a cell on the sign of the velocity there.
template <class T>
__device__ inline void adv_grad_p(const AdvInvFields<T>& F,
                                  const AdvInvParams& P, int it, int ix,
                                  int iy, T uc, T vc, T pc, T& px, T& py) {
  const T dx = T(P.dx), dy = T(P.dy);
  px = uc > T(0) ? (pc - adv_at(F.p, P, it, ix - 1, iy)) / dx
                 : (adv_at(F.p, P, it, ix + 1, iy) - pc) / dx;
  py = vc > T(0) ? (pc - adv_at(F.p, P, it, ix, iy - 1)) / dy
                 : (adv_at(F.p, P, it, ix, iy + 1) - pc) / dy;
}

template <class T>
__device__ inline T adv_fp(const AdvInvFields<T>& F, const AdvInvParams& P,
                           int it, int ix, int iy) {
  const size_t c = ((size_t)it * P.nx + ix) * P.ny + iy;
  const T uc = F.u[c], vc = F.v[c], pc = F.p[c];
  const T pt = it == 0 ? F.pex[c] : F.p[c - (size_t)P.nx * P.ny];
  T px, py;
  adv_grad_p(F, P, it, ix, iy, uc, vc, pc, px, py);
  return (pc - pt) / T(P.dt) + uc * px + vc * py;
}

template <class T>
__device__ inline T adv_lap(const T* a, const AdvInvParams& P, int it, int ix,
                            int iy) {
  const T c = adv_at(a, P, it, ix, iy);
  const T idx2 = T(1) / T(P.dx * P.dx), idy2 = T(1) / T(P.dy * P.dy);
  return (adv_at(a, P, it, ix + 1, iy) - T(2) * c +
          adv_at(a, P, it, ix - 1, iy)) *
             idx2 +
         (adv_at(a, P, it, ix, iy + 1) - T(2) * c +
          adv_at(a, P, it, ix, iy - 1)) *
             idy2;
}

template <class T>
__device__ inline T adv_ft(const T* a, const AdvInvParams& P, int it, int ix,
                           int iy) {
  if (it == 0) return T(0);
  const size_t c = ((size_t)it * P.nx + ix) * P.ny + iy;
  return T(P.ktreg) * (a[c] - a[c - (size_t)P.nx * P.ny]) / T(P.dt);
}

template <class T>
__device__ inline T adv_div(const AdvInvFields<T>& F, const AdvInvParams& P,
                            int it, int ix, int iy) {
  return T(P.divfree) *
         ((adv_at(F.u, P, it, ix + 1, iy) - adv_at(F.u, P, it, ix - 1, iy)) /
              T(2 * P.dx) +
          (adv_at(F.v, P, it, ix, iy + 1) - adv_at(F.v, P, it, ix, iy - 1)) /
              T(2 * P.dy));
}

template <class T>
__device__ inline T adv_mask(const AdvInvParams& P, int it) {
  return (!P.onlyfinal || it == 0 || it == P.nt - 1) ? T(1) : T(0);
}

// Per-thread values live in fixed slots (fe, fp, fsu, fsv, ftu, ftv, fdiv)
// so that register arrays are indexed at compile time; slot -> equation
// index in the reference's order, or -1 when the term is off.
__device__ inline int adv_eq(const AdvInvParams& P, int slot) {
  switch (slot) {
    case 0: return 0;
    case 1: return 1;
    case 2: return P.ek;
    case 3: return P.ek < 0 ? -1 : P.ek + 1;
    case 4: return P.et;
    case 5: return P.et < 0 ? -1 : P.et + 1;
    default: return P.ed;
  }
}

// Block sums of the slots, written by thread 0 as partial[eq][block].
__device__ inline void adv_partials(double (&ssq)[kAdvInvMaxEq],
                                    const AdvInvParams& P, double* partial,
                                    double* red, unsigned b) {
  block_sum_n<kAdvInvMaxEq>(ssq, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < kAdvInvMaxEq; ++q) {
      const int e = adv_eq(P, q);
      if (e >= 0) partial[(size_t)e * gridDim.x + b] = ssq[q];
    }
  }
}

template <class T>
__global__ void __launch_bounds__(256)
    k_advi_loss_grad(AdvInvFields<T> F, AdvInvParams P, T* gu, T* gv, T* gp,
                     double* __restrict__ partial, T scale) {
  __shared__ double red[kAdvInvMaxEq * 4];
  const size_t N = (size_t)P.nt * P.nx * P.ny;
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const size_t c = (size_t)b * 256 + threadIdx.x;
  double ssq[kAdvInvMaxEq];
  for (int e = 0; e < kAdvInvMaxEq; ++e) ssq[e] = 0;
  if (c < N) {
    const int iy = (int)(c % P.ny);
    const size_t r = c / P.ny;
    const int ix = (int)(r % P.nx);
    const int it = (int)(r / P.nx);
    const T dt = T(P.dt), dx = T(P.dx), dy = T(P.dy);
    const T uc = F.u[c], vc = F.v[c], pc = F.p[c];
    // Equations at the cell (also their squares).
    const T m = adv_mask<T>(P, it);
    const T fe = (pc - F.pex[c]) * T(P.alpha) * m;
    T px, py;
    adv_grad_p(F, P, it, ix, iy, uc, vc, pc, px, py);
    const T ptm = it == 0 ? F.pex[c] : F.p[c - (size_t)P.nx * P.ny];
    const T fp = (pc - ptm) / dt + uc * px + vc * py;
    ssq[0] = (double)fe * fe;
    ssq[1] = (double)fp * fp;
    T g_p = T(P.alpha) * m * fe +
            fp * (T(1) / dt + fabs(uc) / dx + fabs(vc) / dy);
    T g_u = fp * px, g_v = fp * py;
    if (it + 1 < P.nt) g_p -= adv_fp(F, P, it + 1, ix, iy) / dt;
    // x and y neighbours whose upwind difference reads p(c).
    {
      int jx = ix + 1, jy = iy;
      if (adv_has(P, jx, jy)) {
        adv_wrap<T>(P, jx, jy);
        const T un = F.u[((size_t)it * P.nx + jx) * P.ny + jy];
        if (un > T(0)) g_p -= un / dx * adv_fp(F, P, it, jx, jy);
      }
      jx = ix - 1, jy = iy;
      if (adv_has(P, jx, jy)) {
        adv_wrap<T>(P, jx, jy);
        const T un = F.u[((size_t)it * P.nx + jx) * P.ny + jy];
        if (!(un > T(0))) g_p += un / dx * adv_fp(F, P, it, jx, jy);
      }
      jx = ix, jy = iy + 1;
      if (adv_has(P, jx, jy)) {
        adv_wrap<T>(P, jx, jy);
        const T vn = F.v[((size_t)it * P.nx + jx) * P.ny + jy];
        if (vn > T(0)) g_p -= vn / dy * adv_fp(F, P, it, jx, jy);
      }
      jx = ix, jy = iy - 1;
      if (adv_has(P, jx, jy)) {
        adv_wrap<T>(P, jx, jy);
        const T vn = F.v[((size_t)it * P.nx + jx) * P.ny + jy];
        if (!(vn > T(0))) g_p += vn / dy * adv_fp(F, P, it, jx, jy);
      }
    }
    if (P.ek >= 0) {
      // kreg lap(fsu) with fsu = kreg lap(u): the Laplacian is symmetric
      // under both boundary rules (wrap, or zero outside with the outside
      // equations absent).
      const T k2 = T(P.kreg) * T(P.kreg);
      const T lu = adv_lap(F.u, P, it, ix, iy);
      const T lv = adv_lap(F.v, P, it, ix, iy);
      ssq[2] = (double)(T(P.kreg) * lu) * (double)(T(P.kreg) * lu);
      ssq[3] = (double)(T(P.kreg) * lv) * (double)(T(P.kreg) * lv);
      const T idx2 = T(1) / (dx * dx), idy2 = T(1) / (dy * dy);
      T su = -T(2) * (idx2 + idy2) * lu, sv = -T(2) * (idx2 + idy2) * lv;
      const int dxs[4] = {1, -1, 0, 0}, dys[4] = {0, 0, 1, -1};
      for (int q = 0; q < 4; ++q) {
        int jx = ix + dxs[q], jy = iy + dys[q];
        if (!adv_has(P, jx, jy)) continue;
        adv_wrap<T>(P, jx, jy);
        const T w = q < 2 ? idx2 : idy2;
        su += w * adv_lap(F.u, P, it, jx, jy);
        sv += w * adv_lap(F.v, P, it, jx, jy);
      }
      g_u += k2 * su;
      g_v += k2 * sv;
    }
    if (P.et >= 0) {
      const T fu = adv_ft(F.u, P, it, ix, iy), fv = adv_ft(F.v, P, it, ix, iy);
      ssq[4] = (double)fu * fu;
      ssq[5] = (double)fv * fv;
      T du = fu, dv = fv;
      if (it + 1 < P.nt) {
        du -= adv_ft(F.u, P, it + 1, ix, iy);
        dv -= adv_ft(F.v, P, it + 1, ix, iy);
      }
      g_u += T(P.ktreg) / dt * du;
      g_v += T(P.ktreg) / dt * dv;
    }
    if (P.ed >= 0) {
      const T fd = adv_div(F, P, it, ix, iy);
      ssq[6] = (double)fd * fd;
      const T D = T(P.divfree);
      int jx = ix - 1, jy = iy;
      if (adv_has(P, jx, jy)) {
        adv_wrap<T>(P, jx, jy);
        g_u += D / (T(2) * dx) * adv_div(F, P, it, jx, jy);
      }
      jx = ix + 1, jy = iy;
      if (adv_has(P, jx, jy)) {
        adv_wrap<T>(P, jx, jy);
        g_u -= D / (T(2) * dx) * adv_div(F, P, it, jx, jy);
      }
      jx = ix, jy = iy - 1;
      if (adv_has(P, jx, jy)) {
        adv_wrap<T>(P, jx, jy);
        g_v += D / (T(2) * dy) * adv_div(F, P, it, jx, jy);
      }
      jx = ix, jy = iy + 1;
      if (adv_has(P, jx, jy)) {
        adv_wrap<T>(P, jx, jy);
        g_v -= D / (T(2) * dy) * adv_div(F, P, it, jx, jy);
      }
    }
    gu[c] = g_u * scale;
    gv[c] = g_v * scale;
    gp[c] = g_p * scale;
  }
  if (partial) adv_partials(ssq, P, partial, red, b);
}

}  // namespace odil

using namespace odil;

namespace {
int advi_params(int nt, int nx, int ny, double dt, double dx, double dy,
                double alpha, int onlyfinal, double kreg, double ktreg,
                double divfree, int periodic, const void* const* fields,
                const void* pex, AdvInvParams* P) {
  if (nt < 1 || nx < 3 || ny < 3 || !fields || !fields[0] || !fields[1] ||
      !fields[2] || !pex) {
    set_error("advection_inverse: need nt >= 1, nx, ny >= 3, fields u, v, "
              "p and p_exact");
    return kErrArg;
  }
  *P = AdvInvParams{nt, nx, ny, onlyfinal ? 1 : 0, periodic ? 1 : 0, -1, -1,
                    -1, 2, dt, dx, dy, alpha, kreg, ktreg, divfree};
  if (kreg != 0) {
    P->ek = P->ne;
    P->ne += 2;
  }
  if (ktreg != 0) {
    P->et = P->ne;
    P->ne += 2;
  }
  if (divfree != 0) {
    P->ed = P->ne;
    P->ne += 1;
  }
  return kOk;
}

template <class T>
AdvInvFields<T> advi_fields(const void* const* fields, const void* pex) {
  return AdvInvFields<T>{static_cast<const T*>(fields[0]),
                         static_cast<const T*>(fields[1]),
                         static_cast<const T*>(fields[2]),
                         static_cast<const T*>(pex)};
}

// Residual output pointers by equation index, passed by value.
template <class T>
struct ResPtrs {
  T* r[kAdvInvMaxEq];
};

template <class T>
__global__ void __launch_bounds__(256)
    k_advi_residual_args(AdvInvFields<T> F, AdvInvParams P, ResPtrs<T> R,
                         int any, double* __restrict__ partial) {
  __shared__ double red[kAdvInvMaxEq * 4];
  const size_t N = (size_t)P.nt * P.nx * P.ny;
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const size_t c = (size_t)b * 256 + threadIdx.x;
  double ssq[kAdvInvMaxEq];
  for (int e = 0; e < kAdvInvMaxEq; ++e) ssq[e] = 0;
  if (c < N) {
    const int iy = (int)(c % P.ny);
    const size_t r = c / P.ny;
    const int ix = (int)(r % P.nx);
    const int it = (int)(r / P.nx);
    T f[kAdvInvMaxEq];
    for (int q = 0; q < kAdvInvMaxEq; ++q) f[q] = 0;
    f[0] = (F.p[c] - F.pex[c]) * T(P.alpha) * adv_mask<T>(P, it);
    f[1] = adv_fp(F, P, it, ix, iy);
    if (P.ek >= 0) {
      f[2] = T(P.kreg) * adv_lap(F.u, P, it, ix, iy);
      f[3] = T(P.kreg) * adv_lap(F.v, P, it, ix, iy);
    }
    if (P.et >= 0) {
      f[4] = adv_ft(F.u, P, it, ix, iy);
      f[5] = adv_ft(F.v, P, it, ix, iy);
    }
    if (P.ed >= 0) f[6] = adv_div(F, P, it, ix, iy);
#pragma unroll
    for (int q = 0; q < kAdvInvMaxEq; ++q) {
      const int e = adv_eq(P, q);
      if (e >= 0) {
        if (any && R.r[e]) R.r[e][c] = f[q];
        ssq[q] = (double)f[q] * (double)f[q];
      }
    }
  }
  if (partial) adv_partials(ssq, P, partial, red, b);
}
}  // namespace

extern "C" {

// Number of equations for these regularisation weights: 2 (fe, fp) + 2 if
// kreg + 2 if ktreg + 1 if divfree (advection_2d.py:224-250).
int odil_advection_inverse_neq(double kreg, double ktreg, double divfree) {
  return 2 + (kreg != 0 ? 2 : 0) + (ktreg != 0 ? 2 : 0) + (divfree != 0);
}

// Partial sums per equation: one per 256-cell block.
int odil_advection_inverse_nblocks(int nt, int nx, int ny) {
  return ceil_div(nt * nx * ny, 256);
}

// Residuals of the active equations (res[e] may be NULL; res may be NULL)
// and per-block sums of squares `partial` [neq][nblocks].
int odil_advection_inverse_residual(int dtype, int nt, int nx, int ny,
                                    double dt, double dx, double dy,
                                    double alpha, int onlyfinal, double kreg,
                                    double ktreg, double divfree,
                                    int periodic, const void* p_exact,
                                    const void* const* fields,
                                    void* const* res, double* partial,
                                    void* stream) {
  AdvInvParams P;
  int st = advi_params(nt, nx, ny, dt, dx, dy, alpha, onlyfinal, kreg, ktreg,
                       divfree, periodic, fields, p_exact, &P);
  if (st != kOk) return st;
  dim3 grid(ceil_div(nt * nx * ny, 256));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64) {
    ResPtrs<double> R{};
    for (int e = 0; e < P.ne && res; ++e) R.r[e] = static_cast<double*>(res[e]);
    hipLaunchKernelGGL(k_advi_residual_args<double>, grid, dim3(256), 0, s,
                       advi_fields<double>(fields, p_exact), P, R,
                       res ? 1 : 0, partial);
  } else if (dtype == kF32) {
    ResPtrs<float> R{};
    for (int e = 0; e < P.ne && res; ++e) R.r[e] = static_cast<float*>(res[e]);
    hipLaunchKernelGGL(k_advi_residual_args<float>, grid, dim3(256), 0, s,
                       advi_fields<float>(fields, p_exact), P, R, res ? 1 : 0,
                       partial);
  } else {
    set_error("advection_inverse_residual: bad dtype %d", dtype);
    return kErrArg;
  }
  return check_launch("advection_inverse_residual");
}

// Fused loss + adjoint: grad (u, v, p) = (2/N) sum_e J_e^T f_e with the
// upwind direction frozen, and the per-block sums of f_e^2.
int odil_advection_inverse_loss_grad(int dtype, int nt, int nx, int ny,
                                     double dt, double dx, double dy,
                                     double alpha, int onlyfinal, double kreg,
                                     double ktreg, double divfree,
                                     int periodic, const void* p_exact,
                                     const void* const* fields,
                                     void* const* grad, double* partial,
                                     void* stream) {
  AdvInvParams P;
  int st = advi_params(nt, nx, ny, dt, dx, dy, alpha, onlyfinal, kreg, ktreg,
                       divfree, periodic, fields, p_exact, &P);
  if (st != kOk) return st;
  if (!grad || !grad[0] || !grad[1] || !grad[2]) {
    set_error("advection_inverse_loss_grad: need three gradient arrays");
    return kErrArg;
  }
  dim3 grid(ceil_div(nt * nx * ny, 256));
  const double scale = 2.0 / ((double)nt * nx * ny);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    hipLaunchKernelGGL(k_advi_loss_grad<double>, grid, dim3(256), 0, s,
                       advi_fields<double>(fields, p_exact), P,
                       static_cast<double*>(grad[0]),
                       static_cast<double*>(grad[1]),
                       static_cast<double*>(grad[2]), partial, scale);
  else if (dtype == kF32)
    hipLaunchKernelGGL(k_advi_loss_grad<float>, grid, dim3(256), 0, s,
                       advi_fields<float>(fields, p_exact), P,
                       static_cast<float*>(grad[0]),
                       static_cast<float*>(grad[1]),
                       static_cast<float*>(grad[2]), partial, (float)scale);
  else {
    set_error("advection_inverse_loss_grad: bad dtype %d", dtype);
    return kErrArg;
  }
  return check_launch("advection_inverse_loss_grad");
}

}  // extern "C"
