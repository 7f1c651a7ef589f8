// Space-time wave operator u_tt = c(x)^2 u_xx on an (nt, nx) grid
// (reference: wave.py:67-122, where c = 1), shared by the forward wave
// entry points and the inverse (speed recovery) problem.
//
// Axis 0 is t (nt cells), axis 1 is x (nx cells, contiguous). Row it = 0
// imposes u = u0 + dt/2 ut0; row it = 1 uses ut0 for the backward time
// difference; the x-neighbours at t-1 next to the walls come from quadratic
// extrapolation through the boundary value at t-1 (periodic index in t).
#pragma once

#include "stencil2d.h"

namespace odil {

struct WaveParams {
  double dt, dx;
  const double* init_u;   // [nx]
  const double* init_ut;  // [nx]
  const double* left_u;   // [nt], value at x = lower
  const double* right_u;  // [nt], value at x = upper
  // Optional wave speed c(x) [nx] in the field dtype (NULL: c = 1); the
  // spatial term is c(x)^2 u_xx.
  const void* c;
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
    T c2 = T(1);
    if (P.c) {
      const T cx = static_cast<const T*>(P.c)[ix];
      c2 = cx * cx;
    }
    f[0] = u_tt - c2 * u_xx;
    if (!want_jac) return;
    const T idt2 = T(1) / (dt * dt), idx2 = c2 / (dx * dx);
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
