// Poisson operator f = lap(u) - src on a 2D cell-centred grid with
// Dirichlet u = 0 walls (halos by quadratic extrapolation through the wall
// value, the no-slip halo rule of ns_steady.py:219-225 applied to a
// scalar), shared by the single-grid entry points and the batched launch.
#pragma once

#include "stencil2d.h"

namespace odil {


struct PoissonParams {
  double dx, dy;
  const void* src;  // same dtype as the fields, [n0][n1]
};

template <class T>
__device__ inline void lap5(T c, T rxm, T rxp, T rym, T ryp, int i, int j,
                            int n0, int n1, int kind, T dx, T dy, T* out,
                            T* coef) {
  const T xm = (i == 0) ? ghost(kind, rxp, c, T(0)) : rxm;
  const T xp = (i == n0 - 1) ? ghost(kind, rxm, c, T(0)) : rxp;
  const T ym = (j == 0) ? ghost(kind, ryp, c, T(0)) : rym;
  const T yp = (j == n1 - 1) ? ghost(kind, rym, c, T(0)) : ryp;
  *out = (xp - T(2) * c + xm) / (dx * dx) + (yp - T(2) * c + ym) / (dy * dy);
  if (!coef) return;
  const T ix2 = T(1) / (dx * dx), iy2 = T(1) / (dy * dy);
  T e[5] = {-T(2) * ix2 - T(2) * iy2, ix2, ix2, iy2, iy2};
  coef[0] = e[0];
  coef[1] = e[1];
  coef[2] = e[2];
  coef[3] = e[3];
  coef[4] = e[4];
  if (i == 0) {
    coef[2] += e[1] * ghost_dfar<T>(kind);
    coef[0] += e[1] * ghost_dnear<T>(kind);
    coef[1] = 0;
  }
  if (i == n0 - 1) {
    coef[1] += e[2] * ghost_dfar<T>(kind);
    coef[0] += e[2] * ghost_dnear<T>(kind);
    coef[2] = 0;
  }
  if (j == 0) {
    coef[4] += e[3] * ghost_dfar<T>(kind);
    coef[0] += e[3] * ghost_dnear<T>(kind);
    coef[3] = 0;
  }
  if (j == n1 - 1) {
    coef[3] += e[4] * ghost_dfar<T>(kind);
    coef[0] += e[4] * ghost_dnear<T>(kind);
    coef[4] = 0;
  }
}

template <class T>
struct PoissonOp {
  using Params = PoissonParams;
  static constexpr int NF = 1, NE = 1, NS = 5, R = 1, RJ = 1;
  static constexpr int SX[NS] = {0, -1, 1, 0, 0};
  static constexpr int SY[NS] = {0, 0, 0, -1, 1};
  static constexpr int NJ = 5;
  static constexpr int JE[NJ] = {0, 0, 0, 0, 0};
  static constexpr int JV[NJ] = {0, 0, 0, 0, 0};
  static constexpr int JS[NJ] = {0, 1, 2, 3, 4};

  template <class Acc>
  __device__ static inline void go(const Acc& q, int i, int j, int n0,
                                   int n1, const Params& P, T* f,
                                   T (*J)[NF][NS]) {
    T l;
    lap5<T>(q(0, 0, 0), q(0, -1, 0), q(0, 1, 0), q(0, 0, -1), q(0, 0, 1), i,
            j, n0, n1, QUADH, T(P.dx), T(P.dy), &l, J ? J[0][0] : nullptr);
    f[0] = l - static_cast<const T*>(P.src)[(size_t)i * n1 + j];
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

}  // namespace odil
