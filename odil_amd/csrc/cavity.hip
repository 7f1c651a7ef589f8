// Steady lid-driven cavity operator: residual, stencil Jacobian and fused loss
// adjoint (reference: ns_steady.py:142-394, defaults bc=0, rhiechow=1).
//
// Unknowns u, v, p on a cell-centred (nx, ny) grid, x along axis 0, y along
// axis 1 (contiguous). Residuals:
//   fu = uf u_x + vf u_y - mu lap(u) + p_x
//   fv = uf v_x + vf v_y - mu lap(v) + p_y
//   fp = Rhie-Chow face-flux divergence
// with deferred-correction upwinding u_x = hi2(uf) + lo1(u) - lo1(uf)
// (ns_steady.py:171-204,282-283,320-323). Frozen copies (uf, vf, pf) carry
// no derivative (util_op.py:541-543), so the live Jacobian is
//   fu <- u: lo1 upwind + Laplacian (5-point), fu <- p: central (3-point)
//   fp <- u, v: face averages, fp <- p: Rhie-Chow pressure Laplacian
// and every equation is linear in the live fields.
#include <mutex>

#include "stencil2d.h"

namespace odil {

struct CavityParams {
  double dx, dy, mu;
  int bc;        // 0: no-slip + lid, 1: no-penetration, 2: extrapolation
  int rhiechow;  // 1: Rhie-Chow continuity, 0: central divergence
  // Grid constants computed once on the host (FP64): 1/dx, 1/dy,
  // 1/(2dx), 1/(2dy), mu (2/dx^2 + 2/dy^2), mu/dx^2, mu/dy^2.
  double idx, idy, h2x, h2y, gdiag, lxx, lyy;
};

inline CavityParams cavity_params(double dx, double dy, double mu, int bc,
                                  int rhiechow) {
  return CavityParams{dx,           dy,
                      mu,           bc,
                      rhiechow,     1.0 / dx,
                      1.0 / dy,     1.0 / (2.0 * dx),
                      1.0 / (2.0 * dy),
                      mu * (2.0 / (dx * dx) + 2.0 / (dy * dy)),
                      mu / (dx * dx), mu / (dy * dy)};
}

// Halo rules per field and face (ns_steady.py:219-257).
struct FaceRule {
  int kx, ky;       // extrapolation kind on x-faces and y-faces
  double wxm, wxp;  // wall values on x = 0 and x = 1
  double wym, wyp;  // wall values on y = 0 and y = 1
};

__host__ __device__ inline FaceRule cavity_rule(int bc, int field) {
  if (bc == 0) return FaceRule{QUADH, QUADH, 0, 0, 0, field == 0 ? 1. : 0.};
  if (bc == 1)
    return field == 0 ? FaceRule{QUADH, LINEAR, 0, 0, 0, 0}
                      : FaceRule{LINEAR, QUADH, 0, 0, 0, 0};
  return FaceRule{LINEAR, LINEAR, 0, 0, 0, 0};
}

// Halo-corrected neighbours of one field around cell (i, j): distance-1
// values and the distance-2 values derived from them (stencil5,
// ns_steady.py:269-280).
template <class T>
struct Nbr {
  T c, xm, xp, ym, yp, xmm, xpp, ymm, ypp;
};

template <bool INT, class T, class Acc>
__device__ inline Nbr<T> neighbours(const Acc& q, int k, int i, int j, int nx,
                                    int ny, const FaceRule& r) {
  if (INT) {  // no wall within two cells: plain neighbours
    return Nbr<T>{q(k, 0, 0),  q(k, -1, 0), q(k, 1, 0),
                  q(k, 0, -1), q(k, 0, 1),  q(k, -2, 0),
                  q(k, 2, 0),  q(k, 0, -2), q(k, 0, 2)};
  }
  Nbr<T> s;
  const T c = q(k, 0, 0);
  const T rxm = q(k, -1, 0), rxp = q(k, 1, 0);
  const T rym = q(k, 0, -1), ryp = q(k, 0, 1);
  const T wxm = T(r.wxm), wxp = T(r.wxp), wym = T(r.wym), wyp = T(r.wyp);
  s.c = c;
  s.xm = (i == 0) ? ghost(r.kx, rxp, c, wxm) : rxm;
  s.xp = (i == nx - 1) ? ghost(r.kx, rxm, c, wxp) : rxp;
  s.ym = (j == 0) ? ghost(r.ky, ryp, c, wym) : rym;
  s.yp = (j == ny - 1) ? ghost(r.ky, rym, c, wyp) : ryp;
  s.xmm = (i == 0)   ? s.xm
          : (i == 1) ? ghost(r.kx, c, rxm, wxm)
                     : q(k, -2, 0);
  s.xpp = (i == nx - 1)   ? s.xp
          : (i == nx - 2) ? ghost(r.kx, c, rxp, wxp)
                          : q(k, 2, 0);
  s.ymm = (j == 0)   ? s.ym
          : (j == 1) ? ghost(r.ky, c, rym, wym)
                     : q(k, 0, -2);
  s.ypp = (j == ny - 1)   ? s.yp
          : (j == ny - 2) ? ghost(r.ky, c, ryp, wyp)
                          : q(k, 0, 2);
  return s;
}

// Effective-neighbour coefficients folded onto raw values: where a
// neighbour is a wall halo, its coefficient moves to the far and near inner
// values of the extrapolation.
template <bool INT, class T>
__device__ inline void fold(T* raw, const T* eff, int i, int j, int nx, int ny,
                            int kx, int ky) {
  // raw/eff order: c, xm, xp, ym, yp
  raw[0] = eff[0];
  raw[1] = eff[1];
  raw[2] = eff[2];
  raw[3] = eff[3];
  raw[4] = eff[4];
  if (INT) return;
  if (i == 0) {
    raw[2] += eff[1] * ghost_dfar<T>(kx);
    raw[0] += eff[1] * ghost_dnear<T>(kx);
    raw[1] = 0;
  }
  if (i == nx - 1) {
    raw[1] += eff[2] * ghost_dfar<T>(kx);
    raw[0] += eff[2] * ghost_dnear<T>(kx);
    raw[2] = 0;
  }
  if (j == 0) {
    raw[4] += eff[3] * ghost_dfar<T>(ky);
    raw[0] += eff[3] * ghost_dnear<T>(ky);
    raw[3] = 0;
  }
  if (j == ny - 1) {
    raw[3] += eff[4] * ghost_dfar<T>(ky);
    raw[0] += eff[4] * ghost_dnear<T>(ky);
    raw[4] = 0;
  }
}

template <class T>
struct CavityOp {
  using Params = CavityParams;
  static constexpr int NF = 3, NE = 3, NS = 5, R = 2, RJ = 1;
  // Shift order: centre, x-1, x+1, y-1, y+1.
  static constexpr int SX[NS] = {0, -1, 1, 0, 0};
  static constexpr int SY[NS] = {0, 0, 0, -1, 1};
  // Non-zero Jacobian entries (equation, field, shift): 27.
  static constexpr int NJ = 27;
  static constexpr int JE[NJ] = {0, 0, 0, 0, 0, 0, 0, 0,        //
                                 1, 1, 1, 1, 1, 1, 1, 1,        //
                                 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2};
  static constexpr int JV[NJ] = {0, 0, 0, 0, 0, 2, 2, 2,        //
                                 1, 1, 1, 1, 1, 2, 2, 2,        //
                                 0, 0, 0, 1, 1, 1, 2, 2, 2, 2, 2};
  static constexpr int JS[NJ] = {0, 1, 2, 3, 4, 0, 1, 2,        //
                                 0, 1, 2, 3, 4, 0, 3, 4,        //
                                 0, 1, 2, 0, 3, 4, 0, 1, 2, 3, 4};

  // 1/g with the hardware reciprocal: FP32 as is (1 ulp); FP64 refined by
  // two Newton-Raphson steps y <- y + y (1 - g y) (to ~1 ulp).
  __device__ static inline T recip(T g) {
    if constexpr (sizeof(T) == 4) {
      return __builtin_amdgcn_rcpf(g);
    } else {
      T y = __builtin_amdgcn_rcp(g);
      y = fma(y, fma(-g, y, T(1)), y);
      return fma(y, fma(-g, y, T(1)), y);
    }
  }

  // INT: the cell and its distance-2 neighbours are inside the domain and
  // away from the walls, so every halo rule drops out at compile time.
  template <bool INT, class Acc>
  __device__ static inline void compute(const Acc& q, int i, int j, int nx,
                                        int ny, const Params& P, T* f,
                                        T (*J)[NF][NS], bool want_jac) {
    const T mu = T(P.mu);
    const FaceRule ru = cavity_rule(P.bc, 0), rv = cavity_rule(P.bc, 1);
    const FaceRule rp{LINEAR, LINEAR, 0, 0, 0, 0};
    const Nbr<T> U = neighbours<INT, T>(q, 0, i, j, nx, ny, ru);
    const Nbr<T> V = neighbours<INT, T>(q, 1, i, j, nx, ny, rv);
    const Nbr<T> Pn = neighbours<INT, T>(q, 2, i, j, nx, ny, rp);
    const bool iw0 = !INT && i == 0, iw1 = !INT && i == nx - 1;
    const bool jw0 = !INT && j == 0, jw1 = !INT && j == ny - 1;
    const T a = U.c, b = V.c;  // frozen velocities at the cell
    const T idx = T(P.idx), idy = T(P.idy);
    const T h2x = T(P.h2x), h2y = T(P.h2y);

    // Second-order upwind of the frozen field, first order at walls.
    auto hi2x = [&](const Nbr<T>& s) {
      const T m = iw0 ? (s.c - s.xm) * idx
                           : (s.xmm - T(4) * s.xm + T(3) * s.c) * h2x;
      const T p = iw1 ? (s.xp - s.c) * idx
                                : (-T(3) * s.c + T(4) * s.xp - s.xpp) * h2x;
      return a > T(0) ? m : p;
    };
    auto hi2y = [&](const Nbr<T>& s) {
      const T m = jw0 ? (s.c - s.ym) * idy
                           : (s.ymm - T(4) * s.ym + T(3) * s.c) * h2y;
      const T p = jw1 ? (s.yp - s.c) * idy
                                : (-T(3) * s.c + T(4) * s.yp - s.ypp) * h2y;
      return b > T(0) ? m : p;
    };
    const T idx2 = idx * idx, idy2 = idy * idy;
    auto lap = [&](const Nbr<T>& s) {
      return (s.xp - T(2) * s.c + s.xm) * idx2 +
             (s.yp - T(2) * s.c + s.ym) * idy2;
    };
    {
      // u_x = hi2(uf) + lo1(u) - lo1(uf); live and frozen values coincide,
      // so the low-order terms cancel.
      f[0] = a * hi2x(U) + b * hi2y(U) - mu * lap(U) + (Pn.xp - Pn.xm) * h2x;
      f[1] = a * hi2x(V) + b * hi2y(V) - mu * lap(V) + (Pn.yp - Pn.ym) * h2y;
    }
    const T gdiag = T(P.gdiag);
    // Rhie-Chow coefficient g from frozen velocities (roll: periodic index,
    // only used where the face flux is kept).
    auto gat = [&](int di, int dj) {
      return fabs(q(0, di, dj)) * idx + fabs(q(1, di, dj)) * idy + gdiag;
    };
    const T g = fabs(a) * idx + fabs(b) * idy + gdiag;
    const bool wall_keep = (P.bc == 2);
    T kxm = 0, kxp = 0, kym = 0, kyp = 0;
    if (P.rhiechow) {
      const T ig = recip(g);
      T qxm = 0, qxp = 0, qym = 0, qyp = 0;
      const T cx = ig * (Pn.xp - Pn.xm) * h2x * T(0.5);
      const T cy = ig * (Pn.yp - Pn.ym) * h2y * T(0.5);
      if (!iw0) {
        const T igm = recip(gat(-1, 0));
        kxm = (igm + ig) * T(0.5) * idx;
        qxm = (U.c + U.xm) * T(0.5) + (igm * (Pn.c - Pn.xmm) * h2x) * T(0.5) +
              cx - (igm + ig) * T(0.5) * (Pn.c - Pn.xm) * idx;
      } else if (wall_keep) {
        qxm = (U.c + U.xm) * T(0.5);
      }
      if (!iw1) {
        const T igp = recip(gat(1, 0));
        kxp = (igp + ig) * T(0.5) * idx;
        qxp = (U.c + U.xp) * T(0.5) + (igp * (Pn.xpp - Pn.c) * h2x) * T(0.5) +
              cx - (igp + ig) * T(0.5) * (Pn.xp - Pn.c) * idx;
      } else if (wall_keep) {
        qxp = (U.c + U.xp) * T(0.5);
      }
      if (!jw0) {
        const T igm = recip(gat(0, -1));
        kym = (igm + ig) * T(0.5) * idy;
        qym = (V.c + V.ym) * T(0.5) + (igm * (Pn.c - Pn.ymm) * h2y) * T(0.5) +
              cy - (igm + ig) * T(0.5) * (Pn.c - Pn.ym) * idy;
      } else if (wall_keep) {
        qym = (V.c + V.ym) * T(0.5);
      }
      if (!jw1) {
        const T igp = recip(gat(0, 1));
        kyp = (igp + ig) * T(0.5) * idy;
        qyp = (V.c + V.yp) * T(0.5) + (igp * (Pn.ypp - Pn.c) * h2y) * T(0.5) +
              cy - (igp + ig) * T(0.5) * (Pn.yp - Pn.c) * idy;
      } else if (wall_keep) {
        qyp = (V.c + V.yp) * T(0.5);
      }
      f[2] = (qxp - qxm) * idx + (qyp - qym) * idy;
    } else {
      f[2] = (U.xp - U.xm) * h2x + (V.yp - V.ym) * h2y;
    }
    if (!want_jac) return;

    // Live Jacobian on effective neighbours, then folded onto raw values.
    T e[5];
    const T lxx = T(P.lxx), lyy = T(P.lyy);
    // fu <- u and fv <- v share the operator uf d/dx + vf d/dy - mu lap.
    e[0] = T(2) * lxx + T(2) * lyy;
    e[1] = -lxx;
    e[2] = -lxx;
    e[3] = -lyy;
    e[4] = -lyy;
    if (a > T(0)) {
      e[0] += a * idx;
      e[1] -= a * idx;
    } else {
      e[2] += a * idx;
      e[0] -= a * idx;
    }
    if (b > T(0)) {
      e[0] += b * idy;
      e[3] -= b * idy;
    } else {
      e[4] += b * idy;
      e[0] -= b * idy;
    }
    for (int k = 0; k < 3; ++k)
      for (int s = 0; s < 5; ++s) J[0][k][s] = J[1][k][s] = 0;
    fold<INT>(J[0][0], e, i, j, nx, ny, ru.kx, ru.ky);
    fold<INT>(J[1][1], e, i, j, nx, ny, rv.kx, rv.ky);
    // fu <- p: (p_xp - p_xm) / (2 dx); fv <- p along y.
    {
      T ep[5] = {0, -h2x, h2x, 0, 0};
      fold<INT>(J[0][2], ep, i, j, nx, ny, LINEAR, LINEAR);
      T eq[5] = {0, 0, 0, -h2y, h2y};
      fold<INT>(J[1][2], eq, i, j, nx, ny, LINEAR, LINEAR);
    }
    // fp <- u, v, p.
    {
      T eu[5] = {0, 0, 0, 0, 0}, ev[5] = {0, 0, 0, 0, 0},
        ep[5] = {0, 0, 0, 0, 0};
      const T hx = T(0.5) * idx, hy = T(0.5) * idy;
      if (P.rhiechow) {
        if (!iw0 || wall_keep) {
          eu[0] -= hx;
          eu[1] -= hx;
        }
        if (!iw1 || wall_keep) {
          eu[0] += hx;
          eu[2] += hx;
        }
        if (!jw0 || wall_keep) {
          ev[0] -= hy;
          ev[3] -= hy;
        }
        if (!jw1 || wall_keep) {
          ev[0] += hy;
          ev[4] += hy;
        }
        // Pressure part (inner faces only; kxm etc. are zero at walls).
        ep[0] = (kxm + kxp) * idx + (kym + kyp) * idy;
        ep[1] = -kxm * idx;
        ep[2] = -kxp * idx;
        ep[3] = -kym * idy;
        ep[4] = -kyp * idy;
      } else {
        eu[1] = -h2x;
        eu[2] = h2x;
        ev[3] = -h2y;
        ev[4] = h2y;
      }
      fold<INT>(J[2][0], eu, i, j, nx, ny, ru.kx, ru.ky);
      fold<INT>(J[2][1], ev, i, j, nx, ny, rv.kx, rv.ky);
      fold<INT>(J[2][2], ep, i, j, nx, ny, LINEAR, LINEAR);
    }
  }

  template <bool INT = false, class Acc>
  __device__ static inline void residual(const Acc& q, int i, int j, int nx,
                                         int ny, const Params& P, T* f) {
    compute<INT>(q, i, j, nx, ny, P, f, nullptr, false);
  }
  template <bool INT = false, class Acc>
  __device__ static inline void eval(const Acc& q, int i, int j, int nx,
                                     int ny, const Params& P, T* f,
                                     T (*J)[NF][NS]) {
    compute<INT>(q, i, j, nx, ny, P, f, J, true);
  }

  // ---- Interior tiles -------------------------------------------------
  // Where a cell and its distance-2 neighbours are away from the walls,
  // compute<true>'s live Jacobian is
  //   fu <- u, fv <- v: row coefficients of the frozen velocities (a, b) at
  //     the row's cell, e0 = 2 lxx + 2 lyy + |a| idx + |b| idy,
  //     e1 = -lxx - max(a, 0) idx (x-1), e2 = -lxx + min(a, 0) idx (x+1),
  //     e3 = -lyy - max(b, 0) idy (y-1), e4 = -lyy + min(b, 0) idy (y+1);
  //   fu <- p: -/+ h2x at x-1 / x+1;  fv <- p: -/+ h2y at y-1 / y+1;
  //   fp <- u: -/+ h2x at x-1 / x+1;  fp <- v: -/+ h2y at y-1 / y+1;
  //   fp <- p (Rhie-Chow): sum over the four faces of K_f (p_c - p_nbr),
  //     K = (ig_c + ig_nbr) idx^2 / 2 across x-faces (idy^2 / 2 across
  //     y-faces), ig = 1 / g of the cell's frozen velocities.
  // J^T w at a cell is then a 5-point column formula in w and the row
  // coefficients of the cell's neighbours, with no Jacobian array and no
  // product buffer. The interior tile kernels below use it for the loss
  // adjoint (w = f) and the normal operator (w = J x, itself a 5-point
  // row formula). Results equal the generic path's up to rounding.
  static constexpr bool INTERIOR_TILES = true;

  struct Coefs {
    T idx, idy, lxx, lyy, h2x, h2y, e0c, ckx, cky, gdiag;
    bool rc;
    __device__ explicit Coefs(const Params& P)
        : idx(T(P.idx)), idy(T(P.idy)), lxx(T(P.lxx)), lyy(T(P.lyy)),
          h2x(T(P.h2x)), h2y(T(P.h2y)),
          e0c(T(2) * T(P.lxx) + T(2) * T(P.lyy)),
          ckx(T(0.5) * T(P.idx) * T(P.idx)),
          cky(T(0.5) * T(P.idy) * T(P.idy)), gdiag(T(P.gdiag)),
          rc(P.rhiechow != 0) {}
    __device__ inline T ig(T a, T b) const {
      return recip(fabs(a) * idx + fabs(b) * idy + gdiag);
    }
  };

  // y = J x at an interior cell. x: field 0 at the cell, fields XN apart,
  // rows X1 apart; uv: u at the cell (v UN further), rows U1 apart; ig:
  // ig at the cell, rows G1 apart.
  template <int X1, int XN, int U1, int UN, int G1>
  __device__ static inline void jx_row(const T* x, const T* uv, const T* ig,
                                       const Coefs& K, T* y) {
    const T a = uv[0], b = uv[UN];
    const T e0 = K.e0c + fabs(a) * K.idx + fabs(b) * K.idy;
    const T e1 = -K.lxx - fmax(a, T(0)) * K.idx;
    const T e2 = -K.lxx + fmin(a, T(0)) * K.idx;
    const T e3 = -K.lyy - fmax(b, T(0)) * K.idy;
    const T e4 = -K.lyy + fmin(b, T(0)) * K.idy;
    const T* xu = x;
    const T* xv = x + XN;
    const T* xp = x + 2 * XN;
    y[0] = e0 * xu[0] + e1 * xu[-X1] + e2 * xu[X1] + e3 * xu[-1] +
           e4 * xu[1] + K.h2x * (xp[X1] - xp[-X1]);
    y[1] = e0 * xv[0] + e1 * xv[-X1] + e2 * xv[X1] + e3 * xv[-1] +
           e4 * xv[1] + K.h2y * (xp[1] - xp[-1]);
    T yp = K.h2x * (xu[X1] - xu[-X1]) + K.h2y * (xv[1] - xv[-1]);
    if (K.rc) {
      const T g = ig[0];
      const T kxm = (ig[-G1] + g) * K.ckx, kxp = (ig[G1] + g) * K.ckx;
      const T kym = (ig[-1] + g) * K.cky, kyp = (ig[1] + g) * K.cky;
      const T pc = xp[0];
      yp += kxm * (pc - xp[-X1]) + kxp * (pc - xp[X1]) +
            kym * (pc - xp[-1]) + kyp * (pc - xp[1]);
    }
    y[2] = yp;
  }

  // g = J^T w at an interior cell; w, uv and ig addressed as in jx_row.
  // Needs w and ig on the cell and its 4 neighbours, u at x-1, x, x+1 and
  // v at y-1, y, y+1.
  template <int W1, int WN, int U1, int UN, int G1>
  __device__ static inline void jt_column(const T* w, const T* uv,
                                          const T* ig, const Coefs& K,
                                          T* g) {
    const T uc = uv[0], uxp = uv[U1], uxm = uv[-U1];
    const T vc = uv[UN], vyp = uv[UN + 1], vym = uv[UN - 1];
    const T e0 = K.e0c + fabs(uc) * K.idx + fabs(vc) * K.idy;
    const T e1n = -K.lxx - fmax(uxp, T(0)) * K.idx;  // row x+1, this column
    const T e2n = -K.lxx + fmin(uxm, T(0)) * K.idx;  // row x-1
    const T e3n = -K.lyy - fmax(vyp, T(0)) * K.idy;  // row y+1
    const T e4n = -K.lyy + fmin(vym, T(0)) * K.idy;  // row y-1
    const T* wu = w;
    const T* wv = w + WN;
    const T* wp = w + 2 * WN;
    g[0] = e0 * wu[0] + e1n * wu[W1] + e2n * wu[-W1] + e3n * wu[1] +
           e4n * wu[-1] + K.h2x * (wp[-W1] - wp[W1]);
    g[1] = e0 * wv[0] + e1n * wv[W1] + e2n * wv[-W1] + e3n * wv[1] +
           e4n * wv[-1] + K.h2y * (wp[-1] - wp[1]);
    T gp = K.h2x * (wu[-W1] - wu[W1]) + K.h2y * (wv[-1] - wv[1]);
    if (K.rc) {
      const T c = ig[0];
      const T kxm = (ig[-G1] + c) * K.ckx, kxp = (ig[G1] + c) * K.ckx;
      const T kym = (ig[-1] + c) * K.cky, kyp = (ig[1] + c) * K.cky;
      const T pc = wp[0];
      gp += kxm * (pc - wp[-W1]) + kxp * (pc - wp[W1]) +
            kym * (pc - wp[-1]) + kyp * (pc - wp[1]);
    }
    g[2] = gp;
  }

  template <int TX, int TY>
  static constexpr size_t interior_adjoint_smem() {
    return sizeof(T) *
           (size_t)(3 * (TX + 6) * (TY + 6) + 4 * (TX + 2) * (TY + 2));
  }

  // Loss adjoint grad = scale J^T f of one interior TX x TY tile at
  // (x0, y0) (the tile grown by 3 lies inside the grid): the state is
  // staged with halo 3, f and ig are evaluated on the tile grown by 1, then
  // every cell applies jt_column. Adds the tile's f_e^2 to ssq[e].
  template <int TX, int TY>
  __device__ static inline void interior_adjoint_tile(
      const FieldPtrs<T, 3>& in, const OutPtrs<T, 3>& grad, int n0, int n1,
      int x0, int y0, T scale, const Params& P, T* smem, double* ssq,
      int v0, int v1) {
    constexpr int H = 3, S0 = TX + 2 * H, S1 = TY + 2 * H, NS1 = S0 * S1;
    constexpr int E0 = TX + 2, E1 = TY + 2, NE1 = E0 * E1;
    constexpr int NPT = TX * TY / 256;
    static_assert((TX * TY) % 256 == 0, "tile must be a multiple of 256");
    T* st = smem;           // state [3][S0][S1]
    T* F = st + 3 * NS1;    // residuals [3][E0][E1]
    T* IG = F + 3 * NE1;    // 1 / g [E0][E1]
    stage_tile<T, 3, TX, TY, H>(in, st, x0, y0, n0, n1);
    __syncthreads();
    const Coefs K(P);
    for (int q = threadIdx.x; q < NE1; q += 256) {
      const int ea = q / E1, eb = q - ea * E1;
      const T* cell = st + (ea + H - 1) * S1 + (eb + H - 1);
      LdsAcc<T, S0, S1> acc{cell};
      T f[3];
      residual<true>(acc, x0 - 1 + ea, y0 - 1 + eb, n0, n1, P, f);
      if (K.rc) IG[q] = K.ig(cell[0], cell[NS1]);
      F[q] = f[0];
      F[NE1 + q] = f[1];
      F[2 * NE1 + q] = f[2];
      if (ea >= 1 && ea <= TX && eb >= 1 && eb <= TY &&
          x0 - 1 + ea < v0 && y0 - 1 + eb < v1) {
#pragma unroll
        for (int e = 0; e < 3; ++e) ssq[e] += (double)f[e] * (double)f[e];
      }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < NPT; ++p) {
      const int q = threadIdx.x + 256 * p;
      const int a = q / TY, b = q - a * TY;
      if (x0 + a >= v0 || y0 + b >= v1) continue;
      const int e = (a + 1) * E1 + (b + 1);
      T g[3];
      jt_column<E1, NE1, S1, NS1, E1>(F + e, st + (a + H) * S1 + (b + H),
                                      IG + e, K, g);
      const size_t c = (size_t)(x0 + a) * n1 + (y0 + b);
#pragma unroll
      for (int k = 0; k < 3; ++k)
        if (grad.f[k]) grad.f[k][c] = g[k] * scale;
    }
  }

  template <int TX, int TY>
  static constexpr size_t interior_normal_smem() {
    return sizeof(T) *
           (size_t)(6 * (TX + 4) * (TY + 4) + 3 * (TX + 2) * (TY + 2));
  }

  // Normal operator N = J^T J of one interior tile (modes 0, 1, 2 of
  // k_normal_apply). u, v and x are staged with halo 2, and ig is computed
  // from the staged u, v on the way into LDS; y = J x on the tile grown by
  // 1; then jt_column and the mode's epilogue (b and d are loaded before
  // the first barrier, so their latency hides behind the staging).
  template <int TX, int TY>
  __device__ static inline void interior_normal_tile(
      const FieldPtrs<T, 3>& st_in, const FieldPtrs<T, 3>& x_in,
      const FieldPtrs<T, 3>& b_in, const FieldPtrs<T, 3>& d_in,
      const OutPtrs<T, 3>& y, int n0, int n1, T omega, int mode,
      const Params& P, int x0, int y0, T* smem, int v0, int v1, bool dot,
      double& dacc) {
    constexpr int A0 = TX + 4, A1 = TY + 4, NA = A0 * A1;
    constexpr int E0 = TX + 2, E1 = TY + 2, NE1 = E0 * E1;
    constexpr int NPT = TX * TY / 256;
    constexpr int NL = (NA + 255) / 256;
    static_assert((TX * TY) % 256 == 0, "tile must be a multiple of 256");
    T* UV = smem;        // u, v [2][A0][A1]
    T* X = UV + 2 * NA;  // x [3][A0][A1]
    T* IG = X + 3 * NA;  // 1 / g [A0][A1]
    T* Y = IG + NA;      // J x [3][E0][E1]
    const Coefs K(P);
    T bv[NPT][3], dv[NPT][3];
    if (mode != 0) {
#pragma unroll
      for (int p = 0; p < NPT; ++p) {
        const int q = threadIdx.x + 256 * p;
        const int a = q / TY, b = q - a * TY;
        const bool ok = x0 + a < v0 && y0 + b < v1;
        const size_t c = ok ? (size_t)(x0 + a) * n1 + (y0 + b) : 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          bv[p][k] = b_in.f[k][c];
          dv[p][k] = mode == 2 ? d_in.f[k][c] : T(1);
        }
      }
    }
    {
      // The tile grown by 2 starts inside the grid (x0, y0 >= 3); a partial
      // tile's far side is clamped (its values only reach the products of
      // cells outside [.., v0) x [.., v1), which are not stored).
      const bool inside = x0 + TX + 2 <= n0 && y0 + TY + 2 <= n1;
      T vu[NL], vv[NL], vx[3][NL];
#pragma unroll
      for (int r = 0; r < NL; ++r) {
        const int q = threadIdx.x + 256 * r;
        if (r + 1 < NL || q < NA) {
          const int a = q / A1, b = q - a * A1;
          const int gi = inside ? x0 - 2 + a : min(x0 - 2 + a, n0 - 1);
          const int gj = inside ? y0 - 2 + b : min(y0 - 2 + b, n1 - 1);
          const size_t o = (size_t)gi * n1 + gj;
          vu[r] = st_in.f[0][o];
          vv[r] = st_in.f[1][o];
#pragma unroll
          for (int k = 0; k < 3; ++k) vx[k][r] = x_in.f[k][o];
        }
      }
#pragma unroll
      for (int r = 0; r < NL; ++r) {
        const int q = threadIdx.x + 256 * r;
        if (r + 1 < NL || q < NA) {
          UV[q] = vu[r];
          UV[NA + q] = vv[r];
#pragma unroll
          for (int k = 0; k < 3; ++k) X[k * NA + q] = vx[k][r];
          if (K.rc) IG[q] = K.ig(vu[r], vv[r]);
        }
      }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < NE1; q += 256) {
      const int ea = q / E1, eb = q - ea * E1;
      const int s = (ea + 1) * A1 + (eb + 1);
      T v[3];
      jx_row<A1, NA, A1, NA, A1>(X + s, UV + s, IG + s, K, v);
      Y[q] = v[0];
      Y[NE1 + q] = v[1];
      Y[2 * NE1 + q] = v[2];
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < NPT; ++p) {
      const int q = threadIdx.x + 256 * p;
      const int a = q / TY, b = q - a * TY;
      if (x0 + a >= v0 || y0 + b >= v1) continue;
      const int e = (a + 1) * E1 + (b + 1);
      const int s = (a + 2) * A1 + (b + 2);
      T g[3];
      jt_column<E1, NE1, A1, NA, A1>(Y + e, UV + s, IG + s, K, g);
      const size_t c = (size_t)(x0 + a) * n1 + (y0 + b);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        T out;
        if (mode == 0)
          out = g[k];
        else if (mode == 1)
          out = bv[p][k] - g[k];
        else
          out = X[k * NA + s] + omega * (bv[p][k] - g[k]) / dv[p][k];
        y.f[k][c] = out;
        if (dot) dacc += (double)X[k * NA + s] * (double)out;
      }
    }
  }
};

// Extra loss terms of the cavity's inverse problems (ns_steady.py:396-411),
// as additional equations on the unknowns (u, v, p):
//   e0 = mask (u - u_imp) cimp,  e1 = mask (v - v_imp) cimp   (imposed
//        points, cimp = alpha N / n_imp),
//   e2..e5 = kreg (u_xx, u_yy, v_xx, v_yy)   (laplace_split of the live,
//        halo-corrected 5-point stencils).
// Terms that are off (no mask, kreg = 0) are zero. Its adjoint is added to
// the cavity's (ACCUM): both kernels together give the reference's loss
// gradient with the extra residuals appended.
struct CavityExtrasParams {
  double idx2, idy2;   // 1/dx^2, 1/dy^2
  double kreg, cimp;   // regularisation factor, imposed-point weight
  int bc;
  const void* mask;    // T arrays (nx, ny), or null: no imposed points
  const void* uimp;
  const void* vimp;
};

template <class T>
struct CavityExtrasOp {
  using Params = CavityExtrasParams;
  static constexpr int NF = 3, NE = 6, NS = 5, R = 1, RJ = 1;
  static constexpr bool ACCUM = true;
  static constexpr int SX[NS] = {0, -1, 1, 0, 0};
  static constexpr int SY[NS] = {0, 0, 0, -1, 1};
  static constexpr int NJ = 14;
  static constexpr int JE[NJ] = {0, 1, 2, 2, 2, 3, 3, 3, 4, 4, 4, 5, 5, 5};
  static constexpr int JV[NJ] = {0, 1, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1};
  static constexpr int JS[NJ] = {0, 0, 0, 1, 2, 0, 3, 4, 0, 1, 2, 0, 3, 4};

  // Halo-corrected distance-1 neighbours (c, xm, xp, ym, yp) of field k.
  template <bool INT, class Acc>
  __device__ static inline void nbr5(const Acc& q, int k, int i, int j,
                                     int nx, int ny, const FaceRule& r,
                                     T* s) {
    const T c = q(k, 0, 0);
    const T rxm = q(k, -1, 0), rxp = q(k, 1, 0);
    const T rym = q(k, 0, -1), ryp = q(k, 0, 1);
    s[0] = c;
    if (INT) {
      s[1] = rxm;
      s[2] = rxp;
      s[3] = rym;
      s[4] = ryp;
      return;
    }
    s[1] = (i == 0) ? ghost(r.kx, rxp, c, T(r.wxm)) : rxm;
    s[2] = (i == nx - 1) ? ghost(r.kx, rxm, c, T(r.wxp)) : rxp;
    s[3] = (j == 0) ? ghost(r.ky, ryp, c, T(r.wym)) : rym;
    s[4] = (j == ny - 1) ? ghost(r.ky, rym, c, T(r.wyp)) : ryp;
  }

  template <bool INT, class Acc>
  __device__ static inline void compute(const Acc& q, int i, int j, int nx,
                                        int ny, const Params& P, T* f,
                                        T (*J)[NF][NS], bool want_jac) {
    const FaceRule ru = cavity_rule(P.bc, 0), rv = cavity_rule(P.bc, 1);
    T U[5], V[5];
    nbr5<INT>(q, 0, i, j, nx, ny, ru, U);
    nbr5<INT>(q, 1, i, j, nx, ny, rv, V);
    T m = 0, ui = 0, vi = 0;
    if (P.mask) {
      const size_t c = (size_t)i * ny + j;
      m = static_cast<const T*>(P.mask)[c];
      ui = static_cast<const T*>(P.uimp)[c];
      vi = static_cast<const T*>(P.vimp)[c];
    }
    const T ci = T(P.cimp), kr = T(P.kreg);
    const T ix2 = T(P.idx2), iy2 = T(P.idy2);
    f[0] = m * (U[0] - ui) * ci;
    f[1] = m * (V[0] - vi) * ci;
    f[2] = (U[2] - T(2) * U[0] + U[1]) * ix2 * kr;
    f[3] = (U[4] - T(2) * U[0] + U[3]) * iy2 * kr;
    f[4] = (V[2] - T(2) * V[0] + V[1]) * ix2 * kr;
    f[5] = (V[4] - T(2) * V[0] + V[3]) * iy2 * kr;
    if (!want_jac) return;
    for (int e = 0; e < NE; ++e)
      for (int k = 0; k < NF; ++k)
        for (int s = 0; s < NS; ++s) J[e][k][s] = 0;
    J[0][0][0] = m * ci;
    J[1][1][0] = m * ci;
    const T ax = ix2 * kr, ay = iy2 * kr;
    const T exx[5] = {-T(2) * ax, ax, ax, 0, 0};
    const T eyy[5] = {-T(2) * ay, 0, 0, ay, ay};
    fold<INT>(J[2][0], exx, i, j, nx, ny, ru.kx, ru.ky);
    fold<INT>(J[3][0], eyy, i, j, nx, ny, ru.kx, ru.ky);
    fold<INT>(J[4][1], exx, i, j, nx, ny, rv.kx, rv.ky);
    fold<INT>(J[5][1], eyy, i, j, nx, ny, rv.kx, rv.ky);
  }

  template <bool INT = false, class Acc>
  __device__ static inline void residual(const Acc& q, int i, int j, int nx,
                                         int ny, const Params& P, T* f) {
    compute<INT>(q, i, j, nx, ny, P, f, nullptr, false);
  }
  template <bool INT = false, class Acc>
  __device__ static inline void eval(const Acc& q, int i, int j, int nx,
                                     int ny, const Params& P, T* f,
                                     T (*J)[NF][NS]) {
    compute<INT>(q, i, j, nx, ny, P, f, J, true);
  }
};

inline CavityExtrasParams extras_params(double dx, double dy, int bc,
                                        double kreg, double cimp,
                                        const void* mask, const void* uimp,
                                        const void* vimp) {
  return CavityExtrasParams{1.0 / (dx * dx), 1.0 / (dy * dy), kreg, cimp,
                            bc, mask, uimp, vimp};
}

}  // namespace odil

using namespace odil;

namespace {
template <class T>
int cavity_dispatch_grad(const void* const* fields, void* const* grad,
                         double* partial, int nx, int ny,
                         const CavityParams& P, hipStream_t s, int* nb) {
  const T scale = T(2.0 / ((double)nx * ny));
  if constexpr (sizeof(T) == 4)
    return launch_loss_grad_t<T, CavityOp<T>, kGradTx32, kGradTy32>(
        fields, grad, partial, nx, ny, scale, P, s, nb);
  else
    return launch_loss_grad_t<T, CavityOp<T>, kGradTx64, kGradTy64>(
        fields, grad, partial, nx, ny, scale, P, s, nb);
}
}  // namespace

extern "C" {

// Residuals fu, fv, fp of the cavity at state (u, v, p); `res` may be NULL
// (sums of squares only). `partial` receives per-block sums of squares
// (3 x nblocks doubles, see odil_cavity_nblocks).
int odil_cavity_residual(int dtype, int nx, int ny, double dx, double dy,
                         double mu, int bc, int rhiechow,
                         const void* const* fields, void* const* res,
                         double* partial, void* stream) {
  if (nx < 3 || ny < 3 || !fields) {
    set_error("cavity_residual: need nx, ny >= 3 and fields");
    return kErrArg;
  }
  CavityParams P = cavity_params(dx, dy, mu, bc, rhiechow);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    return launch_residual<double, CavityOp<double>>(fields, res, partial, nx,
                                                     ny, P, s, nullptr);
  if (dtype == kF32)
    return launch_residual<float, CavityOp<float>>(fields, res, partial, nx,
                                                   ny, P, s, nullptr);
  set_error("cavity_residual: bad dtype %d", dtype);
  return kErrArg;
}

// Jacobian coefficient fields, 27 entries in the order of the
// odil_cavity_jac_entries table, each a full (nx, ny) field.
int odil_cavity_jac(int dtype, int nx, int ny, double dx, double dy,
                    double mu, int bc, int rhiechow, const void* const* fields,
                    void* coef, void* const* res, void* stream) {
  if (nx < 3 || ny < 3 || !fields || !coef) {
    set_error("cavity_jac: need nx, ny >= 3, fields and coef");
    return kErrArg;
  }
  CavityParams P = cavity_params(dx, dy, mu, bc, rhiechow);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    return launch_jac<double, CavityOp<double>>(fields, coef, res, nx, ny, P,
                                                s);
  if (dtype == kF32)
    return launch_jac<float, CavityOp<float>>(fields, coef, res, nx, ny, P, s);
  set_error("cavity_jac: bad dtype %d", dtype);
  return kErrArg;
}

// Table of the Jacobian entries: for q < 27, ent[3q..3q+2] =
// (equation, field, shift index); shifts[2s..2s+1] = (dx, dy).
int odil_cavity_jac_entries(int* ent, int* shifts) {
  if (ent)
    for (int q = 0; q < CavityOp<double>::NJ; ++q) {
      ent[3 * q] = CavityOp<double>::JE[q];
      ent[3 * q + 1] = CavityOp<double>::JV[q];
      ent[3 * q + 2] = CavityOp<double>::JS[q];
    }
  if (shifts)
    for (int s = 0; s < CavityOp<double>::NS; ++s) {
      shifts[2 * s] = CavityOp<double>::SX[s];
      shifts[2 * s + 1] = CavityOp<double>::SY[s];
    }
  return CavityOp<double>::NJ;
}

// Fused loss and adjoint: grad_k = (2/N) sum_e J_e,k^T f_e for (u, v, p);
// `partial` receives per-block sums of f_e^2 (3 x nblocks doubles).
int odil_cavity_loss_grad(int dtype, int nx, int ny, double dx, double dy,
                          double mu, int bc, int rhiechow,
                          const void* const* fields, void* const* grad,
                          double* partial, int* nblocks, void* stream) {
  if (nx < 3 || ny < 3 || !fields || !grad) {
    set_error("cavity_loss_grad: need nx, ny >= 3, fields and grad");
    return kErrArg;
  }
  CavityParams P = cavity_params(dx, dy, mu, bc, rhiechow);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    return cavity_dispatch_grad<double>(fields, grad, partial, nx, ny, P, s,
                                        nblocks);
  if (dtype == kF32)
    return cavity_dispatch_grad<float>(fields, grad, partial, nx, ny, P, s,
                                       nblocks);
  set_error("cavity_loss_grad: bad dtype %d", dtype);
  return kErrArg;
}

// Extra cavity loss terms (imposed points, regularisation; see
// CavityExtrasOp): residual fields (6, any may be NULL) with per-block sums
// of squares; mask/uimp/vimp NULL disables the imposed-point terms.
int odil_cavity_extras_residual(int dtype, int nx, int ny, double dx,
                                double dy, int bc, double kreg, double cimp,
                                const void* mask, const void* uimp,
                                const void* vimp, const void* const* fields,
                                void* const* res, double* partial,
                                void* stream) {
  if (nx < 3 || ny < 3 || !fields || (mask && (!uimp || !vimp))) {
    set_error("cavity_extras_residual: bad arguments");
    return kErrArg;
  }
  const CavityExtrasParams P =
      extras_params(dx, dy, bc, kreg, cimp, mask, uimp, vimp);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    return launch_residual<double, CavityExtrasOp<double>>(
        fields, res, partial, nx, ny, P, s, nullptr);
  if (dtype == kF32)
    return launch_residual<float, CavityExtrasOp<float>>(
        fields, res, partial, nx, ny, P, s, nullptr);
  set_error("cavity_extras_residual: bad dtype %d", dtype);
  return kErrArg;
}

// Adds the extra terms' adjoint (2/N) J_x^T f_x to grad (u, v, p arrays;
// the p gradient gets zeros added) with per-block sums of squares of the
// six extra equations in `partial` (6 x odil_cavity_nblocks(1, ...)).
int odil_cavity_extras_loss_grad(int dtype, int nx, int ny, double dx,
                                 double dy, int bc, double kreg, double cimp,
                                 const void* mask, const void* uimp,
                                 const void* vimp, const void* const* fields,
                                 void* const* grad, double* partial,
                                 int* nblocks, void* stream) {
  if (nx < 3 || ny < 3 || !fields || !grad || (mask && (!uimp || !vimp))) {
    set_error("cavity_extras_loss_grad: bad arguments");
    return kErrArg;
  }
  const CavityExtrasParams P =
      extras_params(dx, dy, bc, kreg, cimp, mask, uimp, vimp);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    return launch_loss_grad_t<double, CavityExtrasOp<double>, kGradTx64,
                              kGradTy64>(fields, grad, partial, nx, ny,
                                       2.0 / ((double)nx * ny), P, s,
                                       nblocks);
  if (dtype == kF32)
    return launch_loss_grad_t<float, CavityExtrasOp<float>, kGradTx32,
                              kGradTy32>(fields, grad, partial, nx, ny,
                                       (float)(2.0 / ((double)nx * ny)), P,
                                       s, nblocks);
  set_error("cavity_extras_loss_grad: bad dtype %d", dtype);
  return kErrArg;
}

// Jacobian coefficient fields of the extra equations (14 x N), entries as
// odil_cavity_extras_jac_entries lists them; `res` (6 fields or NULL)
// optionally receives the residuals.
int odil_cavity_extras_jac(int dtype, int nx, int ny, double dx, double dy,
                           int bc, double kreg, double cimp, const void* mask,
                           const void* uimp, const void* vimp,
                           const void* const* fields, void* coef,
                           void* const* res, void* stream) {
  if (nx < 3 || ny < 3 || !fields || !coef || (mask && (!uimp || !vimp))) {
    set_error("cavity_extras_jac: bad arguments");
    return kErrArg;
  }
  const CavityExtrasParams P =
      extras_params(dx, dy, bc, kreg, cimp, mask, uimp, vimp);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    return launch_jac<double, CavityExtrasOp<double>>(fields, coef, res,
                                                      nx, ny, P, s);
  if (dtype == kF32)
    return launch_jac<float, CavityExtrasOp<float>>(fields, coef, res, nx,
                                                    ny, P, s);
  set_error("cavity_extras_jac: bad dtype %d", dtype);
  return kErrArg;
}

// Per-block partial counts of the extras kernels (the cavity's tiles).
int odil_cavity_extras_nblocks(int kind, int dtype, int nx, int ny) {
  if (kind == 0) return residual_nblocks(nx, ny, CavityExtrasOp<float>::R);
  return grad_nblocks(dtype == kF32, nx, ny);
}

int odil_cavity_extras_jac_entries(int* ent, int* shifts) {
  using E = CavityExtrasOp<double>;
  if (ent)
    for (int q = 0; q < E::NJ; ++q) {
      ent[3 * q] = E::JE[q];
      ent[3 * q + 1] = E::JV[q];
      ent[3 * q + 2] = E::JS[q];
    }
  if (shifts)
    for (int s = 0; s < E::NS; ++s) {
      shifts[2 * s] = E::SX[s];
      shifts[2 * s + 1] = E::SY[s];
    }
  return E::NJ;
}

// Number of per-block partials written by the residual (kind 0),
// loss-grad (kind 1) and normal-operator product with dot (kind 2) kernels
// for the given dtype.
int odil_cavity_nblocks(int kind, int dtype, int nx, int ny) {
  if (kind == 0) return residual_nblocks(nx, ny, CavityOp<float>::R);
  if (kind == 2) return normal_nblocks<CavityOp<float>>(dtype == kF32, nx, ny);
  return loss_grad_nblocks<CavityOp<float>>(dtype == kF32, nx, ny);
}

// Matrix-free normal operator J^T J with J recomputed from `state`
// (mode 0: y = N x, 1: y = b - N x, 2: Jacobi sweep with diagonal d,
// 3: y = diag N). Fields are (u, v, p) triples of the given dtype.
int odil_cavity_normal_apply(int dtype, int nx, int ny, double dx, double dy,
                             double mu, int bc, int rhiechow,
                             const void* const* state, const void* const* x,
                             const void* const* b, const void* const* d,
                             void* const* y, double omega, int mode,
                             void* stream) {
  if (nx < 3 || ny < 3) {
    set_error("cavity_normal_apply: need nx, ny >= 3");
    return kErrArg;
  }
  CavityParams P = cavity_params(dx, dy, mu, bc, rhiechow);
  return dispatch_normal_apply<CavityOp<double>, CavityOp<float>>(
      dtype, state, x, b, d, y, nx, ny, omega, mode, P,
      static_cast<hipStream_t>(stream));
}

// y = N x (mode 0 of odil_cavity_normal_apply) and *dot = sum over the
// cells and fields of x * y (FP64; per-wavefront partials, fixed-order sum).
int odil_cavity_normal_apply_dot(int dtype, int nx, int ny, double dx,
                                 double dy, double mu, int bc, int rhiechow,
                                 const void* const* state,
                                 const void* const* x, void* const* y,
                                 double* partial, double* dot,
                                 void* stream) {
  if (nx < 3 || ny < 3 || !state || !x || !y || !partial || !dot) {
    set_error("cavity_normal_apply_dot: need nx, ny >= 3, state, x, y, "
              "partial and dot");
    return kErrArg;
  }
  CavityParams P = cavity_params(dx, dy, mu, bc, rhiechow);
  hipStream_t s = static_cast<hipStream_t>(stream);
  int nb = 0;
  const int st = dispatch_normal_apply<CavityOp<double>, CavityOp<float>>(
      dtype, state, x, nullptr, nullptr, y, nx, ny, 0.0, 0, P, s, partial,
      &nb);
  if (st != kOk) return st;
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, s, partial,
                     nb, dot);
  return check_launch("normal_apply_dot");
}

// out = scale * J^T f(state) plus per-block sums of f^2 (as loss_grad).
int odil_cavity_jtf(int dtype, int nx, int ny, double dx, double dy,
                    double mu, int bc, int rhiechow,
                    const void* const* fields, void* const* out,
                    double scale, double* partial, int* nblocks,
                    void* stream) {
  if (nx < 3 || ny < 3 || !fields || !out) {
    set_error("cavity_jtf: bad arguments");
    return kErrArg;
  }
  CavityParams P = cavity_params(dx, dy, mu, bc, rhiechow);
  return dispatch_jtf<CavityOp<double>, CavityOp<float>>(
      dtype, fields, out, partial, nx, ny, scale, P,
      static_cast<hipStream_t>(stream), nblocks);
}

}  // extern "C"

// ---- allocation-free loss + adjoint ----------------------------------------
// The per-wavefront partial sums of squares of odil_cavity_loss_grad_noalloc
// live in this file-scope device array (one copy per device, held until the
// process ends): 3 equations x kNoAllocParts doubles = 1.5 MiB. Host code
// never passes `scratch` itself to a kernel (its host-side name is not a
// device address); it uses the per-device address from hipGetSymbolAddress.
namespace odil {
namespace {
constexpr int kNoAllocParts = 65536;
constexpr int kNoAllocMaxDevices = 64;
}  // namespace
}  // namespace odil

static __device__ __align__(256) double scratch[3 * odil::kNoAllocParts];

namespace odil {
namespace {
std::mutex g_scratch_mu;
double* g_scratch_ptr[kNoAllocMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
int scratch_address(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error("cavity_loss_grad_noalloc: hipGetDevice: %s",
              hipGetErrorString(e));
    return kErrLaunch;
  }
  if (dev < 0 || dev >= kNoAllocMaxDevices) {
    set_error("cavity_loss_grad_noalloc: device %d beyond %d", dev,
              kNoAllocMaxDevices);
    return kErrCapacity;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("cavity_loss_grad_noalloc: hipGetSymbolAddress: %s",
                hipGetErrorString(e));
      return kErrLaunch;
    }
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return kOk;
}
}  // namespace
}  // namespace odil

extern "C" {

// As odil_cavity_loss_grad, but allocates nothing and needs no workspace:
// the per-wavefront partials go to the library's device array and their
// fixed-order sums to ssq[0..2] (device; sum of f_e^2 per equation, the same
// values odil_reduce_partials gives for odil_cavity_loss_grad's partials).
// Grids needing more than 65536 partials per equation (square grids above
// 2022 x 2022 cells in FP64, 2822 x 2822 in FP32) are rejected with
// ODIL_ERR_CAPACITY. Calls that overlap
// on one device race on the shared array.
int odil_cavity_loss_grad_noalloc(int dtype, int nx, int ny, double dx,
                                  double dy, double mu, int bc, int rhiechow,
                                  const void* const* fields,
                                  void* const* grad, double* ssq,
                                  void* stream) {
  if (nx < 3 || ny < 3 || !fields || !grad || !ssq) {
    set_error("cavity_loss_grad_noalloc: need nx, ny >= 3, fields, grad "
              "and ssq");
    return kErrArg;
  }
  if (dtype != kF32 && dtype != kF64) {
    set_error("cavity_loss_grad_noalloc: bad dtype %d", dtype);
    return kErrArg;
  }
  const int need = loss_grad_nblocks<CavityOp<float>>(dtype == kF32, nx, ny);
  if (need > kNoAllocParts) {
    set_error("cavity_loss_grad_noalloc: %d x %d needs %d partials per "
              "equation, more than the %d of the library's scratch",
              nx, ny, need, kNoAllocParts);
    return kErrCapacity;
  }
  double* part = nullptr;
  int st = scratch_address(&part);
  if (st != kOk) return st;
  CavityParams P = cavity_params(dx, dy, mu, bc, rhiechow);
  hipStream_t s = static_cast<hipStream_t>(stream);
  int nb = 0;
  st = dtype == kF64 ? cavity_dispatch_grad<double>(fields, grad, part, nx,
                                                    ny, P, s, &nb)
                     : cavity_dispatch_grad<float>(fields, grad, part, nx,
                                                   ny, P, s, &nb);
  if (st != kOk) return st;
  if (nb != need) {
    set_error("cavity_loss_grad_noalloc: partial count %d != %d", nb, need);
    return kErrArg;
  }
  hipLaunchKernelGGL(k_reduce_partials, dim3(3), dim3(256), 0, s, part, nb,
                     ssq);
  return check_launch("cavity_loss_grad_noalloc reduce");
}

}  // extern "C"
