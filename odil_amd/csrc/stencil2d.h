// Generic 2D machinery for ODIL discrete operators that are linear in their
// live fields once the frozen fields are fixed.
//
// An operator type `Op<T>` provides
//   NF, NE          number of fields and equations
//   NS, SX[], SY[]  the shift set of its stencil Jacobian; entry s multiplies
//                   the field value at (i + SX[s], j + SY[s])
//   R               radius of state values read by eval() around a cell
//   eval(acc, i, j, P, f[NE], J[NE][NF][NS])
//                   residual and live Jacobian coefficients of one cell;
//                   acc(k, di, dj) returns field k at (i+di, j+dj)
//   NJ, JE[], JV[], JS[]   the structurally non-zero Jacobian entries
//                   (equation, field, shift index) written by jac_coeffs.
//
// Kernels here: residual (+ per-equation sums of squares), Jacobian
// coefficient fields, and the fused loss adjoint
//   grad_k = (2/N) sum_e J_e,k^T f_e
// (the tape gradient of sum_e mean(f_e^2), util_op.py:627-654, for operators
// whose live dependence is linear).
#pragma once

#include "common.h"

namespace odil {

template <class T, int NF>
struct GlobalAcc {
  const T* const* f;
  int n0, n1, i, j;
  __device__ inline T operator()(int k, int di, int dj) const {
    int ii = min(max(i + di, 0), n0 - 1);
    int jj = min(max(j + dj, 0), n1 - 1);
    return f[k][(size_t)ii * n1 + jj];
  }
};

template <class T>
struct LdsAcc {
  const T* s;
  int e0, e1;  // extents of one field's tile
  int li, lj;  // local position of the cell
  __device__ inline T operator()(int k, int di, int dj) const {
    return s[(k * e0 + li + di) * e1 + lj + dj];
  }
};

template <class T, int NF>
struct FieldPtrs {
  const T* f[NF];
};
template <class T, int NF>
struct OutPtrs {
  T* f[NF];
};

// Residual f_e at every cell, and per-block partial sums of f_e^2.
template <class T, class Op>
__global__ void __launch_bounds__(256)
    k_residual(FieldPtrs<T, Op::NF> in, OutPtrs<T, Op::NE> out,
               double* __restrict__ partial, int n0, int n1,
               typename Op::Params P) {
  __shared__ double red[4];
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  double ssq[Op::NE];
  for (int e = 0; e < Op::NE; ++e) ssq[e] = 0;
  if (i < n0 && j < n1) {
    GlobalAcc<T, Op::NF> acc{in.f, n0, n1, i, j};
    T f[Op::NE];
    Op::residual(acc, i, j, n0, n1, P, f);
    const size_t c = (size_t)i * n1 + j;
    for (int e = 0; e < Op::NE; ++e) {
      if (out.f[e]) out.f[e][c] = f[e];
      ssq[e] = (double)f[e] * (double)f[e];
    }
  }
  if (partial) {
    const int b = blockIdx.y * gridDim.x + blockIdx.x;
    const int nb = gridDim.x * gridDim.y;
    for (int e = 0; e < Op::NE; ++e) {
      double s = block_sum(ssq[e], red);
      if (threadIdx.x == 0) partial[(size_t)e * nb + b] = s;
    }
  }
}

// Jacobian coefficient fields: coef[q] for q < NJ holds entry
// (JE[q], JV[q], JS[q]); optionally also the residual.
template <class T, class Op>
__global__ void __launch_bounds__(256)
    k_jac(FieldPtrs<T, Op::NF> in, T* __restrict__ coef, OutPtrs<T, Op::NE> res,
          int n0, int n1, typename Op::Params P) {
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= n0 || j >= n1) return;
  GlobalAcc<T, Op::NF> acc{in.f, n0, n1, i, j};
  T f[Op::NE];
  T J[Op::NE][Op::NF][Op::NS];
  Op::eval(acc, i, j, n0, n1, P, f, J);
  const size_t N = (size_t)n0 * n1;
  const size_t c = (size_t)i * n1 + j;
#pragma unroll
  for (int q = 0; q < Op::NJ; ++q)
    coef[q * N + c] = J[Op::JE[q]][Op::JV[q]][Op::JS[q]];
  for (int e = 0; e < Op::NE; ++e)
    if (res.f[e]) res.f[e][c] = f[e];
}

// Fused loss + adjoint. One workgroup owns a TX x TY tile of output cells.
//  1. load state for the tile grown by RJ + R into LDS;
//  2. evaluate f and J on the tile grown by RJ (RJ = Jacobian reach); every
//     off-centre product J[e][k][s] f_e goes to LDS, the centre product of
//     inner cells stays in registers;
//  3. gather grad_k(c) = sum_s out_{k,s}(c - shift_s).
template <class T, class Op, int TX, int TY>
__global__ void __launch_bounds__(256)
    k_loss_grad(FieldPtrs<T, Op::NF> in, OutPtrs<T, Op::NF> grad,
                double* __restrict__ partial, int n0, int n1, T scale,
                typename Op::Params P) {
  constexpr int RJ = Op::RJ;
  constexpr int H = RJ + Op::R;
  constexpr int S0 = TX + 2 * H, S1 = TY + 2 * H;
  constexpr int E0 = TX + 2 * RJ, E1 = TY + 2 * RJ;
  constexpr int NT = 256;
  constexpr int NPT = (TX * TY) / NT;  // inner cells per thread
  static_assert((TX * TY) % NT == 0, "tile must be a multiple of 256");
  extern __shared__ unsigned char smem_raw[];
  T* st = reinterpret_cast<T*>(smem_raw);           // [NF][S0][S1]
  T* ot = st + Op::NF * S0 * S1;                    // [NF][NS-1][E0][E1]
  __shared__ double red[4];

  const int x0 = blockIdx.y * TX;
  const int y0 = blockIdx.x * TY;
  const int tid = threadIdx.x;

  // 1. state tile (values outside the domain are never selected; zero-fill).
  for (int q = tid; q < Op::NF * S0 * S1; q += NT) {
    const int k = q / (S0 * S1);
    const int r = q - k * S0 * S1;
    const int a = r / S1, b = r - a * S1;
    const int gi = x0 - H + a, gj = y0 - H + b;
    T v = 0;
    if (gi >= 0 && gi < n0 && gj >= 0 && gj < n1)
      v = in.f[k][(size_t)gi * n1 + gj];
    st[q] = v;
  }
  __syncthreads();

  // 2a. inner cells: keep centre products and squares in registers.
  T cen[NPT][Op::NF];
  double ssq[Op::NE];
  for (int e = 0; e < Op::NE; ++e) ssq[e] = 0;
  for (int p = 0; p < NPT; ++p) {
    const int q = tid + p * NT;
    const int a = q / TY, b = q - a * TY;
    const int gi = x0 + a, gj = y0 + b;
    for (int k = 0; k < Op::NF; ++k) cen[p][k] = 0;
    const int ea = a + RJ, eb = b + RJ;
    if (gi < n0 && gj < n1) {
      LdsAcc<T> acc{st, S0, S1, a + H, b + H};
      T f[Op::NE];
      T J[Op::NE][Op::NF][Op::NS];
      Op::eval(acc, gi, gj, n0, n1, P, f, J);
      for (int e = 0; e < Op::NE; ++e) ssq[e] += (double)f[e] * (double)f[e];
#pragma unroll
      for (int k = 0; k < Op::NF; ++k)
#pragma unroll
        for (int s = 0; s < Op::NS; ++s) {
          T o = 0;
#pragma unroll
          for (int e = 0; e < Op::NE; ++e) o += J[e][k][s] * f[e];
          if (s == 0)
            cen[p][k] = o;
          else
            ot[((k * (Op::NS - 1) + s - 1) * E0 + ea) * E1 + eb] = o;
        }
    } else {
      for (int k = 0; k < Op::NF; ++k)
        for (int s = 1; s < Op::NS; ++s)
          ot[((k * (Op::NS - 1) + s - 1) * E0 + ea) * E1 + eb] = 0;
    }
  }
  // 2b. ring cells around the tile (only off-centre products are needed).
  constexpr int NRING = E0 * E1 - TX * TY;
  for (int q = tid; q < NRING; q += NT) {
    int ea, eb;
    if (q < RJ * E1) {  // top rows
      ea = q / E1;
      eb = q - ea * E1;
    } else if (q < 2 * RJ * E1) {  // bottom rows
      const int r = q - RJ * E1;
      ea = TX + RJ + r / E1;
      eb = r % E1;
    } else {  // side columns
      const int r = q - 2 * RJ * E1;
      const int row = r / (2 * RJ);
      const int col = r - row * (2 * RJ);
      ea = RJ + row;
      eb = col < RJ ? col : TY + col;
    }
    const int gi = x0 - RJ + ea, gj = y0 - RJ + eb;
    if (gi >= 0 && gi < n0 && gj >= 0 && gj < n1) {
      LdsAcc<T> acc{st, S0, S1, ea + Op::R, eb + Op::R};
      T f[Op::NE];
      T J[Op::NE][Op::NF][Op::NS];
      Op::eval(acc, gi, gj, n0, n1, P, f, J);
#pragma unroll
      for (int k = 0; k < Op::NF; ++k)
#pragma unroll
        for (int s = 1; s < Op::NS; ++s) {
          T o = 0;
#pragma unroll
          for (int e = 0; e < Op::NE; ++e) o += J[e][k][s] * f[e];
          ot[((k * (Op::NS - 1) + s - 1) * E0 + ea) * E1 + eb] = o;
        }
    } else {
      for (int k = 0; k < Op::NF; ++k)
        for (int s = 1; s < Op::NS; ++s)
          ot[((k * (Op::NS - 1) + s - 1) * E0 + ea) * E1 + eb] = 0;
    }
  }
  __syncthreads();

  // 3. gather.
  for (int p = 0; p < NPT; ++p) {
    const int q = tid + p * NT;
    const int a = q / TY, b = q - a * TY;
    const int gi = x0 + a, gj = y0 + b;
    if (gi >= n0 || gj >= n1) continue;
    const size_t c = (size_t)gi * n1 + gj;
#pragma unroll
    for (int k = 0; k < Op::NF; ++k) {
      T g = cen[p][k];
#pragma unroll
      for (int s = 1; s < Op::NS; ++s) {
        // Cell n = c - shift_s contributed J[.][k][s] f(n) to c.
        const int ea = a + RJ - Op::SX[s], eb = b + RJ - Op::SY[s];
        g += ot[((k * (Op::NS - 1) + s - 1) * E0 + ea) * E1 + eb];
      }
      if (grad.f[k]) grad.f[k][c] = g * scale;
    }
  }
  if (partial) {
    const int bidx = blockIdx.y * gridDim.x + blockIdx.x;
    const int nb = gridDim.x * gridDim.y;
    for (int e = 0; e < Op::NE; ++e) {
      double s = block_sum(ssq[e], red);
      if (tid == 0) partial[(size_t)e * nb + bidx] = s;
    }
  }
}

template <class T, class Op, int TX, int TY>
constexpr size_t loss_grad_smem() {
  constexpr int H = Op::RJ + Op::R;
  return sizeof(T) * (size_t)(Op::NF * (TX + 2 * H) * (TY + 2 * H) +
                              Op::NF * (Op::NS - 1) * (TX + 2 * Op::RJ) *
                                  (TY + 2 * Op::RJ));
}

// Fused matrix-free normal operator N = J^T J with the Jacobian recomputed
// from the state in LDS (no stored coefficients):
//   mode 0: y = N x          mode 1: y = b - N x
//   mode 2: y = x + omega (b - N x) / d   (weighted Jacobi, d = diag N)
//   mode 3: y = diag(N) = sum_e sum_s J[e][k][s](c - shift_s)^2
// Tile layout as in k_loss_grad; x is staged with halo 2 RJ because
// (J x)(n) reads x at n + shift for every n in the tile grown by RJ.
template <class T, class Op, int TX, int TY>
__global__ void __launch_bounds__(256)
    k_normal_apply(FieldPtrs<T, Op::NF> st_in, FieldPtrs<T, Op::NF> x_in,
                   FieldPtrs<T, Op::NF> b_in, FieldPtrs<T, Op::NF> d_in,
                   OutPtrs<T, Op::NF> y, int n0, int n1, T omega, int mode,
                   typename Op::Params P) {
  constexpr int NF = Op::NF, NE = Op::NE, NS = Op::NS;
  constexpr int RJ = Op::RJ;
  constexpr int H = RJ + Op::R;
  constexpr int HX = 2 * RJ;
  constexpr int S0 = TX + 2 * H, S1 = TY + 2 * H;
  constexpr int X0 = TX + 2 * HX, X1 = TY + 2 * HX;
  constexpr int E0 = TX + 2 * RJ, E1 = TY + 2 * RJ;
  constexpr int NT = 256;
  constexpr int NPT = (TX * TY) / NT;
  static_assert((TX * TY) % NT == 0, "tile must be a multiple of 256");
  extern __shared__ unsigned char smem_raw[];
  T* st = reinterpret_cast<T*>(smem_raw);  // [NF][S0][S1]
  T* xs = st + NF * S0 * S1;               // [NF][X0][X1]
  T* ot = xs + NF * X0 * X1;               // [NF][NS-1][E0][E1]

  const int x0 = blockIdx.y * TX;
  const int y0 = blockIdx.x * TY;
  const int tid = threadIdx.x;
  const bool diag = (mode == 3);

  for (int q = tid; q < NF * S0 * S1; q += NT) {
    const int k = q / (S0 * S1);
    const int r = q - k * S0 * S1;
    const int a = r / S1, b = r - a * S1;
    const int gi = x0 - H + a, gj = y0 - H + b;
    T v = 0;
    if (gi >= 0 && gi < n0 && gj >= 0 && gj < n1)
      v = st_in.f[k][(size_t)gi * n1 + gj];
    st[q] = v;
  }
  if (!diag)
    for (int q = tid; q < NF * X0 * X1; q += NT) {
      const int k = q / (X0 * X1);
      const int r = q - k * X0 * X1;
      const int a = r / X1, b = r - a * X1;
      const int gi = x0 - HX + a, gj = y0 - HX + b;
      T v = 0;
      if (gi >= 0 && gi < n0 && gj >= 0 && gj < n1)
        v = x_in.f[k][(size_t)gi * n1 + gj];
      xs[q] = v;
    }
  __syncthreads();

  // Products of one E-region cell (E-local ea, eb): o[k][s] =
  // sum_e J[e][k][s] v_e with v = J x (or J[e][k][s] itself for diag).
  auto cell = [&](int ea, int eb, int gi, int gj, T (*o)[NS]) {
    LdsAcc<T> acc{st, S0, S1, ea + Op::R, eb + Op::R};
    T f[NE];
    T J[NE][NF][NS];
    Op::eval(acc, gi, gj, n0, n1, P, f, J);
    if (diag) {
#pragma unroll
      for (int k = 0; k < NF; ++k)
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          T a = 0;
#pragma unroll
          for (int e = 0; e < NE; ++e) a += J[e][k][s] * J[e][k][s];
          o[k][s] = a;
        }
      return;
    }
    T v[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      T a = 0;
#pragma unroll
      for (int k = 0; k < NF; ++k)
#pragma unroll
        for (int s = 0; s < NS; ++s)
          a += J[e][k][s] *
               xs[(k * X0 + ea + RJ + Op::SX[s]) * X1 + eb + RJ + Op::SY[s]];
      v[e] = a;
    }
#pragma unroll
    for (int k = 0; k < NF; ++k)
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        T a = 0;
#pragma unroll
        for (int e = 0; e < NE; ++e) a += J[e][k][s] * v[e];
        o[k][s] = a;
      }
  };

  T cen[NPT][NF];
  for (int p = 0; p < NPT; ++p) {
    const int q = tid + p * NT;
    const int a = q / TY, b = q - a * TY;
    const int gi = x0 + a, gj = y0 + b;
    const int ea = a + RJ, eb = b + RJ;
    T o[NF][NS];
    if (gi < n0 && gj < n1) {
      cell(ea, eb, gi, gj, o);
    } else {
      for (int k = 0; k < NF; ++k)
        for (int s = 0; s < NS; ++s) o[k][s] = 0;
    }
    for (int k = 0; k < NF; ++k) {
      cen[p][k] = o[k][0];
      for (int s = 1; s < NS; ++s)
        ot[((k * (NS - 1) + s - 1) * E0 + ea) * E1 + eb] = o[k][s];
    }
  }
  constexpr int NRING = E0 * E1 - TX * TY;
  for (int q = tid; q < NRING; q += NT) {
    int ea, eb;
    if (q < RJ * E1) {
      ea = q / E1;
      eb = q - ea * E1;
    } else if (q < 2 * RJ * E1) {
      const int r = q - RJ * E1;
      ea = TX + RJ + r / E1;
      eb = r % E1;
    } else {
      const int r = q - 2 * RJ * E1;
      const int row = r / (2 * RJ);
      const int col = r - row * (2 * RJ);
      ea = RJ + row;
      eb = col < RJ ? col : TY + col;
    }
    const int gi = x0 - RJ + ea, gj = y0 - RJ + eb;
    T o[NF][NS];
    if (gi >= 0 && gi < n0 && gj >= 0 && gj < n1) {
      cell(ea, eb, gi, gj, o);
    } else {
      for (int k = 0; k < NF; ++k)
        for (int s = 0; s < NS; ++s) o[k][s] = 0;
    }
    for (int k = 0; k < NF; ++k)
      for (int s = 1; s < NS; ++s)
        ot[((k * (NS - 1) + s - 1) * E0 + ea) * E1 + eb] = o[k][s];
  }
  __syncthreads();

  for (int p = 0; p < NPT; ++p) {
    const int q = tid + p * NT;
    const int a = q / TY, b = q - a * TY;
    const int gi = x0 + a, gj = y0 + b;
    if (gi >= n0 || gj >= n1) continue;
    const size_t c = (size_t)gi * n1 + gj;
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      T g = cen[p][k];
#pragma unroll
      for (int s = 1; s < NS; ++s) {
        const int ea = a + RJ - Op::SX[s], eb = b + RJ - Op::SY[s];
        g += ot[((k * (NS - 1) + s - 1) * E0 + ea) * E1 + eb];
      }
      T out;
      if (mode == 0 || mode == 3)
        out = g;
      else if (mode == 1)
        out = b_in.f[k][c] - g;
      else
        out = xs[(k * X0 + a + HX) * X1 + b + HX] +
              omega * (b_in.f[k][c] - g) / d_in.f[k][c];
      y.f[k][c] = out;
    }
  }
}

template <class T, class Op, int TX, int TY>
constexpr size_t normal_apply_smem() {
  constexpr int H = Op::RJ + Op::R, HX = 2 * Op::RJ;
  return sizeof(T) *
         (size_t)(Op::NF * (TX + 2 * H) * (TY + 2 * H) +
                  Op::NF * (TX + 2 * HX) * (TY + 2 * HX) +
                  Op::NF * (Op::NS - 1) * (TX + 2 * Op::RJ) *
                      (TY + 2 * Op::RJ));
}

// Deterministic final reduction of per-block partials: out[e] = sum_b
// partial[e*nb + b]. One block per equation.
__global__ void k_reduce_partials(const double* __restrict__ partial, int nb,
                                  double* __restrict__ out);

// Host-side launch helpers shared by every 2D operator.
template <class T, class Op>
int launch_residual(const void* const* fields, void* const* res,
                    double* partial, int n0, int n1,
                    const typename Op::Params& P, hipStream_t stream,
                    int* nblocks) {
  FieldPtrs<T, Op::NF> in;
  OutPtrs<T, Op::NE> out;
  for (int k = 0; k < Op::NF; ++k) in.f[k] = static_cast<const T*>(fields[k]);
  for (int e = 0; e < Op::NE; ++e)
    out.f[e] = res ? static_cast<T*>(res[e]) : nullptr;
  dim3 grid(ceil_div(n1, 64), ceil_div(n0, 4));
  if (nblocks) *nblocks = grid.x * grid.y;
  hipLaunchKernelGGL((k_residual<T, Op>), grid, dim3(256), 0, stream, in, out,
                     partial, n0, n1, P);
  return check_launch("residual");
}

template <class T, class Op>
int launch_jac(const void* const* fields, void* coef, void* const* res,
               int n0, int n1, const typename Op::Params& P,
               hipStream_t stream) {
  FieldPtrs<T, Op::NF> in;
  OutPtrs<T, Op::NE> out;
  for (int k = 0; k < Op::NF; ++k) in.f[k] = static_cast<const T*>(fields[k]);
  for (int e = 0; e < Op::NE; ++e)
    out.f[e] = res ? static_cast<T*>(res[e]) : nullptr;
  dim3 grid(ceil_div(n1, 64), ceil_div(n0, 4));
  hipLaunchKernelGGL((k_jac<T, Op>), grid, dim3(256), 0, stream, in,
                     static_cast<T*>(coef), out, n0, n1, P);
  return check_launch("jac_coeffs");
}

template <class T, class Op, int TX, int TY>
int launch_loss_grad_t(const void* const* fields, void* const* grad,
                       double* partial, int n0, int n1, T scale,
                       const typename Op::Params& P, hipStream_t stream,
                       int* nblocks) {
  FieldPtrs<T, Op::NF> in;
  OutPtrs<T, Op::NF> out;
  for (int k = 0; k < Op::NF; ++k) {
    in.f[k] = static_cast<const T*>(fields[k]);
    out.f[k] = static_cast<T*>(grad[k]);
  }
  dim3 grid(ceil_div(n1, TY), ceil_div(n0, TX));
  if (nblocks) *nblocks = grid.x * grid.y;
  constexpr size_t smem = loss_grad_smem<T, Op, TX, TY>();
  static_assert(smem <= 160 * 1024, "LDS tile too large");
  auto kern = k_loss_grad<T, Op, TX, TY>;
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  hipLaunchKernelGGL(kern, grid, dim3(256), smem, stream, in, out, partial,
                     n0, n1, scale, P);
  return check_launch("loss_grad");
}

template <class T, class Op, int TX, int TY>
int launch_normal_apply_t(const void* const* state, const void* const* x,
                          const void* const* b, const void* const* d,
                          void* const* y, int n0, int n1, double omega,
                          int mode, const typename Op::Params& P,
                          hipStream_t stream) {
  if (mode < 0 || mode > 3) {
    set_error("normal_apply: mode must be 0..3");
    return kErrArg;
  }
  if ((mode != 3 && !x) || ((mode == 1 || mode == 2) && !b) ||
      (mode == 2 && !d) || !state || !y) {
    set_error("normal_apply: missing operand for mode %d", mode);
    return kErrArg;
  }
  FieldPtrs<T, Op::NF> S{}, X{}, B{}, D{};
  OutPtrs<T, Op::NF> Y{};
  for (int k = 0; k < Op::NF; ++k) {
    S.f[k] = static_cast<const T*>(state[k]);
    X.f[k] = x ? static_cast<const T*>(x[k]) : nullptr;
    B.f[k] = b ? static_cast<const T*>(b[k]) : nullptr;
    D.f[k] = d ? static_cast<const T*>(d[k]) : nullptr;
    Y.f[k] = static_cast<T*>(y[k]);
  }
  dim3 grid(ceil_div(n1, TY), ceil_div(n0, TX));
  constexpr size_t smem = normal_apply_smem<T, Op, TX, TY>();
  static_assert(smem <= 160 * 1024, "LDS tile too large");
  auto kern = k_normal_apply<T, Op, TX, TY>;
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)smem);
  hipLaunchKernelGGL(kern, grid, dim3(256), smem, stream, S, X, B, D, Y, n0,
                     n1, (T)omega, mode, P);
  return check_launch("normal_apply");
}

template <class Op64, class Op32>
int dispatch_normal_apply(int dtype, const void* const* state,
                          const void* const* x, const void* const* b,
                          const void* const* d, void* const* y, int n0,
                          int n1, double omega, int mode,
                          const typename Op64::Params& P, hipStream_t s) {
  if (dtype == kF64)
    return launch_normal_apply_t<double, Op64, 8, 32>(state, x, b, d, y, n0,
                                                      n1, omega, mode, P, s);
  if (dtype == kF32)
    return launch_normal_apply_t<float, Op32, 16, 32>(state, x, b, d, y, n0,
                                                      n1, omega, mode, P, s);
  set_error("normal_apply: bad dtype %d", dtype);
  return kErrArg;
}

// scale * J^T f with the residual's per-block sums of squares (the loss
// adjoint kernel with a caller-chosen scale, e.g. -1 for a Newton
// right-hand side).
template <class Op64, class Op32>
int dispatch_jtf(int dtype, const void* const* fields, void* const* out,
                 double* partial, int n0, int n1, double scale,
                 const typename Op64::Params& P, hipStream_t s,
                 int* nblocks) {
  if (dtype == kF64)
    return launch_loss_grad_t<double, Op64, 8, 64>(fields, out, partial, n0,
                                                   n1, scale, P, s, nblocks);
  if (dtype == kF32)
    return launch_loss_grad_t<float, Op32, 16, 64>(
        fields, out, partial, n0, n1, (float)scale, P, s, nblocks);
  set_error("jtf: bad dtype %d", dtype);
  return kErrArg;
}

}  // namespace odil
