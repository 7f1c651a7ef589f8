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

}  // namespace odil
