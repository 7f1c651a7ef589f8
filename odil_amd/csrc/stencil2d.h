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

#include <algorithm>
#include <type_traits>

#include "common.h"

namespace odil {

template <class T, int NF, bool CLAMP = true>
struct GlobalAcc {
  const T* const* f;
  int n0, n1, i, j;
  __device__ inline T operator()(int k, int di, int dj) const {
    int ii = i + di, jj = j + dj;
    if (CLAMP) {
      ii = min(max(ii, 0), n0 - 1);
      jj = min(max(jj, 0), n1 - 1);
    }
    return f[k][(size_t)ii * n1 + jj];
  }
};

// Unclamped accessor of cells whose stencil stays inside the grid: one
// base pointer per field at the cell, neighbours at 32-bit offsets
// di * n1 + dj (same-row neighbours fold into the load's immediate).
template <class T, int NF>
struct GlobalAcc<T, NF, false> {
  const T* b[NF];
  int n1;
  __device__ inline GlobalAcc(const T* const* f, int n0_, int n1_, int i,
                              int j)
      : n1(n1_) {
    const size_t c = (size_t)i * n1_ + j;
#pragma unroll
    for (int k = 0; k < NF; ++k) b[k] = f[k] + c;
  }
  __device__ inline T operator()(int k, int di, int dj) const {
    return b[k][di * n1 + dj];
  }
};

// True when every cell of rows [a0, a1) x cols [b0, b1) is at least R cells
// from the domain edge, so operators may take their interior fast path
// (no wall rules, no clamping).
__device__ inline bool interior_box(int a0, int a1, int b0, int b1, int n0,
                                    int n1, int R) {
  return a0 >= R && a1 <= n0 - R && b0 >= R && b1 <= n1 - R;
}

// Operators with `static constexpr bool ACCUM = true` add their adjoint to
// the gradient arrays (extra loss terms of another operator's unknowns);
// all others overwrite them.
template <class Op, class = void>
struct op_accumulates {
  static constexpr bool value = false;
};
template <class Op>
struct op_accumulates<Op, std::void_t<decltype(Op::ACCUM)>> {
  static constexpr bool value = Op::ACCUM;
};

template <class T, int NF>
struct FieldPtrs {
  const T* f[NF];
};
template <class T, int NF>
struct OutPtrs {
  T* f[NF];
};

// Accessor of a field tile staged in LDS: `t` points at cell (0, 0) of
// field 0 in a [NF][E0][E1] tile, so neighbour (di, dj) of field k is one
// ds_read at a compile-time offset from the cell's address.
template <class T, int E0, int E1>
struct LdsAcc {
  const T* t;
  __device__ inline T operator()(int k, int di, int dj) const {
    return t[k * (E0 * E1) + di * E1 + dj];
  }
};

// Stages fields[k] over rows [x0 - H, x0 + TX + H) x cols [y0 - H, y0 + TY +
// H) into `tile` ([NF][E0][E1]), indices clamped into the grid exactly as
// the clamping global accessor reads them (wall rules override the clamped
// values where they apply). Consecutive threads read consecutive columns.
// All of a thread's loads are issued before any LDS write, so they are in
// flight together (a load-then-store loop would wait for each in turn);
// blocks whose grown tile lies inside the grid skip the clamps.
template <class T, int NF, int TX, int TY, int H>
__device__ inline void stage_tile(const FieldPtrs<T, NF>& in, T* tile,
                                  int x0, int y0, int n0, int n1) {
  constexpr int E0 = TX + 2 * H, E1 = TY + 2 * H;
  constexpr int PER = E0 * E1;
  constexpr int NL = (PER + 255) / 256;
  const bool inside = x0 - H >= 0 && x0 + TX + H <= n0 && y0 - H >= 0 &&
                      y0 + TY + H <= n1;
  T v[NF][NL];
  if (inside) {
    const T* base[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k)
      base[k] = in.f[k] + (size_t)(x0 - H) * n1 + (y0 - H);
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      const int idx = threadIdx.x + 256 * r;
      if (r + 1 < NL || idx < PER) {
        const int a = idx / E1, b = idx - a * E1;
        const int off = a * n1 + b;
#pragma unroll
        for (int k = 0; k < NF; ++k) v[k][r] = base[k][off];
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      const int idx = threadIdx.x + 256 * r;
      if (r + 1 < NL || idx < PER) {
        const int a = idx / E1, b = idx - a * E1;
        const int gi = min(max(x0 - H + a, 0), n0 - 1);
        const int gj = min(max(y0 - H + b, 0), n1 - 1);
        const size_t off = (size_t)gi * n1 + gj;
#pragma unroll
        for (int k = 0; k < NF; ++k) v[k][r] = in.f[k][off];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < NL; ++r) {
    const int idx = threadIdx.x + 256 * r;
    if (r + 1 < NL || idx < PER) {
#pragma unroll
      for (int k = 0; k < NF; ++k) tile[k * PER + idx] = v[k][r];
    }
  }
}

// Tiles of the residual (TX rows x TY columns per 256-thread workgroup) and
// of the fused loss-adjoint kernels (cells along axis 0 per precision x 32
// along axis 1), and the partial-sum counts every operator's *_nblocks
// reports: one partial per wavefront (4 per 256-thread workgroup).
// (On the 1025^2 cavity 16 x 32 (FP32) and 8 x 32 (FP64) were the fastest
// loss-adjoint tiles of 4..32 x 32..128; residual tiles of 4, 8, 16 and 32
// rows took 16.7, 15.7, 16.9 and 21.3 us in FP32 and 22.0, 21.2, 23.2 and
// 31.6 us in FP64, inputs and outputs HBM-resident.)
constexpr int kResTx = 8, kResTy = 64;
constexpr int kGradTx32 = 16, kGradTy32 = 32, kGradTx64 = 8, kGradTy64 = 32;
// LDS staging of the matrix-free normal operator (k_normal_apply STG),
// measured on the 1025^2 cavity with HBM-resident inputs: FP32 state and x
// staged 28.8 us (nothing staged 30.1); FP64 nothing staged 47.7 us (state
// and x staged 61.4: three FP64 tiles leave two workgroups per CU).
constexpr int kNormStage32 = 2, kNormStage64 = 0;
// Its tiles (FP32 with that staging: 16 x 16 20.7 us, 8 x 32 22.6, 32 x 16
// 25.7, 16 x 32 25.3, 4 x 64 26.5, 8 x 64 26.7; FP64 unstaged: 8 x 32 34.9,
// 16 x 16 35.7, 4 x 64 37.9, 16 x 32 46.7; mode 2 at 1025^2, cache-resident
// inputs).
constexpr int kNormTx32 = 16, kNormTy32 = 16, kNormTx64 = 8, kNormTy64 = 32;

constexpr int kWavesPerBlock = 4;

// Thin-frame launch geometry: TX x TY tiles over the interior cells (at
// least H from every face), the last tile of a row or column partial, and
// frame blocks over the cells within H of a face, `per` frame cells per
// block. With no interior every cell is a frame cell. Used by the residual
// (H = R, every tile on the operator's interior path) and by the operators
// with interior tiles (H = R + RJ).
struct ThinSplit {
  int ntx, nty, nint, nframe, nfb;
  bool has_inner;
  // Frame cell f -> (i, j): the first and last H rows, then the first and
  // last H columns of the middle rows.
  __device__ inline void frame_cell(int f, int n0, int n1, int H, int& i,
                                    int& j) const {
    if (!has_inner) {
      i = f / n1;
      j = f - i * n1;
      return;
    }
    const int band = H * n1;
    if (f < 2 * band) {
      const int r = f / n1;
      i = r < H ? r : n0 - 2 * H + r;
      j = f - r * n1;
      return;
    }
    f -= 2 * band;
    const int r = f / (2 * H), c = f - r * (2 * H);
    i = H + r;
    j = c < H ? c : n1 - 2 * H + c;
  }
};

inline ThinSplit thin_split(int n0, int n1, int H, int TX, int TY, int per) {
  ThinSplit s;
  s.has_inner = n0 > 2 * H && n1 > 2 * H;
  s.ntx = s.has_inner ? ceil_div(n1 - 2 * H, TY) : 0;
  s.nty = s.has_inner ? ceil_div(n0 - 2 * H, TX) : 0;
  s.nint = s.ntx * s.nty;
  s.nframe = s.has_inner ? n0 * n1 - (n0 - 2 * H) * (n1 - 2 * H) : n0 * n1;
  // Frame blocks come first in dispatch order, padded to a multiple of the
  // XCD count: they are dealt round-robin over the XCDs, and the interior
  // tiles after them keep xcd_block's alignment.
  s.nfb = ceil_div(ceil_div(s.nframe, per), (int)kNumXcd) * (int)kNumXcd;
  return s;
}

inline ThinSplit residual_split(int n0, int n1, int R) {
  return thin_split(n0, n1, R, kResTx, kResTy, 256);
}

inline int residual_nblocks(int n0, int n1, int R) {
  const ThinSplit s = residual_split(n0, n1, R);
  return (s.nint + s.nfb) * kWavesPerBlock;
}
inline int grad_nblocks(bool fp32, int n0, int n1) {
  return ceil_div(n1, fp32 ? kGradTy32 : kGradTy64) *
         ceil_div(n0, fp32 ? kGradTx32 : kGradTx64) *
         kWavesPerBlock;
}

template <class T, class Op>
constexpr size_t residual_smem() {
  return sizeof(T) * (size_t)Op::NF * (kResTx + 2 * Op::R) *
         (kResTy + 2 * Op::R);
}

// Residual f_e at every cell, and per-wavefront partial sums of f_e^2.
// Frame blocks (the first nfb in dispatch order, spread over the XCDs)
// evaluate one frame cell per
// thread from global memory on the operator's generic path (wall rules).
// Interior blocks stage their TX x TY tile of the fields,
// grown by the operator reach R, in LDS with one coalesced read per value
// (clamped where a partial tile's halo leaves the grid) and evaluate the
// interior path of the operator from LDS, one wavefront per 64-cell row.
// Partial slot of block b, wave w:
// b * 4 + w (fixed order, bitwise reproducible sums).
template <class T, class Op>
__global__ void __launch_bounds__(256)
    k_residual(FieldPtrs<T, Op::NF> in, OutPtrs<T, Op::NE> out,
               double* __restrict__ partial, int n0, int n1,
               typename Op::Params P, ThinSplit sp) {
  constexpr int R = Op::R;
  constexpr int E0 = kResTx + 2 * R, E1 = kResTy + 2 * R;
  constexpr int NPT = kResTx * kResTy / 256;
  static_assert(kResTy == 64, "one wavefront per tile row");
  extern __shared__ unsigned char smem_raw[];
  T* tile = reinterpret_cast<T*>(smem_raw);
  const unsigned b = blockIdx.x;
  double ssq[Op::NE];
  for (int e = 0; e < Op::NE; ++e) ssq[e] = 0;
  if ((int)b >= sp.nfb) {
    const unsigned t = xcd_block(b - (unsigned)sp.nfb, (unsigned)sp.nint);
    const int x0 = R + (int)(t / (unsigned)sp.ntx) * kResTx;
    const int y0 = R + (int)(t % (unsigned)sp.ntx) * kResTy;
    stage_tile<T, Op::NF, kResTx, kResTy, R>(in, tile, x0, y0, n0, n1);
    __syncthreads();
    const int bb = threadIdx.x & 63;
    const int j = y0 + bb;
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
      const int a = (threadIdx.x >> 6) + 4 * r;
      const int i = x0 + a;
      if (i >= n0 - R || j >= n1 - R) continue;
      LdsAcc<T, E0, E1> acc{tile + (a + R) * E1 + (bb + R)};
      T f[Op::NE];
      Op::template residual<true>(acc, i, j, n0, n1, P, f);
      const size_t c = (size_t)i * n1 + j;
      for (int e = 0; e < Op::NE; ++e) {
        if (out.f[e]) out.f[e][c] = f[e];
        ssq[e] += (double)f[e] * (double)f[e];
      }
    }
  } else {
    const int fc = (int)b * 256 + (int)threadIdx.x;
    if (fc < sp.nframe) {
      int i, j;
      sp.frame_cell(fc, n0, n1, R, i, j);
      GlobalAcc<T, Op::NF> acc{in.f, n0, n1, i, j};
      T f[Op::NE];
      Op::template residual<false>(acc, i, j, n0, n1, P, f);
      const size_t c = (size_t)i * n1 + j;
      for (int e = 0; e < Op::NE; ++e) {
        if (out.f[e]) out.f[e][c] = f[e];
        ssq[e] += (double)f[e] * (double)f[e];
      }
    }
  }
  if (partial)
    wave_partials<Op::NE>(ssq, partial, (size_t)gridDim.x * kWavesPerBlock,
                          (size_t)b * kWavesPerBlock + (threadIdx.x >> 6));
}

// Jacobian coefficient fields: coef[q] for q < NJ holds entry
// (JE[q], JV[q], JS[q]); optionally also the residual.
template <class T, class Op>
__global__ void __launch_bounds__(256)
    k_jac(FieldPtrs<T, Op::NF> in, T* __restrict__ coef, OutPtrs<T, Op::NE> res,
          int n0, int n1, typename Op::Params P) {
  unsigned bx, by;
  xcd_block2(bx, by);
  const int j = bx * 64 + (threadIdx.x & 63);
  const int i = by * 4 + (threadIdx.x >> 6);
  if (i >= n0 || j >= n1) return;
  T f[Op::NE];
  T J[Op::NE][Op::NF][Op::NS];
  if (interior_box(by * 4, by * 4 + 4, bx * 64, bx * 64 + 64, n0, n1,
                   Op::R)) {
    GlobalAcc<T, Op::NF, false> acc{in.f, n0, n1, i, j};
    Op::template eval<true>(acc, i, j, n0, n1, P, f, J);
  } else {
    GlobalAcc<T, Op::NF> acc{in.f, n0, n1, i, j};
    Op::template eval<false>(acc, i, j, n0, n1, P, f, J);
  }
  const size_t N = (size_t)n0 * n1;
  const size_t c = (size_t)i * n1 + j;
#pragma unroll
  for (int q = 0; q < Op::NJ; ++q)
    coef[q * N + c] = J[Op::JE[q]][Op::JV[q]][Op::JS[q]];
  for (int e = 0; e < Op::NE; ++e)
    if (res.f[e]) res.f[e][c] = f[e];
}

// Ring cell q of the tile grown by RJ (the cells around the TX x TY inner
// tile): returns its E-local coordinates.
template <int TX, int TY, int RJ>
__device__ inline void ring_cell(int q, int& ea, int& eb) {
  constexpr int E1 = TY + 2 * RJ;
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
}

// Residual and Jacobian of cell (gi, gj) reading the state from global
// memory (neighbour reuse through L1/L2); INT blocks skip clamping and wall
// rules.
template <class T, class Op>
__device__ inline void eval_global(const FieldPtrs<T, Op::NF>& in, int gi,
                                   int gj, int n0, int n1, bool inner,
                                   const typename Op::Params& P, T* f,
                                   T (*J)[Op::NF][Op::NS]) {
  if (inner) {
    GlobalAcc<T, Op::NF, false> acc{in.f, n0, n1, gi, gj};
    Op::template eval<true>(acc, gi, gj, n0, n1, P, f, J);
  } else {
    GlobalAcc<T, Op::NF> acc{in.f, n0, n1, gi, gj};
    Op::template eval<false>(acc, gi, gj, n0, n1, P, f, J);
  }
}

// Residual and Jacobian of one cell from a state tile staged in LDS
// ([NF][S0][S1], see stage_tile); INT blocks skip wall rules:
// `cell` points at the cell's own position in field 0 of the tile.
template <class T, class Op, int S0, int S1>
__device__ inline void eval_lds(const T* cell, int gi, int gj, int n0, int n1,
                                bool inner, const typename Op::Params& P,
                                T* f, T (*J)[Op::NF][Op::NS]) {
  LdsAcc<T, S0, S1> acc{cell};
  if (inner)
    Op::template eval<true>(acc, gi, gj, n0, n1, P, f, J);
  else
    Op::template eval<false>(acc, gi, gj, n0, n1, P, f, J);
}

// Fused loss + adjoint of one TX x TY tile of output cells at (x0, y0),
// computed by the whole 256-thread workgroup:
//  0. stage the state over the tile grown by RJ + R in LDS (after the
//     products buffer `ot`);
//  1. evaluate f and J on the tile grown by RJ (RJ = Jacobian reach) from
//     that tile; every off-centre product J[e][k][s] f_e goes to LDS `ot`,
//     the centre product of inner cells stays in registers;
//  2. gather grad_k(c) = scale * sum_s out_{k,s}(c - shift_s).
// Adds this thread's f_e^2 to ssq[e]. Starts with a barrier so a workgroup
// may run it on several tiles in turn.
template <class T, class Op, int TX, int TY>
__device__ inline void loss_grad_tile(const FieldPtrs<T, Op::NF>& in,
                                      const OutPtrs<T, Op::NF>& grad, int n0,
                                      int n1, int x0, int y0, T scale,
                                      const typename Op::Params& P, T* ot,
                                      double* ssq) {
  constexpr int NF = Op::NF, NE = Op::NE, NS = Op::NS, RJ = Op::RJ;
  constexpr int E0 = TX + 2 * RJ, E1 = TY + 2 * RJ;
  constexpr int H = RJ + Op::R, S0 = TX + 2 * H, S1 = TY + 2 * H;
  constexpr int NT = 256;
  constexpr int NPT = (TX * TY) / NT;  // inner cells per thread
  static_assert((TX * TY) % NT == 0, "tile must be a multiple of 256");
  const int tid = threadIdx.x;
  const bool inner = interior_box(x0 - RJ, x0 + TX + RJ, y0 - RJ,
                                  y0 + TY + RJ, n0, n1, Op::R);
  auto OT = [&](int k, int s, int ea, int eb) -> T& {
    return ot[((k * (NS - 1) + s - 1) * E0 + ea) * E1 + eb];
  };
  T* st = ot + NF * (NS - 1) * E0 * E1;  // state tile [NF][S0][S1]
  __syncthreads();  // previous tile's gathers are done with `ot`, `st`
  stage_tile<T, NF, TX, TY, H>(in, st, x0, y0, n0, n1);
  __syncthreads();

  T cen[NPT][NF];
  for (int p = 0; p < NPT; ++p) {
    const int q = tid + p * NT;
    const int a = q / TY, b = q - a * TY;
    const int gi = x0 + a, gj = y0 + b;
    const int ea = a + RJ, eb = b + RJ;
    for (int k = 0; k < NF; ++k) cen[p][k] = 0;
    if (gi < n0 && gj < n1) {
      T f[NE];
      T J[NE][NF][NS];
      eval_lds<T, Op, S0, S1>(st + (a + H) * S1 + (b + H), gi, gj, n0, n1,
                              inner, P, f, J);
      for (int e = 0; e < NE; ++e) ssq[e] += (double)f[e] * (double)f[e];
#pragma unroll
      for (int k = 0; k < NF; ++k)
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          T o = 0;
#pragma unroll
          for (int e = 0; e < NE; ++e) o += J[e][k][s] * f[e];
          if (s == 0)
            cen[p][k] = o;
          else
            OT(k, s, ea, eb) = o;
        }
    } else {
      for (int k = 0; k < NF; ++k)
        for (int s = 1; s < NS; ++s) OT(k, s, ea, eb) = 0;
    }
  }
  constexpr int NRING = E0 * E1 - TX * TY;
  for (int q = tid; q < NRING; q += NT) {
    int ea, eb;
    ring_cell<TX, TY, RJ>(q, ea, eb);
    const int gi = x0 - RJ + ea, gj = y0 - RJ + eb;
    if (gi >= 0 && gi < n0 && gj >= 0 && gj < n1) {
      T f[NE];
      T J[NE][NF][NS];
      eval_lds<T, Op, S0, S1>(st + (ea + Op::R) * S1 + (eb + Op::R), gi, gj,
                              n0, n1, inner, P, f, J);
#pragma unroll
      for (int k = 0; k < NF; ++k)
#pragma unroll
        for (int s = 1; s < NS; ++s) {
          T o = 0;
#pragma unroll
          for (int e = 0; e < NE; ++e) o += J[e][k][s] * f[e];
          OT(k, s, ea, eb) = o;
        }
    } else {
      for (int k = 0; k < NF; ++k)
        for (int s = 1; s < NS; ++s) OT(k, s, ea, eb) = 0;
    }
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
      for (int s = 1; s < NS; ++s)
        g += OT(k, s, a + RJ - Op::SX[s], b + RJ - Op::SY[s]);
      if (grad.f[k]) {
        if constexpr (op_accumulates<Op>::value)
          grad.f[k][c] += g * scale;
        else
          grad.f[k][c] = g * scale;
      }
    }
  }
}

// Fused loss + adjoint over the grid: one workgroup per TX x TY tile
// (XCD-aware order), per-wavefront partial sums of f_e^2 in `partial`.
template <class T, class Op, int TX, int TY>
__global__ void __launch_bounds__(256)
    k_loss_grad(FieldPtrs<T, Op::NF> in, OutPtrs<T, Op::NF> grad,
                double* __restrict__ partial, int n0, int n1, T scale,
                typename Op::Params P) {
  constexpr int NE = Op::NE;
  extern __shared__ unsigned char smem_raw[];
  T* ot = reinterpret_cast<T*>(smem_raw);  // [NF][NS-1][E0][E1]
  unsigned bx, by;
  xcd_block2(bx, by);
  double ssq[NE];
  for (int e = 0; e < NE; ++e) ssq[e] = 0;
  loss_grad_tile<T, Op, TX, TY>(in, grad, n0, n1, by * TX, bx * TY, scale, P,
                                ot, ssq);
  if (partial)
    wave_partials<NE>(
        ssq, partial, (size_t)gridDim.x * gridDim.y * kWavesPerBlock,
        (size_t)(by * gridDim.x + bx) * kWavesPerBlock + (threadIdx.x >> 6));
}

template <class T, class Op, int TX, int TY>
constexpr size_t loss_grad_smem() {
  constexpr int H = Op::RJ + Op::R;
  return sizeof(T) * (size_t)(Op::NF * (Op::NS - 1) * (TX + 2 * Op::RJ) *
                                  (TY + 2 * Op::RJ) +
                              Op::NF * (TX + 2 * H) * (TY + 2 * H));
}

// Fused matrix-free normal operator N = J^T J with the Jacobian recomputed
// from the state (no stored coefficients):
//   mode 0: y = N x          mode 1: y = b - N x
//   mode 2: y = x + omega (b - N x) / d   (weighted Jacobi, d = diag N)
//   mode 3: y = diag(N) = sum_e sum_s J[e][k][s](c - shift_s)^2
// Same tile scheme as k_loss_grad with f replaced by v = J x. STG selects
// what is staged in LDS after the products: 0 nothing (state and x read
// through the cache), 1 the state (halo RJ + R), 2 the state and x (halo
// 2 RJ). Staging cuts global loads but costs LDS, hence resident
// workgroups; the dispatch picks the measured best per precision.
// normal_apply_tile computes one TX x TY tile at (x0, y0) with the whole
// workgroup; `ot` is its LDS buffer (normal_apply_smem bytes).
template <class T, class Op, int TX, int TY, int STG>
__device__ inline void normal_apply_tile(
    const FieldPtrs<T, Op::NF>& st_in, const FieldPtrs<T, Op::NF>& x_in,
    const FieldPtrs<T, Op::NF>& b_in, const FieldPtrs<T, Op::NF>& d_in,
    const OutPtrs<T, Op::NF>& y, int n0, int n1, T omega, int mode,
    const typename Op::Params& P, int x0, int y0, T* ot) {
  constexpr int NF = Op::NF, NE = Op::NE, NS = Op::NS, RJ = Op::RJ;
  constexpr int E0 = TX + 2 * RJ, E1 = TY + 2 * RJ;
  constexpr int NT = 256;
  constexpr int NPT = (TX * TY) / NT;
  static_assert((TX * TY) % NT == 0, "tile must be a multiple of 256");
  constexpr int H = RJ + Op::R, S0 = TX + 2 * H, S1 = TY + 2 * H;
  constexpr int X0 = TX + 4 * RJ, X1 = TY + 4 * RJ;
  // ot: products [NF][NS-1][E0][E1], then state [NF][S0][S1], x [NF][X0][X1]
  T* st = ot + NF * (NS - 1) * E0 * E1;
  T* xt = st + NF * S0 * S1;
  const int tid = threadIdx.x;
  const bool diag = (mode == 3);
  const bool inner = interior_box(x0 - RJ, x0 + TX + RJ, y0 - RJ,
                                  y0 + TY + RJ, n0, n1, Op::R);
  auto OT = [&](int k, int s, int ea, int eb) -> T& {
    return ot[((k * (NS - 1) + s - 1) * E0 + ea) * E1 + eb];
  };
  // Coefficients reaching outside the grid are zero, so the clamped values
  // the staging stores there are never weighted.
  if constexpr (STG >= 1) stage_tile<T, NF, TX, TY, H>(st_in, st, x0, y0, n0,
                                                        n1);
  if constexpr (STG >= 2)
    if (!diag) stage_tile<T, NF, TX, TY, 2 * RJ>(x_in, xt, x0, y0, n0, n1);
  if constexpr (STG >= 1) __syncthreads();
  auto xat = [&](int k, int i, int j) -> T {
    if (!inner) {  // coefficients reaching outside are zero; clamp reads
      i = min(max(i, 0), n0 - 1);
      j = min(max(j, 0), n1 - 1);
    }
    return x_in.f[k][(size_t)i * n1 + j];
  };

  // Products of the cell at grown-tile position (ea, eb):
  // o[k][s] = sum_e J[e][k][s] v_e with v = J x (or J[e][k][s]^2 summed
  // over e for the diagonal).
  auto cell = [&](int ea, int eb, int gi, int gj, T (*o)[NS]) {
    T f[NE];
    T J[NE][NF][NS];
    if constexpr (STG >= 1)
      eval_lds<T, Op, S0, S1>(st + (ea + Op::R) * S1 + (eb + Op::R), gi, gj,
                              n0, n1, inner, P, f, J);
    else
      eval_global<T, Op>(st_in, gi, gj, n0, n1, inner, P, f, J);
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
    T xv[NF][NS];
    if constexpr (STG >= 2) {
      const T* xc = xt + (ea + RJ) * X1 + (eb + RJ);
#pragma unroll
      for (int k = 0; k < NF; ++k)
#pragma unroll
        for (int s = 0; s < NS; ++s)
          xv[k][s] = xc[k * (X0 * X1) + Op::SX[s] * X1 + Op::SY[s]];
    } else {
#pragma unroll
      for (int k = 0; k < NF; ++k)
#pragma unroll
        for (int s = 0; s < NS; ++s)
          xv[k][s] = xat(k, gi + Op::SX[s], gj + Op::SY[s]);
    }
    T v[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      T a = 0;
#pragma unroll
      for (int k = 0; k < NF; ++k)
#pragma unroll
        for (int s = 0; s < NS; ++s) a += J[e][k][s] * xv[k][s];
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
      for (int s = 1; s < NS; ++s) OT(k, s, ea, eb) = o[k][s];
    }
  }
  constexpr int NRING = E0 * E1 - TX * TY;
  for (int q = tid; q < NRING; q += NT) {
    int ea, eb;
    ring_cell<TX, TY, RJ>(q, ea, eb);
    const int gi = x0 - RJ + ea, gj = y0 - RJ + eb;
    T o[NF][NS];
    if (gi >= 0 && gi < n0 && gj >= 0 && gj < n1) {
      cell(ea, eb, gi, gj, o);
    } else {
      for (int k = 0; k < NF; ++k)
        for (int s = 0; s < NS; ++s) o[k][s] = 0;
    }
    for (int k = 0; k < NF; ++k)
      for (int s = 1; s < NS; ++s) OT(k, s, ea, eb) = o[k][s];
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
      for (int s = 1; s < NS; ++s)
        g += OT(k, s, a + RJ - Op::SX[s], b + RJ - Op::SY[s]);
      T out;
      if (mode == 0 || mode == 3)
        out = g;
      else if (mode == 1)
        out = b_in.f[k][c] - g;
      else
        out = (STG >= 2
                   ? xt[k * (X0 * X1) + (a + 2 * RJ) * X1 + (b + 2 * RJ)]
                   : x_in.f[k][c]) +
              omega * (b_in.f[k][c] - g) / d_in.f[k][c];
      y.f[k][c] = out;
    }
  }
}

template <class T, class Op, int TX, int TY, int STG>
__global__ void __launch_bounds__(256)
    k_normal_apply(FieldPtrs<T, Op::NF> st_in, FieldPtrs<T, Op::NF> x_in,
                   FieldPtrs<T, Op::NF> b_in, FieldPtrs<T, Op::NF> d_in,
                   OutPtrs<T, Op::NF> y, int n0, int n1, T omega, int mode,
                   typename Op::Params P) {
  extern __shared__ unsigned char smem_raw[];
  unsigned bx, by;
  xcd_block2(bx, by);
  normal_apply_tile<T, Op, TX, TY, STG>(st_in, x_in, b_in, d_in, y, n0, n1,
                                        omega, mode, P, by * TX, bx * TY,
                                        reinterpret_cast<T*>(smem_raw));
}

template <class T, class Op, int TX, int TY, int STG>
constexpr size_t normal_apply_smem() {
  constexpr int H = Op::RJ + Op::R;
  return sizeof(T) *
         (size_t)(Op::NF * (Op::NS - 1) * (TX + 2 * Op::RJ) *
                      (TY + 2 * Op::RJ) +
                  (STG >= 1 ? Op::NF * (TX + 2 * H) * (TY + 2 * H) : 0) +
                  (STG >= 2 ? Op::NF * (TX + 4 * Op::RJ) * (TY + 4 * Op::RJ)
                            : 0));
}

// Operators may supply faster code for tiles of interior cells (cells at
// least R + RJ from every wall: no wall rule applies to them or to the
// Jacobian rows they gather from):
//   static constexpr bool INTERIOR_TILES = true;
//   interior_adjoint_smem<TX, TY>(), interior_adjoint_tile<TX, TY>(...)
//   interior_normal_smem<TX, TY>(),  interior_normal_tile<TX, TY>(...)
// taking the bounds (v0, v1) of the interior cells a partial tile may
// store. The loss-adjoint and normal-operator launches then run those tiles
// over the interior and per-cell frame blocks over the frame of width
// R + RJ (generic path; a thin split).
template <class Op, class = void>
struct has_interior_tiles : std::false_type {};
template <class Op>
struct has_interior_tiles<Op, std::void_t<decltype(Op::INTERIOR_TILES)>>
    : std::bool_constant<Op::INTERIOR_TILES> {};

// Loss adjoint of frame cells (within H = R + RJ of a face) on the
// operator's generic path, for frame block fb of a thin split: each thread
// evaluates f and J at one neighbour m = c - shift_s of one frame cell c
// (NS threads per cell, 256 / NS cells per block) and keeps
// out_{k,s} = sum_e J_e(m)[k][s] f_e(m) in LDS; one thread per cell then
// sums grad_k(c) = scale sum_s out_{k,s} in the order of loss_grad_tile
// (centre first). The thread of shift 0 adds f_e(c)^2 to ssq[e].
template <class T, class Op>
__device__ inline void frame_adjoint(const FieldPtrs<T, Op::NF>& in,
                                     const OutPtrs<T, Op::NF>& grad, int n0,
                                     int n1, T scale,
                                     const typename Op::Params& P,
                                     const ThinSplit& sp, int fb, T* o,
                                     double* ssq) {
  constexpr int NF = Op::NF, NE = Op::NE, NS = Op::NS;
  constexpr int H = Op::R + Op::RJ, PER = 256 / NS;
  const int t = threadIdx.x;
  const int cl = t / NS, sh = t - cl * NS;
  const int f = fb * PER + cl;
  if (cl < PER && f < sp.nframe) {
    int i, j;
    sp.frame_cell(f, n0, n1, H, i, j);
    // The shift is selected by unrolled compares so that J stays in
    // registers (a run-time index would move it to scratch memory).
    int mi = i, mj = j;
#pragma unroll
    for (int q = 0; q < NS; ++q)
      if (q == sh) {
        mi = i - Op::SX[q];
        mj = j - Op::SY[q];
      }
    T out[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) out[k] = 0;
    if (mi >= 0 && mi < n0 && mj >= 0 && mj < n1) {
      GlobalAcc<T, NF> acc{in.f, n0, n1, mi, mj};
      T fe[NE];
      T J[NE][NF][NS];
      Op::template eval<false>(acc, mi, mj, n0, n1, P, fe, J);
#pragma unroll
      for (int q = 0; q < NS; ++q)
        if (q == sh)
#pragma unroll
          for (int k = 0; k < NF; ++k) {
            T v = 0;
#pragma unroll
            for (int e = 0; e < NE; ++e) v += J[e][k][q] * fe[e];
            out[k] = v;
          }
      if (sh == 0)
        for (int e = 0; e < NE; ++e) ssq[e] += (double)fe[e] * (double)fe[e];
    }
#pragma unroll
    for (int k = 0; k < NF; ++k) o[(cl * NS + sh) * NF + k] = out[k];
  }
  __syncthreads();
  const int fc = fb * PER + t;
  if (t < PER && fc < sp.nframe) {
    int i, j;
    sp.frame_cell(fc, n0, n1, H, i, j);
    const size_t c = (size_t)i * n1 + j;
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      T g = o[(t * NS) * NF + k];
#pragma unroll
      for (int q = 1; q < NS; ++q) g += o[(t * NS + q) * NF + k];
      if (grad.f[k]) {
        if constexpr (op_accumulates<Op>::value)
          grad.f[k][c] += g * scale;
        else
          grad.f[k][c] = g * scale;
      }
    }
  }
}

// k_loss_grad over a thin split: the first nfb blocks in dispatch order run
// frame_adjoint (dispatched first and spread over the XCDs, so that their
// dependent global loads overlap the tiles), the
// others the operator's interior_adjoint_tile on tiles of interior cells
// (H = R + RJ from every face; the last tile of a row or column partial).
// Partial slot of block b, wave w: b * 4 + w.
template <class T, class Op, int TX, int TY>
__global__ void __launch_bounds__(256)
    k_loss_grad_split(FieldPtrs<T, Op::NF> in, OutPtrs<T, Op::NF> grad,
                      double* __restrict__ partial, int n0, int n1, T scale,
                      typename Op::Params P, ThinSplit sp) {
  extern __shared__ unsigned char smem_raw[];
  T* smem = reinterpret_cast<T*>(smem_raw);
  constexpr int H = Op::R + Op::RJ;
  const unsigned b = blockIdx.x;
  double ssq[Op::NE];
  for (int e = 0; e < Op::NE; ++e) ssq[e] = 0;
  if ((int)b >= sp.nfb) {
    const unsigned t = xcd_block(b - (unsigned)sp.nfb, (unsigned)sp.nint);
    const int x0 = H + (int)(t / (unsigned)sp.ntx) * TX;
    const int y0 = H + (int)(t % (unsigned)sp.ntx) * TY;
    Op::template interior_adjoint_tile<TX, TY>(in, grad, n0, n1, x0, y0,
                                               scale, P, smem, ssq, n0 - H,
                                               n1 - H);
  } else {
    frame_adjoint<T, Op>(in, grad, n0, n1, scale, P, sp, (int)b, smem, ssq);
  }
  if (partial)
    wave_partials<Op::NE>(ssq, partial, (size_t)gridDim.x * kWavesPerBlock,
                          (size_t)b * kWavesPerBlock + (threadIdx.x >> 6));
}

// Partial count of the loss-adjoint launch of operator Op (TX x TY tiles).
template <class Op, int TX, int TY>
inline int loss_grad_nblocks_t(int n0, int n1) {
  if constexpr (has_interior_tiles<Op>::value) {
    const ThinSplit sp =
        thin_split(n0, n1, Op::R + Op::RJ, TX, TY, 256 / Op::NS);
    return (sp.nint + sp.nfb) * kWavesPerBlock;
  }
  return ceil_div(n1, TY) * ceil_div(n0, TX) * kWavesPerBlock;
}
template <class Op>
inline int loss_grad_nblocks(bool fp32, int n0, int n1) {
  return fp32 ? loss_grad_nblocks_t<Op, kGradTx32, kGradTy32>(n0, n1)
              : loss_grad_nblocks_t<Op, kGradTx64, kGradTy64>(n0, n1);
}
// Partial count of the normal-operator launch with a dot product (the thin
// split of operators with interior tiles, mode 0): one per workgroup.
template <class Op>
inline int normal_nblocks(bool fp32, int n0, int n1) {
  const ThinSplit sp =
      fp32 ? thin_split(n0, n1, Op::R + Op::RJ, kNormTx32, kNormTy32,
                        256 / Op::NS)
           : thin_split(n0, n1, Op::R + Op::RJ, kNormTx64, kNormTy64,
                        256 / Op::NS);
  return sp.nint + sp.nfb;
}

// Normal operator N = J^T J (modes 0, 1, 2 of normal_apply_tile) at frame
// cells (within H = R + RJ of a face) on the operator's generic path, for
// frame block fb of a thin split: each thread evaluates J at one neighbour
// m = c - shift_s of one frame cell c (NS threads per cell), v = (J x)(m)
// from clamped global reads of x (coefficients reaching outside are zero),
// and keeps out_{k,s} = sum_e J_e(m)[k][s] v_e in LDS; one thread per cell
// sums g_k = sum_s out_{k,s} (centre first, as normal_apply_tile) and
// stores g, b - g or the Jacobi update.
template <class T, class Op>
__device__ inline void frame_normal(
    const FieldPtrs<T, Op::NF>& st_in, const FieldPtrs<T, Op::NF>& x_in,
    const FieldPtrs<T, Op::NF>& b_in, const FieldPtrs<T, Op::NF>& d_in,
    const OutPtrs<T, Op::NF>& y, int n0, int n1, T omega, int mode,
    const typename Op::Params& P, const ThinSplit& sp, int fb, T* o,
    bool dot, double& dacc) {
  constexpr int NF = Op::NF, NE = Op::NE, NS = Op::NS;
  constexpr int H = Op::R + Op::RJ, PER = 256 / NS;
  const int t = threadIdx.x;
  const int cl = t / NS, sh = t - cl * NS;
  const int f = fb * PER + cl;
  if (cl < PER && f < sp.nframe) {
    int i, j;
    sp.frame_cell(f, n0, n1, H, i, j);
    // The shift is selected by unrolled compares so that J stays in
    // registers.
    int mi = i, mj = j;
#pragma unroll
    for (int q = 0; q < NS; ++q)
      if (q == sh) {
        mi = i - Op::SX[q];
        mj = j - Op::SY[q];
      }
    T out[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) out[k] = 0;
    if (mi >= 0 && mi < n0 && mj >= 0 && mj < n1) {
      GlobalAcc<T, NF> acc{st_in.f, n0, n1, mi, mj};
      T fe[NE];
      T J[NE][NF][NS];
      Op::template eval<false>(acc, mi, mj, n0, n1, P, fe, J);
      T v[NE];
#pragma unroll
      for (int e = 0; e < NE; ++e) v[e] = 0;
#pragma unroll
      for (int k = 0; k < NF; ++k)
#pragma unroll
        for (int q = 0; q < NS; ++q) {
          const int xi = min(max(mi + Op::SX[q], 0), n0 - 1);
          const int xj = min(max(mj + Op::SY[q], 0), n1 - 1);
          const T xv = x_in.f[k][(size_t)xi * n1 + xj];
#pragma unroll
          for (int e = 0; e < NE; ++e) v[e] += J[e][k][q] * xv;
        }
#pragma unroll
      for (int q = 0; q < NS; ++q)
        if (q == sh)
#pragma unroll
          for (int k = 0; k < NF; ++k) {
            T a = 0;
#pragma unroll
            for (int e = 0; e < NE; ++e) a += J[e][k][q] * v[e];
            out[k] = a;
          }
    }
#pragma unroll
    for (int k = 0; k < NF; ++k) o[(cl * NS + sh) * NF + k] = out[k];
  }
  __syncthreads();
  const int fc = fb * PER + t;
  if (t < PER && fc < sp.nframe) {
    int i, j;
    sp.frame_cell(fc, n0, n1, H, i, j);
    const size_t c = (size_t)i * n1 + j;
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      T g = o[(t * NS) * NF + k];
#pragma unroll
      for (int q = 1; q < NS; ++q) g += o[(t * NS + q) * NF + k];
      T r;
      if (mode == 0)
        r = g;
      else if (mode == 1)
        r = b_in.f[k][c] - g;
      else
        r = x_in.f[k][c] + omega * (b_in.f[k][c] - g) / d_in.f[k][c];
      y.f[k][c] = r;
      if (dot) dacc += (double)x_in.f[k][c] * (double)r;
    }
  }
}

// k_normal_apply (modes 0, 1, 2) over a thin split: the first nfb blocks in
// dispatch order run frame_normal, the others the operator's
// interior_normal_tile on tiles of interior cells (the last tile of a row
// or column partial). With `dotp` (mode 0) each workgroup also writes its
// FP64 sum of x * y over its cells to dotp[b] (fixed order; one partial per
// workgroup keeps the final sum short).
template <class T, class Op, int TX, int TY, int STG>
__global__ void __launch_bounds__(256)
    k_normal_apply_split(FieldPtrs<T, Op::NF> st_in,
                         FieldPtrs<T, Op::NF> x_in, FieldPtrs<T, Op::NF> b_in,
                         FieldPtrs<T, Op::NF> d_in, OutPtrs<T, Op::NF> y,
                         int n0, int n1, T omega, int mode,
                         typename Op::Params P, ThinSplit sp,
                         double* __restrict__ dotp) {
  extern __shared__ unsigned char smem_raw[];
  T* smem = reinterpret_cast<T*>(smem_raw);
  constexpr int H = Op::R + Op::RJ;
  const unsigned b = blockIdx.x;
  const bool dot = dotp != nullptr;
  double dacc = 0;
  if ((int)b >= sp.nfb) {
    const unsigned t = xcd_block(b - (unsigned)sp.nfb, (unsigned)sp.nint);
    const int x0 = H + (int)(t / (unsigned)sp.ntx) * TX;
    const int y0 = H + (int)(t % (unsigned)sp.ntx) * TY;
    Op::template interior_normal_tile<TX, TY>(st_in, x_in, b_in, d_in, y,
                                              n0, n1, omega, mode, P, x0, y0,
                                              smem, n0 - H, n1 - H, dot,
                                              dacc);
  } else {
    frame_normal<T, Op>(st_in, x_in, b_in, d_in, y, n0, n1, omega, mode, P,
                        sp, (int)b, smem, dot, dacc);
  }
  if (dot) {
    __shared__ double red[kWavesPerBlock];
    const double v = block_sum(dacc, red);
    if (threadIdx.x == 0) dotp[b] = v;
  }
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
  const ThinSplit sp = residual_split(n0, n1, Op::R);
  const int nb = sp.nint + sp.nfb;
  if (nblocks) *nblocks = nb * kWavesPerBlock;
  constexpr size_t smem = residual_smem<T, Op>();
  static_assert(smem <= 64 * 1024, "residual tile too large");
  hipLaunchKernelGGL((k_residual<T, Op>), dim3(nb), dim3(256), smem, stream,
                     in, out, partial, n0, n1, P, sp);
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
  constexpr size_t smem = loss_grad_smem<T, Op, TX, TY>();
  static_assert(smem <= 160 * 1024, "LDS tile too large");
  if constexpr (has_interior_tiles<Op>::value) {
    const ThinSplit sp =
        thin_split(n0, n1, Op::R + Op::RJ, TX, TY, 256 / Op::NS);
    const int nb = sp.nint + sp.nfb;
    if (nblocks) *nblocks = nb * kWavesPerBlock;
    constexpr size_t smem2 = std::max(
        Op::template interior_adjoint_smem<TX, TY>(),
        sizeof(T) * (size_t)(256 / Op::NS) * Op::NS * Op::NF);
    static_assert(smem2 <= 160 * 1024, "LDS tile too large");
    auto kern = k_loss_grad_split<T, Op, TX, TY>;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)smem2);
    hipLaunchKernelGGL(kern, dim3(nb), dim3(256), smem2, stream, in, out,
                       partial, n0, n1, scale, P, sp);
    return check_launch("loss_grad");
  }
  if (nblocks) *nblocks = grid.x * grid.y * kWavesPerBlock;
  auto kern = k_loss_grad<T, Op, TX, TY>;
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  hipLaunchKernelGGL(kern, grid, dim3(256), smem, stream, in, out, partial,
                     n0, n1, scale, P);
  return check_launch("loss_grad");
}

template <class T, class Op, int TX, int TY, int STG>
int launch_normal_apply_t(const void* const* state, const void* const* x,
                          const void* const* b, const void* const* d,
                          void* const* y, int n0, int n1, double omega,
                          int mode, const typename Op::Params& P,
                          hipStream_t stream, double* dotp = nullptr,
                          int* nblocks = nullptr) {
  if (mode < 0 || mode > 3) {
    set_error("normal_apply: mode must be 0..3");
    return kErrArg;
  }
  if (dotp && (mode != 0 || !has_interior_tiles<Op>::value)) {
    set_error("normal_apply: the dot product needs mode 0 and an operator "
              "with interior tiles");
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
  constexpr size_t smem = normal_apply_smem<T, Op, TX, TY, STG>();
  static_assert(smem <= 160 * 1024, "LDS tile too large");
  if constexpr (has_interior_tiles<Op>::value) {
    if (mode != 3) {
      const ThinSplit sp =
          thin_split(n0, n1, Op::R + Op::RJ, TX, TY, 256 / Op::NS);
      constexpr size_t smem2 = std::max(
          Op::template interior_normal_smem<TX, TY>(),
          sizeof(T) * (size_t)(256 / Op::NS) * Op::NS * Op::NF);
      static_assert(smem2 <= 160 * 1024, "LDS tile too large");
      auto kern = k_normal_apply_split<T, Op, TX, TY, STG>;
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)smem2);
      if (nblocks) *nblocks = sp.nfb + sp.nint;
      hipLaunchKernelGGL(kern, dim3(sp.nfb + sp.nint), dim3(256), smem2,
                         stream, S, X, B, D, Y, n0, n1, (T)omega, mode, P,
                         sp, dotp);
      return check_launch("normal_apply");
    }
  }
  auto kern = k_normal_apply<T, Op, TX, TY, STG>;
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
                          const typename Op64::Params& P, hipStream_t s,
                          double* dotp = nullptr, int* nblocks = nullptr) {
  if (dtype == kF64)
    return launch_normal_apply_t<double, Op64, kNormTx64, kNormTy64,
                                 kNormStage64>(
        state, x, b, d, y, n0, n1, omega, mode, P, s, dotp, nblocks);
  if (dtype == kF32)
    return launch_normal_apply_t<float, Op32, kNormTx32, kNormTy32,
                                 kNormStage32>(
        state, x, b, d, y, n0, n1, omega, mode, P, s, dotp, nblocks);
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
    return launch_loss_grad_t<double, Op64, kGradTx64, kGradTy64>(
        fields, out, partial, n0, n1, scale, P, s, nblocks);
  if (dtype == kF32)
    return launch_loss_grad_t<float, Op32, kGradTx32, kGradTy32>(
        fields, out, partial, n0, n1, (float)scale, P, s, nblocks);
  set_error("jtf: bad dtype %d", dtype);
  return kErrArg;
}

}  // namespace odil
