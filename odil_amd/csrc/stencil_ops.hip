// Block stencil operators on structured grids and the geometric-multigrid
// building blocks (reference: SparseOperator / MultigridOp, multigrid.py:
// 342-602, 877-1214).
//
// A block stencil operator A maps nvar_in input fields on a grid of nw points
// (2D or 3D, C-order) to nvar_out output fields on nwh points
// (nwh = (nw - 1) / stride + 1):
//     (A x)_o[I] = sum_{e : out(e) = o} coef_e[I] * x_{in(e)}[stride*I + s_e]
// with zero fill where stride*I + s_e leaves the grid (non-circular shift,
// multigrid.py:941-947). Coefficients are one contiguous array
// coef[nent][nwh]; the entry table is int32 [nent][5] = (out, in, s0, s1, s2)
// sorted by `out`, with ostart[nvar_out + 1] delimiting each output field.
#include <algorithm>
#include <type_traits>

#include "common.h"

namespace odil {

constexpr int kMaxVar = 4;

template <class T>
struct VecPtrs {
  T* p[kMaxVar];
};
template <class T>
struct CVecPtrs {
  const T* p[kMaxVar];
};

struct Grid3 {
  int n0, n1, n2;
  __host__ __device__ size_t size() const { return (size_t)n0 * n1 * n2; }
};

// Grids hold fewer than 2^31 points (checked on the host), so index
// arithmetic is 32-bit (64-bit division is a long instruction sequence).
__device__ inline void unflat(size_t c, const Grid3& g, int& i0, int& i1,
                              int& i2) {
  const unsigned cc = (unsigned)c;
  i2 = (int)(cc % (unsigned)g.n2);
  const unsigned r = cc / (unsigned)g.n2;
  i1 = (int)(r % (unsigned)g.n1);
  i0 = (int)(r / (unsigned)g.n1);
}

__device__ inline bool inside(int a, int b, int c, const Grid3& g) {
  return a >= 0 && a < g.n0 && b >= 0 && b < g.n1 && c >= 0 && c < g.n2;
}

// mode 0: y = A x; mode 1: y = f - A x;
// mode 2: weighted Jacobi y = x + omega (f - A x) / diag(A)
//         (smoother_jacobi, multigrid.py:543-552; diag = entry dent[o]).
// Workgroup (bx, o) serves output field o and stages that field's entries
// in LDS. A group of LANES consecutive threads owns one cell, cells fastest
// so coefficient reads coalesce; the lanes split the entries and combine by
// shuffles. Small grids use wide groups so the serial chain of dependent
// loads per thread stays short. Loads use clamped in-grid indices and a 0/1
// weight instead of branches so they can all be issued back to back.
constexpr int kMaxEntries = 512;

template <class T, int LANES>
__global__ void __launch_bounds__(256)
    k_apply(const int* __restrict__ ent, const int* __restrict__ ostart,
            const int* __restrict__ dent, const T* __restrict__ coef,
            CVecPtrs<T> x, VecPtrs<T> y, CVecPtrs<T> f, Grid3 nw, Grid3 nwh,
            int stride, int mode, T omega) {
  __shared__ int4 tab[kMaxEntries];  // (in, s0, s1, s2)
  // Logical block: tile bx, output field o fastest, so every field of one
  // tile runs on the same XCD (shared x reads in its L2).
  const unsigned g = xcd_block(blockIdx.x + blockIdx.y * gridDim.x,
                               gridDim.x * gridDim.y);
  const int o = g % gridDim.y;
  const unsigned bx = g / gridDim.y;
  const int e0 = ostart[o];
  const int ne = ostart[o + 1] - e0;
  for (int t = threadIdx.x; t < ne; t += blockDim.x) {
    const int* it = ent + 5 * (size_t)(e0 + t);
    tab[t] = make_int4(it[1], it[2], it[3], it[4]);
  }
  __syncthreads();
  const size_t Nh = nwh.size();
  const size_t c = ((size_t)bx * blockDim.x + threadIdx.x) / LANES;
  const int lane = threadIdx.x % LANES;
  if (c >= Nh) return;  // whole groups exit together
  int i0, i1, i2;
  unflat(c, nwh, i0, i1, i2);
  const int b0 = stride * i0, b1 = stride * i1, b2 = stride * i2;
  T acc = 0;
#pragma unroll 4
  for (int e = lane; e < ne; e += LANES) {
    const int4 it = tab[e];
    const int p0 = b0 + it.y, p1 = b1 + it.z, p2 = b2 + it.w;
    const bool ok = inside(p0, p1, p2, nw);
    const int q0 = min(max(p0, 0), nw.n0 - 1);
    const int q1 = min(max(p1, 0), nw.n1 - 1);
    const int q2 = min(max(p2, 0), nw.n2 - 1);
    const T cv = coef[(size_t)(e0 + e) * Nh + c];
    const T xv = x.p[it.x][((size_t)q0 * nw.n1 + q1) * nw.n2 + q2];
    acc += (ok ? cv : T(0)) * xv;
  }
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1)
    acc += __shfl_xor(acc, off, LANES);
  if (lane != 0) return;
  if (mode == 0) {
    y.p[o][c] = acc;
  } else if (mode == 1) {
    y.p[o][c] = f.p[o][c] - acc;
  } else {
    const T d = coef[(size_t)dent[o] * Nh + c];
    y.p[o][c] = x.p[o][c] + omega * (f.p[o][c] - acc) / d;
  }
}

// Symmetric square operator (stride 1) stored as its upper half: entries
// (o, i, s) with o < i, or o == i and s >= 0 (lexicographic). The lower
// half follows from A_(i,o,-s)(c) = A_(o,i,s)(c - s). Items list, per output
// field o, every product that reaches it:
//   direct:     coef_q[c]     * x_var[c + s]
//   transposed: coef_q[c - s] * x_var[c - s]
// as int32 [nitem][10] = (q, var, cs_lin, xs_lin, cs0, cs1, cs2, xs0, xs1,
// xs2): coefficient point c + cs, x point c + xs (cs_lin, xs_lin the same
// offsets flattened). `reach` bounds |cs|, |xs| per axis; cells at least
// that far from every face use the flattened offsets without bounds checks.
// Modes as k_apply.
// Workgroup (bx, o) serves output field o: its item list is staged in LDS
// once as flat offsets (q N + cs_lin, var N + xs_lin), so an inner cell
// reads coef[c + a] and xb[c + b] with 32-bit index arithmetic. CONTIG: the
// x fields are one allocation, x_k = x_0 + k N (otherwise the field pointer
// is selected per item). A group of LANES threads owns one cell.
// The per-axis shifts are staged beside the flat offsets, packed (three
// 10-bit fields biased by 512, |shift| <= 511 checked on the host; var in
// bits 30-31 of the x word), so that a cell near a face runs the same
// unrolled, load-batched loop as an inner cell: a product whose coefficient
// or x point falls outside the grid reads a valid address (element 0) and is
// masked to zero. Branching out per product instead serialises one memory
// round trip per item in every wavefront that holds such a cell.
constexpr int kMaxItems = 512;
constexpr int kMaxShift = 511;

__device__ inline int pack_shift(int a, int b, int c) {
  return (a + 512) | ((b + 512) << 10) | ((c + 512) << 20);
}

// Is (i0, i1, i2) + packed shift p inside the grid?
__device__ inline bool shifted_inside(int i0, int i1, int i2, int p,
                                      const Grid3& nw) {
  // Bitwise & (no short circuit) keeps the item loop free of branches.
  return ((unsigned)(i0 + (p & 1023) - 512) < (unsigned)nw.n0) &
         ((unsigned)(i1 + ((p >> 10) & 1023) - 512) < (unsigned)nw.n1) &
         ((unsigned)(i2 + ((p >> 20) & 1023) - 512) < (unsigned)nw.n2);
}

// Stage items [i0e, i0e + ne) of the global table into LDS (see above),
// flat offsets multiplied by `scale` (1: elements; sizeof(T): bytes, taken
// modulo 2^32 so that base + (unsigned)(cell bytes + offset) addresses the
// element).
template <bool CONTIG>
__device__ inline void stage_items(const int* __restrict__ items, int i0e,
                                   int ne, int N, int2* fast, int2* shf,
                                   unsigned scale = 1) {
  for (int t = threadIdx.x; t < ne; t += blockDim.x) {
    const int* it = items + 10 * (i0e + t);
    fast[t] = make_int2(
        (int)((unsigned)(it[0] * N + it[2]) * scale),
        (int)((unsigned)((CONTIG ? it[1] * N : 0) + it[3]) * scale));
    shf[t] = make_int2(pack_shift(it[4], it[5], it[6]),
                       pack_shift(it[7], it[8], it[9]) |
                           (int)((unsigned)it[1] << 30));
  }
}

template <class T>
__device__ inline const T* pick(const CVecPtrs<T>& x, int k) {
  const T* p = x.p[0];
  p = k == 1 ? x.p[1] : p;
  p = k == 2 ? x.p[2] : p;
  p = k == 3 ? x.p[3] : p;
  return p;
}

// Element at `base` + `bytes` (a 32-bit unsigned byte offset): the load
// takes the base from scalar registers and the offset from one vector
// register, one add per address instead of a 64-bit multiply-add.
template <class T>
__device__ inline T ld_b32(const T* base, unsigned bytes) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                     bytes);
}

// B32: every coefficient and x offset fits in 32 bits as bytes (checked on
// the host); the staged offsets are then byte offsets. The item table of
// the workgroup's output field is staged in LDS once (flat offsets and
// packed shifts); the epilogue operands (f, x, diagonal) are loaded before
// the products so that their latency overlaps them.
template <class T, int LANES, bool CONTIG, bool B32>
__global__ void __launch_bounds__(256)
    k_apply_sym(const int* __restrict__ items, const int* __restrict__ ostart,
                const int* __restrict__ dent, const T* __restrict__ coef,
                CVecPtrs<T> x, VecPtrs<T> y, CVecPtrs<T> f, Grid3 nw,
                Grid3 reach, int mode, T omega, int per) {
  __shared__ int2 fast[kMaxItems];
  __shared__ int2 shf[kMaxItems];
  // Logical block: a chunk of `per` consecutive tiles, output field o
  // fastest (one XCD per chunk); the item list is staged once per chunk.
  const unsigned g = xcd_block(blockIdx.x + blockIdx.y * gridDim.x,
                               gridDim.x * gridDim.y);
  const int o = g % gridDim.y;
  const unsigned bx = g / gridDim.y;
  const int i0e = ostart[o];
  const int ne = ostart[o + 1] - i0e;
  const int N = (int)nw.size();
  constexpr unsigned S = B32 ? sizeof(T) : 1;
  stage_items<CONTIG>(items, i0e, ne, N, fast, shf, S);
  __syncthreads();
  // Flat (scaled) offsets and packed shifts of item e.
  auto item = [&](int e) -> int2 { return fast[e]; };
  auto shifts = [&](int e) -> int2 { return shf[e]; };
  auto var_of = [&](int e) -> int {
    return (int)((unsigned)shf[e].y >> 30);
  };
  const int lane = threadIdx.x % LANES;
  for (int r = 0; r < per; ++r) {
    const int c = (int)((((size_t)bx * per + r) * blockDim.x + threadIdx.x) /
                        LANES);
    if (c >= N) return;  // whole groups exit together; later tiles too
    // Epilogue operands first (their loads overlap the products).
    T fo = 0, xo = 0, dd = 1;
    if (lane == 0) {
      if (mode >= 1) fo = f.p[o][c];
      if (mode == 2) {
        xo = x.p[o][c];
        dd = coef[(size_t)dent[o] * N + c];
      }
    }
    // Staged offset `off` of array `a` relative to this cell.
    const unsigned cs = (unsigned)c * S;
    auto at = [&](const T* a, int off) -> T {
      return B32 ? ld_b32(a, cs + (unsigned)off) : a[(unsigned)(c + off)];
    };
    int i0, i1, i2;
    unflat(c, nw, i0, i1, i2);
    const bool inner = i0 >= reach.n0 && i0 < nw.n0 - reach.n0 &&
                       i1 >= reach.n1 && i1 < nw.n1 - reach.n1 &&
                       i2 >= reach.n2 && i2 < nw.n2 - reach.n2;
    T acc = 0;
    // One loop per wavefront: one that holds any cell near a face runs the
    // masked loop for all its cells instead of both loops in turn. Items go
    // in batches of U per lane, every load of a batch issued before its
    // products: full batches first, then one predicated batch (items past
    // the end reread the last item with a zero coefficient). A remainder
    // loop would wait for each item's loads in turn, one memory round trip
    // per item.
    constexpr int U = LANES <= 2 ? 16 : (LANES == 4 ? 12 : (LANES == 8 ? 8 : 4));
    constexpr int UM = LANES <= 4 ? 8 : U;
    if (__all(inner)) {
      if (CONTIG) {
        const T* __restrict__ xb = x.p[0];
        auto batch = [&](int e0, auto tail) {
          T pa[U], pb[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int2 it = item(decltype(tail)::value ? min(e0 + u * LANES, ne - 1)
                                                       : e0 + u * LANES);
            pa[u] = at(coef, it.x);
            pb[u] = at(xb, it.y);
          }
#pragma unroll
          for (int u = 0; u < U; ++u)
            acc += (!decltype(tail)::value || e0 + u * LANES < ne ? pa[u] : T(0)) * pb[u];
        };
        int e0 = lane;
        for (; e0 + (U - 1) * LANES < ne; e0 += U * LANES)
          batch(e0, std::false_type{});
        if (e0 < ne) batch(e0, std::true_type{});
      } else {
        auto batch = [&](int e0, auto tail) {
          T pa[UM], pb[UM];
#pragma unroll
          for (int u = 0; u < UM; ++u) {
            const int e = decltype(tail)::value ? min(e0 + u * LANES, ne - 1)
                                                : e0 + u * LANES;
            const int2 it = item(e);
            pa[u] = at(coef, it.x);
            pb[u] = at(pick(x, var_of(e)), it.y);
          }
#pragma unroll
          for (int u = 0; u < UM; ++u)
            acc += (!decltype(tail)::value || e0 + u * LANES < ne ? pa[u] : T(0)) * pb[u];
        };
        int e0 = lane;
        for (; e0 + (UM - 1) * LANES < ne; e0 += UM * LANES)
          batch(e0, std::false_type{});
        if (e0 < ne) batch(e0, std::true_type{});
      }
    } else {
      // Cells near a face: products outside the grid masked to zero (their
      // loads read element 0).
      for (int e0 = lane; e0 < ne; e0 += UM * LANES) {
        T pa[UM], pb[UM];
        bool pk[UM];
#pragma unroll
        for (int u = 0; u < UM; ++u) {
          const int e = min(e0 + u * LANES, ne - 1);
          const int2 it = item(e);
          const int2 sh = shifts(e);
          const bool ok = (e0 + u * LANES < ne) &
                          shifted_inside(i0, i1, i2, sh.x, nw) &
                          shifted_inside(i0, i1, i2, sh.y, nw);
          const T* xp = CONTIG ? x.p[0] : pick(x, (unsigned)sh.y >> 30);
          pa[u] = B32 ? ld_b32(coef, ok ? cs + (unsigned)it.x : 0u)
                      : coef[ok ? (unsigned)(c + it.x) : 0u];
          pb[u] = B32 ? ld_b32(xp, ok ? cs + (unsigned)it.y : 0u)
                      : xp[ok ? (unsigned)(c + it.y) : 0u];
          pk[u] = ok;
        }
#pragma unroll
        for (int u = 0; u < UM; ++u) acc += pk[u] ? pa[u] * pb[u] : T(0);
      }
    }
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1)
      acc += __shfl_xor(acc, off, LANES);
    if (lane != 0) continue;
    if (mode == 0) {
      y.p[o][c] = acc;
    } else if (mode == 1) {
      y.p[o][c] = fo - acc;
    } else {
      y.p[o][c] = xo + omega * (fo - acc) / dd;
    }
  }
}

// Two cells per thread for contiguous x (x_k = x_0 + k N), one thread per
// cell otherwise as k_apply_sym: cells c and c + 256 (the same lane of two
// consecutive tiles) with their products interleaved, so a wavefront keeps
// twice the loads in flight and the grid has half the wavefronts (one
// resident round where single cells would need two). A second cell past
// the end of the grid mirrors the first and is not written.
template <class T, bool B32>
__global__ void __launch_bounds__(256)
    k_apply_sym2(const int* __restrict__ items, const int* __restrict__ ostart,
                 const int* __restrict__ dent, const T* __restrict__ coef,
                 CVecPtrs<T> x, VecPtrs<T> y, CVecPtrs<T> f, Grid3 nw,
                 Grid3 reach, int mode, T omega) {
  __shared__ int2 fast[kMaxItems];
  __shared__ int2 shf[kMaxItems];
  const unsigned g = xcd_block(blockIdx.x + blockIdx.y * gridDim.x,
                               gridDim.x * gridDim.y);
  const int o = g % gridDim.y;
  const unsigned bx = g / gridDim.y;
  const int i0e = ostart[o];
  const int ne = ostart[o + 1] - i0e;
  const int N = (int)nw.size();
  constexpr unsigned S = B32 ? sizeof(T) : 1;
  stage_items<true>(items, i0e, ne, N, fast, shf, S);
  __syncthreads();
  const int c0 = (int)(bx * 2 * blockDim.x + threadIdx.x);
  if (c0 >= N) return;
  const bool has1 = c0 + (int)blockDim.x < N;
  const int c1 = has1 ? c0 + (int)blockDim.x : c0;
  const unsigned s0 = (unsigned)c0 * S, s1 = (unsigned)c1 * S;
  auto at = [&](const T* a, unsigned cs, int c, int off) -> T {
    return B32 ? ld_b32(a, cs + (unsigned)off) : a[(unsigned)(c + off)];
  };
  int a0, a1, a2, b0, b1, b2;
  unflat(c0, nw, a0, a1, a2);
  unflat(c1, nw, b0, b1, b2);
  auto interior = [&](int i0, int i1, int i2) {
    return i0 >= reach.n0 && i0 < nw.n0 - reach.n0 && i1 >= reach.n1 &&
           i1 < nw.n1 - reach.n1 && i2 >= reach.n2 && i2 < nw.n2 - reach.n2;
  };
  const bool inner = interior(a0, a1, a2) && interior(b0, b1, b2);
  const T* __restrict__ xb = x.p[0];
  T acc0 = 0, acc1 = 0;
  if (__all(inner)) {
#pragma unroll 8
    for (int e = 0; e < ne; ++e) {
      const int2 it = fast[e];
      acc0 += at(coef, s0, c0, it.x) * at(xb, s0, c0, it.y);
      acc1 += at(coef, s1, c1, it.x) * at(xb, s1, c1, it.y);
    }
  } else {
    // Near a face: products outside the grid masked to zero.
#pragma unroll 2
    for (int e = 0; e < ne; ++e) {
      const int2 it = fast[e];
      const int2 sh = shf[e];
      const bool ok0 = shifted_inside(a0, a1, a2, sh.x, nw) &
                       shifted_inside(a0, a1, a2, sh.y, nw);
      const bool ok1 = shifted_inside(b0, b1, b2, sh.x, nw) &
                       shifted_inside(b0, b1, b2, sh.y, nw);
      const T p0 = B32 ? ld_b32(coef, ok0 ? s0 + (unsigned)it.x : 0u)
                       : coef[ok0 ? (unsigned)(c0 + it.x) : 0u];
      const T q0 = B32 ? ld_b32(xb, ok0 ? s0 + (unsigned)it.y : 0u)
                       : xb[ok0 ? (unsigned)(c0 + it.y) : 0u];
      const T p1 = B32 ? ld_b32(coef, ok1 ? s1 + (unsigned)it.x : 0u)
                       : coef[ok1 ? (unsigned)(c1 + it.x) : 0u];
      const T q1 = B32 ? ld_b32(xb, ok1 ? s1 + (unsigned)it.y : 0u)
                       : xb[ok1 ? (unsigned)(c1 + it.y) : 0u];
      acc0 += ok0 ? p0 * q0 : T(0);
      acc1 += ok1 ? p1 * q1 : T(0);
    }
  }
  auto out = [&](int c, T acc) {
    if (mode == 0) {
      y.p[o][c] = acc;
    } else if (mode == 1) {
      y.p[o][c] = f.p[o][c] - acc;
    } else {
      const T d = coef[(size_t)dent[o] * N + c];
      y.p[o][c] = x.p[o][c] + omega * (f.p[o][c] - acc) / d;
    }
  };
  out(c0, acc0);
  if (has1) out(c1, acc1);
}

// Vector form of k_apply_sym for contiguous x (x_k = x_0 + k N): each
// thread owns V = 16 / sizeof(T) consecutive nodes and reads every
// coefficient and x span with one 16-byte load (unaligned), cutting the
// memory instructions and address arithmetic by V. A span that is wholly
// interior (all V nodes at least `reach` from every face, in one row) takes
// the vector path; other spans take the per-node bounds-checked path.
template <class VT, class T>
__device__ inline VT vld(const T* p) {
  VT v;
  __builtin_memcpy(&v, p, sizeof(VT));
  return v;
}
template <class VT, class T>
__device__ inline void vst(T* p, const VT& v) {
  __builtin_memcpy(p, &v, sizeof(VT));
}

template <class T>
__global__ void __launch_bounds__(256)
    k_apply_sym_vec(const int* __restrict__ items,
                    const int* __restrict__ ostart,
                    const int* __restrict__ dent, const T* __restrict__ coef,
                    CVecPtrs<T> x, VecPtrs<T> y, CVecPtrs<T> f, Grid3 nw,
                    Grid3 reach, int mode, T omega) {
  constexpr int V = 16 / sizeof(T);
  typedef T VT __attribute__((ext_vector_type(V)));
  __shared__ int2 fast[kMaxItems];
  __shared__ int2 shf[kMaxItems];
  const unsigned g = xcd_block(blockIdx.x + blockIdx.y * gridDim.x,
                               gridDim.x * gridDim.y);
  const int o = g % gridDim.y;
  const unsigned bx = g / gridDim.y;
  const int i0e = ostart[o];
  const int ne = ostart[o + 1] - i0e;
  const int N = (int)nw.size();
  stage_items<true>(items, i0e, ne, N, fast, shf);
  __syncthreads();
  const int c0 = (int)(((size_t)bx * blockDim.x + threadIdx.x) * V);
  if (c0 >= N) return;
  const int nc = min(V, N - c0);
  int a0, a1, a2, b0, b1, b2;
  unflat(c0, nw, a0, a1, a2);
  unflat(c0 + nc - 1, nw, b0, b1, b2);
  // The span lies in one row of the contiguous (last non-trivial) axis and
  // every node of it is at least `reach` from every face.
  bool inner;
  if (nw.n2 > 1)
    inner = nc == V && a0 == b0 && a1 == b1 && a0 >= reach.n0 &&
            a0 < nw.n0 - reach.n0 && a1 >= reach.n1 &&
            a1 < nw.n1 - reach.n1 && a2 >= reach.n2 &&
            b2 < nw.n2 - reach.n2;
  else
    inner = nc == V && a0 == b0 && a0 >= reach.n0 &&
            a0 < nw.n0 - reach.n0 && a1 >= reach.n1 &&
            b1 < nw.n1 - reach.n1;
  const T* __restrict__ xb = x.p[0];
  if (inner) {
    VT acc = 0;
#pragma unroll 8
    for (int e = 0; e < ne; ++e) {
      const int2 it = fast[e];
      acc += vld<VT>(coef + (unsigned)(c0 + it.x)) *
             vld<VT>(xb + (unsigned)(c0 + it.y));
    }
    VT out;
    if (mode == 0) {
      out = acc;
    } else if (mode == 1) {
      out = vld<VT>(f.p[o] + c0) - acc;
    } else {
      const VT d = vld<VT>(coef + (size_t)dent[o] * N + c0);
      out = vld<VT>(xb + (size_t)o * N + c0) +
            omega * (vld<VT>(f.p[o] + c0) - acc) / d;
    }
    vst(y.p[o] + c0, out);
    return;
  }
  // Span near a face or across a row end: per node, products outside the
  // grid masked to zero (nodes past the end of the grid are skipped).
  int ii0[V], ii1[V], ii2[V];
  T accs[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    unflat(min(c0 + v, N - 1), nw, ii0[v], ii1[v], ii2[v]);
    accs[v] = 0;
  }
#pragma unroll 2
  for (int e = 0; e < ne; ++e) {
    const int2 it = fast[e];
    const int2 sh = shf[e];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const bool ok = (v < nc) &
                      shifted_inside(ii0[v], ii1[v], ii2[v], sh.x, nw) &
                      shifted_inside(ii0[v], ii1[v], ii2[v], sh.y, nw);
      const T a = coef[ok ? (unsigned)(c0 + v + it.x) : 0u];
      const T b = xb[ok ? (unsigned)(c0 + v + it.y) : 0u];
      accs[v] += ok ? a * b : T(0);
    }
  }
  for (int v = 0; v < nc; ++v) {
    const int c = c0 + v;
    const T acc = accs[v];
    T out;
    if (mode == 0)
      out = acc;
    else if (mode == 1)
      out = f.p[o][c] - acc;
    else
      out = xb[(size_t)o * N + c] +
            omega * (f.p[o][c] - acc) / coef[(size_t)dent[o] * N + c];
    y.p[o][c] = out;
  }
}

// LDS-tiled form of k_apply_sym for 2D grids (|shift| <= 2 per axis, x
// fields in one allocation). One workgroup owns a tile of tx x ty nodes and
// computes every output field there. The x fields on the tile grown by two
// nodes per side are staged in LDS once; the stored coefficient arrays
// stream through LDS K at a time on the same grown tile, in row segments
// (coalesced, each array read once per tile plus its ring instead of once
// per product), the next chunk's loads in flight while the current chunk's
// products run. Stored entry q = (o, i, s) adds the direct product
//   A_q(c) x_i(c + s)              to y_o(c)
// and, unless it is a diagonal centre (o == i, s == 0), the mirrored one
//   A_q(c - s) x_o(c - s)          to y_i(c)
// (A_(i,o,-s)(c) = A_(o,i,s)(c - s)). Points outside the grid are zero in
// LDS, so a product reaching outside vanishes as in k_apply_sym. The entry
// loop has no branches: the x field of a product is an LDS offset and its
// output field a uniform 0/1 weight on every accumulator, so the LDS reads
// of a whole chunk issue together. The diagonal d_o(c) of the Jacobi mode is
// taken from LDS while its chunk is resident. Tiles are dealt to XCDs in
// bands (xcd_block) so that rings are read from the XCD's L2.
constexpr int kTileReach = 2;
constexpr int kTileMaxEnt = 256;  // stored entries (LDS rows)

template <class T, int NV, int TX, int TY, int K>
__global__ void __launch_bounds__(256)
    k_apply_sym_tile(const int* __restrict__ ent, int nent,
                     const int* __restrict__ dent, const T* __restrict__ coef,
                     const T* __restrict__ xg, VecPtrs<T> y, CVecPtrs<T> f,
                     int n0, int n1, int tx, int ty, int ntc, int mode,
                     T omega) {
  constexpr int R = kTileReach;
  constexpr int W = TY + 2 * R;              // LDS row stride
  constexpr int G = (TX + 2 * R) * W;        // grown-tile capacity
  constexpr int GP = G + 1;                  // per array, with a spare slot
  constexpr int CPT = (TX * TY + 255) / 256; // nodes per thread
  constexpr int SPT = (G + 255) / 256;       // staged points per thread
  // The spare slot of each array takes the stores of dead staging points,
  // so the store loops have no branches (a branch there makes the compiler
  // wait for every load in flight).
  __shared__ T xs[NV * GP];
  __shared__ T cs[K][GP];
  __shared__ int4 erow[kTileMaxEnt];
  const unsigned g = xcd_block(blockIdx.x, gridDim.x);
  const int r0 = (int)(g / (unsigned)ntc) * tx;
  const int c0 = (int)(g % (unsigned)ntc) * ty;
  const int rows = min(tx, n0 - r0), cols = min(ty, n1 - c0);
  const int gw = cols + 2 * R;          // grown width
  const int gn = (rows + 2 * R) * gw;   // grown points
  const unsigned N = (unsigned)n0 * (unsigned)n1;
  // Staging map: point m of this thread -> byte offset of its node in a
  // field (0 and masked when outside the grid) and its LDS slot (G: the
  // spare slot).
  unsigned soff[SPT];
  int sl[SPT];
  bool sok[SPT];
#pragma unroll
  for (int m = 0; m < SPT; ++m) {
    const int p = (int)threadIdx.x + 256 * m;
    const int gr = p / gw, gc = p - gr * gw;
    const int a = r0 - R + gr, b = c0 - R + gc;
    const bool live = p < gn;
    sok[m] = live && a >= 0 && a < n0 && b >= 0 && b < n1;
    soff[m] = sok[m] ? ((unsigned)a * (unsigned)n1 + (unsigned)b) *
                           (unsigned)sizeof(T)
                     : 0u;
    sl[m] = live ? gr * W + gc : G;
  }
  // Nodes of this thread: LDS centre slot and flat index.
  int cl[CPT];
  unsigned gi[CPT];
  bool cv[CPT];
#pragma unroll
  for (int j = 0; j < CPT; ++j) {
    const int k = (int)threadIdx.x + 256 * j;
    cv[j] = k < rows * cols;
    const int r = cv[j] ? k / cols : 0, cc = cv[j] ? k - r * cols : 0;
    cl[j] = (r + R) * W + cc + R;
    gi[j] = (unsigned)(r0 + r) * (unsigned)n1 + (unsigned)(c0 + cc);
  }
  // x tile (all fields), loads first, then the LDS stores.
  {
    T v[NV][SPT];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const T* a = xg + (size_t)k * N;
#pragma unroll
      for (int m = 0; m < SPT; ++m) v[k][m] = ld_b32(a, soff[m]);
    }
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
      for (int m = 0; m < SPT; ++m)
        xs[k * GP + sl[m]] = sok[m] ? v[k][m] : T(0);
  }
  // Coefficient chunk prefetch; the in-grid mask is applied at the LDS
  // store, so nothing waits for these loads before the products.
  T pf[K][SPT];
  auto load_chunk = [&](int q0) {
#pragma unroll
    for (int kk = 0; kk < K; ++kk) {
      const int q = min(q0 + kk, nent - 1);
      const T* a = coef + (size_t)q * N;
#pragma unroll
      for (int m = 0; m < SPT; ++m) pf[kk][m] = ld_b32(a, soff[m]);
    }
  };
  // Entry rows, precomputed once per workgroup in LDS: x offsets of the
  // direct and mirrored products, the shift offset, and the output fields
  // (+1, 0: no product) of both. Read back as uniform values.
  for (int q = threadIdx.x; q < nent; q += 256) {
    const int* e = ent + 5 * q;
    const int o = e[0], i = e[1], so = e[2] * W + e[3];
    const int om = o == i && so == 0 ? -1 : i;
    erow[q] = make_int4(i * GP + so, o * GP - so, so,
                        (o + 1) | ((om + 1) << 8));
  }
  T acc[CPT][NV], dd[CPT][NV];
#pragma unroll
  for (int j = 0; j < CPT; ++j)
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      acc[j][k] = 0;
      dd[j][k] = 1;
    }
  // Diagonal entries (Jacobi mode), read once.
  int dk[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) dk[k] = mode == 2 ? dent[k] : -1;
  load_chunk(0);
  for (int q0 = 0; q0 < nent; q0 += K) {
    __syncthreads();  // the previous chunk's products are done
#pragma unroll
    for (int kk = 0; kk < K; ++kk)
#pragma unroll
      for (int m = 0; m < SPT; ++m)
        cs[kk][sl[m]] = sok[m] ? pf[kk][m] : T(0);
    __syncthreads();
    if (q0 + K < nent) load_chunk(q0 + K);
#pragma unroll
    for (int kk = 0; kk < K; ++kk) {
      const int q = q0 + kk;
      const int4 r = erow[min(q, nent - 1)];
      const int xd = __builtin_amdgcn_readfirstlane(r.x);
      const int xm = __builtin_amdgcn_readfirstlane(r.y);
      const int sh = __builtin_amdgcn_readfirstlane(r.z);
      const int fl = q < nent ? __builtin_amdgcn_readfirstlane(r.w) : 0;
      const int od = (fl & 255) - 1, om = (fl >> 8) - 1;
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        const T a = cs[kk][cl[j]];
        const T pd = a * xs[xd + cl[j]];
        const T pm = cs[kk][cl[j] - sh] * xs[xm + cl[j]];
#pragma unroll
        for (int k = 0; k < NV; ++k) {
          const T wd = od == k ? T(1) : T(0);
          const T wm = om == k ? T(1) : T(0);
          acc[j][k] += wd * pd + wm * pm;
          if (mode == 2) dd[j][k] = q == dk[k] ? a : dd[j][k];
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < CPT; ++j) {
    if (!cv[j]) continue;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      T out;
      if (mode == 0)
        out = acc[j][k];
      else if (mode == 1)
        out = f.p[k][gi[j]] - acc[j][k];
      else
        out = xs[k * GP + cl[j]] + omega * (f.p[k][gi[j]] - acc[j][k]) / dd[j][k];
      y.p[k][gi[j]] = out;
    }
  }
}

// Zero every coefficient that couples a node to a point outside the grid
// (entry q = (o, i, s) at node c with c + s outside). Such coefficients never
// enter a product (zero fill, multigrid.py:941-947); once they are zero, a
// flat read that runs past the end of a grid line lands on a real node whose
// coefficient for that shift is zero, so the sweeps can skip the per-product
// test along the contiguous axis. One thread per node; only nodes within
// `reach` of a face have work.
template <class T>
__global__ void __launch_bounds__(256)
    k_zero_outside(T* __restrict__ coef, const int* __restrict__ ent,
                   int nent, Grid3 nw, Grid3 reach) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t N = nw.size();
  if (c >= N) return;
  int i0, i1, i2;
  unflat(c, nw, i0, i1, i2);
  const bool frame = i0 < reach.n0 || i0 >= nw.n0 - reach.n0 ||
                     i1 < reach.n1 || i1 >= nw.n1 - reach.n1 ||
                     i2 < reach.n2 || i2 >= nw.n2 - reach.n2;
  if (!frame) return;
  for (int q = 0; q < nent; ++q) {
    const int* e = ent + 5 * q;
    if (!inside(i0 + e[2], i1 + e[3], i2 + e[4], nw))
      coef[(size_t)q * N + c] = T(0);
  }
}

// Transpose: y_k[p] = sum_{e: in(e)=k} sum_{I: stride*I + s_e = p}
//                      coef_e[I] x_{out(e)}[I].
// `tent` lists entry indices grouped by input field, delimited by istart.
template <class T>
__global__ void __launch_bounds__(256)
    k_apply_t(const int* __restrict__ ent, const int* __restrict__ tent,
              const int* __restrict__ istart, int nin,
              const T* __restrict__ coef, CVecPtrs<T> x, VecPtrs<T> y,
              Grid3 nw, Grid3 nwh, int stride, int accumulate) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nw.size()) return;
  int p0, p1, p2;
  unflat(p, nw, p0, p1, p2);
  const size_t Nh = nwh.size();
  for (int k = 0; k < nin; ++k) {
    T acc = 0;
    for (int t = istart[k]; t < istart[k + 1]; ++t) {
      const int e = tent[t];
      const int o = ent[5 * e];
      int q0 = p0 - ent[5 * e + 2], q1 = p1 - ent[5 * e + 3],
          q2 = p2 - ent[5 * e + 4];
      if (stride == 2) {
        if ((q0 | q1 | q2) & 1) continue;
        q0 >>= 1;
        q1 >>= 1;
        q2 >>= 1;
      }
      if (!inside(q0, q1, q2, nwh)) continue;
      const size_t I = ((size_t)q0 * nwh.n1 + q1) * nwh.n2 + q2;
      acc += coef[(size_t)e * Nh + I] * x.p[o][I];
    }
    if (accumulate)
      y.p[k][p] += acc;
    else
      y.p[k][p] = acc;
  }
}

// Normal operator N = sum_e J_e^T J_e in stencil form
// (mul_transpose_op + add_elementwise, multigrid.py:1035-1071,1138-1184):
//   N_o[m] = sum_{terms (a, b, sa)} J_a[m - sa] * J_b[m - sa]
// where the output entry o couples (var(a), var(b)) at shift s_b - s_a.
// Terms are int32 [nterm][5] = (a, b, sa0, sa1, sa2) grouped by tstart.
template <class Tin, class Tout>
__global__ void __launch_bounds__(256)
    k_normal(const int* __restrict__ terms, const int* __restrict__ tstart,
             int nout, const Tin* __restrict__ jc, Tout* __restrict__ out,
             Grid3 g) {
  // Workgroup (bx, o): output entry o (uniform, so the term table is read
  // with scalar loads), cells m along x.
  // Logical block: cell tile, output entry fastest, so all entries of one
  // tile are formed on the same XCD from the same Jacobian lines in its L2.
  const unsigned lb = xcd_block(blockIdx.x + blockIdx.y * gridDim.x,
                                gridDim.x * gridDim.y);
  const int o = lb % gridDim.y;
  const size_t N = g.size();
  const size_t m = (size_t)(lb / gridDim.y) * blockDim.x + threadIdx.x;
  if (m >= N) return;
  int m0, m1, m2;
  unflat(m, g, m0, m1, m2);
  double acc = 0;
  // A term whose cell falls outside reads element 0 and adds an exact zero
  // (same order as skipping it): the loads are not behind branches.
#pragma unroll 8
  for (int q = tstart[o]; q < tstart[o + 1]; ++q) {
    const int q0 = m0 - terms[5 * q + 2], q1 = m1 - terms[5 * q + 3],
              q2 = m2 - terms[5 * q + 4];
    const bool ok = inside(q0, q1, q2, g);
    const size_t n = ((size_t)q0 * g.n1 + q1) * g.n2 + q2;
    const double x = (double)jc[ok ? (size_t)terms[5 * q] * N + n : 0];
    const double y = (double)jc[ok ? (size_t)terms[5 * q + 1] * N + n : 0];
    acc += ok ? x * y : 0.0;
  }
  out[(size_t)o * N + m] = (Tout)acc;
}

// Galerkin coarse operator on node-based grids nw = 2 nwh - 1.
// sigma = +1: A_c = R A T. sigma = -1: the arithmetic of the reference's
// coarsen_sparse_operator (multigrid.py:390-427), which couples fine node a
// to a - s and coarse node I to I - dc, i.e. reflect(R reflect(A) T).
// Output entry o = (i, j, dc); terms int32 [nterm][9] =
// (q, r0, r1, r2, wt_id, s0, s1, s2, tr): fine coefficient array q, logical
// fine entry of block (i, j) at shift s, restriction offset r,
// interpolation weight wtab[wt_id] = 0.5^|t| for t = r + sigma (s - 2 dc).
// tr = 1 reads the logical entry from the stored upper half of a symmetric
// operator, A_(i,j,s)(a) = A_(j,i,-s)(a + s) (sigma = +1 only).
// Restriction weights rw[(r0+1)*9+(r1+1)*3+(r2+1)] at inner coarse nodes;
// at coarse boundary nodes rmode 0 injects (r = 0, weight 1, the reference
// get_R, multigrid.py:525-532) and rmode 1 keeps the weights (fine points
// outside the grid dropped), i.e. R = T^T / 2^dim.
template <class T>
__global__ void __launch_bounds__(256)
    k_galerkin(const int* __restrict__ oent, const int* __restrict__ terms,
               const int* __restrict__ tstart, int nout,
               const double* __restrict__ rw, const double* __restrict__ wtab,
               const T* __restrict__ fc, T* __restrict__ cc, Grid3 nw,
               Grid3 nwh, int dim, int sigma, int rmode) {
  // Workgroup (bx, o): output entry o (uniform, so the term table is read
  // with scalar loads), coarse nodes I along x.
  // Logical block: coarse tile, output entry fastest (one XCD per tile).
  const unsigned lb = xcd_block(blockIdx.x + blockIdx.y * gridDim.x,
                                gridDim.x * gridDim.y);
  const int o = lb % gridDim.y;
  const size_t Nc = nwh.size();
  const size_t I = (size_t)(lb / gridDim.y) * blockDim.x + threadIdx.x;
  if (I >= Nc) return;
  int I0, I1, I2;
  unflat(I, nwh, I0, I1, I2);
  const bool inner = (I0 > 0 && I0 < nwh.n0 - 1) &&
                     (I1 > 0 && I1 < nwh.n1 - 1) &&
                     (dim < 3 || (I2 > 0 && I2 < nwh.n2 - 1));
  const size_t Nf = nw.size();
  const int J0 = I0 + sigma * oent[5 * o + 2],
            J1 = I1 + sigma * oent[5 * o + 3],
            J2 = I2 + sigma * oent[5 * o + 4];
  double acc = 0;
  if (inside(J0, J1, J2, nwh)) {
    // Terms that drop out (zero weight, fine point outside the grid) read
    // element 0 and add an exact zero, in the same order: no load waits on
    // a branch.
    const bool wfull = inner || rmode == 1;
#pragma unroll 8
    for (int t = tstart[o]; t < tstart[o + 1]; ++t) {
      const int* tm = terms + 9 * t;
      const int r0 = tm[1], r1 = tm[2], r2 = tm[3];
      const double w = wfull ? rw[(r0 + 1) * 9 + (r1 + 1) * 3 + (r2 + 1)]
                             : ((r0 == 0 && r1 == 0 && r2 == 0) ? 1.0 : 0.0);
      const int a0 = 2 * I0 + r0, a1 = 2 * I1 + r1, a2 = 2 * I2 + r2;
      const int b0 = a0 + sigma * tm[5], b1 = a1 + sigma * tm[6],
                b2 = a2 + sigma * tm[7];
      const bool ok = (w != 0.0) & inside(a0, a1, a2, nw) &
                      inside(b0, b1, b2, nw);
      const size_t a = tm[8] ? ((size_t)b0 * nw.n1 + b1) * nw.n2 + b2
                             : ((size_t)a0 * nw.n1 + a1) * nw.n2 + a2;
      const double v = (double)fc[ok ? (size_t)tm[0] * Nf + a : 0];
      acc += ok ? w * wtab[tm[4]] * v : 0.0;
    }
  }
  cc[(size_t)o * Nc + I] = (T)acc;
}

// Restriction of NV fields (R.mul_field, multigrid.py:496-534). The 3^DIM
// taps are fully unrolled with every field's load issued before the sums: a
// tap with zero weight or outside the fine grid reads element 0 and adds an
// exact zero (same sum order as skipping it), so no load waits on a branch.
// With sm.p[0] != NULL, also the coarse level's first Jacobi sweep from a
// zero iterate, u_k = omega coarse_k / d_k (d: the coarse diagonal), the
// same rounding as odil_jacobi_zero on the stored coarse values.
template <class T, int NV, int DIM>
__global__ void __launch_bounds__(256)
    k_restrict(CVecPtrs<T> fine, VecPtrs<T> coarse,
               const double* __restrict__ rw, Grid3 nw, Grid3 nwh,
               int rmode, CVecPtrs<T> d, VecPtrs<T> sm, T omega) {
  const unsigned I = xcd_block(blockIdx.x, gridDim.x) * blockDim.x +
                     threadIdx.x;
  if (I >= nwh.size()) return;
  int I0, I1, I2;
  unflat(I, nwh, I0, I1, I2);
  const bool inner = (I0 > 0 && I0 < nwh.n0 - 1) &&
                     (I1 > 0 && I1 < nwh.n1 - 1) &&
                     (DIM < 3 || (I2 > 0 && I2 < nwh.n2 - 1));
  const int c0 = 2 * I0, c1 = 2 * I1, c2 = 2 * I2;
  // Boundary nodes of the injecting restriction (rmode 0) take the centre
  // value: weight 1 there and 0 elsewhere in the same branch-free sum.
  const bool inj = !inner && rmode == 0;
  T acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0;
  constexpr int R2 = DIM == 3 ? 1 : 0;
#pragma unroll
  for (int r0 = -1; r0 <= 1; ++r0)
#pragma unroll
    for (int r1 = -1; r1 <= 1; ++r1)
#pragma unroll
      for (int r2 = -R2; r2 <= R2; ++r2) {
        // Addresses from the bounds alone, so that the fine-grid loads
        // issue together with the weight loads instead of after them.
        const double wr = rw[(r0 + 1) * 9 + (r1 + 1) * 3 + (r2 + 1)];
        const bool centre = r0 == 0 && r1 == 0 && r2 == 0;
        const double w = inj ? (centre ? 1.0 : 0.0) : wr;
        const bool in = inside(c0 + r0, c1 + r1, c2 + r2, nw);
        const unsigned q =
            in ? ((unsigned)(c0 + r0) * nw.n1 + (c1 + r1)) * nw.n2 +
                     (c2 + r2)
               : 0u;
        const bool ok = in & (w != 0.0);
#pragma unroll
        for (int k = 0; k < NV; ++k) {
          const T v = fine.p[k][q];
          acc[k] += ok ? T(w) * v : T(0);
        }
      }
  // The sweep's diagonal, loaded unconditionally (without a sweep, from
  // the coarse output at the same index) so that its loads overlap the
  // restriction's; a conditional load would be waited for on its own.
  const bool smooth = sm.p[0] != nullptr;
  T dv[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) dv[k] = (smooth ? d.p[k] : coarse.p[k])[I];
#pragma unroll
  for (int k = 0; k < NV; ++k) coarse.p[k][I] = acc[k];
  if (smooth) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sm.p[k][I] = omega * acc[k] / dv[k];
  }
}

// Prolongation with add: u_fine += T e_coarse, multilinear interpolation on
// node-based grids (TT.mul_transpose_field, multigrid.py:471-494). The
// 2^DIM coarse neighbours are read unconditionally (a tap that the node's
// parity excludes reads its base node and adds an exact zero).
template <class T, int NV, int DIM>
__global__ void __launch_bounds__(256)
    k_prolong_add(CVecPtrs<T> coarse, VecPtrs<T> fine, Grid3 nw,
                  Grid3 nwh) {
  const unsigned p = xcd_block(blockIdx.x, gridDim.x) * blockDim.x +
                     threadIdx.x;
  if (p >= nw.size()) return;
  int p0, p1, p2;
  unflat(p, nw, p0, p1, p2);
  const int a0 = p0 >> 1, a1 = p1 >> 1, a2 = p2 >> 1;
  const int o0 = p0 & 1, o1 = p1 & 1, o2 = p2 & 1;
  const T w = T(1) / T(1 << (o0 + o1 + o2));
  const unsigned base = ((unsigned)a0 * nwh.n1 + a1) * nwh.n2 + a2;
  T acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0;
  constexpr int D2 = DIM == 3 ? 1 : 0;
#pragma unroll
  for (int d0 = 0; d0 <= 1; ++d0)
#pragma unroll
    for (int d1 = 0; d1 <= 1; ++d1)
#pragma unroll
      for (int d2 = 0; d2 <= D2; ++d2) {
        const bool ok = (d0 <= o0) & (d1 <= o1) & (d2 <= o2);
        const unsigned q =
            ok ? base + ((unsigned)d0 * nwh.n1 + d1) * nwh.n2 + d2 : base;
#pragma unroll
        for (int k = 0; k < NV; ++k) {
          const T v = coarse.p[k][q];
          acc[k] += ok ? v : T(0);
        }
      }
#pragma unroll
  for (int k = 0; k < NV; ++k) fine.p[k][p] += w * acc[k];
}

// Cell-centred interpolation n -> 2n per axis (MultigridDecomp.
// interp_cell_field, multigrid.py:1300-1360): fine cell p takes weights
// 3/4 on its coarse parent a = p / 2 and 1/4 on the neighbour a -/+ 1
// (even/odd p), tensor product over axes, from the coarse field padded by
// one halo cell whose value is 2 u[sym(q)] - u[refl(q)] (sym repeats the
// edge cell, refl mirrors about it, on every axis at once: a corner halo is
// 2 u_00 - u_11, as the reference's REFLECT/SYMMETRIC pads give).
__device__ inline int pad_sym(int q, int n) { return q < 0 ? 0 : (q >= n ? n - 1 : q); }
__device__ inline int pad_refl(int q, int n) {
  return q < 0 ? (n > 1 ? 1 : 0) : (q >= n ? (n > 1 ? n - 2 : 0) : q);
}

template <class T, int DIM>
__device__ inline T cell_padded(const T* u, const Grid3& c, int q0, int q1,
                                int q2) {
  const int s0 = pad_sym(q0, c.n0), s1 = DIM > 1 ? pad_sym(q1, c.n1) : 0,
            s2 = DIM > 2 ? pad_sym(q2, c.n2) : 0;
  const int r0 = pad_refl(q0, c.n0), r1 = DIM > 1 ? pad_refl(q1, c.n1) : 0,
            r2 = DIM > 2 ? pad_refl(q2, c.n2) : 0;
  const T us = u[((size_t)s0 * c.n1 + s1) * c.n2 + s2];
  if (s0 == r0 && s1 == r1 && s2 == r2) return us;
  const T ur = u[((size_t)r0 * c.n1 + r1) * c.n2 + r2];
  return T(2) * us - ur;
}

// fine (+)= I coarse, fine grid 2 c per axis.
template <class T, int DIM>
__global__ void __launch_bounds__(256)
    k_interp_cell(const T* __restrict__ u, T* __restrict__ fine, Grid3 c,
                  Grid3 f, int accumulate) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= f.size()) return;
  int p0, p1, p2;
  unflat(p, f, p0, p1, p2);
  const int a0 = p0 >> 1, a1 = p1 >> 1, a2 = p2 >> 1;
  const int b0 = a0 + ((p0 & 1) ? 1 : -1), b1 = a1 + ((p1 & 1) ? 1 : -1),
            b2 = a2 + ((p2 & 1) ? 1 : -1);
  T acc = 0;
#pragma unroll
  for (int m = 0; m < (1 << DIM); ++m) {
    const bool n0 = m & 1, n1 = DIM > 1 && (m & 2), n2 = DIM > 2 && (m & 4);
    const T w = T((n0 ? 1 : 3) * (DIM > 1 ? (n1 ? 1 : 3) : 1) *
                  (DIM > 2 ? (n2 ? 1 : 3) : 1));
    acc += w * cell_padded<T, DIM>(u, c, n0 ? b0 : a0,
                                   DIM > 1 ? (n1 ? b1 : a1) : 0,
                                   DIM > 2 ? (n2 ? b2 : a2) : 0);
  }
  acc *= T(1) / T(1 << (2 * DIM));
  fine[p] = accumulate ? fine[p] + acc : acc;
}

// Adjoint: coarse = I^T fine. Gathered per coarse cell k (no atomics):
// every fine cell within two cells of 2k in each axis re-derives its
// stencil and adds the weight it gives u[k], including through the halo
// (coefficient 2 at sym(q) == k, -1 at refl(q) == k).
template <class T, int DIM>
__global__ void __launch_bounds__(256)
    k_interp_cell_t(const T* __restrict__ fine, T* __restrict__ u, Grid3 c,
                    Grid3 f) {
  const size_t kk = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (kk >= c.size()) return;
  int k0, k1, k2;
  unflat(kk, c, k0, k1, k2);
  double acc = 0;
  const int lo1 = DIM > 1 ? 2 * k1 - 3 : 0, hi1 = DIM > 1 ? 2 * k1 + 4 : 0;
  const int lo2 = DIM > 2 ? 2 * k2 - 3 : 0, hi2 = DIM > 2 ? 2 * k2 + 4 : 0;
  for (int p0 = max(2 * k0 - 3, 0); p0 <= min(2 * k0 + 4, f.n0 - 1); ++p0)
    for (int p1 = max(lo1, 0); p1 <= min(hi1, f.n1 - 1); ++p1)
      for (int p2 = max(lo2, 0); p2 <= min(hi2, f.n2 - 1); ++p2) {
        const int a0 = p0 >> 1, a1 = p1 >> 1, a2 = p2 >> 1;
        const int b0 = a0 + ((p0 & 1) ? 1 : -1),
                  b1 = a1 + ((p1 & 1) ? 1 : -1),
                  b2 = a2 + ((p2 & 1) ? 1 : -1);
        double coef = 0;
        for (int m = 0; m < (1 << DIM); ++m) {
          const bool n0 = m & 1, n1 = DIM > 1 && (m & 2),
                     n2 = DIM > 2 && (m & 4);
          const int q0 = n0 ? b0 : a0, q1 = DIM > 1 ? (n1 ? b1 : a1) : 0,
                    q2 = DIM > 2 ? (n2 ? b2 : a2) : 0;
          const double w = (n0 ? 1 : 3) * (DIM > 1 ? (n1 ? 1 : 3) : 1) *
                           (DIM > 2 ? (n2 ? 1 : 3) : 1);
          const int s0 = pad_sym(q0, c.n0),
                    s1 = DIM > 1 ? pad_sym(q1, c.n1) : 0,
                    s2 = DIM > 2 ? pad_sym(q2, c.n2) : 0;
          const int r0 = pad_refl(q0, c.n0),
                    r1 = DIM > 1 ? pad_refl(q1, c.n1) : 0,
                    r2 = DIM > 2 ? pad_refl(q2, c.n2) : 0;
          const bool inner = s0 == r0 && s1 == r1 && s2 == r2;
          if (s0 == k0 && s1 == k1 && s2 == k2) coef += inner ? w : 2 * w;
          if (!inner && r0 == k0 && r1 == k1 && r2 == k2) coef -= w;
        }
        if (coef != 0)
          acc += coef * (double)fine[((size_t)p0 * f.n1 + p1) * f.n2 + p2];
      }
  u[kk] = (T)(acc / (double)(1 << (2 * DIM)));
}

// Dense matrix of a square block stencil operator (stride 1), row-major,
// n = nvar * N; used for the coarsest-level direct solve. sym != 0: the
// entries are the upper half of a symmetric operator.
template <class T>
__global__ void __launch_bounds__(256)
    k_to_dense(const int* __restrict__ ent, int nent, int nvar,
               const T* __restrict__ coef, double* __restrict__ M, Grid3 g,
               int sym) {
  const size_t N = g.size();
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * nent) return;
  const int e = (int)(t / N);
  const size_t c = t - (size_t)e * N;
  int c0, c1, c2;
  unflat(c, g, c0, c1, c2);
  const int q0 = c0 + ent[5 * e + 2], q1 = c1 + ent[5 * e + 3],
            q2 = c2 + ent[5 * e + 4];
  if (!inside(q0, q1, q2, g)) return;
  const size_t col = (size_t)ent[5 * e + 1] * N +
                     ((size_t)q0 * g.n1 + q1) * g.n2 + q2;
  const size_t row = (size_t)ent[5 * e] * N + c;
  // Entries of one operator have distinct (out, in, shift), so every matrix
  // element is written at most once (an upper-half symmetric operator
  // writes the mirrored element of each off-diagonal entry too).
  const double v = (double)coef[(size_t)e * N + c];
  M[row * (nvar * N) + col] = v;
  if (sym && row != col) M[col * (nvar * N) + row] = v;
}

// CSR of a stride-1 block stencil operator in the layout of the reference's
// Problem.linearize (util_op.py:877-894): row e N + c, column
// k N + flat((c + s) mod shape) (periodic wrap, as np.roll), value
// coef_q[c]. Entries are sorted by output field; row (e, c) holds the
// cnt_e = ostart[e + 1] - ostart[e] entries of field e in entry order, so
// rowptr[e N + c] = N ostart[e] + c cnt_e. One thread per (e, c).
template <class T>
__global__ void __launch_bounds__(256)
    k_to_csr(const int* __restrict__ ent, const int* __restrict__ ostart,
             int nout, const T* __restrict__ coef, int* __restrict__ rowptr,
             int* __restrict__ col, T* __restrict__ val, Grid3 g) {
  const size_t N = g.size();
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * (size_t)nout) return;
  const int e = (int)(t / N);
  const size_t c = t - (size_t)e * N;
  int c0, c1, c2;
  unflat(c, g, c0, c1, c2);
  const int q0 = ostart[e], cnt = ostart[e + 1] - q0;
  const size_t base = N * (size_t)q0 + c * (size_t)cnt;
  rowptr[t] = (int)base;
  if (t == N * (size_t)nout - 1) rowptr[t + 1] = (int)(N * (size_t)ostart[nout]);
  for (int j = 0; j < cnt; ++j) {
    const int* en = ent + 5 * (q0 + j);
    int a0 = (c0 + en[2]) % g.n0, a1 = (c1 + en[3]) % g.n1,
        a2 = (c2 + en[4]) % g.n2;
    a0 += a0 < 0 ? g.n0 : 0;
    a1 += a1 < 0 ? g.n1 : 0;
    a2 += a2 < 0 ? g.n2 : 0;
    col[base + j] = (int)((size_t)en[1] * N +
                          ((size_t)a0 * g.n1 + a1) * g.n2 + a2);
    val[base + j] = coef[(size_t)(q0 + j) * N + c];
  }
}

// y = M x for a dense row-major FP64 matrix, one wavefront per row; x and y
// may be FP32 or FP64 (the coarse solve stays FP64 inside). CONTIG: the x
// fields are one allocation (x_k = x_0 + k N), so the row and x are both
// read as contiguous vectors of nvar N values in chunks of 16 columns per
// lane, fully unrolled with masked tails (a column past the row reads
// element 0 and adds an exact zero): all of a chunk's loads are in flight
// at once, and a coarsest cavity row of 867 columns is one chunk. Otherwise
// field by field.
template <class Tx, class Ty, bool CONTIG>
__global__ void __launch_bounds__(256)
    k_gemv(const double* __restrict__ M, CVecPtrs<Tx> x, VecPtrs<Ty> y,
           int nvar, int N) {
  const int n = nvar * N;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  double acc = 0;
  const double* __restrict__ m = M + (size_t)row * n;
  if (CONTIG) {
    const Tx* __restrict__ xv = x.p[0];
    for (int j0 = 0; j0 < n; j0 += 16 * 64) {
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const int j = j0 + t * 64 + lane;
        const bool ok = j < n;
        const double a = m[ok ? j : 0];
        const double b = (double)xv[ok ? j : 0];
        acc += ok ? a * b : 0.0;
      }
    }
  } else {
    for (int k = 0; k < nvar; ++k) {
      const double* __restrict__ mk = m + (size_t)k * N;
      const Tx* __restrict__ xk = x.p[k];
#pragma unroll 8
      for (int j = lane; j < N; j += 64) acc += mk[j] * (double)xk[j];
    }
  }
  acc = wave_sum(acc);
  if (lane == 0) {
    const int k = row / N;
    y.p[k][row - k * N] = (Ty)acc;
  }
}

// Block-diagonal Jacobi (smoother_jacobi with full=True, multigrid.py:
// 289-294). Cell c has the NV x NV matrix D_c[o][i] = shift-0 coefficient of
// block (o, i) at c (every block's diagonal, the reference's extdiag). It is
// factored once per operator by LU with partial pivoting (the method of
// numpy.linalg.solve, LAPACK gesv) into lu[o NV + i][c] (unit lower factor
// below the diagonal) and piv[k][c] (the row exchanged with row k at step
// k); an exactly zero pivot flags the block singular, where numpy raises.
// The sweep solves D_c du = r_c with those factors and adds omega du to u.
struct BlockIdx {
  int q[kMaxVar * kMaxVar];  // entry of block (o, i) at shift 0; -1: none
};

template <class T, int NV>
__global__ void __launch_bounds__(256)
    k_block_factor(BlockIdx bi, const T* __restrict__ coef, T* __restrict__ lu,
                   int* __restrict__ piv, int* __restrict__ singular,
                   size_t N) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  T a[NV][NV];
#pragma unroll
  for (int o = 0; o < NV; ++o)
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int q = bi.q[o * kMaxVar + i];
      a[o][i] = q >= 0 ? coef[(size_t)q * N + c] : T(0);
    }
  bool sing = false;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    int m = k;
    T best = fabs(a[k][k]);
#pragma unroll
    for (int r = k + 1; r < NV; ++r)
      if (fabs(a[r][k]) > best) {
        best = fabs(a[r][k]);
        m = r;
      }
    piv[(size_t)k * N + c] = m;
#pragma unroll
    for (int r = k + 1; r < NV; ++r)
      if (r == m) {
#pragma unroll
        for (int j = 0; j < NV; ++j) {
          const T t = a[k][j];
          a[k][j] = a[r][j];
          a[r][j] = t;
        }
      }
    if (a[k][k] == T(0)) {
      sing = true;
      continue;
    }
#pragma unroll
    for (int r = k + 1; r < NV; ++r) {
      a[r][k] /= a[k][k];
#pragma unroll
      for (int j = k + 1; j < NV; ++j) a[r][j] -= a[r][k] * a[k][j];
    }
  }
#pragma unroll
  for (int o = 0; o < NV; ++o)
#pragma unroll
    for (int i = 0; i < NV; ++i) lu[(size_t)(o * NV + i) * N + c] = a[o][i];
  if (sing) *singular = 1;
}

template <class T, int NV>
__global__ void __launch_bounds__(256)
    k_block_jacobi(const T* __restrict__ lu, const int* __restrict__ piv,
                   CVecPtrs<T> r, VecPtrs<T> u, size_t N, T omega) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  T b[NV];
#pragma unroll
  for (int o = 0; o < NV; ++o) b[o] = r.p[o][c];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const int m = piv[(size_t)k * N + c];
#pragma unroll
    for (int q = k + 1; q < NV; ++q)
      if (q == m) {
        const T t = b[k];
        b[k] = b[q];
        b[q] = t;
      }
  }
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int q = k + 1; q < NV; ++q)
      b[q] -= lu[(size_t)(q * NV + k) * N + c] * b[k];
#pragma unroll
  for (int k = NV - 1; k >= 0; --k) {
#pragma unroll
    for (int j = k + 1; j < NV; ++j)
      b[k] -= lu[(size_t)(k * NV + j) * N + c] * b[j];
    b[k] /= lu[(size_t)(k * NV + k) * N + c];
  }
#pragma unroll
  for (int o = 0; o < NV; ++o) u.p[o][c] += omega * b[o];
}

// Forward Gauss-Seidel (smoother_gauss_seidel, multigrid.py:301-305):
// x = L^-1 (f - (A - L) u) with L the lower triangle of A, diagonal
// included, in the flat numbering (field-major, cells in C order); then
// u += omega (x - u). Cell c of field o needs x at the cells of field o that
// precede it (entries (o, o, s) with s < 0 lexicographically) and at every
// cell of the fields i < o; the other entries read the unchanged u. Cells of
// one field with equal level a0 i0 + a1 i1 + i2 are independent when
// a0 s0 + a1 s1 + s2 <= -1 for every such s (the host picks the smallest
// a0, a1), so each field is a0 (n0 - 1) + a1 (n1 - 1) + n2 waves, one launch
// each in level order. A wave's threads walk i0 from `lo` (2D grids) or the
// (i0, i1) plane from row `lo` (3D), each deriving the last index.
template <class T>
__global__ void __launch_bounds__(256)
    k_gs_wave(const int* __restrict__ ent, int e0, int e1, int qd,
              const T* __restrict__ coef, CVecPtrs<T> u, VecPtrs<T> x,
              const T* __restrict__ f, Grid3 nw, int o, int w, int a0, int a1,
              int lo, int cnt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  int i0, i1, i2;
  if (nw.n2 == 1) {
    i0 = lo + t;
    const int rem = w - a0 * i0;
    if (rem < 0 || rem % a1) return;
    i1 = rem / a1;
    i2 = 0;
    if (i1 >= nw.n1) return;
  } else {
    i0 = lo + t / nw.n1;
    i1 = t % nw.n1;
    i2 = w - a0 * i0 - a1 * i1;
    if (i2 < 0 || i2 >= nw.n2) return;
  }
  const size_t N = nw.size();
  const size_t c = ((size_t)i0 * nw.n1 + i1) * nw.n2 + i2;
  T acc = 0;
  for (int e = e0; e < e1; ++e) {
    if (e == qd) continue;
    const int* it = ent + 5 * (size_t)e;
    const int in = it[1], s0 = it[2], s1 = it[3], s2 = it[4];
    const int p0 = i0 + s0, p1 = i1 + s1, p2 = i2 + s2;
    if (!inside(p0, p1, p2, nw)) continue;
    const size_t p = ((size_t)p0 * nw.n1 + p1) * nw.n2 + p2;
    const bool lower =
        in < o ||
        (in == o && (s0 < 0 || (s0 == 0 && (s1 < 0 || (s1 == 0 && s2 < 0)))));
    const T v = lower ? x.p[in][p] : u.p[in][p];
    acc += coef[(size_t)e * N + c] * v;
  }
  x.p[o][c] = (f[c] - acc) / coef[(size_t)qd * N + c];
}

template <class T>
__global__ void __launch_bounds__(256)
    k_relax(CVecPtrs<T> x, VecPtrs<T> u, size_t N, T omega) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const int o = blockIdx.y;
  u.p[o][c] += (x.p[o][c] - u.p[o][c]) * omega;
}

}  // namespace odil

using namespace odil;

namespace {

Grid3 mk_grid(int ndim, const int* n) {
  Grid3 g{1, 1, 1};
  if (ndim == 2) {
    g.n0 = n[0];
    g.n1 = n[1];
  } else {
    g.n0 = n[0];
    g.n1 = n[1];
    g.n2 = n[2];
  }
  return g;
}

template <class T>
VecPtrs<T> vp(void* const* a, int n) {
  VecPtrs<T> v{};
  for (int k = 0; k < n && a; ++k) v.p[k] = static_cast<T*>(a[k]);
  return v;
}
template <class T>
CVecPtrs<T> cvp(const void* const* a, int n) {
  CVecPtrs<T> v{};
  for (int k = 0; k < n && a; ++k) v.p[k] = static_cast<const T*>(a[k]);
  return v;
}

bool bad_var(int n, const char* what) {
  if (n < 1 || n > kMaxVar) {
    set_error("%s: number of fields %d outside [1, %d]", what, n, kMaxVar);
    return true;
  }
  return false;
}

bool bad_dim(int ndim, const int* nw, const char* what) {
  if (ndim != 2 && ndim != 3) {
    set_error("%s: ndim must be 2 or 3, got %d", what, ndim);
    return true;
  }
  if (!nw) {
    set_error("%s: null grid size array", what);
    return true;
  }
  double cells = 1;
  for (int d = 0; d < ndim; ++d) {
    if (nw[d] < 1) {
      set_error("%s: grid size %d along axis %d", what, nw[d], d);
      return true;
    }
    cells *= nw[d];
  }
  if (cells >= 2147483648.0) {
    set_error("%s: grids need fewer than 2^31 points", what);
    return true;
  }
  return false;
}

unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

extern "C" {

// y = A x (mode 0), y = f - A x (mode 1), or one weighted-Jacobi sweep
// y = x + omega (f - A x) / diag (mode 2, stride 1, nvar_in == nvar_out).
// `ent`, `ostart`, `dent` are device int32 arrays (see file comment);
// `max_entries` is the largest entry count of one output field (<= 512).
int odil_stencil_apply(int dtype, int ndim, const int* nw, int stride,
                       int nvar_in, int nvar_out, const int* ent,
                       const int* ostart, const int* dent, int max_entries,
                       const void* coef, const void* const* x, void* const* y,
                       const void* const* f, int mode, double omega,
                       void* stream) {
  if (bad_dim(ndim, nw, "stencil_apply") || bad_var(nvar_in, "stencil_apply") ||
      bad_var(nvar_out, "stencil_apply"))
    return kErrArg;
  if (stride != 1 && stride != 2) {
    set_error("stencil_apply: stride must be 1 or 2");
    return kErrArg;
  }
  if (max_entries < 0 || max_entries > kMaxEntries) {
    set_error("stencil_apply: %d entries per field exceed %d", max_entries,
              kMaxEntries);
    return kErrCapacity;
  }
  if (mode == 2 && (stride != 1 || nvar_in != nvar_out || !dent)) {
    set_error("stencil_apply: Jacobi needs stride 1, square blocks, diag");
    return kErrArg;
  }
  if ((mode == 1 || mode == 2) && !f) {
    set_error("stencil_apply: mode %d needs f", mode);
    return kErrArg;
  }
  Grid3 g = mk_grid(ndim, nw);
  Grid3 gh{(g.n0 - 1) / stride + 1, (g.n1 - 1) / stride + 1,
           ndim == 3 ? (g.n2 - 1) / stride + 1 : 1};
  hipStream_t s = static_cast<hipStream_t>(stream);
  // Lanes per (field, cell): enough groups to fill the chip (~1M threads),
  // at most 16 lanes.
  const size_t groups = gh.size() * (size_t)nvar_out;
  int lanes = 1;
  while (lanes < 16 && groups * lanes < (1u << 20)) lanes *= 2;
  if (dtype != kF64 && dtype != kF32) {
    set_error("stencil_apply: bad dtype %d", dtype);
    return kErrArg;
  }
  dim3 grid(nblk(gh.size() * lanes), nvar_out);
#define ODIL_APPLY(T, L)                                                    \
  hipLaunchKernelGGL((k_apply<T, L>), grid, dim3(256), 0, s, ent, ostart,   \
                     dent, static_cast<const T*>(coef), cvp<T>(x, nvar_in), \
                     vp<T>(y, nvar_out), cvp<T>(f, nvar_out), g, gh,        \
                     stride, mode, (T)omega)
#define ODIL_APPLY_L(T)              \
  switch (lanes) {                   \
    case 1: ODIL_APPLY(T, 1); break; \
    case 2: ODIL_APPLY(T, 2); break; \
    case 4: ODIL_APPLY(T, 4); break; \
    case 8: ODIL_APPLY(T, 8); break; \
    default: ODIL_APPLY(T, 16); break; \
  }
  if (dtype == kF64) {
    ODIL_APPLY_L(double);
  } else {
    ODIL_APPLY_L(float);
  }
#undef ODIL_APPLY_L
#undef ODIL_APPLY
  return check_launch("stencil_apply");
}

// Symmetric operator stored as its upper half (stride 1, square): same
// modes as odil_stencil_apply. `items` int32 [nitem][10] and `ostart`
// [nvar + 1] list the products reaching each output field (see
// k_apply_sym); `reach` (host, 3 ints) bounds |cs|, |xs| per axis;
// `max_items` is the longest per-field list (at most 512); `nent` the number
// of coefficient arrays; `lanes` threads per cell (1, 2, 4, 8, 16, or 0 for
// a size-based default), -1: one thread per 16 bytes of consecutive cells
// with vector loads, or -2: two interleaved cells per thread (x in one
// allocation only for -1 and -2); `per` tiles of 256 threads
// per workgroup (1..64, scalar paths; the item list is staged once per
// workgroup). x fields laid out as one allocation (x_k = x_0 + k N) take the
// fast paths.
int odil_stencil_apply_sym(int dtype, int ndim, const int* nw, int nvar,
                           const int* items, const int* ostart,
                           const int* dent, const int* reach, int max_items,
                           int nent, int lanes, int per, const void* coef,
                           const void* const* x, void* const* y,
                           const void* const* f, int mode, double omega,
                           void* stream) {
  if (bad_dim(ndim, nw, "stencil_apply_sym") ||
      bad_var(nvar, "stencil_apply_sym"))
    return kErrArg;
  if (!items || !ostart || !coef || !x || !y) {
    set_error("stencil_apply_sym: null table, coefficient or field array");
    return kErrArg;
  }
  if (mode == 2 && !dent) {
    set_error("stencil_apply_sym: Jacobi needs the diagonal entries");
    return kErrArg;
  }
  if ((mode == 1 || mode == 2) && !f) {
    set_error("stencil_apply_sym: mode %d needs f", mode);
    return kErrArg;
  }
  if (per < 1 || per > 64) {
    set_error("stencil_apply_sym: tiles per workgroup must be in [1, 64]");
    return kErrArg;
  }
  if (max_items < 0 || max_items > kMaxItems) {
    set_error("stencil_apply_sym: %d products per field exceed %d",
              max_items, kMaxItems);
    return kErrCapacity;
  }
  if (!reach || reach[0] < 0 || reach[1] < 0 || reach[2] < 0 ||
      (ndim == 2 && reach[2] != 0)) {
    set_error("stencil_apply_sym: bad reach");
    return kErrArg;
  }
  if (reach[0] > kMaxShift || reach[1] > kMaxShift || reach[2] > kMaxShift) {
    set_error("stencil_apply_sym: shifts beyond %d per axis", kMaxShift);
    return kErrArg;
  }
  Grid3 g = mk_grid(ndim, nw);
  Grid3 rc{reach[0], reach[1], reach[2]};
  // Flat 32-bit offsets: every coefficient and x element index < 2^31.
  if ((double)g.size() * (double)(nent > nvar ? nent : nvar) > 2.0e9) {
    set_error("stencil_apply_sym: grid too large for 32-bit offsets");
    return kErrCapacity;
  }
  const size_t esz = dtype == kF64 ? 8 : 4;
  bool contig = true;
  for (int k = 1; k < nvar; ++k)
    contig = contig && static_cast<const char*>(x[k]) ==
                           static_cast<const char*>(x[0]) + k * g.size() * esz;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (lanes == 0) {  // default: enough threads to fill the chip
    const size_t groups = g.size() * (size_t)nvar;
    lanes = 1;
    while (lanes < 16 && groups * lanes < (1u << 20)) lanes *= 2;
  }
  if (lanes != -2 && lanes != -1 && lanes != 1 && lanes != 2 && lanes != 4 &&
      lanes != 8 && lanes != 16) {
    set_error("stencil_apply_sym: lanes must be -2, -1, 0, 1, 2, 4, 8 or 16");
    return kErrArg;
  }
  if (dtype != kF64 && dtype != kF32) {
    set_error("stencil_apply_sym: bad dtype %d", dtype);
    return kErrArg;
  }
  // 32-bit byte offsets when the coefficient and x arrays are below 4 GiB.
  const bool b32 = (size_t)nent * g.size() * esz < (1ull << 32) &&
                   (size_t)nvar * g.size() * esz < (1ull << 32);
  if (lanes == -2) {
    if (!contig) {
      set_error("stencil_apply_sym: two cells per thread needs x in one "
                "allocation");
      return kErrArg;
    }
    dim3 grid2((unsigned)ceil_div((int)nblk(g.size()), 2), nvar);
#define ODIL_APPLY_2C(T, B)                                                  \
  hipLaunchKernelGGL((k_apply_sym2<T, B>), grid2, dim3(256), 0, s, items,    \
                     ostart, dent, static_cast<const T*>(coef),              \
                     cvp<T>(x, nvar), vp<T>(y, nvar), cvp<T>(f, nvar), g,    \
                     rc, mode, (T)omega)
    if (dtype == kF64) {
      if (b32) ODIL_APPLY_2C(double, true); else ODIL_APPLY_2C(double, false);
    } else {
      if (b32) ODIL_APPLY_2C(float, true); else ODIL_APPLY_2C(float, false);
    }
#undef ODIL_APPLY_2C
    return check_launch("stencil_apply_sym");
  }
  if (lanes == -1) {
    if (!contig) {
      set_error("stencil_apply_sym: the vector path needs x in one "
                "allocation");
      return kErrArg;
    }
    const size_t V = dtype == kF64 ? 2 : 4;
    dim3 vgrid(nblk((g.size() + V - 1) / V), nvar);
    if (dtype == kF64)
      hipLaunchKernelGGL(k_apply_sym_vec<double>, vgrid, dim3(256), 0, s,
                         items, ostart, dent, static_cast<const double*>(coef),
                         cvp<double>(x, nvar), vp<double>(y, nvar),
                         cvp<double>(f, nvar), g, rc, mode, omega);
    else
      hipLaunchKernelGGL(k_apply_sym_vec<float>, vgrid, dim3(256), 0, s,
                         items, ostart, dent, static_cast<const float*>(coef),
                         cvp<float>(x, nvar), vp<float>(y, nvar),
                         cvp<float>(f, nvar), g, rc, mode, (float)omega);
    return check_launch("stencil_apply_sym");
  }
  dim3 grid(ceil_div((int)nblk(g.size() * lanes), per), nvar);
#define ODIL_APPLY3(T, L, C, B)                                               \
  hipLaunchKernelGGL((k_apply_sym<T, L, C, B>), grid, dim3(256), 0, s, items, \
                     ostart, dent, static_cast<const T*>(coef),               \
                     cvp<T>(x, nvar), vp<T>(y, nvar), cvp<T>(f, nvar), g,     \
                     rc, mode, (T)omega, per)
#define ODIL_APPLY2(T, L, C)      \
  if (b32)                        \
    ODIL_APPLY3(T, L, C, true);   \
  else                            \
    ODIL_APPLY3(T, L, C, false)
#define ODIL_APPLY(T, L)      \
  if (contig)                 \
    ODIL_APPLY2(T, L, true);  \
  else                        \
    ODIL_APPLY2(T, L, false)
#define ODIL_APPLY_L(T)              \
  switch (lanes) {                   \
    case 1: ODIL_APPLY(T, 1); break; \
    case 2: ODIL_APPLY(T, 2); break; \
    case 4: ODIL_APPLY(T, 4); break; \
    case 8: ODIL_APPLY(T, 8); break; \
    default: ODIL_APPLY(T, 16); break; \
  }
  if (dtype == kF64) {
    ODIL_APPLY_L(double);
  } else {
    ODIL_APPLY_L(float);
  }
#undef ODIL_APPLY_L
#undef ODIL_APPLY
#undef ODIL_APPLY2
#undef ODIL_APPLY3
  return check_launch("stencil_apply_sym");
}

}  // extern "C"

namespace {

// Tile shapes of k_apply_sym_tile (rows, columns along the contiguous axis,
// coefficient arrays per LDS chunk), selected by `tile` = 1..kNumTiles.
struct TileShape {
  int tx, ty, k;
};
constexpr int kNumTiles = 6;
constexpr TileShape kTiles[kNumTiles] = {{8, 64, 4},  {16, 32, 4},
                                         {8, 32, 4},  {8, 64, 8},
                                         {16, 64, 4}, {4, 64, 8}};

template <class T, int NV, int TX, int TY, int K>
static void launch_tile(const int* ent, int nent, const int* dent,
                        const void* coef, const void* const* x,
                        void* const* y, const void* const* f, int n0, int n1,
                        int mode, double omega, hipStream_t s) {
  // Shapes above 64 KiB of LDS are not built (the host refuses them).
  if constexpr (((size_t)(TX + 2 * kTileReach) * (TY + 2 * kTileReach) + 1) *
                        (NV + K) * sizeof(T) +
                    kTileMaxEnt * 16 <=
                65536) {
    // Balanced tiles: as many as TX x TY tiles need, sized evenly.
    const int ntr = ceil_div(n0, TX), ntc = ceil_div(n1, TY);
    const int tx = ceil_div(n0, ntr), ty = ceil_div(n1, ntc);
    hipLaunchKernelGGL((k_apply_sym_tile<T, NV, TX, TY, K>),
                       dim3((unsigned)(ntr * ntc)), dim3(256), 0, s, ent,
                       nent, dent, static_cast<const T*>(coef),
                       static_cast<const T*>(x[0]), vp<T>(y, NV),
                       cvp<T>(f, NV), n0, n1, tx, ty, ntc, mode, (T)omega);
  }
}

template <class T, int NV>
static void launch_tile_nv(int tile, const int* ent, int nent,
                           const int* dent, const void* coef,
                           const void* const* x, void* const* y,
                           const void* const* f, int n0, int n1, int mode,
                           double omega, hipStream_t s) {
#define ODIL_TILE(C, TX, TY, K)                                              \
  case C:                                                                    \
    launch_tile<T, NV, TX, TY, K>(ent, nent, dent, coef, x, y, f, n0, n1,    \
                                  mode, omega, s);                           \
    break
  switch (tile) {
    ODIL_TILE(1, 8, 64, 4);
    ODIL_TILE(2, 16, 32, 4);
    ODIL_TILE(3, 8, 32, 4);
    ODIL_TILE(4, 8, 64, 8);
    ODIL_TILE(5, 16, 64, 4);
    default:
      ODIL_TILE(6, 4, 64, 8);
  }
#undef ODIL_TILE
}

// LDS bytes of tile shape `tile` for field count nvar and element size esz.
static size_t tile_lds(int tile, int nvar, size_t esz) {
  const TileShape& t = kTiles[tile - 1];
  return ((size_t)(t.tx + 2 * kTileReach) * (t.ty + 2 * kTileReach) + 1) *
             (size_t)(nvar + t.k) * esz +
         kTileMaxEnt * 16;
}

}  // namespace

extern "C" {

int odil_stencil_tile_lds(int dtype, int nvar, int tile) {
  if (tile < 1 || tile > kNumTiles || nvar < 1 || nvar > kMaxVar)
    return -1;
  return (int)tile_lds(tile, nvar, dtype == kF64 ? 8 : 4);
}

int odil_stencil_apply_sym_tiled(int dtype, const int* nw, int nvar,
                                 const int* ent, int nent, const int* dent,
                                 const int* reach, int tile, const void* coef,
                                 const void* const* x, void* const* y,
                                 const void* const* f, int mode, double omega,
                                 void* stream) {
  if (bad_dim(2, nw, "stencil_apply_sym_tiled")) return kErrArg;
  if (nvar != 1 && nvar != 3 && nvar != 4) {
    set_error("stencil_apply_sym_tiled: %d fields (1, 3 or 4 supported)",
              nvar);
    return kErrArg;
  }
  if (!ent || !coef || !x || !y || nent < 1) {
    set_error("stencil_apply_sym_tiled: null table, coefficient or field "
              "array, or no entries");
    return kErrArg;
  }
  if (mode < 0 || mode > 2 || (mode == 2 && !dent) || (mode >= 1 && !f)) {
    set_error("stencil_apply_sym_tiled: mode %d with missing operands", mode);
    return kErrArg;
  }
  if (!reach || reach[0] < 0 || reach[1] < 0 || reach[0] > kTileReach ||
      reach[1] > kTileReach) {
    set_error("stencil_apply_sym_tiled: shifts beyond %d per axis",
              kTileReach);
    return kErrArg;
  }
  if (tile < 1 || tile > kNumTiles) {
    set_error("stencil_apply_sym_tiled: tile must be in [1, %d]", kNumTiles);
    return kErrArg;
  }
  if (nent > kTileMaxEnt) {
    set_error("stencil_apply_sym_tiled: %d stored entries (at most %d)",
              nent, kTileMaxEnt);
    return kErrCapacity;
  }
  if (dtype != kF64 && dtype != kF32) {
    set_error("stencil_apply_sym_tiled: bad dtype %d", dtype);
    return kErrArg;
  }
  const size_t esz = dtype == kF64 ? 8 : 4;
  if (tile_lds(tile, nvar, esz) > 65536) {
    set_error("stencil_apply_sym_tiled: tile %d needs %zu bytes of LDS "
              "(64 KiB per workgroup)", tile, tile_lds(tile, nvar, esz));
    return kErrCapacity;
  }
  const size_t n = (size_t)nw[0] * (size_t)nw[1];
  // 32-bit byte offsets inside one field.
  if (n * esz >= (1ull << 32)) {
    set_error("stencil_apply_sym_tiled: field above 4 GiB");
    return kErrCapacity;
  }
  for (int k = 1; k < nvar; ++k)
    if (static_cast<const char*>(x[k]) !=
        static_cast<const char*>(x[0]) + k * n * esz) {
      set_error("stencil_apply_sym_tiled: x fields must be one allocation");
      return kErrArg;
    }
  hipStream_t s = static_cast<hipStream_t>(stream);
#define ODIL_TILE_NV(T)                                                     \
  switch (nvar) {                                                           \
    case 1:                                                                 \
      launch_tile_nv<T, 1>(tile, ent, nent, dent, coef, x, y, f, nw[0],     \
                           nw[1], mode, omega, s);                          \
      break;                                                                \
    case 3:                                                                 \
      launch_tile_nv<T, 3>(tile, ent, nent, dent, coef, x, y, f, nw[0],     \
                           nw[1], mode, omega, s);                          \
      break;                                                                \
    default:                                                                \
      launch_tile_nv<T, 4>(tile, ent, nent, dent, coef, x, y, f, nw[0],     \
                           nw[1], mode, omega, s);                          \
  }
  if (dtype == kF64) {
    ODIL_TILE_NV(double);
  } else {
    ODIL_TILE_NV(float);
  }
#undef ODIL_TILE_NV
  return check_launch("stencil_apply_sym_tiled");
}

// y = A^T x (accumulate != 0: y += A^T x). `tent`/`istart` group entry
// indices by input field.
int odil_stencil_apply_t(int dtype, int ndim, const int* nw, int stride,
                         int nvar_in, int nvar_out, const int* ent,
                         const int* tent, const int* istart, const void* coef,
                         const void* const* x, void* const* y, int accumulate,
                         void* stream) {
  if (bad_dim(ndim, nw, "stencil_apply_t") ||
      bad_var(nvar_in, "stencil_apply_t") ||
      bad_var(nvar_out, "stencil_apply_t"))
    return kErrArg;
  if (stride != 1 && stride != 2) {
    set_error("stencil_apply_t: stride must be 1 or 2");
    return kErrArg;
  }
  Grid3 g = mk_grid(ndim, nw);
  Grid3 gh{(g.n0 - 1) / stride + 1, (g.n1 - 1) / stride + 1,
           ndim == 3 ? (g.n2 - 1) / stride + 1 : 1};
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    hipLaunchKernelGGL(k_apply_t<double>, dim3(nblk(g.size())), dim3(256), 0,
                       s, ent, tent, istart, nvar_in,
                       static_cast<const double*>(coef),
                       cvp<double>(x, nvar_out), vp<double>(y, nvar_in), g, gh,
                       stride, accumulate);
  else if (dtype == kF32)
    hipLaunchKernelGGL(k_apply_t<float>, dim3(nblk(g.size())), dim3(256), 0, s,
                       ent, tent, istart, nvar_in,
                       static_cast<const float*>(coef),
                       cvp<float>(x, nvar_out), vp<float>(y, nvar_in), g, gh,
                       stride, accumulate);
  else {
    set_error("stencil_apply_t: bad dtype %d", dtype);
    return kErrArg;
  }
  return check_launch("stencil_apply_t");
}

int odil_stencil_zero_outside(int dtype, int ndim, const int* nw,
                              const int* ent, int nent, const int* reach,
                              void* coef, void* stream) {
  if (bad_dim(ndim, nw, "stencil_zero_outside")) return kErrArg;
  if (!ent || !coef || nent < 0 || !reach) {
    set_error("stencil_zero_outside: null table or coefficients");
    return kErrArg;
  }
  if (nent == 0) return kOk;
  Grid3 g = mk_grid(ndim, nw);
  Grid3 rc{reach[0], reach[1], ndim == 3 ? reach[2] : 0};
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == kF64)
    hipLaunchKernelGGL(k_zero_outside<double>, dim3(nblk(g.size())),
                       dim3(256), 0, s, static_cast<double*>(coef), ent, nent,
                       g, rc);
  else if (dtype == kF32)
    hipLaunchKernelGGL(k_zero_outside<float>, dim3(nblk(g.size())),
                       dim3(256), 0, s, static_cast<float*>(coef), ent, nent,
                       g, rc);
  else {
    set_error("stencil_zero_outside: bad dtype %d", dtype);
    return kErrArg;
  }
  return check_launch("stencil_zero_outside");
}

// Normal operator sum_e J_e^T J_e from Jacobian coefficients `jc`
// (dtype_in) into `out` (dtype_out), accumulated in FP64.
int odil_stencil_normal(int dtype_in, int dtype_out, int ndim, const int* nw,
                        const int* terms, const int* tstart, int nout,
                        const void* jc, void* out, void* stream) {
  if (bad_dim(ndim, nw, "stencil_normal")) return kErrArg;
  if (nout < 1 || nout > 65535) {
    set_error("stencil_normal: bad number of output entries %d", nout);
    return kErrArg;
  }
  Grid3 g = mk_grid(ndim, nw);
  hipStream_t s = static_cast<hipStream_t>(stream);
  dim3 grid(nblk(g.size()), nout);
  if (dtype_in == kF64 && dtype_out == kF64)
    hipLaunchKernelGGL((k_normal<double, double>), grid, dim3(256), 0, s,
                       terms, tstart, nout, static_cast<const double*>(jc),
                       static_cast<double*>(out), g);
  else if (dtype_in == kF64 && dtype_out == kF32)
    hipLaunchKernelGGL((k_normal<double, float>), grid, dim3(256), 0, s, terms,
                       tstart, nout, static_cast<const double*>(jc),
                       static_cast<float*>(out), g);
  else if (dtype_in == kF32 && dtype_out == kF32)
    hipLaunchKernelGGL((k_normal<float, float>), grid, dim3(256), 0, s, terms,
                       tstart, nout, static_cast<const float*>(jc),
                       static_cast<float*>(out), g);
  else {
    set_error("stencil_normal: unsupported dtype pair %d -> %d", dtype_in,
              dtype_out);
    return kErrArg;
  }
  return check_launch("stencil_normal");
}

// Galerkin coarse operator; `rw` (27 doubles) inner restriction weights,
// `wtab` interpolation weights indexed by the terms (int32 [nterm][9], see
// k_galerkin); sigma = +1 exact R A T, -1 the reference's stencil
// coarsening.
int odil_stencil_galerkin(int dtype, int ndim, const int* nw, const int* oent,
                          const int* terms, const int* tstart, int nout,
                          const double* rw, const double* wtab, const void* fc,
                          void* cc, int sigma, int rmode, void* stream) {
  if (sigma != 1 && sigma != -1) {
    set_error("stencil_galerkin: sigma must be +1 or -1");
    return kErrArg;
  }
  if (bad_dim(ndim, nw, "stencil_galerkin")) return kErrArg;
  Grid3 g = mk_grid(ndim, nw);
  if (g.n0 % 2 == 0 || g.n1 % 2 == 0 || (ndim == 3 && g.n2 % 2 == 0)) {
    set_error("stencil_galerkin: node grids need odd sizes");
    return kErrArg;
  }
  if (nout < 1 || nout > 65535) {
    set_error("stencil_galerkin: bad number of output entries %d", nout);
    return kErrArg;
  }
  Grid3 gh{g.n0 / 2 + 1, g.n1 / 2 + 1, ndim == 3 ? g.n2 / 2 + 1 : 1};
  hipStream_t s = static_cast<hipStream_t>(stream);
  dim3 grid(nblk(gh.size()), nout);
  if (dtype == kF64)
    hipLaunchKernelGGL(k_galerkin<double>, grid, dim3(256), 0, s, oent, terms,
                       tstart, nout, rw, wtab, static_cast<const double*>(fc),
                       static_cast<double*>(cc), g, gh, ndim, sigma, rmode);
  else if (dtype == kF32)
    hipLaunchKernelGGL(k_galerkin<float>, grid, dim3(256), 0, s, oent, terms,
                       tstart, nout, rw, wtab, static_cast<const float*>(fc),
                       static_cast<float*>(cc), g, gh, ndim, sigma, rmode);
  else {
    set_error("stencil_galerkin: bad dtype %d", dtype);
    return kErrArg;
  }
  return check_launch("stencil_galerkin");
}

static int mg_restrict(int dtype, int ndim, const int* nw, int nvar,
                       const double* rw, int rmode, const void* const* fine,
                       void* const* coarse, const void* const* d,
                       void* const* u, double omega, hipStream_t s,
                       const char* what) {
  if (bad_dim(ndim, nw, what) || bad_var(nvar, what)) return kErrArg;
  Grid3 g = mk_grid(ndim, nw);
  Grid3 gh{g.n0 / 2 + 1, g.n1 / 2 + 1, ndim == 3 ? g.n2 / 2 + 1 : 1};
  if (dtype != kF64 && dtype != kF32) {
    set_error("%s: bad dtype %d", what, dtype);
    return kErrArg;
  }
#define ODIL_RESTRICT(T, NV, D)                                             \
  hipLaunchKernelGGL((k_restrict<T, NV, D>), dim3(nblk(gh.size())),         \
                     dim3(256), 0, s, cvp<T>(fine, nvar),                   \
                     vp<T>(coarse, nvar), rw, g, gh, rmode,                 \
                     d ? cvp<T>(d, nvar) : CVecPtrs<T>{},                   \
                     u ? vp<T>(u, nvar) : VecPtrs<T>{}, (T)omega)
#define ODIL_RESTRICT_NV(T, D)        \
  switch (nvar) {                     \
    case 1: ODIL_RESTRICT(T, 1, D); break; \
    case 2: ODIL_RESTRICT(T, 2, D); break; \
    case 3: ODIL_RESTRICT(T, 3, D); break; \
    default: ODIL_RESTRICT(T, 4, D); break; \
  }
  if (dtype == kF64) {
    if (ndim == 3) ODIL_RESTRICT_NV(double, 3) else ODIL_RESTRICT_NV(double, 2)
  } else {
    if (ndim == 3) ODIL_RESTRICT_NV(float, 3) else ODIL_RESTRICT_NV(float, 2)
  }
#undef ODIL_RESTRICT_NV
#undef ODIL_RESTRICT
  return check_launch(what);
}

int odil_mg_restrict(int dtype, int ndim, const int* nw, int nvar,
                     const double* rw, int rmode, const void* const* fine,
                     void* const* coarse, void* stream) {
  return mg_restrict(dtype, ndim, nw, nvar, rw, rmode, fine, coarse, nullptr,
                     nullptr, 0.0, static_cast<hipStream_t>(stream),
                     "mg_restrict");
}

// Restriction followed by the coarse level's first weighted-Jacobi sweep
// from a zero iterate: coarse = R fine, u_k = omega coarse_k / d_k.
int odil_mg_restrict_smooth(int dtype, int ndim, const int* nw, int nvar,
                            const double* rw, int rmode,
                            const void* const* fine, void* const* coarse,
                            const void* const* d, void* const* u,
                            double omega, void* stream) {
  if (!d || !u) {
    set_error("mg_restrict_smooth: missing diagonal or output");
    return kErrArg;
  }
  return mg_restrict(dtype, ndim, nw, nvar, rw, rmode, fine, coarse, d, u,
                     omega, static_cast<hipStream_t>(stream),
                     "mg_restrict_smooth");
}

int odil_mg_prolong_add(int dtype, int ndim, const int* nw, int nvar,
                        const void* const* coarse, void* const* fine,
                        void* stream) {
  if (bad_dim(ndim, nw, "mg_prolong_add") || bad_var(nvar, "mg_prolong_add"))
    return kErrArg;
  Grid3 g = mk_grid(ndim, nw);
  if (g.n0 % 2 == 0 || g.n1 % 2 == 0 || (ndim == 3 && g.n2 % 2 == 0)) {
    set_error("mg_prolong_add: node grids need odd sizes");
    return kErrArg;
  }
  Grid3 gh{g.n0 / 2 + 1, g.n1 / 2 + 1, ndim == 3 ? g.n2 / 2 + 1 : 1};
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype != kF64 && dtype != kF32) {
    set_error("mg_prolong_add: bad dtype %d", dtype);
    return kErrArg;
  }
#define ODIL_PROLONG(T, NV, D)                                               \
  hipLaunchKernelGGL((k_prolong_add<T, NV, D>), dim3(nblk(g.size())),        \
                     dim3(256), 0, s, cvp<T>(coarse, nvar), vp<T>(fine, nvar), \
                     g, gh)
#define ODIL_PROLONG_NV(T, D)        \
  switch (nvar) {                    \
    case 1: ODIL_PROLONG(T, 1, D); break; \
    case 2: ODIL_PROLONG(T, 2, D); break; \
    case 3: ODIL_PROLONG(T, 3, D); break; \
    default: ODIL_PROLONG(T, 4, D); break; \
  }
  if (dtype == kF64) {
    if (ndim == 3) ODIL_PROLONG_NV(double, 3) else ODIL_PROLONG_NV(double, 2)
  } else {
    if (ndim == 3) ODIL_PROLONG_NV(float, 3) else ODIL_PROLONG_NV(float, 2)
  }
#undef ODIL_PROLONG_NV
#undef ODIL_PROLONG
  return check_launch("mg_prolong_add");
}

// Cell-centred interpolation (MultigridDecomp.interp_cell_field) of one
// field from the coarse grid nwc to 2 nwc per axis: fine = I coarse, or
// fine += I coarse with accumulate != 0.
int odil_interp_cell(int dtype, int ndim, const int* nwc, const void* coarse,
                     void* fine, int accumulate, void* stream) {
  if (ndim < 1 || ndim > 3 || !nwc || !coarse || !fine) {
    set_error("interp_cell: bad arguments");
    return kErrArg;
  }
  Grid3 c{nwc[0], ndim > 1 ? nwc[1] : 1, ndim > 2 ? nwc[2] : 1};
  if (c.n0 < 1 || c.n1 < 1 || c.n2 < 1 || (double)c.size() * (1 << ndim) > 2.0e9) {
    set_error("interp_cell: bad grid");
    return kErrArg;
  }
  Grid3 f{2 * c.n0, ndim > 1 ? 2 * c.n1 : 1, ndim > 2 ? 2 * c.n2 : 1};
  hipStream_t s = static_cast<hipStream_t>(stream);
#define ODIL_IC(T, D)                                                        \
  hipLaunchKernelGGL((k_interp_cell<T, D>), dim3(nblk(f.size())), dim3(256), \
                     0, s, static_cast<const T*>(coarse),                    \
                     static_cast<T*>(fine), c, f, accumulate)
#define ODIL_IC_D(T)  \
  if (ndim == 1) ODIL_IC(T, 1); else if (ndim == 2) ODIL_IC(T, 2); else ODIL_IC(T, 3)
  if (dtype == kF64) {
    ODIL_IC_D(double);
  } else if (dtype == kF32) {
    ODIL_IC_D(float);
  } else {
    set_error("interp_cell: bad dtype %d", dtype);
    return kErrArg;
  }
#undef ODIL_IC_D
#undef ODIL_IC
  return check_launch("interp_cell");
}

// Adjoint of odil_interp_cell: coarse = I^T fine (fine grid 2 nwc).
int odil_interp_cell_t(int dtype, int ndim, const int* nwc, const void* fine,
                       void* coarse, void* stream) {
  if (ndim < 1 || ndim > 3 || !nwc || !coarse || !fine) {
    set_error("interp_cell_t: bad arguments");
    return kErrArg;
  }
  Grid3 c{nwc[0], ndim > 1 ? nwc[1] : 1, ndim > 2 ? nwc[2] : 1};
  if (c.n0 < 1 || c.n1 < 1 || c.n2 < 1 || (double)c.size() * (1 << ndim) > 2.0e9) {
    set_error("interp_cell_t: bad grid");
    return kErrArg;
  }
  Grid3 f{2 * c.n0, ndim > 1 ? 2 * c.n1 : 1, ndim > 2 ? 2 * c.n2 : 1};
  hipStream_t s = static_cast<hipStream_t>(stream);
#define ODIL_ICT(T, D)                                                        \
  hipLaunchKernelGGL((k_interp_cell_t<T, D>), dim3(nblk(c.size())),           \
                     dim3(256), 0, s, static_cast<const T*>(fine),            \
                     static_cast<T*>(coarse), c, f)
#define ODIL_ICT_D(T) \
  if (ndim == 1) ODIL_ICT(T, 1); else if (ndim == 2) ODIL_ICT(T, 2); else ODIL_ICT(T, 3)
  if (dtype == kF64) {
    ODIL_ICT_D(double);
  } else if (dtype == kF32) {
    ODIL_ICT_D(float);
  } else {
    set_error("interp_cell_t: bad dtype %d", dtype);
    return kErrArg;
  }
#undef ODIL_ICT_D
#undef ODIL_ICT
  return check_launch("interp_cell_t");
}

// Dense FP64 matrix (zero-initialised by the caller) of a square stride-1
// block stencil operator; sym != 0 when `ent` is the upper half of a
// symmetric operator.
int odil_stencil_to_dense(int dtype, int ndim, const int* nw, int nvar,
                          const int* ent, int nent, const void* coef, int sym,
                          double* M, void* stream) {
  if (bad_dim(ndim, nw, "stencil_to_dense") || bad_var(nvar, "stencil_to_dense"))
    return kErrArg;
  Grid3 g = mk_grid(ndim, nw);
  hipStream_t s = static_cast<hipStream_t>(stream);
  dim3 grid(nblk(g.size() * nent));
  if (dtype == kF64)
    hipLaunchKernelGGL(k_to_dense<double>, grid, dim3(256), 0, s, ent, nent,
                       nvar, static_cast<const double*>(coef), M, g, sym);
  else if (dtype == kF32)
    hipLaunchKernelGGL(k_to_dense<float>, grid, dim3(256), 0, s, ent, nent,
                       nvar, static_cast<const float*>(coef), M, g, sym);
  else {
    set_error("stencil_to_dense: bad dtype %d", dtype);
    return kErrArg;
  }
  return check_launch("stencil_to_dense");
}

// CSR (int32 rowptr [nout N + 1], col / val [nent N]) of a stride-1 block
// stencil operator in the reference linearize layout; `ent` must be sorted
// by output field with `ostart` delimiting the fields.
int odil_stencil_to_csr(int dtype, int ndim, const int* nw, int nout,
                        const int* ent, const int* ostart, int nent,
                        const void* coef, int* rowptr, int* col, void* val,
                        void* stream) {
  if (bad_dim(ndim, nw, "stencil_to_csr")) return kErrArg;
  if (nout < 1) {
    set_error("stencil_to_csr: nout must be positive");
    return kErrArg;
  }
  Grid3 g = mk_grid(ndim, nw);
  if ((double)g.size() * (nent > nout ? nent : nout) >= 2147483647.0) {
    set_error("stencil_to_csr: more than 2^31 - 1 non-zeros");
    return kErrCapacity;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  dim3 grid(nblk(g.size() * (size_t)nout));
  if (dtype == kF64)
    hipLaunchKernelGGL(k_to_csr<double>, grid, dim3(256), 0, s, ent, ostart,
                       nout, static_cast<const double*>(coef), rowptr, col,
                       static_cast<double*>(val), g);
  else if (dtype == kF32)
    hipLaunchKernelGGL(k_to_csr<float>, grid, dim3(256), 0, s, ent, ostart,
                       nout, static_cast<const float*>(coef), rowptr, col,
                       static_cast<float*>(val), g);
  else {
    set_error("stencil_to_csr: bad dtype %d", dtype);
    return kErrArg;
  }
  return check_launch("stencil_to_csr");
}

// y = M x with M dense FP64 (nvar*N square); x, y split into nvar fields of
// N values each, in dtype_x / dtype_y.
int odil_dense_gemv(int dtype_x, int dtype_y, int nvar, int N,
                    const double* M, const void* const* x, void* const* y,
                    void* stream) {
  if (bad_var(nvar, "dense_gemv")) return kErrArg;
  if (!M || !x || !y) {
    set_error("dense_gemv: null matrix or field array");
    return kErrArg;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  dim3 grid(ceil_div(nvar * N, 4));
  const size_t esz = dtype_x == kF64 ? 8 : 4;
  bool contig = true;
  for (int k = 1; k < nvar; ++k)
    contig = contig && static_cast<const char*>(x[k]) ==
                           static_cast<const char*>(x[0]) + k * (size_t)N * esz;
#define ODIL_GEMV(TX, TY, C)                                                  \
  hipLaunchKernelGGL((k_gemv<TX, TY, C>), grid, dim3(256), 0, s, M,          \
                     cvp<TX>(x, nvar), vp<TY>(y, nvar), nvar, N)
  if (dtype_x == kF64 && dtype_y == kF64) {
    if (contig) ODIL_GEMV(double, double, true);
    else ODIL_GEMV(double, double, false);
  } else if (dtype_x == kF32 && dtype_y == kF32) {
    if (contig) ODIL_GEMV(float, float, true);
    else ODIL_GEMV(float, float, false);
  } else {
    set_error("dense_gemv: unsupported dtypes");
    return kErrArg;
  }
#undef ODIL_GEMV
  return check_launch("dense_gemv");
}

// Factor the per-cell nvar x nvar shift-0 blocks of a stride-1 operator
// (block-diagonal Jacobi). `dq` (host, nvar * nvar) gives the entry of block
// (o, i) at shift 0 among the nent coefficient arrays of `coef`, -1 where
// the block has none. lu [nvar^2][N] in dtype,
// piv int32 [nvar][N]; *singular (device int32, cleared by the caller) is
// set to 1 when a block has an exactly zero pivot.
int odil_block_diag_factor(int dtype, int nvar, int N, const int* dq,
                           int nent, const void* coef, void* lu, int* piv,
                           int* singular, void* stream) {
  if (bad_var(nvar, "block_diag_factor")) return kErrArg;
  if (N < 1 || !dq || !coef || !lu || !piv || !singular) {
    set_error("block_diag_factor: empty grid or null argument");
    return kErrArg;
  }
  BlockIdx bi;
  for (int k = 0; k < kMaxVar * kMaxVar; ++k) bi.q[k] = -1;
  for (int o = 0; o < nvar; ++o)
    for (int i = 0; i < nvar; ++i) {
      const int q = dq[o * nvar + i];
      if (q < -1 || q >= nent) {
        set_error("block_diag_factor: block (%d, %d) entry %d outside "
                  "[-1, %d)", o, i, q, nent);
        return kErrArg;
      }
      bi.q[o * kMaxVar + i] = q;
    }
  hipStream_t s = static_cast<hipStream_t>(stream);
#define ODIL_BF(T, NV)                                                       \
  hipLaunchKernelGGL((k_block_factor<T, NV>), dim3(nblk(N)), dim3(256), 0, s, \
                     bi, static_cast<const T*>(coef), static_cast<T*>(lu),   \
                     piv, singular, (size_t)N)
#define ODIL_BF_NV(T)           \
  switch (nvar) {               \
    case 1: ODIL_BF(T, 1); break; \
    case 2: ODIL_BF(T, 2); break; \
    case 3: ODIL_BF(T, 3); break; \
    default: ODIL_BF(T, 4); break; \
  }
  if (dtype == kF64) {
    ODIL_BF_NV(double);
  } else if (dtype == kF32) {
    ODIL_BF_NV(float);
  } else {
    set_error("block_diag_factor: bad dtype %d", dtype);
    return kErrArg;
  }
#undef ODIL_BF_NV
#undef ODIL_BF
  return check_launch("block_diag_factor");
}

// u += omega D^-1 r per cell with the factors of odil_block_diag_factor
// (r = f - A u computed beforehand, e.g. by odil_stencil_apply mode 1).
int odil_block_jacobi(int dtype, int nvar, int N, const void* lu,
                      const int* piv, const void* const* r, void* const* u,
                      double omega, void* stream) {
  if (bad_var(nvar, "block_jacobi")) return kErrArg;
  if (N < 1 || !lu || !piv || !r || !u) {
    set_error("block_jacobi: empty grid or null argument");
    return kErrArg;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
#define ODIL_BJ(T, NV)                                                        \
  hipLaunchKernelGGL((k_block_jacobi<T, NV>), dim3(nblk(N)), dim3(256), 0, s, \
                     static_cast<const T*>(lu), piv, cvp<T>(r, nvar),         \
                     vp<T>(u, nvar), (size_t)N, (T)omega)
#define ODIL_BJ_NV(T)           \
  switch (nvar) {               \
    case 1: ODIL_BJ(T, 1); break; \
    case 2: ODIL_BJ(T, 2); break; \
    case 3: ODIL_BJ(T, 3); break; \
    default: ODIL_BJ(T, 4); break; \
  }
  if (dtype == kF64) {
    ODIL_BJ_NV(double);
  } else if (dtype == kF32) {
    ODIL_BJ_NV(float);
  } else {
    set_error("block_jacobi: bad dtype %d", dtype);
    return kErrArg;
  }
#undef ODIL_BJ_NV
#undef ODIL_BJ
  return check_launch("block_jacobi");
}

// One forward Gauss-Seidel sweep of a square stride-1 operator in full
// (non-symmetric) storage, u updated in place, x (nvar fields of N values)
// working storage. `ent` is the device entry table and `ent_host` the same
// nent x 5 table in host memory (sorted by output field), from which the
// wave levels are derived.
int odil_stencil_gauss_seidel(int dtype, int ndim, const int* nw, int nvar,
                              const int* ent, const int* ent_host, int nent,
                              const void* coef, void* const* u,
                              const void* const* f, void* const* x,
                              double omega, void* stream) {
  if (bad_dim(ndim, nw, "gauss_seidel") || bad_var(nvar, "gauss_seidel"))
    return kErrArg;
  if (!ent || !ent_host || nent < 1 || !coef || !u || !f || !x) {
    set_error("gauss_seidel: null argument or empty operator");
    return kErrArg;
  }
  int ostart[kMaxVar + 1] = {0};
  int qd[kMaxVar];
  for (int o = 0; o < nvar; ++o) qd[o] = -1;
  int a1 = 1, a0 = 1;
  for (int e = 0; e < nent; ++e) {
    const int* it = ent_host + 5 * (size_t)e;
    const int o = it[0], in = it[1];
    if (o < 0 || o >= nvar || in < 0 || in >= nvar ||
        (e > 0 && o < ent_host[5 * (size_t)(e - 1)])) {
      set_error("gauss_seidel: entry %d: fields (%d, %d) out of range or "
                "entries not sorted by output field", e, o, in);
      return kErrArg;
    }
    if (ndim == 2 && it[4] != 0) {
      set_error("gauss_seidel: entry %d has a third shift on a 2D grid", e);
      return kErrArg;
    }
    ostart[o + 1] = e + 1;
    const int s0 = it[2], s1 = it[3], s2 = it[4];
    if (o == in && s0 == 0 && s1 == 0 && s2 == 0) qd[o] = e;
    // Level weights: a1 |s1| >= 1 + s2 for s = (0, s1 < 0, s2).
    if (o == in && s0 == 0 && s1 < 0)
      a1 = std::max(a1, ceil_div(1 + s2, -s1));
  }
  for (int e = 0; e < nent; ++e) {
    const int* it = ent_host + 5 * (size_t)e;
    const int s0 = it[2], s1 = it[3], s2 = it[4];
    // a0 |s0| >= 1 + a1 s1 + s2 for s = (s0 < 0, s1, s2).
    if (it[0] == it[1] && s0 < 0) {
      const int need = 1 + a1 * s1 + s2;
      if (need > 0) a0 = std::max(a0, ceil_div(need, -s0));
    }
  }
  for (int o = 0; o < nvar; ++o) {
    if (qd[o] < 0) {
      set_error("gauss_seidel: field %d has no diagonal entry", o);
      return kErrArg;
    }
    if (ostart[o + 1] < ostart[o]) ostart[o + 1] = ostart[o];
  }
  Grid3 g = mk_grid(ndim, nw);
  const size_t N = g.size();
  const int span = a1 * (g.n1 - 1) + (g.n2 - 1);
  const int nwave = a0 * (g.n0 - 1) + span + 1;
  const int row = g.n2 == 1 ? 1 : g.n1;
  hipStream_t s = static_cast<hipStream_t>(stream);
#define ODIL_GS(T)                                                            \
  for (int o = 0; o < nvar; ++o)                                              \
    for (int w = 0; w < nwave; ++w) {                                         \
      const int lo = std::max(0, ceil_div(std::max(w - span, 0), a0));        \
      const int hi = std::min(g.n0 - 1, w / a0);                              \
      if (hi < lo) continue;                                                  \
      const int cnt = (hi - lo + 1) * row;                                    \
      hipLaunchKernelGGL((k_gs_wave<T>), dim3(nblk(cnt)), dim3(256), 0, s,    \
                         ent, ostart[o], ostart[o + 1], qd[o],                \
                         static_cast<const T*>(coef),                         \
                         cvp<T>((const void* const*)u, nvar), vp<T>(x, nvar), \
                         static_cast<const T*>(f[o]), g, o, w, a0, a1, lo,    \
                         cnt);                                                \
    }                                                                         \
  hipLaunchKernelGGL((k_relax<T>), dim3(nblk(N), nvar), dim3(256), 0, s,      \
                     cvp<T>((const void* const*)x, nvar), vp<T>(u, nvar), N,  \
                     (T)omega)
  if (dtype == kF64) {
    ODIL_GS(double);
  } else if (dtype == kF32) {
    ODIL_GS(float);
  } else {
    set_error("gauss_seidel: bad dtype %d", dtype);
    return kErrArg;
  }
#undef ODIL_GS
  return check_launch("gauss_seidel");
}

}  // extern "C"
