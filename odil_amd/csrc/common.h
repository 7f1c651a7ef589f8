// Shared definitions for the ODIL HIP kernels (gfx950).
//
// Grids are cell-centred and stored C-order: a 2D field of shape (n0, n1) has
// element (i, j) at i*n1 + j, so the last axis is contiguous in HBM. Fields of
// one problem are separate arrays (the packed state is their concatenation,
// field-major), matching the flat layout of the reference's pack_state.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>

namespace odil {

enum Dtype : int { kF32 = 0, kF64 = 1 };

// Status codes returned across the C ABI. Nothing throws across it.
enum Status : int {
  kOk = 0,
  kErrArg = 1,      // invalid argument (shape, dtype, null pointer)
  kErrLaunch = 2,   // kernel launch failed
  kErrCapacity = 3  // size above a documented maximum
};

void set_error(const char* fmt, ...);
int check_launch(const char* what);

// Wave-level sum over 64 lanes (CDNA wavefront) using DPP-capable shuffles.
// Workgroups are dealt to the 8 XCDs of the MI355X round-robin in dispatch
// order, and each XCD has its own L2. xcd_block maps the dispatch index L of
// a grid of T workgroups to a logical index such that XCD k serves one
// contiguous range of logical blocks: neighbouring tiles (whose halos
// overlap) then share an L2 instead of fetching the halo again from HBM.
constexpr unsigned kNumXcd = 8;
__device__ inline unsigned xcd_block(unsigned L, unsigned T) {
  const unsigned xcd = L % kNumXcd, idx = L / kNumXcd;
  const unsigned q = T / kNumXcd, r = T % kNumXcd;
  return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}
// Logical (bx, by) of a 2D grid under xcd_block, bx fastest.
__device__ inline void xcd_block2(unsigned& bx, unsigned& by) {
  const unsigned g = xcd_block(blockIdx.x + blockIdx.y * gridDim.x,
                               gridDim.x * gridDim.y);
  bx = g % gridDim.x;
  by = g / gridDim.x;
}

// Lane exchange inside rows of 16 lanes by a DPP control (quad_perm,
// row_ror, ...), on 32- or 64-bit values (a 64-bit value moves as two
// 32-bit halves).
template <int CTRL>
__device__ inline float dpp_move(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL,
                                         0xf, 0xf, false));
}
template <int CTRL>
__device__ inline double dpp_move(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
  const int hi =
      __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __builtin_bit_cast(
      double, ((long long)hi << 32) | (unsigned long long)(unsigned)lo);
}
__device__ inline float read_lane(float v, int l) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ inline double read_lane(double v, int l) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(
      double, ((long long)hi << 32) | (unsigned long long)(unsigned)lo);
}

// Sum over the 64 lanes of a wavefront, returned in every lane. Every lane
// must be active. Inside each row of 16 lanes: swap neighbours
// (quad_perm [1,0,3,2]), swap pairs ([2,3,0,1]), rotate by 4 and by 8
// (DPP moves, no LDS round trip); then the four row sums are read as
// uniform values and added in a fixed order.
template <class T>
__device__ inline T wave_sum(T v) {
  v += dpp_move<0xB1>(v);
  v += dpp_move<0x4E>(v);
  v += dpp_move<0x124>(v);
  v += dpp_move<0x128>(v);
  return (read_lane(v, 0) + read_lane(v, 16)) +
         (read_lane(v, 32) + read_lane(v, 48));
}

// Block-level sum: wavefront reduction, then across waves through LDS.
// `buf` must hold one element per wave. Result valid in thread 0.
template <class T>
__device__ inline T block_sum(T v, T* buf) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) buf[wid] = v;
  __syncthreads();
  T s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < nw; ++w) s += buf[w];
  return s;
}

// Block-level sums of N values at once (one LDS round trip and two
// barriers for all of them). `buf` holds N * 4 elements (up to 4 waves);
// results valid in thread 0, written back into v.
template <int N, class T>
__device__ inline void block_sum_n(T (&v)[N], T* buf) {
#pragma unroll
  for (int e = 0; e < N; ++e) v[e] = wave_sum(v[e]);
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int e = 0; e < N; ++e) buf[e * 4 + wid] = v[e];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int e = 0; e < N; ++e) {
      T s = 0;
      for (int w = 0; w < nw; ++w) s += buf[e * 4 + w];
      v[e] = s;
    }
  }
}

// Wavefront sums of N values stored as one partial per wavefront,
// partial[e * nparts + part] (written by lane 0): no LDS round trip and no
// barrier, so the waves of a workgroup never wait for each other here.
template <int N>
__device__ inline void wave_partials(double (&v)[N], double* partial,
                                     size_t nparts, size_t part) {
#pragma unroll
  for (int e = 0; e < N; ++e) v[e] = wave_sum(v[e]);
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int e = 0; e < N; ++e) partial[(size_t)e * nparts + part] = v[e];
  }
}

__host__ __device__ inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

// Halo rule at a wall: the value outside the domain is extrapolated from the
// nearest inner value `near`, the next one `far` and the wall value `w`.
//   QUADH:  quadratic through far, near and the wall half-point,
//           (far - 6 near + 8 w) / 3                       (util_op.py:1017)
//   LINEAR: 2 near - far                                    (util_op.py:1032)
enum Extrap : int { QUADH = 0, LINEAR = 1 };

template <class T>
__host__ __device__ inline T ghost(int kind, T far, T near, T w) {
  return kind == QUADH ? (far - T(6) * near + T(8) * w) / T(3)
                       : T(2) * near - far;
}
// Derivatives of ghost() with respect to `far` and `near`.
template <class T>
__host__ __device__ inline T ghost_dfar(int kind) {
  return kind == QUADH ? T(1) / T(3) : T(-1);
}
template <class T>
__host__ __device__ inline T ghost_dnear(int kind) {
  return kind == QUADH ? T(-2) : T(2);
}

}  // namespace odil
