// Batched Poisson residual + loss + adjoint over many independent instances
// of uneven size in one persistent launch.
//
// Instance i owns the cells [off[i], off[i] + n0[i] n1[i]) of the packed
// u / src / grad buffers (C order within the instance) and spacings
// dx[i], dy[i]. The host lists every TX x TY tile of every instance in
// `tiles` (int32 [ntiles][4] = (instance, x0, y0, 0), grouped by instance,
// tstart[i] its first tile). Each workgroup takes one contiguous chunk of
// tiles (XCD-aware, so neighbouring tiles share an L2) and runs the fused
// tile adjoint of the single-grid kernel on each; per-tile sums of f^2 then
// reduce per instance in fixed order:
//   loss[i] = mean(f_i^2),  grad_i = (2 / N_i) J_i^T f_i.
#include "poisson_op.h"

namespace odil {

template <class T, int TX, int TY>
__global__ void __launch_bounds__(256)
    k_poisson_batch(const int4* __restrict__ tiles, int ntiles, int per_block,
                    const long long* __restrict__ off,
                    const int* __restrict__ n0s, const int* __restrict__ n1s,
                    const double* __restrict__ dxs,
                    const double* __restrict__ dys, const T* __restrict__ u,
                    const T* __restrict__ src, T* __restrict__ grad,
                    double* __restrict__ partial) {
  extern __shared__ unsigned char smem_raw[];
  T* ot = reinterpret_cast<T*>(smem_raw);
  __shared__ double red[4];
  const unsigned g = xcd_block(blockIdx.x, gridDim.x);
  const int t0 = (int)g * per_block;
  const int t1 = min(t0 + per_block, ntiles);
  for (int t = t0; t < t1; ++t) {
    const int4 tr = tiles[t];
    const int i = tr.x;
    const long long o = off[i];
    const int n0 = n0s[i], n1 = n1s[i];
    FieldPtrs<T, 1> in{{u + o}};
    OutPtrs<T, 1> out{{grad + o}};
    PoissonParams P{dxs[i], dys[i], src + o};
    const T scale = T(2.0 / ((double)n0 * n1));
    double ssq[1] = {0.0};
    loss_grad_tile<T, PoissonOp<T>, TX, TY>(in, out, n0, n1, tr.y, tr.z,
                                            scale, P, ot, ssq);
    const double s = block_sum(ssq[0], red);
    if (threadIdx.x == 0) partial[t] = s;
  }
}

// loss[i] = sum of partial[tstart[i] .. tstart[i+1]) / (n0 n1), one
// workgroup per instance.
__global__ void __launch_bounds__(256)
    k_batch_loss(const double* __restrict__ partial,
                 const int* __restrict__ tstart, const int* __restrict__ n0s,
                 const int* __restrict__ n1s, double* __restrict__ loss) {
  __shared__ double red[4];
  const int i = blockIdx.x;
  double s = 0;
  for (int t = tstart[i] + threadIdx.x; t < tstart[i + 1]; t += 256)
    s += partial[t];
  s = block_sum(s, red);
  if (threadIdx.x == 0) loss[i] = s / ((double)n0s[i] * n1s[i]);
}

template <class T, int TX, int TY>
int batch_launch(const int* tiles, int ntiles, int nblocks,
                 const long long* off, const int* n0, const int* n1,
                 const double* dx, const double* dy, const void* u,
                 const void* src, void* grad, double* partial,
                 hipStream_t s) {
  using Op = PoissonOp<T>;
  constexpr size_t smem = loss_grad_smem<T, Op, TX, TY>();
  auto kern = k_poisson_batch<T, TX, TY>;
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)smem);
  const int per_block = ceil_div(ntiles, nblocks);
  const int grid = ceil_div(ntiles, per_block);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(256), smem, s,
                     reinterpret_cast<const int4*>(tiles), ntiles, per_block,
                     off, n0, n1, dx, dy, static_cast<const T*>(u),
                     static_cast<const T*>(src), static_cast<T*>(grad),
                     partial);
  return check_launch("poisson_batch_loss_grad");
}

}  // namespace odil

using namespace odil;

extern "C" {

// Tile shapes available to odil_poisson_batch_loss_grad: writes up to
// `cap` (tx, ty) pairs into shapes and returns their number.
int odil_poisson_batch_shapes(int dtype, int* shapes, int cap) {
  static const int f32[][2] = {{8, 32}, {8, 64}, {16, 32}, {16, 64}, {32, 32}};
  static const int f64[][2] = {{8, 32}, {8, 64}, {16, 32}, {16, 64}};
  const int(*tab)[2] = dtype == kF32 ? f32 : f64;
  const int n = dtype == kF32 ? 5 : 4;
  for (int k = 0; shapes && k < n && k < cap; ++k) {
    shapes[2 * k] = tab[k][0];
    shapes[2 * k + 1] = tab[k][1];
  }
  return n;
}

int odil_poisson_batch_loss_grad(int dtype, int ninst, const long long* off,
                                 const int* n0, const int* n1,
                                 const double* dx, const double* dy,
                                 const int* tiles, int ntiles,
                                 const int* tstart, int tx, int ty,
                                 int nblocks, const void* u, const void* src,
                                 void* grad, double* partial, double* loss,
                                 void* stream) {
  if (ninst < 1 || ntiles < 1 || nblocks < 1 || !off || !n0 || !n1 || !dx ||
      !dy || !tiles || !tstart || !u || !src || !grad || !partial || !loss) {
    set_error("poisson_batch_loss_grad: bad arguments");
    return kErrArg;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  int st = -1;
#define ODIL_BATCH(T, X, Y)                                                 \
  if (tx == X && ty == Y)                                                   \
    st = batch_launch<T, X, Y>(tiles, ntiles, nblocks, off, n0, n1, dx, dy, \
                               u, src, grad, partial, s);
  if (dtype == kF32) {
    ODIL_BATCH(float, 8, 32)
    ODIL_BATCH(float, 8, 64)
    ODIL_BATCH(float, 16, 32)
    ODIL_BATCH(float, 16, 64)
    ODIL_BATCH(float, 32, 32)
  } else if (dtype == kF64) {
    ODIL_BATCH(double, 8, 32)
    ODIL_BATCH(double, 8, 64)
    ODIL_BATCH(double, 16, 32)
    ODIL_BATCH(double, 16, 64)
  } else {
    set_error("poisson_batch_loss_grad: bad dtype %d", dtype);
    return kErrArg;
  }
#undef ODIL_BATCH
  if (st < 0) {
    set_error("poisson_batch_loss_grad: tile %dx%d not available", tx, ty);
    return kErrArg;
  }
  if (st) return st;
  hipLaunchKernelGGL(k_batch_loss, dim3(ninst), dim3(256), 0, s, partial,
                     tstart, n0, n1, loss);
  return check_launch("poisson_batch_loss");
}

}  // extern "C"
