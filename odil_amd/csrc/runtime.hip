// Error reporting and reductions shared by all entry points.
#include <cstdarg>
#include <mutex>

#include "stencil2d.h"

namespace odil {

namespace {
thread_local char g_err[512] = {0};
}

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: launch failed: %s", what, hipGetErrorString(e));
    return kErrLaunch;
  }
  return kOk;
}

// One block per row e of `partial`: fixed-order tree sum, so the result is
// bitwise reproducible for a given launch shape. Each thread adds partials
// b, b + 256, ... in turn, loading them 8 at a time so that the loads are
// in flight together (one load per step would wait for each in turn: tens
// of round trips for the ~17 000 partials of a 1025^2 launch).
__global__ void __launch_bounds__(256)
    k_reduce_partials(const double* __restrict__ partial, int nb,
                      double* __restrict__ out) {
  __shared__ double red[4];
  const int e = blockIdx.x;
  const double* row = partial + (size_t)e * nb;
  double s = 0;
  int b = threadIdx.x;
  for (; b + 7 * 256 < nb; b += 8 * 256) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = row[b + u * 256];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  for (; b < nb; b += 256) s += row[b];
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[e] = s;
}

}  // namespace odil

using namespace odil;

extern "C" {

// Message of the last failed call on this thread ("" if none).
const char* odil_last_error(void) { return g_err; }

// out[e] = sum_{b < nb} partial[e*nb + b] for e < nrows.
int odil_reduce_partials(const double* partial, int nrows, int nb,
                         double* out, void* stream) {
  if (!partial || !out || nrows <= 0 || nb <= 0) {
    set_error("reduce_partials: bad arguments");
    return kErrArg;
  }
  hipLaunchKernelGGL(k_reduce_partials, dim3(nrows), dim3(256), 0,
                     static_cast<hipStream_t>(stream), partial, nb, out);
  return check_launch("reduce_partials");
}

int odil_abi_version(void) { return 7; }

}  // extern "C"
