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
// bitwise reproducible for a given launch shape.
__global__ void __launch_bounds__(256)
    k_reduce_partials(const double* __restrict__ partial, int nb,
                      double* __restrict__ out) {
  __shared__ double red[4];
  const int e = blockIdx.x;
  double s = 0;
  for (int b = threadIdx.x; b < nb; b += 256) s += partial[(size_t)e * nb + b];
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

int odil_abi_version(void) { return 5; }

}  // extern "C"
