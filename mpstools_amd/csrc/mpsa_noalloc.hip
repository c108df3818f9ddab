// mpsa_noalloc.hip -- the dot product of mpsa_dot for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc): the partial sums live in a file-scope
// device array instead of a buffer the library allocates, and the result is
// written to a device pointer the caller passes, with no host synchronisation.
//
// The array holds the fixed partition of mpsa_dot (1024 partial pairs,
// 16 KiB) whatever the vector length, so every n is accepted and the result is
// bitwise the one mpsa_dot returns.  Each GPU holds its own copy until the
// process ends; its device address is looked up once per device (with that
// device current) and kept.  Every call on one GPU shares that GPU's array:
// calls that overlap on one GPU (two streams at once) race on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/mpstools_amd.h"
#include "mpsa_internal.h"

namespace {

constexpr int kParts = 1024;           // partial pairs (re, im) per call, as mpsa_dot
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[2 * kParts];

std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};   // device address of `scratch` per device

// Address of this device's scratch array, looked up on the first call on it.
hipError_t scratch_ptr(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[dev] = (double*)p;
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

}  // namespace

extern "C" int mpsa_dot_noalloc(const double* dx, const double* dy, int64_t n, int is_complex, double* dout,
                                mpsa_stream stream) {
  if (n < 0 || !dout || (n > 0 && (!dx || !dy)) || (is_complex != 0 && is_complex != 1)) {
    mpsa::set_last_error("dot_noalloc: bad arguments");
    return MPSA_EINVAL;
  }
  double* part = nullptr;
  hipError_t e = scratch_ptr(&part);
  if (e == hipSuccess) e = mpsa::launch_dot(dx, dy, n, is_complex, part, kParts, dout, (hipStream_t)stream);
  if (e != hipSuccess) {
    mpsa::set_last_error(hipGetErrorString(e));
    return MPSA_EHIP;
  }
  return MPSA_OK;
}
