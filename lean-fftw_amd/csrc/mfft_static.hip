// mfft_execute_static: plan execution with the library's static device scratch as the
// intermediate of two-stage plans, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid allocation). The call allocates
// nothing: the scratch is a file-scope __device__ array, one copy per GPU, held until the
// process ends.
//
// In host code the array's name is only a host-side stand-in; its device address comes from
// hipGetSymbolAddress, looked up once per device (with that device current) on the entry
// point's first call there, and only that address is ever passed to kernels.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/mfft.h"
#include "mfft_internal.h"

namespace {

constexpr int64_t kScratchElems = (int64_t)MFFT_STATIC_SCRATCH_BYTES / (int64_t)sizeof(double2);

static __device__ __align__(256) double2 scratch[kScratchElems];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double2 *g_ptr[kMaxDevices] = {};

// The device address of `scratch` on `dev`, looked up on first use (nullptr on failure).
double2 *scratch_on(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return nullptr;
    if (cur != dev && hipSetDevice(dev) != hipSuccess) return nullptr;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) return nullptr;
    g_ptr[dev] = static_cast<double2 *>(p);
  }
  return g_ptr[dev];
}

}  // namespace

extern "C" int mfft_execute_static(const mfft_plan *p, const void *in, void *out, void *stream) {
  if (!p || !in || !out) return MFFT_ERR_INVALID;
  const int64_t need = mfft::workspace_elems(p);
  if (need > kScratchElems) return MFFT_ERR_WORKSPACE;  // one transform does not fit
  int dev = 0;
  if (mfft::plan_device(p, &dev) != MFFT_SUCCESS) return MFFT_ERR_INVALID;
  double2 *ptr = scratch_on(dev);
  if (!ptr) return MFFT_ERR_HIP;
  const mfft::Workspace ws{ptr, kScratchElems};
  return mfft::execute_in(p, in, out, &ws, (hipStream_t)stream);
}
