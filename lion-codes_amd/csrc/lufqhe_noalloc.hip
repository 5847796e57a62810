// lufqhe_noalloc.hip -- lufqhe_det_inplace_static: the reference's in-place
// determinant call (lufqhe_det_inplace) with its workspace -- one mantissa
// and one exponent per matrix, 20 bytes -- in a fixed device array instead
// of an allocation, for callers that cannot allocate (code under HIP graph
// capture, serving loops that forbid hipMalloc).
//
// The array `scratch` is a file-scope __device__ object: every GPU the
// library runs on holds its own copy until the process ends.  Host code
// never uses its name as an address (on the host it is only a stand-in);
// the device address is looked up with hipGetSymbolAddress once per device,
// on the first call with that device current, and kept per device.  Calls
// on one GPU share that GPU's array, so calls that overlap on one GPU (on
// different streams) race; calls ordered on one stream do not.  Every
// matrix's entry is written before it is read, so nothing needs clearing.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/lufqhe.h"

namespace lufqhe {
int32_t hip_fail(hipError_t e);  // lufqhe_api.hip
}

constexpr int64_t kStaticMaxCount = LUFQHE_STATIC_MAX_MATRICES;
constexpr int64_t kStaticBytes = kStaticMaxCount * (16 + 4);  // mant + exp per matrix

static __device__ __align__(256) unsigned char scratch[kStaticBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];  // device address of `scratch`, per device

// The current device's address of `scratch`, looked up on its first use.
hipError_t scratch_address(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

}  // namespace

extern "C" int32_t lufqhe_det_inplace_static(void* d_A, int64_t count, int32_t n, int64_t ld,
                                             int64_t stride, int32_t pair_product,
                                             int32_t policy, hipStream_t stream) {
  if (count < 0 || n < 1 || n > 64 || ld < n || stride < 0) return LUFQHE_ERROR_INVALID_VALUE;
  if (policy != LUFQHE_POLICY_CROUT_REF && policy != LUFQHE_POLICY_ABS)
    return LUFQHE_ERROR_UNSUPPORTED;
  if (count > kStaticMaxCount) return LUFQHE_ERROR_TOO_LARGE;
  if (count == 0) return LUFQHE_SUCCESS;
  if (!d_A) return LUFQHE_ERROR_INVALID_VALUE;
  void* work = nullptr;
  const hipError_t e = scratch_address(&work);
  if (e != hipSuccess) return lufqhe::hip_fail(e);
  // with a workspace given, the in-place call allocates nothing
  return lufqhe_det_inplace(d_A, count, n, ld, stride, pair_product, policy, work, stream);
}
