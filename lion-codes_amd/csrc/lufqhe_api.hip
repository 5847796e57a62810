// lufqhe_api.hip -- the C ABI (include/lufqhe.h) over the kernel tables.
#include <limits.h>

#include "../../include/lufqhe.h"
#include "lup_dispatch.hpp"

namespace lufqhe {

// Filled at load time by the lup_inst.hip objects.
KernelEntry g_kernels[kNumVariants][kMaxN + 1];

namespace {
thread_local hipError_t t_last_hip = hipSuccess;
}

int32_t hip_fail(hipError_t e) {
  t_last_hip = e;
  return LUFQHE_ERROR_HIP;
}

// Workgroups for one plain (non-persistent) launch of `count` matrices.
int64_t plain_blocks(const KernelEntry& e, int64_t count) {
  const int64_t per_wave = 64 / e.shape.lanes;
  const int64_t tasks = (count + per_wave - 1) / per_wave;
  return (tasks + e.shape.waves_per_block - 1) / e.shape.waves_per_block;
}

}  // namespace lufqhe

using namespace lufqhe;

extern "C" {

int32_t lufqhe_version(void) { return 100; }

const char* lufqhe_status_string(int32_t status) {
  switch (status) {
    case LUFQHE_SUCCESS: return "success";
    case LUFQHE_ERROR_INVALID_VALUE: return "invalid value";
    case LUFQHE_ERROR_UNSUPPORTED: return "unsupported combination";
    case LUFQHE_ERROR_HIP: return "HIP runtime error";
    case LUFQHE_ERROR_NOT_BUILT: return "kernel variant not built";
    case LUFQHE_ERROR_OUT_OF_BUDGET: return "memory budget exceeded";
    default: return "unknown status";
  }
}

const char* lufqhe_last_hip_error(void) {
  return t_last_hip == hipSuccess ? "" : hipGetErrorString(t_last_hip);
}

int32_t lufqhe_det_batched(const void* d_A, int64_t batch, int32_t n,
                           int64_t ld, int64_t stride, void* d_mant,
                           int32_t* d_exp, double* d_logabs, double* d_phase,
                           int32_t* d_info, int8_t* d_piv, int32_t policy,
                           int32_t arith, hipStream_t stream) {
  if (batch < 0 || n < 1 || n > kMaxN || ld < n || stride < 0)
    return LUFQHE_ERROR_INVALID_VALUE;
  const int v = variant_of(policy, arith);
  if (v < 0) return LUFQHE_ERROR_UNSUPPORTED;
  if (batch == 0) return LUFQHE_SUCCESS;
  if (!d_A || !d_mant) return LUFQHE_ERROR_INVALID_VALUE;
  const KernelEntry& e = g_kernels[v][n];
  if (!e.launch) return LUFQHE_ERROR_NOT_BUILT;
  const int64_t blocks = plain_blocks(e, batch);
  if (blocks > INT_MAX) return LUFQHE_ERROR_INVALID_VALUE;

  BatchArgs args{};
  args.A = static_cast<const double2*>(d_A);
  args.count = batch;
  args.ld = ld;
  args.stride = stride;
  args.piv_stride = n;
  args.out = DetOut{static_cast<double2*>(d_mant), d_exp, d_logabs, d_phase,
                    d_info, d_piv};
  const hipError_t err = e.launch(args, blocks, stream);
  return err == hipSuccess ? LUFQHE_SUCCESS : hip_fail(err);
}

}  // extern "C"
