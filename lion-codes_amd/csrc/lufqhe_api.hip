// lufqhe_api.hip -- the C ABI (include/lufqhe.h) over the kernel tables.
#include <limits.h>

#include <atomic>

#include "../../include/lufqhe.h"
#include "lup_dispatch.hpp"

namespace lufqhe {

// Filled at load time by the lup_inst.hip and lup_persist_inst.hip objects.
KernelEntry g_kernels[kNumVariants][kMaxN + 1][kMaxShapes];
PersistEntry g_persist[kPersistVariants][kPersistClasses];

namespace {
thread_local hipError_t t_last_hip = hipSuccess;
std::atomic<int> g_selected[kNumVariants][kMaxN + 1];  // zero-initialised: shape 0
}

int selected_shape(int variant, int n) {
  return g_selected[variant][n].load(std::memory_order_relaxed);
}
void set_selected_shape(int variant, int n, int shape) {
  g_selected[variant][n].store(shape, std::memory_order_relaxed);
}

// The kernel a launch of (variant, n) uses: the selected shape if it is
// built, else shape 0.
const KernelEntry& entry_for(int variant, int n) {
  const KernelEntry& e = g_kernels[variant][n][selected_shape(variant, n)];
  return e.launch ? e : g_kernels[variant][n][0];
}

int32_t hip_fail(hipError_t e) {
  t_last_hip = e;
  return LUFQHE_ERROR_HIP;
}

hipError_t launch_inplace_write(double2* A, int64_t count, int64_t stride, const double2* mant,
                                const int32_t* expo, int pair_product, hipStream_t stream);
hipError_t launch_pair_sum_plain(const double2* A, int64_t pairs, int64_t stride,
                                 double2* partial, double2* out, hipStream_t stream);
hipError_t launch_pair_sum_scaled(const double2* mant, const int32_t* expo, int64_t pairs,
                                  double2* part_m, int32_t* part_e, double2* out_m,
                                  int32_t* out_e, hipStream_t stream);
hipError_t launch_fqhe_build(const double2* z, int64_t count, int n, int64_t z_stride,
                             const uint8_t* row_perm, double2* out, hipStream_t stream);
constexpr int64_t kPairSumParts = 1024;  // kSumMaxBlocks in lufqhe_epilogue.hip
constexpr int64_t kPairSumWork = kPairSumParts * (16 + 4);

// Device workspace given by the caller, or allocated stream-ordered.
struct Workspace {
  void* ptr = nullptr;
  bool owned = false;
  hipStream_t stream = nullptr;
  hipError_t acquire(void* given, int64_t bytes, hipStream_t s) {
    stream = s;
    if (given) {
      ptr = given;
      return hipSuccess;
    }
    owned = true;
    return hipMallocAsync(&ptr, (size_t)bytes, s);
  }
  ~Workspace() {
    if (owned && ptr) (void)hipFreeAsync(ptr, stream);
  }
};

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
    case LUFQHE_ERROR_IO: return "file cannot be read";
    case LUFQHE_ERROR_TOO_LARGE: return "batch exceeds the fixed workspace of the static entry point";
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
  const KernelEntry& e = entry_for(v, n);
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

int32_t lufqhe_fqhe_det(const void* d_z, int64_t count, int32_t n, int64_t z_stride,
                        const uint8_t* d_row_perm, void* d_mant, int32_t* d_exp,
                        double* d_logabs, double* d_phase, int32_t* d_info, int8_t* d_piv,
                        int32_t policy, int32_t arith, hipStream_t stream) {
  if (count < 0 || n < 1 || n > kMaxN || z_stride < 0 || (z_stride > 0 && z_stride < n))
    return LUFQHE_ERROR_INVALID_VALUE;
  const int v = variant_of(policy, arith);
  if (v < 0) return LUFQHE_ERROR_UNSUPPORTED;
  if (count == 0) return LUFQHE_SUCCESS;
  if (!d_z || !d_mant) return LUFQHE_ERROR_INVALID_VALUE;
  const KernelEntry& e = entry_for(v, n);
  if (!e.launch) return LUFQHE_ERROR_NOT_BUILT;
  const int64_t blocks = plain_blocks(e, count);
  if (blocks > INT_MAX) return LUFQHE_ERROR_INVALID_VALUE;

  BatchArgs args{};
  args.A = static_cast<const double2*>(d_z);
  args.count = count;
  args.ld = n;
  args.stride = z_stride;
  args.row_perm = d_row_perm;
  args.piv_stride = n;
  args.src = SRC_FQHE;
  args.out = DetOut{static_cast<double2*>(d_mant), d_exp, d_logabs, d_phase, d_info, d_piv};
  const hipError_t err = e.launch(args, blocks, stream);
  return err == hipSuccess ? LUFQHE_SUCCESS : hip_fail(err);
}

int32_t lufqhe_fqhe_build(const void* d_z, int64_t count, int32_t n, int64_t z_stride,
                          const uint8_t* d_row_perm, void* d_A, hipStream_t stream) {
  if (count < 0 || n < 1 || n > kMaxN || z_stride < 0 || (z_stride > 0 && z_stride < n))
    return LUFQHE_ERROR_INVALID_VALUE;
  if (count == 0) return LUFQHE_SUCCESS;
  if (!d_z || !d_A) return LUFQHE_ERROR_INVALID_VALUE;
  if ((count * n + 255) / 256 > INT_MAX) return LUFQHE_ERROR_INVALID_VALUE;
  const hipError_t err = launch_fqhe_build(static_cast<const double2*>(d_z), count, n, z_stride,
                                           d_row_perm, static_cast<double2*>(d_A), stream);
  return err == hipSuccess ? LUFQHE_SUCCESS : hip_fail(err);
}

int32_t lufqhe_shape_count(int32_t n) {
  if (n < 1 || n > kMaxN) return 0;
  int c = 0;
  while (c < kMaxShapes && g_kernels[0][n][c].launch) ++c;
  return c;
}

int32_t lufqhe_shape_info(int32_t n, int32_t shape, int32_t* lanes, int32_t* lane_cols,
                          int32_t* min_waves) {
  if (n < 1 || n > kMaxN || shape < 0 || shape >= kMaxShapes || !lanes || !lane_cols)
    return LUFQHE_ERROR_INVALID_VALUE;
  const KernelEntry& e = g_kernels[0][n][shape];
  if (!e.launch) return LUFQHE_ERROR_NOT_BUILT;
  *lanes = e.shape.lanes;
  *lane_cols = e.shape.lane_cols;
  if (min_waves) *min_waves = e.shape.min_waves;
  return LUFQHE_SUCCESS;
}

int32_t lufqhe_set_shape(int32_t policy, int32_t arith, int32_t n, int32_t shape) {
  const int v = variant_of(policy, arith);
  if (v < 0) return LUFQHE_ERROR_UNSUPPORTED;
  if (n < 1 || n > kMaxN || shape < 0 || shape >= kMaxShapes) return LUFQHE_ERROR_INVALID_VALUE;
  if (!g_kernels[v][n][shape].launch) return LUFQHE_ERROR_NOT_BUILT;
  set_selected_shape(v, n, shape);
  return LUFQHE_SUCCESS;
}

int32_t lufqhe_get_shape(int32_t policy, int32_t arith, int32_t n) {
  const int v = variant_of(policy, arith);
  if (v < 0 || n < 1 || n > kMaxN) return -1;
  const int sh = selected_shape(v, n);
  return g_kernels[v][n][sh].launch ? sh : 0;
}

int64_t lufqhe_det_inplace_workspace_bytes(int64_t count) {
  return count <= 0 ? 0 : count * (16 + 4);
}

int32_t lufqhe_det_inplace(void* d_A, int64_t count, int32_t n, int64_t ld, int64_t stride,
                           int32_t pair_product, int32_t policy, void* d_work,
                           hipStream_t stream) {
  if (count < 0 || n < 1 || n > kMaxN || ld < n || stride < 0) return LUFQHE_ERROR_INVALID_VALUE;
  if (variant_of(policy, ARITH_FUSED) < 0) return LUFQHE_ERROR_UNSUPPORTED;
  if (count == 0) return LUFQHE_SUCCESS;
  if (!d_A) return LUFQHE_ERROR_INVALID_VALUE;
  Workspace ws;
  hipError_t err = ws.acquire(d_work, lufqhe_det_inplace_workspace_bytes(count), stream);
  if (err != hipSuccess) return hip_fail(err);
  double2* mant = static_cast<double2*>(ws.ptr);
  int32_t* expo = reinterpret_cast<int32_t*>(mant + count);
  const int32_t st = lufqhe_det_batched(d_A, count, n, ld, stride, mant, expo, nullptr, nullptr,
                                        nullptr, nullptr, policy, ARITH_FUSED, stream);
  if (st != LUFQHE_SUCCESS) return st;
  err = launch_inplace_write(static_cast<double2*>(d_A), count, stride, mant, expo,
                             pair_product, stream);
  return err == hipSuccess ? LUFQHE_SUCCESS : hip_fail(err);
}

int64_t lufqhe_pair_sum_workspace_bytes(void) { return kPairSumWork; }

int32_t lufqhe_global_reduction(const void* d_A, int64_t count, int64_t stride, void* d_work,
                                void* d_out, hipStream_t stream) {
  if (count < 0 || stride < 0 || !d_out) return LUFQHE_ERROR_INVALID_VALUE;
  if (count >= 2 && !d_A) return LUFQHE_ERROR_INVALID_VALUE;
  Workspace ws;
  hipError_t err = ws.acquire(d_work, kPairSumWork, stream);
  if (err != hipSuccess) return hip_fail(err);
  err = launch_pair_sum_plain(static_cast<const double2*>(d_A), count / 2, stride,
                              static_cast<double2*>(ws.ptr), static_cast<double2*>(d_out), stream);
  return err == hipSuccess ? LUFQHE_SUCCESS : hip_fail(err);
}

int32_t lufqhe_pair_sum(const void* d_mant, const int32_t* d_exp, int64_t count, void* d_work,
                        void* d_out_mant, int32_t* d_out_exp, hipStream_t stream) {
  if (count < 0 || !d_out_mant || !d_out_exp) return LUFQHE_ERROR_INVALID_VALUE;
  if (count >= 2 && (!d_mant || !d_exp)) return LUFQHE_ERROR_INVALID_VALUE;
  Workspace ws;
  hipError_t err = ws.acquire(d_work, kPairSumWork, stream);
  if (err != hipSuccess) return hip_fail(err);
  double2* part_m = static_cast<double2*>(ws.ptr);
  int32_t* part_e = reinterpret_cast<int32_t*>(part_m + kPairSumParts);
  err = launch_pair_sum_scaled(static_cast<const double2*>(d_mant), d_exp, count / 2, part_m,
                               part_e, static_cast<double2*>(d_out_mant), d_out_exp, stream);
  return err == hipSuccess ? LUFQHE_SUCCESS : hip_fail(err);
}

}  // extern "C"
