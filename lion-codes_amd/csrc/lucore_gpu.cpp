// lucore_gpu.cpp -- reference-shaped C++ host API over the C ABI.
#include "../../include/lucore_gpu.hpp"

#include <mutex>
#include <string>

namespace LUCoreGPU {
namespace {
thread_local int t_policy = LUFQHE_POLICY_CROUT_REF;
// The matrix stride of this thread's last CALL: globalReduction, like the
// reference's, takes no RunInfo (LU-fqhe/src/LUCoreGPU.cu:113).
thread_local long long t_stride = 0;

void check(int status, const char* what, int rank, int th_num) {
  if (status == LUFQHE_SUCCESS) return;
  std::string msg = std::string("LUCoreGPU::") + what + " on node " + std::to_string(rank) +
                    " thread " + std::to_string(th_num) + ": " + lufqhe_status_string(status);
  const char* hip = lufqhe_last_hip_error();
  if (hip && *hip) msg += std::string(" (") + hip + ")";
  throw Error(msg, status);
}
}  // namespace

void setPivotPolicy(int policy) { t_policy = policy; }

void CALL(double2** d_Matrices, int nMat, RunInfo* rInfo, int rank, int th_num,
          hipStream_t* stream) {
  if (!d_Matrices || !rInfo || !stream) check(LUFQHE_ERROR_INVALID_VALUE, "CALL", rank, th_num);
  // the tuned launch shapes, once per process ($LUFQHE_SHAPE_TABLE or the
  // table next to the library; without one the default shapes run)
  static std::once_flag table_once;
  std::call_once(table_once, [] { (void)lufqhe_load_shape_table(nullptr); });
  const int n = rInfo->getMatSide();
  const int ld = rInfo->getLoadSize();
  t_stride = (long long)n * ld;
  check(lufqhe_det_inplace(d_Matrices[th_num], nMat, n, ld, t_stride, /*pair_product=*/1,
                           t_policy, nullptr, *stream),
        "CALL", rank, th_num);
}

void globalReduction(double2** d_Matrices, double2** d_Reduction, int nMat, int rank,
                     int th_num, hipStream_t stream) {
  if (!d_Matrices || !d_Reduction)
    check(LUFQHE_ERROR_INVALID_VALUE, "globalReduction", rank, th_num);
  check(lufqhe_global_reduction(d_Matrices[th_num], nMat, t_stride, nullptr, d_Reduction[th_num],
                                stream),
        "globalReduction", rank, th_num);
}

}  // namespace LUCoreGPU
