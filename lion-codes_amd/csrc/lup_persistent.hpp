// lup_persistent.hpp -- persistent kernels for mixed-size batches.
//
// A mixed batch is bucketed by n on the host (lufqhe_mixed.hip).  The sizes
// are grouped into classes of similar register demand; each class has one
// persistent kernel, launched with only as many one-wave workgroups as stay
// resident on the device.  Every wave loops: it takes the next task of its
// class from a device work counter (one vector atomic per task), finds the
// task's bucket -- buckets are ordered largest n first, so the large
// matrices start first and the small ones fill the tail -- and factors the
// task's matrices with that n's fully unrolled code (det_task, the same code
// the per-size kernels run, so the results are bit-identical).  A wave leaves
// the loop when the counter has passed the class's last task.
//
// A class kernel carries the register allocation of its largest case, so
// the classes are cut where the per-size kernels change occupancy.  n <= 16
// and 17..32 run two waves per SIMD.  33..48, 49..57 and 58..64 run one:
// under the 256-register cap of two waves the class kernels spill far more
// than the per-size kernels of the same n (528-952 B per lane, the task loop
// holding values across its body), and measured slower than one wave.
#pragma once

#include "lup_kernel.hpp"

namespace lufqhe {

constexpr int kPersistClasses = 5;
constexpr int kPersistLo[kPersistClasses] = {1, 17, 33, 49, 58};
constexpr int kPersistHi[kPersistClasses] = {16, 32, 48, 57, 64};
constexpr int kPersistOcc[kPersistClasses] = {2, 2, 1, 1, 1};

constexpr int persist_class(int n) {
  int c = 0;
  while (c + 1 < kPersistClasses && n > kPersistHi[c]) ++c;
  return c;
}

// The tile each size runs with inside its class kernel (lanes per matrix,
// lane-columns): the shapes the autotuner chose on MI355X for most sizes of
// each range.  Fixed at compile time; any shape gives the same bits.
struct PersistShape {
  int lanes;
  int lane_cols;
};
constexpr PersistShape persist_shape(int n) {
  if (n <= 8) return {4, 1};
  if (n <= 16) return {8, 1};
  if (n <= 24) return {16, 2};
  if (n <= 32) return {32, 2};
  if (n <= 48) return {64, 4};
  return {64, 2};
}
constexpr int persist_group(int n) { return 64 / persist_shape(n).lanes; }  // matrices per task

// LDS one task of size n uses (G groups of group_lds).
constexpr int persist_lds(int n) {
  const PersistShape s = persist_shape(n);
  return (64 / s.lanes) * group_lds(n, s.lanes, s.lane_cols);
}
constexpr int persist_lds_max(int lo, int hi) {
  int m = 0;
  for (int n = lo; n <= hi; ++n) m = persist_lds(n) > m ? persist_lds(n) : m;
  return m;
}

// One size bucket of a class queue: tasks task_start .. task_start +
// ceil(count / persist_group(n)) - 1 of the class.
struct PersistBucket {
  int32_t n;
  int32_t pad;
  int64_t task_start;
};

// A class launch: its buckets (largest n first) with, per bucket, the
// arguments a per-size launch of that bucket would take (indexed
// addressing into the plan's tables).  Both arrays live in device memory,
// so each task reads its arguments where it needs them instead of holding
// them in registers through the factorisation.
struct PersistArgs {
  const PersistBucket* buckets;
  const BatchArgs* bargs;
  int32_t nbuckets;
  int32_t pad;
  uint32_t* counter;  // this class's work counter, zero before the launch
  int64_t total_tasks;
};

template <int N, int HI, int POLICY>
__device__ __forceinline__ void persist_dispatch(int n, const BatchArgs& args, int64_t task,
                                                 double2* __restrict__ bc, int lane) {
  if (n == N) {
    constexpr PersistShape s = persist_shape(N);
    det_task<N, s.lanes, s.lane_cols, POLICY, ARITH_FUSED>(args, task, bc, lane);
    return;
  }
  if constexpr (N < HI) persist_dispatch<N + 1, HI, POLICY>(n, args, task, bc, lane);
}

template <int CLASS, int POLICY>
__global__ void __launch_bounds__(64, kPersistOcc[CLASS]) lup_persistent_kernel(PersistArgs pa) {
  static_assert(kPersistLo[CLASS] >= 1 && kPersistHi[CLASS] <= 64, "class range");
  constexpr int LO = kPersistLo[CLASS], HI = kPersistHi[CLASS];
  __shared__ double2 bcast[persist_lds_max(LO, HI)];
  for (;;) {
    // a fresh lane id each task: nothing lane-dependent (LDS addresses, row
    // and column indices of every size's tile) is hoisted out of the loop
    // and held across tasks
    int lane = (int)(threadIdx.x & 63);
    asm volatile("" : "+v"(lane));
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(pa.counter, 1u);
    t = __builtin_amdgcn_readfirstlane(t);
    if ((int64_t)t >= pa.total_tasks) break;  // every wave ends here once the queue is empty
    int b = 0;
    while (b + 1 < pa.nbuckets && pa.buckets[b + 1].task_start <= (int64_t)t) ++b;
    const int n = pa.buckets[b].n;
    const int64_t task = (int64_t)t - pa.buckets[b].task_start;
    persist_dispatch<LO, HI, POLICY>(n, pa.bargs[b], task, bcast, lane);
  }
}

}  // namespace lufqhe
