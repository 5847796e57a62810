// lup_dispatch.hpp -- launch tables filled by the per-shard instantiation
// units (lup_inst.hip compiled once per kernel variant and shard).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lup_kernel.hpp"
#include "lup_persistent.hpp"

namespace lufqhe {

// Kernel variants compiled into the library.
//   0: CROUT_REF pivots, fused FP64 arithmetic (production default)
//   1: ABS pivots, fused FP64 arithmetic
//   2: CROUT_REF pivots, the reference's own formulas (bit-exact parity mode)
constexpr int kNumVariants = 3;
constexpr int variant_policy(int v) { return v == 1 ? POLICY_ABS : POLICY_CROUT_REF; }
constexpr int variant_arith(int v) { return v == 2 ? ARITH_REF : ARITH_FUSED; }
inline int variant_of(int policy, int arith) {
  if (policy == POLICY_CROUT_REF && arith == ARITH_FUSED) return 0;
  if (policy == POLICY_ABS && arith == ARITH_FUSED) return 1;
  if (policy == POLICY_CROUT_REF && arith == ARITH_REF) return 2;
  return -1;
}

constexpr int kMaxN = 64;

// Launch shape of one (variant, n) kernel: `lanes` lanes per matrix split into
// `lane_cols` lane-columns, WPB waves per workgroup.  The kernel is compiled
// for exactly one shape.
struct KernelShape {
  int lanes;
  int lane_cols;
  int waves_per_block;
  int min_waves;  // waves per SIMD the register allocation allows at least
};

// Launches the (variant, n) kernel on `blocks` workgroups.
using BatchLaunch = hipError_t (*)(const BatchArgs&, int64_t blocks,
                                   hipStream_t stream);

struct KernelEntry {
  BatchLaunch launch;
  KernelShape shape;
  const void* kernel;  // for occupancy queries
};

// g_kernels[variant][n][shape]: the fused variants carry every candidate
// shape of candidate_shape(); the reference-arithmetic variant only shape 0.
extern KernelEntry g_kernels[kNumVariants][kMaxN + 1][kMaxShapes];

// Shape chosen per (variant, n); 0 until the autotuner's table is applied.
int selected_shape(int variant, int n);
void set_selected_shape(int variant, int n, int shape);

// Persistent mixed-size kernels (lup_persistent.hpp), one per fused variant
// (0: CROUT_REF, 1: ABS) and size class, filled by lup_persist_inst.hip.
using PersistLaunch = hipError_t (*)(const PersistArgs&, int64_t blocks, hipStream_t stream);
struct PersistEntry {
  PersistLaunch launch;
  const void* kernel;  // for occupancy queries
};
constexpr int kPersistVariants = 2;
extern PersistEntry g_persist[kPersistVariants][kPersistClasses];

}  // namespace lufqhe
