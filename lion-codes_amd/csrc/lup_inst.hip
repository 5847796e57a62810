// lup_inst.hip -- explicit kernel instantiations for one variant and one
// shard of the sizes n = 1..64 (n % LUFQHE_SHARDS == LUFQHE_SHARD).  The
// build compiles this file once per (variant, shard) so the 192 fully
// unrolled kernels compile in parallel; each object registers its kernels in
// g_kernels at load time.
#include <utility>

#include "lup_dispatch.hpp"
#include "lup_kernel.hpp"

#if !defined(LUFQHE_VARIANT) || !defined(LUFQHE_SHARD) || !defined(LUFQHE_SHARDS)
#error "compile with -DLUFQHE_VARIANT=v -DLUFQHE_SHARD=s -DLUFQHE_SHARDS=k"
#endif

namespace lufqhe {
namespace {

constexpr int kPolicy = variant_policy(LUFQHE_VARIANT);
constexpr int kArith = variant_arith(LUFQHE_VARIANT);
// One wave per workgroup: measured 3-11% faster than 2 or 4 on MI355X
// (finer-grained placement of the waves as others finish).
constexpr int kWavesPerBlock = 1;

template <int N, int SH>
hipError_t launch(const BatchArgs& args, int64_t blocks, hipStream_t stream) {
  constexpr ShapeLC sh = candidate_shape(N, SH);
  hipLaunchKernelGGL((lup_det_kernel<N, sh.lanes, sh.lane_cols, kPolicy, kArith, kWavesPerBlock, sh.waves>),
                     dim3((unsigned)blocks), dim3(64 * kWavesPerBlock), 0, stream, args);
  return hipGetLastError();
}

template <int N, int SH>
void register_shape() {
  constexpr ShapeLC sh = candidate_shape(N, SH);
  // reference arithmetic is a parity mode: only its default shape is built
  if constexpr (sh.lanes > 0 && (kArith == ARITH_FUSED || SH == 0)) {
    KernelEntry& e = g_kernels[LUFQHE_VARIANT][N][SH];
    e.launch = &launch<N, SH>;
    e.shape = KernelShape{sh.lanes, sh.lane_cols, kWavesPerBlock, sh.waves};
    e.kernel = reinterpret_cast<const void*>(
        &lup_det_kernel<N, sh.lanes, sh.lane_cols, kPolicy, kArith, kWavesPerBlock, sh.waves>);
  }
}

template <int N>
void register_one() {
  if constexpr (N % LUFQHE_SHARDS == LUFQHE_SHARD) {
    register_shape<N, 0>();
    register_shape<N, 1>();
    register_shape<N, 2>();
    register_shape<N, 3>();
  }
}

template <int... I>
void register_all(std::integer_sequence<int, I...>) {
  (register_one<I + 1>(), ...);
}

struct Registrar {
  Registrar() { register_all(std::make_integer_sequence<int, kMaxN>{}); }
};
Registrar registrar;

}  // namespace
}  // namespace lufqhe
