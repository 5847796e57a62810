// lup_persist_inst.hip -- one persistent class kernel (lup_persistent.hpp)
// for one fused variant.  The build compiles this file once per (variant,
// class), so the class kernels -- each holding the fully unrolled code of
// up to 16 sizes -- compile in parallel; each object registers its kernel in
// g_persist at load time.
#include "lup_dispatch.hpp"
#include "lup_persistent.hpp"

#if !defined(LUFQHE_VARIANT) || !defined(LUFQHE_CLASS)
#error "compile with -DLUFQHE_VARIANT=v (0 or 1) -DLUFQHE_CLASS=c"
#endif

namespace lufqhe {
namespace {

constexpr int kPolicy = variant_policy(LUFQHE_VARIANT);
static_assert(variant_arith(LUFQHE_VARIANT) == ARITH_FUSED, "persistent kernels are fused only");

hipError_t launch(const PersistArgs& pa, int64_t blocks, hipStream_t stream) {
  hipLaunchKernelGGL((lup_persistent_kernel<LUFQHE_CLASS, kPolicy>), dim3((unsigned)blocks),
                     dim3(64), 0, stream, pa);
  return hipGetLastError();
}

struct Registrar {
  Registrar() {
    PersistEntry& e = g_persist[LUFQHE_VARIANT][LUFQHE_CLASS];
    e.launch = &launch;
    e.kernel = reinterpret_cast<const void*>(&lup_persistent_kernel<LUFQHE_CLASS, kPolicy>);
  }
};
Registrar registrar;

}  // namespace
}  // namespace lufqhe
