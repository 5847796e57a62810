// lufqhe_mc.cpp -- the reference's Monte Carlo loop in LU mode on one GPU
// (MonteCarloLoop::perform, LU-fqhe/src/MonteCarloLoop.cu:121-401): the
// node's matrices are split between the GPU and the host cores by the
// reference's rule (LU-fqhe/src/MonteCarloLoop.cu:139-148), and every
// iteration runs the GPU share through the in-place determinant call and
// the global reduction (LU-fqhe/src/MonteCarloLoop.cu:299-335) while the
// host cores pair-sum their share (LUCoreCPU::CALL,
// LU-fqhe/src/MonteCarloLoop.cu:343); the iteration's wavefunction value is
// the sum of the two (LU-fqhe/src/MonteCarloLoop.cu:360-364).
//
// Unlike the reference, the GPU share is staged once in pinned host memory
// so that each iteration's copy to the device is asynchronous and overlaps
// the host cores' work, and the stream is created once, not per iteration.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>

#include "../../include/lufqhe.h"

namespace lufqhe {
int32_t hip_fail(hipError_t e);  // lufqhe_api.hip
}

namespace {

struct Cplx {
  double re, im;
};

// Device and pinned buffers of one run, released on every exit path.
struct McBuffers {
  void* d_A = nullptr;
  void* d_work = nullptr;
  void* d_sum = nullptr;
  void* h_stage = nullptr;
  Cplx* h_sum = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  ~McBuffers() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
    if (d_A) (void)hipFree(d_A);
    if (d_work) (void)hipFree(d_work);
    if (d_sum) (void)hipFree(d_sum);
    if (h_stage) (void)hipHostFree(h_stage);
    if (h_sum) (void)hipHostFree(h_sum);
  }
};

}  // namespace

extern "C" int32_t lufqhe_mc_split(int64_t nelem, int32_t cpu_count, int64_t* gpu_matrices,
                                   int64_t* cpu_matrices) {
  if (nelem < 0 || cpu_count < 0 || !gpu_matrices || !cpu_matrices)
    return LUFQHE_ERROR_INVALID_VALUE;
  // one GPU counts as 32 host cores; the GPU share is even (pair products)
  const float total = 1.0f + (float)cpu_count / 32.f;
  int64_t g = (int64_t)std::floor((float)nelem / total);
  if (g > nelem) g = nelem;
  if (g % 2 != 0) --g;
  *gpu_matrices = g;
  *cpu_matrices = nelem - g;
  return LUFQHE_SUCCESS;
}

extern "C" int32_t lufqhe_mc_lu_run(const void* h_matrices, int64_t nelem, int32_t n,
                                    int32_t cpu_count, int32_t iterations, int32_t policy,
                                    void* h_results, lufqhe_mc_stats* stats) {
  if (nelem < 0 || n < 1 || n > 64 || cpu_count < 0 || iterations < 0 || !h_results)
    return LUFQHE_ERROR_INVALID_VALUE;
  if (policy != LUFQHE_POLICY_CROUT_REF && policy != LUFQHE_POLICY_ABS)
    return LUFQHE_ERROR_UNSUPPORTED;
  if (nelem > 0 && !h_matrices) return LUFQHE_ERROR_INVALID_VALUE;
  int64_t ng = 0, nc = 0;
  (void)lufqhe_mc_split(nelem, cpu_count, &ng, &nc);
  const int64_t elems = (int64_t)n * n;
  const size_t gpu_bytes = (size_t)(ng * elems) * sizeof(Cplx);
  const Cplx* A = static_cast<const Cplx*>(h_matrices);
  Cplx* results = static_cast<Cplx*>(h_results);
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->gpu_matrices = ng;
    stats->cpu_matrices = nc;
  }

  McBuffers b;
  hipError_t err = hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking);
  if (err == hipSuccess) err = hipEventCreate(&b.ev0);
  if (err == hipSuccess) err = hipEventCreate(&b.ev1);
  if (err == hipSuccess && ng > 0) err = hipMalloc(&b.d_A, gpu_bytes);
  const int64_t ws_inplace = lufqhe_det_inplace_workspace_bytes(ng);
  const int64_t ws_sum = lufqhe_pair_sum_workspace_bytes();
  if (err == hipSuccess && ng > 0)
    err = hipMalloc(&b.d_work, (size_t)(ws_inplace > ws_sum ? ws_inplace : ws_sum));
  if (err == hipSuccess) err = hipMalloc(&b.d_sum, sizeof(Cplx));
  if (err == hipSuccess) err = hipHostMalloc(reinterpret_cast<void**>(&b.h_sum), sizeof(Cplx));
  if (err == hipSuccess && ng > 0) err = hipHostMalloc(&b.h_stage, gpu_bytes);
  if (err != hipSuccess) return lufqhe::hip_fail(err);
  if (ng > 0) std::memcpy(b.h_stage, A, gpu_bytes);

  using clock = std::chrono::steady_clock;
  const auto t_start = clock::now();
  int32_t status = LUFQHE_SUCCESS;
  for (int32_t it = 0; it < iterations; ++it) {
    Cplx gpu_sum{0.0, 0.0};
    if (ng > 0) {
      // queue the GPU share: copy, determinants with pair products, sum
      err = hipEventRecord(b.ev0, b.stream);
      if (err == hipSuccess)
        err = hipMemcpyAsync(b.d_A, b.h_stage, gpu_bytes, hipMemcpyHostToDevice, b.stream);
      if (err != hipSuccess) return lufqhe::hip_fail(err);
      int32_t st = lufqhe_det_inplace(b.d_A, ng, n, n, elems, /*pair_product=*/1, policy,
                                      b.d_work, b.stream);
      if (st == LUFQHE_SUCCESS)
        st = lufqhe_global_reduction(b.d_A, ng, elems, b.d_work, b.d_sum, b.stream);
      if (st != LUFQHE_SUCCESS) return st;
      err = hipMemcpyAsync(b.h_sum, b.d_sum, sizeof(Cplx), hipMemcpyDeviceToHost, b.stream);
      if (err == hipSuccess) err = hipEventRecord(b.ev1, b.stream);
      if (err != hipSuccess) return lufqhe::hip_fail(err);
    }
    // the host cores' share meanwhile
    Cplx cpu_sum{0.0, 0.0};
    const auto c0 = clock::now();
    int32_t first_bad = -1;
    const int32_t st = lufqhe_cpu_pair_sum(A + ng * elems, nc, n, n, elems, &cpu_sum, &first_bad,
                                           cpu_count > 0 ? cpu_count : 1);
    if (st != LUFQHE_SUCCESS) return st;
    if (first_bad >= 0) status = LUFQHE_ERROR_INVALID_VALUE;  // the reference exit(0)s there
    const auto c1 = clock::now();
    if (ng > 0) {
      err = hipStreamSynchronize(b.stream);
      if (err != hipSuccess) return lufqhe::hip_fail(err);
      gpu_sum = *b.h_sum;
      float ms = 0.f;
      if (stats && hipEventElapsedTime(&ms, b.ev0, b.ev1) == hipSuccess) stats->gpu_s += ms * 1e-3;
    }
    if (stats) stats->cpu_s += std::chrono::duration<double>(c1 - c0).count();
    // LU-fqhe/src/MonteCarloLoop.cu:360-364: the node's sum over its devices
    results[it] = Cplx{gpu_sum.re + cpu_sum.re, gpu_sum.im + cpu_sum.im};
  }
  if (stats) stats->wall_s = std::chrono::duration<double>(clock::now() - t_start).count();
  return status;
}
