// lufqhe_epilogue.hip -- the reference's determinant epilogues.
//
//  * in-place write-back: A_m[0][0] = det_m, and with the pair product
//    A_2i[0][0] = det_2i * det_2i+1 (LU-fqhe/src/LUCoreGPU.cu:384-436 with
//    WRITE_DET and PERFORM_PROD, the reference Makefile's build,
//    LU-fqhe/Makefile:3);
//  * the wavefunction sum  sum_i det_2i * det_2i+1  (LU-fqhe/src/LUCoreGPU.cu:
//    113-156,447-557; LU-fqhe/src/LUCoreCPU.cu:83-98), in two forms:
//      plain  -- complex128 sum of the A_2i[0][0] entries, like the reference;
//      scaled -- from (mant, exp) determinants, carried as (mant, exp) so it
//                cannot overflow.
//
// Both sums are deterministic: a fixed per-thread slice order, a fixed
// in-block tree and a second single-block pass over the block partials (the
// reference's multi-block pass lets every block write output[0],
// LU-fqhe/src/LUCoreGPU.cu:553-556).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lup_kernel.hpp"
#include "scaled_sum.hpp"

namespace lufqhe {

constexpr int kSumThreads = 256;
constexpr int kSumMaxBlocks = 1024;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  // same formula as LU-fqhe/src/LUCoreCPU.cu:88-89
  return make_double2(a.x * b.x - a.y * b.y, a.y * b.x + a.x * b.y);
}

// (mant, exp) sums: Scaled, scaled_norm, scaled_add (scaled_sum.hpp).

// ---------------------------------------------------------------------------
// In-place write-back of plain determinants into A_m[0][0].
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) inplace_write_kernel(
    double2* __restrict__ A, int64_t count, int64_t stride,
    const double2* __restrict__ mant, const int32_t* __restrict__ expo,
    int pair_product) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= count) return;
  const double2 d = make_double2(ldexp(mant[m].x, expo[m]), ldexp(mant[m].y, expo[m]));
  if (pair_product && (m % 2) == 0 && m + 1 < count) {
    const double2 d1 = make_double2(ldexp(mant[m + 1].x, expo[m + 1]),
                                    ldexp(mant[m + 1].y, expo[m + 1]));
    A[m * stride] = cmul(d, d1);
  } else {
    A[m * stride] = d;
  }
}

// ---------------------------------------------------------------------------
// Plain pair sum: sum over i < count/2 of A_2i[0][0] (which holds the pair
// product after the in-place write-back).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 block_sum(double2 v, double2* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = kSumThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sh[threadIdx.x].x += sh[threadIdx.x + w].x;
      sh[threadIdx.x].y += sh[threadIdx.x + w].y;
    }
    __syncthreads();
  }
  return sh[0];
}

__global__ void __launch_bounds__(kSumThreads) pair_sum_plain_kernel(
    const double2* __restrict__ A, int64_t pairs, int64_t stride,
    double2* __restrict__ partial) {
  __shared__ double2 sh[kSumThreads];
  const int64_t per = (pairs + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * per;
  const int64_t hi = min(pairs, lo + per);
  double2 acc = make_double2(0.0, 0.0);
  for (int64_t i = lo + threadIdx.x; i < hi; i += kSumThreads) {
    const double2 v = A[2 * i * stride];
    acc.x += v.x;
    acc.y += v.y;
  }
  const double2 s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kSumThreads) sum_partials_plain_kernel(
    const double2* __restrict__ partial, int nparts, double2* __restrict__ out) {
  __shared__ double2 sh[kSumThreads];
  double2 acc = make_double2(0.0, 0.0);
  for (int i = threadIdx.x; i < nparts; i += kSumThreads) {
    acc.x += partial[i].x;
    acc.y += partial[i].y;
  }
  const double2 s = block_sum(acc, sh);
  if (threadIdx.x == 0) out[0] = s;
}

// ---------------------------------------------------------------------------
// Scaled pair sum: sum_i det_2i * det_2i+1 from (mant, exp) determinants.
// ---------------------------------------------------------------------------
__device__ __forceinline__ Scaled block_sum_scaled(Scaled v, double2* shm, int* she) {
  shm[threadIdx.x] = v.m;
  she[threadIdx.x] = v.e;
  __syncthreads();
  for (int w = kSumThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const Scaled r = scaled_add(Scaled{shm[threadIdx.x], she[threadIdx.x]},
                                  Scaled{shm[threadIdx.x + w], she[threadIdx.x + w]});
      shm[threadIdx.x] = r.m;
      she[threadIdx.x] = r.e;
    }
    __syncthreads();
  }
  return Scaled{shm[0], she[0]};
}

__global__ void __launch_bounds__(kSumThreads) pair_sum_scaled_kernel(
    const double2* __restrict__ mant, const int32_t* __restrict__ expo, int64_t pairs,
    double2* __restrict__ part_m, int32_t* __restrict__ part_e) {
  __shared__ double2 shm[kSumThreads];
  __shared__ int she[kSumThreads];
  const int64_t per = (pairs + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * per;
  const int64_t hi = min(pairs, lo + per);
  Scaled acc{make_double2(0.0, 0.0), 0};
  for (int64_t i = lo + threadIdx.x; i < hi; i += kSumThreads) {
    const double2 p = cmul(mant[2 * i], mant[2 * i + 1]);
    acc = scaled_add(acc, scaled_norm(p, expo[2 * i] + expo[2 * i + 1]));
  }
  const Scaled s = block_sum_scaled(acc, shm, she);
  if (threadIdx.x == 0) {
    part_m[blockIdx.x] = s.m;
    part_e[blockIdx.x] = s.e;
  }
}

__global__ void __launch_bounds__(kSumThreads) sum_partials_scaled_kernel(
    const double2* __restrict__ part_m, const int32_t* __restrict__ part_e, int nparts,
    double2* __restrict__ out_m, int32_t* __restrict__ out_e) {
  __shared__ double2 shm[kSumThreads];
  __shared__ int she[kSumThreads];
  Scaled acc{make_double2(0.0, 0.0), 0};
  for (int i = threadIdx.x; i < nparts; i += kSumThreads)
    acc = scaled_add(acc, Scaled{part_m[i], part_e[i]});
  const Scaled s = block_sum_scaled(acc, shm, she);
  if (threadIdx.x == 0) {
    out_m[0] = s.m;
    out_e[0] = s.e;
  }
}

int sum_blocks(int64_t pairs) {
  const int64_t b = (pairs + kSumThreads * 8 - 1) / (kSumThreads * 8);
  return (int)(b < 1 ? 1 : (b > kSumMaxBlocks ? kSumMaxBlocks : b));
}

hipError_t launch_inplace_write(double2* A, int64_t count, int64_t stride, const double2* mant,
                                const int32_t* expo, int pair_product, hipStream_t stream) {
  const int64_t blocks = (count + 255) / 256;
  hipLaunchKernelGGL(inplace_write_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, A,
                     count, stride, mant, expo, pair_product);
  return hipGetLastError();
}

hipError_t launch_pair_sum_plain(const double2* A, int64_t pairs, int64_t stride,
                                 double2* partial, double2* out, hipStream_t stream) {
  const int nb = sum_blocks(pairs);
  hipLaunchKernelGGL(pair_sum_plain_kernel, dim3(nb), dim3(kSumThreads), 0, stream, A, pairs,
                     stride, partial);
  hipLaunchKernelGGL(sum_partials_plain_kernel, dim3(1), dim3(kSumThreads), 0, stream, partial,
                     nb, out);
  return hipGetLastError();
}

hipError_t launch_pair_sum_scaled(const double2* mant, const int32_t* expo, int64_t pairs,
                                  double2* part_m, int32_t* part_e, double2* out_m,
                                  int32_t* out_e, hipStream_t stream) {
  const int nb = sum_blocks(pairs);
  hipLaunchKernelGGL(pair_sum_scaled_kernel, dim3(nb), dim3(kSumThreads), 0, stream, mant, expo,
                     pairs, part_m, part_e);
  hipLaunchKernelGGL(sum_partials_scaled_kernel, dim3(1), dim3(kSumThreads), 0, stream, part_m,
                     part_e, nb, out_m, out_e);
  return hipGetLastError();
}

}  // namespace lufqhe
