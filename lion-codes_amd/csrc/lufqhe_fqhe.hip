// lufqhe_fqhe.hip -- FQHE wavefunction matrices on the device.
//
// The batch kernels build A_ij = z_i^j exp(-|z_i|^2/4) from particle
// positions in their load phase (BatchArgs::src == SRC_FQHE), so an FQHE
// batch reads 16n bytes of positions per matrix instead of 16n^2 bytes of
// matrix.  This file holds the kernel that writes the same matrices out:
// entry for entry the values the load phase computes (fqhe_gauss and the
// sequential power product are shared), which lets tests compare the fused
// path bit for bit with a determinant of the materialised batch.
#include <hip/hip_runtime.h>

#include "lup_kernel.hpp"

namespace lufqhe {

// Matrix i's row r comes from position z[i*z_stride + p], p = row_perm[i*n+r]
// (or r); a permutation entry >= n gives a row of NaN (the determinant
// kernels report such a matrix with info kInfoBadPerm).  Output matrix i at
// out + i*n*n, ld = n.  One thread per row.
__global__ void __launch_bounds__(256) fqhe_build_kernel(const double2* __restrict__ z,
                                                         int64_t count, int n, int64_t z_stride,
                                                         const uint8_t* __restrict__ row_perm,
                                                         double2* __restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= count * n) return;
  const int64_t i = r / n;
  const int row = (int)(r - i * n);
  int src = row;
  if (row_perm) src = row_perm[r];
  double2* o = out + r * n;
  if (src >= n) {
    for (int j = 0; j < n; ++j) o[j] = make_double2(__builtin_nan(""), __builtin_nan(""));
    return;
  }
  const double2 zi = z[i * z_stride + src];
  const double g = fqhe_gauss(zi);
  double2 pw = make_double2(1.0, 0.0);
  for (int j = 0; j < n; ++j) {
    o[j] = make_double2(pw.x * g, pw.y * g);
    pw = cmul_ref(pw, zi);
  }
}

hipError_t launch_fqhe_build(const double2* z, int64_t count, int n, int64_t z_stride,
                             const uint8_t* row_perm, double2* out, hipStream_t stream) {
  const int64_t rows = count * n;
  const int64_t blocks = (rows + 255) / 256;
  hipLaunchKernelGGL(fqhe_build_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, z, count,
                     n, z_stride, row_perm, out);
  return hipGetLastError();
}

}  // namespace lufqhe
