// Drives the reference-shaped C++ API the way LU-fqhe's Monte Carlo loop
// does (LU-fqhe/src/MonteCarloLoop.cu:191-212,299-335,349-371): one device
// buffer per GPU thread, H2D copy, LUCoreGPU::CALL, globalReduction, D2H of
// the sum.  Test harness: reads nMat complex128 n x n matrices (unpadded,
// row major) from argv[4], writes the nMat [0][0] entries and then the sum
// (complex128 each) to argv[5].
//   lucore_gpu_driver nMat n padded in.bin out.bin
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/lucore_gpu.hpp"

#define CK(x)                                                         \
  do {                                                                \
    hipError_t e_ = (x);                                              \
    if (e_ != hipSuccess) {                                           \
      std::fprintf(stderr, "HIP error %s\n", hipGetErrorString(e_)); \
      return 2;                                                       \
    }                                                                 \
  } while (0)

int main(int argc, char** argv) {
  if (argc != 6) {
    std::fprintf(stderr, "usage: %s nMat n padded in.bin out.bin\n", argv[0]);
    return 1;
  }
  const int nMat = std::atoi(argv[1]), n = std::atoi(argv[2]), padded = std::atoi(argv[3]);
  RunInfo info(/*nMperBl=*/4, n, nMat, padded != 0);
  const int ld = info.getLoadSize();
  std::vector<double2> host((size_t)nMat * n * n), dev_img((size_t)nMat * n * ld);
  FILE* f = std::fopen(argv[4], "rb");
  if (!f || std::fread(host.data(), sizeof(double2), host.size(), f) != host.size()) return 1;
  std::fclose(f);
  for (int m = 0; m < nMat; ++m)
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < ld; ++c)
        dev_img[((size_t)m * n + r) * ld + c] =
            c < n ? host[((size_t)m * n + r) * n + c] : make_double2(0.0, 0.0);

  const int th_num = 0;
  double2* d_MatricesHolder[1];
  double2* d_ReductionHolder[1];
  CK(hipMalloc(&d_MatricesHolder[th_num], dev_img.size() * sizeof(double2)));
  CK(hipMalloc(&d_ReductionHolder[th_num], sizeof(double2)));
  hipStream_t stream1;
  CK(hipStreamCreate(&stream1));
  CK(hipMemcpyAsync(d_MatricesHolder[th_num], dev_img.data(), dev_img.size() * sizeof(double2),
                    hipMemcpyHostToDevice, stream1));
  try {
    LUCoreGPU::CALL(d_MatricesHolder, nMat, &info, /*rank=*/0, th_num, &stream1);
    LUCoreGPU::globalReduction(d_MatricesHolder, d_ReductionHolder, nMat, 0, th_num, stream1);
  } catch (const LUCoreGPU::Error& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 3;
  }
  double2 sum;
  CK(hipMemcpyAsync(&sum, d_ReductionHolder[th_num], sizeof(double2), hipMemcpyDeviceToHost,
                    stream1));
  CK(hipMemcpyAsync(dev_img.data(), d_MatricesHolder[th_num], dev_img.size() * sizeof(double2),
                    hipMemcpyDeviceToHost, stream1));
  CK(hipStreamSynchronize(stream1));
  std::vector<double2> out(nMat + 1);
  for (int m = 0; m < nMat; ++m) out[m] = dev_img[(size_t)m * n * ld];
  out[nMat] = sum;
  f = std::fopen(argv[5], "wb");
  if (!f || std::fwrite(out.data(), sizeof(double2), out.size(), f) != out.size()) return 1;
  std::fclose(f);
  CK(hipFree(d_MatricesHolder[th_num]));
  CK(hipFree(d_ReductionHolder[th_num]));
  CK(hipStreamDestroy(stream1));
  return 0;
}
