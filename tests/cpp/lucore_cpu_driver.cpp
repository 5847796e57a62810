// Drives the reference-shaped CPU API the way LU-fqhe's Monte Carlo loop
// does for its CPU share (LU-fqhe/src/MonteCarloLoop.cu:343): LUCoreCPU::CALL
// over a batch, plus LUCoreCPU::determinant per matrix.  Test harness: reads
// nMat complex128 n x n matrices (unpadded, row major) from argv[1] and
// prints, as hex floats, the nMat determinants (re im), the pair sum of
// CALL (re im) and 1 if a row with no real part makes CALL throw.
//   lucore_cpu_driver in.bin nMat n cpuCnt
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/lucore_cpu.hpp"

int main(int argc, char** argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s in.bin nMat n cpuCnt\n", argv[0]);
    return 1;
  }
  const int nMat = std::atoi(argv[2]), n = std::atoi(argv[3]), cpu = std::atoi(argv[4]);
  std::vector<double2> m((size_t)nMat * n * n);
  FILE* f = std::fopen(argv[1], "rb");
  if (!f || std::fread(m.data(), sizeof(double2), m.size(), f) != m.size()) return 2;
  std::fclose(f);
  RunInfo info(/*nMperBl=*/4, n, nMat);

  double2 sum;
  LUCoreCPU::CALL(m.data(), &sum, nMat, &info, /*rank=*/0, cpu);
  for (int i = 0; i < nMat; ++i) {
    std::vector<double2> a(m.begin() + (size_t)i * n * n, m.begin() + (size_t)(i + 1) * n * n);
    double2 det;
    LUCoreCPU::determinant(a.data(), &det, n, n);
    std::printf("%a %a\n", det.x, det.y);
  }
  std::printf("%a %a\n", sum.x, sum.y);

  m[0].x = 0.0;  // row 0 of matrix 0 keeps no real part: the reference exit(0)s
  for (int c = 1; c < n; ++c) m[c].x = 0.0;
  int threw = 0;
  try {
    LUCoreCPU::CALL(m.data(), &sum, nMat, &info, 0, cpu);
  } catch (const LUCoreCPU::Error& e) {
    threw = 1;
  }
  std::printf("%a\n", (double)threw);
  return 0;
}
