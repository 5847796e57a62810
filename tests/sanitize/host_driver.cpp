// Exercises the library's host-only code -- the CPU Crout (lucore_cpu.cpp),
// RunInfo (run_info.cpp) and the shape-table reader (lufqhe_table.cpp) --
// under AddressSanitizer and UndefinedBehaviorSanitizer, without a GPU.  The
// few library functions those files call are stubbed here.  Exit status 0
// when every call returns; the sanitizers abort on the first finding.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "../../include/lucore_cpu.hpp"
#include "../../include/lufqhe.h"

extern "C" {
// stubs of the kernel tables: 4 candidate shapes per n
int32_t lufqhe_shape_count(int32_t n) { return (n >= 1 && n <= 64) ? 4 : 0; }
int32_t lufqhe_shape_info(int32_t n, int32_t shape, int32_t* lanes, int32_t* cols, int32_t* waves) {
  if (n < 1 || n > 64 || shape < 0 || shape > 3 || !lanes || !cols) return LUFQHE_ERROR_INVALID_VALUE;
  *lanes = 64;
  *cols = 2;
  if (waves) *waves = 1;
  return LUFQHE_SUCCESS;
}
int32_t lufqhe_set_shape(int32_t, int32_t, int32_t, int32_t) { return LUFQHE_SUCCESS; }
const char* lufqhe_status_string(int32_t) { return "stub"; }
}

static std::string write_tmp(const char* name, const std::string& text) {
  const char* dir = std::getenv("TMPDIR");
  std::string p = std::string(dir ? dir : "/tmp") + "/" + name;
  FILE* f = std::fopen(p.c_str(), "wb");
  if (!f) std::exit(3);
  std::fwrite(text.data(), 1, text.size(), f);
  std::fclose(f);
  return p;
}

int main(int argc, char** argv) {
  std::mt19937_64 rng(7);
  std::normal_distribution<double> g;
  for (int n = 1; n <= 64; ++n) {
    const int count = 7, ld = n + (n & 1);
    const int64_t stride = (int64_t)n * ld + 3;
    std::vector<double2> A((size_t)(count * stride + ld));
    for (auto& v : A) v = make_double2(g(rng), g(rng));
    for (int c = 0; c < n; ++c) A[(size_t)(2 * stride + c)].x = 0.0;  // a rejected row
    A[(size_t)(3 * stride)] = make_double2(NAN, INFINITY);
    std::vector<double2> det(count), mant(count);
    std::vector<int32_t> e(count), info(count);
    std::vector<int8_t> piv((size_t)count * n);
    for (int th : {1, 3})
      if (lufqhe_cpu_crout_det(A.data(), count, n, ld, stride, det.data(), mant.data(), e.data(),
                               info.data(), piv.data(), th) != LUFQHE_SUCCESS)
        return 4;
    double2 sum;
    int32_t bad;
    if (lufqhe_cpu_pair_sum(A.data(), count, n, ld, stride, &sum, &bad, 2) != LUFQHE_SUCCESS) return 5;
    std::vector<double2> one(A.begin(), A.begin() + (size_t)n * ld);
    double2 d1;
    LUCoreCPU::determinant(one.data(), &d1, n, ld);
    RunInfo ri(4, n, count, true);
    try {
      std::vector<double2> B((size_t)count * n * ri.getLoadSize());
      for (auto& v : B) v = make_double2(g(rng), g(rng));
      LUCoreCPU::CALL(B.data(), &sum, count, &ri, 0, 2);
    } catch (const LUCoreCPU::Error&) {
    }
  }
  // invalid arguments are refused without touching memory
  if (lufqhe_cpu_crout_det(nullptr, 3, 0, 1, 1, nullptr, nullptr, nullptr, nullptr, nullptr, 1) ==
      LUFQHE_SUCCESS)
    return 6;
  // the shape table: the committed one, then malformed files
  if (argc > 1 && lufqhe_load_shape_table(argv[1]) < 0) return 7;
  const char* bad_tables[] = {"", "{", "{\"entries\": [", "{\"entries\": [{\"n\": 99999999999}]}",
                              "{\"entries\": [{\"policy\": \"abs\", \"n\": \"32\", \"shape\": -4}]}",
                              "{\"entries\": [{\"n\": 8, \"shape\": 1, \"lanes\": 64, \"lane_cols\": 2",
                              "{\"entries\": [{}, {}, {\"n\":,}]}"};
  int i = 0;
  for (const char* t : bad_tables)
    (void)lufqhe_load_shape_table(write_tmp(("bad_table_" + std::to_string(i++)).c_str(), t).c_str());
  (void)lufqhe_load_shape_table("/nonexistent/shapes.json");
  std::printf("host driver: ok\n");
  return 0;
}
