// scaled_sum.hpp -- complex numbers carried as (mantissa, exponent) and
// their sums, for the pair sums of the epilogue and the walker server's
// per-request sums: determinants of FQHE matrices leave the double range
// long before their sums stop mattering.
#pragma once

#include <hip/hip_runtime.h>

namespace lufqhe {

// value = m * 2^e; max(|Re m|, |Im m|) in [0.5, 1), or m == 0 with e == 0.
struct Scaled {
  double2 m;
  int e;
};

// Normalises m * 2^e (exact: power-of-two scaling).
__device__ __forceinline__ Scaled scaled_norm(double2 m, int e) {
  const double a = fmax(fabs(m.x), fabs(m.y));
  if (a == 0.0) return Scaled{make_double2(0.0, 0.0), 0};
  const int k = __builtin_amdgcn_frexp_exp(a);
  return Scaled{make_double2(ldexp(m.x, -k), ldexp(m.y, -k)), e + k};
}

// a + b, aligned to the larger exponent; the smaller term is scaled exactly
// unless it drops below the double range, where it is negligible anyway.
__device__ __forceinline__ Scaled scaled_add(Scaled a, Scaled b) {
  if (a.m.x == 0.0 && a.m.y == 0.0) return b;
  if (b.m.x == 0.0 && b.m.y == 0.0) return a;
  if (a.e < b.e) {
    const Scaled t = a;
    a = b;
    b = t;
  }
  const int d = a.e - b.e;
  const double2 bs = d > 1100 ? make_double2(0.0, 0.0)
                              : make_double2(ldexp(b.m.x, -d), ldexp(b.m.y, -d));
  return scaled_norm(make_double2(a.m.x + bs.x, a.m.y + bs.y), a.e);
}

}  // namespace lufqhe
