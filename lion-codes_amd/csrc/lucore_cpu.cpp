// lucore_cpu.cpp -- the reference's CPU determinant path on the host: the C
// ABI entry points lufqhe_cpu_crout_det / lufqhe_cpu_pair_sum and the C++
// API LUCoreCPU::CALL / determinant (include/lucore_cpu.hpp).
//
// The algorithm is LUCoreCPU::determinant (LU-fqhe/src/LUCoreCPU.cu:112-267)
// at complex128: a left-looking Crout with implicit row scaling, every
// product, sum and quotient rounded as the reference writes it (no fused
// multiply-adds), so that the results are those of an FP64 build of the
// reference and of the GPU kernels' REF arithmetic, bit for bit.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lucore_cpu.hpp"
#include "../../include/lufqhe.h"

#pragma clang fp contract(off)

namespace {

constexpr int kMaxSide = 64;
constexpr double kRowScaleMin = 1e-24;  // LU-fqhe/src/LUCoreCPU.cu:158

// Interleaved complex128 element, the layout of double2.
struct Cplx {
  double re, im;
};

// a - b*c with the reference's rounding steps (LU-fqhe/inc/complex.cuh
// products; LU-fqhe/src/LUCoreCPU.cu:173-179).
inline Cplx sub_prod(Cplx a, Cplx b, Cplx c) {
  const double tx = b.re * c.re - b.im * c.im;
  const double ty = b.im * c.re + b.re * c.im;
  return Cplx{a.re - tx, a.im - ty};
}

// mant * 2^e: scale so that max(|re|, |im|) is in [0.5, 1) (exact).
inline void renormalise(Cplx& m, int32_t& e) {
  const double big = fmax(fabs(m.re), fabs(m.im));
  if (big == 0.0 || !isfinite(big)) return;
  int k;
  (void)frexp(big, &k);
  m.re = ldexp(m.re, -k);
  m.im = ldexp(m.im, -k);
  e += k;
}

struct CroutResult {
  Cplx det;      // the reference's plain product
  Cplx mant;     // the same product carried as mant * 2^exp
  int32_t exp;
  int32_t info;  // 0; j+1 zero pivot at column j; -(i+1) row i rejected
};

// Factors the n x n matrix at m (row stride ld) in place and returns its
// determinant.  piv (n entries, may be null) receives the row swapped into
// position j at column j.
CroutResult crout_in_place(Cplx* m, int n, int64_t ld, int8_t* piv) {
  auto at = [m, ld](int r, int c) -> Cplx& { return m[(int64_t)r * ld + c]; };
  CroutResult res{{1.0, 0.0}, {1.0, 0.0}, 0, 0};
  double scale[kMaxSide];

  // implicit row scaling: the largest |Re| of each row (LUCoreCPU.cu:145-160)
  for (int r = 0; r < n; ++r) {
    double big = 0.0;
    for (int c = 0; c < n; ++c) big = std::max(big, fabs(at(r, c).re));
    if (big < kRowScaleMin) {  // the reference exit(0)s here
      const double nan = __builtin_nan("");
      res.det = res.mant = Cplx{nan, nan};
      res.info = -(r + 1);
      return res;
    }
    scale[r] = big;
  }

  double sign = 1.0;
  int imax = 0;
  for (int j = 0; j < n; ++j) {
    // U part of column j (LUCoreCPU.cu:164-182)
    for (int r = 0; r < j; ++r) {
      Cplx s = at(r, j);
      for (int k = 0; k < r; ++k) s = sub_prod(s, at(r, k), at(k, j));
      at(r, j) = s;
    }
    // L part and the pivot: the largest |Re| / scale, the last row among
    // equals (LUCoreCPU.cu:188-213)
    double best = 0.0;
    for (int r = j; r < n; ++r) {
      Cplx s = at(r, j);
      for (int k = 0; k < j; ++k) s = sub_prod(s, at(r, k), at(k, j));
      at(r, j) = s;
      const double crit = fabs(s.re) / scale[r];
      if (crit >= best) {
        best = crit;
        imax = r;
      }
    }
    // row swap, sign flip and scale move (LUCoreCPU.cu:215-232)
    if (imax != j) {
      for (int c = 0; c < n; ++c) std::swap(at(imax, c), at(j, c));
      sign = -sign;
      scale[imax] = scale[j];
    }
    if (piv) piv[j] = (int8_t)imax;
    // column scaling by the complex divide (LUCoreCPU.cu:234-251); a zero
    // |p|^2 (where the reference divides by zero) zeroes the column
    const Cplx p = at(j, j);
    const double den = p.re * p.re + p.im * p.im;
    if (den == 0.0) {
      if (res.info == 0) res.info = j + 1;
      for (int r = j + 1; r < n; ++r) at(r, j) = Cplx{0.0, 0.0};
    } else if (j != n - 1) {
      for (int r = j + 1; r < n; ++r) {
        const Cplx a = at(r, j);
        at(r, j) = Cplx{(a.re * p.re + a.im * p.im) / den, (a.im * p.re - a.re * p.im) / den};
      }
    }
  }

  // the determinant: diagonal product left to right, times the swap sign
  // (LUCoreCPU.cu:256-265)
  for (int i = 0; i < n; ++i) {
    const Cplx u = at(i, i);
    res.det = Cplx{res.det.re * u.re - res.det.im * u.im, res.det.im * u.re + res.det.re * u.im};
    res.mant =
        Cplx{res.mant.re * u.re - res.mant.im * u.im, res.mant.im * u.re + res.mant.re * u.im};
    renormalise(res.mant, res.exp);
  }
  res.det = Cplx{sign * res.det.re, sign * res.det.im};
  res.mant = Cplx{sign * res.mant.re, sign * res.mant.im};
  return res;
}

// Runs body(first, last) over [0, count) split into contiguous ranges, one
// per thread.
template <class F>
void parallel_ranges(int64_t count, int nthreads, F body) {
  int t = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
  t = (int)std::min<int64_t>(t, std::max<int64_t>(count, 1));
  if (t <= 1) {
    body(int64_t{0}, count);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(t);
  for (int i = 0; i < t; ++i) {
    const int64_t first = count * i / t, last = count * (i + 1) / t;
    pool.emplace_back([=] { body(first, last); });
  }
  for (auto& th : pool) th.join();
}

// Determinant of the matrix at src (row stride ld), factored in a private copy.
CroutResult crout_copy(const Cplx* src, int n, int64_t ld, int8_t* piv) {
  Cplx work[kMaxSide * kMaxSide];
  for (int r = 0; r < n; ++r) memcpy(work + r * n, src + (int64_t)r * ld, sizeof(Cplx) * n);
  return crout_in_place(work, n, n, piv);
}

}  // namespace

extern "C" {

int32_t lufqhe_cpu_crout_det(const void* h_A, int64_t count, int32_t n, int64_t ld,
                             int64_t stride, void* h_det, void* h_mant, int32_t* h_exp,
                             int32_t* h_info, int8_t* h_piv, int32_t nthreads) {
  if (count < 0 || n < 1 || n > kMaxSide || ld < n || stride < 0)
    return LUFQHE_ERROR_INVALID_VALUE;
  if (count == 0) return LUFQHE_SUCCESS;
  if (!h_A) return LUFQHE_ERROR_INVALID_VALUE;
  const Cplx* A = static_cast<const Cplx*>(h_A);
  Cplx* det = static_cast<Cplx*>(h_det);
  Cplx* mant = static_cast<Cplx*>(h_mant);
  parallel_ranges(count, nthreads, [&](int64_t first, int64_t last) {
    for (int64_t i = first; i < last; ++i) {
      int8_t* piv = h_piv ? h_piv + i * n : nullptr;
      const CroutResult r = crout_copy(A + i * stride, n, ld, piv);
      if (r.info < 0 && piv) memset(piv, 0xff, n);
      if (det) det[i] = r.det;
      if (mant) mant[i] = r.mant;
      if (h_exp) h_exp[i] = r.exp;
      if (h_info) h_info[i] = r.info;
    }
  });
  return LUFQHE_SUCCESS;
}

int32_t lufqhe_cpu_pair_sum(const void* h_A, int64_t num, int32_t n, int64_t ld, int64_t stride,
                            void* h_out, int32_t* h_first_bad, int32_t nthreads) {
  if (num < 0 || n < 1 || n > kMaxSide || ld < n || stride < 0 || !h_out)
    return LUFQHE_ERROR_INVALID_VALUE;
  const int64_t pairs = num / 2;
  if (pairs > 0 && !h_A) return LUFQHE_ERROR_INVALID_VALUE;
  const Cplx* A = static_cast<const Cplx*>(h_A);
  std::vector<Cplx> prod(pairs);
  std::vector<int32_t> bad(pairs, 0);
  parallel_ranges(pairs, nthreads, [&](int64_t first, int64_t last) {
    for (int64_t i = first; i < last; ++i) {
      const CroutResult d1 = crout_copy(A + 2 * i * stride, n, ld, nullptr);
      const CroutResult d2 = crout_copy(A + (2 * i + 1) * stride, n, ld, nullptr);
      // LUCoreCPU.cu:88-89
      prod[i] = Cplx{d1.det.re * d2.det.re - d1.det.im * d2.det.im,
                     d1.det.im * d2.det.re + d1.det.re * d2.det.im};
      bad[i] = (d1.info < 0 || d2.info < 0);
    }
  });
  // added in matrix order (the reference adds per-thread partial sums in
  // the order the threads finish, LUCoreCPU.cu:94-98)
  Cplx sum{0.0, 0.0};
  int32_t first_bad = -1;
  for (int64_t i = 0; i < pairs; ++i) {
    if (bad[i] && first_bad < 0) first_bad = (int32_t)std::min<int64_t>(i, INT32_MAX);
    sum = Cplx{sum.re + prod[i].re, sum.im + prod[i].im};
  }
  *static_cast<Cplx*>(h_out) = sum;
  if (h_first_bad) *h_first_bad = first_bad;
  return LUFQHE_SUCCESS;
}

}  // extern "C"

namespace LUCoreCPU {

void CALL(double2* matrices, double2* result, int num, RunInfo* rInfo, int rank, int cpuCnt) {
  if (!rInfo || !result || (num >= 2 && !matrices))
    throw Error("LUCoreCPU::CALL on node " + std::to_string(rank) + ": invalid arguments",
                LUFQHE_ERROR_INVALID_VALUE);
  const int n = rInfo->getMatSide();
  const int ld = rInfo->getLoadSize();
  int32_t first_bad = -1;
  const int32_t st = lufqhe_cpu_pair_sum(matrices, num, n, ld, (int64_t)n * ld, result,
                                         &first_bad, cpuCnt);
  if (st != LUFQHE_SUCCESS)
    throw Error("LUCoreCPU::CALL on node " + std::to_string(rank) + ": " +
                    lufqhe_status_string(st), st);
  if (first_bad >= 0)
    throw Error("LUCoreCPU::CALL on node " + std::to_string(rank) + ": pair " +
                    std::to_string(first_bad) + " holds a row with max |Re| < 1e-24",
                LUFQHE_ERROR_INVALID_VALUE);
}

void determinant(double2* inputMatrix, double2* det, int matSide, int loadSize) {
  if (!inputMatrix || !det || matSide < 1 || matSide > kMaxSide || loadSize < matSide)
    throw Error("LUCoreCPU::determinant: invalid arguments", LUFQHE_ERROR_INVALID_VALUE);
  const CroutResult r =
      crout_in_place(reinterpret_cast<Cplx*>(inputMatrix), matSide, loadSize, nullptr);
  if (r.info < 0)
    throw Error("LUCoreCPU::determinant: row " + std::to_string(-r.info - 1) +
                    " has max |Re| < 1e-24",
                LUFQHE_ERROR_INVALID_VALUE);
  det->x = r.det.re;
  det->y = r.det.im;
}

}  // namespace LUCoreCPU
