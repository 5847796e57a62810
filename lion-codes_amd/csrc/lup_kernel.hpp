// lup_kernel.hpp -- batched LUP determinant kernels for gfx950 (CDNA4).
//
// Replaces LUCoreGPU::LUDecomposition (LU-fqhe/src/LUCoreGPU.cu:159-437) and
// its reductions vOP::* (LU-fqhe/src/vectorOP.cu:23-372).
//
// Layout (see Tile): a group of L lanes factors one matrix, a wave64 holds
// G = 64/L groups.  The group's lanes form L/CG lane-rows by CG lane-columns;
// rows and columns are dealt to them cyclically and the whole matrix stays in
// VGPRs.  Rows never move: each row slot carries the row's *reference
// position*, which is what the reference's physical row swaps
// (LU-fqhe/src/LUCoreCPU.cu:215-232) would have made of it.
//
// Step k of the right-looking factorisation (factor_tile):
//   1. criterion c on column k (per policy), -1 on finished rows;
//   2. exact pivot choice: cmax = group max of c, then the largest position
//      among candidates with c == cmax (the reference's `>=` scan keeps the
//      last position among exact ties, LU-fqhe/src/LUCoreCPU.cu:207-212);
//      butterflies use DPP (quad_perm, row_half_mirror, row_mirror) and the
//      gfx950 v_permlane16_swap / v_permlane32_swap across 16-lane rows;
//      this search is woven into step k-1's update (update_and_search);
//   3. the pivot's lane-row writes the pivot row (columns k..n-1) to the
//      group's LDS slot and every lane reads its columns back;
//   4. det *= pivot (ARITH_FUSED: the pivot scaled by a power of two, the
//      exponents summed, one renormalisation to mant * 2^exp at the end);
//   5. multipliers l = a_ik / p (lane-row DPP broadcast of column k) and the
//      rank-1 update of columns k+1..n-1 (complex FP64, 4 FMAs per element in
//      ARITH_FUSED);
//   6. the position bookkeeping of the reference swap; the pivot row retires.
// The k loop is fully unrolled (N is a template constant) so that finished
// column slots are skipped statically and the tile stays in registers; once
// the active rows fit in fewer row slots the tile is compacted into them
// (compact).  The load phase reads the matrix from HBM, or builds an FQHE
// matrix from particle positions (SRC_FQHE).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

namespace lufqhe {

enum Policy { POLICY_CROUT_REF = 0, POLICY_ABS = 1 };
enum Arith { ARITH_FUSED = 0, ARITH_REF = 1 };

constexpr double kDblMax = 1.7976931348623157e308;
constexpr double kRowScaleMin = 1e-24;  // LU-fqhe/src/LUCoreCPU.cu:158
constexpr double kLn2 = 0.69314718055994530942;
// Range in which the corrected product x * RN(1/s) of quot_rn below is the
// correctly rounded x / s: the reciprocal, the first product and the
// residual stay normal.  Outside it (and for x == 0) the quotient is an
// IEEE division.
constexpr double kQuotLo = 0x1p-880;
constexpr double kQuotHi = 0x1p960;
// info code of a matrix whose row-permutation table has an entry >= n.
constexpr int kInfoBadPerm = -1000;

// Output pointers of one launch.  Any of them may be null except mant.
struct DetOut {
  double2* mant;     // det = mant * 2^exp, max(|re|,|im|) of mant in [0.5,1)
  int32_t* expo;
  double* logabs;    // log|det|
  double* phase;     // arg det in (-pi, pi]
  int32_t* info;     // 0, k+1 (zero pivot at step k), -(i+1) (row i rejected),
                     // kInfoBadPerm (row-permutation entry out of range)
  int8_t* piv;       // n bytes per matrix: reference position swapped at step k
};

// ---------------------------------------------------------------------------
// cross-lane helpers
// ---------------------------------------------------------------------------
// Every pattern used here reads a valid source lane for every lane, so the
// no-old-value form applies (no register has to be cleared first).
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = dpp_i<CTRL>(__double2loint(v));
  const int hi = dpp_i<CTRL>(__double2hiint(v));
  return __hiloint2double(hi, lo);
}

// DPP controls: xor 1 and xor 2 inside a quad, then the mirrors, which reach
// the other half of an 8-lane and a 16-lane row once the smaller blocks agree.
constexpr int kQuadXor1 = 0xB1;   // quad_perm [1,0,3,2]
constexpr int kQuadXor2 = 0x4E;   // quad_perm [2,3,0,1]
constexpr int kHalfMirror = 0x141;
constexpr int kMirror = 0x140;

// v_max_f64 without the NaN-quieting canonicalisations the compiler adds
// around fmax in IEEE mode: every criterion is finite (clamped) or -1.
__device__ __forceinline__ double vmax_f64(double a, double b) {
  double r;
  asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// max(|a|, |b|) in one instruction (abs source modifiers).
__device__ __forceinline__ double vmax_abs_f64(double a, double b) {
  double r;
  asm("v_max_f64 %0, |%1|, |%2|" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// Max over the L lanes of a group (L a power of two, groups aligned).
template <int L>
__device__ __forceinline__ double group_max_d(double v) {
  if constexpr (L > 1) v = vmax_f64(v, dpp_d<kQuadXor1>(v));
  if constexpr (L > 2) v = vmax_f64(v, dpp_d<kQuadXor2>(v));
  if constexpr (L > 4) v = vmax_f64(v, dpp_d<kHalfMirror>(v));
  if constexpr (L > 8) v = vmax_f64(v, dpp_d<kMirror>(v));
  if constexpr (L > 16) {
    const unsigned lo = __double2loint(v), hi = __double2hiint(v);
    const auto sl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto sh = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    v = vmax_f64(__hiloint2double(sh[0], sl[0]),
                 __hiloint2double(sh[1], sl[1]));
  }
  if constexpr (L > 32) {
    const unsigned lo = __double2loint(v), hi = __double2hiint(v);
    const auto sl = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto sh = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    v = vmax_f64(__hiloint2double(sh[0], sl[0]),
                 __hiloint2double(sh[1], sl[1]));
  }
  return v;
}

template <int L>
__device__ __forceinline__ int group_max_i(int v) {
  if constexpr (L > 1) v = max(v, dpp_i<kQuadXor1>(v));
  if constexpr (L > 2) v = max(v, dpp_i<kQuadXor2>(v));
  if constexpr (L > 4) v = max(v, dpp_i<kHalfMirror>(v));
  if constexpr (L > 8) v = max(v, dpp_i<kMirror>(v));
  if constexpr (L > 16) {
    const auto s = __builtin_amdgcn_permlane16_swap((unsigned)v, (unsigned)v,
                                                    false, false);
    v = max((int)s[0], (int)s[1]);
  }
  if constexpr (L > 32) {
    const auto s = __builtin_amdgcn_permlane32_swap((unsigned)v, (unsigned)v,
                                                    false, false);
    v = max((int)s[0], (int)s[1]);
  }
  return v;
}

// Orders LDS traffic between lanes of one wave: the wave executes its LDS
// instructions in order, so only the compiler has to be kept from moving
// accesses across this point.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Select by value.  The arguments are copies, so no pointer into the tile is
// ever selected (a select of two addresses in the tile would keep the whole
// tile out of registers).
__device__ __forceinline__ double2 sel2(bool c, double2 x, double2 y) {
  return make_double2(c ? x.x : y.x, c ? x.y : y.y);
}

// Correctly rounded x / s for x >= 0 from y = RN(1/s): q0 = RN(x*y) is within
// one ulp of RN(x/s), and one residual correction, r = x - s*q0 (exact in
// an FMA) and q = RN(q0 + r*y), gives RN(x/s) (Markstein's correction,
// P. Markstein, IBM J. Res. Dev. 34(1), 1990).  Valid while y, q0 and the
// residual stay normal: y in [2^-960, 2^960] (the caller passes y = NaN for
// a row scale outside it) and q0 in [kQuotLo, kQuotHi], which bounds x from
// below through y <= 2^80 (s >= 1e-24).  Anything else -- x == 0, tiny or
// huge quotients, inf, NaN -- takes an IEEE division, a branch no lane of an
// ordinary batch enters.  The result equals the reference's `fabs(x)/s`
// (LU-fqhe/src/LUCoreCPU.cu:205) bit for bit, so exact criterion ties break
// as the reference's `>=` scan does.
__device__ __forceinline__ double quot_rn(double x, double s, double y, bool need) {
  const double q0 = x * y;
  const double r = fma(-q0, s, x);
  double q = fma(r, y, q0);
  const bool ok = (q0 >= kQuotLo) & (q0 <= kQuotHi);
  if (__builtin_expect(need && !ok, 0)) q = x / s;
  return q;
}

// RN(1/d) for the scaled |p_s|^2 of a pivot, d in [0.25, 2), and 0 for
// d == 0 (a zero pivot) or d == inf: Newton steps from the hardware
// reciprocal estimate and a final residual correction -- the operations of
// the IEEE division 1.0 / d without its range scaling and special-case
// fixup, which are the identity for d in [0.25, 2).  NaN gives NaN.
__device__ __forceinline__ double recip_scaled(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  const double rem = fma(-d, r, 1.0);
  const double q = fma(rem, r, r);
  // class mask: -inf, -0, +0, +inf
  return __builtin_amdgcn_class(d, (1 << 2) | (1 << 5) | (1 << 6) | (1 << 9)) ? 0.0 : q;
}

// min(a, b) without the NaN-quieting canonicalisation fmin adds; a NaN
// operand yields the other.
__device__ __forceinline__ double vmin_f64(double a, double b) {
  double r;
  asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// mant * 2^e renormalisation, exact (power-of-two scaling).
__device__ __forceinline__ void renorm(double& x, double& y, int& e) {
  const double m = vmax_abs_f64(x, y);           // a NaN operand yields the other
  const int k = __builtin_amdgcn_frexp_exp(m);  // 0 for 0, inf, nan
  x = ldexp(x, -k);
  y = ldexp(y, -k);
  e += k;
}

// ---------------------------------------------------------------------------
// Tile geometry.  A group of L lanes factors one matrix.  The lanes form
// RG = L/CG lane-rows by CG lane-columns (lane = rb*CG + cb, so a lane-row is
// CG adjacent lanes).  Lane (rb, cb) holds the R row slots rb + t*RG and the C
// column slots cb + s*CG, i.e. rows and columns are dealt cyclically, so the
// remaining work stays spread over the lanes as the factorisation shrinks.
// CG = 1 is one row per lane; CG > 1 spreads each row over CG lanes, which
// lets CG lanes write the pivot row at once and each broadcast element feed
// R rows.
// ---------------------------------------------------------------------------
template <int N, int L, int CG>
struct Tile {
  static_assert(L >= 1 && L <= 64 && (L & (L - 1)) == 0, "L must be a power of two");
  static_assert(CG == 1 || CG == 2 || CG == 4, "CG must be 1, 2 or 4");
  static constexpr int RG = L / CG;
  static constexpr int R = (N + RG - 1) / RG;
  static constexpr int C = (N + CG - 1) / CG;
  static constexpr int NPAD = C * CG;
  static constexpr int G = 64 / L;
};

constexpr int quad_perm(int a, int b, int c, int d) {
  return a | (b << 2) | (c << 4) | (d << 6);
}

// Value held by lane `src` of this lane's lane-row (CG adjacent lanes).  The
// callers pass a constant once the step loop is unrolled, so the branches
// fold to one DPP move per dword.
template <int CG>
__device__ __forceinline__ double lanerow_bcast(double v, int src) {
  if constexpr (CG == 1) {
    return v;
  } else if constexpr (CG == 2) {
    return src == 0 ? dpp_d<quad_perm(0, 0, 2, 2)>(v) : dpp_d<quad_perm(1, 1, 3, 3)>(v);
  } else {
    switch (src) {
      case 0: return dpp_d<quad_perm(0, 0, 0, 0)>(v);
      case 1: return dpp_d<quad_perm(1, 1, 1, 1)>(v);
      case 2: return dpp_d<quad_perm(2, 2, 2, 2)>(v);
      default: return dpp_d<quad_perm(3, 3, 3, 3)>(v);
    }
  }
}

// Max over the CG lanes of a lane-row.
template <int CG>
__device__ __forceinline__ double lanerow_max(double v) {
  if constexpr (CG > 1) v = vmax_f64(v, dpp_d<kQuadXor1>(v));
  if constexpr (CG > 2) v = vmax_f64(v, dpp_d<kQuadXor2>(v));
  return v;
}

struct GroupResult {
  double mx, my;
  int e;
  int info;
};

// Pivot search of one step as a sequence of stages, so that it can be
// spread over the update of the previous step (see update_and_search).
//   stage 0                 criterion of this lane's candidates, best first
//                           (ties: the larger position)
//   stages 1 .. LOG         butterfly max of the criterion over the group
//   stage LOG+1             candidates equal to the max keep their position
//   stages LOG+2 .. 2LOG+1  butterfly max of those positions
// The result, the position of the exact arg-max with ties to the larger
// position, is in pm after the last stage.
struct SearchState {
  double cl;  // this lane's best criterion
  int pl;     // its position
  double cm;  // running group max of the criterion
  int pm;     // running group max of the candidate positions
};

constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v / 2); }

template <int L>
constexpr int search_stages() { return 2 * ilog2(L) + 2; }

// One butterfly level i (0-based) of the group max, for doubles and ints.
template <int L>
__device__ __forceinline__ double max_level_d(int i, double v) {
  switch (i) {
    case 0: return vmax_f64(v, dpp_d<kQuadXor1>(v));
    case 1: return vmax_f64(v, dpp_d<kQuadXor2>(v));
    case 2: return vmax_f64(v, dpp_d<kHalfMirror>(v));
    case 3: return vmax_f64(v, dpp_d<kMirror>(v));
    case 4: {
      const unsigned lo = __double2loint(v), hi = __double2hiint(v);
      const auto sl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
      const auto sh = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
      return vmax_f64(__hiloint2double(sh[0], sl[0]), __hiloint2double(sh[1], sl[1]));
    }
    default: {
      const unsigned lo = __double2loint(v), hi = __double2hiint(v);
      const auto sl = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
      const auto sh = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
      return vmax_f64(__hiloint2double(sh[0], sl[0]), __hiloint2double(sh[1], sl[1]));
    }
  }
}
template <int L>
__device__ __forceinline__ int max_level_i(int i, int v) {
  switch (i) {
    case 0: return max(v, dpp_i<kQuadXor1>(v));
    case 1: return max(v, dpp_i<kQuadXor2>(v));
    case 2: return max(v, dpp_i<kHalfMirror>(v));
    case 3: return max(v, dpp_i<kMirror>(v));
    case 4: {
      const auto s = __builtin_amdgcn_permlane16_swap((unsigned)v, (unsigned)v, false, false);
      return max((int)s[0], (int)s[1]);
    }
    default: {
      const auto s = __builtin_amdgcn_permlane32_swap((unsigned)v, (unsigned)v, false, false);
      return max((int)s[0], (int)s[1]);
    }
  }
}

// Tie key of a candidate position: CROUT_REF keeps the last position among
// exact ties (the reference's `>=` scan, LU-fqhe/src/LUCoreCPU.cu:207-212),
// ABS the first (LAPACK izamax's strict `>`); the search takes the largest
// key.
template <int POLICY>
__device__ __forceinline__ int tie_key(int pos) {
  return POLICY == POLICY_ABS ? 255 - pos : pos;
}

template <int L, int CG, int POLICY, int ARITH, int R, int C>
__device__ __forceinline__ void search_stage(int j, SearchState& st, const double2 (&a)[R][C],
                                             const int (&pos)[R], const bool (&act)[R],
                                             const double (&sc)[R], const double (&rs)[R],
                                             int cb, int k) {
  constexpr int LOG = ilog2(L);
  if (j == 0) {
    const int ck = k % CG, sk = k / CG;
    double cbest = -1.0;
    int pbest = -1;
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const double2 v = a[t][sk];
      const bool cand = act[t] && cb == ck;
      double c;
      if constexpr (POLICY == POLICY_CROUT_REF) {
        if constexpr (ARITH == ARITH_FUSED)
          c = quot_rn(fabs(v.x), sc[t], rs[t], cand);
        else
          c = fabs(v.x) / sc[t];
      } else {
        // |Re| + |Im| (LAPACK's cabs1): no overflow below DBL_MAX / 2 and
        // invariant under power-of-two scaling
        c = fabs(v.x) + fabs(v.y);
      }
      c = vmin_f64(c, kDblMax);  // NaN -> DBL_MAX: exactly one pivot always exists
      c = cand ? c : -1.0;
      const int key = tie_key<POLICY>(pos[t]);
      if (t == 0) {
        cbest = c;
        pbest = key;
      } else {
        const bool better = (c > cbest) || (c == cbest && key > pbest);
        cbest = better ? c : cbest;
        pbest = better ? key : pbest;
      }
    }
    st.cl = cbest;
    st.pl = pbest;
    st.cm = cbest;
  } else if (j <= LOG) {
    st.cm = max_level_d<L>(j - 1, st.cm);
  } else if (j == LOG + 1) {
    st.pm = (st.cl == st.cm) ? st.pl : -1;
  } else {
    st.pm = max_level_i<L>(j - LOG - 2, st.pm);
  }
}

__device__ __forceinline__ void pin(SearchState& st) {
  asm volatile("" : "+v"(st.cl), "+v"(st.pl), "+v"(st.cm), "+v"(st.pm)::"memory");
}

// v_fma_f64 with source negation: d = -a*b + c, and d = a*b + c.  Written
// out because the compiler's accumulating v_fmac form has no modifiers and
// would need negated copies of the multipliers.
__device__ __forceinline__ double fma_na(double a, double b, double c) {
  double d;
  asm("v_fma_f64 %0, -%1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}
__device__ __forceinline__ double fma_pa(double a, double b, double c) {
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}

// a[t][s] -= l_t * u_s for the column slots s0 .. s0+M-1 below C (4 FMAs per
// complex element).  In fused arithmetic the first FMA of every element
// (x - lx*ux, y - lx*uy) comes before the second (+ ly*uy, - ly*ux), so
// dependent FMAs are never adjacent.
template <int ARITH, int R, int C, int M>
__device__ __forceinline__ void update_slots(double2 (&a)[R][C], const double (&lx)[R],
                                             const double (&ly)[R], int s0,
                                             const double2 (&u)[M]) {
  if constexpr (ARITH == ARITH_FUSED) {
    double xs[M][R], ys[M][R];
#pragma unroll
    for (int c = 0; c < M; ++c) {
      if (s0 + c >= C) continue;
#pragma unroll
      for (int t = 0; t < R; ++t) {
        xs[c][t] = fma_na(lx[t], u[c].x, a[t][s0 + c].x);
        ys[c][t] = fma_na(lx[t], u[c].y, a[t][s0 + c].y);
      }
    }
#pragma unroll
    for (int c = 0; c < M; ++c) {
      if (s0 + c >= C) continue;
#pragma unroll
      for (int t = 0; t < R; ++t) {
        a[t][s0 + c].x = fma_pa(ly[t], u[c].y, xs[c][t]);
        a[t][s0 + c].y = fma_na(ly[t], u[c].x, ys[c][t]);
      }
    }
  } else {
#pragma unroll
    for (int c = 0; c < M; ++c) {
      if (s0 + c >= C) continue;
#pragma unroll
      for (int t = 0; t < R; ++t) {
        double2& e = a[t][s0 + c];
        const double tx = lx[t] * u[c].x - ly[t] * u[c].y;
        const double ty = ly[t] * u[c].x + lx[t] * u[c].y;
        e.x = e.x - tx;
        e.y = e.y - ty;
      }
    }
  }
}

template <int R, int C>
__device__ __forceinline__ void pin_slot(double2 (&a)[R][C], int s) {
#pragma unroll
  for (int t = 0; t < R; ++t) asm volatile("" : "+v"(a[t][s].x), "+v"(a[t][s].y)::"memory");
}

// The rank-1 update of step k (column slots s0 .. C-1, s0 = (k+1)/CG) with
// step k+1's pivot search woven into it.  Slot s0, which holds column k+1,
// is updated first; then the remaining slots stream from LDS in double-
// buffered chunks of CH slots, and after each chunk's FMAs come the next few
// search stages.  A pin after each chunk holds both the chunk's results and
// the search state, so the search's dependent chain of cross-lane moves
// sits between independent FMAs instead of stalling the wave, and the
// compiler cannot hoist every read of the step (which would hold 4*(n-k)
// extra registers).  Each element sees the same operations in the same
// order as in a plain step-by-step elimination.
template <int L, int CG, int POLICY, int ARITH, int R, int C>
__device__ __forceinline__ int update_and_search(
    double2 (&a)[R][C], const double (&lx)[R], const double (&ly)[R],
    const double2* __restrict__ bc, int cb, int k, const int (&pos)[R], const bool (&act)[R],
    const double (&sc)[R], const double (&rs)[R]) {
  constexpr int NS = search_stages<L>();
  constexpr int CH = 4;
  const int s0 = (k + 1) / CG;
  {
    const double2 u0[1] = {bc[cb + s0 * CG]};
    update_slots<ARITH>(a, lx, ly, s0, u0);
  }
  pin_slot(a, s0);

  SearchState st{-1.0, -1, -1.0, -1};
  const int s1 = s0 + 1;
  const int nch = (C - s1 + CH - 1) / CH;
  if (nch == 0) {
#pragma unroll
    for (int j = 0; j < NS; ++j) search_stage<L, CG, POLICY, ARITH>(j, st, a, pos, act, sc, rs, cb, k + 1);
    return tie_key<POLICY>(st.pm);
  }
  double2 ub[2][CH];
#pragma unroll
  for (int c = 0; c < CH; ++c)
    if (s1 + c < C) ub[0][c] = bc[cb + (s1 + c) * CG];
#pragma unroll
  for (int ci = 0; ci < nch; ++ci) {
    const int c0 = s1 + ci * CH;
    if (ci + 1 < nch) {
#pragma unroll
      for (int c = 0; c < CH; ++c)
        if (c0 + CH + c < C) ub[(ci + 1) & 1][c] = bc[cb + (c0 + CH + c) * CG];
    }
    update_slots<ARITH>(a, lx, ly, c0, ub[ci & 1]);
#pragma unroll
    for (int j = 0; j < NS; ++j)
      if (j * nch / NS == ci) search_stage<L, CG, POLICY, ARITH>(j, st, a, pos, act, sc, rs, cb, k + 1);
#pragma unroll
    for (int c = 0; c < CH; ++c)
      if (c0 + c < C) pin_slot(a, c0 + c);
    pin(st);
  }
  return tie_key<POLICY>(st.pm);
}

// Per row slot state of a lane: the row's reference position, whether it is
// still a pivot candidate, and its row scale s_i with RN(1/s_i).
template <int R>
struct RowSlots {
  int pos[R];
  bool act[R];
  double sc[R], rs[R];
};

// Running state of one factorisation: det = sg * (dx, dy) * 2^de, the first
// zero pivot's status, and the position of the next pivot.
struct StepState {
  double dx, dy, sg;
  int de, info, ppos;
};

// Compaction schedule.  Rows retire as they become pivots, so once the
// active rows fit in R-1 row slots per lane-row, the group moves them there
// (compact below): the remaining steps then update, search and select over
// R-1 slots instead of R.  Compaction happens where the active rows first
// fit, if at least kMinCompactSteps steps remain to pay for it.
constexpr int kMinCompactSteps = 6;

constexpr int compact_step(int N, int RG, int R) {
  return (R > 1 && (R - 1) * RG >= kMinCompactSteps) ? N - (R - 1) * RG : N;
}

// LDS budget of one wave's compactions, in double2 (20 KiB: eight waves of
// a CU then fit in its 160 KiB).  A compaction moves its columns in as many
// rounds as it takes to stay within the budget or within the pivot-row
// slots, whichever is larger.
constexpr int kCompactBudget = 1280;

// LDS a compaction at step K needs per group (double2 units) when it moves
// its columns in `rounds` rounds: per row the round's columns plus one
// padding element (rows on distinct banks), then the rows' scales and
// reciprocals and their positions.
constexpr int compact_lds_rounds(int N, int CG, int C, int K, int rounds) {
  const int rows = N - K;
  const int slots = C - K / CG;
  const int per = (slots + rounds - 1) / rounds;
  return rows * (per * CG + 1) + rows + (rows + 3) / 4;
}
constexpr int compact_rounds(int N, int L, int CG, int K) {
  const int C = (N + CG - 1) / CG;
  const int G = 64 / L;
  const int cap0 = kCompactBudget / G;
  const int np = C * CG + 1;
  const int cap = cap0 > np ? cap0 : np;
  const int slots = C - K / CG;
  int r = 1;
  while (r < slots && compact_lds_rounds(N, CG, C, K, r) > cap) ++r;
  return r;
}

// LDS per group: the pivot-row slot (NPAD + 1) or the largest compaction
// area of the schedule, whichever is larger (they are never in use at once).
constexpr int group_lds(int N, int L, int CG) {
  const int RG = L / CG;
  const int C = (N + CG - 1) / CG;
  int m = C * CG + 1;
  int R = (N + RG - 1) / RG;
  for (;;) {
    const int K = compact_step(N, RG, R);
    if (K >= N) break;
    const int need = compact_lds_rounds(N, CG, C, K, compact_rounds(N, L, CG, K));
    m = need > m ? need : m;
    R = (N - K + RG - 1) / RG;
  }
  return m | 1;  // odd: the groups' slots start on different LDS banks
}

// 64-bit mask of the lanes below `lane`.
__device__ __forceinline__ uint64_t lanes_below(int lane) {
  return lane >= 64 ? ~0ull : ((1ull << lane) - 1ull);
}

// Moves the group's active rows from R row slots into R2 < R: the active
// rows, ordered by (slot, lane-row), become rows j = 0 .. N-K-1 of the new
// layout, at lane-row j % RG, slot j / RG (the tile's cyclic deal again).
// Columns below K's slot boundary are finished and not moved.  The rows
// carry their positions and scales; every element keeps its bits.
template <int N, int L, int CG, int K, int R, int R2, int C>
__device__ __forceinline__ void compact(const double2 (&a)[R][C], const RowSlots<R>& rw,
                                        double2 (&b)[R2][C], RowSlots<R2>& rw2,
                                        double2* __restrict__ area, int gl, int g) {
  constexpr int RG = L / CG;
  constexpr int S0 = K / CG;
  constexpr int ROWS = N - K;
  constexpr int NR = compact_rounds(N, L, CG, K);
  constexpr int SR = (C - S0 + NR - 1) / NR;   // column slots per round
  constexpr int STRIDE = SR * CG + 1;
  static_assert(ROWS <= R2 * RG && ROWS > (R2 - 1) * RG, "compaction slot count");
  const int rb = gl / CG, cb = gl % CG;
  double2* meta = area + ROWS * STRIDE;                    // (sc, rs) per row
  int* posv = reinterpret_cast<int*>(meta + ROWS);          // position per row
  // new row index of each active slot: active rows of earlier slots, then
  // active lane-rows of this slot below this one (ballot over lane-column 0)
  const uint64_t gmask = (L == 64) ? ~0ull : (((1ull << L) - 1ull) << (g * L));
  const uint64_t below = lanes_below(g * L + rb * CG);
  int dst[R];
  int base = 0;
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const uint64_t m = __ballot(rw.act[t] && cb == 0) & gmask;
    dst[t] = base + __popcll(m & below);
    base += __popcll(m);
  }
#pragma unroll
  for (int rd = 0; rd < NR; ++rd) {
    const int s_lo = S0 + rd * SR;
    const int s_hi = (s_lo + SR < C) ? s_lo + SR : C;
#pragma unroll
    for (int t = 0; t < R; ++t) {
      if (rw.act[t]) {
#pragma unroll
        for (int s = s_lo; s < s_hi; ++s) area[dst[t] * STRIDE + cb + (s - s_lo) * CG] = a[t][s];
        if (rd == 0 && cb == 0) {
          meta[dst[t]] = make_double2(rw.sc[t], rw.rs[t]);
          posv[dst[t]] = rw.pos[t];
        }
      }
    }
    wave_lds_sync();
#pragma unroll
    for (int t2 = 0; t2 < R2; ++t2) {
      // slots past the last row load a copy of it: they are inactive, so
      // their (finite) contents never reach a pivot, a multiplier of an
      // active row or an output
      const int j = rb + t2 * RG;
      const bool live = j < ROWS;
      const int jr = live ? j : ROWS - 1;
#pragma unroll
      for (int s = s_lo; s < s_hi; ++s) b[t2][s] = area[jr * STRIDE + cb + (s - s_lo) * CG];
      if (rd == 0) {
        const double2 mr = meta[jr];
        rw2.sc[t2] = live ? mr.x : 1.0;
        rw2.rs[t2] = live ? mr.y : 1.0;
        rw2.pos[t2] = live ? posv[jr] : N + j;  // never a position: never the pivot nor k
        rw2.act[t2] = live;
      }
    }
    wave_lds_sync();
  }
}

// One step k of the factorisation over R row slots (see the file comment):
// pivot-row write and broadcast, det product, multipliers, the reference
// swap as position bookkeeping, and the update woven with step k+1's search.
template <int N, int L, int CG, int POLICY, int ARITH, int R, int C, int PR>
__device__ __forceinline__ void factor_step(int k, double2 (&a)[R][C], RowSlots<R>& rw,
                                            StepState& st, double2* __restrict__ bc, int gl,
                                            int (&piv_rec)[PR]) {
  using T = Tile<N, L, CG>;
  const int cb = gl % CG;
  const int ck = k % CG, sk = k / CG;
  bool is_piv[R];
#pragma unroll
  for (int t = 0; t < R; ++t) is_piv[t] = (rw.pos[t] == st.ppos);

  // The pivot's lane-row writes the pivot row (columns >= k) to LDS.  A
  // ds_write costs the same however few lanes are active, so:
  //  * one matrix per wave, or one row slot: one exec-masked region per row
  //    slot; only the pivot's region runs (empty regions are branched over);
  //  * several matrices per wave and several slots: their pivots sit in
  //    different row slots, so the pivot lanes first select their row's slot
  //    values and one region writes every group's pivot row at once.
  // The wave executes its LDS instructions in order, so the reads below see
  // these writes.
  if constexpr (T::G > 1 && R > 1) {
    bool any = false;
#pragma unroll
    for (int t = 0; t < R; ++t) any = any || is_piv[t];
    if (any) {
#pragma unroll
      for (int s = sk; s < C; ++s) {
        double2 v = a[0][s];
#pragma unroll
        for (int t = 1; t < R; ++t) v = sel2(is_piv[t], a[t][s], v);
        bc[cb + s * CG] = v;
      }
    }
  } else {
#pragma unroll
    for (int t = 0; t < R; ++t) {
      if (is_piv[t]) {
#pragma unroll
        for (int s = sk; s < C; ++s) bc[cb + s * CG] = a[t][s];
      }
    }
  }
  wave_lds_sync();
  const double2 p = bc[k];

  if (gl == k % L) piv_rec[k / L] = st.ppos;
  if (st.ppos != k) st.sg = -st.sg;

  // |p|^2 and the determinant product.  Fused arithmetic scales p by
  // 2^-kp, kp the exponent of max(|Re p|, |Im p|), so that p_s = p 2^-kp has
  // |p_s| in [0.5, 2^0.5): |p_s|^2 neither overflows nor underflows (it is 0
  // exactly when p == 0), and the running product of the p_s stays within
  // 2^+-64 of 1 without renormalisation (det = prod p_s * 2^(sum kp)).  The
  // scalings are exact, so in the normal range the multipliers have the
  // bits of the unscaled formula.  Reference arithmetic keeps the plain
  // product, renormalised every step.
  double den, psx = p.x, psy = p.y;
  int kp = 0;
  if constexpr (ARITH == ARITH_FUSED) {
    kp = __builtin_amdgcn_frexp_exp(vmax_abs_f64(p.x, p.y));  // 0 for 0, inf, nan
    psx = ldexp(p.x, -kp);
    psy = ldexp(p.y, -kp);
    den = fma(psx, psx, psy * psy);
    const double nx = fma(st.dx, psx, -(st.dy * psy));
    const double ny = fma(st.dx, psy, st.dy * psx);
    st.dx = nx;
    st.dy = ny;
    st.de += kp;
    // keep the exponent sum running: left alone, the compiler holds every
    // step's kp live and adds them at the end
    asm volatile("" : "+v"(st.de));
  } else {
    den = p.x * p.x + p.y * p.y;
    const double nx = st.dx * p.x - st.dy * p.y;
    const double ny = st.dy * p.x + st.dx * p.y;
    st.dx = nx;
    st.dy = ny;
    renorm(st.dx, st.dy, st.de);
  }
  const bool zero = (den == 0.0);
  if (zero && st.info == 0) st.info = k + 1;
  // the same for the sign and the status
  asm volatile("" : "+v"(st.sg), "+v"(st.info));

  // multipliers l_t = a_tk / p of this lane's rows, column k broadcast
  // along the lane-row
  double lx[R], ly[R];
  if (k + 1 < N) {
    double ir = 0.0, ii = 0.0;
    if constexpr (ARITH == ARITH_FUSED) {
      // 1/p = conj(p_s) / |p_s|^2 * 2^-kp; a zero pivot gets exactly zero
      // multipliers: the step changes no row
      const double rden = recip_scaled(den);
      ir = ldexp(psx * rden, -kp);
      ii = ldexp(-(psy * rden), -kp);
    }
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const double ax = lanerow_bcast<CG>(a[t][sk].x, ck);
      const double ay = lanerow_bcast<CG>(a[t][sk].y, ck);
      if constexpr (ARITH == ARITH_FUSED) {
        lx[t] = fma(ax, ir, -(ay * ii));
        ly[t] = fma(ay, ir, ax * ii);
      } else {
        lx[t] = zero ? 0.0 : (ax * p.x + ay * p.y) / den;
        ly[t] = zero ? 0.0 : (ay * p.x - ax * p.y) / den;
      }
    }
  }

  // the reference's swap, as position bookkeeping
#pragma unroll
  for (int t = 0; t < R; ++t) {
    if (rw.pos[t] == k) rw.pos[t] = st.ppos;
    if (is_piv[t]) {
      rw.pos[t] = k;
      rw.act[t] = false;
    }
  }

  if (k + 1 < N) {
    // Every row slot is updated, with no divergent branch.  Finished rows
    // and the pivot row are never read again after this step, so what the
    // update leaves in them does not matter.  After a zero pivot the
    // multipliers are exactly zero, and fma(-0, u, e) == e for finite u up
    // to the sign of a zero element, which no output sees.  (An infinite
    // entry in such a step's pivot row or column turns into NaN here, where
    // the reference-order elimination skips the step.)
    st.ppos = update_and_search<L, CG, POLICY, ARITH>(a, lx, ly, bc, cb, k, rw.pos, rw.act,
                                                      rw.sc, rw.rs);
  }
  wave_lds_sync();
}

// Steps K0 .. N-1 with the tile in R row slots, compacting to fewer slots
// at the schedule's next step (compact_step).
template <int N, int L, int CG, int POLICY, int ARITH, int K0, int R, int C, int PR>
__device__ __forceinline__ void factor_phase(double2 (&a)[R][C], RowSlots<R>& rw, StepState& st,
                                             double2* __restrict__ bc, int gl, int g,
                                             int (&piv_rec)[PR]) {
  constexpr int RG = L / CG;
  constexpr int K1c = compact_step(N, RG, R);
  constexpr int K1 = K1c > K0 ? K1c : N;
#pragma unroll
  for (int k = K0; k < K1; ++k) factor_step<N, L, CG, POLICY, ARITH>(k, a, rw, st, bc, gl, piv_rec);
  if constexpr (K1 < N) {
    constexpr int R2 = (N - K1 + RG - 1) / RG;
    double2 b[R2][C];
    RowSlots<R2> rw2;
    compact<N, L, CG, K1>(a, rw, b, rw2, bc, gl, g);
    factor_phase<N, L, CG, POLICY, ARITH, K1>(b, rw2, st, bc, gl, g, piv_rec);
  }
}

// Factorisation of one matrix held by a group as a[R][C] (see Tile).  Rows
// never move between lanes except at compactions; each row slot carries its
// reference position.  piv_rec[q] receives the pivot position chosen at
// step gl + q*L.  bc: the group's LDS area (group_lds elements).
template <int N, int L, int CG, int POLICY, int ARITH, int R, int C, int PR>
__device__ __forceinline__ GroupResult factor_tile(double2 (&a)[R][C], int gl, int g,
                                                   double2* __restrict__ bc,
                                                   int (&piv_rec)[PR]) {
  using T = Tile<N, L, CG>;
  static_assert(R == T::R && C == T::C && PR == (N + L - 1) / L, "tile shape");
  constexpr int RG = T::RG;
  const int rb = gl / CG, cb = gl % CG;

  RowSlots<R> rw;
#pragma unroll
  for (int t = 0; t < R; ++t) {
    rw.pos[t] = rb + t * RG;
    rw.act[t] = rw.pos[t] < N;
  }

  // Row scale s_i = max_j |Re a_ij| (LU-fqhe/src/LUCoreCPU.cu:145-160).
  int rej = -1;
#pragma unroll
  for (int t = 0; t < R; ++t) {
    rw.sc[t] = 1.0;
    rw.rs[t] = 1.0;
  }
  if constexpr (POLICY == POLICY_CROUT_REF) {
    int worst = -1;  // 255 - (first rejected row) over this lane's slots
#pragma unroll
    for (int t = 0; t < R; ++t) {
      double m = fabs(a[t][0].x);
#pragma unroll
      for (int s = 1; s < C; ++s)
        if (cb + s * CG < N) m = vmax_f64(m, fabs(a[t][s].x));
      m = lanerow_max<CG>(m);
      const bool bad = rw.act[t] && m < kRowScaleMin;
      if (bad) worst = max(worst, 255 - rw.pos[t]);
      rw.sc[t] = (rw.act[t] && !bad) ? m : 1.0;
      // RN(1/s) for quot_rn; NaN sends a row scale above 2^960 (whose
      // reciprocal would not be a normal number) to the division
      rw.rs[t] = rw.sc[t] <= 0x1p960 ? 1.0 / rw.sc[t] : __builtin_nan("");
    }
    const int first_bad = group_max_i<L>(worst);
    if (first_bad >= 0) rej = 255 - first_bad;
  }

  StepState st{1.0, 0.0, 1.0, 0, 0, -1};
  {
    SearchState ss{-1.0, -1, -1.0, -1};
#pragma unroll
    for (int j = 0; j < search_stages<L>(); ++j)
      search_stage<L, CG, POLICY, ARITH>(j, ss, a, rw.pos, rw.act, rw.sc, rw.rs, cb, 0);
    st.ppos = tie_key<POLICY>(ss.pm);
  }

  factor_phase<N, L, CG, POLICY, ARITH, 0>(a, rw, st, bc, gl, g, piv_rec);

  if constexpr (ARITH == ARITH_FUSED) renorm(st.dx, st.dy, st.de);
  GroupResult res;
  res.mx = st.sg * st.dx;
  res.my = st.sg * st.dy;
  res.e = st.de;
  res.info = (rej >= 0) ? -(rej + 1) : st.info;
  return res;
}

__device__ __forceinline__ void write_result(const DetOut& out, int64_t m,
                                             const GroupResult& res) {
  double mx = res.mx, my = res.my, la, ph;
  int e = res.e;
  if (res.info < 0) {
    mx = my = __builtin_nan("");
    e = 0;
    la = ph = __builtin_nan("");
  } else if (mx == 0.0 && my == 0.0) {
    la = -__builtin_inf();
    ph = 0.0;
    mx = my = 0.0;
    e = 0;
  } else {
    la = 0.5 * log(fma(mx, mx, my * my)) + (double)e * kLn2;
    ph = atan2(my, mx);
  }
  out.mant[m] = make_double2(mx, my);
  if (out.expo) out.expo[m] = e;
  if (out.logabs) out.logabs[m] = la;
  if (out.phase) out.phase[m] = ph;
  if (out.info) out.info[m] = res.info;
}

// ---------------------------------------------------------------------------
// Batch kernel.  A task is one wave's G matrices; matrix i of the launch is
//   uniform addressing : A + i*stride + r*ld            (offs == nullptr)
//   indexed addressing : A + offs[i] + r*N              (mixed-size buckets)
// optionally with its rows gathered through row_perm (row r of matrix i is
// source row p[r], p = row_perm + perm_offs[i] or row_perm + i*N), and its
// results written to output slot out_idx[i] (or i).  One task per wave: a loop over tasks would keep every
// step's loop-invariant addresses live across iterations and double the
// register footprint.
// ---------------------------------------------------------------------------
enum Source { SRC_MATRIX = 0, SRC_FQHE = 1 };

struct BatchArgs {
  const double2* A;  // SRC_MATRIX: the matrices; SRC_FQHE: particle positions
  int64_t count;
  int64_t ld, stride;
  const int64_t* offs;
  const int64_t* out_idx;
  const uint8_t* row_perm;
  const int64_t* perm_offs;  // row perm of matrix i at row_perm + perm_offs[i] (else + i*N)
  int32_t piv_stride;
  int32_t src;
  DetOut out;
};

// FQHE wavefunction matrix row (LU-fqhe's physics; the reference builds its
// matrices on the host or in the disabled PERM builder,
// LU-fqhe/src/mkMatrixElements.cu:169-358): A_ij = z_i^j * exp(-|z_i|^2/4).
// Powers are a sequential product z^j = z^(j-1) * z from z^0 = 1 and every
// entry is computed the same way whatever the launch shape, so a matrix
// built in a kernel's load phase equals the one fqhe_build_kernel writes.
__device__ __forceinline__ double fqhe_gauss(double2 z) {
  return exp(-0.25 * fma(z.x, z.x, z.y * z.y));
}
__device__ __forceinline__ double2 cmul_ref(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.y * b.x + a.x * b.y);
}

// One task: the G matrices task*G .. task*G + G-1 of `args`, factored by one
// wave whose LDS broadcast area is `bc_wave` (G slots of NPAD+1 elements).
// The caller has checked that the task holds at least one matrix.
template <int N, int L, int CG, int POLICY, int ARITH>
__device__ __forceinline__ void det_task(const BatchArgs& args, int64_t task,
                                         double2* __restrict__ bc_wave, int lane) {
  using T = Tile<N, L, CG>;
  constexpr int R = T::R, C = T::C, RG = T::RG, G = T::G;
  constexpr int NP = group_lds(N, L, CG);
  constexpr int PR = (N + L - 1) / L;
  const int g = lane / L;
  const int gl = lane % L;
  const int rb = gl / CG, cb = gl % CG;
  double2* bc = bc_wave + g * NP;
  const int64_t i = task * G + g;
  const bool mat_ok = i < args.count;

  double2 a[R][C];
  const double2* __restrict__ base = nullptr;
  if (mat_ok) base = args.offs ? args.A + args.offs[i] : args.A + i * args.stride;
  const int64_t ld = args.offs ? (int64_t)N : args.ld;
  const uint8_t* __restrict__ perm = nullptr;
  if (mat_ok && args.row_perm)
    perm = args.row_perm + (args.perm_offs ? args.perm_offs[i] : i * N);
  int bad_perm = 0;  // a permutation entry >= N: the matrix gets kInfoBadPerm
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const int row = rb + t * RG;
    if (perm && row < N && perm[row] >= N) bad_perm = 1;
  }
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const int row = rb + t * RG;
    if (args.src == SRC_FQHE && mat_ok && row < N) {
      // positions: matrix i's z at A + offs[i] (or i*stride), one per row
      int src = row;
      if (perm) src = min((int)perm[row], N - 1);
      const double2 z = base[src];
      const double g = fqhe_gauss(z);
      double2 pw = make_double2(1.0, 0.0);
#pragma unroll
      for (int s = 0; s < C; ++s) a[t][s] = make_double2(0.0, 0.0);
#pragma unroll
      for (int j = 0; j < N; ++j) {
        if (j % CG == cb) a[t][j / CG] = make_double2(pw.x * g, pw.y * g);
        pw = cmul_ref(pw, z);
      }
    } else if (mat_ok && row < N) {
      int src = row;
      if (perm) src = min((int)perm[row], N - 1);
      const double2* __restrict__ rp = base + (int64_t)src * ld;
#pragma unroll
      for (int s = 0; s < C; ++s) {
        const int col = cb + s * CG;
        a[t][s] = (col < N) ? rp[col] : make_double2(0.0, 0.0);
      }
    } else {
#pragma unroll
      for (int s = 0; s < C; ++s)
        a[t][s] = make_double2((cb + s * CG) == row ? 1.0 : 0.0, 0.0);
    }
  }

  int piv_rec[PR];
#pragma unroll
  for (int q = 0; q < PR; ++q) piv_rec[q] = -1;
  GroupResult res = factor_tile<N, L, CG, POLICY, ARITH>(a, gl, g, bc, piv_rec);
  if (args.row_perm && group_max_i<L>(bad_perm) > 0) res.info = kInfoBadPerm;

  if (mat_ok) {
    const int64_t o = args.out_idx ? args.out_idx[i] : i;
    if (args.out.piv) {
#pragma unroll
      for (int q = 0; q < PR; ++q) {
        const int k = gl + q * L;
        if (k < N)
          args.out.piv[o * args.piv_stride + k] = (int8_t)(res.info < 0 ? -1 : piv_rec[q]);
      }
    }
    if (gl == 0) write_result(args.out, o, res);
  }
}

template <int N, int L, int CG, int POLICY, int ARITH, int WPB, int OCC>
__global__ void __launch_bounds__(64 * WPB, OCC) lup_det_kernel(BatchArgs args) {
  using T = Tile<N, L, CG>;
  constexpr int G = T::G;
  constexpr int NP = group_lds(N, L, CG);
  __shared__ double2 bcast[WPB * G * NP];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t task = (int64_t)blockIdx.x * WPB + wave;
  if (task * G >= args.count) return;  // whole wave idle (uniform branch)
  det_task<N, L, CG, POLICY, ARITH>(args, task, bcast + wave * G * NP, lane);
}

// Candidate launch shapes per n, as (lanes per matrix L, lane-columns CG,
// waves per SIMD the register allocation must allow).  Every candidate is
// compiled; the autotuner picks one per n on the device and candidate 0 is
// the default before tuning.  The lists come from shape sweeps on MI355X:
// small n wants many matrices per wave (even two rows per lane), mid n a
// split of each row over lane-columns, n > 40 one matrix per wave.  Up to
// n = 40 the natural allocation sits just under or just over 256 VGPRs, so
// those candidates are capped at 256 (two waves per SIMD).  From n = 41 the
// tile alone needs 100-256 VGPRs and the natural allocation leaves one wave
// per SIMD, which cannot hide the step's dependent latencies; the waves = 2
// candidates cap the allocation at 256 registers (spilling the coldest
// values to scratch), which measured up to 1.6x faster at n = 48.
struct ShapeLC {
  int lanes;
  int lane_cols;
  int waves;
};
constexpr int kMaxShapes = 4;
constexpr ShapeLC candidate_shape(int n, int idx) {
  constexpr ShapeLC none{0, 0, 0};
  if (n <= 8) {
    constexpr ShapeLC c[] = {{4, 1, 1}, {8, 1, 1}, none, none};
    return c[idx];
  }
  if (n <= 16) {
    constexpr ShapeLC c[] = {{16, 1, 1}, {8, 1, 1}, none, none};
    return c[idx];
  }
  if (n <= 24) {
    constexpr ShapeLC c[] = {{32, 4, 2}, {16, 2, 2}, {32, 2, 2}, none};
    return c[idx];
  }
  if (n <= 32) {
    constexpr ShapeLC c[] = {{32, 2, 2}, {32, 4, 2}, {32, 1, 2}, none};
    return c[idx];
  }
  if (n <= 40) {
    constexpr ShapeLC c[] = {{64, 4, 2}, {64, 1, 2}, {64, 2, 2}, none};
    return c[idx];
  }
  if (n <= 48) {
    constexpr ShapeLC c[] = {{64, 4, 2}, {64, 2, 2}, {64, 4, 1}, {32, 4, 1}};
    return c[idx];
  }
  constexpr ShapeLC c[] = {{64, 2, 1}, {64, 4, 1}, {64, 4, 2}, {64, 2, 2}};
  return c[idx];
}

}  // namespace lufqhe
