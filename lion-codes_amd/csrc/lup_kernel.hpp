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
//   4. det *= pivot (complex, renormalised to mant * 2^exp);
//   5. multipliers l = a_ik / p (lane-row DPP broadcast of column k) and the
//      rank-1 update of columns k+1..n-1 (complex FP64, 4 FMAs per element in
//      ARITH_FUSED);
//   6. the position bookkeeping of the reference swap; the pivot row retires.
// The k loop is fully unrolled (N is a template constant) so that finished
// column slots are skipped statically and the tile stays in registers.  The
// load phase reads the matrix from HBM, or builds an FQHE matrix from
// particle positions (SRC_FQHE).
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
      c = fmin(c, kDblMax);  // NaN -> DBL_MAX: exactly one pivot always exists
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

// a[t][s] -= l_t * u_s for one column slot s (4 FMAs per complex element).
template <int ARITH, int R, int C>
__device__ __forceinline__ void update_slot(double2 (&a)[R][C], const double (&lx)[R],
                                            const double (&ly)[R], int s, double2 u) {
#pragma unroll
  for (int t = 0; t < R; ++t) {
    double2& e = a[t][s];
    if constexpr (ARITH == ARITH_FUSED) {
      double x = fma(-lx[t], u.x, e.x);
      x = fma(ly[t], u.y, x);
      double y = fma(-lx[t], u.y, e.y);
      y = fma(-ly[t], u.x, y);
      e.x = x;
      e.y = y;
    } else {
      const double tx = lx[t] * u.x - ly[t] * u.y;
      const double ty = ly[t] * u.x + lx[t] * u.y;
      e.x = e.x - tx;
      e.y = e.y - ty;
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
  update_slot<ARITH>(a, lx, ly, s0, bc[cb + s0 * CG]);
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
#pragma unroll
    for (int c = 0; c < CH; ++c)
      if (c0 + c < C) update_slot<ARITH>(a, lx, ly, c0 + c, ub[ci & 1][c]);
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

// Factorisation of one matrix held by a group as a[R][C] (see Tile).  Rows
// never move; each row slot carries its reference position.  piv_rec[q]
// receives the pivot position chosen at step gl + q*L.
template <int N, int L, int CG, int POLICY, int ARITH, int R, int C, int PR>
__device__ __forceinline__ GroupResult factor_tile(double2 (&a)[R][C], int gl,
                                                   double2* __restrict__ bc,
                                                   int (&piv_rec)[PR]) {
  using T = Tile<N, L, CG>;
  static_assert(R == T::R && C == T::C && PR == (N + L - 1) / L, "tile shape");
  constexpr int RG = T::RG;
  const int rb = gl / CG, cb = gl % CG;

  int pos[R];
  bool act[R];
#pragma unroll
  for (int t = 0; t < R; ++t) {
    pos[t] = rb + t * RG;
    act[t] = pos[t] < N;
  }

  // Row scale s_i = max_j |Re a_ij| (LU-fqhe/src/LUCoreCPU.cu:145-160).
  double sc[R], rs[R];
  int rej = -1;
#pragma unroll
  for (int t = 0; t < R; ++t) {
    sc[t] = 1.0;
    rs[t] = 1.0;
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
      const bool bad = act[t] && m < kRowScaleMin;
      if (bad) worst = max(worst, 255 - pos[t]);
      sc[t] = (act[t] && !bad) ? m : 1.0;
      // RN(1/s) for quot_rn; NaN sends a row scale above 2^960 (whose
      // reciprocal would not be a normal number) to the division
      rs[t] = sc[t] <= 0x1p960 ? 1.0 / sc[t] : __builtin_nan("");
    }
    const int first_bad = group_max_i<L>(worst);
    if (first_bad >= 0) rej = 255 - first_bad;
  }

  double dx = 1.0, dy = 0.0, sg = 1.0;
  int de = 0, info = 0;
  int ppos;
  {
    SearchState st{-1.0, -1, -1.0, -1};
#pragma unroll
    for (int j = 0; j < search_stages<L>(); ++j)
      search_stage<L, CG, POLICY, ARITH>(j, st, a, pos, act, sc, rs, cb, 0);
    ppos = tie_key<POLICY>(st.pm);
  }

#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int ck = k % CG, sk = k / CG;
    bool is_piv[R];
#pragma unroll
    for (int t = 0; t < R; ++t) is_piv[t] = (pos[t] == ppos);

    // The pivot's lane-row writes the pivot row (columns >= k) to LDS.  A
    // ds_write costs the same however few lanes are active, so:
    //  * one matrix per wave: one exec-masked region per row slot; only the
    //    pivot's region runs (empty regions are branched over);
    //  * several matrices per wave: their pivots sit in different row slots,
    //    so the pivot lanes first select their row's slot values and one
    //    region writes every group's pivot row at once.
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

    if (gl == k % L) piv_rec[k / L] = ppos;
    if (ppos != k) sg = -sg;

    // determinant product
    {
      double nx, ny;
      if constexpr (ARITH == ARITH_FUSED) {
        nx = fma(dx, p.x, -(dy * p.y));
        ny = fma(dx, p.y, dy * p.x);
      } else {
        nx = dx * p.x - dy * p.y;
        ny = dy * p.x + dx * p.y;
      }
      dx = nx;
      dy = ny;
      renorm(dx, dy, de);
    }

    // |p|^2.  Fused arithmetic first scales p by 2^-kp, kp the exponent of
    // max(|Re p|, |Im p|), so that |p|^2 neither overflows nor underflows:
    // it is 0 exactly when p == 0.  The scaling is exact and the reciprocal
    // is scaled back below, so in the normal range the multipliers have the
    // bits of the unscaled formula.
    double den, psx = p.x, psy = p.y;
    int kp = 0;
    if constexpr (ARITH == ARITH_FUSED) {
      kp = __builtin_amdgcn_frexp_exp(vmax_abs_f64(p.x, p.y));  // 0 for 0, inf, nan
      psx = ldexp(p.x, -kp);
      psy = ldexp(p.y, -kp);
      den = fma(psx, psx, psy * psy);
    } else {
      den = p.x * p.x + p.y * p.y;
    }
    const bool zero = (den == 0.0);
    if (zero && info == 0) info = k + 1;

    // multipliers l_t = a_tk / p of this lane's rows, column k broadcast
    // along the lane-row
    double lx[R], ly[R];
    if (k + 1 < N) {
      double ir = 0.0, ii = 0.0;
      if constexpr (ARITH == ARITH_FUSED) {
        // 1/p = conj(p_s) / |p_s|^2 * 2^-kp; a zero pivot gets exactly zero
        // multipliers: the step changes no row
        const double rden = zero ? 0.0 : 1.0 / den;
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
      if (pos[t] == k) pos[t] = ppos;
      if (is_piv[t]) {
        pos[t] = k;
        act[t] = false;
      }
    }

    if (k + 1 < N) {
      // Every row slot is updated, with no divergent branch.  Finished rows
      // and the pivot row are never read again after this step, so what the
      // update leaves in them does not matter.  After a zero pivot the
      // multipliers are exactly zero (ir = ii = 0 in fused arithmetic, an
      // explicit zero in reference arithmetic), and fma(-0, u, e) == e for
      // finite u up to the sign of a zero element, which no output sees.
      // (An infinite entry in such a step's pivot row or column turns into
      // NaN here, where the reference-order elimination skips the step.)
      ppos = update_and_search<L, CG, POLICY, ARITH>(a, lx, ly, bc, cb, k, pos, act, sc, rs);
    }
    wave_lds_sync();
  }

  GroupResult res;
  res.mx = sg * dx;
  res.my = sg * dy;
  res.e = de;
  res.info = (rej >= 0) ? -(rej + 1) : info;
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
  constexpr int NP = T::NPAD + 1;  // padded LDS slot: groups on distinct banks
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
  GroupResult res = factor_tile<N, L, CG, POLICY, ARITH>(a, gl, bc, piv_rec);
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
  constexpr int NP = T::NPAD + 1;
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
// split of each row over lane-columns, n > 40 one matrix per wave.  From
// n = 41 the tile alone needs 100-256 VGPRs and the natural allocation
// leaves one wave per SIMD, which cannot hide the step's dependent
// latencies; the waves = 2 candidates cap the allocation at 256 registers
// (spilling the coldest values to scratch), which measured up to 1.6x
// faster at n = 48.
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
    constexpr ShapeLC c[] = {{32, 4, 1}, {16, 2, 1}, {32, 2, 1}, none};
    return c[idx];
  }
  if (n <= 32) {
    constexpr ShapeLC c[] = {{32, 2, 1}, {32, 4, 1}, {32, 1, 1}, none};
    return c[idx];
  }
  if (n <= 40) {
    constexpr ShapeLC c[] = {{64, 4, 1}, {64, 1, 1}, {64, 4, 2}, none};
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
