// lup_kernel.hpp -- batched LUP determinant kernels for gfx950 (CDNA4).
//
// Replaces LUCoreGPU::LUDecomposition (LU-fqhe/src/LUCoreGPU.cu:159-437) and
// its reductions vOP::* (LU-fqhe/src/vectorOP.cu:23-372).
//
// Layout: one matrix row per lane, the whole row resident in VGPRs.  A wave64
// is split into G = 64/L groups of L lanes (L = 8, 16, 32 or 64, L >= n); group
// g of the wave factors one matrix, lane r of the group owns row r.  Rows are
// never moved: a lane keeps its row for the whole factorisation and carries
// the row's *reference position* `pos`, which is what the reference's physical
// row swaps (LU-fqhe/src/LUCoreCPU.cu:215-232) would have made of it.
//
// Step k of the right-looking factorisation:
//   1. criterion c on column k (per policy), -1 on finished rows;
//   2. exact pivot choice: cmax = group max of c, then the largest `pos`
//      among lanes with c == cmax (the reference's `>=` scan keeps the last
//      position among exact ties, LU-fqhe/src/LUCoreCPU.cu:207-212);
//      butterflies use DPP (quad_perm, row_half_mirror, row_mirror) and the
//      gfx950 v_permlane16_swap / v_permlane32_swap across 16-lane rows;
//   3. the pivot lane writes its row (columns k..n-1) to the group's LDS slot,
//      every lane reads it back (LDS broadcast);
//   4. det *= pivot (complex, renormalised to mant * 2^exp);
//   5. remaining rows form their multiplier and apply the rank-1 update to
//      columns k+1..n-1 (complex FP64, 4 FMAs per element in ARITH_FUSED);
//   6. the position bookkeeping of the reference swap; the pivot lane retires.
// The k loop is fully unrolled (N is a template constant) so that done
// columns are skipped statically and the row array stays in registers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lufqhe {

enum Policy { POLICY_CROUT_REF = 0, POLICY_ABS = 1 };
enum Arith { ARITH_FUSED = 0, ARITH_REF = 1 };

constexpr double kDblMax = 1.7976931348623157e308;
constexpr double kRowScaleMin = 1e-24;  // LU-fqhe/src/LUCoreCPU.cu:158
constexpr double kLn2 = 0.69314718055994530942;

// Output pointers of one launch.  Any of them may be null except mant.
struct DetOut {
  double2* mant;     // det = mant * 2^exp, max(|re|,|im|) of mant in [0.5,1)
  int32_t* expo;
  double* logabs;    // log|det|
  double* phase;     // arg det in (-pi, pi]
  int32_t* info;     // 0, k+1 (zero pivot at step k), -(i+1) (row i rejected)
  int8_t* piv;       // n bytes per matrix: reference position swapped at step k
};

// ---------------------------------------------------------------------------
// cross-lane helpers
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
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

// mant * 2^e renormalisation, exact (power-of-two scaling).
__device__ __forceinline__ void renorm(double& x, double& y, int& e) {
  const double m = fmax(fabs(x), fabs(y));
  const int k = __builtin_amdgcn_frexp_exp(m);  // 0 for 0, inf, nan
  x = ldexp(x, -k);
  y = ldexp(y, -k);
  e += k;
}

// ---------------------------------------------------------------------------
// One group's factorisation of one matrix whose row r is in a[0..N-1].
// Shared by the uniform kernel and the mixed-size kernels.
// ---------------------------------------------------------------------------
struct GroupResult {
  double mx, my;
  int e;
  int info;
  int my_piv;  // the pivot position recorded at step k == r
};

template <int N, int L, int POLICY, int ARITH>
__device__ __forceinline__ GroupResult factor_rows(double2 (&a)[N], int r,
                                                   bool row_live,
                                                   double2* __restrict__ bc) {
  // Row scale s_i = max_j |Re a_ij| (LU-fqhe/src/LUCoreCPU.cu:145-160).
  double s = 1.0, rs = 1.0;
  int rej = -1;
  if constexpr (POLICY == POLICY_CROUT_REF) {
    s = fabs(a[0].x);
#pragma unroll
    for (int j = 1; j < N; ++j) s = fmax(s, fabs(a[j].x));
    const bool bad = row_live && (s < kRowScaleMin);
    const int first_bad = group_max_i<L>(bad ? (63 - r) : -1);
    if (first_bad >= 0) rej = 63 - first_bad;
    if (!row_live || bad) s = 1.0;
    rs = 1.0 / s;
  }

  bool act = row_live;
  int pos = r;
  double dx = 1.0, dy = 0.0, sg = 1.0;
  int de = 0, info = 0, my_piv = -1;

#pragma unroll
  for (int k = 0; k < N; ++k) {
    // 1. criterion
    double c;
    if constexpr (POLICY == POLICY_CROUT_REF) {
      if constexpr (ARITH == ARITH_FUSED)
        c = fabs(a[k].x) * rs;
      else
        c = fabs(a[k].x) / s;
    } else {
      if constexpr (ARITH == ARITH_FUSED)
        c = fma(a[k].x, a[k].x, a[k].y * a[k].y);
      else
        c = a[k].x * a[k].x + a[k].y * a[k].y;
    }
    c = fmin(c, kDblMax);  // NaN -> DBL_MAX: exactly one pivot always exists
    c = act ? c : -1.0;

    // 2. exact arg-max, ties to the larger reference position
    const double cmax = group_max_d<L>(c);
    const int ppos = group_max_i<L>(c == cmax ? pos : -1);
    const bool is_piv = (pos == ppos);

    // 3. LDS broadcast of the pivot row
    if (is_piv) {
#pragma unroll
      for (int j = k; j < N; ++j) bc[j] = a[j];
    }
    wave_lds_sync();
    const double2 p = bc[k];

    if (r == k) my_piv = ppos;
    if (ppos != k) sg = -sg;

    // 4. determinant product
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

    // 5. multipliers and rank-1 update
    double den;
    if constexpr (ARITH == ARITH_FUSED)
      den = fma(p.x, p.x, p.y * p.y);
    else
      den = p.x * p.x + p.y * p.y;
    const bool zero = (den == 0.0);
    if (zero && info == 0) info = k + 1;
    if (k + 1 < N && act && !is_piv && !zero) {
      double lx, ly;
      if constexpr (ARITH == ARITH_FUSED) {
        const double rden = 1.0 / den;
        const double ir = p.x * rden;
        const double ii = -(p.y * rden);
        lx = fma(a[k].x, ir, -(a[k].y * ii));
        ly = fma(a[k].y, ir, a[k].x * ii);
      } else {
        lx = (a[k].x * p.x + a[k].y * p.y) / den;
        ly = (a[k].y * p.x - a[k].x * p.y) / den;
      }
#pragma unroll
      for (int j = k + 1; j < N; ++j) {
        const double2 u = bc[j];
        if constexpr (ARITH == ARITH_FUSED) {
          double x = fma(-lx, u.x, a[j].x);
          x = fma(ly, u.y, x);
          double y = fma(-lx, u.y, a[j].y);
          y = fma(-ly, u.x, y);
          a[j].x = x;
          a[j].y = y;
        } else {
          const double tx = lx * u.x - ly * u.y;
          const double ty = ly * u.x + lx * u.y;
          a[j].x = a[j].x - tx;
          a[j].y = a[j].y - ty;
        }
      }
    }
    wave_lds_sync();

    // 6. the reference's swap, as position bookkeeping
    if (pos == k) pos = ppos;
    if (is_piv) {
      pos = k;
      act = false;
    }
  }

  GroupResult res;
  res.mx = sg * dx;
  res.my = sg * dy;
  res.e = de;
  res.info = (rej >= 0) ? -(rej + 1) : info;
  res.my_piv = (rej >= 0) ? -1 : my_piv;
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
// source row row_perm[i*N + r]), and its results written to output slot
// out_idx[i] (or i).  With next_task != nullptr the grid is persistent: after
// its first task a wave takes task grid_tasks + atomicAdd(next_task, 1) until
// the tasks run out, so every wave reaches the exit.
// ---------------------------------------------------------------------------
struct BatchArgs {
  const double2* A;
  int64_t count;
  int64_t ld, stride;
  const int64_t* offs;
  const int64_t* out_idx;
  const uint8_t* row_perm;
  int32_t* next_task;
  int64_t grid_tasks;
  int32_t piv_stride;
  DetOut out;
};

template <int N, int L, int POLICY, int ARITH, int WPB>
__global__ void __launch_bounds__(64 * WPB) lup_det_kernel(BatchArgs args) {
  static_assert(L >= N && L <= 64 && (L & (L - 1)) == 0, "bad group size");
  constexpr int G = 64 / L;
  constexpr int NP = N + 1;  // padded LDS slot: groups land on distinct banks
  __shared__ double2 bcast[WPB * G * NP];

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane / L;
  const int r = lane % L;
  double2* bc = bcast + (wave * G + g) * NP;
  const int64_t ntasks = (args.count + G - 1) / G;

  int64_t task = (int64_t)blockIdx.x * WPB + wave;
  while (task < ntasks) {
    const int64_t i = task * G + g;
    const bool mat_ok = i < args.count;
    const bool row_live = mat_ok && r < N;

    double2 a[N];
    if (row_live) {
      int src = r;
      if (args.row_perm) src = min((int)args.row_perm[i * N + r], N - 1);
      const double2* __restrict__ row =
          args.offs ? args.A + args.offs[i] + (int64_t)src * N
                    : args.A + i * args.stride + (int64_t)src * args.ld;
#pragma unroll
      for (int j = 0; j < N; ++j) a[j] = row[j];
    } else {
#pragma unroll
      for (int j = 0; j < N; ++j) a[j] = make_double2(j == r ? 1.0 : 0.0, 0.0);
    }

    const GroupResult res =
        factor_rows<N, L, POLICY, ARITH>(a, r, row_live, bc);

    if (mat_ok) {
      const int64_t o = args.out_idx ? args.out_idx[i] : i;
      if (args.out.piv && r < N)
        args.out.piv[o * args.piv_stride + r] = (int8_t)res.my_piv;
      if (r == 0) write_result(args.out, o, res);
    }

    if (!args.next_task) break;
    int32_t t = 0;
    if (lane == 0) t = atomicAdd(args.next_task, 1);
    task = args.grid_tasks + __builtin_amdgcn_readfirstlane(t);
  }
}

template <int N>
constexpr int default_group_lanes() {
  return N <= 8 ? 8 : N <= 16 ? 16 : N <= 32 ? 32 : 64;
}

}  // namespace lufqhe
