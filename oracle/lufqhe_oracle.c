/*
 * lufqhe_oracle.c -- CPU oracle for the batched LUP-determinant path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, by the smoke check and by the
 * benchmark's CPU-baseline leg, never by the product library.
 *
 * Three independent restatements live here:
 *
 *  1. crout_ref  -- the reference CPU Crout, LUCoreCPU::determinant
 *                   (LU-fqhe/src/LUCoreCPU.cu:112-267): left-looking Crout,
 *                   implicit row scaling on |Re| (LUCoreCPU.cu:145-160),
 *                   pivot = argmax |Re u_ij| / s_i with ties to the last row
 *                   (`>=`, LUCoreCPU.cu:207-212), full row swap + scale move
 *                   (LUCoreCPU.cu:215-232), column scaling by a complex
 *                   divide (LUCoreCPU.cu:234-251), sequential diagonal
 *                   product times the swap sign (LUCoreCPU.cu:256-265).
 *                   Instantiated at complex128 and complex64.
 *  2. lup_mirror -- right-looking LUP in the exact operation order of the HIP
 *                   kernel, so kernel and oracle can be compared bit for bit.
 *                   Its ARITH_REF mode uses the reference's formulas and
 *                   must reproduce crout_ref bit for bit (a left-looking and
 *                   a right-looking order apply the same updates to each
 *                   element in the same ascending-k sequence).
 *  3. lup_ld     -- partial pivoting on |z| in long double; the accuracy
 *                   yardstick (log|det|, arg det).
 *
 * Deviations from the reference, identical in all three and in the kernel:
 *   - a row with max |Re| < 1e-24 returns -(row+1) instead of exit(0);
 *   - a pivot with |p|^2 == 0 returns k+1 and its multipliers are zero
 *     instead of dividing by zero;
 *   - the determinant is also carried as mant * 2^exp (renormalised after
 *     every product with an exact power-of-two scaling), so it cannot
 *     overflow; the bits equal the plain product whenever that is normal.
 *
 * Build with -ffp-contract=off: every rounding step is spelled out.
 */
#include "lufqhe_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXN 64

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* mant * 2^e renormalisation: scale so max(|x|,|y|) is in [0.5, 1).  Power of
 * two scalings are exact, so the rounding of later products is unchanged. */
static void renorm(double *x, double *y, int32_t *e) {
    double m = fmax(fabs(*x), fabs(*y));
    if (m != 0.0 && isfinite(m)) {
        int k;
        (void)frexp(m, &k);
        *x = ldexp(*x, -k);
        *y = ldexp(*y, -k);
        *e += k;
    }
}

/* ------------------------------------------------------------------------ */
/* 1. Restatement of LUCoreCPU::determinant                                 */
/* ------------------------------------------------------------------------ */

/* Body shared by the complex64 and complex128 instantiations.  `a` is a
 * private copy of the matrix (the reference factors in place), stride ld. */
#define CROUT_BODY(T, FABS)                                                   \
    int imax = 0, info = 0;                                                   \
    T d = 1;                                                                  \
    T holder[MAXN];                                                           \
    /* row scaling, LUCoreCPU.cu:145-160 */                                   \
    for (int i = 0; i < n; ++i) {                                             \
        T big = 0;                                                            \
        for (int j = 0; j < n; ++j) {                                         \
            T temp = FABS(RE(i, j));                                          \
            if (temp > big) big = temp;                                       \
        }                                                                     \
        if (big < 1e-24) return -(i + 1); /* reference: exit(0) */           \
        holder[i] = big;                                                      \
    }                                                                         \
    for (int j = 0; j < n; ++j) {                                             \
        /* U part, LUCoreCPU.cu:164-182 */                                    \
        for (int i = 0; i < j; ++i) {                                         \
            T sx = RE(i, j), sy = IM(i, j);                                   \
            for (int k = 0; k < i; ++k) {                                     \
                T tx = RE(i, k) * RE(k, j) - IM(i, k) * IM(k, j);             \
                T ty = IM(i, k) * RE(k, j) + RE(i, k) * IM(k, j);             \
                sx = sx - tx;                                                 \
                sy = sy - ty;                                                 \
            }                                                                 \
            RE(i, j) = sx;                                                    \
            IM(i, j) = sy;                                                    \
        }                                                                     \
        /* L part and pivot search, LUCoreCPU.cu:188-213 */                   \
        T big = 0, second = -1;                                               \
        for (int i = j; i < n; ++i) {                                         \
            T sx = RE(i, j), sy = IM(i, j);                                   \
            for (int k = 0; k < j; ++k) {                                     \
                T tx = RE(i, k) * RE(k, j) - IM(i, k) * IM(k, j);             \
                T ty = IM(i, k) * RE(k, j) + RE(i, k) * IM(k, j);             \
                sx = sx - tx;                                                 \
                sy = sy - ty;                                                 \
            }                                                                 \
            RE(i, j) = sx;                                                    \
            IM(i, j) = sy;                                                    \
            T dum = FABS(sx) / holder[i];                                     \
            if (dum >= big) {                                                 \
                second = (i > j) ? big : second;                              \
                big = dum;                                                    \
                imax = i;                                                     \
            } else if (dum > second) {                                        \
                second = dum;                                                 \
            }                                                                 \
        }                                                                     \
        /* relative margin of the choice over the best other candidate */    \
        if (gap) gap[j] = (j == n - 1) ? INFINITY                             \
                          : (big > 0 ? (double)((big - second) / big) : 0.0); \
        /* row swap, LUCoreCPU.cu:215-232 */                                  \
        if (j != imax) {                                                      \
            for (int k = 0; k < n; ++k) {                                     \
                T tx = RE(imax, k), ty = IM(imax, k);                         \
                RE(imax, k) = RE(j, k);                                       \
                IM(imax, k) = IM(j, k);                                       \
                RE(j, k) = tx;                                                \
                IM(j, k) = ty;                                                \
            }                                                                 \
            d = -d;                                                           \
            holder[imax] = holder[j];                                         \
        }                                                                     \
        if (piv) piv[j] = imax;                                               \
        /* column scaling, LUCoreCPU.cu:234-251 */                            \
        {                                                                     \
            T px = RE(j, j), py = IM(j, j);                                   \
            T den = px * px + py * py;                                        \
            if (den == 0) {                                                   \
                if (!info) info = j + 1;                                      \
                for (int i = j + 1; i < n; ++i) RE(i, j) = IM(i, j) = 0;      \
            } else if (j != n - 1) {                                          \
                for (int i = j + 1; i < n; ++i) {                             \
                    T ax = RE(i, j), ay = IM(i, j);                           \
                    T tx = (ax * px + ay * py) / den;                         \
                    T ty = (ay * px - ax * py) / den;                         \
                    RE(i, j) = tx;                                            \
                    IM(i, j) = ty;                                            \
                }                                                             \
            }                                                                 \
        }                                                                     \
    }

#define RE(i, j) a[2 * ((i) * ld + (j))]
#define IM(i, j) a[2 * ((i) * ld + (j)) + 1]

static int crout_ref_z_inplace(double *a, int n, int ld, double *det_plain,
                               double *mant, int32_t *exp2, int32_t *piv,
                               double *gap) {
    CROUT_BODY(double, fabs)
    /* determinant, LUCoreCPU.cu:256-265 (plain), and the same sequence
     * carried as mant * 2^e */
    double dx = 1.0, dy = 0.0, mx = 1.0, my = 0.0;
    int32_t e = 0;
    for (int i = 0; i < n; ++i) {
        double ux = RE(i, i), uy = IM(i, i);
        double tx = dx * ux - dy * uy;
        double ty = dy * ux + dx * uy;
        dx = tx;
        dy = ty;
        tx = mx * ux - my * uy;
        ty = my * ux + mx * uy;
        mx = tx;
        my = ty;
        renorm(&mx, &my, &e);
    }
    if (det_plain) {
        det_plain[0] = d * dx;
        det_plain[1] = d * dy;
    }
    if (mant) {
        mant[0] = d * mx;
        mant[1] = d * my;
    }
    if (exp2) *exp2 = e;
    return info;
}

static int crout_ref_c_inplace(float *a, int n, int ld, float *det_plain,
                               int32_t *piv) {
    double *gap = NULL;
    CROUT_BODY(float, fabsf)
    float dx = 1.0f, dy = 0.0f;
    for (int i = 0; i < n; ++i) {
        float ux = RE(i, i), uy = IM(i, i);
        float tx = dx * ux - dy * uy;
        float ty = dy * ux + dx * uy;
        dx = tx;
        dy = ty;
    }
    det_plain[0] = d * dx;
    det_plain[1] = d * dy;
    return info;
}
#undef RE
#undef IM

static void nan_outputs(double *det_plain, double *mant, int32_t *exp2,
                        int32_t *piv, int n) {
    if (det_plain) det_plain[0] = det_plain[1] = NAN;
    if (mant) mant[0] = mant[1] = NAN;
    if (exp2) *exp2 = 0;
    if (piv)
        for (int k = 0; k < n; ++k) piv[k] = -1;
}

static int crout_ref_z_gap(const double *A, int n, int ld, double *det_plain,
                           double *mant, int32_t *exp2, int32_t *piv,
                           double *gap) {
    if (n < 1 || n > MAXN || ld < n) return -1000;
    double a[2 * MAXN * MAXN];
    for (int i = 0; i < n; ++i)
        memcpy(&a[2 * i * n], &A[2 * (size_t)i * ld], sizeof(double) * 2 * n);
    int info = crout_ref_z_inplace(a, n, n, det_plain, mant, exp2, piv, gap);
    if (info < 0) nan_outputs(det_plain, mant, exp2, piv, n);
    return info;
}

int oracle_crout_ref_z(const double *A, int n, int ld, double *det_plain,
                       double *mant, int32_t *exp2, int32_t *piv) {
    return crout_ref_z_gap(A, n, ld, det_plain, mant, exp2, piv, NULL);
}

int oracle_crout_ref_c(const float *A, int n, int ld, float *det_plain,
                       int32_t *piv) {
    if (n < 1 || n > MAXN || ld < n) return -1000;
    float a[2 * MAXN * MAXN];
    for (int i = 0; i < n; ++i)
        memcpy(&a[2 * i * n], &A[2 * (size_t)i * ld], sizeof(float) * 2 * n);
    int info = crout_ref_c_inplace(a, n, n, det_plain, piv);
    if (info < 0) {
        det_plain[0] = det_plain[1] = NAN;
        if (piv)
            for (int k = 0; k < n; ++k) piv[k] = -1;
    }
    return info;
}

/* ------------------------------------------------------------------------ */
/* 2. Mirror of the HIP kernel's operation order                            */
/* ------------------------------------------------------------------------ */

/* Exponent k of m with m = f * 2^k, f in [0.5, 1); 0 for 0, inf and NaN (the
 * convention of the kernel's frexp-exponent instruction). */
static int frexp_exp0(double m) {
    int k = 0;
    if (m != 0.0 && isfinite(m)) (void)frexp(m, &k);
    return k;
}

/* The kernel's CROUT_REF criterion in fused arithmetic: x / s from
 * y = RN(1/s) by one residual correction (Markstein), with an IEEE division
 * outside the range where that is exact.  Both branches give RN(x / s), the
 * reference's `fabs(sum.x) / holder[i]` (LU-fqhe/src/LUCoreCPU.cu:205);
 * tests/test_oracle.py checks the corrected product against the division
 * on constructed hardest-to-round quotients. */
double oracle_quot_rn(double x, double s, double y) {
    const double q0 = x * y;
    const double r = fma(-q0, s, x);
    double q = fma(r, y, q0);
    if (!(q0 >= 0x1p-880 && q0 <= 0x1p960)) q = x / s;
    return q;
}

/* RN(1/s) as the kernel keeps it per row: NaN for s > 2^960 sends every
 * quotient of that row to the division. */
static double row_recip(double s) { return s <= 0x1p960 ? 1.0 / s : NAN; }

int oracle_lup_mirror_z(const double *A, int n, int ld, int policy, int arith,
                        double *mant, int32_t *exp2, int32_t *piv) {
    if (n < 1 || n > MAXN || ld < n) return -1000;
    double a[MAXN][MAXN][2];
    double s[MAXN], rs[MAXN];
    int pos[MAXN], act[MAXN];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            a[i][j][0] = A[2 * ((size_t)i * ld + j)];
            a[i][j][1] = A[2 * ((size_t)i * ld + j) + 1];
        }
    if (policy == ORACLE_POLICY_CROUT_REF) {
        /* s_i = max_j |Re a_ij| (exact: max is order independent) */
        int bad = -1;
        for (int i = 0; i < n; ++i) {
            double big = fabs(a[i][0][0]);
            for (int j = 1; j < n; ++j) big = fmax(big, fabs(a[i][j][0]));
            if (big < 1e-24 && bad < 0) bad = i;
            s[i] = big;
            rs[i] = row_recip(big);
        }
        if (bad >= 0) {
            nan_outputs(NULL, mant, exp2, piv, n);
            return -(bad + 1);
        }
    }
    for (int i = 0; i < n; ++i) {
        pos[i] = i;
        act[i] = 1;
    }
    double dx = 1.0, dy = 0.0, sg = 1.0;
    int32_t de = 0;
    int info = 0;
    for (int k = 0; k < n; ++k) {
        /* pivot search: max criterion; exact ties go to the larger position
         * under CROUT_REF (the reference's `>=`), to the smaller under ABS
         * (LAPACK izamax) */
        double cmax = -INFINITY;
        int ppos = -1, pr = -1;
        for (int i = 0; i < n; ++i) {
            if (!act[i]) continue;
            double x = a[i][k][0], y = a[i][k][1], c;
            if (policy == ORACLE_POLICY_CROUT_REF)
                c = (arith == ORACLE_ARITH_FUSED)
                        ? oracle_quot_rn(fabs(x), s[i], rs[i])
                        : fabs(x) / s[i];
            else
                c = fabs(x) + fabs(y); /* LAPACK's cabs1 */
            c = fmin(c, DBL_MAX);
            const int later = (policy == ORACLE_POLICY_ABS) ? pos[i] < ppos
                                                            : pos[i] > ppos;
            if (c > cmax || (c == cmax && later)) {
                cmax = c;
                ppos = pos[i];
                pr = i;
            }
        }
        if (piv) piv[k] = ppos;
        const double px = a[pr][k][0], py = a[pr][k][1];
        if (ppos != k) sg = -sg;
        /* determinant product and |p|^2.  Fused arithmetic scales p by
         * 2^-kp (kp the exponent of max(|Re p|, |Im p|)): |p_s|^2 is 0 only
         * for p == 0, the reciprocal is scaled back below, and the running
         * product of the p_s (within 2^+-64 of 1) is renormalised once at
         * the end, det = prod p_s * 2^(sum kp).  Reference arithmetic keeps
         * the plain product, renormalised every step. */
        double den;
        int kp = 0;
        double psx = px, psy = py;
        if (arith == ORACLE_ARITH_FUSED) {
            kp = frexp_exp0(fmax(fabs(px), fabs(py)));
            psx = ldexp(px, -kp);
            psy = ldexp(py, -kp);
            den = fma(psx, psx, psy * psy);
            double t = dy * psy;
            const double nx = fma(dx, psx, -t);
            t = dy * psx;
            const double ny = fma(dx, psy, t);
            dx = nx;
            dy = ny;
            de += kp;
        } else {
            den = px * px + py * py;
            const double nx = dx * px - dy * py;
            const double ny = dy * px + dx * py;
            dx = nx;
            dy = ny;
            renorm(&dx, &dy, &de);
        }
        if (den == 0.0) {
            if (!info) info = k + 1;
        } else {
            double ir = 0.0, ii = 0.0;
            if (arith == ORACLE_ARITH_FUSED) {
                double rden = 1.0 / den;
                ir = ldexp(psx * rden, -kp);
                ii = ldexp(-(psy * rden), -kp);
            }
            for (int i = 0; i < n; ++i) {
                if (!act[i] || i == pr) continue;
                double ax = a[i][k][0], ay = a[i][k][1], lx, ly;
                if (arith == ORACLE_ARITH_FUSED) {
                    lx = fma(ax, ir, -(ay * ii));
                    ly = fma(ay, ir, ax * ii);
                } else {
                    lx = (ax * px + ay * py) / den;
                    ly = (ay * px - ax * py) / den;
                }
                for (int j = k + 1; j < n; ++j) {
                    const double ux = a[pr][j][0], uy = a[pr][j][1];
                    double x = a[i][j][0], y = a[i][j][1];
                    if (arith == ORACLE_ARITH_FUSED) {
                        x = fma(-lx, ux, x);
                        x = fma(ly, uy, x);
                        y = fma(-lx, uy, y);
                        y = fma(-ly, ux, y);
                    } else {
                        double tx = lx * ux - ly * uy;
                        double ty = ly * ux + lx * uy;
                        x = x - tx;
                        y = y - ty;
                    }
                    a[i][j][0] = x;
                    a[i][j][1] = y;
                }
            }
        }
        /* position bookkeeping (the reference's physical swap) */
        for (int i = 0; i < n; ++i)
            if (pos[i] == k) pos[i] = ppos;
        pos[pr] = k;
        act[pr] = 0;
    }
    if (arith == ORACLE_ARITH_FUSED) renorm(&dx, &dy, &de);
    if (dx == 0.0 && dy == 0.0) de = 0; /* a zero determinant is (0, 2^0) */
    if (mant) {
        mant[0] = sg * dx;
        mant[1] = sg * dy;
    }
    if (exp2) *exp2 = de;
    return info;
}

/* ------------------------------------------------------------------------ */
/* 3. Long-double truth                                                     */
/* ------------------------------------------------------------------------ */

int oracle_lup_ld_z(const double *A, int n, int ld, double *logabs,
                    double *phase) {
    if (n < 1 || n > MAXN || ld < n) return -1000;
    long double a[MAXN][MAXN][2];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            a[i][j][0] = A[2 * ((size_t)i * ld + j)];
            a[i][j][1] = A[2 * ((size_t)i * ld + j) + 1];
        }
    long double la = 0.0L, ph = 0.0L;
    int info = 0;
    for (int k = 0; k < n; ++k) {
        int p = k;
        long double best = -1.0L;
        for (int i = k; i < n; ++i) {
            long double m = a[i][k][0] * a[i][k][0] + a[i][k][1] * a[i][k][1];
            if (m > best) {
                best = m;
                p = i;
            }
        }
        if (p != k) {
            for (int j = 0; j < n; ++j) {
                long double tx = a[p][j][0], ty = a[p][j][1];
                a[p][j][0] = a[k][j][0];
                a[p][j][1] = a[k][j][1];
                a[k][j][0] = tx;
                a[k][j][1] = ty;
            }
            ph += 3.14159265358979323846264338327950288L;
        }
        long double px = a[k][k][0], py = a[k][k][1];
        long double den = px * px + py * py;
        if (den == 0.0L) {
            info = k + 1;
            break;
        }
        la += 0.5L * logl(den);
        ph += atan2l(py, px);
        for (int i = k + 1; i < n; ++i) {
            long double lx = (a[i][k][0] * px + a[i][k][1] * py) / den;
            long double ly = (a[i][k][1] * px - a[i][k][0] * py) / den;
            for (int j = k + 1; j < n; ++j) {
                long double ux = a[k][j][0], uy = a[k][j][1];
                a[i][j][0] -= lx * ux - ly * uy;
                a[i][j][1] -= ly * ux + lx * uy;
            }
        }
    }
    if (info) {
        *logabs = -INFINITY;
        *phase = 0.0;
        return info;
    }
    /* wrap the phase into (-pi, pi] */
    const long double twopi = 6.28318530717958647692528676655900577L;
    ph = fmodl(ph, twopi);
    if (ph > twopi / 2) ph -= twopi;
    if (ph <= -twopi / 2) ph += twopi;
    *logabs = (double)la;
    *phase = (double)ph;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Batched drivers                                                          */
/* ------------------------------------------------------------------------ */

void oracle_crout_ref_batch_z(const double *A, int64_t count, int n, int ld,
                              int64_t stride, double *det_plain, double *mant,
                              int32_t *exp2, int32_t *info, int32_t *piv,
                              int nthreads) {
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < count; ++m) {
        int r = oracle_crout_ref_z(A + 2 * m * stride, n, ld,
                                   det_plain ? det_plain + 2 * m : NULL,
                                   mant ? mant + 2 * m : NULL,
                                   exp2 ? exp2 + m : NULL,
                                   piv ? piv + m * n : NULL);
        if (info) info[m] = r;
    }
}

void oracle_crout_ref_gap_batch_z(const double *A, int64_t count, int n,
                                  int ld, int64_t stride, int32_t *piv,
                                  double *gap, int32_t *info, int nthreads) {
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < count; ++m) {
        int r = crout_ref_z_gap(A + 2 * m * stride, n, ld, NULL, NULL, NULL,
                                piv + m * n, gap + m * n);
        if (info) info[m] = r;
    }
}

void oracle_crout_ref_batch_c(const float *A, int64_t count, int n, int ld,
                              int64_t stride, float *det_plain, int32_t *info,
                              int nthreads) {
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < count; ++m) {
        int r = oracle_crout_ref_c(A + 2 * m * stride, n, ld,
                                   det_plain + 2 * m, NULL);
        if (info) info[m] = r;
    }
}

void oracle_lup_mirror_batch_z(const double *A, int64_t count, int n, int ld,
                               int64_t stride, int policy, int arith,
                               double *mant, int32_t *exp2, int32_t *info,
                               int32_t *piv, int nthreads) {
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < count; ++m) {
        int r = oracle_lup_mirror_z(A + 2 * m * stride, n, ld, policy, arith,
                                    mant + 2 * m, exp2 + m,
                                    piv ? piv + m * n : NULL);
        if (info) info[m] = r;
    }
}

void oracle_lup_ld_batch_z(const double *A, int64_t count, int n, int ld,
                           int64_t stride, double *logabs, double *phase,
                           int32_t *info, int nthreads) {
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < count; ++m) {
        int r = oracle_lup_ld_z(A + 2 * m * stride, n, ld, logabs + m,
                                phase + m);
        if (info) info[m] = r;
    }
}

void oracle_lup_mirror_mixed_z(const double *A, int64_t count,
                               const int32_t *ns, const int64_t *offs,
                               int policy, int arith, double *mant,
                               int32_t *exp2, int32_t *info, int nthreads) {
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t m = 0; m < count; ++m) {
        int r = oracle_lup_mirror_z(A + 2 * offs[m], ns[m], ns[m], policy,
                                    arith, mant + 2 * m, exp2 + m, NULL);
        if (info) info[m] = r;
    }
}

void oracle_lup_ld_mixed_z(const double *A, int64_t count, const int32_t *ns,
                           const int64_t *offs, double *logabs, double *phase,
                           int32_t *info, int nthreads) {
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t m = 0; m < count; ++m) {
        int r = oracle_lup_ld_z(A + 2 * offs[m], ns[m], ns[m], logabs + m,
                                phase + m);
        if (info) info[m] = r;
    }
}

void oracle_crout_ref_pairsum_z(const double *A, int64_t num, int n, int ld,
                                double *sum) {
    double sx = 0.0, sy = 0.0;
    for (int64_t i = 0; i < num / 2; ++i) {
        double d1[2], d2[2];
        oracle_crout_ref_z(A + 2 * (2 * i) * (int64_t)n * ld, n, ld, d1, NULL,
                           NULL, NULL);
        oracle_crout_ref_z(A + 2 * (2 * i + 1) * (int64_t)n * ld, n, ld, d2,
                           NULL, NULL, NULL);
        /* LUCoreCPU.cu:88-92 */
        double tx = d1[0] * d2[0] - d1[1] * d2[1];
        double ty = d1[1] * d2[0] + d1[0] * d2[1];
        sx = sx + tx;
        sy = sy + ty;
    }
    sum[0] = sx;
    sum[1] = sy;
}
