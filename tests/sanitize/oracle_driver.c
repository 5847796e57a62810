/* Exercises every entry point of the CPU oracle (oracle/lufqhe_oracle.c)
 * under AddressSanitizer and UndefinedBehaviorSanitizer: ordinary, padded,
 * singular, non-finite and extreme inputs for n = 1..64.  Exit status 0 when
 * every call returns; the sanitizers abort on the first finding. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/lufqhe_oracle.h"

static uint64_t st = 0x9E3779B97F4A7C15ull;
static double unif(void) {
    st ^= st << 13; st ^= st >> 7; st ^= st << 17;
    return (double)(st >> 11) * 0x1p-53;
}
static double gauss(void) {
    double u = unif(), v = unif();
    if (u < 1e-300) u = 1e-300;
    return sqrt(-2.0 * log(u)) * cos(6.283185307179586 * v);
}

int main(void) {
    const int count = 5;
    for (int n = 1; n <= 64; ++n) {
        const int ld = n + (n % 3); /* padded rows for some n */
        const int64_t stride = (int64_t)n * ld + 1;
        double *A = malloc(sizeof(double) * 2 * (size_t)(count * stride + ld));
        for (int64_t i = 0; i < 2 * (count * stride + ld); ++i) A[i] = gauss();
        /* matrix 1: a zero column; 2: a row without real part; 3: NaN/inf */
        for (int r = 0; r < n; ++r) A[2 * (1 * stride + (int64_t)r * ld)] = A[2 * (1 * stride + (int64_t)r * ld) + 1] = 0.0;
        for (int c = 0; c < n; ++c) A[2 * (2 * stride + (int64_t)(n / 2) * ld + c)] = 0.0;
        A[2 * (3 * stride)] = NAN;
        A[2 * (3 * stride) + 1] = INFINITY;
        for (int c = 0; c < n; ++c) A[2 * (4 * stride + c)] *= 0x1p600;
        double det[2 * 5], mant[2 * 5], la[5], ph[5];
        int32_t e[5], info[5], piv[5 * 64];
        double gap[5 * 64];
        oracle_crout_ref_batch_z(A, count, n, ld, stride, det, mant, e, info, piv, 2);
        oracle_crout_ref_gap_batch_z(A, count, n, ld, stride, piv, gap, info, 2);
        for (int pol = 0; pol < 2; ++pol)
            for (int ar = 0; ar < 2; ++ar)
                oracle_lup_mirror_batch_z(A, count, n, ld, stride, pol, ar, mant, e, info, piv, 2);
        oracle_lup_ld_batch_z(A, count, n, ld, stride, la, ph, info, 2);
        float *Af = malloc(sizeof(float) * 2 * (size_t)(count * stride + ld));
        for (int64_t i = 0; i < 2 * (count * stride + ld); ++i) Af[i] = (float)A[i];
        float detf[2 * 5];
        oracle_crout_ref_batch_c(Af, count, n, ld, stride, detf, info, 2);
        free(Af);
        double sum[2];
        double *P = malloc(sizeof(double) * 2 * (size_t)(6 * n * n));
        for (int64_t i = 0; i < 2 * 6 * n * n; ++i) P[i] = gauss();
        oracle_crout_ref_pairsum_z(P, 6, n, n, sum);
        free(P);
        free(A);
    }
    /* mixed sizes */
    int32_t ns[40];
    int64_t offs[40], tot = 0;
    for (int m = 0; m < 40; ++m) { ns[m] = 1 + m % 64; offs[m] = tot; tot += (int64_t)ns[m] * ns[m]; }
    double *F = malloc(sizeof(double) * 2 * (size_t)tot);
    for (int64_t i = 0; i < 2 * tot; ++i) F[i] = gauss();
    double mant[80], la[40], ph[40];
    int32_t e[40], info[40];
    oracle_lup_mirror_mixed_z(F, 40, ns, offs, 0, 0, mant, e, info, 3);
    oracle_lup_ld_mixed_z(F, 40, ns, offs, la, ph, info, 3);
    free(F);
    /* the corrected quotient at the edges of its range */
    const double xs[] = {0.0, 5e-324, 0x1p-900, 1.0, 0x1p1000, INFINITY, NAN};
    const double ss[] = {1e-24, 1.0, 3.0, 0x1p959, 0x1p1000};
    for (unsigned i = 0; i < sizeof xs / sizeof *xs; ++i)
        for (unsigned j = 0; j < sizeof ss / sizeof *ss; ++j)
            (void)oracle_quot_rn(xs[i], ss[j], ss[j] <= 0x1p960 ? 1.0 / ss[j] : NAN);
    printf("oracle driver: ok\n");
    return 0;
}
