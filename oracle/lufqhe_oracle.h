/*
 * lufqhe_oracle.h -- CPU oracle for the batched LUP-determinant path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is imported by the tests, by the smoke check and by the
 * benchmark's CPU-baseline leg, and only as the checker / baseline.
 *
 * Matrices are complex, row-major, interleaved (re, im), element (r, c) of
 * matrix m at A[2*(m*stride + r*ld + c)] -- the layout of
 * LU-fqhe/inc/LUCoreGPU.h:24-40 with loadSize == ld.
 *
 * Status codes (per matrix, returned and/or written to info[]):
 *   0      success
 *   k+1    the pivot chosen at step k has |p|^2 == 0 (det is 0; the
 *          reference would divide by zero there, LU-fqhe/src/LUCoreCPU.cu:243)
 *   -(i+1) CROUT_REF policy only: row i has max_j |Re a_ij| < 1e-24; the
 *          reference calls exit(0) there (LU-fqhe/src/LUCoreCPU.cu:158).
 *          mant/det outputs are NaN.
 */
#ifndef LUFQHE_ORACLE_H_
#define LUFQHE_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_POLICY_CROUT_REF = 0, ORACLE_POLICY_ABS = 1 };
enum { ORACLE_ARITH_FUSED = 0, ORACLE_ARITH_REF = 1 };

/* Restatement of LUCoreCPU::determinant (LU-fqhe/src/LUCoreCPU.cu:112-267),
 * complex128.  det_plain: the reference's own product (may overflow);
 * mant/exp2: the same product carried as mant * 2^exp2 (identical bits
 * whenever det_plain is finite and normal).  piv (nullable): imax per step. */
int oracle_crout_ref_z(const double *A, int n, int ld, double *det_plain,
                       double *mant, int32_t *exp2, int32_t *piv);
/* The same restatement instantiated at complex64 (the reference's own build). */
int oracle_crout_ref_c(const float *A, int n, int ld, float *det_plain,
                       int32_t *piv);

/* Right-looking LUP that mirrors the GPU kernel's operation order exactly
 * (both pivot policies, both arithmetic modes).  ARITH_REF uses the
 * reference's formulas and must agree bit-for-bit with oracle_crout_ref_z. */
int oracle_lup_mirror_z(const double *A, int n, int ld, int policy, int arith,
                        double *mant, int32_t *exp2, int32_t *piv);

/* Independent truth: partial pivoting on |z| in long double (x87, 64-bit
 * significand).  Returns log|det| and arg(det); -inf log for a singular
 * matrix. */
int oracle_lup_ld_z(const double *A, int n, int ld, double *logabs,
                    double *phase);

/* Batched drivers (OpenMP parallel for, static schedule, nthreads <= 0 means
 * the OpenMP default).  Outputs are per matrix; NULL outputs are skipped. */
void oracle_crout_ref_batch_z(const double *A, int64_t count, int n, int ld,
                              int64_t stride, double *det_plain, double *mant,
                              int32_t *exp2, int32_t *info, int32_t *piv,
                              int nthreads);
/* Pivot sequence of the restated Crout with, per step j, the relative margin
 * (c_best - c_second) / c_best of the chosen criterion over the best other
 * candidate row (0 for an exact tie, +inf at the last step).  piv and gap
 * hold n entries per matrix. */
void oracle_crout_ref_gap_batch_z(const double *A, int64_t count, int n,
                                  int ld, int64_t stride, int32_t *piv,
                                  double *gap, int32_t *info, int nthreads);
/* The fused CROUT_REF criterion of the kernel, x / s from y = RN(1/s) (see
 * lup_mirror): equal to RN(x / s). */
double oracle_quot_rn(double x, double s, double y);
void oracle_crout_ref_batch_c(const float *A, int64_t count, int n, int ld,
                              int64_t stride, float *det_plain, int32_t *info,
                              int nthreads);
void oracle_lup_mirror_batch_z(const double *A, int64_t count, int n, int ld,
                               int64_t stride, int policy, int arith,
                               double *mant, int32_t *exp2, int32_t *info,
                               int32_t *piv, int nthreads);
void oracle_lup_ld_batch_z(const double *A, int64_t count, int n, int ld,
                           int64_t stride, double *logabs, double *phase,
                           int32_t *info, int nthreads);
/* Mixed-size batch: matrix m has side ns[m] and starts at element offs[m]
 * (ld == ns[m]). */
void oracle_lup_mirror_mixed_z(const double *A, int64_t count,
                               const int32_t *ns, const int64_t *offs,
                               int policy, int arith, double *mant,
                               int32_t *exp2, int32_t *info, int nthreads);
void oracle_lup_ld_mixed_z(const double *A, int64_t count, const int32_t *ns,
                           const int64_t *offs, double *logabs, double *phase,
                           int32_t *info, int nthreads);

/* LUCoreCPU::CALL analogue (LU-fqhe/src/LUCoreCPU.cu:39-107): sum over i of
 * det(M_2i) * det(M_2i+1), accumulated in matrix order (deterministic). */
void oracle_crout_ref_pairsum_z(const double *A, int64_t num, int n, int ld,
                                double *sum);

int oracle_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif /* LUFQHE_ORACLE_H_ */
