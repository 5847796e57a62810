/*
 * lufqhe.h -- C ABI of the MI355X batched LUP-determinant engine.
 *
 * Plain pointers and sizes only.  Device pointers are HIP device memory on the
 * current device; every call is stream ordered on `stream` and allocates
 * nothing unless its comment says so.
 *
 * Matrix layout (LU-fqhe/inc/LUCoreGPU.h:24-40): complex128 elements,
 * interleaved (re, im), row major; element (r, c) of matrix m of a uniform
 * batch is at A[m*stride + r*ld + c] (units: complex elements).  The
 * reference's in-place batch is stride = n*ld (LU-fqhe/src/LUCoreGPU.cu:213).
 *
 * Determinant outputs per matrix (any output but mant may be NULL):
 *   mant   complex128, det = mant * 2^exp, max(|Re|,|Im|) of mant in [0.5,1)
 *   exp    int32
 *   logabs float64 log|det| (-inf for det == 0)
 *   phase  float64 arg det in (-pi, pi]
 *   info   int32: 0 ok; k+1 the pivot of step k had |p|^2 == 0 (det = 0; the
 *          reference divides by zero there, LU-fqhe/src/LUCoreCPU.cu:243);
 *          -(i+1) pivot policy CROUT_REF only: row i has max_j |Re a_ij| <
 *          1e-24, which the reference answers with exit(0)
 *          (LU-fqhe/src/LUCoreCPU.cu:158); outputs are NaN
 *   piv    int8, n per matrix: the row position exchanged with position k at
 *          step k (the reference's imax, LU-fqhe/src/LUCoreCPU.cu:209-216)
 *
 * Status: every entry point returns LUFQHE_SUCCESS or an error code; the
 * reference had none (exit(1) on an oversized grid,
 * LU-fqhe/src/LUCoreGPU.cu:102-105).
 */
#ifndef LUFQHE_H_
#define LUFQHE_H_

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LUFQHE_SUCCESS 0
#define LUFQHE_ERROR_INVALID_VALUE 1
#define LUFQHE_ERROR_UNSUPPORTED 2
#define LUFQHE_ERROR_HIP 3
#define LUFQHE_ERROR_NOT_BUILT 4
#define LUFQHE_ERROR_OUT_OF_BUDGET 5

/* Pivot policies.  CROUT_REF is the reference's rule: argmax over remaining
 * rows of |Re u_ik| / s_i, s_i = max_j |Re a_ij| of the original row, ties to
 * the last position (LU-fqhe/src/LUCoreCPU.cu:145-213).  ABS: argmax |u_ik|^2,
 * ties to the last position. */
#define LUFQHE_POLICY_CROUT_REF 0
#define LUFQHE_POLICY_ABS 1
/* Arithmetic.  FUSED: FP64 with fused multiply-adds (4 FMAs per complex
 * update, one reciprocal of the pivot per step).  REF: the reference's own
 * formulas and rounding steps (LU-fqhe/src/LUCoreCPU.cu:173-176,243-246,
 * 258-259); bit-identical to an FP64 build of LUCoreCPU::determinant. */
#define LUFQHE_ARITH_FUSED 0
#define LUFQHE_ARITH_REF 1

/* Library version (major*10000 + minor*100 + patch). */
int32_t lufqhe_version(void);
/* Static text for a status code. */
const char *lufqhe_status_string(int32_t status);
/* Text of the last HIP error seen by this thread ("" if none). */
const char *lufqhe_last_hip_error(void);

/* Batched determinant of a uniform batch, one plain launch.
 * Replaces LUCoreGPU::CALL + LUDecomposition (LU-fqhe/inc/LUCoreGPU.h:47-48,
 * LU-fqhe/src/LUCoreGPU.cu:38-106,159-437) with out-of-place outputs: the
 * input is read only.  1 <= n <= 64, ld >= n, stride >= 0, batch >= 0
 * (batch == 0 is a no-op).  The (ABS, REF) combination is unsupported. */
int32_t lufqhe_det_batched(const void *d_A, int64_t batch, int32_t n,
                           int64_t ld, int64_t stride, void *d_mant,
                           int32_t *d_exp, double *d_logabs, double *d_phase,
                           int32_t *d_info, int8_t *d_piv, int32_t policy,
                           int32_t arith, hipStream_t stream);

/* Bytes of device workspace lufqhe_det_inplace needs for `count` matrices
 * (one mantissa and one exponent per matrix). */
int64_t lufqhe_det_inplace_workspace_bytes(int64_t count);

/* The reference's in-place contract: LUCoreGPU::CALL built with WRITE_DET
 * and PERFORM_PROD (LU-fqhe/Makefile:3, LU-fqhe/src/LUCoreGPU.cu:384-436):
 * after the call A_m[0][0] holds det(M_m) as a plain complex128 (inf where
 * it overflows, as the reference's product would) and, with pair_product,
 * A_2i[0][0] holds det(M_2i) * det(M_2i+1).  Only the [0][0] entries are
 * written (the reference also leaves its LU factors in A; nothing reads
 * them).  d_work: lufqhe_det_inplace_workspace_bytes(count) bytes of device
 * memory; if NULL the call allocates it stream-ordered (hipMallocAsync /
 * hipFreeAsync on `stream`).  Arithmetic is FUSED. */
int32_t lufqhe_det_inplace(void *d_A, int64_t count, int32_t n, int64_t ld,
                           int64_t stride, int32_t pair_product,
                           int32_t policy, void *d_work, hipStream_t stream);

/* Bytes of device workspace the two pair sums need (fixed). */
int64_t lufqhe_pair_sum_workspace_bytes(void);

/* The reference's wavefunction sum, globalReduction
 * (LU-fqhe/inc/LUCoreGPU.h:50, LU-fqhe/src/LUCoreGPU.cu:113-156,447-557):
 * sum over i < count/2 of A_2i[0][0] (stride = matrix stride in complex
 * elements), written to d_out[0] as complex128.  Deterministic (fixed
 * summation order).  d_work: lufqhe_pair_sum_workspace_bytes() bytes, or
 * NULL to allocate it stream-ordered. */
int32_t lufqhe_global_reduction(const void *d_A, int64_t count, int64_t stride,
                                void *d_work, void *d_out, hipStream_t stream);

/* Overflow-free form of the same sum from (mant, exp) determinants:
 * sum over i < count/2 of det_2i * det_2i+1 = out_mant * 2^out_exp
 * (LU-fqhe/src/LUCoreCPU.cu:83-98 summed the pair products in float).
 * Deterministic.  d_work as above. */
int32_t lufqhe_pair_sum(const void *d_mant, const int32_t *d_exp,
                        int64_t count, void *d_work, void *d_out_mant,
                        int32_t *d_out_exp, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* LUFQHE_H_ */
