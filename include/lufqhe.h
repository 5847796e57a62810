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
 *   info   int32: 0 ok; k+1 the pivot of step k was zero (det = 0; the
 *          reference divides by zero there, LU-fqhe/src/LUCoreCPU.cu:243 --
 *          in REF arithmetic, as there, a pivot whose |p|^2 underflows to 0
 *          counts as zero); -(i+1) pivot policy CROUT_REF only: row i has
 *          max_j |Re a_ij| < 1e-24, which the reference answers with exit(0)
 *          (LU-fqhe/src/LUCoreCPU.cu:158); LUFQHE_INFO_BAD_PERM: a row
 *          permutation entry of this matrix is >= n.  For negative codes the
 *          determinant outputs are NaN and the pivots -1
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
#define LUFQHE_ERROR_IO 6
#define LUFQHE_ERROR_TOO_LARGE 7

/* Pivot policies.  CROUT_REF is the reference's rule: argmax over remaining
 * rows of the correctly rounded |Re u_ik| / s_i, s_i = max_j |Re a_ij| of the
 * original row, exact ties to the last position (LU-fqhe/src/LUCoreCPU.cu:
 * 145-213), in both arithmetic modes.  ABS: LAPACK zgetf2's rule, argmax of
 * |Re u_ik| + |Im u_ik| (izamax's cabs1), exact ties to the first position. */
#define LUFQHE_POLICY_CROUT_REF 0
#define LUFQHE_POLICY_ABS 1
/* Arithmetic.  FUSED: FP64 with fused multiply-adds (4 FMAs per complex
 * update, one reciprocal of the pivot per step, the pivot scaled by a power
 * of two before |p|^2 so that it neither overflows nor underflows).  REF: the
 * reference's own formulas and rounding steps (LU-fqhe/src/LUCoreCPU.cu:
 * 173-176,243-246,258-259); bit-identical to an FP64 build of
 * LUCoreCPU::determinant. */
#define LUFQHE_ARITH_FUSED 0
#define LUFQHE_ARITH_REF 1
/* info code of a matrix whose row-permutation entry is out of range. */
#define LUFQHE_INFO_BAD_PERM (-1000)

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

/* Launch shapes.  Every n has up to 4 compiled kernel shapes (lanes per
 * matrix, lane-columns, waves per SIMD the register allocation allows);
 * the autotuner times them on the device and selects one per (policy,
 * arithmetic, n).  Replaces the reference's fixed sizing, RunInfo::RunInfo
 * (LU-fqhe/src/RunInfo.cpp:46-78) and the matrices-per-block command-line
 * option (LU-fqhe/src/FQHE.cpp:214-265).  Selection is process wide; change
 * it while no launch of that n is being issued.  min_waves may be NULL. */
int32_t lufqhe_shape_count(int32_t n);
int32_t lufqhe_shape_info(int32_t n, int32_t shape, int32_t *lanes,
                          int32_t *lane_cols, int32_t *min_waves);
int32_t lufqhe_set_shape(int32_t policy, int32_t arith, int32_t n,
                         int32_t shape);
/* The selected shape, or -1 for invalid arguments. */
int32_t lufqhe_get_shape(int32_t policy, int32_t arith, int32_t n);
/* Applies a shape table written by the autotuner (JSON; path NULL:
 * $LUFQHE_SHAPE_TABLE, else shapes_gfx950.json next to the library).  An
 * entry is applied only if the candidate at its index still has the
 * recorded (lanes, lane_cols, waves).  Returns the number of entries
 * applied, or -1 if the file cannot be read or holds no entry list. */
int32_t lufqhe_load_shape_table(const char *path);

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

/* No counterpart in the reference: lufqhe_det_inplace for callers that
 * cannot allocate (code under HIP graph capture, serving loops that forbid
 * hipMalloc).  The same result, with the workspace in a fixed device array
 * of LUFQHE_STATIC_MAX_MATRICES * 20 bytes (20 MiB) that each GPU holds
 * until the process ends; batches above LUFQHE_STATIC_MAX_MATRICES matrices
 * return LUFQHE_ERROR_TOO_LARGE.  The first call on a device looks up the
 * array's address (hipGetSymbolAddress); no call allocates.  Calls on one
 * GPU share its array: calls that overlap on one GPU race. */
#define LUFQHE_STATIC_MAX_MATRICES (1 << 20)
int32_t lufqhe_det_inplace_static(void *d_A, int64_t count, int32_t n,
                                  int64_t ld, int64_t stride,
                                  int32_t pair_product, int32_t policy,
                                  hipStream_t stream);

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

/* ---- the reference's CPU Crout, on the host -----------------------------
 * LUCoreCPU::determinant (LU-fqhe/src/LUCoreCPU.cu:112-267) in FP64: left-
 * looking Crout, rows scaled by their largest |Re|, the `>=` pivot scan,
 * full row swaps, complex-divide column scaling, sequential diagonal
 * product, every operation rounded as the reference writes it; the results
 * equal the GPU's REF arithmetic bit for bit.  Host pointers; the input is
 * read only; nthreads <= 0 uses every hardware thread.
 *
 * lufqhe_cpu_crout_det: per matrix (layout as lufqhe_det_batched) the plain
 * product h_det (complex128, may overflow), h_mant * 2^h_exp, h_info (as the
 * GPU's, without the permutation code) and h_piv (n int8); any output may
 * be NULL.  Replaces the loop of LUCoreCPU::CALL (LU-fqhe/src/LUCoreCPU.cu:
 * 69-93) with per-matrix outputs.
 * lufqhe_cpu_pair_sum: *h_out = sum over i < num/2 of det(M_2i) *
 * det(M_2i+1), the value LUCoreCPU::CALL returns (LU-fqhe/inc/LUCoreCPU.h:
 * 26, LU-fqhe/src/LUCoreCPU.cu:39-107), with the pair products added in
 * matrix order (deterministic for any thread count); *h_first_bad (may be
 * NULL) gets the first pair holding a row the reference would exit(0) on,
 * or -1. */
int32_t lufqhe_cpu_crout_det(const void *h_A, int64_t count, int32_t n,
                             int64_t ld, int64_t stride, void *h_det,
                             void *h_mant, int32_t *h_exp, int32_t *h_info,
                             int8_t *h_piv, int32_t nthreads);
int32_t lufqhe_cpu_pair_sum(const void *h_A, int64_t num, int32_t n,
                            int64_t ld, int64_t stride, void *h_out,
                            int32_t *h_first_bad, int32_t nthreads);

/* ---- FQHE matrices built on the device ---------------------------------
 * LU-fqhe's matrices are Laughlin-type wavefunction matrices
 * A_ij = z_i^j exp(-|z_i|^2/4), j = 0..n-1, of particle positions z_i
 * (complex128).  The reference builds them on the host and copies whole
 * matrices to the device each Monte Carlo step (LU-fqhe/src/MonteCarloLoop.cu:
 * 292-321; its on-device builder, LU-fqhe/src/mkMatrixElements.cu:169-358,
 * is disabled).  Here the determinant kernels build each matrix from its
 * positions in their load phase: 16n bytes read per matrix instead of 16n^2.
 *
 * Matrix i's positions start at d_z + i*z_stride (complex elements,
 * z_stride >= n or 0 to share one configuration); with d_row_perm (n bytes
 * per matrix, may be NULL) row r of matrix i is particle d_row_perm[i*n + r];
 * an entry >= n makes that matrix's info LUFQHE_INFO_BAD_PERM (NaN outputs),
 * and lufqhe_fqhe_build writes that row as NaN.  Powers are sequential products
 * z^j = z^(j-1) * z, so entries do not depend on the launch shape.
 *
 * lufqhe_fqhe_det: determinants as lufqhe_det_batched.
 * lufqhe_fqhe_build: the same matrices written out (matrix i at
 * d_A + i*n*n, ld = n), entry for entry what the fused load phase uses. */
int32_t lufqhe_fqhe_det(const void *d_z, int64_t count, int32_t n,
                        int64_t z_stride, const uint8_t *d_row_perm,
                        void *d_mant, int32_t *d_exp, double *d_logabs,
                        double *d_phase, int32_t *d_info, int8_t *d_piv,
                        int32_t policy, int32_t arith, hipStream_t stream);
int32_t lufqhe_fqhe_build(const void *d_z, int64_t count, int32_t n,
                          int64_t z_stride, const uint8_t *d_row_perm,
                          void *d_A, hipStream_t stream);

/* ---- mixed-size batches ------------------------------------------------
 * A batch whose matrix m has side n[m] (1..64) and starts at complex element
 * offs[m] of one flat buffer (ld = n[m], row major).  The reference runs one
 * size per RunInfo (LU-fqhe/src/RunInfo.cpp:46-78); a mixed batch there is
 * one CALL per size on repacked uniform batches.
 *
 * lufqhe_mixed_plan_create buckets the batch by n on the host, uploads the
 * per-bucket offset / output-slot tables (16 B per matrix, allocated here,
 * freed by lufqhe_mixed_plan_destroy) and creates `nstreams` streams for the
 * buckets, assigned largest-cost-first.  Fused arithmetic, the given pivot
 * policy, the launch shape selected for each n.
 *
 * lufqhe_mixed_execute writes det outputs (as lufqhe_det_batched) at each
 * matrix's own index m.  source LUFQHE_SOURCE_MATRIX reads matrix m at
 * d_A + offs[m]; LUFQHE_SOURCE_FQHE reads n[m] particle positions there and
 * builds the FQHE matrix on the device (as lufqhe_fqhe_det).  Modes:
 *   LUFQHE_MIXED_GRAPH    the bucket launches forked over the plan streams,
 *                         captured into one hipGraph on the first call for a
 *                         given set of pointers and source (instantiation
 *                         allocates) and replayed after -- with the launch
 *                         shapes selected at capture time; `stream` must not
 *                         be the legacy default stream (capture is not
 *                         allowed there);
 *   LUFQHE_MIXED_STREAMS  the same forked launches, issued directly;
 *   LUFQHE_MIXED_PLAIN    one launch per size, one after another on `stream`;
 *   LUFQHE_MIXED_PERSISTENT  one persistent kernel per size class (n <= 16,
 *                         17..32, 33..48, 49..57, 58..64) instead of one
 *                         launch per bucket: resident waves drain the class's
 *                         buckets, largest n first, from a device work
 *                         counter that a kernel at the head of the graph
 *                         zeroes.  Captured and replayed like GRAPH (the
 *                         per-bucket arguments are written to device memory,
 *                         synchronously, when the pointers change).  Fused
 *                         arithmetic with either policy; a plan with 2^31 or
 *                         more tasks in one class refuses this mode
 *                         (LUFQHE_ERROR_UNSUPPORTED).
 * A plan is not re-entrant: one execute at a time per plan, on the device
 * that was current when it was created (else LUFQHE_ERROR_INVALID_VALUE). */
#define LUFQHE_MIXED_GRAPH 0
#define LUFQHE_MIXED_STREAMS 1
#define LUFQHE_MIXED_PLAIN 2
#define LUFQHE_MIXED_PERSISTENT 3
#define LUFQHE_MIXED_MAX_STREAMS 8
#define LUFQHE_SOURCE_MATRIX 0
#define LUFQHE_SOURCE_FQHE 1

typedef struct lufqhe_mixed_plan lufqhe_mixed_plan;

int32_t lufqhe_mixed_plan_create(const int32_t *h_n, const int64_t *h_offs,
                                 int64_t count, int32_t policy,
                                 int32_t nstreams, lufqhe_mixed_plan **plan);
int32_t lufqhe_mixed_plan_destroy(lufqhe_mixed_plan *plan);
/* Number of buckets; fills up to `capacity` of (n, count, stream) in launch
 * order (NULL arrays are skipped).  -1 for a NULL plan. */
int32_t lufqhe_mixed_plan_buckets(const lufqhe_mixed_plan *plan, int32_t *n,
                                  int64_t *count, int32_t *stream,
                                  int32_t capacity);
int32_t lufqhe_mixed_execute(lufqhe_mixed_plan *plan, const void *d_A,
                             void *d_mant, int32_t *d_exp, double *d_logabs,
                             double *d_phase, int32_t *d_info, int32_t source,
                             int32_t mode, hipStream_t stream);

/* ---- the reference's matrix input file ---------------------------------
 * LU mode reads `count` matrices of side n from one raw file
 * (LU-fqhe/src/PermHandler.cpp:91-170): elements (re, im) interleaved as
 * float32 (LUFQHE_FILE_COMPLEX64; float64 for LUFQHE_FILE_COMPLEX128, the
 * __DPREC__ build), row major, unpadded, no header.  Rank r of `nodes`
 * reads one run of scalars whose bounds the reference computes in float
 * arithmetic: first = (int)((float)(r*total)/nodes), last =
 * (int)((float)((r+1)*total)/nodes) - 1, total = 2*count*n*n (a run can
 * start or end inside a matrix, as in the reference).
 *
 * lufqhe_file_chunk: that run (first scalar, number of scalars).
 * lufqhe_read_matrices: reads the run into d_out as complex128 scalars
 * (capacity = room in d_out, in float64 scalars); complex64 data crosses
 * the bus as float32 and is widened on the device.  A run that reaches past
 * the end of the file is cut there; *read_out gets the scalars read.
 * Synchronous with respect to the host (returns after the data is on the
 * device); allocates two 32 MiB pinned staging buffers for the call. */
#define LUFQHE_FILE_COMPLEX64 0
#define LUFQHE_FILE_COMPLEX128 1

int32_t lufqhe_file_chunk(int64_t count, int32_t n, int32_t rank,
                          int32_t nodes, int64_t *first, int64_t *length);
int32_t lufqhe_read_matrices(const char *path, int32_t format, int64_t count,
                             int32_t n, int32_t rank, int32_t nodes,
                             void *d_out, int64_t capacity, int64_t *read_out,
                             hipStream_t stream);

/* ---- the reference's Monte Carlo loop, LU mode, on one GPU ---------------
 * MonteCarloLoop::perform (LU-fqhe/src/MonteCarloLoop.cu:121-401) for one
 * node with one GPU: the node's nelem matrices (host memory, complex128,
 * n x n row major, packed) are split by the reference's rule -- one GPU
 * counts as 32 host cores and takes an even share, floor(nelem / (1 +
 * cpu_count/32)) rounded down to even (LU-fqhe/src/MonteCarloLoop.cu:139-148,
 * lufqhe_mc_split) -- and each of `iterations` iterations copies the GPU
 * share to the device, runs lufqhe_det_inplace (pair products) and
 * lufqhe_global_reduction there, while the host cores sum their share's
 * pair products (lufqhe_cpu_pair_sum, LUCoreCPU::CALL); h_results[i]
 * (complex128) gets the node's sum, GPU plus host (LU-fqhe/src/
 * MonteCarloLoop.cu:360-364).  The GPU share is staged once in pinned memory,
 * so each iteration's copy overlaps the host work.  Allocates its device and
 * pinned buffers for the call.  A host share holding a row the reference
 * would exit(0) on gives NaN results and LUFQHE_ERROR_INVALID_VALUE. */
typedef struct {
  int64_t gpu_matrices, cpu_matrices; /* the split */
  double wall_s, gpu_s, cpu_s;        /* totals over the iterations */
} lufqhe_mc_stats;

int32_t lufqhe_mc_split(int64_t nelem, int32_t cpu_count, int64_t *gpu_matrices,
                        int64_t *cpu_matrices);
int32_t lufqhe_mc_lu_run(const void *h_matrices, int64_t nelem, int32_t n,
                         int32_t cpu_count, int32_t iterations, int32_t policy,
                         void *h_results, lufqhe_mc_stats *stats);

/* ---- serving: determinants for Monte Carlo walkers ----------------------
 * The reference's Monte Carlo loop (LU-fqhe/src/MonteCarloLoop.cu:280-400)
 * rebuilds all matrices on the host every iteration and runs one blocking
 * CALL per GPU.  A server keeps `walkers` FQHE configurations (n[w]
 * particles, uniform in the disk of radius sqrt(2n)) and, per walker, a table
 * of `perms_per_walker` row permutations (its symmetrisation terms) on the
 * device.  A request is one Monte Carlo step of one walker: move one particle
 * by a complex Gaussian (sigma 0.3 per component), evaluate k permutation
 * terms det(P_t A(z')) of the moved configuration and their sum psi', and
 * apply the Metropolis rule |psi'|^2 >= u |psi|^2.
 *
 * lufqhe_server_run simulates `requests` Poisson arrivals at `rate` per
 * second (rate <= 0: all at once); request r belongs to walker r mod walkers
 * and k ~ U[k_min, k_max].  Requests are admitted while the HBM budget
 * allows (resident state + batch buffers + 44 B per term in flight) and
 * packed into batches of at most max_batch_requests / max_batch_dets;
 * up to max_inflight batches run at once on their own streams, new ones
 * launched as earlier ones are still running.  Random numbers depend only on
 * (seed, request), so records are bit-identical for any rate or batching.
 * Results per request go to `records` (may be NULL); the walkers keep their
 * final configurations (lufqhe_server_reset restores the initial ones).
 * The run allocates its batch buffers (hipMalloc, pinned host memory) and
 * frees them before returning.  A server runs on the device that was current
 * when it was created (else LUFQHE_ERROR_INVALID_VALUE). */
#define LUFQHE_SERVE_MAX_INFLIGHT 8

typedef struct lufqhe_server lufqhe_server;

typedef struct {
  double rate;                /* requests per second; <= 0: all arrive at t = 0 */
  int64_t requests;
  int32_t k_min, k_max;       /* terms per request ~ U[k_min, k_max] */
  uint64_t seed;
  int32_t max_batch_requests; /* >= 1 */
  int32_t max_inflight;       /* 1 .. LUFQHE_SERVE_MAX_INFLIGHT */
  int64_t max_batch_dets;     /* >= k_max */
  int64_t hbm_budget;         /* bytes */
} lufqhe_serve_config;

typedef struct {
  int64_t requests, dets, batches;
  double wall_s;              /* first arrival to last completion */
  double dets_per_s, requests_per_s, offered_rate;
  double p50_s, p99_s, max_s, mean_s; /* arrival to completion */
  double mean_batch_dets;
  int64_t budget_waits;       /* requests held back by the budget */
  int64_t peak_inflight_bytes, resident_bytes;
} lufqhe_serve_stats;

typedef struct {
  int32_t walker, particle, k, perm_start, accepted, sum_exp;
  double delta[2];            /* the move */
  double sum_mant[2];         /* psi' = sum_mant * 2^sum_exp */
  double latency_s;
} lufqhe_serve_record;

int32_t lufqhe_server_create(int32_t walkers, const int32_t *h_walker_n,
                             int32_t perms_per_walker, uint64_t seed,
                             int32_t policy, lufqhe_server **server);
int32_t lufqhe_server_destroy(lufqhe_server *server);
int32_t lufqhe_server_reset(lufqhe_server *server);
/* Walker w's n and (if h_z) its current configuration (n complex128). */
int32_t lufqhe_server_walker(const lufqhe_server *server, int32_t walker,
                             int32_t *n, void *h_z);
/* Walker w's permutation table: perms_per_walker rows of 64 bytes (the
 * first n used). */
int32_t lufqhe_server_perms(const lufqhe_server *server, int32_t walker,
                            uint8_t *h_perm);
int32_t lufqhe_server_run(lufqhe_server *server,
                          const lufqhe_serve_config *config,
                          lufqhe_serve_stats *stats,
                          lufqhe_serve_record *records);

#ifdef __cplusplus
}
#endif
#endif /* LUFQHE_H_ */
