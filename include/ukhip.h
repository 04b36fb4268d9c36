/*
 * ukhip.h -- libukhip extensions beyond the reference's kalman.h.
 *
 * None of these has a counterpart in the reference (c/kalman.h); they serve
 * callers that keep problems resident on the GPU: the benchmark, batched
 * smoothing of many independent problems, and the serving loop.
 */
#ifndef UKHIP_H
#define UKHIP_H

#include <stdint.h>

#include "kalman.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failed call on this thread ("" if none). The reference
 * aborts instead (assert in c/matrix_ops.c and c/kalman_*.c). */
const char* ukhip_last_error(void);
void ukhip_clear_error(void);

/* Value recorded by parallel_set_blocksize (include/parallel.h); 16 by default. */
int ukhip_get_blocksize(void);

/* One problem in packed form: every per-step matrix column-major, steps
 * concatenated. Step 0 has no evolution equation (its F, H, c, K are ignored). */
typedef struct {
  int32_t n;          /* state dimension, 1 .. 64 */
  int32_t m;          /* observation rows per step (maximum), 0 .. n */
  int64_t k;          /* steps, >= 1 */
  const double* F;    /* k * n*n */
  const double* H;    /* k * n*n, or NULL for H_i = I */
  const double* c;    /* k * n, or NULL for c_i = 0 */
  const double* K;    /* k * n*n, or k * n when K_type == 'w' */
  char K_type;        /* 'W', 'w', 'C', 'U', 'F' (c/kalman_base.c:61-175) */
  const double* G;    /* k * m*n, ld m */
  const double* o;    /* k * m */
  const int32_t* m_i; /* k observation counts (0 = unobserved), or NULL for m on every step */
  const double* C;    /* k * m*m (ld m), or k * m when C_type == 'w' */
  char C_type;
  int32_t F_shared;   /* 1: F is one n*n matrix used on every step (time-invariant dynamics) */
} ukhip_problem_t;

/* A set of independent problems, resident on the GPU. Create uploads the
 * inputs (host -> HBM); smooth runs entirely on the device. Problems of one
 * state dimension share every launch of a level; a batch with several state
 * dimensions keeps one part per dimension, and the parts run concurrently. */
typedef struct ukhip_batch ukhip_batch_t;

ukhip_batch_t* ukhip_batch_create(int32_t nprob, const ukhip_problem_t* problems);
void ukhip_batch_free(ukhip_batch_t* b);
int64_t ukhip_batch_steps(const ukhip_batch_t* b);
int32_t ukhip_batch_levels(const ukhip_batch_t* b);

/* Odd-even smoothing of every problem of the batch (options: KALMAN_NO_COVARIANCE
 * honoured). Enqueued on the library stream; returns 0 or -1 (see ukhip_last_error).
 * The first call per covariance setting captures the launch sequence into a
 * HIP graph, later calls replay it (see ukhip_use_graphs). */
int ukhip_batch_smooth_oddeven(ukhip_batch_t* b, uint32_t options);
/* 0: ukhip_batch_smooth_oddeven launches kernel by kernel instead of replaying
 * a captured graph (default 1). */
void ukhip_use_graphs(int enable);
/* 1 (default): the covariance back substitution for n > 16 runs its products on
 * the FP64 matrix cores; 0: on the vector units, as for n <= 16. Applies to
 * launches and graph captures made afterwards. */
void ukhip_use_mfma(int enable);
/* Launch table of the odd-even QR phase kernels (phase 1..3) for state
 * dimensions n > 16, per precision (fp32 = 0: the FP64 kernels, 1: the FP32
 * elimination of the mixed-precision path). An entry is a compiled variant:
 * the matrix columns each group of lanes holds (columns_per_group; 2 or 3
 * share one register copy of each Householder vector) and the workgroups per
 * CU the register allocation leaves room for: 1, 2 or 4 with one column per
 * group, 2, 3 or 4 with two, 2 or 3 with three; columns_per_group 0 (phase 2,
 * buckets 32, 48, 64, one workgroup per CU) is the blocked-WY phase 2 on the
 * matrix cores. The buckets are n in 17..24, 25..32, 33..48 and 49..64.
 * ukhip_autotune times every variant on a synthetic problem of k steps (0:
 * 4096) per bucket up to n_max and keeps the fastest; the register variants
 * run the same operations on every column, so choosing among them does not
 * change results (the blocked-WY form agrees to rounding). Graphs captured
 * earlier keep their launches. get returns the workgroups per CU (0: no such
 * entry) and stores the columns per group; set and autotune return 0 or -1. */
int ukhip_autotune(int32_t n_max, int32_t k);
int ukhip_qr_launch_get(int fp32, int phase, int n, int* columns_per_group);
int ukhip_qr_launch_set(int fp32, int phase, int n, int columns_per_group, int workgroups_per_cu);
/* Associative-scan smoothing of every problem of the batch. */
int ukhip_batch_smooth_associative(ukhip_batch_t* b, uint32_t options);
int ukhip_synchronize(void);

/* Mixed precision: the odd-even elimination in FP32, then `refinements`
 * steps of FP64 iterative refinement of the estimates (corrected
 * semi-normal equations with the FP32 block factor; each step multiplies
 * the error by about cond^2 * 6e-8). Estimates only: the covariances of the
 * batch are left as they were. Enqueued on the library stream. */
int ukhip_batch_smooth_oddeven_mixed(ukhip_batch_t* b, int32_t refinements);
/* The same with options. fp32_levels: elimination levels 0 .. fp32_levels-1
 * run in FP32 and the levels above in FP64, where the reduced blocks'
 * conditioning compounds and few steps remain (<= 0: the default, 8; level
 * 0 is always FP32). covariances = 1: the covariances are formed from
 * the same block factors by the FP64 selected-inversion back substitution
 * (FP32 factors widened), into the batch's covariance output; they carry the
 * FP32 factors' accuracy, not the refined estimates'. The first call per
 * option set captures the launch sequence (every part of a composite on its
 * own forked stream) into a HIP graph, later calls replay it. */
typedef struct {
  int32_t refinements; /* FP64 refinement steps, 0..64 */
  int32_t fp32_levels;
  int32_t covariances;
} ukhip_mixed_options_t;
int ukhip_batch_smooth_oddeven_mixed_opt(ukhip_batch_t* b, const ukhip_mixed_options_t* options);

/* Copies problem p's estimates (k*n) and covariances (k*n*n, may be NULL) to the host. */
int ukhip_batch_download(ukhip_batch_t* b, int32_t p, double* states, double* covariances);
/* Problem p's covariances in one of three formats, formed on the device so that
 * only that format crosses PCIe: 'C' the n x n matrices (k*n*n), 'D' their
 * diagonals (k*n), 'W' the inverse Cholesky factor W = L^{-1} of S = L L^T
 * (W^T W = S^{-1}; lower triangular, column-major, k*n*n), the reference's 'W'
 * covariance representation (c/kalman_base.c:61-175). */
int ukhip_batch_download_cov(ukhip_batch_t* b, int32_t p, char format, double* out);
/* Option bit (not a reference option) for kalman_smooth_oddeven,
 * kalman_smooth_associative and kalman_create_options: covariances are
 * returned in the 'W' representation, covariance_type 'W', as the reference's
 * UltimateKalman filter returns them. */
#define UKHIP_COVARIANCE_W (1u << 24)
/* Option bit for ukhip_batch_smooth_oddeven(_timed) and ukhip_batch_timed_parts:
 * only the covariances' diagonals are wanted. The back substitution of the
 * first elimination level (every second step, whose covariances no other level
 * reads) then forms diag(S_j) alone and writes n values per step instead of
 * n*n; read the diagonals with ukhip_batch_download_cov(b, p, 'D', out). The
 * other entries of the covariance output are left unspecified. */
#define UKHIP_COVARIANCE_DIAG (1u << 25)

/* Timed odd-even run: device milliseconds per kernel class, measured with HIP
 * events recorded on the library stream around every launch (group) and
 * summed over the levels; the parts of a multi-dimension batch run one after
 * another. Slots: UKHIP_OE_WEIGH (k_weigh / k_weigh_diag, and k_obs_compress),
 * PHASE1..3 (k_phase1..3), BACKSUB (k_backsub, k_backsub_mfma, k_estimate),
 * TOTAL (the span from the first launch to the end of the last); for the
 * mixed-precision path also SOLVE (k_solve_est), NORMAL (k_normal, k_gather,
 * k_add), FWD (k_fwd_even, k_fwd_odd) and CONVERT (k_to32 / k_to64).
 * launches[] (may be NULL) receives the launch count of each slot. */
enum { UKHIP_OE_WEIGH = 0, UKHIP_OE_PHASE1, UKHIP_OE_PHASE2, UKHIP_OE_PHASE3, UKHIP_OE_BACKSUB, UKHIP_OE_TOTAL,
       UKHIP_OE_SOLVE, UKHIP_OE_NORMAL, UKHIP_OE_FWD, UKHIP_OE_CONVERT, UKHIP_OE_SLOTS };
int ukhip_batch_smooth_oddeven_timed(ukhip_batch_t* b, uint32_t options, double ms[UKHIP_OE_SLOTS],
                                     int32_t launches[UKHIP_OE_SLOTS]);
/* The same per part (one per state dimension), for the FP64 odd-even path
 * (mixed NULL; options: KALMAN_NO_COVARIANCE) or the mixed-precision one:
 * ms[p * UKHIP_OE_SLOTS + slot], launches likewise (may be NULL), part_n[p]
 * the part's state dimension (may be NULL). Returns the number of parts
 * (<= max_parts) or -1. */
int ukhip_batch_timed_parts(ukhip_batch_t* b, uint32_t options, const ukhip_mixed_options_t* mixed, double* ms,
                            int32_t* launches, int32_t* part_n, int32_t max_parts);

/* Timed associative run (c/kalman_associative_smoother.c:1199-1240): device
 * milliseconds per kernel class, from HIP events recorded on the library
 * stream after each class, parts one after another; added into ms and
 * launches (may be NULL), which the caller zeroes. Slots: WEIGH (the dense
 * whitening), MODEL (ka_model), PRIOR (ka_prior), FELEM (ka_felem), FSCAN
 * (the filtering scan: ka_scan_reduce, ka_scan_round, ka_scan_down), FILTERED
 * (ka_filtered), SELEM (ka_selem), SSCAN (the smoothing scan), RESULTS
 * (ka_results), TOTAL (first launch to last). */
enum { UKHIP_AS_WEIGH = 0, UKHIP_AS_MODEL, UKHIP_AS_PRIOR, UKHIP_AS_FELEM, UKHIP_AS_FSCAN, UKHIP_AS_FILTERED,
       UKHIP_AS_SELEM, UKHIP_AS_SSCAN, UKHIP_AS_RESULTS, UKHIP_AS_TOTAL, UKHIP_AS_SLOTS };
int ukhip_batch_smooth_associative_timed(ukhip_batch_t* b, uint32_t options, double ms[UKHIP_AS_SLOTS],
                                         int32_t launches[UKHIP_AS_SLOTS]);

/* ---- allocation-free odd-even smoothing ------------------------------------
 * No counterpart in the reference (c/kalman.h): for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). It is ukhip_batch_create +
 * ukhip_batch_smooth_oddeven for one problem whose arrays (every pointer of
 * `problem`) and outputs are already in device memory: states (k*n) and,
 * unless options has KALMAN_NO_COVARIANCE or covariances is NULL,
 * covariances (k*n*n), with the batch path's results bit for bit. Every
 * intermediate buffer comes from a static device array of
 * UKHIP_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; a problem
 * needing more (about 9 n^2 doubles per step) is refused with an error before
 * anything is launched. Needs 1 <= m <= n; m_i outside 0..m is clamped.
 * Enqueued on `stream` (NULL: the library stream); returns 0 or -1
 * (ukhip_last_error). No hipMalloc, hipFree or host allocation is made. All
 * calls on one GPU share its scratch: calls that overlap on one GPU race. */
#define UKHIP_NOALLOC_SCRATCH_BYTES (64u << 20)
int ukhip_smooth_oddeven_noalloc(const ukhip_problem_t* problem, uint32_t options, double* states,
                                 double* covariances, void* stream);

/* ---- serving: continuous batching of smoothing requests -------------------
 * A server owns a device memory arena (the HBM budget) and a scheduler
 * thread. Submitted requests (their inputs are copied) wait in arrival
 * order; whenever a stream slot and arena memory are free, the scheduler
 * takes the oldest request and every queued request of the same state
 * dimension that still fits, uploads them as one batch and runs it, while
 * earlier batches are still on the device. A result depends only on its own
 * request: it is the same alone and in any mix. */
typedef struct ukhip_server ukhip_server_t;
typedef struct {
  size_t hbm_budget_bytes;   /* device arena; 0: 40% of the free device memory */
  int32_t max_batch;         /* requests per batch; 0: 256 */
  int32_t max_inflight;      /* batches on the device at once (one stream each); 0: 4 */
  int32_t mixed_refinements; /* < 0: FP64 odd-even with covariances; >= 0: FP32 elimination plus that many
                                FP64 refinement steps, estimates only (covariances read as NaN) */
} ukhip_server_options_t;
ukhip_server_t* ukhip_server_create(const ukhip_server_options_t* options);
/* Returns the request id (>= 0), or -1. */
int64_t ukhip_server_submit(ukhip_server_t* s, const ukhip_problem_t* problem);
/* 1 done, 0 pending, -1 failed or unknown. */
int ukhip_server_poll(ukhip_server_t* s, int64_t id);
/* Blocks until request id is done; copies its estimates (k*n) and
 * covariances (k*n*n, may be NULL) out and gives the time from submission to
 * completion; forgets the request. 0 or -1 (ukhip_last_error). */
int ukhip_server_wait(ukhip_server_t* s, int64_t id, double* states, double* covariances, double* latency_ms);
/* Finishes every submitted request, then frees the server. */
void ukhip_server_destroy(ukhip_server_t* s);

/* ---- batched sequential filter (UltimateKalman) ------------------------------
 * nfilters independent filters of one state dimension n advance together, one
 * evolve + observe step per call, one launch for all of them: the reference's
 * KALMAN_ALGORITHM_ULTIMATE filter (c/kalman_ultimate.c:140-470: kalman_evolve,
 * kalman_observe, kalman_estimate, kalman_covariance) batched on the device.
 * The filter state stays in HBM; each step's inputs are host arrays holding
 * all filters' matrices, column-major, filter after filter. */
typedef struct ukhip_filter ukhip_filter_t;
typedef struct {
  const double* F;    /* nfilters * n*n, or one n*n when F_shared (ignored on the first step) */
  int32_t F_shared;
  const double* H;    /* nfilters * n*n, or NULL for H = I */
  const double* c;    /* nfilters * n, or NULL for c = 0 */
  const double* K;    /* nfilters * n: the 'w' representation (inverse standard deviations) */
  const double* G;    /* nfilters * m*n, ld m */
  const double* o;    /* nfilters * m */
  const int32_t* m_i; /* nfilters observation counts (0 = unobserved), or NULL for m */
  const double* C;    /* nfilters * m: 'w' representation */
  /* representations of K and C for this step, as kalman_evolve / kalman_observe
   * take them (c/kalman.h:126-127): 0 or 'w' the vectors above; 'W', 'C', 'U'
   * or 'F' a matrix per filter (K: nfilters * n*n; C: nfilters * m*m, ld m),
   * whitened on the device first (c/kalman_base.c:61-175) */
  char K_type;
  char C_type;
} ukhip_filter_inputs_t;
ukhip_filter_t* ukhip_filter_create(int32_t nfilters, int32_t n, int32_t m);
/* One step of every filter (the first call has no evolution equation). */
int ukhip_filter_step(ukhip_filter_t* f, const ukhip_filter_inputs_t* in);
/* The same step with every array of `in` (F, H, c, K, G, o, m_i, C) already in
 * device memory, in the layout above: nothing is copied, the launch reads them
 * in place on the library stream (they must stay valid until the step has
 * run; ukhip_synchronize). G, o and C are required; m_i values outside 0..m are
 * clamped into it. No counterpart in the reference: for callers whose inputs
 * are produced on the GPU. Returns 0 or -1. */
int ukhip_filter_step_device(ukhip_filter_t* f, const ukhip_filter_inputs_t* in);
/* Keeps every step's final factors (Rdiag, Rsupdiag, y of the reference's
 * UltimateKalman step) in device memory for ukhip_filter_smooth: at most
 * max_steps steps (nfilters * (max_steps - 1) * (2 n*n + n) doubles). Call it
 * before the first step. Returns 0 or -1. */
int ukhip_filter_keep_history(ukhip_filter_t* f, int64_t max_steps);
/* The smoother pass of the reference's UltimateKalman (c/kalman_ultimate.c:
 * 559-620) over the steps so far, on the device: states (steps * nfilters *
 * n: step after step, filter after filter) and, if W is not NULL, every
 * step's covariance in the 'W' representation (steps * nfilters * n*n,
 * column-major; covariance = (W^T W)^{-1}, as kalman_covariance returns with
 * type 'W' after kalman_smooth). Needs ukhip_filter_keep_history. The filter
 * state is unchanged (further steps may follow). Returns 0 or -1. */
int ukhip_filter_smooth(ukhip_filter_t* f, double* states, double* W);
/* The latest filtered estimates (nfilters * n; NaN while a filter's state is
 * not yet determined, as the reference returns) and, if R is not NULL, the
 * covariances in the reference's 'W' representation: R (nfilters * n*n, upper
 * triangular, column-major) with R^T R the inverse covariance. */
int ukhip_filter_estimates(ukhip_filter_t* f, double* states, double* R);
void ukhip_filter_free(ukhip_filter_t* f);

/* The HIP stream (hipStream_t) every libukhip launch uses. */
void* ukhip_stream(void);

#ifdef __cplusplus
}
#endif

#endif /* UKHIP_H */
