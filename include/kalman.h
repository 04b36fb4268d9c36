/*
 * kalman.h -- drop-in C ABI of libukhip, the MI355X (gfx950) implementation of
 * UltimateKalman's two parallel-in-time linear smoothers.
 *
 * Every declaration below replaces the reference declaration cited beside it
 * (c/kalman.h). Argument meaning, ownership and return conventions are the
 * reference's; the implementation behind them packs the step equations into
 * device structure-of-arrays buffers and runs hand-written HIP kernels.
 *
 * Errors: the reference aborts on assertion failures (c/matrix_ops.c,
 * c/kalman_*.c). libukhip never aborts: a failing call leaves its outputs
 * untouched (smoothers) or returns NULL / '-' , and records a message readable
 * through ukhip_last_error() (include/ukhip.h).
 *
 * Step indices: kalman_step_index_t is int32_t unless one of the reference's
 * selector macros picks another width (c/kalman.h:24-38).
 */
#ifndef ULTIMATE_KALMAN_H
#define ULTIMATE_KALMAN_H

#include <stdio.h>
#include <stdlib.h>
#include <stdint.h>

#include "matrix_ops.h"

#ifdef __cplusplus
extern "C" {
#endif

extern double kalman_nan; /* c/kalman.h:20 */

#if defined(KALMAN_STEP_INDEX_TYPE_INT64)
typedef int64_t kalman_step_index_t;
#elif defined(KALMAN_STEP_INDEX_TYPE_UINT64)
typedef uint64_t kalman_step_index_t;
#elif defined(KALMAN_STEP_INDEX_TYPE_UINT32)
typedef uint32_t kalman_step_index_t;
#else
typedef int32_t kalman_step_index_t; /* c/build.sh:4 builds with INT32 */
#endif

/* ---- covariance matrices (c/kalman.h:44-45) --------------------------------
 * cov_type 'W': cov is W with W^T W = C^{-1};  weigh(A) = W*A
 *          'w': cov is a column of diagonal weights; weigh(A) = diag(w)*A
 *          'C': cov is C itself;   weigh(A) = L^{-1} A with C = L L^T
 *          'U','F': cov is the lower factor L (C = L L^T); weigh(A) = L^{-1} A */
kalman_matrix_t* kalman_covariance_matrix_weigh(kalman_matrix_t* cov, char cov_type, kalman_matrix_t* A);
kalman_matrix_t* kalman_covariance_matrix_explicit(kalman_matrix_t* cov, char type);

/* ---- step equations (c/kalman.h:56-75) -------------------------------------
 * Step i (i > 0):  H_i u_i = F_i u_{i-1} + c_i + eps_i,  Cov(eps_i) given by (K, K_type)
 * Every step:      o_i     = G_i u_i + delta_i,          Cov(delta_i) given by (C, C_type)
 * o == NULL marks an unobserved step. The smoothers allocate `state`
 * (dimension x 1) and `covariance` (dimension x dimension, type 'C') with
 * matrix_create; the caller frees them with matrix_free. */
typedef struct kalman_step_equations_st {
  kalman_step_index_t step; /* logical step number */
  int32_t dimension;

  char C_type;
  char K_type;

  kalman_matrix_t* H;
  kalman_matrix_t* F;
  kalman_matrix_t* K;
  kalman_matrix_t* c;

  kalman_matrix_t* G;
  kalman_matrix_t* o;
  kalman_matrix_t* C;

  kalman_matrix_t* state;
  kalman_matrix_t* covariance;
  char covariance_type;
} kalman_step_equations_t;

/* ---- options (c/kalman.h:81-88), bit-exact ---------------------------------- */
typedef enum {
  KALMAN_NONE = 0,
  KALMAN_ALGORITHM_ULTIMATE = 1 << 0,
  KALMAN_ALGORITHM_CONVENTIONAL = 1 << 1,
  KALMAN_ALGORITHM_ODDEVEN = 1 << 2,
  KALMAN_ALGORITHM_ASSOCIATIVE = 1 << 3,
  KALMAN_NO_COVARIANCE = 1 << 16
} kalman_options_t;

/* ---- polymorphic filter/smoother object (c/kalman.h:92-142) ----------------
 * Opaque here: clients of the reference only use it through these functions.
 * libukhip implements the explicit-representation algorithms (ODDEVEN,
 * ASSOCIATIVE); creating ULTIMATE or CONVENTIONAL returns NULL with an error. */
typedef struct kalman_st kalman_t;

kalman_t* kalman_create(void);                      /* c/kalman.h:117; default ODDEVEN here */
kalman_t* kalman_create_options(uint32_t options);  /* c/kalman.h:118 */
void kalman_free(kalman_t* kalman);                 /* c/kalman.h:119 */

kalman_step_index_t kalman_earliest(kalman_t* kalman); /* c/kalman.h:121 */
kalman_step_index_t kalman_latest(kalman_t* kalman);   /* c/kalman.h:122 */

/* c/kalman.h:124-127 */
void kalman_evolve(kalman_t* kalman, int32_t n_i, kalman_matrix_t* H_i, kalman_matrix_t* F_i, kalman_matrix_t* c_i,
                   kalman_matrix_t* K_i, char K_type);
void kalman_observe(kalman_t* kalman, kalman_matrix_t* G_i, kalman_matrix_t* o_i, kalman_matrix_t* C_i, char C_type);
void kalman_smooth(kalman_t* kalman);

/* c/kalman.h:129-133; si < 0 means the latest step. Copies; NaN-filled when
 * the step has no estimate (c/kalman_base.c:329-348, 437-460). */
kalman_matrix_t* kalman_estimate(kalman_t* kalman, kalman_step_index_t si);
kalman_matrix_t* kalman_covariance(kalman_t* kalman, kalman_step_index_t si);
char kalman_covariance_type(kalman_t* kalman, kalman_step_index_t si);
void kalman_forget(kalman_t* kalman, kalman_step_index_t si);
void kalman_rollback(kalman_t* kalman, kalman_step_index_t si);

/* c/kalman.h:134-142: filter-throughput test of the sequential filters; those
 * are outside this library, so it returns NULL and records an error. */
kalman_matrix_t* kalman_perftest(kalman_t* kalman, kalman_matrix_t* H, kalman_matrix_t* F, kalman_matrix_t* c,
                                 kalman_matrix_t* K, char K_type, kalman_matrix_t* G, kalman_matrix_t* o,
                                 kalman_matrix_t* C, char C_type, int32_t count, int32_t decimation);

/* ---- parallel smoothers (c/kalman.h:148-149): the hot path -------------------
 * Odd-even (Gargir-Toledo) elimination smoother and associative-scan
 * (Saerkkae-Garcia-Fernandez) smoother, on the GPU. Synchronous: on return the
 * estimates are in equations[i]->state / ->covariance. */
void kalman_smooth_oddeven(kalman_options_t options, kalman_step_equations_t** equations, kalman_step_index_t length);
void kalman_smooth_associative(kalman_options_t options, kalman_step_equations_t** equations,
                               kalman_step_index_t length);

#ifdef __cplusplus
}
#endif

#endif /* ULTIMATE_KALMAN_H */
