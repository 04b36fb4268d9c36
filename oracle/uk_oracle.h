/*
 * uk_oracle.h -- CPU FP64 restatement of UltimateKalman's parallel smoothers.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline; the product (libukhip) never calls it.
 *
 * Functions take the reference's own step-equation structs
 * (c/kalman.h:56-75) so one harness drives the product, this oracle and the
 * reference build in oracle/_ref alike. Outputs (state, covariance) are
 * allocated with oracle_matrix_create and freed with oracle_matrix_free.
 */
#ifndef UK_ORACLE_H
#define UK_ORACLE_H

#include "../include/kalman.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Odd-even smoother, restating c/kalman_oddeven_smoother.c:379-975,1358-1712
 * and matlab/KalmanOddevenSmoother.m:21-358 (D1 fixed as in the MATLAB). */
void oracle_smooth_oddeven(kalman_options_t options, kalman_step_equations_t** eq, kalman_step_index_t k);

/* Associative smoother, restating c/kalman_associative_smoother.c:309-1240 and
 * matlab/KalmanAssociativeSmoother.m:20-262, with D2 (unobserved steps), D3
 * (step-0 prior that does not determine the state), D4 (last covariance) and
 * D5 (k == 1) handled; see the header comment in uk_oracle.c. */
void oracle_smooth_associative(kalman_options_t options, kalman_step_equations_t** eq, kalman_step_index_t k);

/* Global least-squares definition (matlab/KalmanSparse.m:142-306): dense
 * Householder QR of the whitened block-bidiagonal matrix. states: sum(n_i)
 * doubles; covs: sum(n_i^2) doubles (diagonal blocks of (A^T A)^{-1}).
 * Returns 0 on success. Intended for sum(n_i) up to a few hundred. */
int oracle_dense_ls(kalman_step_equations_t** eq, kalman_step_index_t k, double* states, double* covs);

/* LAWN-41 FLOP replay of the reference's LAPACK/BLAS call shapes, accumulated
 * over the last oracle_smooth_* call made with counting enabled. */
void   oracle_count_flops(int enable);
double oracle_flops(void);

/* Threads for the per-pair / per-step loops (OpenMP), like the reference's TBB
 * parallel_for with grain 16 (c/parallel_tbb.cpp:45-75). 1 = sequential. */
void oracle_set_threads(int nthreads);

kalman_matrix_t* oracle_matrix_create(int32_t rows, int32_t cols);
void             oracle_matrix_free(kalman_matrix_t* A);

#ifdef __cplusplus
}
#endif

#endif
