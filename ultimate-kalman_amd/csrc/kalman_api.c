/*
 * kalman_api.c -- the kalman_t API of libukhip (c/kalman.h:117-142) over the
 * explicit step representation, plus the covariance helpers
 * (c/kalman_base.c:61-214) and the parallel.h hints.
 *
 * Behaviour follows c/kalman_base.c:238-512 and
 * c/kalman_explicit_representation.c:84-267: evolve/observe store deep copies,
 * smooth hands the stored steps to kalman_smooth_oddeven/associative (on the
 * GPU), estimate/covariance return copies or NaN-filled matrices.
 * Differences: smooth passes the steps from the first live one after
 * kalman_forget (the reference passes the array from physical slot 0,
 * c/kalman_explicit_representation.c:250-251); ULTIMATE and CONVENTIONAL
 * (sequential filters, not part of this library) are rejected with an error.
 * kalman_create() asks for ULTIMATE, as the reference's does
 * (c/kalman_base.c:231-232), so it too returns NULL with an error: a client
 * written for the default filter never silently gets a smoother.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/kalman.h"
#include "../../include/parallel.h"

void uk_set_error(const char* fmt, ...);

double kalman_nan = 0.0 / 0.0;

static int g_thread_limit = 0;
static int g_blocksize = 16;

void parallel_set_thread_limit(int number_of_threads) {
  if (number_of_threads > 0) g_thread_limit = number_of_threads;
}
void parallel_set_blocksize(int blocksize_in) {
  if (blocksize_in > 0) g_blocksize = blocksize_in;
}
int ukhip_get_blocksize(void) { return g_blocksize; }

/* ------------------------------------------------------------------------ */
/* covariance matrices (c/kalman_base.c:61-214)                              */
/* ------------------------------------------------------------------------ */

kalman_matrix_t* kalman_covariance_matrix_weigh(kalman_matrix_t* cov, char cov_type, kalman_matrix_t* A) {
  if (!cov || !A) {
    uk_set_error("kalman_covariance_matrix_weigh: NULL argument");
    return NULL;
  }
  kalman_matrix_t* WA = NULL;
  switch (cov_type) {
    case 'W':
      WA = matrix_create_multiply(cov, A);
      break;
    case 'w':
      if (cov->row_dim != A->row_dim) break;
      WA = matrix_create_copy(A);
      for (int32_t j = 0; j < A->col_dim; j++)
        for (int32_t i = 0; i < A->row_dim; i++) WA->elements[(size_t)j * WA->ld + i] *= cov->elements[i];
      break;
    case 'C': {
      kalman_matrix_t* L = matrix_create_chol(cov);
      WA = matrix_create_trisolve("L", L, A);
      matrix_free(L);
    } break;
    case 'U':
    case 'F':
      WA = matrix_create_trisolve("L", cov, A);
      break;
    default:
      break;
  }
  if (!WA) uk_set_error("kalman_covariance_matrix_weigh: unknown type '%c' or shapes do not conform", cov_type);
  return WA;
}

kalman_matrix_t* kalman_covariance_matrix_explicit(kalman_matrix_t* cov, char type) {
  if (!cov) return NULL;
  if (type == 'w') {
    kalman_matrix_t* C = matrix_create_constant(cov->row_dim, cov->row_dim, 0.0);
    for (int32_t j = 0; j < cov->row_dim; j++) C->elements[(size_t)j * C->ld + j] = pow(cov->elements[j], -2.0);
    return C;
  }
  if (type == 'W') {
    kalman_matrix_t* WT = matrix_create_transpose(cov);
    kalman_matrix_t* WTW = matrix_create_multiply(WT, cov);
    kalman_matrix_t* C = matrix_create_inverse(WTW);
    matrix_free(WT);
    matrix_free(WTW);
    return C;
  }
  if (type == 'U' || type == 'F') {
    kalman_matrix_t* LT = matrix_create_transpose(cov);
    kalman_matrix_t* C = matrix_create_multiply(cov, LT);
    matrix_free(LT);
    return C;
  }
  if (type == 'C') return matrix_create_copy(cov);
  uk_set_error("kalman_covariance_matrix_explicit: unknown type '%c'", type);
  return NULL;
}

/* ------------------------------------------------------------------------ */
/* kalman_t over explicit steps                                               */
/* ------------------------------------------------------------------------ */

/* The steps live in an ingest object (ukhip_host.cpp): kalman_evolve and
 * kalman_observe write each step's matrices once into the device layout's
 * fields, kalman_smooth uploads them as they are. */
typedef struct uk_ingest uk_ingest_t;
uk_ingest_t* uk_ingest_create(void);
void uk_ingest_free(uk_ingest_t* g);
int64_t uk_ingest_count(const uk_ingest_t* g);
int64_t uk_ingest_first(const uk_ingest_t* g);
int32_t uk_ingest_dimension(const uk_ingest_t* g);
int32_t uk_ingest_step_dimension(const uk_ingest_t* g, int64_t si);
int uk_ingest_evolve(uk_ingest_t* g, int32_t n, const kalman_matrix_t* H, const kalman_matrix_t* F,
                     const kalman_matrix_t* c, const kalman_matrix_t* K, char K_type);
int uk_ingest_observe(uk_ingest_t* g, const kalman_matrix_t* G, const kalman_matrix_t* o, const kalman_matrix_t* C,
                      char C_type);
int uk_ingest_smooth(uk_ingest_t* g, uint32_t options);
int uk_ingest_valid(const uk_ingest_t* g, int64_t si);
const double* uk_ingest_state(const uk_ingest_t* g, int64_t si);
const double* uk_ingest_covariance(const uk_ingest_t* g, int64_t si, char* type);
void uk_ingest_forget(uk_ingest_t* g, int64_t si);
void uk_ingest_rollback(uk_ingest_t* g, int64_t si);

struct kalman_st {
  uk_ingest_t* steps;
  uint32_t options;
};

kalman_t* kalman_create_options(uint32_t options) {
  const uint32_t algo = options & (KALMAN_ALGORITHM_ULTIMATE | KALMAN_ALGORITHM_CONVENTIONAL |
                                   KALMAN_ALGORITHM_ODDEVEN | KALMAN_ALGORITHM_ASSOCIATIVE);
  if (algo != KALMAN_ALGORITHM_ODDEVEN && algo != KALMAN_ALGORITHM_ASSOCIATIVE) {
    uk_set_error("kalman_create_options: libukhip implements KALMAN_ALGORITHM_ODDEVEN and "
                 "KALMAN_ALGORITHM_ASSOCIATIVE only (options 0x%x)", options);
    return NULL;
  }
  kalman_t* k = (kalman_t*)calloc(1, sizeof(kalman_t));
  k->steps = uk_ingest_create();
  k->options = options;
  return k;
}

kalman_t* kalman_create(void) { return kalman_create_options(KALMAN_ALGORITHM_ULTIMATE); }

void kalman_free(kalman_t* k) {
  if (!k) return;
  uk_ingest_free(k->steps);
  free(k);
}

kalman_step_index_t kalman_earliest(kalman_t* k) {
  return (k && uk_ingest_count(k->steps) > 0) ? (kalman_step_index_t)uk_ingest_first(k->steps) : -1;
}
kalman_step_index_t kalman_latest(kalman_t* k) {
  return (k && uk_ingest_count(k->steps) > 0)
             ? (kalman_step_index_t)(uk_ingest_first(k->steps) + uk_ingest_count(k->steps) - 1)
             : -1;
}

void kalman_evolve(kalman_t* k, int32_t n_i, kalman_matrix_t* H_i, kalman_matrix_t* F_i, kalman_matrix_t* c_i,
                   kalman_matrix_t* K_i, char K_type) {
  if (!k) return;
  (void)uk_ingest_evolve(k->steps, n_i, H_i, F_i, c_i, K_i, K_type);
}

void kalman_observe(kalman_t* k, kalman_matrix_t* G_i, kalman_matrix_t* o_i, kalman_matrix_t* C_i, char C_type) {
  if (!k) {
    uk_set_error("kalman_observe: NULL kalman_t");
    return;
  }
  (void)uk_ingest_observe(k->steps, G_i, o_i, C_i, C_type);
}

void kalman_smooth(kalman_t* k) {
  if (!k) return;
  (void)uk_ingest_smooth(k->steps, k->options);
}

kalman_matrix_t* kalman_estimate(kalman_t* k, kalman_step_index_t si) {
  if (!k || !uk_ingest_valid(k->steps, si)) return NULL;
  const int n = uk_ingest_step_dimension(k->steps, si);
  const double* x = uk_ingest_state(k->steps, si);
  if (!x) return matrix_create_constant(n, 1, kalman_nan);
  kalman_matrix_t* m = matrix_create(n, 1);
  memcpy(m->elements, x, (size_t)n * sizeof(double));
  return m;
}

kalman_matrix_t* kalman_covariance(kalman_t* k, kalman_step_index_t si) {
  if (!k || !uk_ingest_valid(k->steps, si)) return NULL;
  const int n = uk_ingest_step_dimension(k->steps, si);
  const double* S = uk_ingest_covariance(k->steps, si, NULL);
  if (!S) return matrix_create_constant(n, n, kalman_nan);
  kalman_matrix_t* m = matrix_create(n, n);
  memcpy(m->elements, S, (size_t)n * n * sizeof(double));
  return m;
}

char kalman_covariance_type(kalman_t* k, kalman_step_index_t si) {
  char t = 'C';
  if (!k || !uk_ingest_valid(k->steps, si)) return '-';
  (void)uk_ingest_covariance(k->steps, si, &t);
  return t;
}

void kalman_forget(kalman_t* k, kalman_step_index_t si) {
  if (k) uk_ingest_forget(k->steps, si);
}

void kalman_rollback(kalman_t* k, kalman_step_index_t si) {
  if (k) uk_ingest_rollback(k->steps, si);
}

kalman_matrix_t* kalman_perftest(kalman_t* kalman, kalman_matrix_t* H, kalman_matrix_t* F, kalman_matrix_t* c,
                                 kalman_matrix_t* K, char K_type, kalman_matrix_t* G, kalman_matrix_t* o,
                                 kalman_matrix_t* C, char C_type, int32_t count, int32_t decimation) {
  (void)kalman; (void)H; (void)F; (void)c; (void)K; (void)K_type; (void)G; (void)o; (void)C; (void)C_type;
  (void)count; (void)decimation;
  uk_set_error("kalman_perftest times the sequential filters, which libukhip does not implement");
  return NULL;
}
