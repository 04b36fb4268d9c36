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
 * (sequential filters, not part of this library) are rejected with an error;
 * kalman_create() therefore defaults to ODDEVEN.
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

struct kalman_st {
  kalman_step_equations_t** steps; /* live steps are steps[start .. start+count-1] */
  int64_t cap, start, count;
  kalman_step_index_t first_index; /* logical index of steps[start] */
  kalman_step_equations_t* current;
  uint32_t options;
};

static kalman_step_equations_t* step_new(void) {
  kalman_step_equations_t* s = (kalman_step_equations_t*)calloc(1, sizeof(kalman_step_equations_t));
  s->step = -1;
  s->dimension = -1;
  s->covariance_type = 'C';
  return s;
}

static void step_clear_observation(kalman_step_equations_t* s) {
  matrix_free(s->G);
  matrix_free(s->o);
  matrix_free(s->C);
  matrix_free(s->state);
  matrix_free(s->covariance);
  s->G = s->o = s->C = s->state = s->covariance = NULL;
}

static void step_free(kalman_step_equations_t* s) {
  if (!s) return;
  matrix_free(s->H);
  matrix_free(s->F);
  matrix_free(s->K);
  matrix_free(s->c);
  step_clear_observation(s);
  free(s);
}

kalman_t* kalman_create_options(uint32_t options) {
  const uint32_t algo = options & (KALMAN_ALGORITHM_ULTIMATE | KALMAN_ALGORITHM_CONVENTIONAL |
                                   KALMAN_ALGORITHM_ODDEVEN | KALMAN_ALGORITHM_ASSOCIATIVE);
  if (algo != KALMAN_ALGORITHM_ODDEVEN && algo != KALMAN_ALGORITHM_ASSOCIATIVE) {
    uk_set_error("kalman_create_options: libukhip implements KALMAN_ALGORITHM_ODDEVEN and "
                 "KALMAN_ALGORITHM_ASSOCIATIVE only (options 0x%x)", options);
    return NULL;
  }
  kalman_t* k = (kalman_t*)calloc(1, sizeof(kalman_t));
  k->cap = 1024;
  k->steps = (kalman_step_equations_t**)calloc((size_t)k->cap, sizeof(*k->steps));
  k->options = options;
  return k;
}

kalman_t* kalman_create(void) { return kalman_create_options(KALMAN_ALGORITHM_ODDEVEN); }

void kalman_free(kalman_t* k) {
  if (!k) return;
  for (int64_t i = 0; i < k->count; i++) step_free(k->steps[k->start + i]);
  step_free(k->current);
  free(k->steps);
  free(k);
}

kalman_step_index_t kalman_earliest(kalman_t* k) { return (k && k->count > 0) ? k->first_index : -1; }
kalman_step_index_t kalman_latest(kalman_t* k) {
  return (k && k->count > 0) ? k->first_index + (kalman_step_index_t)(k->count - 1) : -1;
}

static kalman_step_equations_t* step_at(kalman_t* k, kalman_step_index_t si) {
  if (!k || k->count == 0) return NULL;
  if (si < 0) si = kalman_latest(k);
  if (si < k->first_index || si > kalman_latest(k)) {
    uk_set_error("step %lld is not in memory", (long long)si);
    return NULL;
  }
  return k->steps[k->start + (si - k->first_index)];
}

static void append(kalman_t* k, kalman_step_equations_t* s) {
  if (k->start + k->count == k->cap) {
    if (k->count <= k->cap / 2) { /* shift back */
      memmove(k->steps, k->steps + k->start, (size_t)k->count * sizeof(*k->steps));
      k->start = 0;
    } else {
      k->cap *= 2;
      k->steps = (kalman_step_equations_t**)realloc(k->steps, (size_t)k->cap * sizeof(*k->steps));
    }
  }
  k->steps[k->start + k->count++] = s;
}

void kalman_evolve(kalman_t* k, int32_t n_i, kalman_matrix_t* H_i, kalman_matrix_t* F_i, kalman_matrix_t* c_i,
                   kalman_matrix_t* K_i, char K_type) {
  if (!k) return;
  step_free(k->current);
  kalman_step_equations_t* s = k->current = step_new();
  s->dimension = n_i;
  if (k->count == 0) {
    s->step = 0; /* the first step has no evolution equation */
    return;
  }
  s->step = kalman_latest(k) + 1;
  if (!H_i || !F_i || !c_i) uk_set_error("kalman_evolve: H, F and c are required after the first step");
  s->H = matrix_create_copy(H_i);
  s->F = matrix_create_copy(F_i);
  s->c = matrix_create_copy(c_i);
  s->K = matrix_create_copy(K_i);
  s->K_type = K_type;
}

void kalman_observe(kalman_t* k, kalman_matrix_t* G_i, kalman_matrix_t* o_i, kalman_matrix_t* C_i, char C_type) {
  if (!k || !k->current) {
    uk_set_error("kalman_observe: call kalman_evolve first");
    return;
  }
  kalman_step_equations_t* s = k->current;
  if (o_i != NULL) {
    s->G = matrix_create_copy(G_i);
    s->o = matrix_create_copy(o_i);
    s->C = matrix_create_copy(C_i);
    s->C_type = C_type;
  }
  if (k->count == 0) k->first_index = s->step;
  append(k, s);
  k->current = NULL;
}

void kalman_smooth(kalman_t* k) {
  if (!k || k->count == 0) return;
  kalman_step_equations_t** live = k->steps + k->start;
  if (k->options & KALMAN_ALGORITHM_ODDEVEN)
    kalman_smooth_oddeven((kalman_options_t)k->options, live, (kalman_step_index_t)k->count);
  if (k->options & KALMAN_ALGORITHM_ASSOCIATIVE)
    kalman_smooth_associative((kalman_options_t)k->options, live, (kalman_step_index_t)k->count);
}

kalman_matrix_t* kalman_estimate(kalman_t* k, kalman_step_index_t si) {
  kalman_step_equations_t* s = step_at(k, si);
  if (!s) return NULL;
  if (!s->state) return matrix_create_constant(s->dimension, 1, kalman_nan);
  return matrix_create_copy(s->state);
}

kalman_matrix_t* kalman_covariance(kalman_t* k, kalman_step_index_t si) {
  kalman_step_equations_t* s = step_at(k, si);
  if (!s) return NULL;
  if (!s->covariance) return matrix_create_constant(s->dimension, s->dimension, kalman_nan);
  return matrix_create_copy(s->covariance);
}

char kalman_covariance_type(kalman_t* k, kalman_step_index_t si) {
  kalman_step_equations_t* s = step_at(k, si);
  return s ? s->covariance_type : '-';
}

void kalman_forget(kalman_t* k, kalman_step_index_t si) {
  if (!k || k->count == 0) return;
  if (si < 0) si = kalman_latest(k) - 1;
  if (si > kalman_latest(k) - 1) return; /* the last step is never dropped */
  while (k->count > 0 && k->first_index <= si) {
    step_free(k->steps[k->start]);
    k->start++;
    k->count--;
    k->first_index++;
  }
}

void kalman_rollback(kalman_t* k, kalman_step_index_t si) {
  if (!k || k->count == 0) return;
  if (si > kalman_latest(k) || si < k->first_index) return;
  while (k->count > 0) {
    kalman_step_equations_t* s = k->steps[k->start + k->count - 1];
    k->count--;
    if (s->step == si) { /* back to just after evolve */
      step_clear_observation(s);
      step_free(k->current);
      k->current = s;
      break;
    }
    step_free(s);
  }
}

kalman_matrix_t* kalman_perftest(kalman_t* kalman, kalman_matrix_t* H, kalman_matrix_t* F, kalman_matrix_t* c,
                                 kalman_matrix_t* K, char K_type, kalman_matrix_t* G, kalman_matrix_t* o,
                                 kalman_matrix_t* C, char C_type, int32_t count, int32_t decimation) {
  (void)kalman; (void)H; (void)F; (void)c; (void)K; (void)K_type; (void)G; (void)o; (void)C; (void)C_type;
  (void)count; (void)decimation;
  uk_set_error("kalman_perftest times the sequential filters, which libukhip does not implement");
  return NULL;
}
