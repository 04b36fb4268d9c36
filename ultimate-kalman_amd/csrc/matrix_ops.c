/*
 * matrix_ops.c -- host implementation of the matrix helpers libukhip exports
 * for drop-in clients (names and semantics: c/matrix_ops.h:137-236).
 *
 * These are small host utilities used by client code (building inputs,
 * reading results); the smoothers themselves never call them on the hot path.
 * Errors are recorded with uk_set_error; no call aborts.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/matrix_ops.h"

void uk_set_error(const char* fmt, ...);

#define EL(A, i, j) ((A)->elements[(size_t)(j) * (A)->ld + (i)])
#define MIN(a, b) ((a) < (b) ? (a) : (b))
#define MAX(a, b) ((a) > (b) ? (a) : (b))

kalman_matrix_t* matrix_create(int32_t rows, int32_t cols) {
  if (rows < 0 || cols < 0) {
    uk_set_error("matrix_create: negative size");
    return NULL;
  }
  kalman_matrix_t* A = (kalman_matrix_t*)malloc(sizeof(kalman_matrix_t));
  if (!A) return NULL;
  A->row_dim = rows;
  A->col_dim = cols;
  A->ld = rows;
  A->elements = (rows > 0 && cols > 0) ? (double*)malloc((size_t)rows * cols * sizeof(double)) : NULL;
  return A;
}

void matrix_free(kalman_matrix_t* A) {
  if (!A) return;
  free(A->elements);
  free(A);
}

int32_t matrix_rows(kalman_matrix_t* A) { return A ? A->row_dim : 0; }
int32_t matrix_cols(kalman_matrix_t* A) { return A ? A->col_dim : 0; }
int32_t matrix_ld(kalman_matrix_t* A) { return A ? A->ld : 0; }

void matrix_set(kalman_matrix_t* A, int32_t i, int32_t j, double v) {
  if (!A || i < 0 || j < 0 || i >= A->row_dim || j >= A->col_dim) {
    uk_set_error("matrix_set: index (%d,%d) out of range", i, j);
    return;
  }
  EL(A, i, j) = v;
}

double matrix_get(kalman_matrix_t* A, int32_t i, int32_t j) {
  if (!A || i < 0 || j < 0 || i >= A->row_dim || j >= A->col_dim) {
    uk_set_error("matrix_get: index (%d,%d) out of range", i, j);
    return 0.0 / 0.0;
  }
  return EL(A, i, j);
}

kalman_matrix_t* matrix_create_constant(int32_t rows, int32_t cols, double c) {
  kalman_matrix_t* A = matrix_create(rows, cols);
  if (!A) return NULL;
  for (size_t i = 0; i < (size_t)rows * cols; i++) A->elements[i] = c;
  return A;
}

kalman_matrix_t* matrix_create_identity(int32_t rows, int32_t cols) {
  kalman_matrix_t* A = matrix_create_constant(rows, cols, 0.0);
  if (!A) return NULL;
  for (int32_t i = 0; i < MIN(rows, cols); i++) EL(A, i, i) = 1.0;
  return A;
}

kalman_matrix_t* matrix_create_from_rowwise(double* naked, int32_t rows, int32_t cols) {
  kalman_matrix_t* A = matrix_create(rows, cols);
  if (!A) return NULL;
  for (int32_t i = 0; i < rows; i++)
    for (int32_t j = 0; j < cols; j++) EL(A, i, j) = naked[(size_t)i * cols + j];
  return A;
}

kalman_matrix_t* matrix_create_copy(kalman_matrix_t* A) {
  if (!A) return NULL;
  kalman_matrix_t* C = matrix_create(A->row_dim, A->col_dim);
  if (!C) return NULL;
  for (int32_t j = 0; j < A->col_dim; j++)
    for (int32_t i = 0; i < A->row_dim; i++) EL(C, i, j) = EL(A, i, j);
  return C;
}

kalman_matrix_t* matrix_create_sub(kalman_matrix_t* A, int32_t r0, int32_t rows, int32_t c0, int32_t cols) {
  if (!A || r0 < 0 || c0 < 0 || r0 + rows > A->row_dim || c0 + cols > A->col_dim) {
    uk_set_error("matrix_create_sub: block out of range");
    return NULL;
  }
  kalman_matrix_t* C = matrix_create(rows, cols);
  if (!C) return NULL;
  for (int32_t j = 0; j < cols; j++)
    for (int32_t i = 0; i < rows; i++) EL(C, i, j) = EL(A, r0 + i, c0 + j);
  return C;
}

void matrix_mutate_copy_sub(kalman_matrix_t* C, int32_t r0, int32_t c0, kalman_matrix_t* A) {
  if (!A && !C) return;
  if (!A || !C || C->row_dim < r0 + A->row_dim || C->col_dim < c0 + A->col_dim) {
    uk_set_error("matrix_mutate_copy_sub: block out of range");
    return;
  }
  for (int32_t j = 0; j < A->col_dim; j++)
    for (int32_t i = 0; i < A->row_dim; i++) EL(C, r0 + i, c0 + j) = EL(A, i, j);
}

void matrix_mutate_copy(kalman_matrix_t* C, kalman_matrix_t* A) {
  if (!A && !C) return;
  if (!A || !C || A->row_dim != C->row_dim || A->col_dim != C->col_dim) {
    uk_set_error("matrix_mutate_copy: shapes differ");
    return;
  }
  matrix_mutate_copy_sub(C, 0, 0, A);
}

void matrix_print(kalman_matrix_t* A, char* format) {
  if (!A) {
    printf("matrix_print NULL\n");
    return;
  }
  printf("matrix_print %d %d\n", A->row_dim, A->col_dim);
  if (!format) format = "%f";
  for (int32_t i = 0; i < A->row_dim; i++) {
    for (int32_t j = 0; j < A->col_dim; j++) {
      printf(format, EL(A, i, j));
      printf(" ");
    }
    printf("\n");
  }
}

void matrix_mutate_scale(kalman_matrix_t* A, double s) {
  if (!A) return;
  for (int32_t j = 0; j < A->col_dim; j++)
    for (int32_t i = 0; i < A->row_dim; i++) EL(A, i, j) *= s;
}

/* the reference zeroes below the diagonal of the leading rows x rows block */
void matrix_mutate_triu(kalman_matrix_t* A) {
  if (!A) return;
  for (int32_t j = 0; j < A->col_dim; j++)
    for (int32_t i = j + 1; i < A->row_dim; i++) EL(A, i, j) = 0.0;
}

void matrix_mutate_chop(kalman_matrix_t* A, int32_t rows, int32_t cols) {
  if (!A || rows < 0 || cols < 0 || rows > A->row_dim || cols > A->col_dim) {
    uk_set_error("matrix_mutate_chop: bad size");
    return;
  }
  A->row_dim = rows;
  A->col_dim = cols;
}

void matrix_mutate_gemm(double alpha, kalman_matrix_t* A, kalman_matrix_t* B, double beta, kalman_matrix_t* C) {
  if (!A || !B || !C || A->col_dim != B->row_dim || C->row_dim != A->row_dim || C->col_dim != B->col_dim) {
    uk_set_error("matrix_mutate_gemm: shapes do not conform");
    return;
  }
  const int32_t m = C->row_dim, n = C->col_dim, k = A->col_dim;
  double* acc = (double*)calloc((size_t)m > 0 ? (size_t)m : 1, sizeof(double));
  for (int32_t j = 0; j < n; j++) {
    memset(acc, 0, sizeof(double) * (size_t)m);
    for (int32_t p = 0; p < k; p++) {
      const double b = EL(B, p, j);
      for (int32_t i = 0; i < m; i++) acc[i] += EL(A, i, p) * b;
    }
    for (int32_t i = 0; i < m; i++) EL(C, i, j) = alpha * acc[i] + (beta == 0.0 ? 0.0 : beta * EL(C, i, j));
  }
  free(acc);
}

kalman_matrix_t* matrix_create_mutate_qr(kalman_matrix_t* A) {
  if (!A) {
    uk_set_error("matrix_create_mutate_qr: NULL");
    return NULL;
  }
  const int32_t m = A->row_dim, n = A->col_dim, kk = MIN(m, n);
  kalman_matrix_t* tau = matrix_create_constant(MAX(n, 1), 1, 0.0);
  for (int32_t k = 0; k < kk; k++) {
    double alpha = EL(A, k, k), ss = 0.0;
    for (int32_t i = k + 1; i < m; i++) ss += EL(A, i, k) * EL(A, i, k);
    if (ss == 0.0) continue;
    const double beta = -copysign(sqrt(alpha * alpha + ss), alpha);
    const double t = (beta - alpha) / beta, s = 1.0 / (alpha - beta);
    for (int32_t i = k + 1; i < m; i++) EL(A, i, k) *= s;
    EL(A, k, k) = beta;
    tau->elements[k] = t;
    for (int32_t j = k + 1; j < n; j++) {
      double w = EL(A, k, j);
      for (int32_t i = k + 1; i < m; i++) w += EL(A, i, k) * EL(A, i, j);
      w *= t;
      EL(A, k, j) -= w;
      for (int32_t i = k + 1; i < m; i++) EL(A, i, j) -= w * EL(A, i, k);
    }
  }
  return tau;
}

void matrix_mutate_apply_qt(kalman_matrix_t* QR, kalman_matrix_t* TAU, kalman_matrix_t* C) {
  if (!QR || !TAU || !C || C->row_dim != QR->row_dim) {
    uk_set_error("matrix_mutate_apply_qt: shapes do not conform");
    return;
  }
  const int32_t m = C->row_dim, kk = MIN(QR->row_dim, QR->col_dim);
  for (int32_t k = 0; k < kk; k++) {
    const double t = TAU->elements[k];
    if (t == 0.0) continue;
    for (int32_t j = 0; j < C->col_dim; j++) {
      double w = EL(C, k, j);
      for (int32_t i = k + 1; i < m; i++) w += EL(QR, i, k) * EL(C, i, j);
      w *= t;
      EL(C, k, j) -= w;
      for (int32_t i = k + 1; i < m; i++) EL(C, i, j) -= w * EL(QR, i, k);
    }
  }
}

void matrix_mutate_trisolve(char* triangle, kalman_matrix_t* U, kalman_matrix_t* b) {
  if (!U || !b || U->row_dim != U->col_dim || b->row_dim != U->col_dim) {
    uk_set_error("matrix_mutate_trisolve: shapes do not conform");
    return;
  }
  const int32_t n = U->row_dim;
  const int upper = triangle && (triangle[0] == 'U' || triangle[0] == 'u');
  for (int32_t j = 0; j < b->col_dim; j++) {
    if (upper) {
      for (int32_t i = n - 1; i >= 0; i--) {
        double s = EL(b, i, j);
        for (int32_t q = i + 1; q < n; q++) s -= EL(U, i, q) * EL(b, q, j);
        EL(b, i, j) = s / EL(U, i, i);
      }
    } else {
      for (int32_t i = 0; i < n; i++) {
        double s = EL(b, i, j);
        for (int32_t q = 0; q < i; q++) s -= EL(U, i, q) * EL(b, q, j);
        EL(b, i, j) = s / EL(U, i, i);
      }
    }
  }
}

kalman_matrix_t* matrix_create_trisolve(char* triangle, kalman_matrix_t* U, kalman_matrix_t* b) {
  kalman_matrix_t* x = matrix_create_copy(b);
  matrix_mutate_trisolve(triangle, U, x);
  return x;
}

void matrix_mutate_chol(kalman_matrix_t* A) {
  if (!A || A->row_dim != A->col_dim) {
    uk_set_error("matrix_mutate_chol: not square");
    return;
  }
  const int32_t n = A->row_dim;
  for (int32_t j = 0; j < n; j++) {
    double d = EL(A, j, j);
    for (int32_t q = 0; q < j; q++) d -= EL(A, j, q) * EL(A, j, q);
    if (!(d > 0.0)) {
      uk_set_error("matrix_mutate_chol: matrix is not positive definite");
      return;
    }
    d = sqrt(d);
    EL(A, j, j) = d;
    for (int32_t i = j + 1; i < n; i++) {
      double s = EL(A, i, j);
      for (int32_t q = 0; q < j; q++) s -= EL(A, i, q) * EL(A, j, q);
      EL(A, i, j) = s / d;
    }
  }
  for (int32_t j = 1; j < n; j++)
    for (int32_t i = 0; i < j; i++) EL(A, i, j) = 0.0;
}

kalman_matrix_t* matrix_create_chol(kalman_matrix_t* A) {
  kalman_matrix_t* L = matrix_create_copy(A);
  matrix_mutate_chol(L);
  return L;
}

kalman_matrix_t* matrix_create_vconcat(kalman_matrix_t* A, kalman_matrix_t* B) {
  const int32_t ra = A ? A->row_dim : 0, rb = B ? B->row_dim : 0;
  if (ra + rb == 0) return NULL;
  if (ra == 0) return matrix_create_copy(B);
  if (rb == 0) return matrix_create_copy(A);
  const int32_t cols = MAX(A->col_dim, B->col_dim);
  kalman_matrix_t* X = matrix_create_constant(ra + rb, cols, 0.0);
  for (int32_t j = 0; j < cols; j++) {
    for (int32_t i = 0; i < ra; i++)
      if (j < A->col_dim) EL(X, i, j) = EL(A, i, j);
    for (int32_t i = 0; i < rb; i++)
      if (j < B->col_dim) EL(X, ra + i, j) = EL(B, i, j);
  }
  return X;
}

kalman_matrix_t* matrix_create_transpose(kalman_matrix_t* A) {
  if (!A) return NULL;
  kalman_matrix_t* T = matrix_create(A->col_dim, A->row_dim);
  for (int32_t j = 0; j < A->col_dim; j++)
    for (int32_t i = 0; i < A->row_dim; i++) EL(T, j, i) = EL(A, i, j);
  return T;
}

kalman_matrix_t* matrix_create_multiply(kalman_matrix_t* A, kalman_matrix_t* B) {
  if (!A || !B || A->col_dim != B->row_dim) {
    uk_set_error("matrix_create_multiply: shapes do not conform");
    return NULL;
  }
  kalman_matrix_t* C = matrix_create_constant(A->row_dim, B->col_dim, 0.0);
  matrix_mutate_gemm(1.0, A, B, 0.0, C);
  return C;
}

static kalman_matrix_t* elementwise(kalman_matrix_t* A, kalman_matrix_t* B, double sb, const char* what) {
  if (!A || !B || A->row_dim != B->row_dim || A->col_dim != B->col_dim) {
    uk_set_error("%s: shapes differ", what);
    return NULL;
  }
  kalman_matrix_t* C = matrix_create_copy(A);
  for (int32_t j = 0; j < A->col_dim; j++)
    for (int32_t i = 0; i < A->row_dim; i++) EL(C, i, j) += sb * EL(B, i, j);
  return C;
}

kalman_matrix_t* matrix_create_subtract(kalman_matrix_t* A, kalman_matrix_t* B) {
  return elementwise(A, B, -1.0, "matrix_create_subtract");
}
kalman_matrix_t* matrix_create_add(kalman_matrix_t* A, kalman_matrix_t* B) {
  return elementwise(A, B, 1.0, "matrix_create_add");
}

kalman_matrix_t* matrix_create_mldivide(kalman_matrix_t* A, kalman_matrix_t* B) {
  if (!A || !B || A->row_dim != A->col_dim || B->row_dim != A->row_dim) {
    uk_set_error("matrix_create_mldivide: shapes do not conform");
    return NULL;
  }
  kalman_matrix_t* R = matrix_create_copy(A);
  kalman_matrix_t* tau = matrix_create_mutate_qr(R);
  kalman_matrix_t* X = matrix_create_copy(B);
  matrix_mutate_apply_qt(R, tau, X);
  matrix_mutate_triu(R);
  matrix_mutate_trisolve("U", R, X);
  matrix_free(R);
  matrix_free(tau);
  return X;
}

kalman_matrix_t* matrix_create_inverse(kalman_matrix_t* A) {
  if (!A) return NULL;
  kalman_matrix_t* I = matrix_create_identity(A->row_dim, A->col_dim);
  kalman_matrix_t* X = matrix_create_mldivide(A, I);
  matrix_free(I);
  return X;
}
