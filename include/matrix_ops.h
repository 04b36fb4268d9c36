/*
 * matrix_ops.h -- dense column-major FP64 matrices exported by libukhip.
 *
 * Drop-in for the matrix helpers that UltimateKalman clients call
 * (reference: c/matrix_ops.h:137-236). Names, argument meaning and ownership
 * are those of the reference; the implementation is this library's own
 * (host code only -- the smoothers never call these on the hot path, they
 * pack the matrices into device structure-of-arrays buffers instead).
 *
 * Layout: element (i,j) lives at elements[j*ld + i]; ld == rows for every
 * matrix this library creates (c/matrix_ops.h:150-151).
 */
#ifndef UKHIP_MATRIX_OPS_H
#define UKHIP_MATRIX_OPS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* c/matrix_ops.h:137-147 */
typedef struct matrix_st {
  int32_t row_dim;
  int32_t col_dim;
  int32_t ld;       /* leading dimension */
  double* elements; /* column-major, NULL when rows*cols == 0 */
} kalman_matrix_t;

#ifdef KALMAN_MATRIX_SHORT_TYPE
typedef kalman_matrix_t matrix_t;
#endif

/* c/matrix_ops.h:149-155 (functions, bounds-checked; the reference turns them
 * into macros under NDEBUG -- these are always functions). */
void   matrix_set(kalman_matrix_t* A, int32_t i, int32_t j, double v);
double matrix_get(kalman_matrix_t* A, int32_t i, int32_t j);

/* c/matrix_ops.h:160-170 */
kalman_matrix_t* matrix_create(int32_t rows, int32_t cols);              /* undefined contents */
kalman_matrix_t* matrix_create_constant(int32_t rows, int32_t cols, double c);
kalman_matrix_t* matrix_create_copy(kalman_matrix_t* A);                 /* NULL -> NULL */
kalman_matrix_t* matrix_create_sub(kalman_matrix_t* A, int32_t first_row, int32_t rows,
                                   int32_t first_col, int32_t cols);
void             matrix_mutate_copy_sub(kalman_matrix_t* C, int32_t first_row, int32_t first_col,
                                        kalman_matrix_t* A);
void             matrix_free(kalman_matrix_t* A);                       /* NULL is a no-op */

int32_t matrix_rows(kalman_matrix_t* A); /* 0 for NULL */
int32_t matrix_cols(kalman_matrix_t* A);
int32_t matrix_ld(kalman_matrix_t* A);

/* c/matrix_ops.h:180-185 */
kalman_matrix_t* matrix_create_identity(int32_t rows, int32_t cols);
kalman_matrix_t* matrix_create_from_rowwise(double* naked, int32_t rows, int32_t cols);

void matrix_print(kalman_matrix_t* A, char* format);

/* c/matrix_ops.h:198-236 */
void matrix_mutate_gemm(double ALPHA, kalman_matrix_t* A, kalman_matrix_t* B, double BETA, kalman_matrix_t* C);
void matrix_mutate_scale(kalman_matrix_t* A, double s);
void matrix_mutate_triu(kalman_matrix_t* A);
void matrix_mutate_chop(kalman_matrix_t* A, int32_t rows, int32_t cols);
void matrix_mutate_copy(kalman_matrix_t* C, kalman_matrix_t* A);
void matrix_mutate_trisolve(char* triangle, kalman_matrix_t* U, kalman_matrix_t* b);
void matrix_mutate_chol(kalman_matrix_t* L);

/* Householder QR in LAPACK's compact form: R in the upper triangle, the
 * essential parts of the reflectors below it; returns TAU (min(rows,cols) x 1,
 * the reference returns cols x 1 -- the extra entries are zero). */
kalman_matrix_t* matrix_create_mutate_qr(kalman_matrix_t* A);
/* C <- Q^T C with Q from matrix_create_mutate_qr. */
void             matrix_mutate_apply_qt(kalman_matrix_t* QR, kalman_matrix_t* TAU, kalman_matrix_t* C);

kalman_matrix_t* matrix_create_trisolve(char* triangle, kalman_matrix_t* U, kalman_matrix_t* b);
kalman_matrix_t* matrix_create_chol(kalman_matrix_t* A);
kalman_matrix_t* matrix_create_vconcat(kalman_matrix_t* A, kalman_matrix_t* B);
kalman_matrix_t* matrix_create_inverse(kalman_matrix_t* A);
kalman_matrix_t* matrix_create_transpose(kalman_matrix_t* A);
kalman_matrix_t* matrix_create_mldivide(kalman_matrix_t* A, kalman_matrix_t* B);
kalman_matrix_t* matrix_create_multiply(kalman_matrix_t* A, kalman_matrix_t* B);
kalman_matrix_t* matrix_create_subtract(kalman_matrix_t* A, kalman_matrix_t* B);
kalman_matrix_t* matrix_create_add(kalman_matrix_t* A, kalman_matrix_t* B);

#ifdef __cplusplus
}
#endif

#endif /* UKHIP_MATRIX_OPS_H */
