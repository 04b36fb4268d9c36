/*
 * uk_oracle.c -- CPU FP64 restatement of UltimateKalman's parallel smoothers.
 *
 * TEST INFRASTRUCTURE: the checker for libukhip and the CPU baseline of
 * bench.py. Nothing in the product links or loads this file.
 *
 * What it restates (reference file:line):
 *   whitening         c/kalman_base.c:61-214, matlab/CovarianceMatrix.m
 *   odd-even smoother c/kalman_oddeven_smoother.c:379-975 (phases, back
 *                     substitution, selected-inversion covariances),
 *                     :1358-1553 (recursion), :1620-1712 (weigh/finalize);
 *                     matlab/KalmanOddevenSmoother.m:21-358
 *   associative       c/kalman_associative_smoother.c:309-505 (filtering
 *                     elements), :697-753 (smoothing elements), :889-1016
 *                     (combines), :1199-1240 (driver);
 *                     matlab/KalmanAssociativeSmoother.m:20-262
 *   global LS         matlab/KalmanSparse.m:142-306
 *   dense kernels     LAPACK dgeqrf/dormqr/dtrtrs/dpotrf, BLAS dgemm as called
 *                     from c/matrix_ops.c:360-638 (own unblocked code)
 *
 * The operation sequence of the smoothers mirrors the reference call by call,
 * so the LAWN-41 FLOP replay (oracle_flops) counts the reference's own LAPACK
 * and BLAS call shapes.
 *
 * Reference defects and how this restatement behaves instead:
 *   D1  odd-even fix-last reads a freed matrix (c/kalman_oddeven_smoother.c:
 *       1469-1471): the split is max(columns of the new R, rows of the old
 *       G~), as in matlab/KalmanOddevenSmoother.m:171.
 *   D2  associative with unobserved steps (J, e NULL): e = 0, J = 0.
 *   D3  associative when step 0 does not determine u_0 (m_0 < n): the prior is
 *       carried in information form up to the first step p whose filtered
 *       information is full rank (sum of m_t over t <= p reaches n); smoothing
 *       elements for steps before p use the information form of
 *       E = P F^T (F P F^T + Q)^{-1}, g = x - E(Fx + c), L = P - E F P.
 *   D4  associative last-step covariance: filtered Z (matlab :144-149).
 *   D5  associative k == 1: x = R^{-1} Q^T o, P = (R^T R)^{-1}.
 *   D9  associative unobserved step 1 uses b = m0 + c (c :1203-1206, matlab
 *       :61): the prior mean of u_1 is F m0 + c; that is used here.
 *   D10 odd-even fix-last with fewer stacked rows than n (c :1463) keeps all
 *       rows, like phase 3 (:545-546), instead of an n x n block.
 *   The associative reference ignores H (assumes H = I); this restatement
 *   applies H^{-1} to F, c and K, which is the same when H = I.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "uk_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

typedef kalman_matrix_t M;

#define MIN(a, b) ((a) < (b) ? (a) : (b))
#define MAX(a, b) ((a) > (b) ? (a) : (b))

double kalman_nan = 0.0 / 0.0;

/* ------------------------------------------------------------------------ */
/* bookkeeping                                                               */
/* ------------------------------------------------------------------------ */

static int g_count = 0;
static double g_flops = 0.0;
static int g_threads = 1;
static int g_error = 0;

void oracle_count_flops(int enable) { g_count = enable; g_flops = 0.0; }
double oracle_flops(void) { return g_flops; }
void oracle_set_threads(int n) { g_threads = n > 0 ? n : 1; }

static void flops_add(double f) {
  if (!g_count) return;
#ifdef _OPENMP
#pragma omp atomic
#endif
  g_flops += f;
}

static void fail(const char* msg) {
  if (!g_error) fprintf(stderr, "uk_oracle: %s\n", msg);
  g_error = 1;
}

/* ------------------------------------------------------------------------ */
/* dense column-major matrices                                               */
/* ------------------------------------------------------------------------ */

static M* mnew(int r, int c) {
  M* A = (M*)malloc(sizeof(M));
  A->row_dim = r;
  A->col_dim = c;
  A->ld = r;
  A->elements = (r > 0 && c > 0) ? (double*)calloc((size_t)r * c, sizeof(double)) : NULL;
  return A;
}

kalman_matrix_t* oracle_matrix_create(int32_t rows, int32_t cols) { return mnew(rows, cols); }

static void mfree(M* A) {
  if (!A) return;
  free(A->elements);
  free(A);
}
void oracle_matrix_free(kalman_matrix_t* A) { mfree(A); }

static inline int R_(const M* A) { return A ? A->row_dim : 0; }
static inline int C_(const M* A) { return A ? A->col_dim : 0; }
#define EL(A, i, j) ((A)->elements[(size_t)(j) * (A)->ld + (i)])

static M* mcopy(const M* A) {
  if (!A) return NULL;
  M* B = mnew(A->row_dim, A->col_dim);
  for (int j = 0; j < A->col_dim; j++)
    for (int i = 0; i < A->row_dim; i++) EL(B, i, j) = EL(A, i, j);
  return B;
}

static void mset(M** dst, M* src) {
  mfree(*dst);
  *dst = src;
}

/* Vertical concatenation; NULL or 0-row blocks are empty (c/matrix_ops.c:283-320). */
static M* mvcat(const M* A, const M* B) {
  int ra = R_(A), rb = R_(B);
  if (ra + rb == 0) return NULL;
  if (ra == 0) return mcopy(B);
  if (rb == 0) return mcopy(A);
  int cols = MAX(C_(A), C_(B));
  M* X = mnew(ra + rb, cols);
  for (int j = 0; j < cols; j++) {
    for (int i = 0; i < ra; i++) EL(X, i, j) = j < A->col_dim ? EL(A, i, j) : 0.0;
    for (int i = 0; i < rb; i++) EL(X, ra + i, j) = j < B->col_dim ? EL(B, i, j) : 0.0;
  }
  return X;
}

static M* msub(const M* A, int r0, int nr, int c0, int nc) {
  M* X = mnew(nr, nc);
  for (int j = 0; j < nc; j++)
    for (int i = 0; i < nr; i++) EL(X, i, j) = EL(A, r0 + i, c0 + j);
  return X;
}

static void mtriu(M* A) {
  for (int j = 0; j < A->col_dim; j++)
    for (int i = j + 1; i < A->row_dim; i++) EL(A, i, j) = 0.0;
}

static M* mtrans(const M* A) {
  M* T = mnew(C_(A), R_(A));
  for (int j = 0; j < C_(A); j++)
    for (int i = 0; i < R_(A); i++) EL(T, j, i) = EL(A, i, j);
  return T;
}

static M* madd(const M* A, const M* B, double sb) {
  M* X = mcopy(A);
  for (int j = 0; j < C_(A); j++)
    for (int i = 0; i < R_(A); i++) EL(X, i, j) += sb * EL(B, i, j);
  return X;
}

static void mscale(M* A, double s) {
  for (int j = 0; j < C_(A); j++)
    for (int i = 0; i < R_(A); i++) EL(A, i, j) *= s;
}

static void madd_identity(M* A) {
  for (int i = 0; i < MIN(R_(A), C_(A)); i++) EL(A, i, i) += 1.0;
}

static M* meye(int n) {
  M* I = mnew(n, n);
  for (int i = 0; i < n; i++) EL(I, i, i) = 1.0;
  return I;
}

/* C = A*B (dgemm "N","N"; LAWN-41 count 2MNK). */
static M* mmul(const M* A, const M* B) {
  int m = R_(A), kk = C_(A), n = C_(B);
  if (kk != R_(B)) { fail("mmul: inner dimensions differ"); return mnew(m, n); }
  M* X = mnew(m, n);
  for (int j = 0; j < n; j++)
    for (int p = 0; p < kk; p++) {
      double b = EL(B, p, j);
      if (b == 0.0) continue;
      for (int i = 0; i < m; i++) EL(X, i, j) += EL(A, i, p) * b;
    }
  flops_add(2.0 * m * n * kk);
  return X;
}

/* Householder QR in place (dgeqrf semantics, unblocked). Returns tau with
 * min(rows, cols) entries. H_k = I - tau_k v_k v_k^T, v_k(k) = 1. */
static M* mqr(M* A) {
  int m = R_(A), n = C_(A), kmax = MIN(m, n);
  M* tau = mnew(MAX(kmax, 1), 1);
  for (int k = 0; k < kmax; k++) {
    double alpha = EL(A, k, k), ss = 0.0;
    for (int i = k + 1; i < m; i++) ss += EL(A, i, k) * EL(A, i, k);
    double t = 0.0;
    if (ss != 0.0) {
      double beta = -copysign(sqrt(alpha * alpha + ss), alpha);
      t = (beta - alpha) / beta;
      double s = 1.0 / (alpha - beta);
      for (int i = k + 1; i < m; i++) EL(A, i, k) *= s;
      EL(A, k, k) = beta;
    }
    tau->elements[k] = t;
    if (t == 0.0) continue;
    for (int j = k + 1; j < n; j++) {
      double w = EL(A, k, j);
      for (int i = k + 1; i < m; i++) w += EL(A, i, k) * EL(A, i, j);
      w *= t;
      EL(A, k, j) -= w;
      for (int i = k + 1; i < m; i++) EL(A, i, j) -= w * EL(A, i, k);
    }
  }
  if (m >= n) flops_add(2.0 * m * n * n - 2.0 / 3.0 * n * n * n);
  else flops_add(2.0 * n * m * m - 2.0 / 3.0 * m * m * m);
  return tau;
}

/* C <- Q^T C (dormqr "L","T"; count P(4MK - 2K^2)). */
static void mapply_qt(const M* QR, const M* tau, M* C) {
  if (!C || R_(C) == 0) return;
  int m = R_(C), p = C_(C), kk = MIN(R_(QR), C_(QR));
  for (int k = 0; k < kk; k++) {
    double t = tau->elements[k];
    if (t == 0.0) continue;
    for (int j = 0; j < p; j++) {
      double w = EL(C, k, j);
      for (int i = k + 1; i < m; i++) w += EL(QR, i, k) * EL(C, i, j);
      w *= t;
      EL(C, k, j) -= w;
      for (int i = k + 1; i < m; i++) EL(C, i, j) -= w * EL(QR, i, k);
    }
  }
  flops_add((double)p * (4.0 * m * kk - 2.0 * kk * kk));
}

/* In-place solve with a square triangle (dtrtrs "N","N"; count P N^2). */
static void mtrisolve(char tri, const M* T, M* B) {
  int n = R_(T), p = C_(B);
  if (C_(T) != n || R_(B) != n) { fail("trisolve: shape"); return; }
  for (int j = 0; j < p; j++) {
    if (tri == 'U') {
      for (int i = n - 1; i >= 0; i--) {
        double s = EL(B, i, j);
        for (int q = i + 1; q < n; q++) s -= EL(T, i, q) * EL(B, q, j);
        EL(B, i, j) = s / EL(T, i, i);
      }
    } else {
      for (int i = 0; i < n; i++) {
        double s = EL(B, i, j);
        for (int q = 0; q < i; q++) s -= EL(T, i, q) * EL(B, q, j);
        EL(B, i, j) = s / EL(T, i, i);
      }
    }
  }
  for (int i = 0; i < n; i++)
    if (EL(T, i, i) == 0.0) fail("trisolve: singular triangle");
  flops_add((double)p * n * n);
}

/* Lower Cholesky in place (dpotrf "L"; count n^3/3). */
static void mchol(M* A) {
  int n = R_(A);
  for (int j = 0; j < n; j++) {
    double d = EL(A, j, j);
    for (int q = 0; q < j; q++) d -= EL(A, j, q) * EL(A, j, q);
    if (!(d > 0.0)) { fail("chol: not positive definite"); d = 1.0; }
    d = sqrt(d);
    EL(A, j, j) = d;
    for (int i = j + 1; i < n; i++) {
      double s = EL(A, i, j);
      for (int q = 0; q < j; q++) s -= EL(A, i, q) * EL(A, j, q);
      EL(A, i, j) = s / d;
    }
    for (int i = 0; i < j; i++) EL(A, i, j) = 0.0;
  }
  flops_add(n * (double)n * n / 3.0);
}

/* A \ B through QR, as c/matrix_ops.c:610-630. */
static M* mmldivide(const M* A, const M* B) {
  M* R = mcopy(A);
  M* tau = mqr(R);
  M* X = mcopy(B);
  mapply_qt(R, tau, X);
  mtriu(R);
  mtrisolve('U', R, X);
  mfree(R);
  mfree(tau);
  return X;
}

static M* minverse(const M* A) {
  M* I = meye(R_(A));
  M* X = mmldivide(A, I);
  mfree(I);
  return X;
}

/* Applies Q^T to [*upper; *lower] and splits the result at `top` rows
 * (c/kalman_oddeven_smoother.c:326-347). */
static void apply_split(const M* QR, const M* tau, int top, M** upper, M** lower) {
  M* cat = mvcat(*upper, *lower);
  if (!cat) return;
  mapply_qt(QR, tau, cat);
  M* nu = msub(cat, 0, top, 0, C_(cat));
  M* nl = msub(cat, top, R_(cat) - top, 0, C_(cat));
  mfree(cat);
  mset(upper, nu);
  mset(lower, nl);
}

/* ------------------------------------------------------------------------ */
/* covariance matrices (c/kalman_base.c:61-214)                              */
/* ------------------------------------------------------------------------ */

static M* weigh(const M* cov, char type, const M* A) {
  M* X = NULL;
  M* L;
  switch (type) {
    case 'W':
      X = mmul(cov, A);
      break;
    case 'w':
      X = mcopy(A);
      for (int j = 0; j < C_(A); j++)
        for (int i = 0; i < R_(A); i++) EL(X, i, j) *= cov->elements[i];
      break;
    case 'C':
      L = mcopy(cov);
      mchol(L);
      X = mcopy(A);
      mtrisolve('L', L, X);
      mfree(L);
      break;
    case 'U':
    case 'F':
      X = mcopy(A);
      mtrisolve('L', cov, X);
      break;
    default:
      fail("weigh: unknown covariance type");
      X = mcopy(A);
  }
  return X;
}

static M* explicit_cov(const M* cov, char type) {
  if (type == 'w') {
    int n = R_(cov);
    M* X = mnew(n, n);
    for (int i = 0; i < n; i++) EL(X, i, i) = 1.0 / (cov->elements[i] * cov->elements[i]);
    return X;
  }
  if (type == 'W') {
    M* WT = mtrans(cov);
    M* WTW = mmul(WT, cov);
    M* X = minverse(WTW);
    mfree(WT);
    mfree(WTW);
    return X;
  }
  if (type == 'U' || type == 'F') {
    M* LT = mtrans(cov);
    M* X = mmul(cov, LT);
    mfree(LT);
    return X;
  }
  if (type == 'C') return mcopy(cov);
  fail("explicit: unknown covariance type");
  return NULL;
}

/* ------------------------------------------------------------------------ */
/* odd-even smoother                                                          */
/* ------------------------------------------------------------------------ */

typedef struct {
  int32_t id; /* logical step number, for the off-diagonal bookkeeping */
  int32_t n;
  M *D, *B, *c;  /* reference H, F, c after weighing: V H, -V F, V c */
  M *C, *o;      /* reference G, o after weighing: W G, W o (NULL if unobserved) */
  M *X, *Y, *Z, *R;
  M *Xt, *Bt, *Dt, *Rt, *Gt; /* X~, F~ (= B~), H~ (= D~), R~, G~ */
  M *S, *Soff, *S1off;
  int32_t ioff, i1off;
  M *state, *cov;
} ostep;

static void ostep_release(ostep* s) {
  M** all[] = {&s->D, &s->B, &s->c, &s->C, &s->o, &s->X, &s->Y, &s->Z, &s->R, &s->Xt, &s->Bt,
               &s->Dt, &s->Rt, &s->Gt, &s->S, &s->Soff, &s->S1off};
  for (size_t i = 0; i < sizeof(all) / sizeof(all[0]); i++) mset(all[i], NULL);
}

#define PFOR _Pragma("omp parallel for schedule(dynamic, 16) num_threads(nthr)")

static int nthreads_now(void) { return g_count ? 1 : g_threads; }

/* phase 1: c/kalman_oddeven_smoother.c:379-457 */
static void oe_phase1(ostep** s, int L, int j) {
  ostep* si = s[j];
  if (j == L - 1) {
    if (R_(si->C) == 0) { /* unobserved lone last step: empty R~ */
      mset(&si->Rt, mnew(0, si->n));
      return;
    }
    M* QR = mcopy(si->C);
    M* tau = mqr(QR);
    mapply_qt(QR, tau, si->o);
    mtriu(QR);
    mset(&si->Rt, QR);
    mfree(tau);
    return;
  }
  ostep* sp = s[j + 1];
  M* QR = mvcat(si->C, sp->B);
  int nn = C_(sp->B);
  M* tau = mqr(QR);
  int ll = MIN(nn, R_(QR));
  M* Rt = msub(QR, 0, ll, 0, nn);
  mtriu(Rt);
  mset(&si->Rt, Rt);
  mset(&si->X, mnew(R_(QR) - R_(sp->D), C_(sp->D)));
  mset(&sp->Dt, mcopy(sp->D));
  apply_split(QR, tau, ll, &si->o, &sp->c);
  apply_split(QR, tau, ll, &si->X, &sp->Dt);
  mfree(QR);
  mfree(tau);
}

/* phase 2: c/kalman_oddeven_smoother.c:460-523 */
static void oe_phase2(ostep** s, int L, int j) {
  ostep* si = s[j];
  if (j == 0) {
    mset(&si->R, mcopy(si->Rt));
    return;
  }
  int nn = C_(si->D);
  M* QR = mvcat(si->D, si->Rt);
  M* tau = mqr(QR);
  M* R = msub(QR, 0, nn, 0, nn);
  mtriu(R);
  mset(&si->R, R);
  mset(&si->Z, mnew(R_(si->Rt), C_(si->B)));
  mset(&si->Bt, mcopy(si->B));
  int mm = MIN(nn, R_(QR));
  apply_split(QR, tau, mm, &si->Bt, &si->Z);
  apply_split(QR, tau, mm, &si->c, &si->o);
  if (j + 1 != L) {
    mset(&si->Y, mnew(R_(QR) - R_(si->X), C_(si->X)));
    mset(&si->Xt, mcopy(si->X));
    apply_split(QR, tau, mm, &si->Y, &si->Xt);
  }
  mfree(QR);
  mfree(tau);
}

/* phase 3: c/kalman_oddeven_smoother.c:526-560 */
static void oe_phase3(ostep** s, int j) {
  ostep* sp = s[j + 1];
  M* QR = mvcat(sp->Dt, sp->C);
  M* tau = mqr(QR);
  int ll = MIN(R_(QR), C_(QR));
  M* R = msub(QR, 0, ll, 0, C_(QR));
  mtriu(R);
  mset(&sp->Gt, R);
  apply_split(QR, tau, ll, &sp->c, &sp->o);
  mfree(QR);
  mfree(tau);
}

/* back substitution and selected-inversion covariance:
 * c/kalman_oddeven_smoother.c:614-975, matlab/KalmanOddevenSmoother.m:230-319 */
static void oe_backsub(ostep** s, int L, int j, int cov) {
  ostep* si = s[j];
  M* R = si->R;
  if (j == 0) {
    ostep* sp = s[j + 1];
    M* Xx = mmul(si->X, sp->state);
    M* b = madd(si->o, Xx, -1.0);
    mtriu(R);
    mtrisolve('U', R, b);
    mset(&si->state, b);
    mfree(Xx);
    if (!cov) return;
    M* XS = mmul(si->X, sp->S);
    M* Xt = mtrans(si->X);
    M* XSXt = mmul(XS, Xt);
    madd_identity(XSXt);
    mtrisolve('U', R, XSXt);
    M* T = mtrans(XSXt);
    mtrisolve('U', R, T);
    mset(&si->S, mtrans(T));
    mfree(T);
    mfree(XSXt);
    mfree(Xt);
    mfree(XS);
    M* Soff = mmul(si->X, sp->S);
    mtrisolve('U', R, Soff);
    mscale(Soff, -1.0);
    mset(&si->Soff, Soff);
    si->ioff = sp->id;
    mset(&si->cov, mcopy(si->S));
    return;
  }
  if (j != L - 1) {
    ostep *sp = s[j + 1], *sm = s[j - 1];
    M* m1 = mmul(si->Bt, sm->state);
    M* m2 = mmul(si->Y, sp->state);
    M* b1 = madd(si->c, m1, -1.0);
    M* b = madd(b1, m2, -1.0);
    mtriu(R);
    mtrisolve('U', R, b);
    mset(&si->state, b);
    mfree(m1);
    mfree(m2);
    mfree(b1);
    if (!cov) return;
    M *Smp = NULL, *Smpt = NULL;
    /* c/kalman_oddeven_smoother.c:770-801: locate Cov(u_{j-1}, u_{j+1}) */
    if (sp->ioff == sm->id) { Smpt = mcopy(sp->Soff); Smp = mtrans(Smpt); }
    if (sp->i1off == sm->id) { mfree(Smp); mfree(Smpt); Smpt = mcopy(sp->S1off); Smp = mtrans(Smpt); }
    if (sm->ioff == sp->id) { mfree(Smp); mfree(Smpt); Smp = mcopy(sm->Soff); Smpt = mtrans(Smp); }
    if (sm->i1off == sp->id) { mfree(Smp); mfree(Smpt); Smp = mcopy(sm->S1off); Smpt = mtrans(Smp); }
    if (!Smp) { fail("oddeven: missing cross covariance"); return; }
    M* BS = mmul(si->Bt, sm->S);
    M* Bt_t = mtrans(si->Bt);
    M* BSB = mmul(BS, Bt_t);
    M* YS = mmul(si->Y, sp->S);
    M* Yt = mtrans(si->Y);
    M* YSY = mmul(YS, Yt);
    M* BSo = mmul(si->Bt, Smp);
    M* BSoY = mmul(BSo, Yt);
    M* BSoYt = mtrans(BSoY);
    M* a1 = madd(BSB, YSY, 1.0);
    M* a2 = madd(a1, BSoY, 1.0);
    M* a3 = madd(a2, BSoYt, 1.0);
    madd_identity(a3);
    mtrisolve('U', R, a3);
    M* T = mtrans(a3);
    mtrisolve('U', R, T);
    mset(&si->S, mtrans(T));
    mfree(T);
    mfree(BS); mfree(Bt_t); mfree(BSB); mfree(YS); mfree(YSY);
    mfree(BSo); mfree(BSoY); mfree(BSoYt); mfree(a1); mfree(a2); mfree(a3);
    M* t1 = mmul(si->Bt, sm->S);
    M* t2 = mmul(si->Y, Smpt);
    M* Soff = madd(t1, t2, 1.0);
    mtrisolve('U', R, Soff);
    mscale(Soff, -1.0);
    mset(&si->Soff, Soff);
    M* t3 = mmul(si->Y, sp->S);
    M* t4 = mmul(si->Bt, Smp);
    M* S1off = madd(t3, t4, 1.0);
    mtrisolve('U', R, S1off);
    mscale(S1off, -1.0);
    mset(&si->S1off, S1off);
    mfree(t1); mfree(t2); mfree(t3); mfree(t4); mfree(Yt); mfree(Smp); mfree(Smpt);
    si->ioff = sm->id;
    si->i1off = sp->id;
    mset(&si->cov, mcopy(si->S));
    return;
  }
  /* last even step */
  ostep* sm = s[j - 1];
  M* m1 = mmul(si->Bt, sm->state);
  M* b = madd(si->c, m1, -1.0);
  mtriu(R);
  mtrisolve('U', R, b);
  mset(&si->state, b);
  mfree(m1);
  if (!cov) return;
  M* BS = mmul(si->Bt, sm->S);
  M* Bt_t = mtrans(si->Bt);
  M* BSB = mmul(BS, Bt_t);
  madd_identity(BSB);
  mtrisolve('U', R, BSB);
  M* T = mtrans(BSB);
  mtrisolve('U', R, T);
  mset(&si->S, mtrans(T));
  mfree(T); mfree(BSB); mfree(Bt_t); mfree(BS);
  M* Soff = mmul(si->Bt, sm->S);
  mtrisolve('U', R, Soff);
  mscale(Soff, -1.0);
  mset(&si->Soff, Soff);
  si->ioff = sm->id;
  mset(&si->cov, mcopy(si->S));
}

/* c/kalman_oddeven_smoother.c:1358-1553 */
static void oe_recursive(ostep** s, int L, int cov) {
  int nthr = nthreads_now();
  (void)nthr;
  if (L == 1) { /* :1360-1423 */
    ostep* s0 = s[0];
    if (R_(s0->C) < s0->n) { fail("oddeven: single step does not determine the state"); return; }
    M* QR = mcopy(s0->C);
    M* o = mcopy(s0->o);
    M* tau = mqr(QR);
    mapply_qt(QR, tau, o);
    mtriu(QR);
    M* Rsq = msub(QR, 0, s0->n, 0, s0->n); /* rows beyond n are zero after QR */
    M* x = msub(o, 0, s0->n, 0, 1);
    mtrisolve('U', Rsq, x);
    mset(&s0->state, x);
    M* I = meye(s0->n);
    mtrisolve('U', Rsq, I);
    M* T = mtrans(I);
    mtrisolve('U', Rsq, T);
    if (cov) {
      mset(&s0->S, mtrans(T));
      mset(&s0->cov, mcopy(s0->S));
    }
    mfree(T); mfree(I); mfree(Rsq); mfree(QR); mfree(o); mfree(tau);
    return;
  }
  int half_up = (L + 1) / 2, half = L / 2;
  PFOR for (int u = 0; u < half_up; u++) oe_phase1(s, L, 2 * u);
  PFOR for (int u = 0; u < half_up; u++) oe_phase2(s, L, 2 * u);
  PFOR for (int u = 0; u < half; u++) oe_phase3(s, 2 * u);

  if (L % 2 == 1) { /* fix the last index, :1452-1476 (D1 fixed) */
    ostep *sl = s[L - 1], *sm = s[L - 2];
    M* QR = mvcat(sm->Gt, sl->Z);
    M* tau = mqr(QR);
    int nc = C_(QR);
    /* matlab/KalmanOddevenSmoother.m:171: max(cols of the new R, rows of the old G~) */
    int split = MAX(nc, R_(sm->Gt));
    /* D10: with fewer than n stacked rows the reference takes an n x n block
     * of a shorter matrix (matrix_create_sub asserts); keep all rows then,
     * as phase 3 does. */
    int keep = MIN(nc, R_(QR));
    if (split > keep) split = keep;
    M* R = msub(QR, 0, keep, 0, nc);
    mtriu(R);
    mset(&sm->Gt, R);
    apply_split(QR, tau, split, &sm->c, &sl->o);
    mfree(QR);
    mfree(tau);
  }

  /* variable renaming, :563-597 */
  PFOR for (int u = 0; u < half; u++) {
    int j = 2 * u;
    ostep* sp = s[j + 1];
    M* nG = mcopy(sp->Gt);
    M* no = mcopy(sp->c);
    if (j != 0) {
      ostep* si = s[j];
      mset(&sp->B, mcopy(si->Z));
      mset(&sp->D, mcopy(si->Xt));
      mset(&sp->c, mcopy(si->o));
    }
    mset(&sp->C, nG);
    mset(&sp->o, no);
  }

  ostep** sub = (ostep**)malloc(sizeof(ostep*) * (size_t)half);
  for (int u = 0; u < half; u++) sub[u] = s[2 * u + 1];
  oe_recursive(sub, half, cov);
  free(sub);
  if (g_error) return;

  PFOR for (int u = 0; u < half_up; u++) oe_backsub(s, L, 2 * u, cov);
}

static kalman_matrix_t* out_copy(const M* A) { return mcopy(A); }

void oracle_smooth_oddeven(kalman_options_t options, kalman_step_equations_t** eq, kalman_step_index_t k) {
  g_error = 0;
  if (k <= 0) return;
  int cov = (options & KALMAN_NO_COVARIANCE) == 0;
  ostep* arr = (ostep*)calloc((size_t)k, sizeof(ostep));
  ostep** s = (ostep**)malloc(sizeof(ostep*) * (size_t)k);
  int nthr = nthreads_now();
  (void)nthr;
  /* steps_weigh, c/kalman_oddeven_smoother.c:1620-1659 */
  PFOR for (int i = 0; i < k; i++) {
    s[i] = &arr[i];
    ostep* st = s[i];
    st->id = eq[i]->step;
    st->n = eq[i]->dimension;
    st->ioff = st->i1off = -1;
    if (eq[i]->step > 0) {
      st->D = weigh(eq[i]->K, eq[i]->K_type, eq[i]->H);
      st->B = weigh(eq[i]->K, eq[i]->K_type, eq[i]->F);
      st->c = weigh(eq[i]->K, eq[i]->K_type, eq[i]->c);
      mscale(st->B, -1.0);
    }
    if (eq[i]->o != NULL) {
      st->C = weigh(eq[i]->C, eq[i]->C_type, eq[i]->G);
      st->o = weigh(eq[i]->C, eq[i]->C_type, eq[i]->o);
    }
  }
  oe_recursive(s, (int)k, cov);
  for (int i = 0; i < k; i++) { /* steps_finalize, :1661-1695 */
    if (!g_error) {
      eq[i]->state = out_copy(s[i]->state);
      eq[i]->covariance = cov ? out_copy(s[i]->cov) : NULL;
      eq[i]->covariance_type = 'C';
    }
    mfree(s[i]->state);
    mfree(s[i]->cov);
    ostep_release(s[i]);
  }
  free(s);
  free(arr);
}

/* ------------------------------------------------------------------------ */
/* associative smoother                                                       */
/* ------------------------------------------------------------------------ */

typedef struct {
  int n;
  int has; /* element present (filtering: 1..k-1) */
  M *A, *b, *Z, *e, *J;  /* filtering 5-tuple; Z is the article's C */
  M *E, *g, *L;          /* smoothing 3-tuple */
} elem;

static void elem_release(elem* x) {
  mset(&x->A, NULL); mset(&x->b, NULL); mset(&x->Z, NULL); mset(&x->e, NULL); mset(&x->J, NULL);
  mset(&x->E, NULL); mset(&x->g, NULL); mset(&x->L, NULL);
}

/* c/kalman_associative_smoother.c:889-980 (matlab :240-254) */
static void filt_combine(const elem* si, const elem* sj, elem* out) {
  int n = R_(si->b);
  M* I = meye(n);
  M* ZJ = mmul(si->Z, sj->J);
  M* IZJ = madd(I, ZJ, 1.0);
  M* AjT = mtrans(sj->A);
  M* IZJt = mtrans(IZJ);
  M* XT = mmldivide(IZJt, AjT);
  M* X = mtrans(XT);
  M* JZ = mmul(sj->J, si->Z);
  M* IJZ = madd(I, JZ, 1.0);
  M* IJZt = mtrans(IJZ);
  M* YT = mmldivide(IJZt, si->A);
  M* Y = mtrans(YT);

  out->n = n;
  out->has = 1;
  mset(&out->A, mmul(X, si->A));
  M* Ze = mmul(si->Z, sj->e);
  M* Zeb = madd(Ze, si->b, 1.0);
  M* XZeb = mmul(X, Zeb);
  mset(&out->b, madd(XZeb, sj->b, 1.0));
  M* XZ = mmul(X, si->Z);
  M* XZA = mmul(XZ, AjT);
  mset(&out->Z, madd(XZA, sj->Z, 1.0));
  M* Jb = mmul(sj->J, si->b);
  M* eJb = madd(sj->e, Jb, -1.0);
  M* YeJb = mmul(Y, eJb);
  mset(&out->e, madd(YeJb, si->e, 1.0));
  M* JA = mmul(sj->J, si->A);
  M* YJA = mmul(Y, JA);
  mset(&out->J, madd(YJA, si->J, 1.0));

  mfree(I); mfree(ZJ); mfree(IZJ); mfree(AjT); mfree(IZJt); mfree(XT); mfree(X);
  mfree(JZ); mfree(IJZ); mfree(IJZt); mfree(YT); mfree(Y);
  mfree(Ze); mfree(Zeb); mfree(XZeb); mfree(XZ); mfree(XZA); mfree(Jb); mfree(eJb); mfree(YeJb);
  mfree(JA); mfree(YJA);
}

/* c/kalman_associative_smoother.c:982-1016 (matlab :256-262); si is the
 * running suffix, sj the new (earlier) element. */
static void smooth_combine(const elem* si, const elem* sj, elem* out) {
  out->n = si->n;
  out->has = 1;
  mset(&out->E, mmul(sj->E, si->E));
  M* Eg = mmul(sj->E, si->g);
  mset(&out->g, madd(Eg, sj->g, 1.0));
  M* ET = mtrans(sj->E);
  M* EL_ = mmul(sj->E, si->L);
  M* ELE = mmul(EL_, ET);
  mset(&out->L, madd(ELE, sj->L, 1.0));
  mfree(Eg); mfree(ET); mfree(EL_); mfree(ELE);
}

/* Model data of a step in the form the associative smoother uses:
 * u_i = F u_{i-1} + c + w, Cov w = Q;  o = G u_i + d, Cov d = Cobs. */
typedef struct {
  M *F, *c, *Q;
  M *G, *o, *Cobs;
  M *WG, *Wo; /* whitened observation */
} amodel;

static void amodel_build(kalman_step_equations_t* e, amodel* m) {
  memset(m, 0, sizeof(*m));
  if (e->step > 0) {
    M* Q = explicit_cov(e->K, e->K_type);
    int hid = 1; /* H == I ? */
    for (int j = 0; j < C_(e->H) && hid; j++)
      for (int i = 0; i < R_(e->H); i++)
        if (EL(e->H, i, j) != (i == j ? 1.0 : 0.0)) { hid = 0; break; }
    if (hid) {
      m->F = mcopy(e->F);
      m->c = mcopy(e->c);
      m->Q = Q;
    } else { /* F <- H^{-1} F, c <- H^{-1} c, Q <- H^{-1} Q H^{-T} */
      m->F = mmldivide(e->H, e->F);
      m->c = mmldivide(e->H, e->c);
      M* HQ = mmldivide(e->H, Q);
      M* HQt = mtrans(HQ);
      M* X = mmldivide(e->H, HQt);
      m->Q = mtrans(X);
      mfree(HQ); mfree(HQt); mfree(X); mfree(Q);
    }
  }
  if (e->o != NULL) {
    m->G = mcopy(e->G);
    m->o = mcopy(e->o);
    m->Cobs = explicit_cov(e->C, e->C_type);
    m->WG = weigh(e->C, e->C_type, e->G);
    m->Wo = weigh(e->C, e->C_type, e->o);
  }
}

static void amodel_release(amodel* m) {
  mset(&m->F, NULL); mset(&m->c, NULL); mset(&m->Q, NULL); mset(&m->G, NULL);
  mset(&m->o, NULL); mset(&m->Cobs, NULL); mset(&m->WG, NULL); mset(&m->Wo, NULL);
}

/* Filtering element of step i given the prior (m0, P0) of step i-1 when
 * `first` (c/kalman_associative_smoother.c:370-505, matlab :49-91). */
static void build_filtering(const amodel* md, int n, int first, const M* m0, const M* P0, elem* x) {
  x->n = n;
  x->has = 1;
  M* K = mcopy(md->Q);
  if (first) {
    M* FP = mmul(md->F, P0);
    M* Ft = mtrans(md->F);
    M* FPF = mmul(FP, Ft);
    mset(&K, madd(K, FPF, 1.0));
    mfree(FP); mfree(Ft); mfree(FPF);
  }
  if (md->o == NULL) { /* D2: e = 0, J = 0 */
    x->Z = K;
    if (first) {
      x->A = mnew(n, n);
      M* Fm = mmul(md->F, m0); /* D9 */
      x->b = madd(Fm, md->c, 1.0);
      mfree(Fm);
    } else {
      x->A = mcopy(md->F);
      x->b = mcopy(md->c);
    }
    x->e = mnew(n, 1);
    x->J = mnew(n, n);
    return;
  }
  M* Gt = mtrans(md->G);
  M* KGt = mmul(K, Gt);
  M* GKGt = mmul(md->G, KGt);
  M* S = madd(GKGt, md->Cobs, 1.0);
  M* St = mtrans(S);
  M* GSt = mmldivide(St, md->G); /* (G^T S^{-1})^T */
  M* GS = mtrans(GSt);            /* G^T S^{-1} */
  M* Kg = mmul(K, GS);
  if (first) {
    x->A = mnew(n, n);
    M* Fm = mmul(md->F, m0);
    M* m1 = madd(Fm, md->c, 1.0);
    M* Gm = mmul(md->G, m1);
    M* r = madd(md->o, Gm, -1.0);
    M* Kr = mmul(Kg, r);
    x->b = madd(m1, Kr, 1.0);
    M* KS = mmul(Kg, S);
    M* Kt = mtrans(Kg);
    M* KSK = mmul(KS, Kt);
    x->Z = madd(K, KSK, -1.0);
    mfree(Fm); mfree(m1); mfree(Gm); mfree(r); mfree(Kr); mfree(KS); mfree(Kt); mfree(KSK);
  } else {
    M* GF = mmul(md->G, md->F);
    M* KGF = mmul(Kg, GF);
    x->A = madd(md->F, KGF, -1.0);
    M* Gc = mmul(md->G, md->c);
    M* r = madd(md->o, Gc, -1.0);
    M* Kr = mmul(Kg, r);
    x->b = madd(md->c, Kr, 1.0);
    M* KG = mmul(Kg, md->G);
    M* KGK = mmul(KG, K);
    x->Z = madd(K, KGK, -1.0);
    mfree(GF); mfree(KGF); mfree(Gc); mfree(r); mfree(Kr); mfree(KG); mfree(KGK);
  }
  M* Gc = mmul(md->G, md->c);
  M* r = madd(md->o, Gc, -1.0);
  M* Ft = mtrans(md->F);
  M* FtG = mmul(Ft, GS);
  x->e = mmul(FtG, r);
  M* GF = mmul(md->G, md->F);
  x->J = mmul(FtG, GF);
  mfree(Gc); mfree(r); mfree(Ft); mfree(FtG); mfree(GF);
  mfree(Gt); mfree(KGt); mfree(GKGt); mfree(S); mfree(St); mfree(GSt); mfree(GS); mfree(Kg); mfree(K);
}

/* Smoothing element from a filtered (x, P) and the next step's dynamics
 * (c/kalman_associative_smoother.c:697-753, matlab :94-114). */
static void build_smoothing(const M* x, const M* P, const amodel* next, elem* el) {
  if (!next) {
    int n = R_(x);
    el->E = mnew(n, n);
    el->g = mcopy(x);
    el->L = mcopy(P);
    return;
  }
  M* Ft = mtrans(next->F);
  M* PFt = mmul(P, Ft);
  M* FPFt = mmul(next->F, PFt);
  M* FPFtQ = madd(FPFt, next->Q, 1.0);
  M* PFtT = mtrans(PFt);
  M* FPFtQT = mtrans(FPFtQ);
  M* ET = mmldivide(FPFtQT, PFtT);
  el->E = mtrans(ET);
  M* Fx = mmul(next->F, x);
  M* Fxc = madd(Fx, next->c, 1.0);
  M* EFxc = mmul(el->E, Fxc);
  el->g = madd(x, EFxc, -1.0);
  M* EF = mmul(el->E, next->F);
  M* EFP = mmul(EF, P);
  el->L = madd(P, EFP, -1.0);
  mfree(Ft); mfree(PFt); mfree(FPFt); mfree(FPFtQ); mfree(PFtT); mfree(FPFtQT); mfree(ET);
  mfree(Fx); mfree(Fxc); mfree(EFxc); mfree(EF); mfree(EFP);
}

/* Smoothing element of a step whose filtered distribution is given in
 * information form (Lam, eta) (D3): E = L F^T Q^{-1}, g = L (eta - F^T Q^{-1} c),
 * L = (Lam + F^T Q^{-1} F)^{-1}. */
static void build_smoothing_info(const M* Lam, const M* eta, const amodel* next, elem* el) {
  M* Ft = mtrans(next->F);
  M* QiF = mmldivide(next->Q, next->F);
  M* Qic = mmldivide(next->Q, next->c);
  M* FtQiF = mmul(Ft, QiF);
  M* Mx = madd(Lam, FtQiF, 1.0);
  el->L = minverse(Mx);
  M* QiFt = mtrans(QiF); /* F^T Q^{-1} (Q symmetric) */
  el->E = mmul(el->L, QiFt);
  M* FtQic = mmul(Ft, Qic);
  M* r = madd(eta, FtQic, -1.0);
  el->g = mmul(el->L, r);
  mfree(Ft); mfree(QiF); mfree(Qic); mfree(FtQiF); mfree(Mx); mfree(QiFt); mfree(FtQic); mfree(r);
}

void oracle_smooth_associative(kalman_options_t options, kalman_step_equations_t** eq, kalman_step_index_t k) {
  (void)options; /* NO_COVARIANCE is not honoured by the reference associative path (D8) */
  g_error = 0;
  if (k <= 0) return;
  int n = eq[0]->dimension;
  for (int i = 1; i < k; i++)
    if (eq[i]->dimension != n) { fail("associative: state dimension must be constant"); return; }
  int nthr = nthreads_now();
  (void)nthr;

  amodel* md = (amodel*)calloc((size_t)k, sizeof(amodel));
  PFOR for (int i = 0; i < k; i++) amodel_build(eq[i], &md[i]);

  /* p: first step whose filtered information can be full rank */
  int p = 0, acc = 0;
  for (p = 0; p < k; p++) {
    acc += R_(md[p].WG);
    if (acc >= n) break;
  }
  if (p == k) { fail("associative: observations do not determine the states"); goto done_models; }

  M** xs = (M**)calloc((size_t)k, sizeof(M*)); /* filtered estimates */
  M** Ps = (M**)calloc((size_t)k, sizeof(M*));
  M** Lams = (M**)calloc((size_t)k, sizeof(M*)); /* information form for steps < p */
  M** etas = (M**)calloc((size_t)k, sizeof(M*));

  /* information filter over steps 0..p (the prior, D3) */
  {
    M *Lam = NULL, *eta = NULL;
    for (int i = 0; i <= p; i++) {
      M *Lp, *ep;
      if (i == 0) {
        Lp = mnew(n, n);
        ep = mnew(n, 1);
      } else { /* predict */
        amodel* m = &md[i];
        M* Ft = mtrans(m->F);
        M* QiF = mmldivide(m->Q, m->F);
        M* Qic = mmldivide(m->Q, m->c);
        M* FtQiF = mmul(Ft, QiF);
        M* Mx = madd(Lam, FtQiF, 1.0);
        M* FtQic = mmul(Ft, Qic);
        M* r = madd(eta, FtQic, -1.0);
        M* Mr = mmldivide(Mx, r);
        M* QiFt = mtrans(QiF);
        M* MQ = mmldivide(Mx, QiFt);
        M* Qi = minverse(m->Q);
        M* QFMQ = mmul(QiF, MQ);
        Lp = madd(Qi, QFMQ, -1.0);
        M* QFMr = mmul(QiF, Mr);
        ep = madd(Qic, QFMr, 1.0);
        mfree(Ft); mfree(QiF); mfree(Qic); mfree(FtQiF); mfree(Mx); mfree(FtQic); mfree(r);
        mfree(Mr); mfree(QiFt); mfree(MQ); mfree(Qi); mfree(QFMQ); mfree(QFMr);
      }
      if (md[i].WG) { /* update */
        M* WGt = mtrans(md[i].WG);
        M* GG = mmul(WGt, md[i].WG);
        M* Go = mmul(WGt, md[i].Wo);
        mset(&Lp, madd(Lp, GG, 1.0));
        mset(&ep, madd(ep, Go, 1.0));
        mfree(WGt); mfree(GG); mfree(Go);
      }
      Lam = Lp;
      eta = ep;
      Lams[i] = Lam;
      etas[i] = eta;
    }
    if (p == 0) { /* the reference's own prior, c/kalman_associative_smoother.c:335-368 */
      M* R = mcopy(md[0].WG);
      M* tau = mqr(R);
      M* o = mcopy(md[0].Wo);
      mapply_qt(R, tau, o);
      mtriu(R);
      M* Rs = msub(R, 0, n, 0, n);
      M* x = msub(o, 0, n, 0, 1);
      mtrisolve('U', Rs, x);
      M* RT = mtrans(Rs);
      M* RTR = mmul(RT, Rs);
      xs[0] = x;
      Ps[0] = minverse(RTR);
      mfree(R); mfree(tau); mfree(o); mfree(Rs); mfree(RT); mfree(RTR);
    } else {
      Ps[p] = minverse(Lams[p]);
      xs[p] = mmul(Ps[p], etas[p]);
    }
  }

  /* filtering elements p+1 .. k-1 and the inclusive scan (sequential order,
   * c/parallel_sequential.c:30-50) */
  if (p < k - 1) {
    elem* fe = (elem*)calloc((size_t)k, sizeof(elem));
    PFOR for (int i = p + 1; i < k; i++)
      build_filtering(&md[i], n, i == p + 1, xs[p], Ps[p], &fe[i]);
    elem run;
    memset(&run, 0, sizeof(run));
    run.n = n;
    run.A = mcopy(fe[p + 1].A); run.b = mcopy(fe[p + 1].b); run.Z = mcopy(fe[p + 1].Z);
    run.e = mcopy(fe[p + 1].e); run.J = mcopy(fe[p + 1].J);
    xs[p + 1] = mcopy(run.b);
    Ps[p + 1] = mcopy(run.Z);
    for (int i = p + 2; i < k; i++) {
      elem nx;
      memset(&nx, 0, sizeof(nx));
      filt_combine(&run, &fe[i], &nx);
      elem_release(&run);
      run = nx;
      xs[i] = mcopy(run.b);
      Ps[i] = mcopy(run.Z);
    }
    elem_release(&run);
    for (int i = 0; i < k; i++) elem_release(&fe[i]);
    free(fe);
  }

  /* smoothing elements and the reverse inclusive scan */
  {
    elem* se = (elem*)calloc((size_t)k, sizeof(elem));
    PFOR for (int i = 0; i < k; i++) {
      if (i == k - 1) build_smoothing(xs[i], Ps[i], NULL, &se[i]);
      else if (i < p) build_smoothing_info(Lams[i], etas[i], &md[i + 1], &se[i]);
      else build_smoothing(xs[i], Ps[i], &md[i + 1], &se[i]);
    }
    elem run;
    memset(&run, 0, sizeof(run));
    run.n = n;
    run.E = mcopy(se[k - 1].E); run.g = mcopy(se[k - 1].g); run.L = mcopy(se[k - 1].L);
    if (!g_error) {
      eq[k - 1]->state = mcopy(xs[k - 1]); /* D4: filtered Z of the last step */
      eq[k - 1]->covariance = mcopy(Ps[k - 1]);
      eq[k - 1]->covariance_type = 'C';
    }
    for (int i = k - 2; i >= 0; i--) {
      elem nx;
      memset(&nx, 0, sizeof(nx));
      smooth_combine(&run, &se[i], &nx);
      elem_release(&run);
      run = nx;
      if (!g_error) {
        eq[i]->state = mcopy(run.g);
        eq[i]->covariance = mcopy(run.L);
        eq[i]->covariance_type = 'C';
      }
    }
    elem_release(&run);
    for (int i = 0; i < k; i++) elem_release(&se[i]);
    free(se);
  }

  for (int i = 0; i < k; i++) { mfree(xs[i]); mfree(Ps[i]); mfree(Lams[i]); mfree(etas[i]); }
  free(xs); free(Ps); free(Lams); free(etas);
done_models:
  for (int i = 0; i < k; i++) amodel_release(&md[i]);
  free(md);
}

/* ------------------------------------------------------------------------ */
/* global least squares (matlab/KalmanSparse.m:142-306)                       */
/* ------------------------------------------------------------------------ */

int oracle_dense_ls(kalman_step_equations_t** eq, kalman_step_index_t k, double* states, double* covs) {
  g_error = 0;
  int* off = (int*)malloc(sizeof(int) * ((size_t)k + 1));
  off[0] = 0;
  for (int i = 0; i < k; i++) off[i + 1] = off[i] + eq[i]->dimension;
  int N = off[k];
  int rows = 0;
  for (int i = 0; i < k; i++) {
    if (i > 0) rows += R_(eq[i]->H);
    if (eq[i]->o) rows += R_(eq[i]->o);
  }
  M* A = mnew(rows, N + 1); /* last column: right-hand side */
  int r = 0;
  for (int i = 0; i < k; i++) {
    if (i > 0) {
      M* VH = weigh(eq[i]->K, eq[i]->K_type, eq[i]->H);
      M* VF = weigh(eq[i]->K, eq[i]->K_type, eq[i]->F);
      M* Vc = weigh(eq[i]->K, eq[i]->K_type, eq[i]->c);
      for (int a = 0; a < R_(VH); a++) {
        for (int b = 0; b < C_(VF); b++) EL(A, r + a, off[i - 1] + b) = -EL(VF, a, b);
        for (int b = 0; b < C_(VH); b++) EL(A, r + a, off[i] + b) = EL(VH, a, b);
        EL(A, r + a, N) = Vc->elements[a];
      }
      r += R_(VH);
      mfree(VH); mfree(VF); mfree(Vc);
    }
    if (eq[i]->o) {
      M* WG = weigh(eq[i]->C, eq[i]->C_type, eq[i]->G);
      M* Wo = weigh(eq[i]->C, eq[i]->C_type, eq[i]->o);
      for (int a = 0; a < R_(WG); a++) {
        for (int b = 0; b < C_(WG); b++) EL(A, r + a, off[i] + b) = EL(WG, a, b);
        EL(A, r + a, N) = Wo->elements[a];
      }
      r += R_(WG);
      mfree(WG); mfree(Wo);
    }
  }
  int ok = 0;
  if (rows < N) { fail("dense_ls: fewer equations than unknowns"); ok = -1; goto out; }
  {
    M* tau = mqr(A);
    M* R = msub(A, 0, N, 0, N);
    mtriu(R);
    M* x = msub(A, 0, N, N, 1);
    mtrisolve('U', R, x);
    for (int i = 0; i < N; i++) states[i] = x->elements[i];
    M* Ri = meye(N);
    mtrisolve('U', R, Ri); /* R^{-1}; cov = R^{-1} R^{-T} */
    size_t co = 0;
    for (int s = 0; s < k; s++) {
      int n = eq[s]->dimension;
      for (int b = 0; b < n; b++)
        for (int a = 0; a < n; a++) {
          double acc = 0.0;
          for (int q = 0; q < N; q++) acc += EL(Ri, off[s] + a, q) * EL(Ri, off[s] + b, q);
          covs[co + (size_t)b * n + a] = acc;
        }
      co += (size_t)n * n;
    }
    mfree(tau); mfree(R); mfree(x); mfree(Ri);
    if (g_error) ok = -1;
  }
out:
  mfree(A);
  free(off);
  return ok;
}
