/*
 * ukhip_internal.h -- device data layout shared by the host planner and the
 * HIP kernels of libukhip.
 *
 * A "batch" is a set of independent problems with one state dimension n.
 * Steps of all problems are concatenated; every per-step matrix is stored
 * column-major in structure-of-arrays buffers (one buffer per field, step s
 * at s * stride). The odd-even elimination runs level by level over the whole
 * batch: level l holds, per problem, L_l steps (L_0 = k, L_{l+1} = L_l / 2).
 */
#ifndef UKHIP_INTERNAL_H
#define UKHIP_INTERNAL_H

#include <stdint.h>

#define UKHIP_MAX_N 64

/* covariance representation codes (c/kalman_base.c:61-175) */
enum { UK_COV_W = 0, UK_COV_w = 1, UK_COV_C = 2, UK_COV_L = 3 /* 'U' and 'F' */ };

/* Raw (unweighed) inputs of a batch, as packed by the host. */
typedef struct {
  int32_t n;        /* state dimension */
  int32_t mr;       /* observation rows per step (max over steps) */
  int64_t steps;    /* total steps over all problems */
  int32_t H_ident;  /* 1: every H_i is the identity (H not stored) */
  int32_t K_full;   /* 1: K stored n*n per step, else n (all 'w') */
  int32_t C_full;   /* 1: C stored mr*mr per step, else mr (all 'w') */
  const double* F;  /* F blocks, n*n each: step s uses block Fidx[s] (first step of a problem unused) */
  const int64_t* Fidx; /* steps: F block of the step (a problem with time-invariant F has one block) */
  const double* H;  /* steps * n*n or NULL */
  const double* c;  /* steps * n */
  const double* K;  /* steps * (n*n | n) */
  const double* G;  /* steps * mr*n  (ld = mr; rows >= m_i unused) */
  const double* o;  /* steps * mr */
  const double* C;  /* steps * (mr*mr | mr), ld = mr */
  const int8_t* Kt; /* steps: UK_COV_* */
  const int8_t* Ct; /* steps: UK_COV_* */
  const int32_t* m; /* steps: observation count m_i (0 = unobserved) */
  const int8_t* first; /* steps: 1 on the first step of a problem */
} uk_raw_t;

/* Per-level step arrays: the whitened chain of one elimination level.
 *   B = -V F, D = V H, c = V c   (n x n, n x n, n; absent on a chain's first step)
 *   C = W G (mr x n, ld mr), o = W o (mr)                                        */
typedef struct {
  double* B;
  double* D;
  double* c;
  double* C;
  double* o;
  int32_t mr;
  int32_t d_diag; /* 1: every D is diagonal and stored as its diagonal, the first n entries of its
                     slot (level 0 with H = I and every K and C given as 'w') */
  int32_t c_tri;  /* 1: every C is upper triangular n x n (the levels above 0: phase 3's G~) */
  /* b_synth = 1: B is not stored; B_s = -diag(wk_s) wF[wFidx[s]] (wFidx NULL:
   * every step uses block 0) is formed on load (level 0 with every K and C
   * given as 'w': wk = raw K, wF = raw F) */
  int32_t b_synth;
  const double* wk;
  const double* wF;
  const int64_t* wFidx;
  /* from_raw = 1 (read by k_normal_w; requires d_diag and b_synth): C, o, c
   * and D's diagonal are not read from the chain but formed on load from the
   * raw inputs as k_weigh_diag forms them: C = diag(w) G (rows >= m_s zero),
   * o = w o, c = k c, diag(D) = k, with w = rC, k = wk (the mixed-precision
   * path, whose whitening then writes only the FP32 chain) */
  int32_t from_raw;
  const double* rG;
  const double* rC;
  const double* ro;
  const double* rc;
} uk_chain_t;

/* The same chain in single precision (the FP32 elimination). c_synth = 1
 * (level 0, every C given as 'w'): C is not stored; C_s = diag(rC_s) rG_s
 * (rows >= rm[s] zero) is formed on load, rounded to FP32 as the whitening
 * would store it. */
typedef struct {
  float* B;
  float* D;
  float* c;
  float* C;
  float* o;
  int32_t mr;
  int32_t d_diag;
  int32_t c_tri;
  int32_t b_synth;
  const double* wk;
  const double* wF;
  const int64_t* wFidx;
  int32_t c_synth;
  const double* rG;
  const double* rC;
  const int32_t* rm;
} uk_chain32_t;

/* Per-level factors kept for the back substitution (one slot per even step):
 *   R_j u_j + Bt_j u_{j-1} + Y_j u_{j+1} = cf_j */
typedef struct {
  double* R;
  double* Bt;
  double* Y;
  double* cf;
} uk_factors_t;

typedef struct {
  float* R;
  float* Bt;
  float* Y;
  float* cf;
} uk_factors32_t;

/* Level plan: problem p of the batch has L[p] steps at this level. Work units
 * of the elimination are numbered unit_start[p] .. unit_start[p+1]-1; even
 * steps (back substitution) even_start[p] ..; steps step_start[p] ..; the
 * next level's steps next_start[p] ..; original step numbers orig_start[p] + t. */
typedef struct {
  int32_t nprob;        /* problems active at this level */
  int32_t level;
  int64_t units;        /* elimination work units */
  int64_t evens;        /* even steps */
  const int32_t* prob;  /* nprob: batch problem index (for orig_start) */
  const int64_t* L;
  const int64_t* unit_start; /* nprob + 1 */
  const int64_t* even_start; /* nprob + 1 */
  const int64_t* step_start;
  const int64_t* next_start;
  const int64_t* orig_start; /* indexed by batch problem */
} uk_level_t;

/* Everything one odd-even smoothing of a batch reads and writes on the
 * device, as plain pointers: the raw inputs, the two chains, the factor slots
 * (level l's at fac_base[l]), the cross-covariance buffers, the outputs, the
 * per-problem fix-last scratch and, when observation blocks are reduced
 * (compress), their staging. levels and fac_base are host arrays of nl
 * entries. ukhip_host.cpp fills one from a batch; the allocation-free entry
 * point fills one from its static device scratch array. */
typedef struct {
  uk_raw_t raw;
  uk_chain_t chain[2];
  uk_factors_t fac;
  double* X[2];
  double* XT[2];
  double* xout;
  double* Sout;
  double* fix;
  double* wobs;
  int32_t n, mr0, compress, nl;
  int64_t nF;
  const uk_level_t* levels;
  const int64_t* fac_base;
} uk_oe_view_t;

/* Batched sequential filter (UltimateKalman, c/kalman_ultimate.c:140-470):
 * per filter the information factor R (n x n, upper triangular, zero rows
 * while the state is undetermined), its right-hand side y (n), the number of
 * determined rows z, and the estimate x (n). */
typedef struct {
  double* R;
  double* y;
  int32_t* z;
  double* x;
  /* history for the smoother pass (hist != 0): per step slot s (the step's
   * final factors, fixed by the next step's evolve), Rd (n x n upper
   * triangular: Rdiag), Rs (n x n: Rsupdiag, the coupling to the next step)
   * and hy (n); slot s of filter f at (s * nf + f) * (n*n | n) */
  double* hRd;
  double* hRs;
  double* hy;
  int64_t hslot; /* slot the current launch writes (the previous step) */
  int32_t hist;
  int32_t nf;
} uk_filter_t;

/* One step's inputs for all filters (device pointers): F (n*n per filter, or
 * one when F_shared), H (n*n per filter, or NULL for I), c (n per filter, or
 * NULL for 0), K ('w': n inverse standard deviations per filter), G (mr*n per
 * filter, ld mr), o (mr), m_i (observation rows per filter, or NULL for mr),
 * C ('w': mr per filter). */
typedef struct {
  const double* F;
  const double* H;
  const double* c;
  const double* K;
  const double* G;
  const double* o;
  const int32_t* m_i;
  const double* C;
  int32_t F_shared;
  int32_t mr;
  /* pre != 0: the step is already whitened (k_weigh, for K or C given in a
   * full representation): wB = -V F, wD = V H (n*n per filter), wc = V c (n),
   * wC = W G (ld mr, an n*n slot per filter), wo = W o (an n slot); F, H, c,
   * K, G, o and C are then not read */
  const double* wB;
  const double* wD;
  const double* wc;
  const double* wC;
  const double* wo;
  int32_t pre;
} uk_filter_in_t;

#endif
