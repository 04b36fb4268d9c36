/*
 * oddeven.hip -- gfx950 kernels of the odd-even (Gargir-Toledo) smoother.
 *
 * Reference: c/kalman_oddeven_smoother.c (phases :379-560, fix-last
 * :1452-1476, renaming :563-597, back substitution and covariances
 * :614-975, recursion :1358-1553, weighing :1620-1659) and
 * matlab/KalmanOddevenSmoother.m:21-358.
 *
 * k_weigh      whitening of one step per workgroup (c/kalman_base.c:61-175).
 * k_eliminate  one workgroup per (even, odd) step pair of a level: the three
 *              Householder QRs of the reference's phases fused, writing the
 *              back-substitution factors and the next level's chain.
 * k_backsub    one workgroup per even step: estimate and, optionally, the
 *              selected-inversion covariance and the two cross covariances.
 *
 * Register layout of the QR kernels: every matrix column is owned by a pair
 * of adjacent lanes. The column has two row blocks of NB rows (block 0 holds
 * the pivot rows, block 1 the stacked rows); sub-lane b holds the rows with
 * parity b of both blocks, NB doubles. A Householder reflector is published
 * through double-buffered LDS by the pivot column's lanes; every other lane
 * only runs uniform chunk loops (no dynamic register indexing), combining its
 * two half dot products with one DPP swap.
 *
 * Rows that are zero in every column stay zero under all reflectors (their
 * reflector entries are zero), so padding blocks with zero rows yields an
 * equivalent least-squares problem: the fixed block layout replaces the
 * reference's data-dependent block heights.
 */
#include <hip/hip_runtime.h>

#include "ukhip_internal.h"

namespace {

__device__ __forceinline__ double pair_swap(double v) {
  int lo = __double2loint(v);
  int hi = __double2hiint(v);
  lo = __builtin_amdgcn_update_dpp(0, lo, 0xB1, 0xF, 0xF, false); /* quad_perm [1,0,3,2] */
  hi = __builtin_amdgcn_update_dpp(0, hi, 0xB1, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ int64_t find_problem(const int64_t* start, int nprob, int64_t x) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int NB>
struct Col {
  static constexpr int H = NB / 2; /* locals per block */
  double a[NB];                    /* [0,H): block 0 rows 2l+sub; [H,NB): block 1 rows 2(l-H)+sub */

  __device__ __forceinline__ void zero_block(int blk) {
#pragma unroll
    for (int l = 0; l < H; ++l) a[blk * H + l] = 0.0;
  }
  /* block blk <- rows [0, rows) of a column at src (unit stride), zero elsewhere */
  __device__ __forceinline__ void load_block(int blk, const double* __restrict__ src, int rows, int sub,
                                             double scale = 1.0) {
#pragma unroll
    for (int l = 0; l < H; ++l) {
      int r = 2 * l + sub;
      a[blk * H + l] = (r < rows) ? scale * src[r] : 0.0;
    }
  }
  __device__ __forceinline__ void store_block(int blk, double* __restrict__ dst, int rows, int sub) const {
#pragma unroll
    for (int l = 0; l < H; ++l) {
      int r = 2 * l + sub;
      if (r < rows) dst[r] = a[blk * H + l];
    }
  }
  /* upper-triangular store of block 0: rows > col are written as zero */
  __device__ __forceinline__ void store_triu(double* __restrict__ dst, int rows, int col, int sub) const {
#pragma unroll
    for (int l = 0; l < H; ++l) {
      int r = 2 * l + sub;
      if (r < rows) dst[r] = (r <= col) ? a[l] : 0.0;
    }
  }
  __device__ __forceinline__ void move_0_to_1() {
#pragma unroll
    for (int l = 0; l < H; ++l) a[H + l] = a[l];
  }
  __device__ __forceinline__ void move_1_to_0() {
#pragma unroll
    for (int l = 0; l < H; ++l) a[l] = a[H + l];
  }
};

/* LDS reflector buffers: [buffer][sub][NB + 4] (the +4 skews the two
 * sub-lane arrays onto different banks), then [buffer][2] scalars. */
template <int NB>
struct Refl {
  static constexpr int SUBSTRIDE = NB + 4;
  double v[2][2 * SUBSTRIDE];
  double sc[2][2];
};

/*
 * One Householder step for pivot k (target row k of block 0).
 *   n0: block-0 rows that can be nonzero (pivot rows included), n1: block-1 rows.
 *   is_piv: this lane pair owns pivot column k; upd: this column is updated.
 * Publishes v (v_k = 1, v_i = 0 for i < k) and tau, then applies
 * H = I - tau v v^T to every updated column. One barrier.
 */
template <int NB>
__device__ __forceinline__ void house_step(Col<NB>& col, Refl<NB>& R, int buf, int sub, bool is_piv, bool upd,
                                           int k, int n0, int n1) {
  constexpr int H = NB / 2;
  const int c0 = (k >> 1) & ~1;     /* first chunk (2 locals) of block 0 holding rows >= k */
  const int c0e = (n0 + 3) >> 1;    /* locals up to here may hold block-0 rows < n0 (rounded to chunks) */
  const int c1e = (n1 + 3) >> 1;
  double* vv = &R.v[buf][sub * Refl<NB>::SUBSTRIDE];
  if (is_piv) {
    double s = 0.0, alpha = 0.0;
#pragma unroll
    for (int l = 0; l < H; l += 2) {
      if (l >= c0 && l < c0e) {
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          const int r = 2 * (l + t) + sub;
          const double x = col.a[l + t];
          alpha = (r == k) ? x : alpha;
          s = (r > k) ? fma(x, x, s) : s;
        }
      }
    }
#pragma unroll
    for (int l = H; l < NB; l += 2) {
      if (l - H < c1e) s = fma(col.a[l], col.a[l], fma(col.a[l + 1], col.a[l + 1], s));
    }
    s += pair_swap(s);
    alpha += pair_swap(alpha);
    double tau = 0.0, beta = alpha, scale = 0.0;
    if (s != 0.0) {
      beta = -copysign(sqrt(fma(alpha, alpha, s)), alpha);
      tau = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
#pragma unroll
    for (int l = 0; l < H; l += 2) {
      if (l >= c0 && l < c0e) {
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          const int r = 2 * (l + t) + sub;
          vv[l + t] = (r > k) ? col.a[l + t] * scale : (r == k ? 1.0 : 0.0);
        }
      }
    }
#pragma unroll
    for (int l = H; l < NB; l += 2) {
      if (l - H < c1e) {
        vv[l] = col.a[l] * scale;
        vv[l + 1] = col.a[l + 1] * scale;
      }
    }
    if (sub == 0) {
      R.sc[buf][0] = tau;
      R.sc[buf][1] = beta;
    }
  }
  __syncthreads();
  if (upd) {
    const double tau = R.sc[buf][0];
    if (tau != 0.0) {
      double d = 0.0;
#pragma unroll
      for (int l = 0; l < H; l += 2)
        if (l >= c0 && l < c0e) d = fma(vv[l], col.a[l], fma(vv[l + 1], col.a[l + 1], d));
#pragma unroll
      for (int l = H; l < NB; l += 2)
        if (l - H < c1e) d = fma(vv[l], col.a[l], fma(vv[l + 1], col.a[l + 1], d));
      d += pair_swap(d);
      const double tw = tau * d;
      __asm__ volatile("" ::: "memory"); /* re-read v from LDS instead of holding it in VGPRs */
#pragma unroll
      for (int l = 0; l < H; l += 2)
        if (l >= c0 && l < c0e) {
          col.a[l] = fma(-tw, vv[l], col.a[l]);
          col.a[l + 1] = fma(-tw, vv[l + 1], col.a[l + 1]);
        }
#pragma unroll
      for (int l = H; l < NB; l += 2)
        if (l - H < c1e) {
          col.a[l] = fma(-tw, vv[l], col.a[l]);
          col.a[l + 1] = fma(-tw, vv[l + 1], col.a[l + 1]);
        }
    }
  }
}

/* QR of the stacked matrix over npiv pivot columns piv0 .. piv0+npiv-1. */
template <int NB>
__device__ __forceinline__ void house_qr(Col<NB>& col, Refl<NB>& R, int mycol, int sub, int piv0, int npiv, bool upd,
                                         int n0, int n1) {
  for (int k = 0; k < npiv; ++k) {
    house_step<NB>(col, R, k & 1, sub, mycol == piv0 + k, upd, k, n0, n1);
  }
  __syncthreads();
}

template <int NB>
struct ElimShared {
  Refl<NB> refl;
  double stashD[NB * NB + NB]; /* D~_{j+1} columns (mr each) then c~_{j+1} */
  double stashZ[NB * NB + NB]; /* Z_{L-1} columns then o_{L-1} (fix-last) */
};

/* QR tasks of one elimination work unit */
enum { T_LONE1 = 0, T_LONE2 = 1, T_P1 = 2, T_P2 = 3, T_P3 = 4, T_FIX = 5, T_BASE = 6 };

/*
 * Fused elimination of one level. Work unit u of problem p:
 *   L == 1      : the chain's single step (base case, :1360-1423)
 *   otherwise   : pair (j, j+1), j = 2u; the last unit of an odd-length
 *                 chain also eliminates the lone last step L-1 and folds it
 *                 into the pair's odd step (fix-last, :1452-1476).
 * Columns (lane pairs): P = [0,n) pivots of u_j; Q = [n,2n) u_{j+1};
 * R1 = 2n right-hand side; Bl = [2n+1, 3n+1) u_{j-1}.
 * Tasks: LONE1 = QR of [0; C_{L-1}], LONE2 = QR of [D_{L-1}; R~_{L-1}],
 *        P1 = QR of [B_{j+1}; C_j] (:379-457), P2 = QR of [D_j; R~_j]
 *        (:460-523), P3 = QR of [D~_{j+1}; C_{j+1}] (:526-560),
 *        FIX = QR of [G~_{L-2}; Z_{L-1}], BASE = QR of [0; C_0] when L == 1.
 */
template <int NB>
__global__ void __launch_bounds__(((2 * (3 * NB + 1) + 63) / 64) * 64)
    k_eliminate(uk_level_t lv, uk_chain_t in, uk_chain_t out, uk_factors_t fac, int n) {
  __shared__ ElimShared<NB> sh;
  constexpr int NT = ((2 * (3 * NB + 1) + 63) / 64) * 64;
  const int tid = threadIdx.x;
  const int sub = tid & 1;
  const int c = tid >> 1;
  const int64_t unit = blockIdx.x;
  const int p = (int)find_problem(lv.unit_start, lv.nprob, unit);
  const int64_t u = unit - lv.unit_start[p];
  const int64_t L = lv.L[p];
  const int64_t s0 = lv.step_start[p];
  const int mr = in.mr;
  const int64_t nn = (int64_t)n * n;

  const bool isP = c < n;
  const bool isQ = c >= n && c < 2 * n;
  const bool isR1 = c == 2 * n;
  const bool isBl = c > 2 * n && c <= 3 * n;
  const int qc = isP ? c : (isQ ? c - n : (isBl ? c - 2 * n - 1 : 0)); /* column within its block */

  for (int i = tid; i < (int)(sizeof(Refl<NB>) / sizeof(double)); i += NT)
    reinterpret_cast<double*>(&sh.refl)[i] = 0.0;

  int tasks[6];
  int ntask = 0;
  const int64_t j = 2 * u;
  const bool do_fix = L > 1 && (L & 1) && (u == L / 2 - 1);
  if (L == 1) {
    tasks[ntask++] = T_BASE;
  } else {
    if (do_fix) { tasks[ntask++] = T_LONE1; tasks[ntask++] = T_LONE2; }
    tasks[ntask++] = T_P1;
    if (j != 0) tasks[ntask++] = T_P2; /* the chain's first step has no [D; R~] */
    tasks[ntask++] = T_P3;
    if (do_fix) tasks[ntask++] = T_FIX;
  }
  const int64_t jl = L - 1;
  const int64_t t_next = lv.next_start[p] + u; /* new step for old step j+1 */
  const int64_t e_pair = lv.even_start[p] + u;
  const int64_t e_lone = lv.even_start[p] + jl / 2;

  Col<NB> col;
#pragma unroll
  for (int l = 0; l < NB; ++l) col.a[l] = 0.0;

  for (int ti = 0; ti < ntask; ++ti) {
    const int task = tasks[ti];
    /* ---- set up the stacked matrix ---- */
    int piv0 = 0, n1 = n;
    bool upd = false;
    switch (task) {
      case T_BASE:
      case T_LONE1: {
        const int64_t s = s0 + (task == T_BASE ? 0 : jl);
        if (isP) { col.zero_block(0); col.load_block(1, in.C + s * nn + (int64_t)qc * mr, mr, sub); }
        if (isR1) { col.zero_block(0); col.load_block(1, in.o + s * n, mr, sub); }
        upd = isP || isR1;
        n1 = mr;
      } break;
      case T_LONE2: {
        const int64_t s = s0 + jl;
        if (isP) { col.move_0_to_1(); col.load_block(0, in.D + s * nn + (int64_t)qc * n, n, sub); }
        if (isR1) { col.move_0_to_1(); col.load_block(0, in.c + s * n, n, sub); }
        if (isBl) { col.load_block(0, in.B + s * nn + (int64_t)qc * n, n, sub); col.zero_block(1); }
        upd = isP || isR1 || isBl;
      } break;
      case T_P1: {
        const int64_t sj = s0 + j, sp = s0 + j + 1;
        if (isP) {
          col.load_block(0, in.B + sp * nn + (int64_t)qc * n, n, sub);
          col.load_block(1, in.C + sj * nn + (int64_t)qc * mr, mr, sub);
        }
        if (isQ) { col.load_block(0, in.D + sp * nn + (int64_t)qc * n, n, sub); col.zero_block(1); }
        if (isR1) { col.load_block(0, in.c + sp * n, n, sub); col.load_block(1, in.o + sj * n, mr, sub); }
        upd = isP || isQ || isR1;
        n1 = mr;
      } break;
      case T_P2: {
        const int64_t sj = s0 + j;
        if (isP) { col.move_0_to_1(); col.load_block(0, in.D + sj * nn + (int64_t)qc * n, n, sub); }
        if (isQ) { col.move_0_to_1(); col.zero_block(0); }
        if (isR1) { col.move_0_to_1(); col.load_block(0, in.c + sj * n, n, sub); }
        if (isBl) { col.load_block(0, in.B + sj * nn + (int64_t)qc * n, n, sub); col.zero_block(1); }
        upd = isP || isQ || isR1 || isBl;
      } break;
      case T_P3: {
        const int64_t sp = s0 + j + 1;
        if (isQ) {
          col.load_block(0, sh.stashD + qc * mr, mr, sub);
          col.load_block(1, in.C + sp * nn + (int64_t)qc * mr, mr, sub);
        }
        if (isR1) {
          col.load_block(0, sh.stashD + (int64_t)n * mr, mr, sub);
          col.load_block(1, in.o + sp * n, mr, sub);
        }
        piv0 = n;
        upd = isQ || isR1;
        n1 = mr;
      } break;
      case T_FIX: {
        if (isQ) col.load_block(1, sh.stashZ + qc * n, n, sub);
        if (isR1) col.load_block(1, sh.stashZ + nn, n, sub);
        piv0 = n;
        upd = isQ || isR1;
      } break;
    }
    __syncthreads();

    /* ---- Householder QR over n pivot columns ---- */
    for (int k = 0; k < n; ++k) house_step<NB>(col, sh.refl, k & 1, sub, c == piv0 + k, upd, k, n, n1);
    __syncthreads();

    /* ---- store results ---- */
    switch (task) {
      case T_BASE:
        if (isP) {
          col.store_triu(fac.R + e_pair * nn + (int64_t)qc * n, n, qc, sub);
          for (int r = sub; r < n; r += 2) {
            fac.Bt[e_pair * nn + (int64_t)qc * n + r] = 0.0;
            fac.Y[e_pair * nn + (int64_t)qc * n + r] = 0.0;
          }
        }
        if (isR1) col.store_block(0, fac.cf + e_pair * n, n, sub);
        break;
      case T_LONE2:
        if (isP) {
          col.store_triu(fac.R + e_lone * nn + (int64_t)qc * n, n, qc, sub);
          for (int r = sub; r < n; r += 2) fac.Y[e_lone * nn + (int64_t)qc * n + r] = 0.0;
        }
        if (isBl) {
          col.store_block(0, fac.Bt + e_lone * nn + (int64_t)qc * n, n, sub);
          col.store_block(1, sh.stashZ + qc * n, n, sub); /* Z_{L-1} */
        }
        if (isR1) {
          col.store_block(0, fac.cf + e_lone * n, n, sub);
          col.store_block(1, sh.stashZ + nn, n, sub); /* o''_{L-1} */
        }
        break;
      case T_P1:
        if (isQ) col.store_block(1, sh.stashD + qc * mr, mr, sub);          /* D~_{j+1} */
        if (isR1) col.store_block(1, sh.stashD + (int64_t)n * mr, mr, sub); /* c~_{j+1} */
        if (j == 0) { /* the chain's first step: R_0 = R~_0, Y_0 = X_0, cf_0 = o_0 */
          if (isP) {
            col.store_triu(fac.R + e_pair * nn + (int64_t)qc * n, n, qc, sub);
            for (int r = sub; r < n; r += 2) fac.Bt[e_pair * nn + (int64_t)qc * n + r] = 0.0;
          }
          if (isQ) col.store_block(0, fac.Y + e_pair * nn + (int64_t)qc * n, n, sub);
          if (isR1) col.store_block(0, fac.cf + e_pair * n, n, sub);
        }
        break;
      case T_P2: {
          if (isP) col.store_triu(fac.R + e_pair * nn + (int64_t)qc * n, n, qc, sub);
          if (isBl) {
            col.store_block(0, fac.Bt + e_pair * nn + (int64_t)qc * n, n, sub);
            col.store_block(1, out.B + t_next * nn + (int64_t)qc * n, n, sub); /* Z_j -> new B */
          }
          if (isQ) {
            col.store_block(0, fac.Y + e_pair * nn + (int64_t)qc * n, n, sub);
            col.store_block(1, out.D + t_next * nn + (int64_t)qc * n, n, sub); /* X~_j -> new D */
          }
          if (isR1) {
            col.store_block(0, fac.cf + e_pair * n, n, sub);
            col.store_block(1, out.c + t_next * n, n, sub); /* o_j -> new c */
          }
        }
        break;
      case T_P3:
      case T_FIX:
        if (task == T_FIX || !do_fix) {
          if (isQ) col.store_triu(out.C + t_next * nn + (int64_t)qc * n, n, qc, sub); /* G~ -> new C */
          if (isR1) col.store_block(0, out.o + t_next * n, n, sub);                    /* c_{j+1} -> new o */
        }
        break;
      default:
        break;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* back substitution                                                          */
/* ------------------------------------------------------------------------ */

/* x <- R^{-1} x for upper-triangular R (ld n) in LDS, rinv = 1/diag(R). */
template <int NB>
__device__ __forceinline__ void back_solve(double (&x)[NB], const double* R, const double* rinv, int n) {
#pragma unroll
  for (int pp = NB - 1; pp >= 0; --pp) {
    if (pp < n) {
      const double xq = x[pp] * rinv[pp];
      x[pp] = xq;
#pragma unroll
      for (int i = 0; i < pp; ++i) x[i] = fma(-R[i + pp * n], xq, x[i]);
    }
  }
}

template <int NB>
struct BackShared {
  double R[NB * NB];
  double T[2 * NB * NB]; /* Bt | Y, later W (n x 2n), later K (n x n) */
  double rinv[NB];
  double xm[NB], xp[NB];
};

/*
 * Even step j of a level: x_j = R^{-1}(cf - Bt x_{j-1} - Y x_{j+1}); with
 * covariances, for T = [Bt Y] and Sigma the joint covariance of
 * (u_{j-1}, u_{j+1}) from the level above:
 *   W = T Sigma, U = R^{-1} W  ->  Cov(u_j, u_{j-1}) = -U_1, Cov(u_j, u_{j+1}) = -U_2
 *   S_j = R^{-1} (I + W T^T) R^{-T}.
 * (:614-975; the same quantities as the reference's S, S_off, S1_off.)
 * Cross covariances of a level are kept as X[s] = Cov(u_s, u_{s+1}) and
 * XT[s] = Cov(u_{s+1}, u_s), column-major n x n.
 */
template <int NB>
__global__ void __launch_bounds__(((2 * NB + 63) / 64) * 64)
    k_backsub(uk_level_t lv, uk_factors_t fac, double* __restrict__ xout, double* __restrict__ Sout,
              const double* __restrict__ Xup, const double* __restrict__ XTup, double* __restrict__ Xcur,
              double* __restrict__ XTcur, int n, int cov) {
  __shared__ BackShared<NB> sh;
  constexpr int NT = ((2 * NB + 63) / 64) * 64;
  const int tid = threadIdx.x;
  const int64_t ev = blockIdx.x;
  const int p = (int)find_problem(lv.even_start, lv.nprob, ev);
  const int64_t u = ev - lv.even_start[p];
  const int64_t j = 2 * u;
  const int64_t L = lv.L[p];
  const bool has_m = j > 0, has_p = j + 1 < L;
  const int64_t nn = (int64_t)n * n;
  const int64_t ob = lv.orig_start[lv.prob[p]];
  const int64_t stride = (int64_t)1 << lv.level;
  const int64_t oj = ob + (j + 1) * stride - 1;
  const int64_t om = ob + j * stride - 1;
  const int64_t op = ob + (j + 2) * stride - 1;
  const int64_t sj = lv.step_start[p] + j; /* this level's step index of j */
  const int64_t sup = lv.next_start[p] + (j >> 1) - 1; /* level+1 index of u_{j-1}: pair (j-1, j+1) */

  const double* Rg = fac.R + ev * nn;
  const double* Bg = fac.Bt + ev * nn;
  const double* Yg = fac.Y + ev * nn;
  for (int i = tid; i < n * n; i += NT) {
    sh.R[i] = Rg[i];
    sh.T[i] = has_m ? Bg[i] : 0.0;
    sh.T[nn + i] = has_p ? Yg[i] : 0.0;
  }
  for (int i = tid; i < n; i += NT) {
    sh.xm[i] = has_m ? xout[om * n + i] : 0.0;
    sh.xp[i] = has_p ? xout[op * n + i] : 0.0;
  }
  __syncthreads();
  for (int i = tid; i < n; i += NT) sh.rinv[i] = 1.0 / sh.R[i * n + i];
  __syncthreads();

  /* ---- estimate (lanes 0..n-1 of wave 0) ---- */
  if (tid < 64) {
    double r = 0.0;
    if (tid < n) {
      r = fac.cf[ev * n + tid];
      for (int q = 0; q < n; ++q) r -= sh.T[tid + q * n] * sh.xm[q] + sh.T[nn + tid + q * n] * sh.xp[q];
    }
    for (int q = n - 1; q >= 0; --q) {
      const double xq = __shfl(r, q) * sh.rinv[q];
      if (tid == q) r = xq;
      else if (tid < q) r -= sh.R[tid + q * n] * xq;
    }
    if (tid < n) xout[oj * n + tid] = r;
  }
  if (!cov) return;

  /* ---- W = T Sigma: lane q < 2n owns column q (n rows) ---- */
  const int q = tid;
  const bool act = q < 2 * n;
  const bool left = q < n;
  const int qq = left ? q : q - n;
  double v[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) v[i] = 0.0;
  if (act) {
    if (has_m) { /* rows p < n of Sigma: S_mm (left) or S_mp (right) */
      const double* Smm = Sout + om * nn;
      const double* Spm = XTup + sup * nn; /* Cov(u_{j+1}, u_{j-1}) */
      for (int pp = 0; pp < n; ++pp) {
        double sig;
        if (left) sig = Smm[qq + (int64_t)pp * n];            /* S_mm[pp, q] (symmetric) */
        else sig = has_p ? Spm[qq + (int64_t)pp * n] : 0.0;   /* S_mp[pp, qq] = S_pm[qq, pp] */
        const double* tc = &sh.T[pp * n];
#pragma unroll
        for (int i = 0; i < NB; ++i)
          if (i < n) v[i] = fma(tc[i], sig, v[i]);
      }
    }
    if (has_p) { /* rows p >= n: S_pm (left) or S_pp (right) */
      const double* Spp = Sout + op * nn;
      const double* Smp = Xup + sup * nn; /* Cov(u_{j-1}, u_{j+1}) */
      for (int pp = 0; pp < n; ++pp) {
        double sig;
        if (left) sig = has_m ? Smp[qq + (int64_t)pp * n] : 0.0; /* S_pm[pp, q] = S_mp[q, pp] */
        else sig = Spp[qq + (int64_t)pp * n];
        const double* tc = &sh.T[nn + pp * n];
#pragma unroll
        for (int i = 0; i < NB; ++i)
          if (i < n) v[i] = fma(tc[i], sig, v[i]);
      }
    }
  }
  __syncthreads(); /* T (Bt|Y) no longer needed in LDS */

  /* ---- W to LDS; U = R^{-1} W in place -> cross covariances ---- */
  if (act) {
#pragma unroll
    for (int i = 0; i < NB; ++i)
      if (i < n) sh.T[i + q * n] = v[i];
    back_solve<NB>(v, sh.R, sh.rinv, n);
    if (left && has_m) { /* Cov(u_j, u_{j-1}) = -U_1: XT[j-1] = -U_1, X[j-1] = -U_1^T */
      double* xt = XTcur + (sj - 1) * nn;
      double* x = Xcur + (sj - 1) * nn;
#pragma unroll
      for (int i = 0; i < NB; ++i)
        if (i < n) {
          xt[i + (int64_t)qq * n] = -v[i];
          x[qq + (int64_t)i * n] = -v[i];
        }
    }
    if (!left && has_p) { /* Cov(u_j, u_{j+1}) = -U_2: X[j] = -U_2, XT[j] = -U_2^T */
      double* x = Xcur + sj * nn;
      double* xt = XTcur + sj * nn;
#pragma unroll
      for (int i = 0; i < NB; ++i)
        if (i < n) {
          x[i + (int64_t)qq * n] = -v[i];
          xt[qq + (int64_t)i * n] = -v[i];
        }
    }
  }
  __syncthreads();

  /* ---- N = I + W T^T (lane q < n: column q), K = R^{-1} N ---- */
  if (q < n) {
#pragma unroll
    for (int i = 0; i < NB; ++i) v[i] = (i == q) ? 1.0 : 0.0;
    for (int pp = 0; pp < 2 * n; ++pp) {
      double tq;
      if (pp < n) tq = has_m ? Bg[q + (int64_t)pp * n] : 0.0; /* Bt[q, pp] */
      else tq = has_p ? Yg[q + (int64_t)(pp - n) * n] : 0.0; /* Y[q, pp-n] */
      const double* wc = &sh.T[pp * n];
#pragma unroll
      for (int i = 0; i < NB; ++i)
        if (i < n) v[i] = fma(wc[i], tq, v[i]);
    }
    back_solve<NB>(v, sh.R, sh.rinv, n);
  }
  __syncthreads(); /* W in LDS consumed */
  if (q < n) {
#pragma unroll
    for (int i = 0; i < NB; ++i)
      if (i < n) sh.T[i + q * n] = v[i];
  }
  __syncthreads();
  /* ---- S = R^{-1} K^T: lane q takes row q of K ---- */
  if (q < n) {
#pragma unroll
    for (int i = 0; i < NB; ++i) v[i] = (i < n) ? sh.T[q + i * n] : 0.0;
    back_solve<NB>(v, sh.R, sh.rinv, n);
    double* S = Sout + oj * nn;
#pragma unroll
    for (int i = 0; i < NB; ++i)
      if (i < n) S[q + (int64_t)i * n] = v[i]; /* S symmetric: column q written as row q */
  }
}

/* ------------------------------------------------------------------------ */
/* whitening (level-0 chain)                                                  */
/* ------------------------------------------------------------------------ */

template <int NB>
struct WeighShared {
  double K[NB * NB];
  double C[NB * NB];
};

/* In-place lower Cholesky of an m x m matrix (ld m) in LDS, whole block. */
template <int NT>
__device__ void lds_cholesky(double* A, int m, int tid) {
  for (int jj = 0; jj < m; ++jj) {
    if (tid == 0) A[jj + jj * m] = sqrt(A[jj + jj * m]);
    __syncthreads();
    const double d = A[jj + jj * m];
    for (int i = jj + 1 + tid; i < m; i += NT) A[i + jj * m] /= d;
    __syncthreads();
    const int rem = m - jj - 1;
    for (int idx = tid; idx < rem * rem; idx += NT) {
      const int i = jj + 1 + idx % rem, kk = jj + 1 + idx / rem;
      if (i >= kk) A[i + kk * m] -= A[i + jj * m] * A[kk + jj * m];
    }
    __syncthreads();
  }
}

/* x <- V x for a whitening V given as (code, matrix M in LDS, diag d);
 * src is x's column in global memory (re-read by the dense product). */
template <int NB>
__device__ __forceinline__ void apply_weight(double (&x)[NB], int m, int code, const double* M, const double* d,
                                             const double* src, int src_is_identity_col) {
  if (code == UK_COV_w) {
#pragma unroll
    for (int i = 0; i < NB; ++i)
      if (i < m) x[i] *= d[i];
  } else if (code == UK_COV_W) {
#pragma unroll
    for (int i = 0; i < NB; ++i) x[i] = 0.0;
    for (int pp = 0; pp < m; ++pp) {
      const double xp = src_is_identity_col >= 0 ? (pp == src_is_identity_col ? 1.0 : 0.0) : src[pp];
#pragma unroll
      for (int i = 0; i < NB; ++i)
        if (i < m) x[i] = fma(M[i + pp * m], xp, x[i]);
    }
  } else { /* lower-triangular solve with M (Cholesky factor or given L) */
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      if (i < m) {
        double s = x[i];
#pragma unroll
        for (int pp = 0; pp < i; ++pp) s = fma(-M[i + pp * m], x[pp], s);
        x[i] = s / M[i + i * m];
      }
    }
  }
}

/*
 * One step per workgroup. Lanes: [0,n) columns of F, [n,2n) columns of H,
 * 2n: c, [2n+1, 3n+1) columns of G, 3n+1: o.
 */
template <int NB>
__global__ void __launch_bounds__(((3 * NB + 2 + 63) / 64) * 64) k_weigh(uk_raw_t raw, uk_chain_t out) {
  __shared__ WeighShared<NB> sh;
  constexpr int NT = ((3 * NB + 2 + 63) / 64) * 64;
  const int tid = threadIdx.x;
  const int64_t s = blockIdx.x;
  const int n = raw.n, mr = raw.mr;
  const int64_t nn = (int64_t)n * n;
  const int m = raw.m[s];
  const bool first = raw.first[s] != 0;
  const int kt = raw.Kt[s], ct = raw.Ct[s];

  if (!first) {
    if (raw.K_full) {
      for (int i = tid; i < n * n; i += NT) sh.K[i] = raw.K[s * nn + i];
    } else {
      for (int i = tid; i < n; i += NT) sh.K[i] = raw.K[s * n + i];
    }
  }
  if (m > 0) {
    if (raw.C_full) {
      for (int i = tid; i < m * m; i += NT) sh.C[i] = raw.C[s * mr * mr + (i % m) + (int64_t)(i / m) * mr];
    } else {
      for (int i = tid; i < m; i += NT) sh.C[i] = raw.C[s * mr + i];
    }
  }
  __syncthreads();
  if (!first && kt == UK_COV_C) lds_cholesky<NT>(sh.K, n, tid);
  if (m > 0 && ct == UK_COV_C) lds_cholesky<NT>(sh.C, m, tid);
  __syncthreads();

  double x[NB];
  const int q = tid;
  if (q < 2 * n + 1) {
    if (first) { /* no evolution equation on a chain's first step */
      return;
    }
    if (q < n) {
      const double* src = raw.F + s * nn + (int64_t)q * n;
#pragma unroll
      for (int i = 0; i < NB; ++i) x[i] = (i < n) ? src[i] : 0.0;
    } else if (q < 2 * n) {
      const int qq = q - n;
#pragma unroll
      for (int i = 0; i < NB; ++i) x[i] = (i < n) ? (raw.H_ident ? (i == qq ? 1.0 : 0.0) : raw.H[s * nn + (int64_t)qq * n + i]) : 0.0;
    } else {
#pragma unroll
      for (int i = 0; i < NB; ++i) x[i] = (i < n) ? raw.c[s * n + i] : 0.0;
    }
    const double* src = q < n ? raw.F + s * nn + (int64_t)q * n
                        : (q < 2 * n ? (raw.H_ident ? nullptr : raw.H + s * nn + (int64_t)(q - n) * n) : raw.c + s * n);
    apply_weight<NB>(x, n, kt, sh.K, sh.K, src, (q >= n && q < 2 * n && raw.H_ident) ? q - n : -1);
    if (q < n) {
      double* dst = out.B + s * nn + (int64_t)q * n;
#pragma unroll
      for (int i = 0; i < NB; ++i)
        if (i < n) dst[i] = -x[i];
    } else if (q < 2 * n) {
      double* dst = out.D + s * nn + (int64_t)(q - n) * n;
#pragma unroll
      for (int i = 0; i < NB; ++i)
        if (i < n) dst[i] = x[i];
    } else {
#pragma unroll
      for (int i = 0; i < NB; ++i)
        if (i < n) out.c[s * n + i] = x[i];
    }
  } else if (q < 3 * n + 2) {
    const bool is_o = q == 3 * n + 1;
    const int qq = q - 2 * n - 1;
    if (m > 0) {
      const double* src = is_o ? raw.o + s * mr : raw.G + s * mr * n + (int64_t)qq * mr;
#pragma unroll
      for (int i = 0; i < NB; ++i) x[i] = (i < m) ? src[i] : 0.0;
      apply_weight<NB>(x, m, ct, sh.C, sh.C, src, -1);
    } else {
#pragma unroll
      for (int i = 0; i < NB; ++i) x[i] = 0.0;
    }
    double* dst = is_o ? out.o + s * n : out.C + s * nn + (int64_t)qq * mr;
#pragma unroll
    for (int i = 0; i < NB; ++i)
      if (i < mr) dst[i] = (i < m) ? x[i] : 0.0;
  }
}

} /* namespace */

/* ------------------------------------------------------------------------ */
/* launchers (C linkage, called from ukhip_host.cpp)                          */
/* ------------------------------------------------------------------------ */

#define UK_DISPATCH(NBV, CALL) \
  case NBV: { CALL(NBV); } break;

extern "C" int uk_bucket(int n) {
  if (n <= 8) return 8;
  if (n <= 16) return 16;
  if (n <= 32) return 32;
  if (n <= 48) return 48;
  if (n <= 64) return 64;
  return -1;
}

extern "C" hipError_t uk_launch_weigh(const uk_raw_t* raw, const uk_chain_t* out, hipStream_t st) {
  const int nb = uk_bucket(raw->n > raw->mr ? raw->n : raw->mr);
#define L_(NBV) k_weigh<NBV><<<dim3((unsigned)raw->steps), dim3(((3 * NBV + 2 + 63) / 64) * 64), 0, st>>>(*raw, *out)
  switch (nb) {
    UK_DISPATCH(8, L_) UK_DISPATCH(16, L_) UK_DISPATCH(32, L_) UK_DISPATCH(48, L_) UK_DISPATCH(64, L_)
    default: return hipErrorInvalidValue;
  }
#undef L_
  return hipGetLastError();
}

extern "C" hipError_t uk_launch_eliminate(const uk_level_t* lv, const uk_chain_t* in, const uk_chain_t* out,
                                          const uk_factors_t* fac, int n, hipStream_t st) {
  const int nb = uk_bucket(n > in->mr ? n : in->mr);
#define L_(NBV) \
  k_eliminate<NBV><<<dim3((unsigned)lv->units), dim3(((2 * (3 * NBV + 1) + 63) / 64) * 64), 0, st>>>(*lv, *in, *out, *fac, n)
  switch (nb) {
    UK_DISPATCH(8, L_) UK_DISPATCH(16, L_) UK_DISPATCH(32, L_) UK_DISPATCH(48, L_) UK_DISPATCH(64, L_)
    default: return hipErrorInvalidValue;
  }
#undef L_
  return hipGetLastError();
}

extern "C" hipError_t uk_launch_backsub(const uk_level_t* lv, const uk_factors_t* fac, double* xout, double* Sout,
                                        const double* Xup, const double* XTup, double* Xcur, double* XTcur, int n,
                                        int cov, hipStream_t st) {
  const int nb = uk_bucket(n);
#define L_(NBV)                                                                                                  \
  k_backsub<NBV><<<dim3((unsigned)lv->evens), dim3(((2 * NBV + 63) / 64) * 64), 0, st>>>(*lv, *fac, xout, Sout, Xup, \
                                                                                       XTup, Xcur, XTcur, n, cov)
  switch (nb) {
    UK_DISPATCH(8, L_) UK_DISPATCH(16, L_) UK_DISPATCH(32, L_) UK_DISPATCH(48, L_) UK_DISPATCH(64, L_)
    default: return hipErrorInvalidValue;
  }
#undef L_
  return hipGetLastError();
}
