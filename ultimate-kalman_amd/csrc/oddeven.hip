/*
 * oddeven.hip -- gfx950 kernels of the odd-even (Gargir-Toledo) smoother.
 *
 * Reference: c/kalman_oddeven_smoother.c (phases :379-560, fix-last
 * :1452-1476, renaming :563-597, back substitution and covariances
 * :614-975, recursion :1358-1553, weighing :1620-1659) and
 * matlab/KalmanOddevenSmoother.m:21-358.
 *
 * k_weigh      whitening of one step per workgroup (c/kalman_base.c:61-175).
 * k_phase1..3  the reference's three elimination phases of one level, one
 *              workgroup per step pair (or even step), writing the
 *              back-substitution factors and the next level's chain
 *              (oddeven_qr.h; the buckets above 16 are compiled in
 *              oddeven_qr{32,48,64}.hip).
 * k_backsub    one workgroup per even step: estimate and, optionally, the
 *              selected-inversion covariance and the two cross covariances.
 *
 * Register layout of the QR kernels: every matrix column is owned by SL = 2
 * or 4 adjacent lanes. The column has two row blocks of NB rows (block 0 holds
 * the pivot rows, block 1 the stacked rows); sub-lane b holds the rows
 * r = b (mod SL) of both blocks, 2 NB / SL doubles. A Householder reflector is
 * published through double-buffered LDS by the pivot column's lanes; every
 * other lane runs branch-free loops over its rows (no dynamic register
 * indexing), combining its partial dot products with DPP quad permutations.
 *
 * Rows that are zero in every column stay zero under all reflectors (their
 * reflector entries are zero), so padding blocks with zero rows yields an
 * equivalent least-squares problem: the fixed block layout replaces the
 * reference's data-dependent block heights.
 */
#include <hip/hip_runtime.h>

#include "oddeven_qr.h"
#include "ukhip_internal.h"

namespace {
using ukqr::ChainT;
using ukqr::FactorsT;

/* ------------------------------------------------------------------------ */
/* back substitution                                                          */
/* ------------------------------------------------------------------------ */

/* x <- R^{-1} x for upper-triangular R in LDS (ld NB, identity-padded beyond
 * n, so padded entries of x stay zero), rinv = 1/diag(R). */
template <int NB>
__device__ __forceinline__ void back_solve(double (&x)[NB], const double* R, const double* rinv) {
#pragma unroll
  for (int pp = NB - 1; pp >= 0; --pp) {
    const double xq = x[pp] * rinv[pp];
    x[pp] = xq;
#pragma unroll
    for (int i = 0; i < pp; ++i) x[i] = fma(-R[i + pp * NB], xq, x[i]);
  }
}

template <int NB>
struct BackShared {
  double R[NB * NB];     /* ld NB, identity beyond n */
  double T[2 * NB * NB]; /* Bt | Y (ld NB, zero-padded), later W (n x 2n), later K */
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
 * XT[s] = Cov(u_{s+1}, u_s), column-major n x n; Xcur = NULL (level 0, which
 * no level below reads) skips writing them.
 */
template <int NB>
__global__ void __launch_bounds__(((2 * NB + 63) / 64) * 64)
    k_backsub(uk_level_t lv, uk_factors_t fac, double* __restrict__ xout, double* __restrict__ Sout,
              const double* __restrict__ Xup, const double* __restrict__ XTup, double* __restrict__ Xcur,
              double* __restrict__ XTcur, int n, int cov) {
  __shared__ BackShared<NB> sh;
  constexpr int NT = ((2 * NB + 63) / 64) * 64;
  constexpr int NN = NB * NB;
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
  const int64_t sj = lv.step_start[p] + j;             /* this level's step index of j */
  const int64_t sup = lv.next_start[p] + (j >> 1) - 1; /* level+1 index of u_{j-1} */

  const double* Rg = fac.R + ev * nn;
  const double* Bg = fac.Bt + ev * nn;
  const double* Yg = fac.Y + ev * nn;
  for (int idx = tid; idx < NN; idx += NT) {
    const int i = idx % NB, c = idx / NB;
    const bool in = i < n && c < n;
    sh.R[idx] = in ? Rg[i + c * n] : (i == c ? 1.0 : 0.0);
    sh.T[idx] = (in && has_m) ? Bg[i + c * n] : 0.0;
    sh.T[NN + idx] = (in && has_p) ? Yg[i + c * n] : 0.0;
  }
  for (int i = tid; i < NB; i += NT) {
    sh.xm[i] = (has_m && i < n) ? xout[om * n + i] : 0.0;
    sh.xp[i] = (has_p && i < n) ? xout[op * n + i] : 0.0;
  }
  __syncthreads();
  for (int i = tid; i < NB; i += NT) sh.rinv[i] = 1.0 / sh.R[i * NB + i];
  __syncthreads();

  /* ---- estimate (lanes 0..n-1 of wave 0) ---- */
  if (tid < 64) {
    double r = 0.0;
    if (tid < n) {
      r = fac.cf[ev * n + tid];
      for (int q = 0; q < n; ++q) r -= sh.T[tid + q * NB] * sh.xm[q] + sh.T[NN + tid + q * NB] * sh.xp[q];
    }
    for (int q = n - 1; q >= 0; --q) {
      const double xq = __shfl(r, q) * sh.rinv[q];
      if (tid == q) r = xq;
      else if (tid < q) r -= sh.R[tid + q * NB] * xq;
    }
    if (tid < n) xout[oj * n + tid] = r;
  }
  if (!cov) return;

  /* ---- W = T Sigma: lane q < 2n owns column q (NB rows, zero beyond n) ---- */
  const int q = tid;
  const bool act = q < 2 * n;
  const bool left = q < n;
  const int qq = left ? q : q - n;
  double v[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) v[i] = 0.0;
  if (act) {
    /* Sigma's column q streams from global memory in batches of PF rows,
     * all loads of a batch issued before its FMAs. */
    constexpr int PF = 8;
    if (has_m) { /* rows p < n of Sigma: S_mm (left) or S_mp (right) */
      const double* src = left ? Sout + om * nn : XTup + sup * nn; /* S_mm, or S_pm = Cov(u_{j+1}, u_{j-1}) */
      const bool live = left || has_p;
      for (int p0 = 0; p0 < n; p0 += PF) {
        double sg[PF];
#pragma unroll
        for (int t = 0; t < PF; ++t) sg[t] = (live && p0 + t < n) ? src[qq + (int64_t)(p0 + t) * n] : 0.0;
#pragma unroll
        for (int t = 0; t < PF; ++t) {
          if (p0 + t < n) {
            const double* tc = &sh.T[(p0 + t) * NB];
#pragma unroll
            for (int i = 0; i < NB; ++i) v[i] = fma(tc[i], sg[t], v[i]);
          }
        }
      }
    }
    if (has_p) { /* rows p >= n: S_pm (left) or S_pp (right) */
      const double* src = left ? Xup + sup * nn : Sout + op * nn; /* S_mp = Cov(u_{j-1}, u_{j+1}), or S_pp */
      const bool live = !left || has_m;
      for (int p0 = 0; p0 < n; p0 += PF) {
        double sg[PF];
#pragma unroll
        for (int t = 0; t < PF; ++t) sg[t] = (live && p0 + t < n) ? src[qq + (int64_t)(p0 + t) * n] : 0.0;
#pragma unroll
        for (int t = 0; t < PF; ++t) {
          if (p0 + t < n) {
            const double* tc = &sh.T[NN + (p0 + t) * NB];
#pragma unroll
            for (int i = 0; i < NB; ++i) v[i] = fma(tc[i], sg[t], v[i]);
          }
        }
      }
    }
  }
  __syncthreads(); /* T (Bt|Y) no longer needed in LDS */

  /* ---- W to LDS; U = R^{-1} W in place -> cross covariances ---- */
  if (act) {
#pragma unroll
    for (int i = 0; i < NB; ++i) sh.T[i + q * NB] = v[i];
    back_solve<NB>(v, sh.R, sh.rinv);
    if (Xcur && left && has_m) { /* Cov(u_j, u_{j-1}) = -U_1: XT[j-1] = -U_1, X[j-1] = -U_1^T */
      double* xt = XTcur + (sj - 1) * nn;
      double* x = Xcur + (sj - 1) * nn;
#pragma unroll
      for (int i = 0; i < NB; ++i)
        if (i < n) {
          xt[i + (int64_t)qq * n] = -v[i];
          x[qq + (int64_t)i * n] = -v[i];
        }
    }
    if (Xcur && !left && has_p) { /* Cov(u_j, u_{j+1}) = -U_2: X[j] = -U_2, XT[j] = -U_2^T */
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
    constexpr int PF = 8;
    for (int p0 = 0; p0 < 2 * n; p0 += PF) {
      double tq[PF];
#pragma unroll
      for (int t = 0; t < PF; ++t) {
        const int pp = p0 + t;
        tq[t] = 0.0;
        if (pp < n) tq[t] = has_m ? Bg[q + (int64_t)pp * n] : 0.0;            /* Bt[q, pp] */
        else if (pp < 2 * n) tq[t] = has_p ? Yg[q + (int64_t)(pp - n) * n] : 0.0; /* Y[q, pp-n] */
      }
#pragma unroll
      for (int t = 0; t < PF; ++t) {
        if (p0 + t < 2 * n) {
          const double* wc = &sh.T[(p0 + t) * NB];
#pragma unroll
          for (int i = 0; i < NB; ++i) v[i] = fma(wc[i], tq[t], v[i]);
        }
      }
    }
    back_solve<NB>(v, sh.R, sh.rinv);
  }
  __syncthreads(); /* W in LDS consumed */
  if (q < n) {
#pragma unroll
    for (int i = 0; i < NB; ++i) sh.T[i + q * NB] = v[i];
  }
  __syncthreads();
  /* ---- S = R^{-1} K^T: lane q takes row q of K (rows beyond n read as zero) ---- */
  if (q < n) {
#pragma unroll
    for (int i = 0; i < NB; ++i) v[i] = (i < n) ? sh.T[q + i * NB] : 0.0;
    back_solve<NB>(v, sh.R, sh.rinv);
    double* S = Sout + oj * nn;
    if (cov == 2) { /* diagonal only (level 0 with UKHIP_COVARIANCE_DIAG): S_qq in its place */
#pragma unroll
      for (int i = 0; i < NB; ++i)
        if (i == q) S[q + (int64_t)q * n] = v[i];
    } else {
#pragma unroll
      for (int i = 0; i < NB; ++i)
        if (i < n) S[q + (int64_t)i * n] = v[i]; /* S symmetric: column q written as row q */
    }
  }
}

/* ------------------------------------------------------------------------ */
/* estimates only (KALMAN_NO_COVARIANCE), n > 16                               */
/* ------------------------------------------------------------------------ */

__device__ __forceinline__ double lane_bcast(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

/*
 * x_j = R^{-1} (cf - Bt x_{j-1} - Y x_{j+1}) for one even step per wave: lane
 * i owns row i. Bt and Y stream column by column (coalesced), the
 * neighbours' estimates and each solved x_q reach every lane by readlane,
 * and lane i keeps row i of R in registers for the substitution.
 */
template <int NB>
__global__ void __launch_bounds__(64) k_estimate(uk_level_t lv, uk_factors_t fac, double* __restrict__ xout, int n) {
  const int i = threadIdx.x;
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
  const double* Rg = fac.R + ev * nn;
  const double* Bg = fac.Bt + ev * nn;
  const double* Yg = fac.Y + ev * nn;
  const bool row = i < n;
  double rrow[NB];
#pragma unroll
  for (int q = 0; q < NB; ++q) rrow[q] = (row && q < n && q >= i) ? Rg[i + (int64_t)q * n] : 0.0;
  const double rinv = row ? 1.0 / Rg[i + (int64_t)i * n] : 0.0;
  const double xm = (has_m && row) ? xout[om * n + i] : 0.0;
  const double xp = (has_p && row) ? xout[op * n + i] : 0.0;
  double r = row ? fac.cf[ev * n + i] : 0.0;
#pragma unroll 8
  for (int q = 0; q < NB; ++q) {
    if (q < n) {
      const double b = (has_m && row) ? Bg[i + (int64_t)q * n] : 0.0;
      const double y = (has_p && row) ? Yg[i + (int64_t)q * n] : 0.0;
      r = fma(-b, lane_bcast(xm, q), r);
      r = fma(-y, lane_bcast(xp, q), r);
    }
  }
#pragma unroll
  for (int q = NB - 1; q >= 0; --q) {
    if (q < n) {
      const double xq = lane_bcast(r * rinv, q);
      r = (i == q) ? xq : fma(-rrow[q], xq, r);
    }
  }
  if (row) xout[oj * n + i] = r;
}

/* ------------------------------------------------------------------------ */
/* back substitution on the matrix cores (NB >= 32, with covariances)         */
/* ------------------------------------------------------------------------ */

typedef double dx4 __attribute__((ext_vector_type(4)));

/* One 16x16x4 FP64 matrix-core step, D = A B + D. Lane l supplies A[l%16][l/16]
 * and B[l/16][l%16]; D[4r+l/16][l%16] is element r of the lane's dx4. */
__device__ __forceinline__ dx4 mfma16(double a, double b, dx4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

/* x <- R^{-1} x for upper-triangular R in LDS with leading dimension LD. */
template <int NB, int LD>
__device__ __forceinline__ void back_solve_ld(double (&x)[NB], const double* R, const double* rinv) {
#pragma unroll
  for (int pp = NB - 1; pp >= 0; --pp) {
    const double xq = x[pp] * rinv[pp];
    x[pp] = xq;
#pragma unroll
    for (int i = 0; i < pp; ++i) x[i] = fma(-R[i + pp * LD], xq, x[i]);
  }
}

/* Four NB x NB blocks (column-major, odd leading dimension NB + 1, so that
 * both the unit-stride and the LD-stride fragment reads of the matrix cores
 * spread over the LDS banks) and the step's vectors. Block use over time:
 *   [0] Bt -> R -> R^{-1}   [1] Y -> N = I + W T^T   [2..3] W -> [2] K = R^{-1} N */
template <int NB>
struct BackMShared {
  static constexpr int LD = NB + 1;
  double blk[4][NB * LD];
  double r[NB]; /* cf - Bt x_{j-1} - Y x_{j+1} */
  double xm[NB], xp[NB];
  double rinv[NB];
};

/*
 * The back substitution of one even step per workgroup (NB / 8 waves), the
 * same quantities as k_backsub with covariances:
 *   W^T = Sigma T^T (Sigma symmetric: its rows stream from HBM as matrix-core
 *   A fragments, 16 consecutive doubles per fragment row), N = I + W T^T,
 *   R^{-1} by per-lane back substitution of the unit vectors,
 *   U = R^{-1} W (cross covariances; skipped when Xcur is NULL),
 *   S = (R^{-1} N) R^{-T}.
 * Products with R^{-1} skip its zero lower blocks. The estimate is a
 * substitution with R (lanes of wave 1) beside the inverse (lanes of wave 0).
 */
template <int NB, bool DIAG = false>
__global__ void __launch_bounds__(NB * 8, (NB == 48 ? 3 : 1))
    k_backsub_mfma(uk_level_t lv, uk_factors_t fac, double* __restrict__ xout, double* __restrict__ Sout,
                   const double* __restrict__ Xup, const double* __restrict__ XTup, double* __restrict__ Xcur,
                   double* __restrict__ XTcur, int n) {
  constexpr int LD = NB + 1;
  constexpr int NT = NB * 8; /* NW = NB / 8 waves: one per 16-row tile of the 2 NB rows of W^T */
  constexpr int NW = NT / 64;
  constexpr int TB = NB / 16; /* tiles per block side */
  __shared__ BackMShared<NB> sh;
  double* const T = sh.blk[0]; /* Bt | Y: NB x 2 NB */
  double* const W = sh.blk[2]; /* NB x 2 NB */
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, li = lane & 15, lk = lane >> 4;
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
  const int64_t sj = lv.step_start[p] + j;
  const int64_t sup = lv.next_start[p] + (j >> 1) - 1;
  const double* Rg = fac.R + ev * nn;
  const double* Bg = fac.Bt + ev * nn;
  const double* Yg = fac.Y + ev * nn;

  /* Every global read of the kernel is issued here, so that one memory
   * latency covers them: this wave's rows of Sigma (the A fragments of
   * W^T = Sigma T^T, 16 consecutive doubles per fragment row) and this
   * thread's share of R (stored to LDS once T is dead). */
  constexpr int KS = NB / 2; /* 2 NB columns of Sigma in steps of 4 */
  double sig[KS];
  {
    const int rb = w / TB;             /* row block of Sigma: 0 (u_{j-1}) or 1 (u_{j+1}) */
    const int ri = (w % TB) * 16 + li; /* row within the block */
    const double* src0 = rb == 0 ? Sout + om * nn : XTup + sup * nn; /* S_mm | S_pm */
    const double* src1 = rb == 0 ? Xup + sup * nn : Sout + op * nn;  /* S_mp | S_pp */
    const bool live0 = rb == 0 ? has_m : (has_m && has_p);
    const bool live1 = rb == 0 ? (has_m && has_p) : has_p;
#pragma unroll
    for (int t = 0; t < KS; ++t) {
      const int cb = t * 4 / NB, ci = t * 4 + lk - cb * NB;
      sig[t] = 0.0;
      if ((cb == 0 ? live0 : live1) && ri < n && ci < n) sig[t] = (cb == 0 ? src0 : src1)[ri + (int64_t)ci * n];
    }
  }
  constexpr int RPER = (NB * NB + NT - 1) / NT;
  double rpre[RPER];
#pragma unroll
  for (int e = 0; e < RPER; ++e) {
    const int idx = tid + e * NT;
    const int i = idx % NB, c = idx / NB;
    rpre[e] = (idx < NB * NB && i < n && c < n) ? Rg[i + c * n] : (i == c ? 1.0 : 0.0);
  }
  for (int idx = tid; idx < NB * NB; idx += NT) {
    const int i = idx % NB, c = idx / NB;
    const bool in = i < n && c < n;
    T[i + c * LD] = (in && has_m) ? Bg[i + c * n] : 0.0;
    T[NB * LD + i + c * LD] = (in && has_p) ? Yg[i + c * n] : 0.0;
  }
  for (int i = tid; i < NB; i += NT) {
    sh.xm[i] = (has_m && i < n) ? xout[om * n + i] : 0.0;
    sh.xp[i] = (has_p && i < n) ? xout[op * n + i] : 0.0;
  }
  __syncthreads();
  if (tid < NB) {
    double r = 0.0;
    if (tid < n) {
      r = fac.cf[ev * n + tid];
      for (int q = 0; q < n; ++q) r -= T[tid + q * LD] * sh.xm[q] + T[NB * LD + tid + q * LD] * sh.xp[q];
    }
    sh.r[tid] = r;
  }

  /* ---- W^T = Sigma T^T: wave w owns rows [16w, 16w + 16) of W^T ---- */
  {
    dx4 acc[TB];
#pragma unroll
    for (int t = 0; t < TB; ++t) acc[t] = dx4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < KS; ++t) {
      const int k = t * 4 + lk;
#pragma unroll
      for (int tc = 0; tc < TB; ++tc) acc[tc] = mfma16(sig[t], T[k * LD + tc * 16 + li], acc[tc]);
    }
    __syncthreads(); /* sh.r read T; everyone past the T reads of this phase before W lands */
#pragma unroll
    for (int tc = 0; tc < TB; ++tc)
#pragma unroll
      for (int r = 0; r < 4; ++r) W[(w * 16 + 4 * r + lk) * LD + tc * 16 + li] = acc[tc][r];
  }
  __syncthreads();

  /* ---- N = I + W T^T, kept in registers until T is dead ---- */
  constexpr int NTILE = TB * TB;
  constexpr int PER = (NTILE + NW - 1) / NW;
  dx4 nacc[PER];
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    const int t = w + s * NW;
    nacc[s] = dx4{0.0, 0.0, 0.0, 0.0};
    if (t < NTILE) {
      const int ti = t % TB, tj = t / TB;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        nacc[s][r] = (ti == tj && 4 * r + lk == li && ti * 16 + li < n) ? 1.0 : 0.0;
#pragma unroll 8
      for (int ks = 0; ks < NB / 2; ++ks) {
        const int k = ks * 4 + lk;
        nacc[s] = mfma16(W[k * LD + ti * 16 + li], T[k * LD + tj * 16 + li], nacc[s]);
      }
    }
  }
  __syncthreads(); /* T dead */
  double* const Rs = sh.blk[0];
  double* const Ns = sh.blk[1];
#pragma unroll
  for (int e = 0; e < RPER; ++e) {
    const int idx = tid + e * NT;
    if (idx < NB * NB) Rs[idx % NB + (idx / NB) * LD] = rpre[e];
  }
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    const int t = w + s * NW;
    if (t < NTILE) {
      const int ti = t % TB, tj = t / TB;
#pragma unroll
      for (int r = 0; r < 4; ++r) Ns[(tj * 16 + li) * LD + ti * 16 + 4 * r + lk] = nacc[s][r];
    }
  }
  __syncthreads();
  if (tid < NB) sh.rinv[tid] = 1.0 / Rs[tid * LD + tid];
  __syncthreads();

  /* ---- R^{-1}, blockwise over 16 x 16 blocks; the estimate (wave 1) ----
   * Diagonal blocks: lane t < NB solves column t % 16 of block t / 16 by
   * substitution. Off-diagonal blocks, one superdiagonal d at a time (wave i
   * takes block (i, i + d)): R^{-1}_ij = -R^{-1}_ii sum_{k=i+1..j} R_ik R^{-1}_kj
   * on the matrix cores, the inner sum's D fragments serving as the outer
   * product's B operand. Off-diagonal results are kept transposed in the
   * (unused) lower blocks while R's upper blocks are still read, then moved. */
  {
    double v[16];
    const int bb = tid >> 4, bc = tid & 15;
    if (tid < NB) {
      const double* Rb = Rs + bb * 16 * (LD + 1);
      const double* rb = sh.rinv + bb * 16;
#pragma unroll
      for (int i = 0; i < 16; ++i) v[i] = (i == bc) ? 1.0 : 0.0;
#pragma unroll
      for (int pp = 15; pp >= 0; --pp) {
        const double xq = v[pp] * rb[pp];
        v[pp] = xq;
#pragma unroll
        for (int i = 0; i < pp; ++i) v[i] = fma(-Rb[i + pp * LD], xq, v[i]);
      }
    }
    if (w == 1) {
      double r = lane < n ? sh.r[lane] : 0.0;
      for (int q = n - 1; q >= 0; --q) {
        const double xq = __shfl(r, q) * sh.rinv[q];
        if (lane == q) r = xq;
        else if (lane < q) r -= Rs[lane + q * LD] * xq;
      }
      if (lane < n) xout[oj * n + lane] = r;
    }
    __syncthreads(); /* R's diagonal blocks consumed */
    if (tid < NB) {
      double* Rb = Rs + bb * 16 * (LD + 1);
#pragma unroll
      for (int i = 0; i < 16; ++i) Rb[i + bc * LD] = v[i];
    }
    __syncthreads();
    for (int d = 1; d < TB; ++d) {
      if (w < TB - d) {
        const int i0 = w * 16, j0 = (w + d) * 16;
        dx4 x = dx4{0.0, 0.0, 0.0, 0.0};
        for (int kb = w + 1; kb <= w + d; ++kb) {
          const int k0 = kb * 16;
#pragma unroll
          for (int r4 = 0; r4 < 4; ++r4) {
            const int kc = k0 + 4 * r4 + lk;
            const double a = Rs[(i0 + li) + kc * LD]; /* R_ik */
            const double b = (kb == w + d) ? Rs[kc + (j0 + li) * LD]      /* R^{-1}_jj (diagonal block) */
                                           : Rs[(j0 + li) + kc * LD];     /* R^{-1}_kj, stored transposed */
            x = mfma16(a, b, x);
          }
        }
        dx4 y = dx4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int r4 = 0; r4 < 4; ++r4) y = mfma16(-Rs[(i0 + li) + (i0 + 4 * r4 + lk) * LD], x[r4], y);
#pragma unroll
        for (int r = 0; r < 4; ++r) Rs[(j0 + li) + (i0 + 4 * r + lk) * LD] = y[r]; /* R^{-1}_ij transposed */
      }
      __syncthreads();
    }
    /* move the off-diagonal blocks to their upper positions */
    for (int idx = tid; idx < NB * NB; idx += NT) {
      const int a = idx % NB, c = idx / NB;
      if ((a >> 4) < (c >> 4)) Rs[a + c * LD] = Rs[c + a * LD];
    }
    __syncthreads();
  }
  const double* const Ri = Rs;

  /* ---- U = R^{-1} W -> cross covariances (wave w: tile column w of 2 TB) ---- */
  if (Xcur) {
    const int tc = w;
    const bool left = tc < TB;
    const bool want = left ? has_m : has_p;
    if (want) {
      double* dA = left ? XTcur + (sj - 1) * nn : Xcur + sj * nn; /* -U_1 -> XT[j-1]; -U_2 -> X[j] */
      double* dB = left ? Xcur + (sj - 1) * nn : XTcur + sj * nn; /* transposes */
      const int jc = (left ? tc : tc - TB) * 16 + li;             /* column within the n x n block */
#pragma unroll
      for (int ti = 0; ti < TB; ++ti) {
        dx4 acc = dx4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
        for (int ks = ti * 4; ks < NB / 4; ++ks) {
          const int k = ks * 4 + lk;
          acc = mfma16(Ri[k * LD + ti * 16 + li], W[(tc * 16 + li) * LD + k], acc);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = ti * 16 + 4 * r + lk;
          if (i < n && jc < n) {
            dA[i + (int64_t)jc * n] = -acc[r];
            dB[jc + (int64_t)i * n] = -acc[r];
          }
        }
      }
    }
  }

  /* ---- K = R^{-1} N (registers), then over W's first block ---- */
  dx4 kacc[PER];
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    const int t = w + s * NW;
    kacc[s] = dx4{0.0, 0.0, 0.0, 0.0};
    if (t < NTILE) {
      const int ti = t % TB, tj = t / TB;
#pragma unroll 4
      for (int ks = ti * 4; ks < NB / 4; ++ks) {
        const int k = ks * 4 + lk;
        kacc[s] = mfma16(Ri[k * LD + ti * 16 + li], Ns[(tj * 16 + li) * LD + k], kacc[s]);
      }
    }
  }
  __syncthreads(); /* W consumed by U */
  double* const Ks = sh.blk[2];
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    const int t = w + s * NW;
    if (t < NTILE) {
      const int ti = t % TB, tj = t / TB;
#pragma unroll
      for (int r = 0; r < 4; ++r) Ks[(tj * 16 + li) * LD + ti * 16 + 4 * r + lk] = kacc[s][r];
    }
  }
  __syncthreads();

  double* S = Sout + oj * nn;
  if constexpr (DIAG) {
    /* diagonal only (level 0 with UKHIP_COVARIANCE_DIAG): S_ii = sum_k K_ik R^{-1}_ik
     * over k >= i (R^{-1} upper triangular), written in its place in the slot */
    if (tid < n) {
      double d = 0.0;
      for (int k = tid; k < n; ++k) d = fma(Ks[k * LD + tid], Ri[k * LD + tid], d);
      S[tid + (int64_t)tid * n] = d;
    }
    return;
  }
  /* ---- S = K R^{-T}: B fragment R^{-T}[k][j] = R^{-1}[j][k], zero for k < j ---- */
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    const int t = w + s * NW;
    if (t < NTILE) {
      const int ti = t % TB, tj = t / TB;
      dx4 acc = dx4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
      for (int ks = tj * 4; ks < NB / 4; ++ks) {
        const int k = ks * 4 + lk;
        acc = mfma16(Ks[k * LD + ti * 16 + li], Ri[k * LD + tj * 16 + li], acc);
      }
      const int jc = tj * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = ti * 16 + 4 * r + lk;
        if (i < n && jc < n) S[jc + (int64_t)i * n] = acc[r]; /* S symmetric: S[i][jc] stored as its transpose */
      }
    }
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
 * 2n: c, [2n+1, 3n+1) columns of G, 3n+1: o. The whitened observation block
 * W G (mr rows, ld mr) and W o of step s go to out.C + s * cst and
 * out.o + s * ost: the chain's slots (cst = n * n, ost = n), or a staging
 * area that k_obs_compress reduces to n rows when mr > n.
 */
template <int NB>
__global__ void __launch_bounds__(((3 * NB + 2 + 63) / 64) * 64)
    k_weigh(uk_raw_t raw, uk_chain_t out, int64_t cst, int64_t ost) {
  __shared__ WeighShared<NB> sh;
  constexpr int NT = ((3 * NB + 2 + 63) / 64) * 64;
  const int tid = threadIdx.x;
  const int64_t s = blockIdx.x;
  const int n = raw.n, mr = raw.mr;
  const int64_t nn = (int64_t)n * n;
  const int mi = raw.m[s];
  const int m = mi < 0 ? 0 : (mi > mr ? mr : mi); /* counts from device-resident filter inputs are not host-checked */
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
      const double* src = raw.F + raw.Fidx[s] * nn + (int64_t)q * n;
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
    const double* src = q < n ? raw.F + raw.Fidx[s] * nn + (int64_t)q * n
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
    double* dst = is_o ? out.o + s * ost : out.C + s * cst + (int64_t)qq * mr;
#pragma unroll
    for (int i = 0; i < NB; ++i)
      if (i < mr) dst[i] = (i < m) ? x[i] : 0.0;
  }
}

/*
 * Whitening when every K_i and C_i is given as 'w' (inverse standard
 * deviations): V = diag(k_i), W = diag(w_i), so B = -k o F, D = k o H,
 * c = k o c, C = w o G, o = w o o (row scalings, c/kalman_base.c:61-175).
 * Element-parallel over spw consecutive steps per workgroup: consecutive
 * threads touch consecutive doubles of every array.
 */
/* w64 = 0 (F32 only): the FP64 chain is not written (its one reader, the
 * refinement's k_normal_w, forms it from the raw inputs: from_raw). */
template <bool F32>
__global__ void __launch_bounds__(256) k_weigh_diag(uk_raw_t raw, uk_chain_t out, uk_chain32_t out32, int spw,
                                                    int w64) {
  const int n = raw.n, mr = raw.mr, nn = n * n, mn = mr * n;
  const int64_t s0 = (int64_t)blockIdx.x * spw;
  const int cnt = (int)((raw.steps - s0) < spw ? (raw.steps - s0) : spw);
  const int tid = threadIdx.x;
  /* B formed where it is read and D stored as its diagonal: nothing is left
   * of the n x n pass, and D's n entries per step are written by the c pass
   * below, which reads the same k_i */
  const bool d_only = out.b_synth && out.d_diag;
  if (!d_only) {
    /* element e = c n + i of step ls; the indices advance by 256 elements
     * per iteration without a division */
    const int dls = 256 / nn, de = 256 - dls * nn, dc = de / n, di = de - dc * n;
    int ls = tid / nn, e = tid - ls * nn, c = e / n, i = e - c * n;
    for (int idx = tid; idx < cnt * nn; idx += 256) {
      const int64_t s = s0 + ls;
      if (!raw.first[s]) { /* no evolution equation on a chain's first step */
        const int64_t g = s * nn + e;
        const double kw = raw.K[s * n + i];
        if (!out.b_synth) { /* else B is formed where it is read */
          const double bv = -kw * raw.F[raw.Fidx[s] * nn + e];
          if (w64) out.B[g] = bv;
          if (F32) out32.B[g] = (float)bv;
        }
        if (out.d_diag) { /* D = diag(k): its diagonal in the slot's first n entries (then c == 0, i == e) */
          if (e < n) {
            if (w64) out.D[g] = kw;
            if (F32) out32.D[g] = (float)kw;
          }
        } else {
          const double dv = raw.H_ident ? ((c == i) ? kw : 0.0) : kw * raw.H[g];
          if (w64) out.D[g] = dv;
          if (F32) out32.D[g] = (float)dv;
        }
      }
      ls += dls;
      e += de;
      c += dc;
      i += di;
      if (i >= n) i -= n, ++c;
      if (e >= nn) e -= nn, ++ls, c -= n;
    }
  }
  /* C = W G: not stored in FP32 when the FP32 chain forms it on load (c_synth) */
  const bool w32c = F32 && !out32.c_synth;
  if (w64 || w32c) {
    const int dls = 256 / mn, de = 256 - dls * mn, di = de % mr;
    int ls = tid / mn, e = tid - ls * mn, i = e % mr;
    for (int idx = tid; idx < cnt * mn; idx += 256) {
      const int64_t s = s0 + ls;
      const double v = (i < raw.m[s]) ? raw.C[s * mr + i] * raw.G[s * mn + e] : 0.0;
      if (w64) out.C[s * nn + e] = v;
      if (w32c) out32.C[s * nn + e] = (float)v;
      ls += dls;
      e += de;
      i += di;
      if (i >= mr) i -= mr;
      if (e >= mn) e -= mn, ++ls;
    }
  }
  for (int idx = tid; idx < cnt * n; idx += 256) {
    const int64_t s = s0 + idx / n;
    if (!raw.first[s]) {
      const double kw = raw.K[s0 * n + idx];
      const double v = kw * raw.c[s0 * n + idx];
      if (w64) out.c[s0 * n + idx] = v;
      if (F32) out32.c[s0 * n + idx] = (float)v;
      if (d_only) { /* D = diag(k): the slot's first n entries */
        const int64_t g = s * nn + (idx - (s - s0) * n);
        if (w64) out.D[g] = kw;
        if (F32) out32.D[g] = (float)kw;
      }
    }
  }
  for (int idx = tid; idx < cnt * mr; idx += 256) {
    const int ls = idx / mr, i = idx - ls * mr;
    const int64_t s = s0 + ls;
    const double v = (i < raw.m[s]) ? raw.C[s * mr + i] * raw.o[s * mr + i] : 0.0;
    if (w64) out.o[s * n + i] = v;
    if (F32) out32.o[s * n + i] = (float)v;
  }
}

/*
 * Observation blocks taller than the state (mr > n rows after whitening,
 * at most 2 NB): one step per workgroup, QR of [W G_s | W o_s] (mr x (n+1),
 * staged by k_weigh at wC + s * mr * n, ld mr, and wo + s * mr); R (n x n,
 * upper triangular) and the first n entries of Q^T W o become the chain's C
 * (ld n) and o. The least-squares problem is unchanged: the dropped rows of
 * Q^T [W G | W o] are zero in G's columns, so they add a constant to the
 * residual and nothing to the normal equations. This lets the elimination
 * keep one n-row observation slot per step, where the reference stacks the
 * m_i rows into its phase-1 and phase-3 blocks
 * (c/kalman_oddeven_smoother.c:422-440, :535-546).
 */
template <int NB>
__global__ void __launch_bounds__((QRGeom<NB, NB + 1, 3>::NT))
    k_obs_compress(const double* __restrict__ wC, const double* __restrict__ wo, int mr, int n,
                   double* __restrict__ C, double* __restrict__ o) {
  using G = QRGeom<NB, NB + 1, 3>;
  constexpr int SL = G::SL;
  constexpr int NT = G::NT;
  __shared__ Refl<NB, SL, double> refl;
  const int tid = threadIdx.x, sub = tid & (SL - 1), c = tid / SL;
  const int64_t s = blockIdx.x;
  const int64_t nn = (int64_t)n * n;
  const int hi = mr > NB ? mr - NB : 0; /* rows in block 1 */
  refl_clear<NB, SL, NT, double>(refl, tid);
  Col<NB, SL, double> col;
  const double* src = c < n ? wC + s * (int64_t)mr * n + (int64_t)c * mr : wo + s * mr;
  if (c <= n) {
    col.load_block(0, src, mr < NB ? mr : NB, sub);
    col.load_block(1, src + NB, hi, sub);
  } else {
    col.zero_block(0);
    col.zero_block(1);
  }
  house_qr<NB, SL, double>(col, refl, c, sub, 0, n, c <= n);
  if (c < n) col.store_triu(C + s * nn + (int64_t)c * n, n, c, sub);
  if (c == n) col.store_block(0, o + s * n, n, sub);
}

/*
 * Covariance output formats, one step per workgroup (64 threads):
 *   fmt 1 ('D'): the diagonal, n doubles per step;
 *   fmt 2 ('W'): W = L^{-1} for the Cholesky factor S = L L^T, so that
 *   W^T W = S^{-1}: the reference's 'W' representation of a covariance
 *   (c/kalman_base.c:61-175), lower triangular, n*n per step, column-major.
 */
template <int NB>
__global__ void __launch_bounds__(64) k_cov_format(const double* __restrict__ S, int n, int fmt,
                                                   double* __restrict__ dst) {
  __shared__ double L[NB * NB];
  const int tid = threadIdx.x;
  const int64_t s = blockIdx.x;
  const int64_t nn = (int64_t)n * n;
  const double* Ss = S + s * nn;
  if (fmt == 1) {
    if (tid < n) dst[s * n + tid] = Ss[tid * (n + 1)];
    return;
  }
  for (int i = tid; i < n * n; i += 64) L[i] = Ss[i];
  __syncthreads();
  lds_cholesky<64>(L, n, tid); /* lower factor in place, ld n (ends with a barrier) */
  if (tid < n) {
    const int j = tid;
    double x[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      double v = 0.0;
      if (i < n && i >= j) {
        double acc = (i == j) ? 1.0 : 0.0;
#pragma unroll
        for (int q = 0; q < i; ++q)
          if (q >= j) acc = fma(-L[i + q * n], x[q], acc);
        v = acc / L[i + i * n];
      }
      x[i] = v;
    }
    double* W = dst + s * nn + (int64_t)j * n;
#pragma unroll
    for (int i = 0; i < NB; ++i)
      if (i < n) W[i] = x[i];
  }
}

/*
 * One step (evolve + observe) of every filter of a batch of sequential
 * UltimateKalman filters, one workgroup per filter (c/kalman_ultimate.c:140-470):
 *   evolve:  QR of [R_{i-1}; -V F_i] applied to [0; V H_i] and [y_{i-1}; V c_i];
 *            rows n..2n-1 of the last two are the prediction (R~, y~);
 *   observe: QR of [R~; W G_i] applied to [y~; W o_i] -> (R_i, y_i);
 *   x_i = R_i^{-1} y_i once the state is determined (z_i = n), NaN before.
 * Zero rows stand for the reference's not-yet-square ("flat") factors: they
 * stay zero under the reflectors. Columns (SL lanes each): [0,n) u_{i-1},
 * [n,2n) u_i, 2n right-hand side. V = diag(K), W = diag(C) ('w' covariances).
 */
template <int NB>
struct FilterShared {
  double R[NB * NB];
  double y[NB];
  double rinv[NB];
};

template <int NB>
__global__ void __launch_bounds__((QRGeom<NB, 2 * NB + 1, 1>::NT))
    k_filter_step(uk_filter_t st, uk_filter_in_t in, int n, int first) {
  constexpr int SL = QRGeom<NB, 2 * NB + 1, 1>::SL;
  constexpr int NT = QRGeom<NB, 2 * NB + 1, 1>::NT;
  __shared__ Refl<NB, SL, double> refl;
  __shared__ FilterShared<NB> fs;
  const int tid = threadIdx.x, sub = tid & (SL - 1), c = tid / SL;
  const int64_t f = blockIdx.x;
  const int64_t nn = (int64_t)n * n;
  const int mr = in.mr;
  const bool isP = c < n, isB = c >= n && c < 2 * n, isY = c == 2 * n;
  const int bc = isB ? c - n : 0;
  const double* kw = in.K + f * n;
  const int zprev = first ? 0 : st.z[f];
  refl_clear<NB, SL, NT, double>(refl, tid);
  Col<NB, SL, double> col;
  col.zero_block(0);
  col.zero_block(1);
  if (!first) {
    if (isP) {
      col.load_block(0, st.R + f * nn + (int64_t)c * n, n, sub);
      if (in.pre) {
        col.load_block(1, in.wB + f * nn + (int64_t)c * n, n, sub); /* -V F */
      } else {
        col.load_block(1, in.F + (in.F_shared ? 0 : f * nn) + (int64_t)c * n, n, sub);
        col.scale_block(1, kw, n, sub, -1.0);
      }
    }
    if (isB) {
      if (in.pre) {
        col.load_block(1, in.wD + f * nn + (int64_t)bc * n, n, sub); /* V H */
      } else if (in.H) {
        col.load_block(1, in.H + f * nn + (int64_t)bc * n, n, sub);
        col.scale_block(1, kw, n, sub, 1.0);
      } else {
#pragma unroll
        for (int l = 0; l < NB / SL; ++l) {
          const int r = SL * l + sub;
          col.a[NB / SL + l] = (r == bc) ? kw[r] : 0.0;
        }
      }
    }
    if (isY) {
      col.load_block(0, st.y + f * n, n, sub);
      if (in.pre) {
        col.load_block(1, in.wc + f * n, n, sub); /* V c */
      } else if (in.c) {
        col.load_block(1, in.c + f * n, n, sub);
        col.scale_block(1, kw, n, sub, 1.0);
      }
    }
    house_qr<NB, SL, double>(col, refl, c, sub, 0, n, isP || isB || isY);
    if (st.hist) { /* the previous step's final rows: Rdiag, Rsupdiag, y (c/kalman_ultimate.c:140-260) */
      const int64_t hs = st.hslot * st.nf + f;
      if (isP) col.store_triu(st.hRd + hs * nn + (int64_t)c * n, n, c, sub);
      if (isB) col.store_block(0, st.hRs + hs * nn + (int64_t)bc * n, n, sub);
      if (isY) col.store_block(0, st.hy + hs * n, n, sub);
    }
    if (isB || isY) col.move_1_to_0();
  }
  const int mi = in.m_i ? in.m_i[f] : mr;
  const int m = mi < 0 ? 0 : (mi > mr ? mr : mi); /* device-resident counts are not checked on the host */
  if (m > 0) {
    if (in.pre) {
      if (isB) col.load_block(1, in.wC + f * nn + (int64_t)bc * mr, m, sub); /* W G, ld mr in an n*n slot */
      if (isY) col.load_block(1, in.wo + f * n, m, sub);                      /* W o */
    } else {
      const double* cw = in.C + f * mr;
      if (isB) {
        col.load_block(1, in.G + f * mr * n + (int64_t)bc * mr, m, sub);
        col.scale_block(1, cw, m, sub, 1.0);
      }
      if (isY) {
        col.load_block(1, in.o + f * mr, m, sub);
        col.scale_block(1, cw, m, sub, 1.0);
      }
    }
  }
  house_qr<NB, SL, double>(col, refl, c, sub, n, n, isB || isY);
  if (isB) {
    col.store_triu(st.R + f * nn + (int64_t)bc * n, n, bc, sub);
#pragma unroll
    for (int l = 0; l < NB / SL; ++l) {
      const int r = SL * l + sub;
      fs.R[r + bc * NB] = (r <= bc && r < n) ? col.a[l] : (r == bc ? 1.0 : 0.0);
    }
  }
  if (isY) {
    col.store_block(0, st.y + f * n, n, sub);
#pragma unroll
    for (int l = 0; l < NB / SL; ++l) {
      const int r = SL * l + sub;
      fs.y[r] = r < n ? col.a[l] : 0.0;
    }
  }
  const int z = min(n, zprev + m);
  __syncthreads();
  if (tid < n) fs.rinv[tid] = 1.0 / fs.R[tid + tid * NB];
  __syncthreads();
  if (tid < 64) {
    double r = tid < n ? fs.y[tid] : 0.0;
    if (z == n) {
      for (int q = n - 1; q >= 0; --q) {
        const double xq = __shfl(r, q) * fs.rinv[q];
        if (tid == q) r = xq;
        else if (tid < q) r -= fs.R[tid + q * NB] * xq;
      }
    } else {
      r = __builtin_nan("");
    }
    if (tid < n) st.x[f * n + tid] = r;
  }
  if (tid == 0) st.z[f] = z;
}

/*
 * The smoother pass of the batched UltimateKalman filters over the kept
 * history, one workgroup per filter, steps last .. 0 (c/kalman_ultimate.c:
 * 559-620): x_last = the filtered estimate, x_s = Rdiag_s^{-1} (y_s -
 * Rsupdiag_s x_{s+1}); the covariance factor ('W': covariance = (W^T W)^{-1})
 * W_last = R_last, and W_s = the last n rows of Q^T [Rdiag_s; 0] for the QR
 * [Rsupdiag_s; W_{s+1}] = Q R. Columns (SL lanes each): [0, n) the pivots
 * [Rsupdiag_s; W_{s+1}], [n, 2n) [Rdiag_s; 0]. xs and Ws: slot s of filter f
 * at (s * nf + f) * (n | n*n).
 */
template <int NB>
struct SmoothShared {
  double Wn[NB * NB]; /* W_{s+1}, column-major, ld NB */
  double Rd[NB * NB];
  double rs[NB];      /* y_s - Rsupdiag_s x_{s+1} */
  double xn[NB];      /* x_{s+1} */
  double rinv[NB];
};

template <int NB>
__global__ void __launch_bounds__((QRGeom<NB, 2 * NB, 1>::NT))
    k_filter_smooth(uk_filter_t st, int n, int64_t steps, double* __restrict__ xs, double* __restrict__ Ws) {
  constexpr int SL = QRGeom<NB, 2 * NB, 1>::SL;
  constexpr int NT = QRGeom<NB, 2 * NB, 1>::NT;
  __shared__ Refl<NB, SL, double> refl;
  __shared__ SmoothShared<NB> sh;
  const int tid = threadIdx.x, sub = tid & (SL - 1), c = tid / SL;
  const int64_t f = blockIdx.x, nf = st.nf;
  const int64_t nn = (int64_t)n * n;
  const bool isP = c < n, isS = c >= n && c < 2 * n;
  const int sc = isS ? c - n : 0;
  /* the last step: the filter's own factor and estimate */
  {
    const int64_t o = (steps - 1) * nf + f;
    for (int i = tid; i < NB * NB; i += NT) {
      const int r = i % NB, q = i / NB;
      sh.Wn[i] = (r < n && q < n) ? st.R[f * nn + (int64_t)q * n + r] : 0.0;
    }
    for (int i = tid; i < n; i += NT) sh.xn[i] = xs[o * n + i] = st.x[f * n + i];
    for (int i = tid; i < nn; i += NT) Ws[o * nn + i] = st.R[f * nn + i];
  }
  __syncthreads();
  for (int64_t s = steps - 2; s >= 0; --s) {
    const int64_t hs = s * nf + f;
    const double* Rd = st.hRd + hs * nn;
    const double* Rs = st.hRs + hs * nn;
    /* estimate: rs = y - Rsupdiag x_{s+1} (row per thread), then the triangular solve on wave 0 */
    for (int i = tid; i < NB * NB; i += NT) {
      const int r = i % NB, q = i / NB;
      sh.Rd[i] = (r < n && q < n) ? Rd[(int64_t)q * n + r] : (r == q ? 1.0 : 0.0);
    }
    if (tid < n) {
      double v = st.hy[hs * n + tid];
      for (int j = 0; j < n; ++j) v = fma(-Rs[(int64_t)j * n + tid], sh.xn[j], v);
      sh.rs[tid] = v;
    }
    __syncthreads();
    if (tid < n) sh.rinv[tid] = 1.0 / sh.Rd[tid + tid * NB];
    __syncthreads();
    if (tid < 64) {
      double r = tid < n ? sh.rs[tid] : 0.0;
      for (int q = n - 1; q >= 0; --q) {
        const double xq = __shfl(r, q) * sh.rinv[q];
        if (tid == q) r = xq;
        else if (tid < q) r -= sh.Rd[tid + q * NB] * xq;
      }
      if (tid < n) xs[hs * n + tid] = r;
      if (tid < n) sh.rs[tid] = r; /* x_s, the next iteration's x_{s+1} */
    }
    /* covariance factor: QR of [Rsupdiag_s; W_{s+1}] applied to [Rdiag_s; 0] */
    Col<NB, SL, double> col;
    col.zero_block(0);
    col.zero_block(1);
    if (isP) {
      col.load_block(0, Rs + (int64_t)c * n, n, sub);
#pragma unroll
      for (int l = 0; l < NB / SL; ++l) {
        const int r = SL * l + sub;
        col.a[NB / SL + l] = r < n ? sh.Wn[r + c * NB] : 0.0;
      }
    }
    if (isS) col.load_block(0, Rd + (int64_t)sc * n, n, sub);
    house_qr<NB, SL, 1, double, QF_DONE>(reinterpret_cast<Col<NB, SL, double>(&)[1]>(col), refl, c, sub, 0, n,
                                        (isP || isS) ? 1u : 0u);
    /* house_qr ends with a barrier: every read of Wn and xn above is done */
    if (isS) {
      col.store_block(1, Ws + hs * nn + (int64_t)sc * n, n, sub);
#pragma unroll
      for (int l = 0; l < NB / SL; ++l) {
        const int r = SL * l + sub;
        if (r < n) sh.Wn[r + sc * NB] = col.a[NB / SL + l];
      }
    }
    if (tid < n) sh.xn[tid] = sh.rs[tid];
    __syncthreads();
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
/* The elimination's and the refinement sweeps' buckets: uk_bucket's, and 24
 * for 16 < n <= 24 (three quarters of bucket 32's rows and lanes). The other
 * kernels keep uk_bucket's; every bucket serves any n up to its size on the
 * same layouts (blocks stored n x n). */
extern "C" int uk_bucket_qr(int n) {
  if (n > 16 && n <= 24) return 24;
  return uk_bucket(n);
}

/* Steps per workgroup of k_weigh_diag: about 2048 elements of work per
 * 256-thread workgroup (the n x n pass is skipped when B is formed on load and
 * D is stored as its diagonal). */
static int weigh_diag_spw(const uk_raw_t* raw, const uk_chain_t* out, bool c_pass = true) {
  const int n = raw->n, mr = raw->mr;
  const int per = (out->b_synth && out->d_diag ? 0 : n * n) + (c_pass ? mr * n : 0) + n + mr;
  const int spw = 2048 / (per > 0 ? per : 1);
  return spw < 1 ? 1 : (spw > 64 ? 64 : spw);
}

extern "C" hipError_t uk_launch_weigh(const uk_raw_t* raw, const uk_chain_t* out, hipStream_t st) {
  if (!raw->K_full && !raw->C_full) { /* every covariance given as 'w' */
    const int spw = weigh_diag_spw(raw, out);
    const int64_t blocks = (raw->steps + spw - 1) / spw;
    k_weigh_diag<false><<<dim3((unsigned)blocks), dim3(256), 0, st>>>(*raw, *out, uk_chain32_t{}, spw, 1);
    return hipGetLastError();
  }
  const int nb = uk_bucket(raw->n > raw->mr ? raw->n : raw->mr);
  const int64_t nn = (int64_t)raw->n * raw->n;
#define L_(NBV)                                                                                            \
  k_weigh<NBV><<<dim3((unsigned)raw->steps), dim3(((3 * NBV + 2 + 63) / 64) * 64), 0, st>>>(*raw, *out, nn, \
                                                                                             (int64_t)raw->n)
  switch (nb) {
    UK_DISPATCH(8, L_) UK_DISPATCH(16, L_) UK_DISPATCH(32, L_) UK_DISPATCH(48, L_) UK_DISPATCH(64, L_)
    default: return hipErrorInvalidValue;
  }
#undef L_
  return hipGetLastError();
}

/* Whitening when observation blocks have more rows than the state (raw->mr >
 * n): k_weigh writes the evolution part into the chain and the whitened
 * observations to wobs (mr * n + mr doubles per step), then k_obs_compress
 * reduces every step's observation block to n rows in the chain's C (ld n)
 * and o. raw->mr must not exceed 2 * uk_bucket(n) or 64. */
extern "C" hipError_t uk_launch_weigh_obs(const uk_raw_t* raw, const uk_chain_t* out, double* wobs, hipStream_t st) {
  const int n = raw->n, mr = raw->mr;
  const int nbn = uk_bucket(n), nb = uk_bucket(n > mr ? n : mr);
  if (mr <= n || nbn < 0 || nb < 0 || mr > 2 * nbn) return hipErrorInvalidValue;
  uk_chain_t w = *out;
  w.C = wobs;
  w.o = wobs + raw->steps * (int64_t)mr * n;
  const unsigned g = (unsigned)raw->steps;
#define L_(NBV) k_weigh<NBV><<<dim3(g), dim3(((3 * NBV + 2 + 63) / 64) * 64), 0, st>>>(*raw, w, (int64_t)mr * n, (int64_t)mr)
  switch (nb) {
    UK_DISPATCH(8, L_) UK_DISPATCH(16, L_) UK_DISPATCH(32, L_) UK_DISPATCH(48, L_) UK_DISPATCH(64, L_)
    default: return hipErrorInvalidValue;
  }
#undef L_
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
#define L_(NBV) k_obs_compress<NBV><<<dim3(g), dim3(QRGeom<NBV, NBV + 1, 3>::NT), 0, st>>>(w.C, w.o, mr, n, out->C, out->o)
  switch (nbn) {
    UK_DISPATCH(8, L_) UK_DISPATCH(16, L_) UK_DISPATCH(32, L_) UK_DISPATCH(48, L_) UK_DISPATCH(64, L_)
    default: return hipErrorInvalidValue;
  }
#undef L_
  return hipGetLastError();
}

/* The level-0 chain's FP32 copy (the mixed-precision path; its C only when
 * out32->c_synth is 0) and, written by the same pass when w64 is set, the FP64
 * chain, when every K and C is given as 'w'; returns hipErrorNotSupported
 * otherwise (then the FP64 chain is converted afterwards). */
extern "C" hipError_t uk_launch_weigh_f32(const uk_raw_t* raw, const uk_chain_t* out, const uk_chain32_t* out32,
                                          int w64, hipStream_t st) {
  if (raw->K_full || raw->C_full) return hipErrorNotSupported;
  const int spw = weigh_diag_spw(raw, out, w64 || !out32->c_synth);
  const int64_t blocks = (raw->steps + spw - 1) / spw;
  k_weigh_diag<true><<<dim3((unsigned)blocks), dim3(256), 0, st>>>(*raw, *out, *out32, spw, w64);
  return hipGetLastError();
}

/* Run-time launch table of the phase kernels at NB >= 24 (FP64, and the FP32
 * elimination of the mixed-precision path), per precision, phase and bucket
 * (32, 48, 64, 24): the columns each lane group holds (cpl) and the workgroups per
 * CU asked of the register allocator (wg). Starts from variants measured
 * fastest on MI355X (n = 24, 32, 48 and 64, k = 200000), rewritten by
 * ukhip_autotune. */
struct QrLaunch {
  int cpl, wg;
};
static QrLaunch g_qr[2][3][4] = {
    /* FP64:  bucket 32   48        64        24 */
    {/* phase 1 */ {{1, 2}, {2, 4}, {1, 2}, {1, 2}},
     /* phase 2 */ {{2, 4}, {3, 2}, {3, 2}, {1, 1}},
     /* phase 3 */ {{1, 4}, {1, 4}, {2, 3}, {1, 2}}},
    /* FP32 */
    {/* phase 1 */ {{2, 3}, {2, 4}, {2, 4}, {1, 4}},
     /* phase 2 */ {{2, 3}, {3, 2}, {3, 3}, {2, 4}},
     /* phase 3 */ {{1, 2}, {1, 1}, {2, 3}, {1, 1}}}};
static int qr_bucket_index(int nb) {
  return nb == 32 ? 0 : (nb == 48 ? 1 : (nb == 64 ? 2 : (nb == 24 ? 3 : -1)));
}
/* the compiled variants: one column per group at 1, 2 or 4 workgroups per CU,
 * two columns per group at 2, 3 or 4, three columns per group at 2 or 3 */
/* cpl 0: phase 2 in blocked WY form on the matrix cores (k_phase2_wy), one
 * workgroup per CU asked of the register allocator.
 * Left out: the FP32 phase 2 of bucket 64 with two columns per group at 3
 * workgroups per CU. On MI355X it returned non-finite factors for an
 * ill-conditioned n = 64 problem (F = Q diag(logspace(0, -3, 64)), K = 1e-2 I)
 * where every other variant agrees bit for bit; it is the one variant whose
 * register budget (80 VGPRs, spilling) differs, and no fault in the source
 * explains it. */
static bool qr_variant(int fp32, int nb, int cpl, int wg, int phase) {
  if (fp32 && nb == 64 && phase == 2 && cpl == 2 && wg == 3) return false;
  return (cpl == 1 && (wg == 1 || wg == 2 || wg == 4)) || (cpl == 2 && (wg == 2 || wg == 3 || wg == 4)) ||
         (cpl == 3 && (wg == 2 || wg == 3)) || (cpl == 0 && wg == 1 && phase == 2 && nb % 16 == 0);
}

/* fp32: 0 the FP64 kernels, 1 the FP32 elimination of the mixed-precision path */
extern "C" int ukhip_qr_launch_get(int fp32, int phase, int n, int* columns_per_group) {
  const int bi = qr_bucket_index(uk_bucket_qr(n));
  if (phase < 1 || phase > 3 || bi < 0) return 0;
  const QrLaunch v = g_qr[fp32 ? 1 : 0][phase - 1][bi];
  if (columns_per_group) *columns_per_group = v.cpl;
  return v.wg;
}
extern "C" int ukhip_qr_launch_set(int fp32, int phase, int n, int columns_per_group, int wg) {
  const int bi = qr_bucket_index(uk_bucket_qr(n));
  if (phase < 1 || phase > 3 || bi < 0 || !qr_variant(fp32, uk_bucket_qr(n), columns_per_group, wg, phase)) return -1;
  g_qr[fp32 ? 1 : 0][phase - 1][bi] = QrLaunch{columns_per_group, wg};
  return 0;
}

namespace {
template <int NB, class T>
void phase_launch(const uk_level_t* lv, const ChainT<T>& in, const ChainT<T>& out, const FactorsT<T>& fac, T* fix,
                  int n, hipStream_t st, int phase) {
  if constexpr (NB >= 24) {
    const QrLaunch q = g_qr[sizeof(T) == 8 ? 0 : 1][phase - 1][qr_bucket_index(NB)];
    const ukqr::Launch v{q.cpl, q.wg};
    if constexpr (NB == 24) ukqr::phase_launch_b24(lv, in, out, fac, fix, n, st, phase, v);
    else if constexpr (NB == 32) ukqr::phase_launch_b32(lv, in, out, fac, fix, n, st, phase, v);
    else if constexpr (NB == 48) ukqr::phase_launch_b48(lv, in, out, fac, fix, n, st, phase, v);
    else ukqr::phase_launch_b64(lv, in, out, fac, fix, n, st, phase, v);
  } else {
    phase_launch_impl<NB, T>(lv, in, out, fac, fix, n, st, phase, ukqr::Launch{1, 0});
  }
}

template <class T>
hipError_t launch_eliminate(const uk_level_t* lv, const ChainT<T>& in, const ChainT<T>& out, const FactorsT<T>& fac,
                            T* fix, int n, hipStream_t st, hipEvent_t* ev) {
  const int nb = uk_bucket_qr(n > in.mr ? n : in.mr);
  auto mark = [&](int i) {
    if (ev) (void)hipEventRecord(ev[i], st);
  };
#define L_(NBV)                                                   \
  mark(0);                                                        \
  for (int ph = 1; ph <= 3; ++ph) {                               \
    phase_launch<NBV, T>(lv, in, out, fac, fix, n, st, ph);       \
    mark(ph);                                                     \
  }
  switch (nb) {
    UK_DISPATCH(8, L_) UK_DISPATCH(16, L_) UK_DISPATCH(24, L_) UK_DISPATCH(32, L_) UK_DISPATCH(48, L_)
    UK_DISPATCH(64, L_)
    default: return hipErrorInvalidValue;
  }
#undef L_
  return hipGetLastError();
}
}  // namespace

/* ev (NULL or 4 events) is recorded before phase 1 and after each phase. */
extern "C" hipError_t uk_launch_eliminate(const uk_level_t* lv, const uk_chain_t* in, const uk_chain_t* out,
                                          const uk_factors_t* fac, double* fix, int n, hipStream_t st,
                                          hipEvent_t* ev) {
  const ChainT<double> ci{in->B,      in->D,  in->c,  in->C,     in->o,   in->mr,  in->d_diag, in->c_tri,
                          in->b_synth, in->wk, in->wF, in->wFidx, 0,       nullptr, nullptr,    nullptr};
  const ChainT<double> co{out->B,      out->D,  out->c,  out->C,     out->o,  out->mr, out->d_diag, out->c_tri,
                          out->b_synth, out->wk, out->wF, out->wFidx, 0,       nullptr, nullptr,     nullptr};
  const FactorsT<double> f{fac->R, fac->Bt, fac->Y, fac->cf};
  return launch_eliminate<double>(lv, ci, co, f, fix, n, st, ev);
}

/* The same elimination in single precision (chains, factors and arithmetic). */
extern "C" hipError_t uk_launch_eliminate32(const uk_level_t* lv, const uk_chain32_t* in, const uk_chain32_t* out,
                                            const uk_factors32_t* fac, float* fix, int n, hipStream_t st,
                                            hipEvent_t* ev) {
  const ChainT<float> ci{in->B,      in->D,  in->c,  in->C,     in->o,      in->mr,  in->d_diag, in->c_tri,
                         in->b_synth, in->wk, in->wF, in->wFidx, in->c_synth, in->rG, in->rC,     in->rm};
  const ChainT<float> co{out->B,      out->D,  out->c,  out->C,     out->o,      out->mr, out->d_diag, out->c_tri,
                         out->b_synth, out->wk, out->wF, out->wFidx, out->c_synth, out->rG, out->rC,     out->rm};
  const FactorsT<float> f{fac->R, fac->Bt, fac->Y, fac->cf};
  return launch_eliminate<float>(lv, ci, co, f, fix, n, st, ev);
}

/* Back substitution with covariances for n > 16 on the matrix cores (default)
 * or on the vector units (k_backsub, the kernel used for n <= 16). cov: 0
 * estimates only, 1 covariances, 2 only the covariances' diagonals (written
 * in their places; for level 0, whose covariances no other level reads). */
static int g_backsub_mfma = 1;
static bool uk_backsub_mfma_enabled() { return g_backsub_mfma != 0; }
extern "C" void ukhip_use_mfma(int on) { g_backsub_mfma = on ? 1 : 0; }

extern "C" hipError_t uk_launch_backsub(const uk_level_t* lv, const uk_factors_t* fac, double* xout, double* Sout,
                                        const double* Xup, const double* XTup, double* Xcur, double* XTcur, int n,
                                        int cov, hipStream_t st) {
  const int nb = uk_bucket(n);
  if (!cov && nb >= 32) {
#define E_(NBV) k_estimate<NBV><<<dim3((unsigned)lv->evens), dim3(64), 0, st>>>(*lv, *fac, xout, n)
    switch (nb) {
      case 32: E_(32); break;
      case 48: E_(48); break;
      case 64: E_(64); break;
      default: return hipErrorInvalidValue;
    }
#undef E_
    return hipGetLastError();
  }
  if (cov && nb >= 32 && uk_backsub_mfma_enabled()) {
#define M_(NBV)                                                                                                  \
  if (cov == 2)                                                                                                  \
    k_backsub_mfma<NBV, true><<<dim3((unsigned)lv->evens), dim3(NBV * 8), 0, st>>>(*lv, *fac, xout, Sout, Xup,    \
                                                                                 XTup, Xcur, XTcur, n);          \
  else                                                                                                           \
    k_backsub_mfma<NBV, false><<<dim3((unsigned)lv->evens), dim3(NBV * 8), 0, st>>>(*lv, *fac, xout, Sout, Xup,   \
                                                                                  XTup, Xcur, XTcur, n)
    switch (nb) {
      case 32: M_(32); break;
      case 48: M_(48); break;
      case 64: M_(64); break;
      default: return hipErrorInvalidValue;
    }
#undef M_
    return hipGetLastError();
  }
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

/* Covariance output formats of `steps` consecutive n x n covariances S:
 * fmt 1 diagonal (n per step), 2 inverse Cholesky factor W (n*n per step). */
extern "C" hipError_t uk_launch_cov_format(const double* S, int64_t steps, int n, int fmt, double* dst,
                                           hipStream_t st) {
  if (steps <= 0) return hipSuccess;
  const int nb = uk_bucket(n);
#define C_(NBV) k_cov_format<NBV><<<dim3((unsigned)steps), dim3(64), 0, st>>>(S, n, fmt, dst)
  switch (nb) {
    UK_DISPATCH(8, C_) UK_DISPATCH(16, C_) UK_DISPATCH(32, C_) UK_DISPATCH(48, C_) UK_DISPATCH(64, C_)
    default: return hipErrorInvalidValue;
  }
#undef C_
  return hipGetLastError();
}

/* The smoother pass over every filter's history (k_filter_smooth). */
extern "C" hipError_t uk_launch_filter_smooth(const uk_filter_t* st, int n, int64_t steps, double* xs, double* Ws,
                                              hipStream_t s) {
  const int nb = uk_bucket(n);
#define S_(NBV) k_filter_smooth<NBV><<<dim3((unsigned)st->nf), dim3(QRGeom<NBV, 2 * NBV, 1>::NT), 0, s>>>(*st, n, steps, xs, Ws)
  switch (nb) {
    UK_DISPATCH(8, S_) UK_DISPATCH(16, S_) UK_DISPATCH(32, S_) UK_DISPATCH(48, S_) UK_DISPATCH(64, S_)
    default: return hipErrorInvalidValue;
  }
#undef S_
  return hipGetLastError();
}

/* One step of every filter of a batch (k_filter_step). */
extern "C" hipError_t uk_launch_filter_step(const uk_filter_t* st, const uk_filter_in_t* in, int nf, int n, int first,
                                            hipStream_t s) {
  const int nb = uk_bucket(n > in->mr ? n : in->mr);
#define F_(NBV) \
  k_filter_step<NBV><<<dim3((unsigned)nf), dim3(QRGeom<NBV, 2 * NBV + 1, 1>::NT), 0, s>>>(*st, *in, n, first)
  switch (nb) {
    UK_DISPATCH(8, F_) UK_DISPATCH(16, F_) UK_DISPATCH(32, F_) UK_DISPATCH(48, F_) UK_DISPATCH(64, F_)
    default: return hipErrorInvalidValue;
  }
#undef F_
  return hipGetLastError();
}
