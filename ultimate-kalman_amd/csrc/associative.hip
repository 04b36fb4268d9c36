/*
 * associative.hip -- gfx950 kernels of the associative-scan smoother
 * (Saerkkae and Garcia-Fernandez), reference c/kalman_associative_smoother.c:
 * filtering elements :309-505, smoothing elements :697-753, combines
 * :889-1016, driver :1199-1240; matlab/KalmanAssociativeSmoother.m:20-262.
 *
 * Pipeline over a batch (all problems of one state dimension at once):
 *   k_weigh (shared with odd-even) -> whitened B = -V F, D = V H, c, C = W G, o
 *   ka_model     per step: F = -D^{-1} B, c = D^{-1} c_w, Q = (D^T D)^{-1}
 *                (the covariance-form model with H folded in: H^{-1} F,
 *                H^{-1} c, H^{-1} K H^{-T})
 *   ka_prior     per problem: information filter over steps 0..p, p the first
 *                step whose observations can determine the state (sum of m_t
 *                over t <= p reaches n), giving the filtered (x_p, P_p); the
 *                reference needs p = 0
 *   ka_felem     filtering elements (A, b, Z, e, J) of steps p+1 .. k-1
 *                (unobserved steps: e = 0, J = 0)
 *   filtering scan: chunk reduce, Hillis-Steele rounds over chunk totals,
 *                chunk down-sweep writing every prefix in place
 *   ka_selem     smoothing elements (E, g, L); steps before p from the
 *                information form
 *   smoothing scan (same scheme, last step first) -> estimates, covariances
 *
 * Every element build and every combine runs in one 256-thread workgroup
 * with its blocks in LDS (the filtering work blocks of the 64 bucket go to a
 * per-workgroup global slice: 6 blocks of 64 x 64 exceed the 160 KiB LDS).
 * Observations are used whitened (noise covariance I), which is the
 * reference's S = G K G^T + C after the congruence with the whitening W.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "ukhip_internal.h"

extern "C" void uk_set_error(const char* fmt, ...);
extern "C" int uk_bucket(int n);
extern "C" hipError_t uk_launch_weigh(const uk_raw_t* raw, const uk_chain_t* out, hipStream_t st);
extern "C" void* ukhip_stream(void);

namespace {

constexpr int AT = 256;                   /* threads of every associative kernel */
constexpr int LDS_BYTES = 160 * 1024;     /* gfx950 LDS per workgroup */

/* ------------------------------------------------------------------------ */
/* workgroup-level dense algebra on blocks of order <= NB                    */
/* ------------------------------------------------------------------------ */

/* element (i, k) of op(A) for a column-major A with leading dimension ld */
__device__ __forceinline__ double at(const double* A, int ld, bool t, int i, int k) {
  return t ? A[k + (int64_t)i * ld] : A[i + (int64_t)k * ld];
}

/* C (m x q, ld ldc) = alpha op(A) op(B) + beta C with op(A) m x kk and
 * op(B) kk x q. Each thread accumulates a TD x TD register tile; every
 * thread has read its inputs before any output is written, so C may alias
 * A or B. Ends with a barrier. */
template <int NB>
__device__ void mm(double* C, int ldc, const double* A, int lda, bool tA, const double* B, int ldb, bool tB, int m,
                   int q, int kk, double alpha, double beta) {
  constexpr int TD = (NB + 15) / 16;
  const int ti = threadIdx.x % 16, tj = threadIdx.x / 16;
  double acc[TD][TD];
#pragma unroll
  for (int a = 0; a < TD; ++a)
#pragma unroll
    for (int b = 0; b < TD; ++b) acc[a][b] = 0.0;
  for (int k = 0; k < kk; ++k) {
    double av[TD], bv[TD];
#pragma unroll
    for (int a = 0; a < TD; ++a) {
      const int i = ti + 16 * a;
      av[a] = i < m ? at(A, lda, tA, i, k) : 0.0;
    }
#pragma unroll
    for (int b = 0; b < TD; ++b) {
      const int j = tj + 16 * b;
      bv[b] = j < q ? at(B, ldb, tB, k, j) : 0.0;
    }
#pragma unroll
    for (int a = 0; a < TD; ++a)
#pragma unroll
      for (int b = 0; b < TD; ++b) acc[a][b] = fma(av[a], bv[b], acc[a][b]);
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < TD; ++a)
#pragma unroll
    for (int b = 0; b < TD; ++b) {
      const int i = ti + 16 * a, j = tj + 16 * b;
      if (i < m && j < q) {
        double* c = &C[i + (int64_t)j * ldc];
        *c = beta == 0.0 ? alpha * acc[a][b] : fma(beta, *c, alpha * acc[a][b]);
      }
    }
  __syncthreads();
}

/* y (m) = alpha op(A) x + beta y, one thread per row; y may alias x. */
__device__ void mv(double* y, const double* A, int lda, bool tA, const double* x, int m, int kk, double alpha,
                   double beta) {
  double r = 0.0;
  const int i = threadIdx.x;
  if (i < m)
    for (int k = 0; k < kk; ++k) r = fma(at(A, lda, tA, i, k), x[k], r);
  __syncthreads();
  if (i < m) y[i] = beta == 0.0 ? alpha * r : fma(beta, y[i], alpha * r);
  __syncthreads();
}

/* D = op(S) (m x q; op(S) transposed when t) */
__device__ void copy_mat(double* D, int ldd, const double* S, int lds, bool t, int m, int q) {
  for (int idx = threadIdx.x; idx < m * q; idx += AT) {
    const int i = idx % m, j = idx / m;
    D[i + (int64_t)j * ldd] = at(S, lds, t, i, j);
  }
  __syncthreads();
}

/* D = d I (n x n) */
__device__ void set_identity(double* D, int ld, int n, double d) {
  for (int idx = threadIdx.x; idx < n * n; idx += AT) {
    const int i = idx % n, j = idx / n;
    D[i + (int64_t)j * ld] = (i == j) ? d : 0.0;
  }
  __syncthreads();
}

/* y (m) = a x + y */
__device__ void axpy(double* y, const double* x, int m, double a) {
  for (int i = threadIdx.x; i < m; i += AT) y[i] = fma(a, x[i], y[i]);
  __syncthreads();
}

/* LU with partial pivoting (LAPACK dgetrf order) of M (n x n, ld) in place;
 * row k was swapped with row piv[k]. */
__device__ void lu(double* M, int ld, int n, int* piv) {
  for (int k = 0; k < n; ++k) {
    if (threadIdx.x < 64) { /* one wave finds the pivot row */
      double best = -1.0;
      int bi = k;
      for (int i = k + (int)threadIdx.x; i < n; i += 64) {
        const double v = fabs(M[i + k * ld]);
        if (v > best) {
          best = v;
          bi = i;
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_down(best, off);
        const int oi = __shfl_down(bi, off);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (threadIdx.x == 0) piv[k] = bi;
    }
    __syncthreads();
    const int pr = piv[k];
    if (pr != k)
      for (int j = threadIdx.x; j < n; j += AT) {
        const double t = M[k + j * ld];
        M[k + j * ld] = M[pr + j * ld];
        M[pr + j * ld] = t;
      }
    __syncthreads();
    const double d = M[k + k * ld];
    for (int i = k + 1 + (int)threadIdx.x; i < n; i += AT) M[i + k * ld] /= d;
    __syncthreads();
    const int rem = n - k - 1;
    for (int idx = threadIdx.x; idx < rem * rem; idx += AT) {
      const int i = k + 1 + idx % rem, j = k + 1 + idx / rem;
      M[i + j * ld] = fma(-M[i + k * ld], M[k + j * ld], M[i + j * ld]);
    }
    __syncthreads();
  }
}

/* X (n x q, ld ldx) <- M^{-1} X, or M^{-T} X when trans, for M = P^T L U
 * from lu(). One thread per right-hand side. */
__device__ void lu_solve(const double* M, int ld, const int* piv, int n, double* X, int ldx, int q, bool trans) {
  for (int j = threadIdx.x; j < q; j += AT) {
    double* x = X + (int64_t)j * ldx;
    if (!trans) {
      for (int k = 0; k < n; ++k)
        if (piv[k] != k) {
          const double t = x[k];
          x[k] = x[piv[k]];
          x[piv[k]] = t;
        }
      for (int i = 0; i < n; ++i) {
        double s = x[i];
        for (int k = 0; k < i; ++k) s = fma(-M[i + k * ld], x[k], s);
        x[i] = s;
      }
      for (int i = n - 1; i >= 0; --i) {
        double s = x[i];
        for (int k = i + 1; k < n; ++k) s = fma(-M[i + k * ld], x[k], s);
        x[i] = s / M[i + i * ld];
      }
    } else {
      for (int i = 0; i < n; ++i) {
        double s = x[i];
        for (int k = 0; k < i; ++k) s = fma(-M[k + i * ld], x[k], s);
        x[i] = s / M[i + i * ld];
      }
      for (int i = n - 1; i >= 0; --i) {
        double s = x[i];
        for (int k = i + 1; k < n; ++k) s = fma(-M[k + i * ld], x[k], s);
        x[i] = s;
      }
      for (int k = n - 1; k >= 0; --k)
        if (piv[k] != k) {
          const double t = x[k];
          x[k] = x[piv[k]];
          x[piv[k]] = t;
        }
    }
  }
  __syncthreads();
}

/* ------------------------------------------------------------------------ */
/* batch data                                                                */
/* ------------------------------------------------------------------------ */

struct AsProb {
  int64_t s0; /* first step (batch-global index) */
  int64_t k;  /* steps */
  int64_t p;  /* first step with a proper filtered distribution */
  int64_t e0; /* first filtering element (elements of steps p+1 .. k-1) */
};

/* covariance-form model and filtered distributions, one slot per step */
struct AsModel {
  double* F;   /* n x n (unused on a problem's first step) */
  double* c;   /* n */
  double* Q;   /* n x n */
  double* xf;  /* filtered mean, steps >= p */
  double* Pf;  /* filtered covariance, steps >= p */
  double* Lam; /* filtered information matrix, steps < p */
  double* eta; /* filtered information vector, steps < p */
};

/* Element arrays: filtering (A, b, Z, e, J) or smoothing (E = A, g = b, L = Z). */
struct AsElems {
  double* A;
  double* b;
  double* Z;
  double* e;
  double* J;
};

/* ------------------------------------------------------------------------ */
/* model                                                                     */
/* ------------------------------------------------------------------------ */

template <int NB>
__global__ void __launch_bounds__(AT) ka_model(uk_chain_t wch, const int8_t* first, AsModel md, int n) {
  __shared__ double Dm[NB * NB], Di[NB * NB];
  __shared__ int piv[NB];
  const int64_t s = blockIdx.x;
  if (first[s]) return;
  const int64_t nn = (int64_t)n * n;
  copy_mat(Dm, n, wch.D + s * nn, n, false, n, n);
  lu(Dm, n, n, piv);
  set_identity(Di, n, n, 1.0);
  lu_solve(Dm, n, piv, n, Di, n, n, false);                                   /* Di = D^{-1} */
  mm<NB>(md.F + s * nn, n, Di, n, false, wch.B + s * nn, n, false, n, n, n, -1.0, 0.0);
  mv(md.c + s * n, Di, n, false, wch.c + s * n, n, n, 1.0, 0.0);
  mm<NB>(md.Q + s * nn, n, Di, n, false, Di, n, true, n, n, n, 1.0, 0.0);
}

/* ------------------------------------------------------------------------ */
/* prior: information filter over steps 0..p, one workgroup per problem      */
/* ------------------------------------------------------------------------ */

template <int NB>
__global__ void __launch_bounds__(AT) ka_prior(const AsProb* probs, uk_chain_t wch, const int32_t* m_obs, AsModel md,
                                               int n, int mr) {
  __shared__ double Lm[NB * NB], Mx[NB * NB], Qi[NB * NB], T2[NB * NB];
  __shared__ double et[NB], v1[NB], v2[NB];
  __shared__ int piv[NB];
  const AsProb pr = probs[blockIdx.x];
  const int64_t nn = (int64_t)n * n;
  for (int i = threadIdx.x; i < n * n; i += AT) Lm[i] = 0.0;
  for (int i = threadIdx.x; i < n; i += AT) et[i] = 0.0;
  __syncthreads();
  for (int64_t t = 0; t <= pr.p; ++t) {
    const int64_t s = pr.s0 + t;
    if (t > 0) {
      /* predict: with Mx = Lam + F^T Q^{-1} F and r = eta - F^T Q^{-1} c,
       * Lam <- Q^{-1} - Q^{-1} F Mx^{-1} F^T Q^{-1}, eta <- Q^{-1} c + Q^{-1} F Mx^{-1} r */
      const double* F = md.F + s * nn;
      copy_mat(Mx, n, md.Q + s * nn, n, false, n, n);
      lu(Mx, n, n, piv);
      set_identity(Qi, n, n, 1.0);
      lu_solve(Mx, n, piv, n, Qi, n, n, false);                               /* Qi = Q^{-1} */
      mm<NB>(T2, n, Qi, n, false, F, n, false, n, n, n, 1.0, 0.0);             /* T2 = Q^{-1} F */
      copy_mat(Mx, n, Lm, n, false, n, n);
      mm<NB>(Mx, n, F, n, true, T2, n, false, n, n, n, 1.0, 1.0);              /* Mx = Lam + F^T T2 */
      mv(v1, Qi, n, false, md.c + s * n, n, n, 1.0, 0.0);                     /* v1 = Q^{-1} c */
      mv(v2, F, n, true, v1, n, n, -1.0, 0.0);
      axpy(v2, et, n, 1.0);                                                   /* v2 = r */
      lu(Mx, n, n, piv);
      lu_solve(Mx, n, piv, n, v2, n, 1, false);                               /* v2 = Mx^{-1} r */
      copy_mat(Lm, n, T2, n, true, n, n);                                     /* Lm = T2^T */
      lu_solve(Mx, n, piv, n, Lm, n, n, false);                               /* Lm = Mx^{-1} T2^T */
      mm<NB>(Qi, n, T2, n, false, Lm, n, false, n, n, n, -1.0, 1.0);
      copy_mat(Lm, n, Qi, n, false, n, n);                                    /* Lam- */
      mv(et, T2, n, false, v2, n, n, 1.0, 0.0);
      axpy(et, v1, n, 1.0);                                                   /* eta- */
    }
    const int m = m_obs[s];
    if (m > 0) { /* update: Lam += C^T C, eta += C^T o (whitened) */
      const double* Cw = wch.C + s * nn;
      mm<NB>(Lm, n, Cw, mr, true, Cw, mr, false, n, n, m, 1.0, 1.0);
      mv(et, Cw, mr, true, wch.o + s * n, n, m, 1.0, 1.0);
    }
    if (t < pr.p) {
      copy_mat(md.Lam + s * nn, n, Lm, n, false, n, n);
      for (int i = threadIdx.x; i < n; i += AT) md.eta[s * n + i] = et[i];
      __syncthreads();
    }
  }
  /* P_p = Lam^{-1}, x_p = P_p eta */
  const int64_t sp = pr.s0 + pr.p;
  copy_mat(Mx, n, Lm, n, false, n, n);
  lu(Mx, n, n, piv);
  set_identity(Qi, n, n, 1.0);
  lu_solve(Mx, n, piv, n, Qi, n, n, false);
  copy_mat(md.Pf + sp * nn, n, Qi, n, false, n, n);
  mv(md.xf + sp * n, Qi, n, false, et, n, n, 1.0, 0.0);
}

/* ------------------------------------------------------------------------ */
/* filtering elements (c :370-505, matlab :49-91)                            */
/* ------------------------------------------------------------------------ */

/*
 * Step s with model (F, c, Q) and whitened observation (C, o), noise I:
 *   S = C Q C^T + I, gain K = Q C^T S^{-1} = T^T with T = S^{-1} C Q,
 *   A = F - K C F,  b = c + K (o - C c),  Z = Q - K C Q,
 *   e = (C F)^T S^{-1} (o - C c),  J = (C F)^T S^{-1} C F.
 * The first element after the prior (x_p, P_p) uses Q1 = Q + F P_p F^T in
 * place of Q, A = 0, and b = m1 + K (o - C m1) with m1 = F x_p + c
 * (Z = Q1 - K S K^T = Q1 - K C Q1). Unobserved: A = F (0 if first), b = c
 * (m1 if first), Z = Q (Q1), e = 0, J = 0.
 */
template <int NB>
__global__ void __launch_bounds__(AT) ka_felem(const AsProb* probs, const int32_t* elem_prob, uk_chain_t wch,
                                               const int32_t* m_obs, AsModel md, AsElems el, int n, int mr) {
  __shared__ double Qs[NB * NB], CQ[NB * NB], S[NB * NB], T[NB * NB];
  __shared__ double w1[NB], r[NB];
  __shared__ int piv[NB];
  const int64_t g = blockIdx.x;
  const AsProb pr = probs[elem_prob[g]];
  const int64_t t = g - pr.e0 + pr.p + 1;
  const int64_t s = pr.s0 + t;
  const bool firstel = t == pr.p + 1;
  const int64_t nn = (int64_t)n * n;
  const double* F = md.F + s * nn;
  const double* c = md.c + s * n;
  double* A = el.A + g * nn;
  double* b = el.b + g * n;
  double* Z = el.Z + g * nn;
  double* e = el.e + g * n;
  double* J = el.J + g * nn;
  const int64_t spv = pr.s0 + pr.p;

  copy_mat(Qs, n, md.Q + s * nn, n, false, n, n);
  for (int i = threadIdx.x; i < n; i += AT) w1[i] = c[i];
  __syncthreads();
  if (firstel) { /* Q1 = Q + F P F^T, m1 = F x + c */
    mm<NB>(T, n, F, n, false, md.Pf + spv * nn, n, false, n, n, n, 1.0, 0.0);
    mm<NB>(Qs, n, T, n, false, F, n, true, n, n, n, 1.0, 1.0);
    mv(w1, F, n, false, md.xf + spv * n, n, n, 1.0, 1.0);
  }
  const int m = m_obs[s];
  if (m == 0) {
    for (int idx = threadIdx.x; idx < n * n; idx += AT) {
      A[idx] = firstel ? 0.0 : F[idx];
      Z[idx] = Qs[idx];
      J[idx] = 0.0;
    }
    for (int i = threadIdx.x; i < n; i += AT) {
      b[i] = w1[i];
      e[i] = 0.0;
    }
    return;
  }
  const double* Cw = wch.C + s * nn; /* m x n, ld mr */
  const double* ow = wch.o + s * n;
  mm<NB>(CQ, m, Cw, mr, false, Qs, n, false, m, n, n, 1.0, 0.0);             /* CQ = C Q */
  set_identity(S, m, m, 1.0);
  mm<NB>(S, m, CQ, m, false, Cw, mr, true, m, m, n, 1.0, 1.0);               /* S = C Q C^T + I */
  lu(S, m, m, piv);
  copy_mat(T, m, CQ, m, false, m, n);
  lu_solve(S, m, piv, m, T, m, n, false);                                    /* T = S^{-1} C Q = K^T */
  mv(r, Cw, mr, false, w1, m, n, -1.0, 0.0);
  axpy(r, ow, m, 1.0);                                                       /* r = o - C w1 */
  mv(b, T, m, true, r, n, m, 1.0, 0.0);
  axpy(b, w1, n, 1.0);                                                       /* b = w1 + K r */
  copy_mat(Z, n, Qs, n, false, n, n);
  mm<NB>(Z, n, T, m, true, CQ, m, false, n, n, m, -1.0, 1.0);                /* Z = Q - K C Q */
  mm<NB>(CQ, m, Cw, mr, false, F, n, false, m, n, n, 1.0, 0.0);              /* CQ <- C F */
  if (firstel) {
    for (int idx = threadIdx.x; idx < n * n; idx += AT) A[idx] = 0.0;
    __syncthreads();
  } else {
    copy_mat(A, n, F, n, false, n, n);
    mm<NB>(A, n, T, m, true, CQ, m, false, n, n, m, -1.0, 1.0);              /* A = F - K C F */
  }
  mv(r, Cw, mr, false, c, m, n, -1.0, 0.0);
  axpy(r, ow, m, 1.0);                                                       /* r = o - C c */
  lu_solve(S, m, piv, m, r, m, 1, false);
  mv(e, CQ, m, true, r, n, m, 1.0, 0.0);                                     /* e = (C F)^T S^{-1} r */
  copy_mat(T, m, CQ, m, false, m, n);
  lu_solve(S, m, piv, m, T, m, n, false);                                    /* T = S^{-1} C F */
  mm<NB>(J, n, CQ, m, true, T, m, false, n, n, m, 1.0, 0.0);                 /* J = (C F)^T S^{-1} C F */
}

/* ------------------------------------------------------------------------ */
/* combines (c :889-1016, matlab :240-262)                                   */
/* ------------------------------------------------------------------------ */

/* blocks of an accumulator or of the work area, in LDS or global memory */
struct Blk {
  double *A, *b, *Z, *e, *J;
};
struct Work {
  double *M, *P, *T, *v1, *v2;
  int* piv;
};

/* a <- a (x) element g, filtering:
 *   M = I + Z_i J_j,  Y^T = M^{-1} A_i,  X^T = M^{-T} A_j^T,
 *   e = Y (e_j - J_j b_i) + e_i,  J = Y J_j A_i + J_i,
 *   b = X (b_i + Z_i e_j) + b_j,  Z = X Z_i A_j^T + Z_j,  A = X A_i. */
template <int NB>
__device__ void filt_combine(const Blk& a, const Work& w, const AsElems& el, int64_t g, int n) {
  const int64_t nn = (int64_t)n * n;
  const double *Aj = el.A + g * nn, *bj = el.b + g * n, *Zj = el.Z + g * nn, *ej = el.e + g * n, *Jj = el.J + g * nn;
  set_identity(w.M, n, n, 1.0);
  mm<NB>(w.M, n, a.Z, n, false, Jj, n, false, n, n, n, 1.0, 1.0);
  lu(w.M, n, n, w.piv);
  copy_mat(w.P, n, a.A, n, false, n, n);
  lu_solve(w.M, n, w.piv, n, w.P, n, n, false);                              /* P = Y^T */
  mv(w.v2, Jj, n, false, a.b, n, n, -1.0, 0.0);
  axpy(w.v2, ej, n, 1.0);
  mv(a.e, w.P, n, true, w.v2, n, n, 1.0, 1.0);                               /* e */
  mm<NB>(w.T, n, Jj, n, false, a.A, n, false, n, n, n, 1.0, 0.0);
  mm<NB>(a.J, n, w.P, n, true, w.T, n, false, n, n, n, 1.0, 1.0);            /* J */
  copy_mat(w.P, n, Aj, n, true, n, n);
  lu_solve(w.M, n, w.piv, n, w.P, n, n, true);                               /* P = X^T */
  mv(w.v1, a.Z, n, false, ej, n, n, 1.0, 0.0);
  axpy(w.v1, a.b, n, 1.0);
  mv(a.b, w.P, n, true, w.v1, n, n, 1.0, 0.0);
  axpy(a.b, bj, n, 1.0);                                                     /* b */
  mm<NB>(w.T, n, w.P, n, true, a.Z, n, false, n, n, n, 1.0, 0.0);
  copy_mat(a.Z, n, Zj, n, false, n, n);
  mm<NB>(a.Z, n, w.T, n, false, Aj, n, true, n, n, n, 1.0, 1.0);             /* Z */
  mm<NB>(a.A, n, w.P, n, true, a.A, n, false, n, n, n, 1.0, 0.0);            /* A */
}

/* a <- a (x) element g, smoothing; a is the running suffix and g the next
 * earlier step: E = E_j E_i, g = E_j g_i + g_j, L = E_j L_i E_j^T + L_j. */
template <int NB>
__device__ void smooth_combine(const Blk& a, const Work& w, const AsElems& el, int64_t g, int n) {
  const int64_t nn = (int64_t)n * n;
  const double *Ej = el.A + g * nn, *gj = el.b + g * n, *Lj = el.Z + g * nn;
  mm<NB>(a.A, n, Ej, n, false, a.A, n, false, n, n, n, 1.0, 0.0);
  mv(a.b, Ej, n, false, a.b, n, n, 1.0, 0.0);
  axpy(a.b, gj, n, 1.0);
  mm<NB>(w.T, n, Ej, n, false, a.Z, n, false, n, n, n, 1.0, 0.0);
  copy_mat(a.Z, n, Lj, n, false, n, n);
  mm<NB>(a.Z, n, w.T, n, false, Ej, n, true, n, n, n, 1.0, 1.0);
}

__device__ void blk_load(const Blk& a, const AsElems& el, int64_t g, int n, bool filt) {
  const int64_t nn = (int64_t)n * n;
  for (int idx = threadIdx.x; idx < n * n; idx += AT) {
    a.A[idx] = el.A[g * nn + idx];
    a.Z[idx] = el.Z[g * nn + idx];
    if (filt) a.J[idx] = el.J[g * nn + idx];
  }
  for (int i = threadIdx.x; i < n; i += AT) {
    a.b[i] = el.b[g * n + i];
    if (filt) a.e[i] = el.e[g * n + i];
  }
  __syncthreads();
}

__device__ void blk_store(const Blk& a, const AsElems& el, int64_t g, int n, bool filt) {
  const int64_t nn = (int64_t)n * n;
  for (int idx = threadIdx.x; idx < n * n; idx += AT) {
    el.A[g * nn + idx] = a.A[idx];
    el.Z[g * nn + idx] = a.Z[idx];
    if (filt) el.J[g * nn + idx] = a.J[idx];
  }
  for (int i = threadIdx.x; i < n; i += AT) {
    el.b[g * n + i] = a.b[i];
    if (filt) el.e[g * n + i] = a.e[i];
  }
  __syncthreads();
}

/* LDS plan of a scan kernel: accumulator, then the work area when it fits */
template <int NB, bool FILT>
struct ScanPlan {
  static constexpr int ACC = FILT ? 3 * NB * NB + 2 * NB : 2 * NB * NB + NB;
  static constexpr int WORK = FILT ? 3 * NB * NB + 2 * NB : NB * NB;
  static constexpr bool WORK_GLOBAL = (ACC + WORK) * 8 + NB * 4 + 256 > LDS_BYTES;
  static constexpr int LDS = ACC + (WORK_GLOBAL ? 0 : WORK);
};

template <int NB, bool FILT>
__device__ void scan_setup(double* lds, double* gwork, Blk& a, Work& w) {
  using SP = ScanPlan<NB, FILT>;
  double* q = lds;
  a.A = q; q += NB * NB;
  a.Z = q; q += NB * NB;
  a.J = nullptr;
  a.e = nullptr;
  if (FILT) { a.J = q; q += NB * NB; }
  a.b = q; q += NB;
  if (FILT) { a.e = q; q += NB; }
  double* p = SP::WORK_GLOBAL ? gwork + (int64_t)blockIdx.x * SP::WORK : q;
  w.T = p; p += NB * NB;
  w.M = w.P = w.v1 = w.v2 = nullptr;
  if (FILT) {
    w.M = p; p += NB * NB;
    w.P = p; p += NB * NB;
    w.v1 = p; p += NB;
    w.v2 = p; p += NB;
  }
}

template <int NB, bool FILT>
__device__ __forceinline__ void combine(const Blk& a, const Work& w, const AsElems& el, int64_t g, int n) {
  if (FILT) filt_combine<NB>(a, w, el, g, n);
  else smooth_combine<NB>(a, w, el, g, n);
}

/* ------------------------------------------------------------------------ */
/* chunked scans                                                             */
/* ------------------------------------------------------------------------ */

/* A chunk: `len` elements of one problem in processing order, from `first`
 * in direction `dir` (+1 filtering, -1 smoothing); `tot` is its slot in the
 * chunk-total arrays and `rank` its index within its problem. */
struct Chunk {
  int64_t first;
  int64_t tot;
  int32_t len;
  int32_t dir;
  int32_t rank;
  int32_t pad;
};

/* pass 1: the total of every chunk */
template <int NB, bool FILT>
__global__ void __launch_bounds__(AT) ka_scan_reduce(const Chunk* ch, AsElems el, AsElems tot, int n, double* gwork) {
  using SP = ScanPlan<NB, FILT>;
  __shared__ double lds[SP::LDS];
  __shared__ int piv[NB];
  Blk a;
  Work w;
  scan_setup<NB, FILT>(lds, gwork, a, w);
  w.piv = piv;
  const Chunk c = ch[blockIdx.x];
  blk_load(a, el, c.first, n, FILT);
  for (int i = 1; i < c.len; ++i) combine<NB, FILT>(a, w, el, c.first + (int64_t)i * c.dir, n);
  blk_store(a, tot, c.tot, n, FILT);
}

/* pass 2, one Hillis-Steele round over the chunk totals of every problem:
 * out[t] = in[t - d] (x) in[t] when rank >= d, else in[t]. */
template <int NB, bool FILT>
__global__ void __launch_bounds__(AT) ka_scan_round(const Chunk* ch, AsElems in, AsElems out, int d, int n,
                                                    double* gwork) {
  using SP = ScanPlan<NB, FILT>;
  __shared__ double lds[SP::LDS];
  __shared__ int piv[NB];
  Blk a;
  Work w;
  scan_setup<NB, FILT>(lds, gwork, a, w);
  w.piv = piv;
  const Chunk c = ch[blockIdx.x];
  if (c.rank >= d) {
    blk_load(a, in, c.tot - d, n, FILT);
    combine<NB, FILT>(a, w, in, c.tot, n);
  } else {
    blk_load(a, in, c.tot, n, FILT);
  }
  blk_store(a, out, c.tot, n, FILT);
}

/* pass 3: every prefix within a chunk, seeded with the inclusive total of
 * the chunk before it; results overwrite the elements in place. */
template <int NB, bool FILT>
__global__ void __launch_bounds__(AT) ka_scan_down(const Chunk* ch, AsElems el, AsElems tot, int n, double* gwork) {
  using SP = ScanPlan<NB, FILT>;
  __shared__ double lds[SP::LDS];
  __shared__ int piv[NB];
  Blk a;
  Work w;
  scan_setup<NB, FILT>(lds, gwork, a, w);
  w.piv = piv;
  const Chunk c = ch[blockIdx.x];
  int i0 = 0;
  if (c.rank == 0) {
    blk_load(a, el, c.first, n, FILT);
    i0 = 1;
  } else {
    blk_load(a, tot, c.tot - 1, n, FILT);
  }
  for (int i = i0; i < c.len; ++i) {
    const int64_t g = c.first + (int64_t)i * c.dir;
    combine<NB, FILT>(a, w, el, g, n);
    blk_store(a, el, g, n, FILT);
  }
}

/* ------------------------------------------------------------------------ */
/* filtered distributions, smoothing elements, results                       */
/* ------------------------------------------------------------------------ */

/* filtered (x, P) of steps p+1 .. k-1: (b, Z) of the filtering prefixes */
__global__ void __launch_bounds__(AT) ka_filtered(const AsProb* probs, const int32_t* elem_prob, AsElems el,
                                                  AsModel md, int n) {
  const int64_t g = blockIdx.x;
  const AsProb pr = probs[elem_prob[g]];
  const int64_t s = pr.s0 + (g - pr.e0) + pr.p + 1;
  const int64_t nn = (int64_t)n * n;
  for (int idx = threadIdx.x; idx < n * n; idx += AT) md.Pf[s * nn + idx] = el.Z[g * nn + idx];
  for (int i = threadIdx.x; i < n; i += AT) md.xf[s * n + i] = el.b[g * n + i];
}

/*
 * Smoothing element of step s, stored at slot s (c :697-753, matlab :94-114):
 *   last step: E = 0, g = x, L = P;
 *   t >= p: with (F, c, Q) of step s + 1,
 *           E = P F^T (F P F^T + Q)^{-1}, g = x - E (F x + c), L = P - E F P;
 *   t < p:  from the information form (Lam, eta),
 *           L = (Lam + F^T Q^{-1} F)^{-1}, E = L F^T Q^{-1}, g = L (eta - F^T Q^{-1} c).
 */
template <int NB>
__global__ void __launch_bounds__(AT) ka_selem(const AsProb* probs, const int32_t* step_prob, AsModel md, AsElems el,
                                               int n) {
  __shared__ double M[NB * NB], T[NB * NB], U[NB * NB];
  __shared__ double v1[NB], v2[NB];
  __shared__ int piv[NB];
  const int64_t s = blockIdx.x;
  const AsProb pr = probs[step_prob[s]];
  const int64_t t = s - pr.s0;
  const int64_t nn = (int64_t)n * n;
  double* E = el.A + s * nn;
  double* gv = el.b + s * n;
  double* L = el.Z + s * nn;
  if (t == pr.k - 1) {
    for (int idx = threadIdx.x; idx < n * n; idx += AT) {
      E[idx] = 0.0;
      L[idx] = md.Pf[s * nn + idx];
    }
    for (int i = threadIdx.x; i < n; i += AT) gv[i] = md.xf[s * n + i];
    return;
  }
  const double* F = md.F + (s + 1) * nn;
  const double* c = md.c + (s + 1) * n;
  const double* Q = md.Q + (s + 1) * nn;
  if (t >= pr.p) {
    const double* P = md.Pf + s * nn;
    const double* x = md.xf + s * n;
    mm<NB>(T, n, P, n, false, F, n, true, n, n, n, 1.0, 0.0);                 /* T = P F^T */
    copy_mat(M, n, Q, n, false, n, n);
    mm<NB>(M, n, F, n, false, T, n, false, n, n, n, 1.0, 1.0);                /* M = F P F^T + Q */
    lu(M, n, n, piv);
    copy_mat(U, n, T, n, true, n, n);                                         /* U = F P */
    lu_solve(M, n, piv, n, U, n, n, true);                                    /* U = E^T */
    copy_mat(E, n, U, n, true, n, n);
    mv(v1, F, n, false, x, n, n, 1.0, 0.0);
    axpy(v1, c, n, 1.0);                                                      /* v1 = F x + c */
    mv(gv, U, n, true, v1, n, n, -1.0, 0.0);
    axpy(gv, x, n, 1.0);                                                      /* g = x - E v1 */
    copy_mat(M, n, T, n, true, n, n);                                         /* M = F P */
    copy_mat(L, n, P, n, false, n, n);
    mm<NB>(L, n, U, n, true, M, n, false, n, n, n, -1.0, 1.0);                /* L = P - E F P */
  } else {
    copy_mat(M, n, Q, n, false, n, n);
    lu(M, n, n, piv);
    copy_mat(T, n, F, n, false, n, n);
    lu_solve(M, n, piv, n, T, n, n, false);                                   /* T = Q^{-1} F */
    for (int i = threadIdx.x; i < n; i += AT) v1[i] = c[i];
    __syncthreads();
    lu_solve(M, n, piv, n, v1, n, 1, false);                                  /* v1 = Q^{-1} c */
    copy_mat(U, n, md.Lam + s * nn, n, false, n, n);
    mm<NB>(U, n, F, n, true, T, n, false, n, n, n, 1.0, 1.0);                 /* U = Lam + F^T Q^{-1} F */
    lu(U, n, n, piv);
    set_identity(M, n, n, 1.0);
    lu_solve(U, n, piv, n, M, n, n, false);                                   /* M = L */
    copy_mat(L, n, M, n, false, n, n);
    mm<NB>(E, n, M, n, false, T, n, true, n, n, n, 1.0, 0.0);                 /* E = L (Q^{-1} F)^T */
    mv(v2, F, n, true, v1, n, n, -1.0, 0.0);
    axpy(v2, md.eta + s * n, n, 1.0);
    mv(gv, M, n, false, v2, n, n, 1.0, 0.0);                                  /* g = L (eta - F^T Q^{-1} c) */
  }
}

/* the smoothing prefix of step s (taken from the last step) is (E, g, L):
 * estimate g, covariance L */
__global__ void __launch_bounds__(AT) ka_results(AsElems el, double* xout, double* Sout, int n) {
  const int64_t s = blockIdx.x;
  const int64_t nn = (int64_t)n * n;
  for (int idx = threadIdx.x; idx < n * n; idx += AT) Sout[s * nn + idx] = el.Z[s * nn + idx];
  for (int i = threadIdx.x; i < n; i += AT) xout[s * n + i] = el.b[s * n + i];
}

} /* namespace */

/* ------------------------------------------------------------------------ */
/* host driver                                                               */
/* ------------------------------------------------------------------------ */

struct ukhip_batch;
extern "C" void uk_batch_view(ukhip_batch* b, const uk_raw_t** raw, double** xout, double** Sout, int64_t* steps,
                              int* nprob, const int64_t** k, void** assoc_ws, size_t* assoc_bytes);
extern "C" void uk_batch_set_assoc_ws(ukhip_batch* b, void* p, size_t bytes);
extern "C" void uk_batch_chain0(ukhip_batch* b, uk_chain_t* ch);

namespace {

bool ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  uk_set_error("%s: %s", what, hipGetErrorString(e));
  return false;
}

/* chunks of at most `chunk` elements over `len` elements of one problem;
 * dir = +1 starts at `base`, dir = -1 at base + len - 1. */
int64_t make_chunks(std::vector<Chunk>& out, int64_t base, int64_t len, int dir, int64_t chunk, int64_t& tot0) {
  const int64_t nch = (len + chunk - 1) / chunk;
  for (int64_t r = 0; r < nch; ++r) {
    Chunk c{};
    const int64_t lo = r * chunk;
    c.first = dir > 0 ? base + lo : base + len - 1 - lo;
    c.len = (int32_t)std::min(chunk, len - lo);
    c.dir = dir;
    c.tot = tot0 + r;
    c.rank = (int32_t)r;
    out.push_back(c);
  }
  tot0 += nch;
  return nch;
}

/* chunk length: about CHUNKS_TARGET chunks over the batch, at least 4 */
constexpr int64_t CHUNKS_TARGET = 2048;

template <int NB>
int run_assoc(ukhip_batch* b) {
  hipStream_t st = (hipStream_t)ukhip_stream();
  const uk_raw_t* raw;
  double *xout, *Sout;
  int64_t S;
  int nprob;
  const int64_t* kk;
  void* ws;
  size_t ws_bytes;
  uk_batch_view(b, &raw, &xout, &Sout, &S, &nprob, &kk, &ws, &ws_bytes);
  const int n = raw->n, mr = raw->mr;
  uk_chain_t wch;
  uk_batch_chain0(b, &wch);
  wch.mr = mr;

  /* plan: p of every problem from the observation counts */
  std::vector<int32_t> m_h(S);
  if (!ok(hipMemcpy(m_h.data(), raw->m, S * sizeof(int32_t), hipMemcpyDeviceToHost), "download m")) return -1;
  std::vector<AsProb> probs(nprob);
  std::vector<int32_t> elem_prob, step_prob(S);
  int64_t s0 = 0, ne = 0;
  for (int p = 0; p < nprob; ++p) {
    AsProb pr{};
    pr.s0 = s0;
    pr.k = kk[p];
    int64_t acc = 0, pp = 0;
    for (pp = 0; pp < pr.k; ++pp) {
      acc += m_h[s0 + pp];
      if (acc >= n) break;
    }
    if (pp == pr.k) {
      uk_set_error("kalman_smooth_associative: problem %d: the observations do not determine the states", p);
      return -1;
    }
    pr.p = pp;
    pr.e0 = ne;
    for (int64_t t = pr.p + 1; t < pr.k; ++t) elem_prob.push_back(p);
    ne += pr.k - pr.p - 1;
    for (int64_t t = 0; t < pr.k; ++t) step_prob[s0 + t] = p;
    probs[p] = pr;
    s0 += pr.k;
  }
  const int64_t fchunk = std::max<int64_t>(4, (ne + CHUNKS_TARGET - 1) / CHUNKS_TARGET);
  const int64_t schunk = std::max<int64_t>(4, (S + CHUNKS_TARGET - 1) / CHUNKS_TARGET);
  std::vector<Chunk> fch, sch;
  int64_t ftot = 0, stot = 0, maxr_f = 0, maxr_s = 0;
  for (int p = 0; p < nprob; ++p) {
    const int64_t len = probs[p].k - probs[p].p - 1;
    if (len > 0) maxr_f = std::max(maxr_f, make_chunks(fch, probs[p].e0, len, +1, fchunk, ftot));
    maxr_s = std::max(maxr_s, make_chunks(sch, probs[p].s0, probs[p].k, -1, schunk, stot));
  }

  /* workspace */
  const size_t nn = (size_t)n * n;
  size_t need = 0;
  auto take = [&](size_t bytes) {
    const size_t r = need;
    need += (bytes + 255) & ~(size_t)255;
    return r;
  };
  auto elem_bytes = [&](int64_t cnt, bool filt) { return (size_t)std::max<int64_t>(cnt, 1) * ((filt ? 3 : 2) * nn + (filt ? 2 : 1) * n) * 8; };
  const size_t o_model = take((size_t)S * (4 * nn + 4 * n) * 8);
  const size_t o_fe = take(elem_bytes(ne, true));
  const size_t o_ft0 = take(elem_bytes(ftot, true)), o_ft1 = take(elem_bytes(ftot, true));
  const size_t o_se = take(elem_bytes(S, false));
  const size_t o_st0 = take(elem_bytes(stot, false)), o_st1 = take(elem_bytes(stot, false));
  const int64_t maxch = std::max<int64_t>((int64_t)std::max(fch.size(), sch.size()), 1);
  const size_t o_gw = take((size_t)maxch * ScanPlan<NB, true>::WORK * 8);
  const size_t tab_bytes = probs.size() * sizeof(AsProb) + (elem_prob.size() + step_prob.size()) * 4 +
                           (fch.size() + sch.size()) * sizeof(Chunk) + 6 * 256;
  const size_t o_tab = take(tab_bytes);
  if (ws_bytes < need) {
    if (ws) (void)hipFree(ws);
    uk_batch_set_assoc_ws(b, nullptr, 0);
    ws = nullptr;
    if (!ok(hipMalloc(&ws, need), "hipMalloc(associative workspace)")) return -1;
    uk_batch_set_assoc_ws(b, ws, need);
  }
  char* W = (char*)ws;
  AsModel md;
  {
    double* q = (double*)(W + o_model);
    md.F = q; q += S * nn;
    md.Q = q; q += S * nn;
    md.Pf = q; q += S * nn;
    md.Lam = q; q += S * nn;
    md.c = q; q += S * n;
    md.xf = q; q += S * n;
    md.eta = q;
  }
  auto elems_at = [&](size_t off, int64_t cnt, bool filt) {
    cnt = std::max<int64_t>(cnt, 1);
    AsElems e{};
    double* q = (double*)(W + off);
    e.A = q; q += cnt * nn;
    e.Z = q; q += cnt * nn;
    if (filt) { e.J = q; q += cnt * nn; }
    e.b = q; q += cnt * n;
    if (filt) e.e = q;
    return e;
  };
  const AsElems fe = elems_at(o_fe, ne, true);
  const AsElems ft[2] = {elems_at(o_ft0, ftot, true), elems_at(o_ft1, ftot, true)};
  const AsElems se = elems_at(o_se, S, false);
  const AsElems stt[2] = {elems_at(o_st0, stot, false), elems_at(o_st1, stot, false)};
  double* gw = (double*)(W + o_gw);

  /* tables */
  std::vector<char> tab(tab_bytes, 0);
  size_t to = 0;
  auto put = [&](const void* src, size_t bytes) {
    const size_t at_ = to;
    if (bytes) std::memcpy(tab.data() + to, src, bytes);
    to += (bytes + 255) & ~(size_t)255;
    return W + o_tab + at_;
  };
  const AsProb* d_probs = (const AsProb*)put(probs.data(), probs.size() * sizeof(AsProb));
  const int32_t* d_elem_prob = (const int32_t*)put(elem_prob.data(), elem_prob.size() * 4);
  const int32_t* d_step_prob = (const int32_t*)put(step_prob.data(), step_prob.size() * 4);
  const Chunk* d_fch = (const Chunk*)put(fch.data(), fch.size() * sizeof(Chunk));
  const Chunk* d_sch = (const Chunk*)put(sch.data(), sch.size() * sizeof(Chunk));
  if (!ok(hipMemcpyAsync(W + o_tab, tab.data(), to, hipMemcpyHostToDevice, st), "upload scan tables")) return -1;

  const dim3 blk(AT);
  if (!ok(uk_launch_weigh(raw, &wch, st), "k_weigh")) return -1;
  ka_model<NB><<<dim3((unsigned)S), blk, 0, st>>>(wch, raw->first, md, n);
  ka_prior<NB><<<dim3((unsigned)nprob), blk, 0, st>>>(d_probs, wch, raw->m, md, n, mr);
  if (ne > 0) {
    const unsigned nc = (unsigned)fch.size();
    ka_felem<NB><<<dim3((unsigned)ne), blk, 0, st>>>(d_probs, d_elem_prob, wch, raw->m, md, fe, n, mr);
    ka_scan_reduce<NB, true><<<dim3(nc), blk, 0, st>>>(d_fch, fe, ft[0], n, gw);
    int cur = 0;
    for (int64_t d = 1; d < maxr_f; d <<= 1) {
      ka_scan_round<NB, true><<<dim3(nc), blk, 0, st>>>(d_fch, ft[cur], ft[cur ^ 1], (int)d, n, gw);
      cur ^= 1;
    }
    ka_scan_down<NB, true><<<dim3(nc), blk, 0, st>>>(d_fch, fe, ft[cur], n, gw);
    ka_filtered<<<dim3((unsigned)ne), blk, 0, st>>>(d_probs, d_elem_prob, fe, md, n);
  }
  {
    const unsigned nc = (unsigned)sch.size();
    ka_selem<NB><<<dim3((unsigned)S), blk, 0, st>>>(d_probs, d_step_prob, md, se, n);
    ka_scan_reduce<NB, false><<<dim3(nc), blk, 0, st>>>(d_sch, se, stt[0], n, gw);
    int cur = 0;
    for (int64_t d = 1; d < maxr_s; d <<= 1) {
      ka_scan_round<NB, false><<<dim3(nc), blk, 0, st>>>(d_sch, stt[cur], stt[cur ^ 1], (int)d, n, gw);
      cur ^= 1;
    }
    ka_scan_down<NB, false><<<dim3(nc), blk, 0, st>>>(d_sch, se, stt[cur], n, gw);
    ka_results<<<dim3((unsigned)S), blk, 0, st>>>(se, xout, Sout, n);
  }
  if (!ok(hipGetLastError(), "associative launches")) return -1;
  /* the host tables above must outlive the asynchronous upload */
  return ok(hipStreamSynchronize(st), "associative smoothing") ? 0 : -1;
}

}  // namespace

/* The associative smoother over a batch (c/kalman_associative_smoother.c:
 * 1199-1240). Covariances are always produced, as in the reference, whose
 * associative path does not read KALMAN_NO_COVARIANCE. */
extern "C" int uk_assoc_smooth(struct ukhip_batch* b, uint32_t options) {
  (void)options;
  const uk_raw_t* raw;
  double *xout, *Sout;
  int64_t S;
  int nprob;
  const int64_t* k;
  void* ws;
  size_t wsb;
  uk_batch_view(b, &raw, &xout, &Sout, &S, &nprob, &k, &ws, &wsb);
  switch (uk_bucket(raw->n > raw->mr ? raw->n : raw->mr)) {
    case 8: return run_assoc<8>(b);
    case 16: return run_assoc<16>(b);
    case 32: return run_assoc<32>(b);
    case 48: return run_assoc<48>(b);
    case 64: return run_assoc<64>(b);
    default: uk_set_error("associative: state dimension %d is above 64", raw->n); return -1;
  }
}
