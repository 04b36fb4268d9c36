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
 * Every element build and every combine runs in one 256-thread workgroup.
 * Its blocks live in LDS with the odd leading dimension NB + 1, so that the
 * column-per-thread and transposed accesses spread over the LDS banks; the
 * other operand of a combine is staged from HBM into LDS once. (The
 * filtering work blocks of the 64 bucket go to a per-workgroup global slice:
 * six 64 x 65 blocks exceed the 160 KiB of LDS.) Triangular solves keep each
 * right-hand side in one thread's registers. Observations are used whitened
 * (noise covariance I), which is the reference's S = G K G^T + C after the
 * congruence with the whitening W.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/kalman.h"
#include "../../include/ukhip.h"
#include "ukhip_internal.h"

extern "C" void uk_set_error(const char* fmt, ...);
extern "C" int uk_bucket(int n);
extern "C" void* ukhip_stream(void);

namespace {

constexpr int AT = 256;               /* threads of every associative kernel */
constexpr int LDS_BYTES = 160 * 1024; /* gfx950 LDS per workgroup */

/* ------------------------------------------------------------------------ */
/* workgroup-level dense algebra on blocks of order <= NB                    */
/* ------------------------------------------------------------------------ */

/* element (i, k) of op(A) for a column-major A with leading dimension ld */
__device__ __forceinline__ double at(const double* A, int ld, bool t, int i, int k) {
  return t ? A[k + (int64_t)i * ld] : A[i + (int64_t)k * ld];
}

/* C (m x q, ld ldc) = alpha op(A) op(B) + beta C with op(A) m x kk and
 * op(B) kk x q. Each thread accumulates a TD x TD register tile. The k loop
 * runs in groups of KU whose loads are all issued before the group's FMAs
 * (one wave per SIMD is typical here, so latency has to be hidden by
 * independent loads); indices are clamped into range and the extra terms
 * zeroed, so there are no branches. Every thread has read its inputs before
 * any output is written, so C may alias A or B. Ends with a barrier. */
template <int NB>
__device__ void mm(double* C, int ldc, const double* A, int lda, bool tA, const double* B, int ldb, bool tB, int m,
                   int q, int kk, double alpha, double beta) {
  constexpr int TD = (NB + 15) / 16;
  constexpr int KU = TD >= 4 ? 4 : 8;
  const int ti = threadIdx.x % 16, tj = threadIdx.x / 16;
  int ia[TD], jb[TD];
#pragma unroll
  for (int a = 0; a < TD; ++a) {
    const int i = ti + 16 * a;
    ia[a] = i < m ? i : m - 1;
  }
#pragma unroll
  for (int b = 0; b < TD; ++b) {
    const int j = tj + 16 * b;
    jb[b] = j < q ? j : q - 1;
  }
  double acc[TD][TD];
#pragma unroll
  for (int a = 0; a < TD; ++a)
#pragma unroll
    for (int b = 0; b < TD; ++b) acc[a][b] = 0.0;
  for (int k0 = 0; k0 < kk; k0 += KU) {
    double av[KU][TD], bv[KU][TD];
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      const int k = k0 + u < kk ? k0 + u : kk - 1;
#pragma unroll
      for (int a = 0; a < TD; ++a) av[u][a] = at(A, lda, tA, ia[a], k);
#pragma unroll
      for (int b = 0; b < TD; ++b) bv[u][b] = at(B, ldb, tB, k, jb[b]);
    }
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      if (k0 + u >= kk) /* uniform: only the last group has a tail */
#pragma unroll
        for (int a = 0; a < TD; ++a) av[u][a] = 0.0;
#pragma unroll
      for (int a = 0; a < TD; ++a)
#pragma unroll
        for (int b = 0; b < TD; ++b) acc[a][b] = fma(av[u][a], bv[u][b], acc[a][b]);
    }
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

/* y (m) = alpha op(A) x + beta y, one thread per row, the k loop in groups
 * of 8 loads with two partial sums; y may alias x. */
__device__ void mv(double* y, const double* A, int lda, bool tA, const double* x, int m, int kk, double alpha,
                   double beta) {
  double r0 = 0.0, r1 = 0.0;
  const int i = threadIdx.x;
  if (i < m) {
    for (int k0 = 0; k0 < kk; k0 += 8) {
      double av[8], xv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int k = k0 + u < kk ? k0 + u : kk - 1;
        av[u] = at(A, lda, tA, i, k);
        xv[u] = k0 + u < kk ? x[k] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; u += 2) {
        r0 = fma(av[u], xv[u], r0);
        r1 = fma(av[u + 1], xv[u + 1], r1);
      }
    }
  }
  const double r = r0 + r1;
  __syncthreads();
  if (i < m) y[i] = beta == 0.0 ? alpha * r : fma(beta, y[i], alpha * r);
  __syncthreads();
}

/* D = op(S) (m x q; op(S) transposed when t). Threads walk an NB x NB grid
 * (compile-time strides) so that S is read along its columns; all of a
 * thread's loads are issued before its stores. */
template <int NB>
__device__ void copy_mat(double* D, int ldd, const double* S, int lds, bool t, int m, int q) {
  constexpr int PER = (NB * NB + AT - 1) / AT;
  double v[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int idx = (int)threadIdx.x + u * AT;
    const int r = idx % NB, c = idx / NB;
    int i = t ? c : r, j = t ? r : c;
    i = i < m ? i : m - 1;
    j = j < q ? j : q - 1;
    v[u] = at(S, lds, t, i, j);
  }
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int idx = (int)threadIdx.x + u * AT;
    const int r = idx % NB, c = idx / NB;
    const int i = t ? c : r, j = t ? r : c;
    if (idx < NB * NB && i < m && j < q) D[i + (int64_t)j * ldd] = v[u];
  }
  __syncthreads();
}

/* D = d I on the leading n x n block; the padding rows and columns
 * [n, NB) of the NB x NB block become the identity. */
template <int NB>
__device__ void set_identity(double* D, int ld, int n, double d) {
  constexpr int PER = (NB * NB + AT - 1) / AT;
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int idx = (int)threadIdx.x + u * AT;
    const int i = idx % NB, j = idx / NB;
    if (idx < NB * NB) D[i + (int64_t)j * ld] = (i == j) ? (i < n ? d : 1.0) : 0.0;
  }
  __syncthreads();
}

/* y (m) = a x + y */
__device__ void axpy(double* y, const double* x, int m, double a) {
  for (int i = threadIdx.x; i < m; i += AT) y[i] = fma(a, x[i], y[i]);
  __syncthreads();
}

/* y (m) = x */
__device__ void copy_vec(double* y, const double* x, int m) {
  for (int i = threadIdx.x; i < m; i += AT) y[i] = x[i];
  __syncthreads();
}

/* Pivot keys: |v| with its low 7 mantissa bits replaced by (127 - row), so
 * that an unsigned 64-bit maximum picks the largest magnitude and, among
 * (near-)equal magnitudes, the first row. NaN sorts above every number. */
__device__ __forceinline__ unsigned long long pivot_key(double v, int row) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(v));
  return (b & ~127ull) | (unsigned long long)(127 - row);
}
__device__ __forceinline__ int pivot_row(unsigned long long key) { return 127 - (int)(key & 127ull); }

/* Workgroup scratch of gj_inverse: pivot keys, pivot rows, column order. */
template <int NB>
struct GJScratch {
  unsigned long long key[NB + 1];
  int piv[NB];
  int col[NB];
};

/*
 * Inverse of the leading n x n block of A (ld) by Gauss-Jordan elimination
 * with partial pivoting (row interchanges, undone on the columns at the
 * end). Each step maps the whole matrix from one buffer into the other
 * (A <-> W, both ld) in one pass:
 *   new row k  = (row p, with 1 in column k) / a_pk,
 *   new row r  = (row r', with 0 in column k) - a_r'k * new row k,
 * r' being row r after the interchange of rows k and p. The pivot of the
 * next column is found in the same pass with LDS atomic maxima over the
 * candidates' keys, so a step costs one barrier. Returns the buffer (A or
 * W) holding the inverse; the other one is free. n <= NB <= 127.
 */
template <int NB>
__device__ double* gj_inverse(double* A, double* W, int ld, int n, GJScratch<NB>& g) {
  for (int i = threadIdx.x; i <= NB; i += AT) g.key[i] = 0ull;
  __syncthreads();
  for (int r = threadIdx.x; r < n; r += AT) atomicMax(&g.key[0], pivot_key(A[r], r));
  __syncthreads();
  double* cur = A;
  double* nxt = W;
  constexpr int PER = (NB * NB + AT - 1) / AT;
  for (int k = 0; k < n; ++k) {
    const int p = pivot_row(g.key[k]);
    if (threadIdx.x == 0) g.piv[k] = p;
    const double rp = 1.0 / cur[p + k * ld];
    double pj[PER], fk[PER], aj[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) { /* all loads first, indices clamped into the block */
      const int idx = (int)threadIdx.x + u * AT;
      const int r = idx < NB * NB ? idx % NB : 0, j = idx < NB * NB ? idx / NB : 0;
      const int sr = (r == p) ? k : (r == k ? p : r);
      pj[u] = cur[p + j * ld];
      fk[u] = cur[sr + k * ld];
      aj[u] = cur[sr + j * ld];
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int idx = (int)threadIdx.x + u * AT;
      const int r = idx % NB, j = idx / NB;
      if (idx < NB * NB && r < n && j < n) {
        const double rowk = (j == k ? 1.0 : pj[u]) * rp;
        const double v = r == k ? rowk : fma(-fk[u], rowk, j == k ? 0.0 : aj[u]);
        nxt[r + j * ld] = v;
        if (j == k + 1 && r > k) atomicMax(&g.key[k + 1], pivot_key(v, r));
      }
    }
    __syncthreads();
    double* t = cur;
    cur = nxt;
    nxt = t;
  }
  /* undo the interchanges on the columns, last first */
  if (threadIdx.x == 0) {
    for (int j = 0; j < n; ++j) g.col[j] = j;
    for (int k = n - 1; k >= 0; --k) {
      const int t = g.col[k];
      g.col[k] = g.col[g.piv[k]];
      g.col[g.piv[k]] = t;
    }
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < NB * NB; idx += AT) {
    const int r = idx % NB, j = idx / NB;
    if (r < n && j < n) nxt[r + j * ld] = cur[r + g.col[j] * ld];
  }
  __syncthreads();
  return nxt;
}

/* X = A^{-1} for the leading n x n block of A (ld), A left unchanged (the
 * Gauss-Jordan inverse of a copy). W1 and W2 are scratch blocks; the
 * result is in one of them (returned), the other is free afterwards. */
template <int NB>
__device__ double* inverse(const double* A, double* W1, double* W2, int ld, int n, GJScratch<NB>& g) {
  copy_mat<NB>(W1, ld, A, ld, false, n, n);
  return gj_inverse<NB>(W1, W2, ld, n, g);
}

/* LU with partial pivoting (LAPACK dgetrf order, whole rows swapped) of the
 * leading n x n block of M (ld) in place; row k was interchanged with row
 * piv[k] at step k. Used where the matrices can be ill-conditioned
 * (information-form steps, smoothing gains): solving with the factors keeps
 * the error at the level of a backward-stable solve, where multiplying by an
 * explicit inverse can lose a further factor. */
template <int NB>
__device__ void lu(double* M, int ld, int n, int* piv, GJScratch<NB>& g) {
  for (int i = threadIdx.x; i <= NB; i += AT) g.key[i] = 0ull;
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    for (int r = k + (int)threadIdx.x; r < n; r += AT) atomicMax(&g.key[k], pivot_key(M[r + k * ld], r));
    __syncthreads();
    const int pr = pivot_row(g.key[k]);
    if (pr != k)
      for (int j = threadIdx.x; j < n; j += AT) {
        const double t = M[k + j * ld];
        M[k + j * ld] = M[pr + j * ld];
        M[pr + j * ld] = t;
      }
    if (threadIdx.x == 0) piv[k] = pr;
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

/* X (n x q, ld ldx) <- A^{-1} X, or A^{-T} X when trans, with the factors of
 * lu(A): one thread per right-hand side, in place. */
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

/* Per-thread registers holding an n x n global block (ld n) distributed over
 * the NB x NB grid: entry idx = threadIdx.x + u * AT. */
template <int NB>
struct RegBlk {
  static constexpr int PER = (NB * NB + AT - 1) / AT;
  double v[PER];
  __device__ void load(const double* G, int n) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int idx = (int)threadIdx.x + u * AT;
      const int i = idx % NB, j = idx / NB;
      v[u] = (idx < NB * NB && i < n && j < n) ? G[i + (int64_t)j * n] : 0.0;
    }
  }
  /* D (ld) <- the block, or its transpose; no barrier */
  __device__ void store(double* D, int ld, int n, bool t) const {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int idx = (int)threadIdx.x + u * AT;
      const int i = idx % NB, j = idx / NB;
      if (idx < NB * NB && i < n && j < n) D[t ? j + i * ld : i + j * ld] = v[u];
    }
  }
};

/* ------------------------------------------------------------------------ */
/* batch data                                                                */
/* ------------------------------------------------------------------------ */

struct AsProb {
  int64_t s0; /* first step (batch-global index) */
  int64_t k;  /* steps */
  int64_t p;  /* first step with a proper filtered distribution */
  int64_t e0; /* first filtering element (elements of steps p+1 .. k-1) */
};

/* covariance-form model and filtered distributions, one slot per step (ld n) */
struct AsModel {
  double* F;   /* n x n (unused on a problem's first step) */
  double* c;   /* n */
  double* Q;   /* n x n */
  double* xf;  /* filtered mean, steps >= p */
  double* Pf;  /* filtered covariance, steps >= p */
  double* Lam; /* filtered information matrix, steps < p */
  double* eta; /* filtered information vector, steps < p */
};

/* Element arrays (ld n): filtering (A, b, Z, e, J) or smoothing (E = A, g = b, L = Z). */
struct AsElems {
  double* A;
  double* b;
  double* Z;
  double* e;
  double* J;
};

template <int NB>
struct Ld {
  static constexpr int L = NB + 1; /* LDS leading dimension of an NB x NB block */
  static constexpr int SZ = NB * L;
};

/* ------------------------------------------------------------------------ */
/* model                                                                     */
/* ------------------------------------------------------------------------ */

template <int NB>
__global__ void __launch_bounds__(AT) ka_model(uk_chain_t wch, const int8_t* first, AsModel md, int n) {
  constexpr int L = Ld<NB>::L;
  __shared__ double B0[Ld<NB>::SZ], B1[Ld<NB>::SZ], B2[Ld<NB>::SZ];
  __shared__ GJScratch<NB> g;
  const int64_t s = blockIdx.x;
  if (first[s]) return;
  const int64_t nn = (int64_t)n * n;
  copy_mat<NB>(B0, L, wch.D + s * nn, n, false, n, n);
  double* Di = inverse<NB>(B0, B1, B2, L, n, g);                              /* D^{-1} */
  double* Bw = B0;
  copy_mat<NB>(Bw, L, wch.B + s * nn, n, false, n, n);
  mm<NB>(md.F + s * nn, n, Di, L, false, Bw, L, false, n, n, n, -1.0, 0.0);   /* F = -D^{-1} B */
  mv(md.c + s * n, Di, L, false, wch.c + s * n, n, n, 1.0, 0.0);             /* c = D^{-1} c_w */
  mm<NB>(md.Q + s * nn, n, Di, L, false, Di, L, true, n, n, n, 1.0, 0.0);     /* Q = D^{-1} D^{-T} */
}

/* ------------------------------------------------------------------------ */
/* prior: information filter over steps 0..p, one workgroup per problem      */
/* ------------------------------------------------------------------------ */

template <int NB>
__global__ void __launch_bounds__(AT) ka_prior(const AsProb* probs, uk_chain_t wch, const int32_t* m_obs, AsModel md,
                                               int n, int mr) {
  constexpr int L = Ld<NB>::L;
  __shared__ double Lm[Ld<NB>::SZ], Mx[Ld<NB>::SZ], Qi[Ld<NB>::SZ], T2[Ld<NB>::SZ];
  __shared__ double et[NB], v1[NB], v2[NB];
  __shared__ int piv[NB];
  __shared__ GJScratch<NB> g;
  const AsProb pr = probs[blockIdx.x];
  const int64_t nn = (int64_t)n * n;
  set_identity<NB>(Lm, L, n, 0.0);
  for (int i = threadIdx.x; i < n; i += AT) et[i] = 0.0;
  __syncthreads();
  for (int64_t t = 0; t <= pr.p; ++t) {
    const int64_t s = pr.s0 + t;
    if (t > 0) {
      /* predict: with Mx = Lam + F^T Q^{-1} F and r = eta - F^T Q^{-1} c,
       * Lam <- Q^{-1} - Q^{-1} F Mx^{-1} F^T Q^{-1}, eta <- Q^{-1} c + Q^{-1} F Mx^{-1} r */
      const double* F = md.F + s * nn;
      copy_mat<NB>(Mx, L, md.Q + s * nn, n, false, n, n);
      lu<NB>(Mx, L, n, piv, g);
      set_identity<NB>(Qi, L, n, 1.0);
      lu_solve(Mx, L, piv, n, Qi, L, n, false);                              /* Qi = Q^{-1} */
      mm<NB>(T2, L, Qi, L, false, F, n, false, n, n, n, 1.0, 0.0);            /* T2 = Q^{-1} F */
      copy_mat<NB>(Mx, L, Lm, L, false, n, n);
      mm<NB>(Mx, L, F, n, true, T2, L, false, n, n, n, 1.0, 1.0);             /* Mx = Lam + F^T T2 */
      mv(v1, Qi, L, false, md.c + s * n, n, n, 1.0, 0.0);                    /* v1 = Q^{-1} c */
      mv(v2, F, n, true, v1, n, n, -1.0, 0.0);
      axpy(v2, et, n, 1.0);                                                  /* v2 = r */
      lu<NB>(Mx, L, n, piv, g);
      lu_solve(Mx, L, piv, n, v2, NB, 1, false);                             /* v2 = Mx^{-1} r */
      copy_mat<NB>(Lm, L, T2, L, true, n, n);                                /* Lm = T2^T */
      lu_solve(Mx, L, piv, n, Lm, L, n, false);                              /* Lm = Mx^{-1} T2^T */
      mm<NB>(Qi, L, T2, L, false, Lm, L, false, n, n, n, -1.0, 1.0);
      copy_mat<NB>(Lm, L, Qi, L, false, n, n);                               /* Lam- */
      mv(et, T2, L, false, v2, n, n, 1.0, 0.0);
      axpy(et, v1, n, 1.0);                                                  /* eta- */
    }
    const int m = m_obs[s];
    if (m > 0) { /* update: Lam += C^T C, eta += C^T o (whitened) */
      const double* Cw = wch.C + s * nn;
      mm<NB>(Lm, L, Cw, mr, true, Cw, mr, false, n, n, m, 1.0, 1.0);
      mv(et, Cw, mr, true, wch.o + s * n, n, m, 1.0, 1.0);
    }
    if (t < pr.p) {
      copy_mat<NB>(md.Lam + s * nn, n, Lm, L, false, n, n);
      for (int i = threadIdx.x; i < n; i += AT) md.eta[s * n + i] = et[i];
      __syncthreads();
    }
  }
  /* P_p = Lam^{-1}, x_p = P_p eta */
  const int64_t sp = pr.s0 + pr.p;
  copy_mat<NB>(Mx, L, Lm, L, false, n, n);
  lu<NB>(Mx, L, n, piv, g);
  set_identity<NB>(Qi, L, n, 1.0);
  lu_solve(Mx, L, piv, n, Qi, L, n, false);
  copy_mat<NB>(md.Pf + sp * nn, n, Qi, L, false, n, n);
  mv(md.xf + sp * n, Qi, L, false, et, n, n, 1.0, 0.0);
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
  constexpr int L = Ld<NB>::L;
  __shared__ double Qs[Ld<NB>::SZ], CQ[Ld<NB>::SZ], S[Ld<NB>::SZ], T[Ld<NB>::SZ];
  __shared__ double w1[NB], r[NB], rc[NB];
  __shared__ int piv[NB];
  __shared__ GJScratch<NB> g;
  const int64_t ge = blockIdx.x;
  const AsProb pr = probs[elem_prob[ge]];
  const int64_t t = ge - pr.e0 + pr.p + 1;
  const int64_t s = pr.s0 + t;
  const bool firstel = t == pr.p + 1;
  const int64_t nn = (int64_t)n * n;
  const double* F = md.F + s * nn;
  const double* c = md.c + s * n;
  double* A = el.A + ge * nn;
  double* b = el.b + ge * n;
  double* Z = el.Z + ge * nn;
  double* e = el.e + ge * n;
  double* J = el.J + ge * nn;
  const int64_t spv = pr.s0 + pr.p;

  copy_mat<NB>(Qs, L, md.Q + s * nn, n, false, n, n);                       /* Qs = Q */
  copy_vec(w1, c, n);
  if (firstel) { /* Q1 = Q + F P F^T, m1 = F x + c */
    mm<NB>(T, L, F, n, false, md.Pf + spv * nn, n, false, n, n, n, 1.0, 0.0);
    mm<NB>(Qs, L, T, L, false, F, n, true, n, n, n, 1.0, 1.0);
    mv(w1, F, n, false, md.xf + spv * n, n, n, 1.0, 1.0);
  }
  const int m = m_obs[s];
  if (m == 0) {
    for (int idx = threadIdx.x; idx < NB * NB; idx += AT) {
      const int i = idx % NB, j = idx / NB;
      if (i < n && j < n) {
        A[i + j * n] = firstel ? 0.0 : F[i + j * n];
        Z[i + j * n] = Qs[i + j * L];
        J[i + j * n] = 0.0;
      }
    }
    for (int i = threadIdx.x; i < n; i += AT) {
      b[i] = w1[i];
      e[i] = 0.0;
    }
    return;
  }
  const double* ow = wch.o + s * n;
  copy_mat<NB>(T, L, wch.C + s * nn, mr, false, m, n);                       /* T = C (m x n) */
  mm<NB>(CQ, L, T, L, false, Qs, L, false, m, n, n, 1.0, 0.0);               /* CQ = C Q */
  set_identity<NB>(S, L, m, 1.0);
  mm<NB>(S, L, CQ, L, false, T, L, true, m, m, n, 1.0, 1.0);                 /* S = C Q C^T + I */
  mv(r, T, L, false, w1, m, n, -1.0, 0.0);
  axpy(r, ow, m, 1.0);                                                       /* r = o - C w1 */
  mv(rc, T, L, false, c, m, n, -1.0, 0.0);
  axpy(rc, ow, m, 1.0);                                                      /* rc = o - C c */
  mm<NB>(J, n, T, L, false, F, n, false, m, n, n, 1.0, 0.0);                 /* C F, parked in J's slot */
  copy_mat<NB>(Z, n, Qs, L, false, n, n);                                    /* Z = Q (Qs is free now) */
  /* S can be ill-conditioned on the first element (Q1 holds the prior
   * covariance): LU solves rather than an explicit inverse */
  lu<NB>(S, L, m, piv, g);
  copy_mat<NB>(T, L, CQ, L, false, m, n);
  lu_solve(S, L, piv, m, T, L, n, false);                                    /* T = S^{-1} C Q (the gain^T) */
  mv(b, T, L, true, r, n, m, 1.0, 0.0);
  axpy(b, w1, n, 1.0);                                                       /* b = w1 + gain r */
  mm<NB>(Z, n, T, L, true, CQ, L, false, n, n, m, -1.0, 1.0);                /* Z = Q - gain C Q */
  copy_mat<NB>(CQ, L, J, n, false, m, n);                                    /* CQ <- C F */
  if (firstel) {
    for (int idx = threadIdx.x; idx < NB * NB; idx += AT) {
      const int i = idx % NB, j = idx / NB;
      if (i < n && j < n) A[i + j * n] = 0.0;
    }
    __syncthreads();
  } else {
    copy_mat<NB>(A, n, F, n, false, n, n);
    mm<NB>(A, n, T, L, true, CQ, L, false, n, n, m, -1.0, 1.0);              /* A = F - gain C F */
  }
  lu_solve(S, L, piv, m, rc, NB, 1, false);                                  /* rc = S^{-1} (o - C c) */
  mv(e, CQ, L, true, rc, n, m, 1.0, 0.0);                                    /* e = (C F)^T S^{-1} rc */
  copy_mat<NB>(T, L, CQ, L, false, m, n);
  lu_solve(S, L, piv, m, T, L, n, false);                                    /* T = S^{-1} C F */
  mm<NB>(J, n, CQ, L, true, T, L, false, n, n, m, 1.0, 0.0);                 /* J = (C F)^T S^{-1} C F */
}

/* ------------------------------------------------------------------------ */
/* combines (c :889-1016, matlab :240-262)                                   */
/* ------------------------------------------------------------------------ */

/* blocks of an accumulator or of the work area (LDS, or global for the
 * filtering work of the 64 bucket), leading dimension Ld<NB>::L */
struct Blk {
  double *A, *b, *Z, *e, *J;
};
template <int NB>
struct Work {
  double *M, *W1, *W2, *R, *T, *v1, *v2;
  GJScratch<NB>* g;
};

/* a <- a (x) element g, filtering:
 *   M = I + Z_i J_j,  Y^T = M^{-1} A_i,  X = A_j M^{-1},
 *   e = Y (e_j - J_j b_i) + e_i,  J = Y J_j A_i + J_i,
 *   b = X (b_i + Z_i e_j) + b_j,  Z = X Z_i A_j^T + Z_j,  A = X A_i.
 * Element g's blocks are read from HBM into registers first (all loads in
 * flight together) and written into LDS where they are needed. */
template <int NB>
__device__ void filt_combine(const Blk& a, const Work<NB>& w, const AsElems& el, int64_t g, int n) {
  constexpr int L = Ld<NB>::L;
  const int64_t nn = (int64_t)n * n;
  const double *bj = el.b + g * n, *ej = el.e + g * n;
  /* Up to NB = 48 all three blocks are requested at once; the 64 bucket
   * fetches A_j and Z_j where they are used, to stay within its registers. */
  constexpr bool EARLY = NB <= 48;
  RegBlk<NB> rA, rZ, rJ;
  rJ.load(el.J + g * nn, n);
  if (EARLY) {
    rA.load(el.A + g * nn, n);
    rZ.load(el.Z + g * nn, n);
  }
  rJ.store(w.T, L, n, false);                                                /* T = J_j */
  set_identity<NB>(w.M, L, n, 1.0);
  mm<NB>(w.M, L, a.Z, L, false, w.T, L, false, n, n, n, 1.0, 1.0);           /* M = I + Z_i J_j */
  mv(w.v2, w.T, L, false, a.b, n, n, -1.0, 0.0);
  axpy(w.v2, ej, n, 1.0);                                                    /* v2 = e_j - J_j b_i */
  mm<NB>(w.T, L, w.T, L, false, a.A, L, false, n, n, n, 1.0, 0.0);           /* T = J_j A_i */
  double* Mi = inverse<NB>(w.M, w.W1, w.W2, L, n, *w.g);                     /* M^{-1} */
  double* F = w.M;
  mm<NB>(F, L, Mi, L, false, a.A, L, false, n, n, n, 1.0, 0.0);              /* F = Y^T */
  mv(a.e, F, L, true, w.v2, n, n, 1.0, 1.0);                                 /* e */
  mm<NB>(a.J, L, F, L, true, w.T, L, false, n, n, n, 1.0, 1.0);              /* J */
  if (!EARLY) rA.load(el.A + g * nn, n);
  rA.store(w.R, L, n, false);                                                /* R = A_j */
  __syncthreads();
  mm<NB>(w.T, L, w.R, L, false, Mi, L, false, n, n, n, 1.0, 0.0);            /* T = X */
  mv(w.v1, a.Z, L, false, ej, n, n, 1.0, 0.0);
  axpy(w.v1, a.b, n, 1.0);
  mv(a.b, w.T, L, false, w.v1, n, n, 1.0, 0.0);
  axpy(a.b, bj, n, 1.0);                                                     /* b */
  mm<NB>(F, L, w.T, L, false, a.Z, L, false, n, n, n, 1.0, 0.0);             /* F = X Z_i */
  if (!EARLY) rZ.load(el.Z + g * nn, n);
  rZ.store(a.Z, L, n, false);                                                /* Z = Z_j */
  __syncthreads();
  mm<NB>(a.Z, L, F, L, false, w.R, L, true, n, n, n, 1.0, 1.0);              /* Z */
  mm<NB>(a.A, L, w.T, L, false, a.A, L, false, n, n, n, 1.0, 0.0);           /* A */
}

/* a <- a (x) element g, smoothing; a is the running suffix and g the next
 * earlier step: E = E_j E_i, g = E_j g_i + g_j, L = E_j L_i E_j^T + L_j. */
template <int NB>
__device__ void smooth_combine(const Blk& a, const Work<NB>& w, const AsElems& el, int64_t g, int n) {
  constexpr int L = Ld<NB>::L;
  const int64_t nn = (int64_t)n * n;
  const double* gj = el.b + g * n;
  RegBlk<NB> rE, rL;
  rE.load(el.A + g * nn, n);
  rL.load(el.Z + g * nn, n);
  rE.store(w.W1, L, n, false);                                               /* W1 = E_j */
  __syncthreads();
  mm<NB>(a.A, L, w.W1, L, false, a.A, L, false, n, n, n, 1.0, 0.0);
  mv(a.b, w.W1, L, false, a.b, n, n, 1.0, 0.0);
  axpy(a.b, gj, n, 1.0);
  mm<NB>(w.T, L, w.W1, L, false, a.Z, L, false, n, n, n, 1.0, 0.0);
  rL.store(a.Z, L, n, false);
  __syncthreads();
  mm<NB>(a.Z, L, w.T, L, false, w.W1, L, true, n, n, n, 1.0, 1.0);
}

template <int NB>
__device__ void blk_load(const Blk& a, int ld, const AsElems& el, int64_t g, int n, bool filt) {
  const int64_t nn = (int64_t)n * n;
  RegBlk<NB> r;
  r.load(el.A + g * nn, n);
  r.store(a.A, ld, n, false);
  r.load(el.Z + g * nn, n);
  r.store(a.Z, ld, n, false);
  if (filt) {
    r.load(el.J + g * nn, n);
    r.store(a.J, ld, n, false);
  }
  for (int i = threadIdx.x; i < n; i += AT) {
    a.b[i] = el.b[g * n + i];
    if (filt) a.e[i] = el.e[g * n + i];
  }
  __syncthreads();
}

template <int NB>
__device__ void blk_store(const Blk& a, int ld, const AsElems& el, int64_t g, int n, bool filt) {
  const int64_t nn = (int64_t)n * n;
  for (int idx = threadIdx.x; idx < NB * NB; idx += AT) {
    const int i = idx % NB, j = idx / NB;
    if (i < n && j < n) {
      el.A[g * nn + i + j * n] = a.A[i + j * ld];
      el.Z[g * nn + i + j * n] = a.Z[i + j * ld];
      if (filt) el.J[g * nn + i + j * n] = a.J[i + j * ld];
    }
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
  static constexpr int SZ = Ld<NB>::SZ;
  static constexpr int ACC = FILT ? 3 * SZ + 2 * NB : 2 * SZ + NB;
  static constexpr int WORK = FILT ? 5 * SZ + 2 * NB : 2 * SZ;             /* M, W1, W2, R, T, v1, v2 | W1, T */
  static constexpr bool WORK_GLOBAL = (ACC + WORK) * 8 + NB * 4 + 256 > LDS_BYTES;
  static constexpr int LDS = ACC + (WORK_GLOBAL ? 0 : WORK);
};

template <int NB, bool FILT>
__device__ void scan_setup(double* lds, double* gwork, Blk& a, Work<NB>& w) {
  using SP = ScanPlan<NB, FILT>;
  constexpr int SZ = SP::SZ;
  double* q = lds;
  a.A = q; q += SZ;
  a.Z = q; q += SZ;
  a.J = nullptr;
  a.e = nullptr;
  if (FILT) { a.J = q; q += SZ; }
  a.b = q; q += NB;
  if (FILT) { a.e = q; q += NB; }
  double* p = SP::WORK_GLOBAL ? gwork + (int64_t)blockIdx.x * SP::WORK : q;
  w.W1 = p; p += SZ;
  w.T = p; p += SZ;
  w.M = w.W2 = w.R = w.v1 = w.v2 = nullptr;
  if (FILT) {
    w.M = p; p += SZ;
    w.W2 = p; p += SZ;
    w.R = p; p += SZ;
    w.v1 = p; p += NB;
    w.v2 = p; p += NB;
  }
}

template <int NB, bool FILT>
__device__ __forceinline__ void combine(const Blk& a, const Work<NB>& w, const AsElems& el, int64_t g, int n) {
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
  __shared__ GJScratch<NB> gjs;
  Blk a;
  Work<NB> w;
  scan_setup<NB, FILT>(lds, gwork, a, w);
  w.g = &gjs;
  const Chunk c = ch[blockIdx.x];
  blk_load<NB>(a, Ld<NB>::L, el, c.first, n, FILT);
  for (int i = 1; i < c.len; ++i) combine<NB, FILT>(a, w, el, c.first + (int64_t)i * c.dir, n);
  blk_store<NB>(a, Ld<NB>::L, tot, c.tot, n, FILT);
}

/* pass 2, one Hillis-Steele round over the chunk totals of every problem:
 * out[t] = in[t - d] (x) in[t] when rank >= d, else in[t]. */
template <int NB, bool FILT>
__global__ void __launch_bounds__(AT) ka_scan_round(const Chunk* ch, AsElems in, AsElems out, int d, int n,
                                                    double* gwork) {
  using SP = ScanPlan<NB, FILT>;
  __shared__ double lds[SP::LDS];
  __shared__ GJScratch<NB> gjs;
  Blk a;
  Work<NB> w;
  scan_setup<NB, FILT>(lds, gwork, a, w);
  w.g = &gjs;
  const Chunk c = ch[blockIdx.x];
  if (c.rank >= d) {
    blk_load<NB>(a, Ld<NB>::L, in, c.tot - d, n, FILT);
    combine<NB, FILT>(a, w, in, c.tot, n);
  } else {
    blk_load<NB>(a, Ld<NB>::L, in, c.tot, n, FILT);
  }
  blk_store<NB>(a, Ld<NB>::L, out, c.tot, n, FILT);
}

/* pass 3: every prefix within a chunk, seeded with the inclusive total of
 * the chunk before it; results overwrite the elements in place. */
template <int NB, bool FILT>
__global__ void __launch_bounds__(AT) ka_scan_down(const Chunk* ch, AsElems el, AsElems tot, int n, double* gwork) {
  using SP = ScanPlan<NB, FILT>;
  __shared__ double lds[SP::LDS];
  __shared__ GJScratch<NB> gjs;
  Blk a;
  Work<NB> w;
  scan_setup<NB, FILT>(lds, gwork, a, w);
  w.g = &gjs;
  const Chunk c = ch[blockIdx.x];
  int i0 = 0;
  if (c.rank == 0) {
    blk_load<NB>(a, Ld<NB>::L, el, c.first, n, FILT);
    i0 = 1;
  } else {
    blk_load<NB>(a, Ld<NB>::L, tot, c.tot - 1, n, FILT);
  }
  for (int i = i0; i < c.len; ++i) {
    const int64_t g = c.first + (int64_t)i * c.dir;
    combine<NB, FILT>(a, w, el, g, n);
    blk_store<NB>(a, Ld<NB>::L, el, g, n, FILT);
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
  constexpr int L = Ld<NB>::L;
  __shared__ double M[Ld<NB>::SZ], T[Ld<NB>::SZ], U[Ld<NB>::SZ];
  __shared__ double v1[NB], v2[NB];
  __shared__ int piv[NB];
  __shared__ GJScratch<NB> g;
  const int64_t s = blockIdx.x;
  const AsProb pr = probs[step_prob[s]];
  const int64_t t = s - pr.s0;
  const int64_t nn = (int64_t)n * n;
  double* E = el.A + s * nn;
  double* gv = el.b + s * n;
  double* Lo = el.Z + s * nn;
  if (t == pr.k - 1) {
    for (int idx = threadIdx.x; idx < n * n; idx += AT) {
      E[idx] = 0.0;
      Lo[idx] = md.Pf[s * nn + idx];
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
    copy_mat<NB>(U, L, P, n, false, n, n);                                   /* U = P */
    copy_mat<NB>(M, L, F, n, false, n, n);                                   /* M = F */
    mm<NB>(T, L, U, L, false, M, L, true, n, n, n, 1.0, 0.0);                /* T = P F^T */
    mv(v1, M, L, false, x, n, n, 1.0, 0.0);
    axpy(v1, c, n, 1.0);                                                     /* v1 = F x + c */
    mm<NB>(U, L, M, L, false, T, L, false, n, n, n, 1.0, 0.0);               /* U = F P F^T */
    copy_mat<NB>(M, L, Q, n, false, n, n);
    for (int idx = threadIdx.x; idx < NB * NB; idx += AT) {
      const int i = idx % NB, j = idx / NB;
      if (i < n && j < n) M[i + j * L] += U[i + j * L];                      /* M = F P F^T + Q */
    }
    __syncthreads();
    lu<NB>(M, L, n, piv, g);
    copy_mat<NB>(U, L, T, L, true, n, n);                                    /* U = F P */
    lu_solve(M, L, piv, n, U, L, n, true);                                   /* U = E^T */
    copy_mat<NB>(E, n, U, L, true, n, n);
    mv(gv, U, L, true, v1, n, n, -1.0, 0.0);
    axpy(gv, x, n, 1.0);                                                     /* g = x - E (F x + c) */
    copy_mat<NB>(M, L, T, L, true, n, n);                                    /* M = F P */
    copy_mat<NB>(Lo, n, P, n, false, n, n);
    mm<NB>(Lo, n, U, L, true, M, L, false, n, n, n, -1.0, 1.0);              /* L = P - E F P */
  } else {
    copy_mat<NB>(M, L, Q, n, false, n, n);
    lu<NB>(M, L, n, piv, g);
    copy_mat<NB>(T, L, F, n, false, n, n);
    lu_solve(M, L, piv, n, T, L, n, false);                                  /* T = Q^{-1} F */
    copy_vec(v1, c, n);
    lu_solve(M, L, piv, n, v1, NB, 1, false);                                /* v1 = Q^{-1} c */
    copy_mat<NB>(U, L, md.Lam + s * nn, n, false, n, n);
    mm<NB>(U, L, F, n, true, T, L, false, n, n, n, 1.0, 1.0);                /* U = Lam + F^T Q^{-1} F */
    lu<NB>(U, L, n, piv, g);
    set_identity<NB>(M, L, n, 1.0);
    lu_solve(U, L, piv, n, M, L, n, false);                                  /* M = L */
    copy_mat<NB>(Lo, n, M, L, false, n, n);
    mm<NB>(E, n, M, L, false, T, L, true, n, n, n, 1.0, 0.0);                /* E = L (Q^{-1} F)^T */
    mv(v2, F, n, true, v1, n, n, -1.0, 0.0);
    axpy(v2, md.eta + s * n, n, 1.0);
    mv(gv, M, L, false, v2, n, n, 1.0, 0.0);                                 /* g = L (eta - F^T Q^{-1} c) */
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
extern "C" void uk_batch_chain0(ukhip_batch* b, uk_chain_t* ch, const int32_t** m0);
extern "C" hipError_t uk_batch_weigh_dense0(ukhip_batch* b, hipStream_t st);

namespace {

bool ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  uk_set_error("%s: %s", what, hipGetErrorString(e));
  return false;
}

/* Event timing of one associative smoothing (ukhip_batch_smooth_associative_timed):
 * an event after each kernel class, in launch order; slot i's time is
 * ev[i] -> ev[i + 1]. Set for the duration of one timed call on this thread. */
struct AsTimer {
  hipEvent_t ev[UKHIP_AS_SLOTS + 1]{};
  int32_t launches[UKHIP_AS_SLOTS]{};
  double ms[UKHIP_AS_SLOTS]{};
};
thread_local AsTimer* t_timer = nullptr;

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

/* chunk length: about CHUNKS_TARGET chunks over the batch (two workgroups per
 * compute unit at one workgroup per unit), at least 4 */
constexpr int64_t CHUNKS_TARGET = 512;

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
  if (!xout || !Sout) return -1;
  const int n = raw->n;
  uk_chain_t wch;
  const int32_t* m_obs = nullptr;
  uk_batch_chain0(b, &wch, &m_obs);
  const int mr = wch.mr;

  /* plan: p of every problem from the observation counts */
  std::vector<int32_t> m_h(S);
  if (!ok(hipMemcpy(m_h.data(), m_obs, S * sizeof(int32_t), hipMemcpyDeviceToHost), "download m")) return -1;
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
  AsTimer* tm = t_timer;
  auto mark = [&](int slot, int launches) { /* slot's launches are done: its closing event */
    if (!tm) return;
    tm->launches[slot] += launches;
    (void)hipEventRecord(tm->ev[slot + 1], st);
  };
  if (tm) (void)hipEventRecord(tm->ev[0], st);
  if (!ok(uk_batch_weigh_dense0(b, st), "k_weigh")) return -1;
  mark(UKHIP_AS_WEIGH, 1);
  ka_model<NB><<<dim3((unsigned)S), blk, 0, st>>>(wch, raw->first, md, n);
  mark(UKHIP_AS_MODEL, 1);
  ka_prior<NB><<<dim3((unsigned)nprob), blk, 0, st>>>(d_probs, wch, m_obs, md, n, mr);
  mark(UKHIP_AS_PRIOR, 1);
  if (ne > 0) {
    const unsigned nc = (unsigned)fch.size();
    ka_felem<NB><<<dim3((unsigned)ne), blk, 0, st>>>(d_probs, d_elem_prob, wch, m_obs, md, fe, n, mr);
    mark(UKHIP_AS_FELEM, 1);
    ka_scan_reduce<NB, true><<<dim3(nc), blk, 0, st>>>(d_fch, fe, ft[0], n, gw);
    int cur = 0, rounds = 0;
    for (int64_t d = 1; d < maxr_f; d <<= 1) {
      ka_scan_round<NB, true><<<dim3(nc), blk, 0, st>>>(d_fch, ft[cur], ft[cur ^ 1], (int)d, n, gw);
      cur ^= 1;
      ++rounds;
    }
    ka_scan_down<NB, true><<<dim3(nc), blk, 0, st>>>(d_fch, fe, ft[cur], n, gw);
    mark(UKHIP_AS_FSCAN, 2 + rounds);
    ka_filtered<<<dim3((unsigned)ne), blk, 0, st>>>(d_probs, d_elem_prob, fe, md, n);
    mark(UKHIP_AS_FILTERED, 1);
  } else {
    mark(UKHIP_AS_FELEM, 0);
    mark(UKHIP_AS_FSCAN, 0);
    mark(UKHIP_AS_FILTERED, 0);
  }
  {
    const unsigned nc = (unsigned)sch.size();
    ka_selem<NB><<<dim3((unsigned)S), blk, 0, st>>>(d_probs, d_step_prob, md, se, n);
    mark(UKHIP_AS_SELEM, 1);
    ka_scan_reduce<NB, false><<<dim3(nc), blk, 0, st>>>(d_sch, se, stt[0], n, gw);
    int cur = 0, rounds = 0;
    for (int64_t d = 1; d < maxr_s; d <<= 1) {
      ka_scan_round<NB, false><<<dim3(nc), blk, 0, st>>>(d_sch, stt[cur], stt[cur ^ 1], (int)d, n, gw);
      cur ^= 1;
      ++rounds;
    }
    ka_scan_down<NB, false><<<dim3(nc), blk, 0, st>>>(d_sch, se, stt[cur], n, gw);
    mark(UKHIP_AS_SSCAN, 2 + rounds);
    ka_results<<<dim3((unsigned)S), blk, 0, st>>>(se, xout, Sout, n);
    mark(UKHIP_AS_RESULTS, 1);
  }
  if (!ok(hipGetLastError(), "associative launches")) return -1;
  /* the host tables above must outlive the asynchronous upload */
  if (!ok(hipStreamSynchronize(st), "associative smoothing")) return -1;
  if (tm) {
    for (int i = 0; i < UKHIP_AS_TOTAL; ++i) {
      float f = 0.f;
      if (!ok(hipEventElapsedTime(&f, tm->ev[i], tm->ev[i + 1]), "associative timing")) return -1;
      tm->ms[i] += f;
    }
    float f = 0.f;
    if (!ok(hipEventElapsedTime(&f, tm->ev[0], tm->ev[UKHIP_AS_TOTAL]), "associative timing")) return -1;
    tm->ms[UKHIP_AS_TOTAL] += f;
  }
  return 0;
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
  if (!xout || !Sout) return -1; /* the output storage could not be allocated (error recorded) */
  uk_chain_t ch0;
  const int32_t* m0;
  uk_batch_chain0(b, &ch0, &m0);
  switch (uk_bucket(raw->n > ch0.mr ? raw->n : ch0.mr)) {
    case 8: return run_assoc<8>(b);
    case 16: return run_assoc<16>(b);
    case 32: return run_assoc<32>(b);
    case 48: return run_assoc<48>(b);
    case 64: return run_assoc<64>(b);
    default: uk_set_error("associative: state dimension %d is above 64", raw->n); return -1;
  }
}

/* ukhip_batch_smooth_associative_timed (include/ukhip.h): parts one after
 * another, each with its events; the slot sums are returned. */
extern "C" int uk_assoc_smooth_timed(struct ukhip_batch* b, uint32_t options, double* ms, int32_t* launches) {
  AsTimer tm;
  for (auto& e : tm.ev)
    if (!ok(hipEventCreate(&e), "hipEventCreate")) {
      for (auto& x : tm.ev)
        if (x) (void)hipEventDestroy(x);
      return -1;
    }
  t_timer = &tm;
  const int rc = uk_assoc_smooth(b, options);
  t_timer = nullptr;
  for (auto& e : tm.ev) (void)hipEventDestroy(e);
  if (rc != 0) return rc;
  for (int i = 0; i < UKHIP_AS_SLOTS; ++i) {
    if (ms) ms[i] += tm.ms[i];
    if (launches) launches[i] += tm.launches[i];
  }
  return 0;
}
