/*
 * refine.hip -- the mixed-precision odd-even smoother: the elimination runs in
 * FP32 (k_phase1..3<NB, float>), and FP64 accuracy of the estimates is
 * recovered by iterative refinement on the least-squares problem
 * min || A u - b || of the whitened step equations (the system the
 * reference's odd-even smoother solves, c/kalman_oddeven_smoother.c:379-975).
 *
 * The odd-even elimination is an orthogonal factorization of the permuted
 * block matrix, A P = Q Rs, Rs block upper triangular with one block row
 * per even step j of every level l:  R_j u_j + B~_j u_{j-1} + Y_j u_{j+1}
 * (u_{j-1}, u_{j+1} are the odd steps, eliminated at later levels). So
 * Rs^T Rs = P^T A^T A P, and with the FP32 factors one refinement step is a
 * corrected semi-normal-equations step:
 *   r = b - A x                (FP64, whitened chain kept in FP64)
 *   g = A^T r                  (k_normal)
 *   Rs^T y = g                 (forward, level 0 -> top: k_fwd_even, k_fwd_odd)
 *   Rs d = y                   (backward, top -> level 0: k_solve_est)
 *   x <- x + d                 (k_add)
 * The sweeps compute in FP64 and read the FP32 factors. The first x comes
 * from the FP32 transformed right-hand side (k_solve_est on cf).
 * Each step multiplies the error by about cond(A)^2 eps_32.
 */
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ukhip_internal.h"

namespace {

template <class T>
struct FactorPtrs {
  const T* R;
  const T* Bt;
  const T* Y;
};

/* v of lane `lane`, broadcast to the wave (two readlanes, no LDS) */
__device__ __forceinline__ double lane_bcast(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

/* The sweeps at NB <= 32 give each step a segment of S lanes (S = 16 or 32,
 * 64 / S steps per wave). seg_bcast<S>(v, q): lane q of each segment,
 * broadcast to that segment (q folds to a constant in the unrolled loops):
 * S = 32 reads both halves' lane q and each half keeps its own; S = 16 is the
 * DPP row broadcast (row_newbcast, a row being 16 lanes). */
template <int Q>
__device__ __forceinline__ int row_newbcast(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x150 + Q, 0xf, 0xf, false);
}
__device__ __forceinline__ int row_bcast16(int v, int q) {
  switch (q & 15) {
    case 0: return row_newbcast<0>(v);
    case 1: return row_newbcast<1>(v);
    case 2: return row_newbcast<2>(v);
    case 3: return row_newbcast<3>(v);
    case 4: return row_newbcast<4>(v);
    case 5: return row_newbcast<5>(v);
    case 6: return row_newbcast<6>(v);
    case 7: return row_newbcast<7>(v);
    case 8: return row_newbcast<8>(v);
    case 9: return row_newbcast<9>(v);
    case 10: return row_newbcast<10>(v);
    case 11: return row_newbcast<11>(v);
    case 12: return row_newbcast<12>(v);
    case 13: return row_newbcast<13>(v);
    case 14: return row_newbcast<14>(v);
    default: return row_newbcast<15>(v);
  }
}
template <int S>
__device__ __forceinline__ double seg_bcast(double v, int q) {
  if constexpr (S == 64) {
    return lane_bcast(v, q);
  } else if constexpr (S == 32) {
    const double a = lane_bcast(v, q), b = lane_bcast(v, 32 + q);
    return (threadIdx.x & 32) ? b : a;
  } else {
    static_assert(S == 16, "segments of 16, 32 or 64 lanes");
    return __hiloint2double(row_bcast16(__double2hiint(v), q), row_bcast16(__double2loint(v), q));
  }
}
/* lanes per step of the sweeps in bucket NB: 16 up to NB = 16 (four steps per
 * wave), else the whole wave. Segments of 32 lanes (two steps per wave at
 * NB = 24 and 32) were measured slower on MI355X: each broadcast of the
 * substitutions' serial chain becomes two readlane pairs and a select, and
 * the doubled column loads per wave cut the resident waves. */
#ifndef UK_SWEEP_SEG32
#define UK_SWEEP_SEG32 0
#endif
__host__ __device__ constexpr int sweep_seg(int NB) { return NB <= 16 ? 16 : (NB <= 32 && UK_SWEEP_SEG32 ? 32 : 64); }

__device__ __forceinline__ int64_t find_problem(const int64_t* start, int nprob, int64_t x) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
/* The problem of step x = x0 + d (0 <= d < G) when x0 is the same for the
 * whole wave: the binary search runs once, on the scalar unit, and each
 * segment walks at most G - 1 problems on (every problem has a step). */
template <int G>
__device__ __forceinline__ int find_problem_seg(const int64_t* start, int nprob, int64_t x0, int64_t x) {
  int p = (int)find_problem(start, nprob, x0);
#pragma unroll
  for (int t = 0; t < G - 1; ++t)
    if (p + 1 < nprob && start[p + 1] <= x) ++p;
  return p;
}

/* Row and column of element idx = lane + 64 t (t = 0, 1, ...) of a
 * column-major block with `rows` rows: one division per lane, then 64
 * elements further per step without one (a division by a run-time value is
 * a long instruction sequence). */
struct Walk64 {
  int r, c, dr, dc, rows;
  /* stride: elements per step (64, or a segment's lane count) */
  __device__ __forceinline__ Walk64(int lane, int rows_, int stride = 64) : rows(rows_) {
    dc = stride / rows;
    dr = stride - dc * rows;
    c = lane / rows;
    r = lane - c * rows;
  }
  __device__ __forceinline__ void next() {
    r += dr;
    c += dc;
    if (r >= rows) r -= rows, ++c;
  }
};

/* dst[i] = (float) src[i] */
__global__ void __launch_bounds__(256) k_to32(float* __restrict__ dst, const double* __restrict__ src, int64_t count) {
  for (int64_t i = blockIdx.x * 256LL + threadIdx.x; i < count; i += (int64_t)gridDim.x * 256) dst[i] = (float)src[i];
}

/* dst[i] = src[i] widened (FP32 chain blocks entering the FP64 top levels) */
__global__ void __launch_bounds__(256) k_to64(double* __restrict__ dst, const float* __restrict__ src, int64_t count) {
  for (int64_t i = blockIdx.x * 256LL + threadIdx.x; i < count; i += (int64_t)gridDim.x * 256) dst[i] = (double)src[i];
}

/* x[i] += d[i] */
__global__ void __launch_bounds__(256) k_add(double* __restrict__ x, const double* __restrict__ d, int64_t count) {
  for (int64_t i = blockIdx.x * 256LL + threadIdx.x; i < count; i += (int64_t)gridDim.x * 256) x[i] += d[i];
}

/*
 * Residual and normal-equation right-hand side of step s (one workgroup):
 *   re_s = c_s - B_s x_{s-1} - D_s x_s   (not on a problem's first step)
 *   ro_s = o_s - C_s x_s
 *   gA_s = D_s^T re_s + C_s^T ro_s,   gB_s = B_s^T re_s (belongs to g_{s-1}).
 * Blocks are staged through LDS (odd leading dimension) so that HBM is read
 * along columns and the transposed products read LDS.
 */
template <int NB, bool DD>
__global__ void __launch_bounds__(256) k_normal(uk_chain_t ch, const int8_t* first, const int32_t* m_obs,
                                                const double* __restrict__ x, double* __restrict__ gA,
                                                double* __restrict__ gB, int n) {
  constexpr int L = NB + 1;
  /* DD (the chain stores D as its diagonal): one block, B and then C; the
   * diagonal's zeros are formed in the products, so the arithmetic is the
   * same as with D staged */
  __shared__ double M2[NB * L];
  __shared__ double M1s[DD ? 1 : NB * L];
  double* const M1 = DD ? M2 : M1s;
  __shared__ double xs[NB], xm[NB], re[NB], ro[NB], dd[NB];
  const int tid = threadIdx.x;
  const int64_t s = blockIdx.x;
  const int64_t nn = (int64_t)n * n;
  const int mr = ch.mr;
  const bool evo = first[s] == 0;
  const int m = m_obs[s];
  /* Every block element of the step is read into registers first, at indices
   * clamped into the step's slots (an element outside the block is read but
   * never stored), so the loads go out together instead of one loop
   * iteration's latency after another. C is read before it is staged. */
  constexpr int PS = (NB * NB + 255) / 256;
  const int nb2 = n * n, cn = mr * n;
  double tcv[PS], tbv[PS], tdv[DD ? 1 : PS];
#pragma unroll
  for (int t = 0; t < PS; ++t) {
    const int idx = tid + 256 * t;
    tcv[t] = ch.C[s * nn + (idx < cn ? idx : 0)];
  }
  double ddv = 0.0;
  if (evo) {
    if (ch.b_synth) {
      const double* wFs = ch.wF + (ch.wFidx ? ch.wFidx[s] * nn : 0);
#pragma unroll
      for (int t = 0; t < PS; ++t) {
        const int idx = tid + 256 * t, ic = idx < nb2 ? idx : 0;
        tbv[t] = -ch.wk[s * n + ic % n] * wFs[ic];
      }
    } else {
#pragma unroll
      for (int t = 0; t < PS; ++t) {
        const int idx = tid + 256 * t;
        tbv[t] = ch.B[s * nn + (idx < nb2 ? idx : 0)];
      }
    }
    if (DD) {
      ddv = ch.D[s * nn + (tid < n ? tid : 0)];
    } else {
#pragma unroll
      for (int t = 0; t < (DD ? 1 : PS); ++t) {
        const int idx = tid + 256 * t;
        tdv[t] = ch.D[s * nn + (idx < nb2 ? idx : 0)];
      }
    }
  }
  if (tid < n) {
    xs[tid] = x[s * n + tid];
    xm[tid] = evo ? x[(s - 1) * n + tid] : 0.0;
  }
  if (evo) {
#pragma unroll
    for (int t = 0; t < PS; ++t) {
      const int idx = tid + 256 * t;
      if (idx < nb2) {
        const int i = idx % n, j = idx / n;
        if (!DD) M1[i + j * L] = tdv[DD ? 0 : t];
        M2[i + j * L] = tbv[t];
      }
    }
    if (DD && tid < n) dd[tid] = ddv;
  }
  __syncthreads();
  /* D's entry (i, j): from the staged block, or the diagonal */
  auto Dij = [&](int i, int j) -> double { return DD ? (i == j ? dd[i] : 0.0) : M1[i + j * L]; };
  if (tid < n) {
    double r = 0.0;
    if (evo) {
      r = ch.c[s * n + tid];
      for (int j = 0; j < n; ++j) r -= Dij(tid, j) * xs[j] + M2[tid + j * L] * xm[j];
    }
    re[tid] = r;
  }
  __syncthreads();
  double ga = 0.0, gb = 0.0;
  if (tid < n && evo)
    for (int r = 0; r < n; ++r) {
      ga = fma(Dij(r, tid), re[r], ga);
      gb = fma(M2[r + tid * L], re[r], gb);
    }
  __syncthreads(); /* B consumed: with DD its block takes C */
  if (m > 0) {
#pragma unroll
    for (int t = 0; t < PS; ++t) {
      const int idx = tid + 256 * t;
      if (idx < cn) {
        const int i = idx % mr, j = idx / mr;
        M1[i + j * L] = tcv[t]; /* C is mr x n, ld mr; rows >= m are zero */
      }
    }
    __syncthreads();
    if (tid < mr) {
      double r = ch.o[s * n + tid];
      for (int j = 0; j < n; ++j) r -= M1[tid + j * L] * xs[j];
      ro[tid] = tid < m ? r : 0.0;
    }
    __syncthreads();
    if (tid < n)
      for (int r = 0; r < mr; ++r) ga = fma(M1[r + tid * L], ro[r], ga);
  }
  if (tid < n) {
    gA[s * n + tid] = ga;
    gB[s * n + tid] = gb;
  }
}

/* Orders one wave's LDS accesses (its writes before its later reads of other
 * lanes' values); the compiler moves no access across it. */
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/*
 * The same quantities as k_normal for NB <= 32, one wave per step and four
 * steps per workgroup (no workgroup barrier: each wave orders its own LDS
 * staging). Every product runs on all 64 lanes: lane i + 32 h (h = 0, 1)
 * takes entry i of the result over half h of the summation range, and the
 * halves are added across the wave.
 */
template <int NB, bool DD>
__global__ void __launch_bounds__(256) k_normal_w(uk_chain_t ch, const int8_t* first, const int32_t* m_obs,
                                                  const double* __restrict__ x, double* __restrict__ gA,
                                                  double* __restrict__ gB, int n, int64_t steps) {
  static_assert(NB <= 32, "one half-wave per row or column");
  constexpr int L = NB + 1, WPB = 4, PS = (NB * NB + 63) / 64;
  __shared__ double Ms[WPB][NB * L];            /* B, then C */
  __shared__ double Ds[WPB][DD ? 1 : NB * L];   /* D when it is stored in full */
  __shared__ double vec[WPB][5][NB];            /* x_s, x_{s-1}, re, ro, diag(D) */
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, h = lane >> 5, i = lane & 31;
  const int64_t s = (int64_t)blockIdx.x * WPB + w;
  if (s >= steps) return; /* uniform per wave */
  double* M = Ms[w];
  double* Dm = Ds[w];
  double* xs = vec[w][0];
  double* xm = vec[w][1];
  double* re = vec[w][2];
  double* ro = vec[w][3];
  double* dd = vec[w][4];
  const int64_t nn = (int64_t)n * n;
  const int mr = ch.mr;
  const bool evo = first[s] == 0;
  const int m = m_obs[s];
  const int nb2 = n * n, cn = mr * n;
  /* every element of the step's blocks in flight at once (indices clamped
   * into the step's slots; values outside a block are never stored) */
  double tcv[PS], tbv[PS], tdv[DD ? 1 : PS];
  if (ch.from_raw) { /* C = diag(w) G formed here, the product k_weigh_diag forms */
    Walk64 wr(lane, mr);
#pragma unroll
    for (int t = 0; t < PS; ++t) {
      const int idx = lane + 64 * t, ic = idx < cn ? idx : 0, r = idx < cn ? wr.r : 0;
      const double cw = ch.rC[s * mr + r], gv = ch.rG[s * cn + ic];
      tcv[t] = r < m ? cw * gv : 0.0;
      wr.next();
    }
  } else {
#pragma unroll
    for (int t = 0; t < PS; ++t) {
      const int idx = lane + 64 * t;
      tcv[t] = ch.C[s * nn + (idx < cn ? idx : 0)];
    }
  }
  if (evo) {
    if (ch.b_synth) {
      const double* wFs = ch.wF + (ch.wFidx ? ch.wFidx[s] * nn : 0);
      Walk64 wb(lane, n);
#pragma unroll
      for (int t = 0; t < PS; ++t) {
        const int idx = lane + 64 * t, ic = idx < nb2 ? idx : 0;
        tbv[t] = -ch.wk[s * n + wb.r] * wFs[ic]; /* row wb.r of B (any row for idx >= n^2) */
        wb.next();
      }
    } else {
#pragma unroll
      for (int t = 0; t < PS; ++t) {
        const int idx = lane + 64 * t;
        tbv[t] = ch.B[s * nn + (idx < nb2 ? idx : 0)];
      }
    }
    if (DD) {
      if (lane < n) dd[lane] = ch.from_raw ? ch.wk[s * n + lane] : ch.D[s * nn + lane];
    } else {
#pragma unroll
      for (int t = 0; t < (DD ? 1 : PS); ++t) {
        const int idx = lane + 64 * t;
        tdv[t] = ch.D[s * nn + (idx < nb2 ? idx : 0)];
      }
    }
  }
  if (lane < n) {
    xs[lane] = x[s * n + lane];
    xm[lane] = evo ? x[(s - 1) * n + lane] : 0.0;
  }
  /* summation range of this half: [j0, j0 + cn_h) of n, [r0, r0 + cm_h) of mr.
   * The loops run a compile-time number of iterations (all LDS reads issue
   * back to back) at indices clamped into the staged arrays; iterations past
   * the range leave the sums as they are (a select, so stale LDS values are
   * never used). */
  constexpr int HB = (NB + 1) / 2;
  const int hn = (n + 1) >> 1, hm = (mr + 1) >> 1;
  const int j0 = h * hn, cn_h = h ? n - hn : hn;
  const int r0 = h * hm, cm_h = h ? mr - hm : hm;
  const int ir = i < n ? i : 0;   /* row / column index of this lane, clamped */
  const int irc = i < mr ? i : 0; /* row of C, clamped */
  auto cl = [](int v) { return v < NB ? v : NB - 1; };
  double ga = 0.0, gb = 0.0;
  if (evo) {
    Walk64 wb(lane, n);
#pragma unroll
    for (int t = 0; t < PS; ++t) {
      if (lane + 64 * t < nb2) {
        if (!DD) Dm[wb.r + wb.c * L] = tdv[DD ? 0 : t];
        M[wb.r + wb.c * L] = tbv[t];
      }
      wb.next();
    }
    wave_sync();
    /* re = c - D x_s - B x_{s-1}, row i (D diagonal: its term once, in half 0) */
    double p = (DD && h == 0) ? dd[ir] * xs[ir] : 0.0, q = 0.0;
#pragma unroll
    for (int jj = 0; jj < HB; ++jj) {
      const int j = cl(j0 + jj);
      const bool on = jj < cn_h;
      q = on ? fma(M[ir + j * L], xm[j], q) : q;
      if (!DD) p = on ? fma(Dm[ir + j * L], xs[j], p) : p;
    }
    p += q;
    p += __shfl_xor(p, 32);
    if (h == 0 && i < n) re[i] = (ch.from_raw ? ch.wk[s * n + i] * ch.rc[s * n + i] : ch.c[s * n + i]) - p;
    wave_sync();
    /* gA = D^T re, gB = B^T re, column i */
    if (DD && h == 0) ga = dd[ir] * re[ir];
#pragma unroll
    for (int rr = 0; rr < HB; ++rr) {
      const int r = cl(j0 + rr);
      const bool on = rr < cn_h;
      gb = on ? fma(M[r + ir * L], re[r], gb) : gb;
      if (!DD) ga = on ? fma(Dm[r + ir * L], re[r], ga) : ga;
    }
    wave_sync(); /* B consumed: its block takes C */
  }
  if (m > 0) {
    Walk64 wc(lane, mr);
#pragma unroll
    for (int t = 0; t < PS; ++t) {
      if (lane + 64 * t < cn) M[wc.r + wc.c * L] = tcv[t]; /* C is mr x n, ld mr; rows >= m are zero */
      wc.next();
    }
    wave_sync();
    /* ro = o - C x_s, row i (rows >= m count as zero) */
    double p = 0.0;
#pragma unroll
    for (int jj = 0; jj < HB; ++jj) {
      const int j = cl(j0 + jj);
      p = jj < cn_h ? fma(M[irc + j * L], xs[j], p) : p;
    }
    p += __shfl_xor(p, 32);
    if (h == 0 && i < mr)
      ro[i] = i < m ? (ch.from_raw ? ch.rC[s * mr + i] * ch.ro[s * mr + i] : ch.o[s * n + i]) - p : 0.0;
    wave_sync();
    /* gA += C^T ro, column i */
#pragma unroll
    for (int rr = 0; rr < HB; ++rr) {
      const int r = cl(r0 + rr);
      ga = rr < cm_h ? fma(M[r + ir * L], ro[r], ga) : ga;
    }
  }
  ga += __shfl_xor(ga, 32);
  gb += __shfl_xor(gb, 32);
  if (h == 0 && i < n) {
    gA[s * n + i] = ga;
    gB[s * n + i] = gb;
  }
}

/* g_s = gA_s + gB_{s+1} (when step s+1 belongs to the same problem) */
__global__ void __launch_bounds__(256) k_gather(const double* __restrict__ gA, const double* __restrict__ gB,
                                                const int8_t* first, double* __restrict__ g, int64_t steps, int n) {
  const int64_t total = steps * n;
  for (int64_t i = blockIdx.x * 256LL + threadIdx.x; i < total; i += (int64_t)gridDim.x * 256) {
    const int64_t s = i / n;
    const bool nxt = s + 1 < steps && first[s + 1] == 0;
    g[i] = gA[i] + (nxt ? gB[i + n] : 0.0);
  }
}

/*
 * Forward sweep, even steps of a level: y_j = R_j^{-T} g_j, one segment of
 * S = sweep_seg(NB) lanes per even step (64 / S steps per wave). R is staged
 * through LDS in its own precision; lane i of a segment holds entry i, the
 * solved y_k reach the segment by seg_bcast, and the pivots' reciprocals are
 * formed in parallel before the substitution.
 */
template <int NB, class T>
__global__ void __launch_bounds__(64) k_fwd_even(uk_level_t lv, const T* __restrict__ R, const double* __restrict__ g,
                                                 double* __restrict__ y, int n) {
  constexpr int L = NB + 1, S = sweep_seg(NB), G = 64 / S;
  __shared__ T Rs[G][NB * L];
  const int lane = threadIdx.x, seg = S == 64 ? 0 : lane / S, i = lane - seg * S; /* seg uniform at S = 64 */
  const int64_t e0 = (int64_t)blockIdx.x * G + seg;
  const bool live = e0 < lv.evens; /* a spare segment repeats the last step and stores nothing */
  const int64_t e = live ? e0 : lv.evens - 1;
  const int p = find_problem_seg<G>(lv.even_start, lv.nprob, (int64_t)blockIdx.x * G, e);
  const int64_t j = 2 * (e - lv.even_start[p]);
  const int64_t sj = lv.step_start[p] + j;
  const int64_t nn = (int64_t)n * n;
  T* Rw = Rs[seg];
  {
    constexpr int PER = (NB * NB + S - 1) / S;
    T rr[PER];
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int idx = i + S * t;
      rr[t] = idx < n * n ? R[e * nn + idx] : T(0);
    }
    Walk64 wr(i, n, S);
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      if (i + S * t < n * n) Rw[wr.r + wr.c * L] = rr[t];
      wr.next();
    }
  }
  double v = i < n ? g[sj * n + i] : 0.0;
  __syncthreads();
  const double rinv = i < n ? 1.0 / (double)Rw[i + i * L] : 0.0;
  /* R^T is lower triangular: (R^T)_{ik} = R_{ki} */
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    if (k < n) {
      const double yk = seg_bcast<S>(v * rinv, k);
      if (i == k) v = yk;
      else if (i > k && i < n) v = fma(-(double)Rw[k + i * L], yk, v);
    }
  }
  if (i < n && live) y[e * n + i] = v;
}

/*
 * Forward sweep, odd steps (one segment of S lanes per pair u, odd step
 * j + 1 = 2u + 1): the next level's right-hand side
 *   g'_{j+1} = g_{j+1} - Y_j^T y_j - B~_{j+2}^T y_{j+2}.
 * Y_j and B~_{j+2} are staged through LDS with coalesced loads (all in
 * flight together); lane i then forms entry i from column i.
 */
template <int NB, class T>
__global__ void __launch_bounds__(64) k_fwd_odd(uk_level_t lv, const T* __restrict__ Bt, const T* __restrict__ Y,
                                                const double* __restrict__ g, const double* __restrict__ y,
                                                double* __restrict__ gnext, int n) {
  constexpr int L = NB + 1, S = sweep_seg(NB), G = 64 / S;
  constexpr int PER = (NB * NB + S - 1) / S;
  __shared__ T Ys[G][NB * L], Bs[G][NB * L];
  __shared__ double yj[G][NB], yn[G][NB];
  const int lane = threadIdx.x, seg = S == 64 ? 0 : lane / S, i = lane - seg * S; /* seg uniform at S = 64 */
  const int64_t unit0 = (int64_t)blockIdx.x * G + seg;
  const int64_t unit = unit0 < lv.units ? unit0 : lv.units - 1;
  const int p = find_problem_seg<G>(lv.unit_start, lv.nprob, (int64_t)blockIdx.x * G, unit);
  const int64_t Lp = lv.L[p];
  /* a one-step problem has no odd step; a spare segment repeats the last pair */
  const bool act = unit0 < lv.units && Lp > 1;
  const int64_t u = unit - lv.unit_start[p];
  const int64_t j = 2 * u;
  const int64_t e = lv.even_start[p] + u; /* slot of even step j */
  const bool has_n = j + 2 < Lp;          /* even step j + 2 exists */
  const int64_t nn = (int64_t)n * n;
  T ry[PER], rb[PER];
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int idx = i + S * t;
    const bool in = idx < n * n;
    ry[t] = in ? Y[e * nn + idx] : T(0);
    rb[t] = (in && has_n) ? Bt[(e + 1) * nn + idx] : T(0);
  }
  if (i < n) {
    yj[seg][i] = y[e * n + i];
    yn[seg][i] = has_n ? y[(e + 1) * n + i] : 0.0;
  }
  Walk64 wy(i, n, S);
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    if (i + S * t < n * n) {
      Ys[seg][wy.r + wy.c * L] = ry[t];
      Bs[seg][wy.r + wy.c * L] = rb[t];
    }
    wy.next();
  }
  __syncthreads();
  if (i < n && act) {
    double v0 = g[(lv.step_start[p] + j + 1) * n + i], v1 = 0.0;
    for (int r = 0; r < n; ++r) {
      v0 = fma(-(double)Ys[seg][r + i * L], yj[seg][r], v0);
      v1 = fma(-(double)Bs[seg][r + i * L], yn[seg][r], v1);
    }
    gnext[(lv.next_start[p] + u) * n + i] = v0 + v1;
  }
}

/*
 * Backward sweep (one segment of S = sweep_seg(NB) lanes per even step):
 * x_j = R^{-1} (y_j - B~ x_{j-1} - Y x_{j+1}) with the neighbours' x from the
 * levels above, in the original step order. Lane i of a segment owns row i:
 * B~, Y and R stream column by column (coalesced, no LDS), lane i keeps row i
 * of R in registers, and the neighbours' estimates and each solved x_q reach
 * the segment by seg_bcast.
 */
template <int NB, class T, class TR>
__global__ void __launch_bounds__(64) k_solve_est(uk_level_t lv, FactorPtrs<T> fac, const TR* __restrict__ rhs,
                                                  double* __restrict__ x, int n) {
  constexpr int S = sweep_seg(NB), G = 64 / S;
  const int lane = threadIdx.x, seg = S == 64 ? 0 : lane / S, i = lane - seg * S; /* seg uniform at S = 64 */
  const bool row = i < n;
  const int64_t e0 = (int64_t)blockIdx.x * G + seg;
  const bool live = e0 < lv.evens; /* a spare segment repeats the last step and stores nothing */
  const int64_t e = live ? e0 : lv.evens - 1;
  const int p = find_problem_seg<G>(lv.even_start, lv.nprob, (int64_t)blockIdx.x * G, e);
  const int64_t j = 2 * (e - lv.even_start[p]);
  const int64_t Lp = lv.L[p];
  const bool has_m = j > 0, has_p = j + 1 < Lp;
  const int64_t nn = (int64_t)n * n;
  const int64_t ob = lv.orig_start[lv.prob[p]];
  const int64_t stride = (int64_t)1 << lv.level;
  const int64_t oj = ob + (j + 1) * stride - 1;
  const int64_t om = ob + j * stride - 1;
  const int64_t op = ob + (j + 2) * stride - 1;
  const T* Rg = fac.R + e * nn;
  const T* Bg = fac.Bt + e * nn;
  const T* Yg = fac.Y + e * nn;
  const double xm = (has_m && row) ? x[om * n + i] : 0.0;
  const double xp = (has_p && row) ? x[op * n + i] : 0.0;
  double v = row ? (double)rhs[e * n + i] : 0.0, v1 = 0.0;
  /* The column loads take no branch: row and column indices are clamped into
   * the step's n x n block (every even step of the level has its R, B~ and Y
   * slots) and the values outside it are replaced by zeros, so the compiler
   * issues every load of the loop before the first wait. A replaced term adds
   * -0 * 0 to v, which leaves v as it is. */
  const int ic = row ? i : 0;
  const bool lm = has_m && row, lp = has_p && row;
  T rrow[NB];
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const bool col = q < n;
    const int64_t idx = ic + (int64_t)(col ? q : n - 1) * n;
    const T rv = Rg[idx], bv = Bg[idx], yw = Yg[idx];
    rrow[q] = (row && col) ? rv : T(0);
    const double b = (lm && col) ? (double)bv : 0.0;
    const double yv = (lp && col) ? (double)yw : 0.0;
    v = fma(-b, seg_bcast<S>(xm, q), v);
    v1 = fma(-yv, seg_bcast<S>(xp, q), v1);
  }
  v += v1;
  const double rinv = row ? 1.0 / (double)Rg[i + (int64_t)i * n] : 0.0;
#pragma unroll
  for (int q = NB - 1; q >= 0; --q) {
    if (q < n) {
      const double xq = seg_bcast<S>(v * rinv, q);
      v = (i == q) ? xq : (i < q ? fma(-(double)rrow[q], xq, v) : v);
    }
  }
  if (row && live) x[oj * n + i] = v;
}

}  // namespace

/* ------------------------------------------------------------------------ */
/* launchers                                                                  */
/* ------------------------------------------------------------------------ */

extern "C" int uk_bucket_qr(int n);

#define UK_DISPATCH(NBV, CALL) \
  case NBV: { CALL(NBV); } break;

/* workgroups of a sweep over `count` steps in bucket nb (64 / sweep_seg steps each) */
static unsigned sweep_grid(int64_t count, int nb) {
  const int64_t g = 64 / sweep_seg(nb);
  return (unsigned)((count + g - 1) / g);
}

extern "C" hipError_t uk_launch_to32(float* dst, const double* src, int64_t count, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>((count + 255) / 256, 65536);
  k_to32<<<dim3((unsigned)blocks), dim3(256), 0, st>>>(dst, src, count);
  return hipGetLastError();
}

extern "C" hipError_t uk_launch_add(double* x, const double* d, int64_t count, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>((count + 255) / 256, 65536);
  k_add<<<dim3((unsigned)blocks), dim3(256), 0, st>>>(x, d, count);
  return hipGetLastError();
}

extern "C" hipError_t uk_launch_normal(const uk_chain_t* ch, const int8_t* first, const int32_t* m_obs,
                                       const double* x, double* gA, double* gB, double* g, int64_t steps, int n,
                                       hipStream_t st) {
  const int nb = uk_bucket_qr(n > ch->mr ? n : ch->mr);
  if (steps <= 0) return hipSuccess;
  /* NB <= 32: one wave per step, four steps per workgroup */
  const dim3 gw((unsigned)((steps + 3) / 4));
#define W_(NBV)                                                                                              \
  if (ch->d_diag) k_normal_w<NBV, true><<<gw, dim3(256), 0, st>>>(*ch, first, m_obs, x, gA, gB, n, steps); \
  else k_normal_w<NBV, false><<<gw, dim3(256), 0, st>>>(*ch, first, m_obs, x, gA, gB, n, steps)
#define L_(NBV)                                                                                       \
  if (ch->d_diag) k_normal<NBV, true><<<dim3((unsigned)steps), dim3(256), 0, st>>>(*ch, first, m_obs, x, gA, gB, n); \
  else k_normal<NBV, false><<<dim3((unsigned)steps), dim3(256), 0, st>>>(*ch, first, m_obs, x, gA, gB, n)
  switch (nb) {
    UK_DISPATCH(8, W_) UK_DISPATCH(16, W_) UK_DISPATCH(24, W_) UK_DISPATCH(32, W_) UK_DISPATCH(48, L_)
    UK_DISPATCH(64, L_)
    default: return hipErrorInvalidValue;
  }
#undef W_
#undef L_
  const int64_t blocks = std::min<int64_t>((steps * n + 255) / 256, 65536);
  k_gather<<<dim3((unsigned)blocks), dim3(256), 0, st>>>(gA, gB, first, g, steps, n);
  return hipGetLastError();
}

extern "C" hipError_t uk_launch_to64(double* dst, const float* src, int64_t count, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>((count + 255) / 256, 65536);
  k_to64<<<dim3((unsigned)blocks), dim3(256), 0, st>>>(dst, src, count);
  return hipGetLastError();
}

/* forward sweep of one level: y (per even slot of the level) and the next
 * level's right-hand side, reading the level's factors in precision T */
template <class T>
static hipError_t launch_fwd(const uk_level_t* lv, const T* R, const T* Bt, const T* Y, const double* g, double* y,
                             double* gnext, int n, hipStream_t st) {
  const int nb = uk_bucket_qr(n);
#define L_(NBV)                                                                                   \
  k_fwd_even<NBV, T><<<dim3(sweep_grid(lv->evens, NBV)), dim3(64), 0, st>>>(*lv, R, g, y, n);            \
  k_fwd_odd<NBV, T><<<dim3(sweep_grid(lv->units, NBV)), dim3(64), 0, st>>>(*lv, Bt, Y, g, y, gnext, n)
  switch (nb) {
    UK_DISPATCH(8, L_) UK_DISPATCH(16, L_) UK_DISPATCH(24, L_) UK_DISPATCH(32, L_) UK_DISPATCH(48, L_)
    UK_DISPATCH(64, L_)
    default: return hipErrorInvalidValue;
  }
#undef L_
  return hipGetLastError();
}
extern "C" hipError_t uk_launch_fwd(const uk_level_t* lv, const uk_factors32_t* fac, const double* g, double* y,
                                    double* gnext, int n, hipStream_t st) {
  return launch_fwd<float>(lv, fac->R, fac->Bt, fac->Y, g, y, gnext, n, st);
}
extern "C" hipError_t uk_launch_fwd64(const uk_level_t* lv, const uk_factors_t* fac, const double* g, double* y,
                                      double* gnext, int n, hipStream_t st) {
  return launch_fwd<double>(lv, fac->R, fac->Bt, fac->Y, g, y, gnext, n, st);
}

/* backward sweep of one level; rhs32 (the FP32 transformed right-hand side)
 * or rhs64 (the forward sweep's y, or an FP64 level's transformed right-hand
 * side) */
template <class T>
static hipError_t launch_solve_est(const uk_level_t* lv, const FactorPtrs<T>& f, const float* rhs32,
                                   const double* rhs64, double* x, int n, hipStream_t st) {
  const int nb = uk_bucket_qr(n);
#define L_(NBV)                                                                                                \
  if (rhs32) k_solve_est<NBV, T, float><<<dim3(sweep_grid(lv->evens, NBV)), dim3(64), 0, st>>>(*lv, f, rhs32, x, n); \
  else k_solve_est<NBV, T, double><<<dim3(sweep_grid(lv->evens, NBV)), dim3(64), 0, st>>>(*lv, f, rhs64, x, n)
  switch (nb) {
    UK_DISPATCH(8, L_) UK_DISPATCH(16, L_) UK_DISPATCH(24, L_) UK_DISPATCH(32, L_) UK_DISPATCH(48, L_)
    UK_DISPATCH(64, L_)
    default: return hipErrorInvalidValue;
  }
#undef L_
  return hipGetLastError();
}
extern "C" hipError_t uk_launch_solve_est(const uk_level_t* lv, const uk_factors32_t* fac, const float* rhs32,
                                          const double* rhs64, double* x, int n, hipStream_t st) {
  return launch_solve_est<float>(lv, FactorPtrs<float>{fac->R, fac->Bt, fac->Y}, rhs32, rhs64, x, n, st);
}
extern "C" hipError_t uk_launch_solve_est64(const uk_level_t* lv, const uk_factors_t* fac, const double* rhs64,
                                            double* x, int n, hipStream_t st) {
  return launch_solve_est<double>(lv, FactorPtrs<double>{fac->R, fac->Bt, fac->Y}, nullptr, rhs64, x, n, st);
}
