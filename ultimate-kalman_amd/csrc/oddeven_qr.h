/*
 * oddeven_qr.h -- the QR kernels of the odd-even elimination (phases 1..3),
 * shared by oddeven.hip (n <= 16, and the device helpers its other kernels
 * use) and oddeven_qr{24,32,48,64}.hip (one translation unit per bucket above
 * 16, so that their many launch variants compile in parallel).
 *
 * Reference: c/kalman_oddeven_smoother.c (phases :379-560, fix-last
 * :1452-1476, renaming :563-597).
 */
#ifndef UKHIP_ODDEVEN_QR_H
#define UKHIP_ODDEVEN_QR_H

#include <hip/hip_runtime.h>

#include "ukhip_internal.h"

namespace ukqr {
/* Device views of a whitened chain and of the back-substitution factors in
 * the working precision T (double: the layouts of uk_chain_t and
 * uk_factors_t; float: the same layouts in single precision). */
template <class T>
struct ChainT {
  T *B, *D, *c, *C, *o;
  int32_t mr;
  int32_t d_diag; /* 1: every D is diagonal, stored as its diagonal (the first n entries of its slot) */
  int32_t c_tri;  /* 1: every C is upper triangular n x n */
  int32_t b_synth; /* 1: B_s = -diag(wk_s) wF[wFidx[s]] (wFidx NULL: block 0), formed on load */
  const double* wk;
  const double* wF;
  const int64_t* wFidx;
  int32_t c_synth; /* 1: C_s = diag(rC_s) rG_s (mr x n, ld mr; rows >= rm[s] zero), formed on load */
  const double* rG;
  const double* rC;
  const int32_t* rm;
};
template <class T>
struct FactorsT {
  T *R, *Bt, *Y, *cf;
};
/* A launch variant: columns per lane group and workgroups per CU. */
struct Launch {
  int cpl, wg;
};
/* Phase `phase` of one level for the buckets above 16, defined in
 * oddeven_qr<NB>.hip. */
#define UKQR_DECLARE(NBV)                                                                                            \
  void phase_launch_b##NBV(const uk_level_t* lv, const ChainT<double>& in, const ChainT<double>& out,                 \
                           const FactorsT<double>& fac, double* fix, int n, hipStream_t st, int phase, Launch v);    \
  void phase_launch_b##NBV(const uk_level_t* lv, const ChainT<float>& in, const ChainT<float>& out,                   \
                           const FactorsT<float>& fac, float* fix, int n, hipStream_t st, int phase, Launch v);
UKQR_DECLARE(24)
UKQR_DECLARE(32)
UKQR_DECLARE(48)
UKQR_DECLARE(64)
#undef UKQR_DECLARE
}  // namespace ukqr

namespace {
using ukqr::ChainT;
using ukqr::FactorsT;

/* x + (x of the lane with index ^1) + ... over groups of SL = 2 or 4 adjacent
 * lanes (DPP quad permutations, no LDS). */
__device__ __forceinline__ double dpp_xor1(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_update_dpp(0, lo, 0xB1, 0xF, 0xF, false); /* quad_perm [1,0,3,2] */
  hi = __builtin_amdgcn_update_dpp(0, hi, 0xB1, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double dpp_xor2(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_update_dpp(0, lo, 0x4E, 0xF, 0xF, false); /* quad_perm [2,3,0,1] */
  hi = __builtin_amdgcn_update_dpp(0, hi, 0x4E, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float dpp_xor1(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_xor2(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
}
template <int SL, class T>
__device__ __forceinline__ T group_sum(T v) {
  v += dpp_xor1(v);
  if (SL == 4) v += dpp_xor2(v);
  return v;
}

/* 2-vector of T for paired LDS reads */
template <class T>
struct Vec2;
template <>
struct Vec2<double> {
  using type = double2;
};
template <>
struct Vec2<float> {
  using type = float2;
};

/* The problem of work item x: the last p < nprob with start[p] <= x (start
 * non-decreasing, start[0] = 0 <= x; x the same on every lane). The wave
 * searches in rounds of 64 probes, one per lane, so a search costs
 * ceil(log64 nprob) dependent loads instead of a binary search's
 * ceil(log2 nprob): 2 up to 4096 problems. Called with the whole wave active
 * (at kernel start, every kernel's workgroup a multiple of 64 threads). */
__device__ __forceinline__ int64_t find_problem(const int64_t* __restrict__ start, int nprob, int64_t x) {
  const int lane = (int)(threadIdx.x & 63);
  int lo = 0, cnt = nprob; /* the answer lies in [lo, lo + cnt) and start[lo] <= x */
  while (cnt > 1) {
    const int step = (cnt + 63) >> 6;
    const int off = lane * step;
    const bool ok = off < cnt && start[lo + off] <= x; /* true on a prefix of the lanes, lane 0 included */
    const int j = __popcll(__ballot(ok)) - 1;
    lo += j * step;
    cnt = min(step, cnt - j * step);
  }
  return lo;
}

/* One matrix column as held by one of its SL lanes: rows r = SL*l + sub of
 * block 0 in locals [0,H) and of block 1 in [H,2H), H = NB/SL. */
template <int NB, int SL, class T = double>
struct Col {
  static constexpr int H = NB / SL;
  T a[2 * H];

  __device__ __forceinline__ void zero_block(int blk) {
#pragma unroll
    for (int l = 0; l < H; ++l) a[blk * H + l] = T(0);
  }
  /* block blk <- rows [0, rows) of a column at src (unit stride), zero elsewhere */
  __device__ __forceinline__ void load_block(int blk, const T* __restrict__ src, int rows, int sub) {
#pragma unroll
    for (int l = 0; l < H; ++l) {
      const int r = SL * l + sub;
      a[blk * H + l] = (r < rows) ? src[r] : T(0);
    }
  }
  __device__ __forceinline__ void store_block(int blk, T* __restrict__ dst, int rows, int sub) const {
#pragma unroll
    for (int l = 0; l < H; ++l) {
      const int r = SL * l + sub;
      if (r < rows) dst[r] = a[blk * H + l];
    }
  }
  /* block blk <- column col of B = -diag(k) F (n x n, F column-major), the
   * whitened evolution block formed as the whitening pass forms it */
  __device__ __forceinline__ void load_bsynth(int blk, const double* __restrict__ k, const double* __restrict__ F,
                                              int col, int n, int sub) {
#pragma unroll
    for (int l = 0; l < H; ++l) {
      const int r = SL * l + sub;
      a[blk * H + l] = (r < n) ? T(-k[r] * F[r + (int64_t)col * n]) : T(0);
    }
  }
  /* block blk <- column col of C = diag(w) G (G mr x n, ld mr; rows >= m zero),
   * each entry the FP64 product rounded to T, as the whitening pass stores it */
  __device__ __forceinline__ void load_csynth(int blk, const double* __restrict__ w, const double* __restrict__ G,
                                              int col, int mr, int m, int sub) {
#pragma unroll
    for (int l = 0; l < H; ++l) {
      const int r = SL * l + sub;
      a[blk * H + l] = (r < mr && r < m) ? T(w[r] * G[r + (int64_t)col * mr]) : T(0);
    }
  }
  /* block blk <- column col of a diagonal matrix given by its diagonal d */
  __device__ __forceinline__ void load_diag(int blk, const T* __restrict__ d, int col, int sub) {
    const T v = d[col];
#pragma unroll
    for (int l = 0; l < H; ++l) a[blk * H + l] = (SL * l + sub == col) ? v : T(0);
  }
  /* upper-triangular store of block 0: rows > col are written as zero */
  __device__ __forceinline__ void store_triu(T* __restrict__ dst, int rows, int col, int sub) const {
#pragma unroll
    for (int l = 0; l < H; ++l) {
      const int r = SL * l + sub;
      if (r < rows) dst[r] = (r <= col) ? a[l] : T(0);
    }
  }
  __device__ __forceinline__ void move_0_to_1() {
#pragma unroll
    for (int l = 0; l < H; ++l) a[H + l] = a[l];
  }
  /* block 0 <- block 1, block 1 <- 0 */
  __device__ __forceinline__ void move_1_to_0() {
#pragma unroll
    for (int l = 0; l < H; ++l) {
      a[l] = a[H + l];
      a[H + l] = T(0);
    }
  }
  /* the column's block blk into LDS at st (rows at or above the block's
   * row count are written too and never copied out) */
  __device__ __forceinline__ void write_stage(int blk, T* st, int sub) const {
#pragma unroll
    for (int l = 0; l < H; ++l) st[SL * l + sub] = a[blk * H + l];
  }
  /* block 0 into LDS as an upper-triangular column: rows > col written as zero */
  __device__ __forceinline__ void write_stage_triu(T* st, int col, int sub) const {
#pragma unroll
    for (int l = 0; l < H; ++l) {
      const int r = SL * l + sub;
      st[r] = (r <= col) ? a[l] : T(0);
    }
  }
  /* rows r < rows of block blk scaled by sign * s[r] */
  __device__ __forceinline__ void scale_block(int blk, const T* __restrict__ s, int rows, int sub, T sign) {
#pragma unroll
    for (int l = 0; l < H; ++l) {
      const int r = SL * l + sub;
      if (r < rows) a[blk * H + l] *= sign * s[r];
    }
  }
};

/*
 * LDS staging of an FP32 kernel's output block (at most NB x NB, column-major
 * in HBM with leading dimension ld): the lane groups write their columns into
 * LDS, and the workgroup stores the block with RW consecutive lanes per
 * column (consecutive addresses, whole cache lines) instead of each group's
 * 16-byte pieces of 16 columns per wave store. Stores need no wait, so this
 * adds no latency to the chain; measured 2.8% faster on the mixed-precision
 * smoothing of a 600-problem chunk of the batching workload, and slower in
 * FP64, whose kernels store directly. (Loads through the stage were slower:
 * block after block behind barriers, they wait one HBM round trip each.)
 * Buffer: column c at c * LDP. LDP is RW plus 4 (plus 2 with two columns per
 * group): at one local, the groups of a wave (columns g * CPL + q, rows
 * SL * l + sub) then start on different banks.
 */
template <int NB, int CPL, int NT, class T>
struct Stage {
  static constexpr int RW = NB <= 8 ? 8 : NB <= 16 ? 16 : NB <= 32 ? 32 : 64;
  static constexpr int LDP = RW + (CPL == 2 ? 2 : 4);
  static constexpr int SIZE = NB * LDP;
  static_assert(NT % RW == 0, "a workgroup covers whole staging columns");
  T* buf;
  __device__ __forceinline__ T* col(int c) const { return buf + c * LDP; }
  /* rows r < rows of columns c < cols to dst (after a barrier over the writes) */
  __device__ __forceinline__ void out(T* __restrict__ dst, int ld, int rows, int cols, int tid) const {
    const int r = tid & (RW - 1);
    if (r < rows)
      for (int c = tid / RW; c < cols; c += NT / RW) dst[r + (int64_t)c * ld] = buf[c * LDP + r];
  }
};

/* the n x n block at dst set to zero by the whole workgroup (contiguous) */
template <int NT, class T>
__device__ __forceinline__ void block_zero(T* __restrict__ dst, int64_t nn, int tid) {
  for (int64_t i = tid; i < nn; i += NT) dst[i] = T(0);
}

/* LDS reflector buffers: [buffer][sub][2H + 4] (the +4 skews the sub-lane
 * arrays onto different banks), then [buffer][2] scalars. A stride that
 * lands on bank 0 or 16 of the 32 four-byte banks (2H + 4 = 16 in bucket 24)
 * would put two sub-lanes' reads on the same banks: it is padded by 4 more. */
template <int NB, int SL, class T = double>
struct Refl {
  static constexpr int SUBSTRIDE0 = 2 * (NB / SL) + 4;
  static constexpr int SUBSTRIDE =
      SUBSTRIDE0 + ((SUBSTRIDE0 * (int)sizeof(T) / 4) % 16 == 0 ? 4 : 0);
  T v[2][SL * SUBSTRIDE];
  T sc[2][2];
};

/* Reflector prefetch: with one column per lane group the whole reflector is
 * held in registers when a lane's share has at most UK_QR_FULL entries, else
 * it streams in chunks of UK_QR_CH; with two columns (whose registers leave
 * less room) it streams in chunks of at most UK_QR_CH2 whenever it is longer,
 * read once for the dot products and once for the updates. */
#ifndef UK_QR_FULL
#define UK_QR_FULL 24
#endif
#ifndef UK_QR_CH
#define UK_QR_CH 16
#endif
#ifndef UK_QR_CH2
#define UK_QR_CH2 12
#endif
/* the largest even divisor of nl not above cap (chunks tile the share exactly) */
__host__ __device__ constexpr int qr_chunk(int nl, int cap) {
  int c = cap < nl ? cap : nl;
  c -= c & 1;
  while (c > 2 && nl % c != 0) c -= 2;
  return c;
}

/*
 * Step-dependent row ranges. A column's locals are cut into granules of GL
 * locals (GL * SL rows of one block, NG = H / GL granules per block); the
 * pivot steps k in segment SEG = k / (GL * SL) share which granules can hold
 * a nonzero of the Householder vector u. The QR's structure flags:
 *   block 0: rows >= k only (u_i = 0 above the pivot), so granules >= SEG;
 *     QF_HALF0: rows >= NB / 2 of block 0 start zero in every column, so a
 *     reflector reaches one of them only at its own pivot row: granules
 *     < NG / 2, or SEG.
 *   block 1: all rows;
 *     QF_TRI1: block 1 starts upper triangular in the pivot columns (R~ in
 *     phase 2; C = G~ of the levels above 0 in phases 1 and 3), and every
 *     reflector keeps it so: rows <= k, granules <= SEG;
 *     QF_HALF1: rows >= qr_half_rows of block 1 are zero in every column and
 *     stay so: granules < ceil(NG / 2).
 *     QF_DIAG0: block 0 of the pivot columns starts diagonal (D_j in phase 2
 *     at level 0 when every H is the identity and every K is given as 'w'):
 *     reflector i reaches block 0 only at row i, so u_k's block-0 part is its
 *     pivot row alone: granule SEG.
 *   QF_DONE: a finished pivot column (piv0 <= column < piv0 + k) is not
 *     updated any more; the reflectors leave its rows up to its pivot as they
 *     are (u is zero there), and the rest of it is not stored.
 * The other granules of u are exact zeros: skipping them changes no result
 * bit (their products are exact zeros in every dot product and update), and
 * the step code is specialised per segment, so the skipping costs no branch.
 */
/*   QF_SWAP: the registers of column 0 (group 0, slot 0) hold another column
 *     from the second step on (house_qr_swap): never counted as finished. */
enum { QF_TRI1 = 1, QF_DONE = 2, QF_HALF0 = 4, QF_HALF1 = 8, QF_DIAG0 = 16, QF_SWAP = 32 };
__host__ __device__ constexpr int qr_gl(int H) { return H <= 16 ? 2 : 4; }
/* the granules of a block that QF_HALF0 / QF_HALF1 keep: the first half,
 * rounded up (bucket 24: 2 of 3) */
__host__ __device__ constexpr int qr_half_granules(int H) { return (H / qr_gl(H) + 1) / 2; }
/* rows of a block that QF_HALF0 / QF_HALF1 keep for bucket NB with SL lanes
 * per column (NB / 2, and 16 in bucket 24) */
__host__ __device__ constexpr int qr_half_rows(int NB, int SL) {
  return qr_half_granules(NB / SL) * qr_gl(NB / SL) * SL;
}
__host__ __device__ constexpr bool qr_need(int l, int H, int SEG, int QF) {
  const int hg = qr_half_granules(H);
  if (l < H) {
    const int g = l / qr_gl(H);
    return g >= SEG && (!(QF & QF_HALF0) || g < hg || g == SEG) && (!(QF & QF_DIAG0) || g == SEG);
  }
  const int g = (l - H) / qr_gl(H);
  return (!(QF & QF_TRI1) || g <= SEG) && (!(QF & QF_HALF1) || g < hg);
}

/*
 * One Householder step for pivot k (target row k of block 0), over the CPL
 * columns a lane group holds; k lies in segment SEG.
 *   piv_slot: which of this group's columns is pivot column k (-1: none);
 *   upd: bit s set when column s is updated (uniform over the group's lanes).
 * The pivot lanes publish the reflector (u_i = 0 for i < k and its scalar;
 * the two forms below); then every updated column applies H = I - tau u u^T.
 * The dot and the update run over the segment's granules without branches,
 * so the LDS reads of v issue back to back; the reflector read into registers
 * serves all CPL columns. Each column's operations and their order do not
 * depend on CPL. One barrier.
 */
/* V (optional, leading dimension ldv): the pivot lanes also store the
 * published vector as column k of V and its tau as tau[k] (a blocked-WY
 * accumulation of the reflectors; a column that counts as zero stores zeros,
 * and granules skipped as exact zeros are left as they are, zero). */
template <int NB, int SL, class T, int SEG, int QF>
__device__ __forceinline__ void house_publish(const Col<NB, SL, T>& col, Refl<NB, SL, T>& R, int buf, int sub, int k,
                                              T* V = nullptr, T* tauv = nullptr, int ldv = 0) {
  constexpr int H = NB / SL;
  T* vv = &R.v[buf][sub * Refl<NB, SL, T>::SUBSTRIDE];
  T s0 = T(0), s1 = T(0), alpha = T(0);
#pragma unroll
  for (int l = 0; l < H; ++l) {
    if (!qr_need(l, H, SEG, QF)) continue;
    const int r = SL * l + sub;
    const T x = col.a[l];
    alpha = (r == k) ? x : alpha;
    const T xm = (r > k) ? x : T(0);
    if (l & 1) s1 = fma(xm, xm, s1);
    else s0 = fma(xm, xm, s0);
  }
#pragma unroll
  for (int l = H; l < 2 * H; ++l) {
    if (!qr_need(l, H, SEG, QF)) continue;
    if (l & 1) s1 = fma(col.a[l], col.a[l], s1);
    else s0 = fma(col.a[l], col.a[l], s0);
  }
  const T s = group_sum<SL>(s0 + s1);
  alpha = group_sum<SL>(alpha);
  if constexpr (sizeof(T) == 8) {
    /* H = I - tau' u u^T with u = x below the pivot and u_k = alpha - beta:
     * tau' = -1 / (beta u_k), the unit-pivot form's tau / u_k^2, so the
     * pivot lanes publish their column as it is (one division, no
     * scaling). |beta u_k| >= beta^2 = ||x||^2, so a column whose squared
     * norm is below 1e-280 counts as zero (no reflection). */
    T taup = T(0), ukk = T(0);
    if (s > T(1e-280)) {
      const T beta = -copysign(sqrt(fma(alpha, alpha, s)), alpha);
      ukk = alpha - beta;
      taup = T(-1) / (beta * ukk);
    }
#pragma unroll
    for (int l = 0; l < H; ++l) {
      if (!qr_need(l, H, SEG, QF)) continue;
      const int r = SL * l + sub;
      vv[l] = (r > k) ? col.a[l] : (r == k ? ukk : T(0));
    }
#pragma unroll
    for (int l = H; l < 2 * H; ++l)
      if (qr_need(l, H, SEG, QF)) vv[l] = col.a[l];
    if (sub == 0) R.sc[buf][0] = taup;
    if (V) {
      const T keep = taup != T(0) ? T(1) : T(0);
      T* vc = V + (int64_t)k * ldv;
#pragma unroll
      for (int l = 0; l < H; ++l) {
        if (!qr_need(l, H, SEG, QF)) continue;
        const int r = SL * l + sub;
        vc[r] = keep * ((r > k) ? col.a[l] : (r == k ? ukk : T(0)));
      }
#pragma unroll
      for (int l = H; l < 2 * H; ++l)
        if (qr_need(l, H, SEG, QF)) vc[NB + SL * (l - H) + sub] = keep * col.a[l];
      if (sub == 0) tauv[k] = taup;
    }
  } else {
    /* single precision: the unit-pivot form v = x / (alpha - beta), v_k = 1,
     * tau = (beta - alpha) / beta in [1, 2]; squaring the column's scale, as
     * the form above does, would leave FP32's range on nearly dependent
     * columns (m << n) */
    /* The square root and the two reciprocals are the hardware's 1-ulp
     * instructions (v_sqrt_f32, v_rcp_f32) instead of the correctly rounded
     * sequences: they sit on the step's critical path, and an ulp in tau or
     * in v's scale is the size of the FP32 rounding the factorization makes
     * anyway (the FP64 refinement removes both). Sums of squares below
     * 1e-28 are scaled by 2^64 first, keeping the square root's argument
     * normal. */
    T tau = T(0), scale = T(0);
    if (s != T(0)) {
      const T t2 = fma(alpha, alpha, s);
      const bool tiny = t2 < T(1e-28);
      T nrm = __builtin_amdgcn_sqrtf(tiny ? t2 * T(0x1p64) : t2);
      nrm = tiny ? nrm * T(0x1p-32) : nrm;
      const T beta = -copysign(nrm, alpha);
      tau = fma(-alpha, __builtin_amdgcn_rcpf(beta), T(1)); /* (beta - alpha) / beta */
      scale = __builtin_amdgcn_rcpf(alpha - beta);
    }
#pragma unroll
    for (int l = 0; l < H; ++l) {
      if (!qr_need(l, H, SEG, QF)) continue;
      const int r = SL * l + sub;
      vv[l] = (r > k) ? col.a[l] * scale : (r == k ? T(1) : T(0));
    }
#pragma unroll
    for (int l = H; l < 2 * H; ++l)
      if (qr_need(l, H, SEG, QF)) vv[l] = col.a[l] * scale;
    if (sub == 0) R.sc[buf][0] = tau;
    if (V) {
      const T keep = tau != T(0) ? T(1) : T(0);
      T* vc = V + (int64_t)k * ldv;
#pragma unroll
      for (int l = 0; l < H; ++l) {
        if (!qr_need(l, H, SEG, QF)) continue;
        const int r = SL * l + sub;
        vc[r] = keep * ((r > k) ? col.a[l] * scale : (r == k ? T(1) : T(0)));
      }
#pragma unroll
      for (int l = H; l < 2 * H; ++l)
        if (qr_need(l, H, SEG, QF)) vc[NB + SL * (l - H) + sub] = keep * col.a[l] * scale;
      if (sub == 0) tauv[k] = tau;
    }
  }
}

/* The reflector published in buffer buf applied to the columns whose bit is
 * set in upd (uniform over the group's lanes; with CPL == 1 the caller has
 * checked upd). */
template <int NB, int SL, int CPL, class T, int SEG, int QF>
__device__ __forceinline__ void house_update(Col<NB, SL, T> (&cols)[CPL], const Refl<NB, SL, T>& R, int buf, int sub,
                                             unsigned upd) {
  constexpr int H = NB / SL;
  {
    const T tau = R.sc[buf][0];
    if (tau != T(0)) {
      /* v is read from LDS in chunks of CH locals issued back to back. With
       * CPL > 1 every column of the group runs the same branch-free code (a
       * column that is not updated gets the scale 0 and is left as it is), so
       * the columns' chains interleave. */
      constexpr int NL = 2 * H;
      constexpr int CH = CPL == 1 ? (NL <= UK_QR_FULL ? NL : qr_chunk(NL, UK_QR_CH)) : qr_chunk(NL, UK_QR_CH2);
      static_assert(NL % CH == 0 && CH % 2 == 0, "reflector chunks must tile a lane's share");
      static_assert(H % qr_gl(H) == 0, "granules must tile a block");
      using V2 = typename Vec2<T>::type;
      const V2* v2 = reinterpret_cast<const V2*>(&R.v[buf][sub * Refl<NB, SL, T>::SUBSTRIDE]);
      T d0[CPL], d1[CPL], d2[CPL], d3[CPL];
#pragma unroll
      for (int s = 0; s < CPL; ++s) d0[s] = d1[s] = d2[s] = d3[s] = T(0);
#pragma unroll
      for (int c0 = 0; c0 < NL; c0 += CH) {
        T vr[CH];
#pragma unroll
        for (int l = 0; l < CH; l += 2) {
          if (!qr_need(c0 + l, H, SEG, QF)) continue;
          const V2 t = v2[(c0 + l) / 2];
          vr[l] = t.x;
          vr[l + 1] = t.y;
        }
#pragma unroll
        for (int s = 0; s < CPL; ++s) {
          T* a = cols[s].a;
#pragma unroll
          for (int l = 0; l < CH; l += 4) {
            if (qr_need(c0 + l, H, SEG, QF)) {
              d0[s] = fma(vr[l], a[c0 + l], d0[s]);
              d1[s] = fma(vr[l + 1], a[c0 + l + 1], d1[s]);
            }
            if (l + 2 < CH && qr_need(c0 + l + 2, H, SEG, QF)) {
              d2[s] = fma(vr[l + 2], a[c0 + l + 2], d2[s]);
              d3[s] = fma(vr[l + 3], a[c0 + l + 3], d3[s]);
            }
          }
        }
        if (CH == NL) { /* whole reflector in registers: update without re-reading */
#pragma unroll
          for (int s = 0; s < CPL; ++s) {
            const T d = group_sum<SL>((d0[s] + d1[s]) + (d2[s] + d3[s]));
            const T tw = (CPL == 1 || ((upd >> s) & 1u)) ? tau * d : T(0);
#pragma unroll
            for (int l = 0; l < NL; ++l)
              if (qr_need(l, H, SEG, QF)) cols[s].a[l] = fma(-tw, vr[l], cols[s].a[l]);
          }
          return;
        }
      }
      T tw[CPL];
#pragma unroll
      for (int s = 0; s < CPL; ++s) {
        const T d = group_sum<SL>((d0[s] + d1[s]) + (d2[s] + d3[s]));
        tw[s] = (CPL == 1 || ((upd >> s) & 1u)) ? tau * d : T(0);
      }
      /* compiler-only memory clobber (no instruction): the update pass reads
       * v from LDS again instead of keeping the first pass's values live */
      asm volatile("" ::: "memory");
#pragma unroll
      for (int c0 = 0; c0 < NL; c0 += CH) {
        T vr[CH];
#pragma unroll
        for (int l = 0; l < CH; l += 2) {
          if (!qr_need(c0 + l, H, SEG, QF)) continue;
          const V2 t = v2[(c0 + l) / 2];
          vr[l] = t.x;
          vr[l + 1] = t.y;
        }
#pragma unroll
        for (int s = 0; s < CPL; ++s) {
#pragma unroll
          for (int l = 0; l < CH; ++l)
            if (qr_need(c0 + l, H, SEG, QF)) cols[s].a[c0 + l] = fma(-tw[s], vr[l], cols[s].a[c0 + l]);
        }
      }
    }
  }
}

template <int NB, int SL, int CPL, class T, int SEG, int QF>
__device__ __forceinline__ void house_step(Col<NB, SL, T> (&cols)[CPL], Refl<NB, SL, T>& R, int buf, int sub,
                                           int piv_slot, unsigned upd, int k, T* V = nullptr, T* tauv = nullptr,
                                           int ldv = 0) {
#pragma unroll
  for (int s = 0; s < CPL; ++s)
    if (s == piv_slot) house_publish<NB, SL, T, SEG, QF>(cols[s], R, buf, sub, k, V, tauv, ldv);
  __syncthreads();
  if (upd) house_update<NB, SL, CPL, T, SEG, QF>(cols, R, buf, sub, upd);
}

/* Lanes per column group and threads of a QR kernel over COLS matrix
 * columns: 4 lanes per group up to 1024 threads, else 2. MINW is the waves
 * per SIMD asked of the register allocator so that W workgroups share a CU
 * (every Householder step is a latency chain; more workgroups hide it where
 * the columns and the reflector still fit in the registers they leave). */
/* WG > 0 overrides the table's workgroups per CU (the variants the run-time
 * autotuner chooses among); CPL is the number of columns each lane group
 * holds (GRP = ceil(COLS / CPL) groups). */
/* SW: the kernel can run its QR with house_qr_swap, which needs the
 * registers of COLS - 1 columns; SWAP is set where that saves a wave (the
 * phases' COLS are a multiple of the bucket plus one right-hand side, so at
 * n == NB the last column would otherwise open a wave of its own). */
#ifndef UK_QR_SWAP
#define UK_QR_SWAP 1
#endif
template <int NB, int COLS, int PHASE, int WG = 0, int CPL = 1, bool SW = false>
struct QRGeom {
  static constexpr int GRP0 = (COLS + CPL - 1) / CPL;
  static constexpr int SL = (NB <= 16 || 4 * GRP0 > 1024) ? 2 : 4;
  static constexpr int NT0 = ((SL * GRP0 + 63) / 64) * 64;
  static constexpr int GRP1 = (COLS - 1 + CPL - 1) / CPL;
  static constexpr int NT1 = ((SL * GRP1 + 63) / 64) * 64;
  static constexpr bool SWAP = UK_QR_SWAP && SW && NT1 < NT0;
  static constexpr int GRP = SWAP ? GRP1 : GRP0;
  static constexpr int NT = SWAP ? NT1 : NT0;
  static constexpr int W = WG > 0 ? WG : 1;
  static constexpr int MINW_ = (W * (NT / 64) + 3) / 4;
  static constexpr int MINW = MINW_ < 1 ? 1 : (MINW_ > 8 ? 8 : MINW_);
};

/* Pivot steps of segment SEG and the segments after it. */
template <int NB, int SL, int CPL, class T, int QF, int SEG = 0>
__device__ __forceinline__ void house_qr_segs(Col<NB, SL, T> (&cols)[CPL], Refl<NB, SL, T>& R, int grp, int sub,
                                              int piv0, int n, unsigned upd, T* V = nullptr, T* tauv = nullptr,
                                              int ldv = 0, int kbeg = 0) {
  constexpr int H = NB / SL, ROWS = qr_gl(H) * SL;
  if constexpr (SEG * ROWS < NB) {
    const int k1 = n < (SEG + 1) * ROWS ? n : (SEG + 1) * ROWS;
    for (int k = SEG * ROWS > kbeg ? SEG * ROWS : kbeg; k < k1; ++k) {
      const int pc = piv0 + k;
      unsigned u = upd;
      if (QF & QF_DONE) {
#pragma unroll
        for (int s = 0; s < CPL; ++s) {
          const int cc = grp * CPL + s;
          if (cc >= piv0 && cc < pc && !((QF & QF_SWAP) && cc == 0)) u &= ~(1u << s);
        }
      }
      house_step<NB, SL, CPL, T, SEG, QF>(cols, R, k & 1, sub, pc / CPL == grp ? pc % CPL : -1, u, k, V, tauv, ldv);
    }
    house_qr_segs<NB, SL, CPL, T, QF, SEG + 1>(cols, R, grp, sub, piv0, n, upd, V, tauv, ldv, kbeg);
  }
}

/* QR of the columns of a workgroup: pivots k = 0..n-1 on columns piv0 + k.
 * Lane group grp holds columns grp * CPL + s, s < CPL. QF: the structure
 * flags above (0: every granule of block 1, every column updated). */
template <int NB, int SL, int CPL, class T, int QF = 0>
__device__ __forceinline__ void house_qr(Col<NB, SL, T> (&cols)[CPL], Refl<NB, SL, T>& R, int grp, int sub, int piv0,
                                         int n, unsigned upd, T* V = nullptr, T* tauv = nullptr, int ldv = 0) {
  house_qr_segs<NB, SL, CPL, T, QF>(cols, R, grp, sub, piv0, n, upd, V, tauv, ldv);
  __syncthreads();
}
/* one column per lane group */
template <int NB, int SL, class T>
__device__ __forceinline__ void house_qr(Col<NB, SL, T>& col, Refl<NB, SL, T>& R, int mycol, int sub, int piv0, int n,
                                         bool upd) {
  house_qr<NB, SL, 1, T>(reinterpret_cast<Col<NB, SL, T>(&)[1]>(col), R, mycol, sub, piv0, n, upd ? 1u : 0u);
}

/*
 * QR of pivots k = 0..n-1 on columns 0..n-1 (piv0 = 0) with one column more
 * than the groups' registers hold. Column 0 (group 0, slot 0) is the first
 * pivot: after the first step it is final, and its one nonzero, R(0,0) in
 * row 0 (the lanes' first local; the rows below are what store_triu writes as
 * zero), goes to r00. Its registers then take `extra`, the workgroup's last
 * column (loaded by group 0 at the start), and the first reflector is applied
 * to it there: its LDS buffer is rewritten only by the third step's publish,
 * after a barrier this group has not reached yet. From the second step on the
 * steps run as in house_qr. Every column sees the same operations in the same
 * order as with one group more (house_update's order does not depend on CPL).
 * save0 (LDS, NB entries, or null): receives column 0's block 0 after the
 * first step, for a second QR over the same pivot columns.
 */
template <int NB, int SL, int CPL, class T, int QF>
__device__ __forceinline__ void house_qr_swap(Col<NB, SL, T> (&cols)[CPL], const Col<NB, SL, T>& extra, T& r00,
                                              Refl<NB, SL, T>& R, int grp, int sub, int n, unsigned upd,
                                              T* save0 = nullptr) {
  house_step<NB, SL, CPL, T, 0, QF>(cols, R, 0, sub, grp == 0 ? 0 : -1, upd, 0);
  if (grp == 0) {
    r00 = cols[0].a[0];
    if (save0) { /* column 0's block 0 as it stands (a second QR over the same pivots starts from it) */
#pragma unroll
      for (int l = 0; l < Col<NB, SL, T>::H; ++l) save0[sub * Col<NB, SL, T>::H + l] = cols[0].a[l];
    }
    cols[0] = extra;
    house_update<NB, SL, 1, T, 0, QF>(reinterpret_cast<Col<NB, SL, T>(&)[1]>(cols[0]), R, 0, sub, 1u);
  }
  house_qr_segs<NB, SL, CPL, T, QF | QF_SWAP>(cols, R, grp, sub, 0, n, upd, nullptr, nullptr, 0, 1);
  __syncthreads();
}

/* Column 0 after house_qr_swap: R(0,0) in row 0 (block 0), zeros below. */
template <int NB, int SL, class T>
__device__ __forceinline__ Col<NB, SL, T> swap_col0(T r00, int sub) {
  Col<NB, SL, T> c;
#pragma unroll
  for (int l = 0; l < 2 * Col<NB, SL, T>::H; ++l) c.a[l] = T(0);
  c.a[0] = sub == 0 ? r00 : T(0);
  return c;
}

template <int NB, int SL, int NT, class T>
__device__ __forceinline__ void refl_clear(Refl<NB, SL, T>& R, int tid) {
  for (int i = tid; i < (int)(sizeof(R) / sizeof(T)); i += NT) reinterpret_cast<T*>(&R)[i] = T(0);
  __syncthreads();
}

/*
 * Elimination of one level, three launches (the reference's three phases):
 *
 * k_phase1, one workgroup per even step j (:379-457). Pair (j, j+1): QR of
 *   [B_{j+1}; C_j] applied to [D_{j+1}; 0] and [c_{j+1}; o_j]. A lone last
 *   step (j = L-1) or a one-step chain: QR of [0; C_j] applied to [0; o_j].
 *   R~_j, X_j, o_j go to the factor slot of j (R, Y, cf), D~_{j+1} and
 *   c~_{j+1} to the next level's C and o slots of the new step.
 * k_phase2, one workgroup per even step j > 0 (:460-523): QR of [D_j; R~_j]
 *   applied to [B_j; 0], [c_j; o_j], [0; X_j]: final R, B~, Y, cf; the rest
 *   rows Z_j, X~_j, o_j become the new step's B, D, c (a lone last step's Z
 *   and o go to a per-problem scratch for the fix-last fold).
 * k_phase3, one workgroup per pair (:526-560): QR of [D~_{j+1}; C_{j+1}]
 *   applied to [c~_{j+1}; o_{j+1}]; in the last pair of an odd chain it then
 *   folds the lone step's [Z_{L-1}, o_{L-1}] in (:1452-1476). G~, c give the
 *   new step's C and o.
 * Columns (groups of SL lanes): P = [0,n) u_j, Q = [n,2n) u_{j+1},
 * R1 = 2n right-hand side, Bl = [2n+1,3n+1) u_{j-1}.
 */
/* LV: the level's stacked rows (C) as structure: 0 dense, 1 at most
 * qr_half_rows rows (level 0 with m <= NB / 2; 16 in bucket 24), 2 upper
 * triangular (the levels above 0) */
template <int NB, class T, int WG = 0, int CPL = 1, int LV = 0>
__global__ void __launch_bounds__((QRGeom<NB, 2 * NB + 1, 1, WG, CPL, true>::NT),
                                  (QRGeom<NB, 2 * NB + 1, 1, WG, CPL, true>::MINW))
    k_phase1(uk_level_t lv, ChainT<T> in, ChainT<T> out, FactorsT<T> fac, int n) {
  using G = QRGeom<NB, 2 * NB + 1, 1, WG, CPL, true>;
  constexpr int SL = G::SL;
  constexpr int NT = G::NT;
  constexpr bool SW = G::SWAP; /* R1 (the last column) swapped into column 0's registers */
  __shared__ Refl<NB, SL, T> refl;
  const int tid = threadIdx.x, sub = tid & (SL - 1), grp = tid / SL;
  const int64_t e = blockIdx.x;
  const int p = (int)find_problem(lv.even_start, lv.nprob, e);
  const int64_t u = e - lv.even_start[p];
  const int64_t j = 2 * u;
  const int64_t L = lv.L[p];
  const int64_t s0 = lv.step_start[p];
  const int64_t t = lv.next_start[p] + u;
  const int mr = in.mr;
  const int64_t nn = (int64_t)n * n;
  const bool lone = j + 1 == L;
  refl_clear<NB, SL, NT, T>(refl, tid);

  const int64_t sj = s0 + j, sp = s0 + j + 1;
  using St = Stage<NB, CPL, NT, T>;
  __shared__ T sbuf[sizeof(T) == 4 ? St::SIZE : 1]; /* FP32 kernels stage their stores */
  const St stg{sbuf};
  /* column c of the workgroup's matrix: loads it, returns whether it exists */
  auto load = [&](Col<NB, SL, T>& col, int c) -> bool {
    const bool isP = c < n, isQ = !lone && c >= n && c < 2 * n, isR1 = c == 2 * n;
    const int qc = isP ? c : (isQ ? c - n : 0);
    if (isP) {
      if (lone) col.zero_block(0);
      else if (in.b_synth) col.load_bsynth(0, in.wk + sp * n, in.wF + (in.wFidx ? in.wFidx[sp] * nn : 0), qc, n, sub);
      else col.load_block(0, in.B + sp * nn + (int64_t)qc * n, n, sub);
      if (in.c_synth) col.load_csynth(1, in.rC + sj * mr, in.rG + sj * (int64_t)mr * n, qc, mr, in.rm[sj], sub);
      else col.load_block(1, in.C + sj * nn + (int64_t)qc * mr, mr, sub);
    }
    if (isQ) {
      if (in.d_diag) col.load_diag(0, in.D + sp * nn, qc, sub); /* stored as its diagonal */
      else col.load_block(0, in.D + sp * nn + (int64_t)qc * n, n, sub);
      col.zero_block(1);
    }
    if (isR1) {
      if (lone) col.zero_block(0);
      else col.load_block(0, in.c + sp * n, n, sub);
      col.load_block(1, in.o + sj * n, mr, sub);
    }
    return isP || isQ || isR1;
  };
  auto store = [&](const Col<NB, SL, T>& col, int c) {
    const bool isP = c < n, isQ = !lone && c >= n && c < 2 * n, isR1 = c == 2 * n;
    const int qc = isP ? c : (isQ ? c - n : 0);
    if (isP) {
      col.store_triu(fac.R + e * nn + (int64_t)qc * n, n, qc, sub);
      for (int r = sub; r < n; r += SL) {
        if (lone) fac.Y[e * nn + (int64_t)qc * n + r] = T(0);
        if (j == 0) fac.Bt[e * nn + (int64_t)qc * n + r] = T(0);
      }
    }
    if (isQ) {
      col.store_block(0, fac.Y + e * nn + (int64_t)qc * n, n, sub);     /* X_j */
      col.store_block(1, out.C + t * nn + (int64_t)qc * mr, mr, sub);  /* D~_{j+1} */
    }
    if (isR1) {
      col.store_block(0, fac.cf + e * n, n, sub);                       /* o_j */
      if (!lone) col.store_block(1, out.o + t * n, mr, sub);           /* c~_{j+1} */
    }
  };

  Col<NB, SL, T> cols[CPL];
  Col<NB, SL, T> extra;
  unsigned upd = 0;
#pragma unroll
  for (int q = 0; q < CPL; ++q) {
    const int c = grp * CPL + q;
    if (!(SW && c == 2 * n) && load(cols[q], c)) upd |= 1u << q;
    else if (CPL > 1) cols[q].zero_block(0), cols[q].zero_block(1); /* computed on as zeros, never stored */
  }
  if (SW && grp == 0) load(extra, 2 * n);
  /* LV 1: the stacked rows C_j, o_j fill at most qr_half_rows rows; LV 2: C_j is
   * upper triangular (the pivot columns' block 1) */
  constexpr int QF = QF_DONE | (LV == 1 ? QF_HALF1 : 0) | (LV == 2 ? QF_TRI1 : 0);
  T r00 = T(0);
  if constexpr (SW) house_qr_swap<NB, SL, CPL, T, QF>(cols, extra, r00, refl, grp, sub, n, upd);
  else house_qr<NB, SL, CPL, T, QF>(cols, refl, grp, sub, 0, n, upd);
  if constexpr (sizeof(T) == 4) { /* FP32: the output blocks through the stage */
    /* P: R~_j (upper triangular); a lone step's X_j and the first step's B~ are zero */
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      const int c = grp * CPL + q;
      if (SW && c == 0) swap_col0<NB, SL, T>(r00, sub).write_stage_triu(stg.col(0), 0, sub);
      else if (c < n) cols[q].write_stage_triu(stg.col(c), c, sub);
    }
    __syncthreads();
    stg.out(fac.R + e * nn, n, n, n, tid);
    if (lone) block_zero<NT>(fac.Y + e * nn, nn, tid);
    if (j == 0) block_zero<NT>(fac.Bt + e * nn, nn, tid);
    if (!lone) { /* Q: X_j, then D~_{j+1} (ld mr) */
      __syncthreads();
#pragma unroll
      for (int q = 0; q < CPL; ++q) {
        const int c = grp * CPL + q;
        if (c >= n && c < 2 * n) cols[q].write_stage(0, stg.col(c - n), sub);
      }
      __syncthreads();
      stg.out(fac.Y + e * nn, n, n, n, tid);
      __syncthreads();
#pragma unroll
      for (int q = 0; q < CPL; ++q) {
        const int c = grp * CPL + q;
        if (c >= n && c < 2 * n) cols[q].write_stage(1, stg.col(c - n), sub);
      }
      __syncthreads();
      stg.out(out.C + t * nn, mr, mr, n, tid);
    }
    /* R1: o_j and c~_{j+1} */
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      const int c = grp * CPL + q;
      if (SW ? c == 0 : c == 2 * n) {
        cols[q].store_block(0, fac.cf + e * n, n, sub);
        if (!lone) cols[q].store_block(1, out.o + t * n, mr, sub);
      }
    }
  } else { /* FP64: straight from the registers */
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      const int c = grp * CPL + q;
      if (SW && c == 0) {
        store(cols[0], 2 * n);
        store(swap_col0<NB, SL, T>(r00, sub), 0);
      } else if (!(SW && c == 2 * n)) {
        store(cols[q], c);
      }
    }
  }
}

template <int NB, class T, int WG = 0, int CPL = 1, bool DIAG = false>
__global__ void __launch_bounds__((QRGeom<NB, 3 * NB + 1, 2, WG, CPL, true>::NT),
                                  (QRGeom<NB, 3 * NB + 1, 2, WG, CPL, true>::MINW))
    k_phase2(uk_level_t lv, ChainT<T> in, ChainT<T> out, FactorsT<T> fac, T* __restrict__ fix, int n) {
  using G = QRGeom<NB, 3 * NB + 1, 2, WG, CPL, true>;
  constexpr int SL = G::SL;
  constexpr int NT = G::NT;
  constexpr bool SW = G::SWAP; /* the last column of Bl swapped into column 0's registers */
  __shared__ Refl<NB, SL, T> refl;
  const int tid = threadIdx.x, sub = tid & (SL - 1), grp = tid / SL;
  const int64_t e = blockIdx.x;
  const int p = (int)find_problem(lv.even_start, lv.nprob, e);
  const int64_t u = e - lv.even_start[p];
  const int64_t j = 2 * u;
  if (j == 0) return; /* the chain's first step has no [D; R~] (uniform per block) */
  const int64_t L = lv.L[p];
  const int64_t t = lv.next_start[p] + u;
  const int64_t nn = (int64_t)n * n;
  const bool lone = j + 1 == L;
  refl_clear<NB, SL, NT, T>(refl, tid);

  const int64_t sj = lv.step_start[p] + j;
  T* fz = fix + (int64_t)p * (nn + n);
  using St = Stage<NB, CPL, NT, T>;
  __shared__ T sbuf[sizeof(T) == 4 ? St::SIZE : 1]; /* FP32 kernels stage their stores */
  const St stg{sbuf};
  auto load = [&](Col<NB, SL, T>& col, int c) -> bool {
    const bool isP = c < n, isQ = c >= n && c < 2 * n, isR1 = c == 2 * n, isBl = c > 2 * n && c <= 3 * n;
    const int qc = isP ? c : (isQ ? c - n : (isBl ? c - 2 * n - 1 : 0));
    if (isP) {
      if (DIAG) col.load_diag(0, in.D + sj * nn, qc, sub); /* stored as its diagonal */
      else col.load_block(0, in.D + sj * nn + (int64_t)qc * n, n, sub);
      col.load_block(1, fac.R + e * nn + (int64_t)qc * n, n, sub); /* R~_j */
    }
    if (isQ) {
      col.zero_block(0);
      col.load_block(1, fac.Y + e * nn + (int64_t)qc * n, n, sub); /* X_j (zero for a lone step) */
    }
    if (isR1) {
      col.load_block(0, in.c + sj * n, n, sub);
      col.load_block(1, fac.cf + e * n, n, sub); /* o_j */
    }
    if (isBl) {
      if (in.b_synth) col.load_bsynth(0, in.wk + sj * n, in.wF + (in.wFidx ? in.wFidx[sj] * nn : 0), qc, n, sub);
      else col.load_block(0, in.B + sj * nn + (int64_t)qc * n, n, sub);
      col.zero_block(1);
    }
    return isP || isQ || isR1 || isBl;
  };
  auto store = [&](const Col<NB, SL, T>& col, int c) {
    const bool isP = c < n, isQ = c >= n && c < 2 * n, isR1 = c == 2 * n, isBl = c > 2 * n && c <= 3 * n;
    const int qc = isP ? c : (isQ ? c - n : (isBl ? c - 2 * n - 1 : 0));
    if (isP) col.store_triu(fac.R + e * nn + (int64_t)qc * n, n, qc, sub);
    if (isBl) {
      col.store_block(0, fac.Bt + e * nn + (int64_t)qc * n, n, sub);
      col.store_block(1, (lone ? fz : out.B + t * nn) + (int64_t)qc * n, n, sub); /* Z_j */
    }
    if (isQ) {
      col.store_block(0, fac.Y + e * nn + (int64_t)qc * n, n, sub);
      if (!lone) col.store_block(1, out.D + t * nn + (int64_t)qc * n, n, sub); /* X~_j */
    }
    if (isR1) {
      col.store_block(0, fac.cf + e * n, n, sub);
      col.store_block(1, lone ? fz + nn : out.c + t * n, n, sub); /* o_j */
    }
  };

  Col<NB, SL, T> cols[CPL];
  Col<NB, SL, T> extra;
  unsigned upd = 0;
#pragma unroll
  for (int q = 0; q < CPL; ++q) {
    const int c = grp * CPL + q;
    if (!(SW && c == 3 * n) && load(cols[q], c)) upd |= 1u << q;
    else if (CPL > 1) cols[q].zero_block(0), cols[q].zero_block(1); /* computed on as zeros, never stored */
  }
  if (SW && grp == 0) load(extra, 3 * n);
  /* block 1 of the pivot columns is R~_j, upper triangular */
  /* DIAG: D_j is diagonal (the chain's d_diag) */
  constexpr int QF = QF_TRI1 | QF_DONE | (DIAG ? QF_DIAG0 : 0);
  T r00 = T(0);
  if constexpr (SW) house_qr_swap<NB, SL, CPL, T, QF>(cols, extra, r00, refl, grp, sub, n, upd);
  else house_qr<NB, SL, CPL, T, QF>(cols, refl, grp, sub, 0, n, upd);
  if constexpr (sizeof(T) == 4) { /* FP32: the output blocks through the stage */
    /* the register slot holding Bl column c (with the swap, the last one is in group 0's slot 0) */
    auto bl_slot = [&](int q) -> int {
      const int c = grp * CPL + q;
      if (SW && c == 0) return n - 1;
      if (SW && c == 3 * n) return -1;
      return (c > 2 * n && c <= 3 * n) ? c - 2 * n - 1 : -1;
    };
    /* P: the final R (upper triangular) */
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      const int c = grp * CPL + q;
      if (SW && c == 0) swap_col0<NB, SL, T>(r00, sub).write_stage_triu(stg.col(0), 0, sub);
      else if (c < n) cols[q].write_stage_triu(stg.col(c), c, sub);
    }
    __syncthreads();
    stg.out(fac.R + e * nn, n, n, n, tid);
    __syncthreads();
    /* Bl: B~ (block 0), then Z_j (block 1) to the new step's B or a lone step's scratch */
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      const int b = bl_slot(q);
      if (b >= 0) cols[q].write_stage(0, stg.col(b), sub);
    }
    __syncthreads();
    stg.out(fac.Bt + e * nn, n, n, n, tid);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      const int b = bl_slot(q);
      if (b >= 0) cols[q].write_stage(1, stg.col(b), sub);
    }
    __syncthreads();
    stg.out(lone ? fz : out.B + t * nn, n, n, n, tid);
    __syncthreads();
    /* Q: Y (block 0), then X~_j (block 1) to the new step's D */
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      const int c = grp * CPL + q;
      if (c >= n && c < 2 * n) cols[q].write_stage(0, stg.col(c - n), sub);
    }
    __syncthreads();
    stg.out(fac.Y + e * nn, n, n, n, tid);
    if (!lone) {
      __syncthreads();
#pragma unroll
      for (int q = 0; q < CPL; ++q) {
        const int c = grp * CPL + q;
        if (c >= n && c < 2 * n) cols[q].write_stage(1, stg.col(c - n), sub);
      }
      __syncthreads();
      stg.out(out.D + t * nn, n, n, n, tid);
    }
    /* R1: cf (block 0) and o_j (block 1) */
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      const int c = grp * CPL + q;
      if (c == 2 * n) {
        cols[q].store_block(0, fac.cf + e * n, n, sub);
        cols[q].store_block(1, lone ? fz + nn : out.c + t * n, n, sub);
      }
    }
  } else { /* FP64: straight from the registers */
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      const int c = grp * CPL + q;
      if (SW && c == 0) {
        store(cols[0], 3 * n);
        store(swap_col0<NB, SL, T>(r00, sub), 0);
      } else if (!(SW && c == 3 * n)) {
        store(cols[q], c);
      }
    }
  }
}

template <int NB, class T, int WG = 0, int CPL = 1, int LV = 0>
__global__ void __launch_bounds__((QRGeom<NB, NB + 1, 3, WG, CPL, true>::NT),
                                  (QRGeom<NB, NB + 1, 3, WG, CPL, true>::MINW))
    k_phase3(uk_level_t lv, ChainT<T> in, ChainT<T> out, const T* __restrict__ fix, int n) {
  using G = QRGeom<NB, NB + 1, 3, WG, CPL, true>;
  constexpr int SL = G::SL;
  constexpr int NT = G::NT;
  constexpr bool SW = G::SWAP; /* the right-hand side swapped into column 0's registers */
  __shared__ Refl<NB, SL, T> refl;
  const int tid = threadIdx.x, sub = tid & (SL - 1), grp = tid / SL;
  const int64_t unit = blockIdx.x;
  const int p = (int)find_problem(lv.unit_start, lv.nprob, unit);
  const int64_t L = lv.L[p];
  if (L == 1) return; /* uniform per block */
  const int64_t u = unit - lv.unit_start[p];
  const int64_t j = 2 * u;
  const int64_t t = lv.next_start[p] + u;
  const int mr = in.mr;
  const int64_t nn = (int64_t)n * n;
  const bool do_fix = (L & 1) && (u == L / 2 - 1);
  refl_clear<NB, SL, NT, T>(refl, tid);

  /* Q columns numbered from 0 here */
  const int64_t sp = lv.step_start[p] + j + 1;
  using St = Stage<NB, CPL, NT, T>;
  __shared__ T sbuf[sizeof(T) == 4 ? St::SIZE : 1]; /* FP32 kernels stage their stores */
  const St stg{sbuf};
  auto load = [&](Col<NB, SL, T>& col, int c) -> bool {
    const bool isQ = c < n, isR1 = c == n;
    const int qc = isQ ? c : 0;
    if (isQ) {
      col.load_block(0, out.C + t * nn + (int64_t)qc * mr, mr, sub); /* D~_{j+1} */
      if (in.c_synth) col.load_csynth(1, in.rC + sp * mr, in.rG + sp * (int64_t)mr * n, qc, mr, in.rm[sp], sub);
      else col.load_block(1, in.C + sp * nn + (int64_t)qc * mr, mr, sub);
    }
    if (isR1) {
      col.load_block(0, out.o + t * n, mr, sub); /* c~_{j+1} */
      col.load_block(1, in.o + sp * n, mr, sub);
    }
    return isQ || isR1;
  };
  Col<NB, SL, T> cols[CPL];
  Col<NB, SL, T> extra;
  unsigned upd = 0;
#pragma unroll
  for (int q = 0; q < CPL; ++q) {
    const int c = grp * CPL + q;
    if (!(SW && c == n) && load(cols[q], c)) upd |= 1u << q;
    else if (CPL > 1) cols[q].zero_block(0), cols[q].zero_block(1); /* computed on as zeros, never stored */
  }
  if (SW && grp == 0) load(extra, n);
  /* LV 1: both blocks (D~_{j+1}, C_{j+1} and the right-hand sides) fill at
   * most qr_half_rows rows; LV 2: C_{j+1} (block 1 of the pivot columns) is upper
   * triangular */
  constexpr int QF = QF_DONE | (LV == 1 ? QF_HALF0 | QF_HALF1 : 0) | (LV == 2 ? QF_TRI1 : 0);
  T r00 = T(0);
  __shared__ T p0blk[SW ? NB : 1]; /* column 0's block 0 for the fold */
  if constexpr (SW) house_qr_swap<NB, SL, CPL, T, QF>(cols, extra, r00, refl, grp, sub, n, upd, do_fix ? p0blk : nullptr);
  else house_qr<NB, SL, CPL, T, QF>(cols, refl, grp, sub, 0, n, upd);
  if (do_fix) {
    /* [G~; Z] and [c; o_Z]: block 1 from the lone step's scratch */
    const T* fz = fix + (int64_t)p * (nn + n);
    if constexpr (SW) {
      if (grp == 0) { /* the right-hand side leaves column 0's registers for the second QR */
        extra = cols[0];
        extra.load_block(1, fz + nn, n, sub);
#pragma unroll
        for (int l = 0; l < Col<NB, SL, T>::H; ++l) cols[0].a[l] = p0blk[sub * Col<NB, SL, T>::H + l];
      }
    }
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      const int c = grp * CPL + q;
      const int qc = c < n ? c : 0;
      if (c < n) cols[q].load_block(1, fz + (int64_t)qc * n, n, sub);
      if (!SW && c == n) cols[q].load_block(1, fz + nn, n, sub);
    }
    if constexpr (SW) house_qr_swap<NB, SL, CPL, T, QF_DONE>(cols, extra, r00, refl, grp, sub, n, upd);
    else house_qr<NB, SL, CPL, T, QF_DONE>(cols, refl, grp, sub, 0, n, upd);
  }
  if constexpr (sizeof(T) == 4) { /* FP32: the output blocks through the stage */
    /* G~ (upper triangular) to the new step's C; the right-hand side to its o */
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      const int c = grp * CPL + q;
      if (SW && c == 0) {
        swap_col0<NB, SL, T>(r00, sub).write_stage_triu(stg.col(0), 0, sub);
        cols[0].store_block(0, out.o + t * n, n, sub);
      } else if (c < n) {
        cols[q].write_stage_triu(stg.col(c), c, sub);
      } else if (!SW && c == n) {
        cols[q].store_block(0, out.o + t * n, n, sub);
      }
    }
    __syncthreads();
    stg.out(out.C + t * nn, n, n, n, tid);
  } else { /* FP64: straight from the registers */
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      const int c = grp * CPL + q;
      const int qc = c < n ? c : 0;
      if (SW && c == 0) {
        cols[0].store_block(0, out.o + t * n, n, sub);
        swap_col0<NB, SL, T>(r00, sub).store_triu(out.C + t * nn, n, 0, sub);
      } else if (c < n) {
        cols[q].store_triu(out.C + t * nn + (int64_t)qc * n, n, qc, sub); /* G~ */
      } else if (!SW && c == n) {
        cols[q].store_block(0, out.o + t * n, n, sub);
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* blocked WY on the matrix cores (phase 2, buckets 32..64)                   */
/* ------------------------------------------------------------------------ */

#ifndef UK_MFMA_F32_ROW
#define UK_MFMA_F32_ROW(r, lk) (4 * (lk) + (r))
#endif
typedef double wy_d4 __attribute__((ext_vector_type(4)));
typedef float wy_f4 __attribute__((ext_vector_type(4)));

/* 16 x 16 x 4 matrix-core products in T. Operands: a = A[li][lk] (A 16 x 4),
 * b = B[lk][li] (B 4 x 16), lane = li + 16 lk; the accumulator's element r of
 * lane (li, lk) is D[row(r, lk)][li]. */
template <class T>
struct WyMM;
template <>
struct WyMM<double> {
  using acc = wy_d4;
  static __device__ __forceinline__ acc zero() { return acc{0.0, 0.0, 0.0, 0.0}; }
  static __device__ __forceinline__ acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int r, int lk) { return 4 * r + lk; }
};
template <>
struct WyMM<float> {
  using acc = wy_f4;
  static __device__ __forceinline__ acc zero() { return acc{0.f, 0.f, 0.f, 0.f}; }
  static __device__ __forceinline__ acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int r, int lk) { return UK_MFMA_F32_ROW(r, lk); }
};

/* Orders one wave's LDS accesses: its earlier LDS writes are visible to its
 * later LDS reads (of other lanes), and the compiler moves no access across. */
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* Workgroup of the blocked phase 2: SL = 4 lanes for each of the NB pivot
 * columns and the right-hand side column during the panel; its NW waves then
 * take 16-column blocks of the trailing columns. */
template <int NB>
struct WyGeom {
  static constexpr int SL = 4;
  static constexpr int NT = ((SL * (NB + 1) + 63) / 64) * 64;
  static constexpr int NW = NT / 64;
  static constexpr int LDV = 2 * NB + 1; /* odd: both fragment read patterns spread over the banks */
  static constexpr int LDW = 17;
  static constexpr int NG = NB / 16;     /* reflector groups of 16 */
  static constexpr int RT = 2 * NB / 16; /* row tiles of a trailing column block */
};

template <int NB, class T>
struct WyShared {
  Refl<NB, 4, T> refl;
  T V[WyGeom<NB>::LDV * NB]; /* the reflectors u_k, column k (zero above the pivot row) */
  T tau[NB];
  T Tg[WyGeom<NB>::NG][16 * WyGeom<NB>::LDW]; /* per group g: S_g, then T_g (16 x 16, upper triangular) */
};

/*
 * Phase 2 in blocked WY form (c/kalman_oddeven_smoother.c:460-523), one
 * workgroup per even step j > 0, the same quantities as k_phase2:
 *  1. panel: Householder QR of [D_j; R~_j] (the n pivot columns) with the
 *     right-hand side [c_j; o_j] updated alongside, on the register/DPP
 *     machinery of k_phase2; every reflector u_k and tau_k is also kept in V;
 *  2. per group g of 16 reflectors, the compact WY factor T_g from
 *     T_g^{-1} = striu(V_g^T V_g) + diag(1 / tau): V_g^T V_g on the matrix
 *     cores, the 16 x 16 triangular inverse by substitution (one lane per
 *     column);
 *  3. the 2n trailing columns [B_j; 0] and [0; X_j], in 16-column blocks (one
 *     per wave at a time, held as matrix-core accumulators): for each group in
 *     order, W = V_g^T B, W = T_g^T W, B -= V_g W (so Q^T B = prod_g (I - V_g
 *     T_g^T V_g^T) B), then stored to B~_j, Z_j, Y_j, X~_j.
 * The accumulators serve as the B operands of the next products: a k-step
 * pairs A[.][kappa] with B[kappa][.] for kappa = the accumulator row that
 * lane holds (any assignment of the K indices to the steps gives the same
 * sum), so no operand passes through LDS.
 */
template <int NB, class T, bool DIAG = false>
__global__ void __launch_bounds__(WyGeom<NB>::NT, 1)
    k_phase2_wy(uk_level_t lv, ChainT<T> in, ChainT<T> out, FactorsT<T> fac, T* __restrict__ fix, int n) {
  using Gm = WyGeom<NB>;
  using MM = WyMM<T>;
  using acc_t = typename MM::acc;
  constexpr int SL = Gm::SL, NT = Gm::NT, NW = Gm::NW, LDV = Gm::LDV, LDW = Gm::LDW, RT = Gm::RT;
  __shared__ WyShared<NB, T> sh;
  const int tid = threadIdx.x, sub = tid & (SL - 1), grp = tid / SL;
  const int lane = tid & 63, w = tid >> 6, li = lane & 15, lk = lane >> 4;
  const int64_t e = blockIdx.x;
  const int p = (int)find_problem(lv.even_start, lv.nprob, e);
  const int64_t u = e - lv.even_start[p];
  const int64_t j = 2 * u;
  if (j == 0) return; /* the chain's first step has no [D; R~] (uniform per block) */
  const int64_t L = lv.L[p];
  const int64_t t = lv.next_start[p] + u;
  const int64_t nn = (int64_t)n * n;
  const bool lone = j + 1 == L;
  const int64_t sj = lv.step_start[p] + j;
  T* fz = fix + (int64_t)p * (nn + n);
  for (int i = tid; i < LDV * NB; i += NT) sh.V[i] = T(0);
  for (int i = tid; i < NB; i += NT) sh.tau[i] = T(0);
  refl_clear<NB, SL, NT, T>(sh.refl, tid);

  /* ---- 1. panel ---- */
  {
    Col<NB, SL, T> col;
    const int c = grp;
    const bool isP = c < n, isR1 = c == NB;
    if (isP) {
      if (DIAG) col.load_diag(0, in.D + sj * nn, c, sub);
      else col.load_block(0, in.D + sj * nn + (int64_t)c * n, n, sub);
      col.load_block(1, fac.R + e * nn + (int64_t)c * n, n, sub); /* R~_j */
    } else if (isR1) {
      col.load_block(0, in.c + sj * n, n, sub);
      col.load_block(1, fac.cf + e * n, n, sub); /* o_j */
    } else {
      col.zero_block(0);
      col.zero_block(1);
    }
    Col<NB, SL, T>(&cols)[1] = reinterpret_cast<Col<NB, SL, T>(&)[1]>(col);
    house_qr<NB, SL, 1, T, QF_TRI1 | QF_DONE | (DIAG ? QF_DIAG0 : 0)>(cols, sh.refl, grp, sub, 0, n,
                                                                     (isP || isR1) ? 1u : 0u, sh.V, sh.tau, LDV);
    if (isP) col.store_triu(fac.R + e * nn + (int64_t)c * n, n, c, sub);
    if (isR1) {
      col.store_block(0, fac.cf + e * n, n, sub);
      col.store_block(1, lone ? fz + nn : out.c + t * n, n, sub); /* o_j */
    }
  }
  const int ng = (n + 15) >> 4;

  /* ---- 2. T_g, wave w takes groups w, w + NW, ... ---- */
  for (int g = w; g < ng; g += NW) {
    acc_t acc = MM::zero();
    const T* Vg = sh.V + (int64_t)g * 16 * LDV;
#pragma unroll 4
    for (int ks = 0; ks < 2 * NB / 4; ++ks) {
      const int k = 4 * ks + lk;
      acc = MM::mma(Vg[k + li * LDV], Vg[k + li * LDV], acc); /* A = V_g^T, B = V_g */
    }
    T* S = sh.Tg[g]; /* S = striu(V_g^T V_g) + diag(1 / tau), column-major */
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = MM::row(r, lk), c = li;
      const T tv = sh.tau[g * 16 + i];
      S[i + c * LDW] = i < c ? acc[r] : (i == c ? (tv != T(0) ? T(1) / tv : T(1)) : T(0));
    }
    wave_lds_sync();
    T x[16];
    if (lane < 16) { /* column `lane` of T_g = S^{-1} by back substitution */
#pragma unroll
      for (int i = 0; i < 16; ++i) x[i] = i == lane ? T(1) : T(0);
#pragma unroll
      for (int q = 15; q >= 0; --q) {
        const T xq = x[q] / S[q + q * LDW];
        x[q] = xq;
#pragma unroll
        for (int i = 0; i < q; ++i) x[i] -= S[i + q * LDW] * xq;
      }
    }
    wave_lds_sync();
    if (lane < 16) {
#pragma unroll
      for (int i = 0; i < 16; ++i) S[i + lane * LDW] = x[i];
    }
    wave_lds_sync();
  }
  __syncthreads();

  /* ---- 3. the trailing columns, 16 at a time: block b < nb1 of [B_j; 0], then of [0; X_j] ---- */
  const int nb1 = (n + 15) >> 4;
  for (int blk = w; blk < 2 * nb1; blk += NW) {
    const bool isB = blk < nb1;
    const int cc = (isB ? blk : blk - nb1) * 16 + li; /* this lane's column */
    const bool colok = cc < n;
    /* B: row tile rt, element r = row rt * 16 + row(r, lk), column cc */
    acc_t b[RT];
    const double* kw = in.b_synth ? in.wk + sj * n : nullptr;
    const double* Fs = in.b_synth ? in.wF + (in.wFidx ? in.wFidx[sj] * nn : 0) : nullptr;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rt * 16 + MM::row(r, lk);
        T v = T(0);
        if (colok) {
          if (isB && row < n) {
            v = in.b_synth ? T(-kw[row] * Fs[row + (int64_t)cc * n]) : in.B[sj * nn + (int64_t)cc * n + row];
          } else if (!isB && row >= NB && row - NB < n) {
            v = fac.Y[e * nn + (int64_t)cc * n + (row - NB)]; /* X_j (zero for a lone step) */
          }
        }
        b[rt][r] = v;
      }
    for (int g = 0; g < ng; ++g) {
      const T* Vg = sh.V + (int64_t)g * 16 * LDV;
      const T* Tg = sh.Tg[g];
      /* W = V_g^T B (K = 2 NB; for the first group the block's zero half is skipped) */
      acc_t wv = MM::zero();
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
        if (g == 0 && (isB ? rt >= RT / 2 : rt < RT / 2)) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int k = rt * 16 + MM::row(r, lk);
          wv = MM::mma(Vg[k + li * LDV], b[rt][r], wv);
        }
      }
      /* W <- T_g^T W (A[m][k] = T_g[k][m]) */
      acc_t w2 = MM::zero();
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = MM::row(r, lk);
        w2 = MM::mma(Tg[k + li * LDW], wv[r], w2);
      }
      /* B -= V_g W */
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int k = MM::row(r, lk);
          b[rt] = MM::mma(-Vg[rt * 16 + li + k * LDV], w2[r], b[rt]);
        }
    }
    /* store: block 0 rows -> B~_j / Y_j, block 1 rows -> Z_j / X~_j */
    if (colok) {
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = rt * 16 + MM::row(r, lk);
          const bool top = row < NB;
          const int rr = top ? row : row - NB;
          if (rr >= n) continue;
          const T v = b[rt][r];
          if (isB) {
            if (top) fac.Bt[e * nn + (int64_t)cc * n + rr] = v;
            else (lone ? fz : out.B + t * nn)[(int64_t)cc * n + rr] = v; /* Z_j */
          } else {
            if (top) fac.Y[e * nn + (int64_t)cc * n + rr] = v;
            else if (!lone) out.D[t * nn + (int64_t)cc * n + rr] = v; /* X~_j */
          }
        }
    }
  }
}

/* Phases of one level in bucket NB with launch variant v (NB <= 16: the one
 * variant). Per level: the structure of its C blocks (LV of phases 1 and 3:
 * at most half a block of rows at level 0, upper triangular above), and DIAG:
 * its D blocks are diagonal. */
template <int NB, class T, int WG, int CPL>
void phases_nb(const uk_level_t* lv, const ChainT<T>& in, const ChainT<T>& out, const FactorsT<T>& fac, T* fix, int n,
               hipStream_t st, int phase) {
  const int lvk = in.c_tri ? 2 : (in.mr <= qr_half_rows(NB, QRGeom<NB, 2 * NB + 1, 1, WG, CPL, true>::SL) ? 1 : 0);
  static_assert(QRGeom<NB, 2 * NB + 1, 1, WG, CPL, true>::SL == QRGeom<NB, NB + 1, 3, WG, CPL, true>::SL,
                "phases 1 and 3 share the half-block row count");
  if (phase == 1) {
    const dim3 g((unsigned)lv->evens), b(QRGeom<NB, 2 * NB + 1, 1, WG, CPL, true>::NT);
    if (lvk == 2) k_phase1<NB, T, WG, CPL, 2><<<g, b, 0, st>>>(*lv, in, out, fac, n);
    else if (lvk == 1) k_phase1<NB, T, WG, CPL, 1><<<g, b, 0, st>>>(*lv, in, out, fac, n);
    else k_phase1<NB, T, WG, CPL, 0><<<g, b, 0, st>>>(*lv, in, out, fac, n);
  } else if (phase == 2) {
    const dim3 g((unsigned)lv->evens), b(QRGeom<NB, 3 * NB + 1, 2, WG, CPL, true>::NT);
    if (in.d_diag) k_phase2<NB, T, WG, CPL, true><<<g, b, 0, st>>>(*lv, in, out, fac, fix, n);
    else k_phase2<NB, T, WG, CPL, false><<<g, b, 0, st>>>(*lv, in, out, fac, fix, n);
  } else {
    const dim3 g((unsigned)lv->units), b(QRGeom<NB, NB + 1, 3, WG, CPL, true>::NT);
    if (lvk == 2) k_phase3<NB, T, WG, CPL, 2><<<g, b, 0, st>>>(*lv, in, out, fix, n);
    else if (lvk == 1) k_phase3<NB, T, WG, CPL, 1><<<g, b, 0, st>>>(*lv, in, out, fix, n);
    else k_phase3<NB, T, WG, CPL, 0><<<g, b, 0, st>>>(*lv, in, out, fix, n);
  }
}

template <int NB, class T>
void phase_launch_impl(const uk_level_t* lv, const ChainT<T>& in, const ChainT<T>& out, const FactorsT<T>& fac, T* fix,
                       int n, hipStream_t st, int phase, ukqr::Launch v) {
  if constexpr (NB >= 24) {
    if constexpr (NB % 16 == 0) {
      if (v.cpl == 0 && phase == 2) { /* blocked WY on the matrix cores */
        const dim3 g((unsigned)lv->evens), b(WyGeom<NB>::NT);
        if (in.d_diag) k_phase2_wy<NB, T, true><<<g, b, 0, st>>>(*lv, in, out, fac, fix, n);
        else k_phase2_wy<NB, T, false><<<g, b, 0, st>>>(*lv, in, out, fac, fix, n);
        return;
      }
    }
    if (v.cpl == 3) {
      switch (v.wg) {
        case 2: phases_nb<NB, T, 2, 3>(lv, in, out, fac, fix, n, st, phase); break;
        default: phases_nb<NB, T, 3, 3>(lv, in, out, fac, fix, n, st, phase); break;
      }
    } else if (v.cpl == 2) {
      switch (v.wg) {
        case 2: phases_nb<NB, T, 2, 2>(lv, in, out, fac, fix, n, st, phase); break;
        case 3: phases_nb<NB, T, 3, 2>(lv, in, out, fac, fix, n, st, phase); break;
        default: phases_nb<NB, T, 4, 2>(lv, in, out, fac, fix, n, st, phase); break;
      }
    } else {
      switch (v.wg) {
        case 1: phases_nb<NB, T, 1, 1>(lv, in, out, fac, fix, n, st, phase); break;
        case 2: phases_nb<NB, T, 2, 1>(lv, in, out, fac, fix, n, st, phase); break;
        default: phases_nb<NB, T, 4, 1>(lv, in, out, fac, fix, n, st, phase); break;
      }
    }
  } else {
    (void)v;
    phases_nb<NB, T, 0, 1>(lv, in, out, fac, fix, n, st, phase);
  }
}
}  // namespace

/* The per-bucket launchers' definitions, for the translation unit of bucket NBV. */
#define UKQR_DEFINE(NBV)                                                                                             \
  void ukqr::phase_launch_b##NBV(const uk_level_t* lv, const ChainT<double>& in, const ChainT<double>& out,           \
                                 const FactorsT<double>& fac, double* fix, int n, hipStream_t st, int phase,          \
                                 Launch v) {                                                                          \
    phase_launch_impl<NBV, double>(lv, in, out, fac, fix, n, st, phase, v);                                           \
  }                                                                                                                   \
  void ukqr::phase_launch_b##NBV(const uk_level_t* lv, const ChainT<float>& in, const ChainT<float>& out,             \
                                 const FactorsT<float>& fac, float* fix, int n, hipStream_t st, int phase,            \
                                 Launch v) {                                                                          \
    phase_launch_impl<NBV, float>(lv, in, out, fac, fix, n, st, phase, v);                                            \
  }

#endif /* UKHIP_ODDEVEN_QR_H */
