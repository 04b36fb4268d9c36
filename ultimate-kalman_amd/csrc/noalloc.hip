/*
 * noalloc.hip -- the allocation-free odd-even entry point,
 * ukhip_smooth_oddeven_noalloc (include/ukhip.h). It has no counterpart in
 * the reference: it serves callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).
 *
 * It is ukhip_batch_create + ukhip_batch_smooth_oddeven for one problem
 * whose inputs and outputs already live in device memory: the same whitening,
 * elimination and back-substitution kernels on the same data layout, so its
 * results equal the batch path's bit for bit. Every intermediate buffer (the
 * per-step index arrays of the raw inputs, the level tables, both chains, the
 * factor slots, the cross covariances, the fix-last scratch) is carved from
 * `scratch`, a file-scope device array of UKHIP_NOALLOC_SCRATCH_BYTES that
 * each GPU holds until the process ends. The call makes no hipMalloc, hipFree
 * or host heap allocation.
 *
 * In host code `scratch` names a host-side stand-in, not device memory: its
 * device address is looked up once per device with hipGetSymbolAddress (with
 * that device current) and only that address is handed to kernels.
 *
 * Every call on a GPU shares that GPU's scratch: calls that overlap on one
 * GPU (on different streams, or from several host threads) race. Nothing in
 * scratch has to start at zero: every element a kernel reads is written
 * earlier in the same call (k_noalloc_init writes the index arrays and level
 * tables, the smoothing writes the rest before reading it).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "../../include/ukhip.h"
#include "ukhip_internal.h"

static __device__ __align__(256) unsigned char scratch[UKHIP_NOALLOC_SCRATCH_BYTES];

extern "C" int uk_run_oddeven_view(uk_oe_view_t* v, int cov, hipStream_t st);
extern "C" void uk_set_error(const char* fmt, ...);
extern "C" void* ukhip_stream(void);

namespace {

constexpr int MAX_LEVELS = 64;
constexpr int MAX_DEVICES = 64;
/* level table: [0] orig_start (0), [1] prob (int32 0, twice), then per level
 * 8 int64: L, unit_start[0..1], even_start[0..1], step_start, next_start, pad */
constexpr int TAB_WORDS = 2 + 8 * MAX_LEVELS;

/*
 * The per-step arrays of the raw inputs (the index of each step's F block,
 * covariance type codes, observation counts, first-step flags, and zeros for
 * an absent c) and the level tables of one problem of k steps: level l has
 * L_l steps, L_0 = k, L_{l+1} = floor(L_l / 2) while L_l >= 2 (the plan the
 * batch path builds on the host). m_i outside 0..m is clamped into it, so no
 * step reads past its m-row observation slot.
 */
__global__ void __launch_bounds__(256)
    k_noalloc_init(int64_t k, int32_t m, int32_t F_shared, const int32_t* __restrict__ m_i, int8_t kt, int8_t ct,
                   int64_t* __restrict__ Fidx, int8_t* __restrict__ Kt, int8_t* __restrict__ Ct,
                   int32_t* __restrict__ mo, int8_t* __restrict__ first, double* __restrict__ czero, int n,
                   int64_t* __restrict__ tab, int nl) {
  for (int64_t s = blockIdx.x * 256LL + threadIdx.x; s < k; s += (int64_t)gridDim.x * 256) {
    Fidx[s] = F_shared ? 0 : s;
    Kt[s] = kt;
    Ct[s] = ct;
    const int v = m_i ? m_i[s] : m;
    mo[s] = v < 0 ? 0 : (v > m ? m : v);
    first[s] = s == 0 ? 1 : 0;
    if (czero)
      for (int i = 0; i < n; ++i) czero[s * n + i] = 0.0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    tab[0] = 0;
    tab[1] = 0;
    int64_t L = k;
    for (int l = 0; l < nl; ++l) {
      int64_t* t = tab + 2 + 8 * l;
      t[0] = L;
      t[1] = 0;
      t[2] = L == 1 ? 1 : L / 2;
      t[3] = 0;
      t[4] = (L + 1) / 2;
      t[5] = 0;
      t[6] = 0;
      t[7] = 0;
      L /= 2;
    }
  }
}

int cov_code(char t) {
  switch (t) {
    case 'W': return UK_COV_W;
    case 'w': return UK_COV_w;
    case 'C': return UK_COV_C;
    case 'U':
    case 'F': return UK_COV_L;
    default: return -1;
  }
}

std::mutex g_mu;
void* g_scratch[MAX_DEVICES];

/* scratch's device address on the current device, looked up once per device */
void* scratch_ptr() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
    uk_set_error("ukhip_smooth_oddeven_noalloc: no current HIP device");
    return nullptr;
  }
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      uk_set_error("ukhip_smooth_oddeven_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
      return nullptr;
    }
    g_scratch[dev] = p;
  }
  return g_scratch[dev];
}

}  // namespace

extern "C" int ukhip_smooth_oddeven_noalloc(const ukhip_problem_t* p, uint32_t options, double* states,
                                            double* covariances, void* stream) {
  if (!p || !states) {
    uk_set_error("ukhip_smooth_oddeven_noalloc: NULL problem or states");
    return -1;
  }
  const int n = p->n, m = p->m;
  const int64_t k = p->k;
  if (n < 1 || n > UKHIP_MAX_N || k < 1 || m < 1 || m > n) {
    uk_set_error("ukhip_smooth_oddeven_noalloc: need 1 <= n <= %d, k >= 1 and 1 <= m <= n (n = %d, k = %lld, m = %d)",
                 UKHIP_MAX_N, n, (long long)k, m);
    return -1;
  }
  const int kt = cov_code(p->K_type), ct = cov_code(p->C_type);
  if (kt < 0 || ct < 0 || !p->F || !p->K || !p->G || !p->o || !p->C) {
    uk_set_error("ukhip_smooth_oddeven_noalloc: unknown covariance type or missing input array");
    return -1;
  }
  const bool cov = (options & KALMAN_NO_COVARIANCE) == 0 && covariances != nullptr;

  /* the level plan (integer bookkeeping only, on the stack) */
  int nl = 0;
  int64_t Ls[MAX_LEVELS], evens[MAX_LEVELS], units[MAX_LEVELS], fac_base[MAX_LEVELS];
  int64_t E = 0;
  for (int64_t L = k; nl < MAX_LEVELS; L /= 2) {
    Ls[nl] = L;
    units[nl] = L == 1 ? 1 : L / 2;
    evens[nl] = (L + 1) / 2;
    fac_base[nl] = E;
    E += evens[nl];
    ++nl;
    if (L < 2) break;
  }
  const int64_t L1 = nl > 1 ? Ls[1] : 1, L2 = nl > 2 ? Ls[2] : 1;

  /* the scratch layout, 256-byte aligned pieces */
  const size_t nn = (size_t)n * n, S = (size_t)k;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t r = o;
    o += (bytes + 255) & ~(size_t)255;
    return r;
  };
  const size_t oTab = take(TAB_WORDS * 8), oFidx = take(S * 8), oKt = take(S), oCt = take(S), oM = take(S * 4),
               oFirst = take(S), oC0 = p->c ? 0 : take(S * n * 8);
  size_t oCh[2][5];
  for (int cb = 0; cb < 2; ++cb) {
    const size_t cs = cb == 0 ? S : (size_t)L1;
    oCh[cb][0] = take(cs * nn * 8);
    oCh[cb][1] = take(cs * nn * 8);
    oCh[cb][2] = take(cs * n * 8);
    oCh[cb][3] = take(cs * nn * 8);
    oCh[cb][4] = take(cs * n * 8);
  }
  const size_t oR = take((size_t)E * nn * 8), oB = take((size_t)E * nn * 8), oY = take((size_t)E * nn * 8),
               oCf = take((size_t)E * n * 8);
  size_t oX[2], oXT[2];
  for (int cb = 0; cb < 2; ++cb) {
    const size_t cs = (size_t)(cb == 0 ? L2 : L1);
    oX[cb] = take(cs * nn * 8);
    oXT[cb] = take(cs * nn * 8);
  }
  const size_t oFix = take((nn + n) * 8);
  if (o > (size_t)UKHIP_NOALLOC_SCRATCH_BYTES) {
    uk_set_error("ukhip_smooth_oddeven_noalloc: n = %d, k = %lld needs %zu bytes of device scratch; the "
                 "allocation-free entry point has %zu (use ukhip_batch_create for larger problems)",
                 n, (long long)k, o, (size_t)UKHIP_NOALLOC_SCRATCH_BYTES);
    return -1;
  }
  char* base = (char*)scratch_ptr();
  if (!base) return -1;
  hipStream_t st = stream ? (hipStream_t)stream : (hipStream_t)ukhip_stream();
  if (!st) return -1;

  int64_t* tab = (int64_t*)(base + oTab);
  double* czero = p->c ? nullptr : (double*)(base + oC0);
  const unsigned blocks = (unsigned)std::min<int64_t>((k + 255) / 256, 4096);
  k_noalloc_init<<<dim3(blocks), dim3(256), 0, st>>>(k, m, p->F_shared, p->m_i, (int8_t)kt, (int8_t)ct,
                                                       (int64_t*)(base + oFidx), (int8_t*)(base + oKt),
                                                       (int8_t*)(base + oCt), (int32_t*)(base + oM),
                                                       (int8_t*)(base + oFirst), czero, n, tab, nl);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    uk_set_error("ukhip_smooth_oddeven_noalloc: k_noalloc_init launch: %s", hipGetErrorString(e));
    return -1;
  }

  uk_oe_view_t v{};
  v.raw.n = n;
  v.raw.mr = m;
  v.raw.steps = k;
  v.raw.H_ident = p->H ? 0 : 1;
  v.raw.K_full = p->K_type != 'w';
  v.raw.C_full = p->C_type != 'w';
  v.raw.F = p->F;
  v.raw.Fidx = (const int64_t*)(base + oFidx);
  v.raw.H = p->H;
  v.raw.c = p->c ? p->c : czero;
  v.raw.K = p->K;
  v.raw.G = p->G;
  v.raw.o = p->o;
  v.raw.C = p->C;
  v.raw.Kt = (const int8_t*)(base + oKt);
  v.raw.Ct = (const int8_t*)(base + oCt);
  v.raw.m = (const int32_t*)(base + oM);
  v.raw.first = (const int8_t*)(base + oFirst);
  for (int cb = 0; cb < 2; ++cb) {
    v.chain[cb].B = (double*)(base + oCh[cb][0]);
    v.chain[cb].D = (double*)(base + oCh[cb][1]);
    v.chain[cb].c = (double*)(base + oCh[cb][2]);
    v.chain[cb].C = (double*)(base + oCh[cb][3]);
    v.chain[cb].o = (double*)(base + oCh[cb][4]);
    v.X[cb] = (double*)(base + oX[cb]);
    v.XT[cb] = (double*)(base + oXT[cb]);
  }
  v.fac.R = (double*)(base + oR);
  v.fac.Bt = (double*)(base + oB);
  v.fac.Y = (double*)(base + oY);
  v.fac.cf = (double*)(base + oCf);
  v.xout = states;
  v.Sout = cov ? covariances : nullptr;
  v.fix = (double*)(base + oFix);
  v.wobs = nullptr;
  v.n = n;
  v.mr0 = m;
  v.compress = 0;
  v.nl = nl;
  v.nF = p->F_shared ? 1 : k;
  uk_level_t lv[MAX_LEVELS];
  const int32_t* prob32 = (const int32_t*)(tab + 1);
  for (int l = 0; l < nl; ++l) {
    const int64_t* t = tab + 2 + 8 * l;
    lv[l].nprob = 1;
    lv[l].level = l;
    lv[l].units = units[l];
    lv[l].evens = evens[l];
    lv[l].prob = prob32;
    lv[l].L = t;
    lv[l].unit_start = t + 1;
    lv[l].even_start = t + 3;
    lv[l].step_start = t + 5;
    lv[l].next_start = t + 6;
    lv[l].orig_start = tab;
  }
  v.levels = lv;
  v.fac_base = fac_base;
  return uk_run_oddeven_view(&v, cov ? 1 : 0, st);
}
