/*
 * ukhip_host.cpp -- host side of libukhip: packing step equations into device
 * structure-of-arrays buffers, planning the odd-even levels of a batch, and
 * launching the HIP kernels on the library stream.
 *
 * Takes the place of the reference's parallel backends under the smoothers
 * (c/parallel.h:31-46, c/parallel_tbb.cpp, c/parallel_sequential.c): the
 * per-pair loops of each level run as one kernel launch over all pairs of
 * all problems of a batch.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/kalman.h"
#include "../../include/ukhip.h"
#include "ukhip_internal.h"

extern "C" {
int uk_bucket(int n);
int uk_bucket_qr(int n);
hipError_t uk_launch_weigh(const uk_raw_t* raw, const uk_chain_t* out, hipStream_t st);
hipError_t uk_launch_weigh_obs(const uk_raw_t* raw, const uk_chain_t* out, double* wobs, hipStream_t st);
hipError_t uk_launch_eliminate(const uk_level_t* lv, const uk_chain_t* in, const uk_chain_t* out,
                               const uk_factors_t* fac, double* fix, int n, hipStream_t st, hipEvent_t* ev);
hipError_t uk_launch_filter_step(const uk_filter_t* st, const uk_filter_in_t* in, int nf, int n, int first,
                                 hipStream_t s);
hipError_t uk_launch_filter_smooth(const uk_filter_t* st, int n, int64_t steps, double* xs, double* Ws,
                                   hipStream_t s);
hipError_t uk_launch_cov_format(const double* S, int64_t steps, int n, int fmt, double* dst, hipStream_t st);
hipError_t uk_launch_backsub(const uk_level_t* lv, const uk_factors_t* fac, double* xout, double* Sout,
                             const double* Xup, const double* XTup, double* Xcur, double* XTcur, int n, int cov,
                             hipStream_t st);
int uk_assoc_smooth(struct ukhip_batch* b, uint32_t options);
int uk_assoc_smooth_timed(struct ukhip_batch* b, uint32_t options, double* ms, int32_t* launches);
void uk_set_error(const char* fmt, ...);
hipError_t uk_launch_eliminate32(const uk_level_t* lv, const uk_chain32_t* in, const uk_chain32_t* out,
                                 const uk_factors32_t* fac, float* fix, int n, hipStream_t st, hipEvent_t* ev);
hipError_t uk_launch_to32(float* dst, const double* src, int64_t count, hipStream_t st);
hipError_t uk_launch_weigh_f32(const uk_raw_t* raw, const uk_chain_t* out, const uk_chain32_t* out32, int w64,
                               hipStream_t st);
hipError_t uk_launch_add(double* x, const double* d, int64_t count, hipStream_t st);
hipError_t uk_launch_normal(const uk_chain_t* ch, const int8_t* first, const int32_t* m_obs, const double* x,
                            double* gA, double* gB, double* g, int64_t steps, int n, hipStream_t st);
hipError_t uk_launch_fwd(const uk_level_t* lv, const uk_factors32_t* fac, const double* g, double* y, double* gnext,
                         int n, hipStream_t st);
hipError_t uk_launch_solve_est(const uk_level_t* lv, const uk_factors32_t* fac, const float* rhs32,
                               const double* rhs64, double* x, int n, hipStream_t st);
hipError_t uk_launch_to64(double* dst, const float* src, int64_t count, hipStream_t st);
hipError_t uk_launch_fwd64(const uk_level_t* lv, const uk_factors_t* fac, const double* g, double* y, double* gnext,
                           int n, hipStream_t st);
hipError_t uk_launch_solve_est64(const uk_level_t* lv, const uk_factors_t* fac, const double* rhs64, double* x, int n,
                                 hipStream_t st);
}

/* ------------------------------------------------------------------------ */
/* errors and the library stream                                            */
/* ------------------------------------------------------------------------ */

static thread_local std::string g_err;

extern "C" void uk_set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

extern "C" const char* ukhip_last_error(void) { return g_err.c_str(); }
extern "C" void ukhip_clear_error(void) { g_err.clear(); }

static hipStream_t g_stream = nullptr;
static std::once_flag g_stream_once;
static hipError_t g_stream_err = hipSuccess;

static bool hip_ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  uk_set_error("%s: %s", what, hipGetErrorString(e));
  return false;
}

static bool ensure_stream() {
  std::call_once(g_stream_once, [] {
    int count = 0;
    g_stream_err = hipGetDeviceCount(&count);
    if (g_stream_err == hipSuccess && count == 0) g_stream_err = hipErrorNoDevice;
    if (g_stream_err == hipSuccess) g_stream_err = hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking);
  });
  return hip_ok(g_stream_err, "libukhip needs a HIP device (gfx950)");
}

extern "C" void* ukhip_stream(void) {
  ensure_stream();
  return (void*)g_stream;
}

extern "C" int ukhip_synchronize(void) {
  if (!ensure_stream()) return -1;
  return hip_ok(hipStreamSynchronize(g_stream), "hipStreamSynchronize") ? 0 : -1;
}

/* ------------------------------------------------------------------------ */
/* batch                                                                      */
/* ------------------------------------------------------------------------ */

struct LevelHost {
  int32_t nprob = 0;
  int64_t units = 0, evens = 0, steps = 0, next_steps = 0;
  std::vector<int32_t> prob;
  std::vector<int64_t> L, unit_start, even_start, step_start, next_start;
  int64_t fac_base = 0;
  uk_level_t dev{};
};

/* A device memory arena: one allocation carved first-fit into 256-byte
 * aligned blocks. Batches created for the server take their memory from it,
 * so admitting and retiring batches never calls hipMalloc/hipFree (which
 * would synchronize with the work of other batches in flight). */
struct DevArena {
  char* base = nullptr;
  size_t size = 0;
  std::map<size_t, size_t> free_blocks; /* offset -> size */
  std::map<size_t, size_t> used;        /* offset -> size */
  std::mutex mu;

  void* alloc(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    std::lock_guard<std::mutex> g(mu);
    for (auto it = free_blocks.begin(); it != free_blocks.end(); ++it) {
      if (it->second >= bytes) {
        const size_t off = it->first, sz = it->second;
        free_blocks.erase(it);
        if (sz > bytes) free_blocks[off + bytes] = sz - bytes;
        used[off] = bytes;
        return base + off;
      }
    }
    return nullptr;
  }
  void release(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(mu);
    const size_t off = (size_t)((char*)p - base);
    auto u = used.find(off);
    if (u == used.end()) return;
    size_t start = off, sz = u->second;
    used.erase(u);
    auto nx = free_blocks.lower_bound(start);
    if (nx != free_blocks.end() && nx->first == start + sz) {
      sz += nx->second;
      nx = free_blocks.erase(nx);
    }
    if (nx != free_blocks.begin()) {
      auto pv = std::prev(nx);
      if (pv->first + pv->second == start) {
        start = pv->first;
        sz += pv->second;
        free_blocks.erase(pv);
      }
    }
    free_blocks[start] = sz;
  }
  size_t in_use() {
    std::lock_guard<std::mutex> g(mu);
    size_t t = 0;
    for (auto& kv : used) t += kv.second;
    return t;
  }
};

struct ukhip_batch {
  int64_t nF = 0; /* F blocks in raw.F (1: every step shares one) */
  int n = 0, mr = 0, nprob = 0;
  /* Level-0 observation slots: compress = observation blocks taller than n
   * are reduced to n rows after whitening (staged in wobs, mr * n + mr doubles
   * per step); the level-0 chain's C then has ld mr0 = n and step s holds
   * m0[s] = min(m_s, n) rows. Otherwise mr0 = mr and m0 = raw.m. */
  bool compress = false;
  int mr0 = 0;
  const int32_t* m0 = nullptr;
  double* wobs = nullptr;
  int64_t steps = 0;
  std::vector<int64_t> k, orig_start;
  std::vector<LevelHost> levels;
  /* device memory */
  void* d_raw = nullptr;
  void* d_work = nullptr;
  /* the FP64 elimination's own storage (chain[1], fac, X, XT, Sout), allocated
   * by ensure_elim on the first FP64 smoothing, associative smoothing or
   * covariance request: a batch smoothed only by the mixed-precision path
   * for its estimates never holds it */
  void* d_elim = nullptr;
  size_t elim_bytes = 0;
  void* d_tab = nullptr;
  size_t raw_bytes = 0, work_bytes = 0;
  uk_raw_t raw{};
  uk_chain_t chain[2]{};
  uk_factors_t fac{};
  double* X[2]{};
  double* XT[2]{};
  double* xout = nullptr;
  double* Sout = nullptr;
  double* fix = nullptr; /* per problem: Z (n x n) and o (n) of a lone last step */
  int64_t total_evens = 0;
  /* associative-path workspace, allocated on first use */
  void* d_assoc = nullptr;
  size_t assoc_bytes = 0;
  /* A batch of problems with different state dimensions is a composite: one
   * part per dimension (each a batch of its own); the composite itself then
   * holds no device arrays. where[p] = (part, index in part). */
  std::vector<ukhip_batch*> parts;
  std::vector<std::pair<int32_t, int32_t>> where;
  std::vector<hipStream_t> part_streams;
  /* captured odd-even smoothing, [0] without and [1] with covariances */
  hipGraphExec_t graph[3]{}; /* by covariance mode: none, full, diagonals at level 0 */
  /* mixed-precision workspace, allocated on first use: FP32 chains and
   * factors, FP64 right-hand sides and corrections */
  void* d_mixed = nullptr;
  uk_chain32_t chain32[2]{};
  uk_factors32_t fac32{};
  float* fix32 = nullptr;
  double *gA = nullptr, *gB = nullptr, *gbuf[2]{}, *ys = nullptr, *dx = nullptr;
  /* The FP64 storage of the levels the mixed path eliminates in FP64, per
   * first FP64 level lsw (allocated on first use, kept until the batch is
   * freed, since captured graphs refer to it): the two chains (levels of
   * lsw's parity in ch[lsw & 1]) and the factor slots of the levels from lsw
   * up (level l's at levels[l].fac_base - levels[lsw].fac_base). chain[0]
   * keeps the FP64 level-0 chain for the refinement residuals. */
  struct TopStore {
    void* mem = nullptr;
    uk_chain_t ch[2]{};
    uk_factors_t fac{};
  };
  std::map<int, TopStore> tops;
  /* captured mixed-precision smoothings, by (refinements, FP32 levels, covariances) */
  std::map<int64_t, hipGraphExec_t> mixed_graph;
  /* memory source (NULL: hipMalloc) and the host copy of the level tables
   * (kept until an asynchronous upload has completed) */
  DevArena* arena = nullptr;
  std::vector<int64_t> tab_host;
};

static void* dev_alloc(ukhip_batch* b, size_t bytes, const char* what) {
  if (b->arena) {
    void* p = b->arena->alloc(bytes);
    if (!p) uk_set_error("%s: %zu bytes do not fit in the device arena", what, bytes);
    return p;
  }
  void* p = nullptr;
  if (!hip_ok(hipMalloc(&p, bytes), what)) return nullptr;
  return p;
}
static void dev_free(ukhip_batch* b, void* p) {
  if (!p) return;
  if (b->arena) b->arena->release(p);
  else (void)hipFree(p);
}

static int cov_code(char t) {
  switch (t) {
    case 'W': return UK_COV_W;
    case 'w': return UK_COV_w;
    case 'C': return UK_COV_C;
    case 'U':
    case 'F': return UK_COV_L;
    default: return -1;
  }
}

/* Host staging of the raw inputs in the device layout of uk_raw_t. */
/* Pinned host buffers reused by size class (powers of two from 1 MiB), so
 * that a steady stream of batches does not allocate and free pinned memory
 * (hipHostFree can wait for the device). */
struct PinnedPool {
  bool pageable = false; /* plain host memory (buffers that are never DMA'd) */
  std::mutex mu;
  std::map<size_t, std::vector<void*>> free_;
  static size_t cls(size_t bytes) {
    size_t c = (size_t)1 << 20;
    while (c < bytes) c <<= 1;
    return c;
  }
  void* get(size_t bytes) {
    const size_t c = cls(bytes);
    {
      std::lock_guard<std::mutex> g(mu);
      auto& v = free_[c];
      if (!v.empty()) {
        void* p = v.back();
        v.pop_back();
        return p;
      }
    }
    void* p = nullptr;
    if (pageable) {
      p = std::malloc(c);
      if (!p) uk_set_error("out of host memory (%zu bytes)", c);
      return p;
    }
    if (!hip_ok(hipHostMalloc(&p, c, hipHostMallocDefault), "hipHostMalloc(pinned pool)")) return nullptr;
    return p;
  }
  void put(void* p, size_t bytes) {
    if (!p) return;
    std::lock_guard<std::mutex> g(mu);
    free_[cls(bytes)].push_back(p);
  }
  void clear() {
    std::lock_guard<std::mutex> g(mu);
    for (auto& kv : free_)
      for (void* p : kv.second) {
        if (pageable) std::free(p);
        else (void)hipHostFree(p);
      }
    free_.clear();
  }
};

struct Staging {
  PinnedPool* pool = nullptr; /* host buffer from this pool, or hipHostMalloc */
  int n = 0, mr = 0;
  int64_t steps = 0;
  int H_ident = 1, K_full = 0, C_full = 0;
  /* compress = 1: some step has more observation rows than states (mr > n);
   * the whitened observation blocks are then reduced to n rows on the device
   * (k_obs_compress) and meff holds each step's row count after that,
   * min(m_i, n) */
  int compress = 0;
  int64_t nF = 0; /* F blocks (0: one per step) */
  size_t offFidx = 0, offF = 0, offH = 0, offc = 0, offK = 0, offG = 0, offo = 0, offC = 0, offKt = 0, offCt = 0, offm = 0,
         offfirst = 0, offme = 0, bytes = 0;
  char* host = nullptr;

  void layout() {
    const size_t nn = (size_t)n * n, S = (size_t)steps;
    size_t o = 0;
    auto take = [&](size_t b) {
      size_t r = o;
      o += (b + 255) & ~(size_t)255;
      return r;
    };
    if (nF <= 0) nF = steps;
    offFidx = take(S * 8);
    offF = take((size_t)nF * nn * 8);
    offH = H_ident ? 0 : take(S * nn * 8);
    offc = take(S * n * 8);
    offK = take(S * (K_full ? nn : (size_t)n) * 8);
    offG = take(S * (size_t)mr * n * 8);
    offo = take(S * (size_t)mr * 8);
    offC = take(S * (C_full ? (size_t)mr * mr : (size_t)mr) * 8);
    offKt = take(S);
    offCt = take(S);
    offm = take(S * 4);
    offfirst = take(S);
    offme = compress ? take(S * 4) : 0;
    bytes = o;
  }
  double* F() { return (double*)(host + offF); }
  int64_t* Fidx() { return (int64_t*)(host + offFidx); }
  double* H() { return (double*)(host + offH); }
  double* c() { return (double*)(host + offc); }
  double* K() { return (double*)(host + offK); }
  double* G() { return (double*)(host + offG); }
  double* o() { return (double*)(host + offo); }
  double* C() { return (double*)(host + offC); }
  int8_t* Kt() { return (int8_t*)(host + offKt); }
  int8_t* Ct() { return (int8_t*)(host + offCt); }
  int32_t* m() { return (int32_t*)(host + offm); }
  int8_t* first() { return (int8_t*)(host + offfirst); }
  int32_t* meff() { return (int32_t*)(host + offme); }
  /* sets compress from mr and n, before layout() */
  void decide_compress() { compress = mr > n ? 1 : 0; }
  /* meff from m (compress only) */
  void fill_meff() {
    if (!compress) return;
    for (int64_t s = 0; s < steps; ++s) meff()[s] = std::min(m()[s], n);
  }
};

/* Largest observation row count (after whitening) a step may have for state
 * dimension n: k_obs_compress holds the rows in two blocks of the n bucket,
 * k_weigh in one block of at most 64 rows. */
static int max_obs_rows(int n) { return std::min(2 * uk_bucket(n), UKHIP_MAX_N); }

static void plan_levels(ukhip_batch* b) {
  b->levels.clear();
  std::vector<int64_t> Lcur(b->k.begin(), b->k.end());
  std::vector<int32_t> active(b->nprob);
  for (int i = 0; i < b->nprob; ++i) active[i] = i;
  int level = 0;
  int64_t fac_base = 0;
  while (!active.empty()) {
    LevelHost lv;
    lv.nprob = (int32_t)active.size();
    int64_t us = 0, es = 0, ss = 0, ns = 0;
    std::vector<int32_t> next_active;
    std::vector<int64_t> next_L;
    for (int32_t p : active) {
      const int64_t L = Lcur[p];
      lv.prob.push_back(p);
      lv.L.push_back(L);
      lv.unit_start.push_back(us);
      lv.even_start.push_back(es);
      lv.step_start.push_back(ss);
      lv.next_start.push_back(ns);
      us += (L == 1) ? 1 : L / 2;
      es += (L + 1) / 2;
      ss += L;
      if (L >= 2) {
        ns += L / 2;
        next_active.push_back(p);
      }
    }
    lv.unit_start.push_back(us);
    lv.even_start.push_back(es);
    lv.units = us;
    lv.evens = es;
    lv.steps = ss;
    lv.next_steps = ns;
    lv.fac_base = fac_base;
    fac_base += es;
    b->levels.push_back(std::move(lv));
    for (int32_t p : next_active) Lcur[p] = Lcur[p] / 2;
    active.swap(next_active);
    ++level;
  }
  b->total_evens = fac_base;
}

static void batch_release(ukhip_batch* b) {
  if (!b) return;
  for (auto* q : b->parts) batch_release(q);
  for (auto st : b->part_streams) (void)hipStreamDestroy(st);
  for (auto g : b->graph)
    if (g) (void)hipGraphExecDestroy(g);
  for (auto& kv : b->mixed_graph)
    if (kv.second) (void)hipGraphExecDestroy(kv.second);
  dev_free(b, b->d_raw);
  dev_free(b, b->d_work);
  dev_free(b, b->d_elim);
  dev_free(b, b->d_tab);
  if (b->d_assoc) (void)hipFree(b->d_assoc);
  dev_free(b, b->d_mixed);
  for (auto& kv : b->tops) dev_free(b, kv.second.mem);
  delete b;
}

/* The FP64 elimination's storage beyond the level-0 chain: chain[1] (level-1
 * size), the factor slots of every level, the cross covariances of two
 * levels (level 0's are not needed: no level below reads them, so X[0]
 * holds level 2's at most) and the covariance output. Idempotent; call it
 * before a stream capture, never inside one. */
static bool ensure_elim(ukhip_batch* b) {
  if (b->d_elim) return true;
  const size_t n = b->n, nn = n * n, S = (size_t)b->steps;
  const int64_t L1 = b->levels.size() > 1 ? b->levels[1].steps : 1;
  const int64_t L2 = b->levels.size() > 2 ? b->levels[2].steps : 1;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    size_t rr = o;
    o += (bytes + 255) & ~(size_t)255;
    return rr;
  };
  const size_t c1 = (size_t)std::max<int64_t>(L1, 1);
  const size_t cB = take(c1 * nn * 8), cD = take(c1 * nn * 8), cc = take(c1 * n * 8), cC = take(c1 * nn * 8),
               co = take(c1 * n * 8);
  const size_t E = (size_t)b->total_evens;
  const size_t fR = take(E * nn * 8), fB = take(E * nn * 8), fY = take(E * nn * 8), fc = take(E * n * 8);
  size_t xo[2], xto[2];
  for (int cb = 0; cb < 2; ++cb) {
    const size_t cs = (size_t)std::max<int64_t>(cb == 0 ? L2 : L1, 1);
    xo[cb] = take(cs * nn * 8);
    xto[cb] = take(cs * nn * 8);
  }
  const size_t Soo = take(S * nn * 8);
  void* mem = dev_alloc(b, o, "device memory for the FP64 elimination");
  if (!mem) return false;
  char* w = (char*)mem;
  b->chain[1].B = (double*)(w + cB);
  b->chain[1].D = (double*)(w + cD);
  b->chain[1].c = (double*)(w + cc);
  b->chain[1].C = (double*)(w + cC);
  b->chain[1].o = (double*)(w + co);
  b->fac.R = (double*)(w + fR);
  b->fac.Bt = (double*)(w + fB);
  b->fac.Y = (double*)(w + fY);
  b->fac.cf = (double*)(w + fc);
  for (int cb = 0; cb < 2; ++cb) {
    b->X[cb] = (double*)(w + xo[cb]);
    b->XT[cb] = (double*)(w + xto[cb]);
  }
  b->Sout = (double*)(w + Soo);
  b->d_elim = mem;
  b->elim_bytes = o;
  return true;
}
/* ensure_elim for a batch or every part of a composite */
static bool ensure_elim_all(ukhip_batch* b) {
  if (b->parts.empty()) return ensure_elim(b);
  for (auto* q : b->parts)
    if (!ensure_elim(q)) return false;
  return true;
}

/* Allocates device memory (from `arena`, or hipMalloc), uploads staging and
 * builds the level tables, on stream `us`; with `sync` false the upload may
 * still be in flight on return (st.host and b->tab_host must outlive it).
 * With st.host NULL only the layout is used: the caller uploads the raw
 * fields itself (at st's offsets in b->d_raw). */
static ukhip_batch* batch_from_staging(Staging& st, const std::vector<int64_t>& lengths, hipStream_t us = nullptr,
                                       DevArena* arena = nullptr, bool sync = true) {
  if (!us) us = g_stream;
  ukhip_batch* b = new ukhip_batch();
  b->arena = arena;
  b->n = st.n;
  b->mr = st.mr;
  b->nprob = (int)lengths.size();
  b->steps = st.steps;
  b->k = lengths;
  b->orig_start.resize(b->nprob);
  int64_t acc = 0;
  for (int p = 0; p < b->nprob; ++p) {
    b->orig_start[p] = acc;
    acc += lengths[p];
  }
  plan_levels(b);

  const int n = st.n;
  const size_t nn = (size_t)n * n, S = (size_t)st.steps;
  /* raw inputs */
  b->d_raw = dev_alloc(b, st.bytes, "device memory for the raw inputs");
  if (!b->d_raw) { batch_release(b); return nullptr; }
  b->raw_bytes = st.bytes;
  if (st.host && !hip_ok(hipMemcpyAsync(b->d_raw, st.host, st.bytes, hipMemcpyHostToDevice, us), "upload")) {
    batch_release(b);
    return nullptr;
  }
  char* r = (char*)b->d_raw;
  b->raw.n = n;
  b->raw.mr = st.mr;
  b->raw.steps = st.steps;
  b->raw.H_ident = st.H_ident;
  b->raw.K_full = st.K_full;
  b->raw.C_full = st.C_full;
  b->raw.F = (const double*)(r + st.offF);
  b->raw.Fidx = (const int64_t*)(r + st.offFidx);
  b->nF = st.nF;
  b->raw.H = st.H_ident ? nullptr : (const double*)(r + st.offH);
  b->raw.c = (const double*)(r + st.offc);
  b->raw.K = (const double*)(r + st.offK);
  b->raw.G = (const double*)(r + st.offG);
  b->raw.o = (const double*)(r + st.offo);
  b->raw.C = (const double*)(r + st.offC);
  b->raw.Kt = (const int8_t*)(r + st.offKt);
  b->raw.Ct = (const int8_t*)(r + st.offCt);
  b->raw.m = (const int32_t*)(r + st.offm);
  b->raw.first = (const int8_t*)(r + st.offfirst);
  b->compress = st.compress != 0;
  b->mr0 = b->compress ? n : st.mr;
  b->m0 = b->compress ? (const int32_t*)(r + st.offme) : b->raw.m;

  /* workspace: the level-0 chain, the estimates and the per-problem fix-last
   * scratch; the rest of the FP64 elimination's storage comes with
   * ensure_elim */
  size_t o = 0;
  auto take = [&](size_t bytes) {
    size_t rr = o;
    o += (bytes + 255) & ~(size_t)255;
    return rr;
  };
  size_t chain_off[5];
  chain_off[0] = take(S * nn * 8);
  chain_off[1] = take(S * nn * 8);
  chain_off[2] = take(S * n * 8);
  chain_off[3] = take(S * nn * 8);
  chain_off[4] = take(S * n * 8);
  const size_t xoo = take(S * n * 8);
  const size_t fxo = take((size_t)b->nprob * (nn + n) * 8);
  const size_t wbo = b->compress ? take(S * (size_t)st.mr * (n + 1) * 8) : 0;
  b->work_bytes = o;
  b->d_work = dev_alloc(b, o, "device memory for the workspace");
  if (!b->d_work) { batch_release(b); return nullptr; }
  char* w = (char*)b->d_work;
  b->chain[0].B = (double*)(w + chain_off[0]);
  b->chain[0].D = (double*)(w + chain_off[1]);
  b->chain[0].c = (double*)(w + chain_off[2]);
  b->chain[0].C = (double*)(w + chain_off[3]);
  b->chain[0].o = (double*)(w + chain_off[4]);
  b->xout = (double*)(w + xoo);
  b->fix = (double*)(w + fxo);
  b->wobs = b->compress ? (double*)(w + wbo) : nullptr;
  /* server batches (arena-backed) take all of their storage up front */
  if (arena && !ensure_elim(b)) { batch_release(b); return nullptr; }

  /* level tables: one int64 array; prob indices stored widened */
  std::vector<int64_t> tab;
  struct Off { size_t prob, L, U, E, S, N; };
  std::vector<Off> offs;
  for (auto& lv : b->levels) {
    Off f;
    auto put = [&](const int64_t* v, size_t cnt) {
      size_t at = tab.size();
      tab.insert(tab.end(), v, v + cnt);
      return at;
    };
    std::vector<int64_t> pw(lv.prob.begin(), lv.prob.end());
    f.prob = put(pw.data(), pw.size());
    f.L = put(lv.L.data(), lv.nprob);
    f.U = put(lv.unit_start.data(), lv.nprob + 1);
    f.E = put(lv.even_start.data(), lv.nprob + 1);
    f.S = put(lv.step_start.data(), lv.nprob);
    f.N = put(lv.next_start.data(), lv.nprob);
    offs.push_back(f);
  }
  const size_t oOrig = tab.size();
  tab.insert(tab.end(), b->orig_start.begin(), b->orig_start.end());
  /* prob[] is read as int32 by the kernels: store a packed int32 copy after the int64 data */
  const size_t oProb32 = tab.size();
  size_t nprob_all = 0;
  for (auto& lv : b->levels) nprob_all += lv.nprob;
  tab.resize(tab.size() + (nprob_all + 1) / 2 + 1, 0);
  {
    int32_t* p32 = reinterpret_cast<int32_t*>(tab.data() + oProb32);
    size_t at = 0;
    for (auto& lv : b->levels)
      for (int32_t v : lv.prob) p32[at++] = v;
  }
  b->d_tab = dev_alloc(b, tab.size() * 8, "device memory for the level tables");
  if (!b->d_tab) { batch_release(b); return nullptr; }
  b->tab_host.swap(tab);
  if (!hip_ok(hipMemcpyAsync(b->d_tab, b->tab_host.data(), b->tab_host.size() * 8, hipMemcpyHostToDevice, us),
              "upload tables") ||
      (sync && !hip_ok(hipStreamSynchronize(us), "upload"))) {
    batch_release(b);
    return nullptr;
  }
  const int64_t* d = (const int64_t*)b->d_tab;
  const int32_t* d32 = reinterpret_cast<const int32_t*>(d + oProb32);
  size_t at32 = 0;
  for (size_t l = 0; l < b->levels.size(); ++l) {
    auto& lv = b->levels[l];
    const Off& f = offs[l];
    lv.dev.nprob = lv.nprob;
    lv.dev.level = (int32_t)l;
    lv.dev.units = lv.units;
    lv.dev.evens = lv.evens;
    lv.dev.prob = d32 + at32;
    at32 += lv.nprob;
    lv.dev.L = d + f.L;
    lv.dev.unit_start = d + f.U;
    lv.dev.even_start = d + f.E;
    lv.dev.step_start = d + f.S;
    lv.dev.next_start = d + f.N;
    lv.dev.orig_start = d + oOrig;
  }
  return b;
}

/* ------------------------------------------------------------------------ */
/* packing                                                                    */
/* ------------------------------------------------------------------------ */

static bool staging_alloc(Staging& st) {
  st.layout();
  if (st.pool) {
    st.host = (char*)st.pool->get(st.bytes);
    return st.host != nullptr;
  }
  return hip_ok(hipHostMalloc((void**)&st.host, st.bytes, hipHostMallocDefault), "hipHostMalloc(staging)");
}
static void staging_free(Staging& st) {
  if (st.host) {
    if (st.pool) st.pool->put(st.host, st.bytes);
    else (void)hipHostFree(st.host);
  }
  st.host = nullptr;
}

static void copy_cols(double* dst, int ld_dst, const kalman_matrix_t* A, int rows, int cols) {
  for (int j = 0; j < cols; ++j)
    for (int i = 0; i < rows; ++i) dst[(size_t)j * ld_dst + i] = A->elements[(size_t)j * A->ld + i];
}

static bool is_identity(const kalman_matrix_t* A) {
  for (int j = 0; j < A->col_dim; ++j)
    for (int i = 0; i < A->row_dim; ++i)
      if (A->elements[(size_t)j * A->ld + i] != (i == j ? 1.0 : 0.0)) return false;
  return true;
}

#define CHECK_SHAPE(A, r, c, what)                                                                            \
  do {                                                                                                          \
    if ((A) == NULL || (A)->row_dim != (r) || (A)->col_dim != (c)) {                                           \
      uk_set_error("step %lld: %s must be %d x %d", (long long)i, what, (int)(r), (int)(c));                     \
      return false;                                                                                             \
    }                                                                                                           \
  } while (0)

static bool cov_shape_ok(const kalman_matrix_t* A, char t, int dim) {
  if (!A) return false;
  if (t == 'w') return A->row_dim == dim && A->col_dim == 1;
  return A->row_dim == dim && A->col_dim == dim;
}

/*
 * Validates and packs kalman_step_equations_t (c/kalman.h:56-75) into the
 * device layout, with every step padded to N = max_i n_i state components.
 *
 * The reference lets step i carry its own state dimension n_i, an evolution
 * equation of l_i rows (H_i l_i x n_i, F_i l_i x n_{i-1}) and m_i observation
 * rows, and sizes every block of the elimination from them
 * (c/kalman_oddeven_smoother.c:422-440 ll = MIN(nn, rows(QR)), :480-494
 * ll = rows(H_i), nn = cols(H_i); matlab/KalmanExplicitRepresentation.m:95-101
 * pads H_i as [I 0]). Here each step has N state components, u'_i =
 * [u_i; z_i], and the same least-squares problem in that form:
 *   - evolution: [H_i 0] u'_i = [F_i 0] u'_{i-1} + c_i in the first l_i of N
 *     rows; the remaining rows are zero (zero rows leave every Householder
 *     factorization unchanged);
 *   - observation: the m_i rows [G_i 0] u'_i = o_i, plus N - n_i rows
 *     z_i = 0 of unit weight.
 * The padding components z_i appear in no other equation, so they are
 * determined as zero, the normal equations are block diagonal between u and
 * z, and u's estimates and covariance blocks equal the unpadded problem's.
 * K_i and C_i are padded as block-diag(K_i, I) in their own representation
 * ('w': the vector extended with ones). A step with more observation rows
 * than N (after padding) is reduced to N rows on the device (k_obs_compress).
 */
static bool pack_equations(kalman_step_equations_t** eq, int64_t k, Staging& st) {
  if (k < 1) { uk_set_error("length must be >= 1"); return false; }
  int N = 0;
  for (int64_t i = 0; i < k; ++i) {
    if (!eq[i]) { uk_set_error("step %lld: NULL step equations", (long long)i); return false; }
    N = std::max(N, (int)eq[i]->dimension);
  }
  if (N < 1 || N > UKHIP_MAX_N) { uk_set_error("state dimension %d outside 1..%d", N, UKHIP_MAX_N); return false; }
  int mr = 0;
  st.H_ident = 1;
  st.K_full = 0;
  st.C_full = 0;
  for (int64_t i = 0; i < k; ++i) {
    kalman_step_equations_t* e = eq[i];
    const int n = e->dimension;
    if (n < 1) { uk_set_error("step %lld: state dimension %d < 1", (long long)i, n); return false; }
    if (i > 0) {
      const int np = eq[i - 1]->dimension;
      if (!e->c || e->c->col_dim != 1) { uk_set_error("step %lld: c must be a column vector", (long long)i); return false; }
      const int l = e->c->row_dim;
      if (l > N) { uk_set_error("step %lld: %d evolution rows exceed the state dimension %d", (long long)i, l, N); return false; }
      CHECK_SHAPE(e->H, l, n, "H");
      CHECK_SHAPE(e->F, l, np, "F");
      if (cov_code(e->K_type) < 0 || !cov_shape_ok(e->K, e->K_type, l)) {
        uk_set_error("step %lld: K of type '%c' has the wrong shape or type", (long long)i, e->K_type);
        return false;
      }
      if (e->K_type != 'w') st.K_full = 1;
      if (st.H_ident && (l != N || n != N || !is_identity(e->H))) st.H_ident = 0;
    }
    int m = 0;
    if (e->o != NULL) {
      m = e->o->row_dim;
      CHECK_SHAPE(e->o, m, 1, "o");
      CHECK_SHAPE(e->G, m, n, "G");
      if (cov_code(e->C_type) < 0 || !cov_shape_ok(e->C, e->C_type, m)) {
        uk_set_error("step %lld: C of type '%c' has the wrong shape or type", (long long)i, e->C_type);
        return false;
      }
      if (e->C_type != 'w') st.C_full = 1;
    }
    const int mp = m + (N - n); /* with the padding components' unit rows */
    if (mp > max_obs_rows(N)) {
      uk_set_error("step %lld: %d observation rows (with %d padding rows for state dimension %d of %d) exceed "
                   "the %d this build reduces on the device",
                   (long long)i, mp, N - n, n, N, max_obs_rows(N));
      return false;
    }
    mr = std::max(mr, mp);
  }
  st.n = N;
  st.mr = std::max(mr, 1);
  st.steps = k;
  st.decide_compress();
  if (!staging_alloc(st)) return false;
  const size_t nn = (size_t)N * N;
  const int Kw = st.K_full ? N * N : N, Cw = st.C_full ? st.mr * st.mr : st.mr;
  std::memset(st.host, 0, st.bytes);
  for (int64_t i = 0; i < k; ++i) {
    kalman_step_equations_t* e = eq[i];
    const int n = e->dimension;
    st.first()[i] = i == 0;
    st.Fidx()[i] = i;
    if (i > 0) {
      const int l = e->c->row_dim, np = eq[i - 1]->dimension;
      copy_cols(st.F() + i * nn, N, e->F, l, np);
      if (!st.H_ident) copy_cols(st.H() + i * nn, N, e->H, l, n);
      copy_cols(st.c() + i * N, N, e->c, l, 1);
      double* K = st.K() + i * Kw;
      if (e->K_type == 'w') {
        copy_cols(K, N, e->K, l, 1);
        for (int r = l; r < N; ++r) K[r] = 1.0;
      } else {
        copy_cols(K, N, e->K, l, l);
        for (int r = l; r < N; ++r) K[(size_t)r * N + r] = 1.0;
      }
      st.Kt()[i] = (int8_t)cov_code(e->K_type);
    }
    const int m = e->o ? e->o->row_dim : 0, pad = N - n, mp = m + pad;
    st.m()[i] = mp;
    if (mp == 0) continue;
    double* G = st.G() + i * (size_t)st.mr * N;
    double* o = st.o() + i * (size_t)st.mr;
    double* C = st.C() + i * (size_t)Cw;
    const char ct = m > 0 ? e->C_type : 'w';
    if (m > 0) {
      copy_cols(G, st.mr, e->G, m, n);
      copy_cols(o, st.mr, e->o, m, 1);
    }
    for (int r = 0; r < pad; ++r) G[(size_t)(n + r) * st.mr + m + r] = 1.0; /* z_i = 0, unit weight */
    if (ct == 'w') {
      if (m > 0) copy_cols(C, st.mr, e->C, m, 1);
      for (int r = m; r < mp; ++r) C[r] = 1.0;
    } else {
      copy_cols(C, st.mr, e->C, m, m);
      for (int r = m; r < mp; ++r) C[(size_t)r * st.mr + r] = 1.0;
    }
    st.Ct()[i] = (int8_t)cov_code(ct);
  }
  st.fill_meff();
  return true;
}

/* Checks one ukhip_problem_t before anything is packed or launched: sizes,
 * covariance types, required arrays and every per-step observation count
 * (0 <= m_i <= m, so that a step never reaches past its m-row slot). */
static bool problem_ok(const ukhip_problem_t& q, int p) {
  if (q.n < 1 || q.n > UKHIP_MAX_N) { uk_set_error("problem %d: state dimension %d outside 1..%d", p, q.n, UKHIP_MAX_N); return false; }
  if (q.k < 1) { uk_set_error("problem %d: need k >= 1", p); return false; }
  if (q.m < 0 || q.m > max_obs_rows(q.n)) {
    uk_set_error("problem %d: m = %d outside 0..%d (observation rows the device reduces for n = %d)", p, q.m,
                 max_obs_rows(q.n), q.n);
    return false;
  }
  if (cov_code(q.K_type) < 0 || cov_code(q.C_type) < 0) { uk_set_error("problem %d: unknown covariance type", p); return false; }
  if (!q.F || !q.K || (q.m > 0 && (!q.G || !q.o || !q.C))) { uk_set_error("problem %d: missing input array", p); return false; }
  if (q.m_i) {
    for (int64_t i = 0; i < q.k; ++i)
      if (q.m_i[i] < 0 || q.m_i[i] > q.m) {
        uk_set_error("problem %d: step %lld has m_i = %d outside 0..m = %d", p, (long long)i, q.m_i[i], q.m);
        return false;
      }
  }
  return true;
}

/* Validates and packs ukhip_problem_t arrays. */
static bool pack_problems(int32_t nprob, const ukhip_problem_t* P, Staging& st, std::vector<int64_t>& lengths) {
  if (nprob < 1) { uk_set_error("no problems"); return false; }
  const int n = P[0].n;
  if (n < 1 || n > UKHIP_MAX_N) { uk_set_error("state dimension %d outside 1..%d", n, UKHIP_MAX_N); return false; }
  int mr = 1;
  int64_t S = 0, nF = 0;
  st.H_ident = 1;
  st.K_full = 0;
  st.C_full = 0;
  for (int p = 0; p < nprob; ++p) {
    const ukhip_problem_t& q = P[p];
    if (q.n != n) { uk_set_error("problem %d: n = %d differs from %d (one n per batch)", p, q.n, n); return false; }
    if (!problem_ok(q, p)) return false;
    if (q.H) st.H_ident = 0;
    if (q.K_type != 'w') st.K_full = 1;
    if (q.C_type != 'w') st.C_full = 1;
    mr = std::max(mr, q.m);
    S += q.k;
    nF += q.F_shared ? 1 : q.k;
    lengths.push_back(q.k);
  }
  st.n = n;
  st.mr = mr;
  st.steps = S;
  st.nF = nF;
  st.decide_compress();
  if (!staging_alloc(st)) return false;
  /* Every field a kernel reads is written below (no blanket clearing): rows
   * of G, o, C at or beyond a step's m_i and the evolution fields of a
   * problem's first step are never read. */
  const size_t nn = (size_t)n * n;
  const int Kw = st.K_full ? n * n : n, Cw = st.C_full ? mr * mr : mr;
  /* first step and first F block of every problem */
  std::vector<int64_t> s0v(nprob), fbv(nprob);
  int64_t s_acc = 0, f_acc = 0;
  for (int p = 0; p < nprob; ++p) {
    s0v[p] = s_acc;
    fbv[p] = f_acc;
    s_acc += P[p].k;
    f_acc += P[p].F_shared ? 1 : P[p].k;
  }
  auto pack_one = [&](int p) {
    const ukhip_problem_t& q = P[p];
    const int m = q.m;
    const int64_t k = q.k;
    const int64_t s0 = s0v[p], fb = fbv[p];
    const int kw = q.K_type == 'w' ? n : n * n;
    const int cw = q.C_type == 'w' ? m : m * m;
    const int8_t kt = (int8_t)cov_code(q.K_type), ct = (int8_t)cov_code(q.C_type);
    /* per-step scalars */
    for (int64_t i = 0; i < k; ++i) {
      const int64_t s = s0 + i;
      const int mi = q.m_i ? q.m_i[i] : m;
      st.first()[s] = i == 0;
      st.Fidx()[s] = q.F_shared ? fb : fb + i;
      st.Kt()[s] = kt;
      st.m()[s] = mi;
      st.Ct()[s] = ct;
    }
    /* evolution fields: whole problems at once (the first step's entries are
     * copied along and never read) */
    std::memcpy(st.F() + fb * nn, q.F, (q.F_shared ? 1 : (size_t)k) * nn * 8);
    if (!st.H_ident) {
      if (q.H) {
        std::memcpy(st.H() + s0 * nn, q.H, (size_t)k * nn * 8);
      } else {
        std::memset(st.H() + s0 * nn, 0, (size_t)k * nn * 8);
        for (int64_t i = 0; i < k; ++i)
          for (int d = 0; d < n; ++d) st.H()[(s0 + i) * nn + (size_t)d * n + d] = 1.0;
      }
    }
    if (q.c) std::memcpy(st.c() + s0 * n, q.c, (size_t)k * n * 8);
    else std::memset(st.c() + s0 * n, 0, (size_t)k * n * 8);
    if (kw == Kw) {
      std::memcpy(st.K() + s0 * Kw, q.K, (size_t)k * kw * 8);
    } else {
      for (int64_t i = 0; i < k; ++i) std::memcpy(st.K() + (s0 + i) * Kw, q.K + i * kw, (size_t)kw * 8);
    }
    /* observations: one copy when the request's row count is the batch's
     * stride, else rows 0 .. m_i - 1 step by step (rows beyond m_i are never read) */
    if (m > 0) {
      if (m == mr) {
        std::memcpy(st.G() + s0 * (size_t)mr * n, q.G, (size_t)k * m * n * 8);
        std::memcpy(st.o() + s0 * (size_t)mr, q.o, (size_t)k * m * 8);
        if (cw == Cw) std::memcpy(st.C() + s0 * (size_t)Cw, q.C, (size_t)k * cw * 8);
        else
          for (int64_t i = 0; i < k; ++i) std::memcpy(st.C() + (s0 + i) * (size_t)Cw, q.C + i * (size_t)cw, (size_t)cw * 8);
      } else {
        for (int64_t i = 0; i < k; ++i) {
          const int mi = st.m()[s0 + i];
          if (mi <= 0) continue;
          double* G = st.G() + (s0 + i) * (size_t)mr * n;
          const double* g = q.G + i * (size_t)m * n;
          for (int col = 0; col < n; ++col)
            for (int r = 0; r < mi; ++r) G[(size_t)col * mr + r] = g[(size_t)col * m + r];
          double* o = st.o() + (s0 + i) * (size_t)mr;
          for (int r = 0; r < mi; ++r) o[r] = q.o[i * (size_t)m + r];
          double* C = st.C() + (s0 + i) * (size_t)Cw;
          const double* cc = q.C + i * (size_t)cw;
          if (q.C_type == 'w') {
            for (int r = 0; r < mi; ++r) C[r] = cc[r];
          } else {
            for (int col = 0; col < mi; ++col)
              for (int r = 0; r < mi; ++r) C[(size_t)col * mr + r] = cc[(size_t)col * m + r];
          }
        }
      }
    }
  };
  /* problems write disjoint staging ranges: large batches pack on several threads */
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int nt = (int)std::min<int64_t>({(int64_t)std::min(8u, hw), (int64_t)nprob, std::max<int64_t>(1, S / 16384)});
  if (nt <= 1) {
    for (int p = 0; p < nprob; ++p) pack_one(p);
  } else {
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
      pool.emplace_back([&] {
        for (int p; (p = next.fetch_add(1)) < nprob;) pack_one(p);
      });
    for (auto& th : pool) th.join();
  }
  st.fill_meff();
  return true;
}

/* ------------------------------------------------------------------------ */
/* running                                                                    */
/* ------------------------------------------------------------------------ */

/* Device time per kernel class (the UKHIP_OE_* slots) of runs on one stream:
 * a pair of HIP events around every launch group, summed per slot once the
 * stream has finished. A phase triple of the elimination records four events
 * (before phase 1, after each phase). */
struct SlotTimer {
  hipStream_t st = nullptr;
  struct Rec {
    int slot;
    hipEvent_t a, b;
    int launches;
  };
  std::vector<Rec> rec;
  std::vector<hipEvent_t> all;
  hipEvent_t first = nullptr, last = nullptr;
  bool ok = true;

  hipEvent_t make() {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) {
      ok = false;
      return nullptr;
    }
    all.push_back(e);
    return e;
  }
  hipEvent_t mark() {
    hipEvent_t e = make();
    if (e && hipEventRecord(e, st) != hipSuccess) ok = false;
    if (!first) first = e;
    last = e;
    return e;
  }
  /* begin() ... launches ... end(slot, launches) */
  hipEvent_t pending = nullptr;
  void begin() { pending = mark(); }
  void end(int slot, int launches = 1) { rec.push_back({slot, pending, mark(), launches}); }
  /* four recorded events for uk_launch_eliminate(32) */
  hipEvent_t quad[4]{};
  hipEvent_t* phases() {
    for (auto& e : quad) e = make();
    return quad;
  }
  void phases_done() {
    if (!first) first = quad[0];
    last = quad[3];
    for (int p = 0; p < 3; ++p) rec.push_back({UKHIP_OE_PHASE1 + p, quad[p], quad[p + 1], 1});
  }
  /* waits for the stream, adds the slot sums into ms and launches */
  bool collect(double* ms, int32_t* launches) {
    if (!ok || !last || hipEventSynchronize(last) != hipSuccess) return false;
    for (auto& r : rec) {
      float t = 0;
      if (hipEventElapsedTime(&t, r.a, r.b) != hipSuccess) return false;
      ms[r.slot] += t;
      if (launches) launches[r.slot] += r.launches;
    }
    float t = 0;
    if (hipEventElapsedTime(&t, first, last) != hipSuccess) return false;
    ms[UKHIP_OE_TOTAL] += t;
    if (launches) launches[UKHIP_OE_TOTAL] += 1;
    return true;
  }
  ~SlotTimer() {
    for (auto e : all)
      if (e) (void)hipEventDestroy(e);
  }
};

/* The level-0 chain's D = V H is diagonal when every H is the identity and
 * every K is given as 'w' (V diagonal). When C is 'w' too, k_weigh_diag writes
 * the chain and stores D as its diagonal (d_diag); the phase kernels and
 * k_normal read it so, and phase 2 skips D's exact zeros. With every K and C
 * given as 'w', the level-0 chain's B_s = -diag(k_s) F is not stored either
 * (b_synth): its readers form it from the raw K and F. (B: a batch or a
 * uk_oe_view_t, both carrying raw, nF, compress and mr0.) */
template <class B, class Chain>
static Chain level0_chain(const B* b, Chain c) {
  const bool d_diag = b->raw.H_ident && !b->raw.K_full && !b->raw.C_full && !b->compress;
  c.d_diag = d_diag;
  c.b_synth = !b->raw.K_full && !b->raw.C_full && !b->compress; /* compressed chains come from k_weigh */
  c.mr = b->mr0;
  c.c_tri = b->compress; /* k_obs_compress leaves every C upper triangular */
  c.wk = b->raw.K;
  c.wF = b->raw.F;
  c.wFidx = b->nF == 1 ? nullptr : b->raw.Fidx; /* one F block: no per-step lookup */
  return c;
}

/* The whitened level-0 chain (c/kalman_oddeven_smoother.c:1620-1659): k_weigh
 * or the element-parallel k_weigh_diag; with observation blocks taller than n,
 * k_weigh followed by k_obs_compress. */
template <class B>
static bool weigh_level0(B* b, hipStream_t st) {
  if (b->compress) return hip_ok(uk_launch_weigh_obs(&b->raw, &b->chain[0], b->wobs, st), "k_weigh launch");
  const uk_chain_t w0 = level0_chain(b, b->chain[0]);
  return hip_ok(uk_launch_weigh(&b->raw, &w0, st), "k_weigh launch");
}

/* One whole odd-even smoothing (c/kalman_oddeven_smoother.c:1358-1553,
 * 614-975) over the buffers of v, enqueued on st. */
static bool run_oddeven_view(uk_oe_view_t* v, int cov, SlotTimer* tm, hipStream_t st) {
  const int nl = v->nl, n = v->n;
  const size_t nn = (size_t)n * n;
  if (tm) tm->begin();
  if (!weigh_level0(v, st)) return false;
  if (tm) tm->end(UKHIP_OE_WEIGH, v->compress ? 2 : 1);
  auto fac_at = [&](int l) {
    uk_factors_t f = v->fac;
    const int64_t fb = v->fac_base[l];
    f.R += fb * nn;
    f.Bt += fb * nn;
    f.Y += fb * nn;
    f.cf += fb * n;
    return f;
  };
  for (int l = 0; l < nl; ++l) {
    uk_chain_t in = l == 0 ? level0_chain(v, v->chain[0]) : v->chain[l & 1];
    if (l > 0) {
      in.mr = n;
      in.c_tri = 1;
    }
    uk_chain_t out = v->chain[(l + 1) & 1]; /* its flags stay 0: a dense chain */
    out.mr = n;
    const uk_factors_t f = fac_at(l);
    uk_level_t lv = v->levels[l]; /* even_start is relative to the level's fac_base */
    if (!hip_ok(uk_launch_eliminate(&lv, &in, &out, &f, v->fix, n, st, tm ? tm->phases() : nullptr),
                "elimination launch"))
      return false;
    if (tm) tm->phases_done();
  }
  for (int l = nl - 1; l >= 0; --l) {
    const uk_factors_t f = fac_at(l);
    uk_level_t lv = v->levels[l];
    if (tm) tm->begin();
    double* Xc = l == 0 ? nullptr : v->X[l & 1];
    double* XTc = l == 0 ? nullptr : v->XT[l & 1];
    /* cov 2: only level 0's covariances are reduced to their diagonals (the
     * levels above feed the level below their full blocks) */
    const int cl = cov == 2 ? (l == 0 ? 2 : 1) : cov;
    if (!hip_ok(uk_launch_backsub(&lv, &f, v->xout, v->Sout, v->X[(l + 1) & 1], v->XT[(l + 1) & 1], Xc, XTc, n, cl,
                                  st),
                "k_backsub launch"))
      return false;
    if (tm) tm->end(UKHIP_OE_BACKSUB);
  }
  return true;
}

/* for the allocation-free entry point (noalloc.hip) */
extern "C" int uk_run_oddeven_view(uk_oe_view_t* v, int cov, hipStream_t st) {
  return run_oddeven_view(v, cov, nullptr, st) ? 0 : -1;
}

static bool run_oddeven(ukhip_batch* b, int cov, SlotTimer* tm, hipStream_t st) {
  if (!ensure_elim(b)) return false;
  const int nl = (int)b->levels.size();
  std::vector<uk_level_t> lv(nl);
  std::vector<int64_t> fb(nl);
  for (int l = 0; l < nl; ++l) {
    lv[l] = b->levels[l].dev;
    fb[l] = b->levels[l].fac_base;
  }
  uk_oe_view_t v{};
  v.raw = b->raw;
  v.chain[0] = b->chain[0];
  v.chain[1] = b->chain[1];
  v.fac = b->fac;
  v.X[0] = b->X[0];
  v.X[1] = b->X[1];
  v.XT[0] = b->XT[0];
  v.XT[1] = b->XT[1];
  v.xout = b->xout;
  v.Sout = b->Sout;
  v.fix = b->fix;
  v.wobs = b->wobs;
  v.n = b->n;
  v.mr0 = b->mr0;
  v.compress = b->compress ? 1 : 0;
  v.nl = nl;
  v.nF = b->nF;
  v.levels = lv.data();
  v.fac_base = fb.data();
  return run_oddeven_view(&v, cov, tm, st);
}

static ukhip_batch* batch_create_one(int32_t nprob, const ukhip_problem_t* problems) {
  Staging st;
  std::vector<int64_t> lengths;
  if (!pack_problems(nprob, problems, st, lengths)) {
    staging_free(st);
    return nullptr;
  }
  ukhip_batch* b = batch_from_staging(st, lengths);
  if (b && !hip_ok(hipStreamSynchronize(g_stream), "upload")) {
    batch_release(b);
    b = nullptr;
  }
  staging_free(st);
  return b;
}

extern "C" ukhip_batch_t* ukhip_batch_create(int32_t nprob, const ukhip_problem_t* problems) {
  if (nprob < 1 || !problems) { uk_set_error("ukhip_batch_create: no problems"); return nullptr; }
  for (int32_t p = 0; p < nprob; ++p) /* every input is checked before the device is touched */
    if (!problem_ok(problems[p], p)) return nullptr;
  if (!ensure_stream()) return nullptr;
  bool mixed = false;
  for (int32_t p = 1; p < nprob; ++p) mixed |= problems[p].n != problems[0].n;
  if (!mixed) return batch_create_one(nprob, problems);
  /* one part per state dimension, problems in input order within a part */
  std::vector<int32_t> dims;
  for (int32_t p = 0; p < nprob; ++p)
    if (std::find(dims.begin(), dims.end(), problems[p].n) == dims.end()) dims.push_back(problems[p].n);
  std::sort(dims.begin(), dims.end());
  ukhip_batch* b = new ukhip_batch();
  b->nprob = nprob;
  b->where.resize(nprob);
  for (int32_t d : dims) {
    std::vector<ukhip_problem_t> sub;
    for (int32_t p = 0; p < nprob; ++p)
      if (problems[p].n == d) {
        b->where[p] = {(int32_t)b->parts.size(), (int32_t)sub.size()};
        sub.push_back(problems[p]);
      }
    ukhip_batch* q = batch_create_one((int32_t)sub.size(), sub.data());
    if (!q) {
      batch_release(b);
      return nullptr;
    }
    b->parts.push_back(q);
    b->steps += q->steps;
    b->n = std::max(b->n, q->n);
  }
  for (size_t i = 0; i < b->parts.size(); ++i) {
    hipStream_t st = nullptr;
    if (!hip_ok(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate")) {
      batch_release(b);
      return nullptr;
    }
    b->part_streams.push_back(st);
  }
  return b;
}

extern "C" void ukhip_batch_free(ukhip_batch_t* b) {
  if (b && g_stream) (void)hipStreamSynchronize(g_stream);
  batch_release(b);
}
extern "C" int64_t ukhip_batch_steps(const ukhip_batch_t* b) { return b ? b->steps : 0; }
extern "C" int32_t ukhip_batch_levels(const ukhip_batch_t* b) {
  if (!b) return 0;
  int32_t l = (int32_t)b->levels.size();
  for (auto* q : b->parts) l = std::max(l, (int32_t)q->levels.size());
  return l;
}

static bool g_use_graphs = true;
extern "C" void ukhip_use_graphs(int enable) { g_use_graphs = enable != 0; }

/* Enqueues a whole odd-even smoothing of b on the library stream: the parts
 * of a composite run concurrently, each on a stream forked from it. */
static bool enqueue_oddeven(ukhip_batch* b, int cov) {
  if (b->parts.empty()) return run_oddeven(b, cov, nullptr, g_stream);
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> joins(b->parts.size(), nullptr);
  bool ok = hip_ok(hipEventCreateWithFlags(&fork, hipEventDisableTiming), "hipEventCreate") &&
            hip_ok(hipEventRecord(fork, g_stream), "hipEventRecord");
  for (size_t i = 0; ok && i < b->parts.size(); ++i) {
    ok = hip_ok(hipStreamWaitEvent(b->part_streams[i], fork, 0), "hipStreamWaitEvent") &&
         run_oddeven(b->parts[i], cov, nullptr, b->part_streams[i]) &&
         hip_ok(hipEventCreateWithFlags(&joins[i], hipEventDisableTiming), "hipEventCreate") &&
         hip_ok(hipEventRecord(joins[i], b->part_streams[i]), "hipEventRecord") &&
         hip_ok(hipStreamWaitEvent(g_stream, joins[i], 0), "hipStreamWaitEvent");
  }
  if (fork) (void)hipEventDestroy(fork);
  for (auto e : joins)
    if (e) (void)hipEventDestroy(e);
  return ok;
}

/* Captures enqueue_oddeven into a graph once per covariance setting; replays
 * it afterwards (the captured launches reference b's device buffers, which
 * live as long as b). */
/* Covariance mode of options: 0 none (KALMAN_NO_COVARIANCE), 2 diagonals
 * (UKHIP_COVARIANCE_DIAG), 1 full. */
static int cov_mode(uint32_t options) {
  if (options & KALMAN_NO_COVARIANCE) return 0;
  return (options & UKHIP_COVARIANCE_DIAG) ? 2 : 1;
}

static bool graph_oddeven(ukhip_batch* b, int cov) {
  if (!b->graph[cov]) {
    if (!ensure_elim_all(b)) return false; /* no allocation inside the capture */
    if (!hip_ok(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"))
      return false;
    const bool ok = enqueue_oddeven(b, cov);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(g_stream, &g);
    if (!ok || !hip_ok(e, "hipStreamEndCapture")) {
      if (g) (void)hipGraphDestroy(g);
      return false;
    }
    const hipError_t ei = hipGraphInstantiate(&b->graph[cov], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (!hip_ok(ei, "hipGraphInstantiate")) {
      b->graph[cov] = nullptr;
      return false;
    }
  }
  return hip_ok(hipGraphLaunch(b->graph[cov], g_stream), "hipGraphLaunch");
}

extern "C" int ukhip_batch_smooth_oddeven(ukhip_batch_t* b, uint32_t options) {
  if (!b) { uk_set_error("NULL batch"); return -1; }
  if (!ensure_stream()) return -1;
  const int cov = cov_mode(options);
  return (g_use_graphs ? graph_oddeven(b, cov) : enqueue_oddeven(b, cov)) ? 0 : -1;
}

extern "C" int ukhip_batch_smooth_associative(ukhip_batch_t* b, uint32_t options) {
  if (!b) { uk_set_error("NULL batch"); return -1; }
  if (!ensure_stream()) return -1;
  if (!ensure_elim_all(b)) return -1; /* the covariance output */
  if (b->parts.empty()) return uk_assoc_smooth(b, options);
  for (auto* q : b->parts)
    if (uk_assoc_smooth(q, options) != 0) return -1;
  return 0;
}

extern "C" int ukhip_batch_smooth_associative_timed(ukhip_batch_t* b, uint32_t options, double* ms,
                                                    int32_t* launches) {
  if (!b) { uk_set_error("NULL batch"); return -1; }
  if (!ensure_stream()) return -1;
  if (!ensure_elim_all(b)) return -1;
  if (b->parts.empty()) return uk_assoc_smooth_timed(b, options, ms, launches);
  for (auto* q : b->parts)
    if (uk_assoc_smooth_timed(q, options, ms, launches) != 0) return -1;
  return 0;
}

static bool run_mixed(ukhip_batch* b, int refinements, int fp32_levels, bool cov, SlotTimer* tm, hipStream_t st);
static int mixed_fp32_levels(const ukhip_mixed_options_t* o);

/* One part (a batch without parts) run eagerly with a slot timer; the slot
 * sums are added into ms and launches. */
static bool timed_part(ukhip_batch* q, uint32_t options, const ukhip_mixed_options_t* mixed, double* ms,
                       int32_t* launches) {
  SlotTimer tm;
  tm.st = g_stream;
  const bool ok = mixed ? run_mixed(q, mixed->refinements, mixed_fp32_levels(mixed), mixed->covariances != 0, &tm,
                                    g_stream)
                        : run_oddeven(q, cov_mode(options), &tm, g_stream);
  if (!ok) return false;
  if (!tm.collect(ms, launches)) return hip_ok(hipErrorUnknown, "timed smoothing events");
  return true;
}

extern "C" int ukhip_batch_timed_parts(ukhip_batch_t* b, uint32_t options, const ukhip_mixed_options_t* mixed,
                                       double* ms, int32_t* launches, int32_t* part_n, int32_t max_parts) {
  if (!b || !ms || max_parts < 1) { uk_set_error("NULL batch or result array"); return -1; }
  if (!ensure_stream()) return -1;
  if (mixed && (mixed->refinements < 0 || mixed->refinements > 64)) { uk_set_error("refinements must be in 0..64"); return -1; }
  std::vector<ukhip_batch*> parts = b->parts;
  if (parts.empty()) parts.push_back(b);
  if ((int)parts.size() > max_parts) { uk_set_error("the batch has %zu parts, more than max_parts", parts.size()); return -1; }
  for (size_t i = 0; i < parts.size(); ++i) {
    double* m = ms + i * UKHIP_OE_SLOTS;
    int32_t* l = launches ? launches + i * UKHIP_OE_SLOTS : nullptr;
    for (int t = 0; t < UKHIP_OE_SLOTS; ++t) {
      m[t] = 0.0;
      if (l) l[t] = 0;
    }
    if (part_n) part_n[i] = parts[i]->n;
    if (!timed_part(parts[i], options, mixed, m, l)) return -1;
  }
  return (int)parts.size();
}

extern "C" int ukhip_batch_smooth_oddeven_timed(ukhip_batch_t* b, uint32_t options, double ms[UKHIP_OE_SLOTS],
                                                int32_t launches[UKHIP_OE_SLOTS]) {
  if (!b || !ms) { uk_set_error("NULL batch or result array"); return -1; }
  const int np = std::max<int>(1, (int)b->parts.size());
  std::vector<double> m(np * UKHIP_OE_SLOTS);
  std::vector<int32_t> l(np * UKHIP_OE_SLOTS);
  if (ukhip_batch_timed_parts(b, options, nullptr, m.data(), l.data(), nullptr, np) < 0) return -1;
  for (int t = 0; t < UKHIP_OE_SLOTS; ++t) {
    ms[t] = 0.0;
    if (launches) launches[t] = 0;
    for (int p = 0; p < np; ++p) {
      ms[t] += m[p * UKHIP_OE_SLOTS + t];
      if (launches) launches[t] += l[p * UKHIP_OE_SLOTS + t];
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* mixed precision: FP32 elimination + FP64 refinement (refine.hip)          */
/* ------------------------------------------------------------------------ */

static bool mixed_workspace(ukhip_batch* b) {
  if (b->d_mixed) return true;
  const size_t n = b->n, nn = n * n, S = (size_t)b->steps;
  const size_t L1 = b->levels.size() > 1 ? (size_t)b->levels[1].steps : 1;
  const size_t E = (size_t)b->total_evens;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t r = o;
    o += (bytes + 255) & ~(size_t)255;
    return r;
  };
  size_t ch[2][5];
  for (int cb = 0; cb < 2; ++cb) {
    const size_t cs = cb == 0 ? S : std::max<size_t>(L1, 1);
    ch[cb][0] = take(cs * nn * 4);
    ch[cb][1] = take(cs * nn * 4);
    ch[cb][2] = take(cs * n * 4);
    ch[cb][3] = take(cs * nn * 4);
    ch[cb][4] = take(cs * n * 4);
  }
  const size_t fR = take(E * nn * 4), fB = take(E * nn * 4), fY = take(E * nn * 4), fc = take(E * n * 4);
  const size_t ofx = take((size_t)b->nprob * (nn + n) * 4);
  const size_t ogA = take(S * n * 8), ogB = take(S * n * 8), og0 = take(S * n * 8),
               og1 = take(std::max<size_t>(L1, 1) * n * 8), oys = take(E * n * 8), odx = take(S * n * 8);
  b->d_mixed = dev_alloc(b, o, "device memory for the mixed-precision workspace");
  if (!b->d_mixed) return false;
  char* w = (char*)b->d_mixed;
  for (int cb = 0; cb < 2; ++cb) {
    b->chain32[cb].B = (float*)(w + ch[cb][0]);
    b->chain32[cb].D = (float*)(w + ch[cb][1]);
    b->chain32[cb].c = (float*)(w + ch[cb][2]);
    b->chain32[cb].C = (float*)(w + ch[cb][3]);
    b->chain32[cb].o = (float*)(w + ch[cb][4]);
  }
  b->fac32.R = (float*)(w + fR);
  b->fac32.Bt = (float*)(w + fB);
  b->fac32.Y = (float*)(w + fY);
  b->fac32.cf = (float*)(w + fc);
  b->fix32 = (float*)(w + ofx);
  b->gA = (double*)(w + ogA);
  b->gB = (double*)(w + ogB);
  b->gbuf[0] = (double*)(w + og0);
  b->gbuf[1] = (double*)(w + og1);
  b->ys = (double*)(w + oys);
  b->dx = (double*)(w + odx);
  return true;
}

/* The FP64 storage of levels lsw .. top for the mixed path (see TopStore);
 * idempotent per lsw, never inside a stream capture. */
static const ukhip_batch::TopStore* ensure_top(ukhip_batch* b, int lsw) {
  auto it = b->tops.find(lsw);
  if (it != b->tops.end()) return &it->second;
  const int nl = (int)b->levels.size();
  const size_t n = b->n, nn = n * n;
  const size_t E = (size_t)(b->total_evens - b->levels[lsw].fac_base);
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t r = o;
    o += (bytes + 255) & ~(size_t)255;
    return r;
  };
  size_t ch[2][5];
  for (int par = 0; par < 2; ++par) { /* ch[lsw & 1] holds level lsw, the other level lsw + 1 */
    const int l = (par == (lsw & 1)) ? lsw : lsw + 1;
    const size_t cs = l < nl ? (size_t)std::max<int64_t>(b->levels[l].steps, 1) : 1;
    ch[par][0] = take(cs * nn * 8);
    ch[par][1] = take(cs * nn * 8);
    ch[par][2] = take(cs * n * 8);
    ch[par][3] = take(cs * nn * 8);
    ch[par][4] = take(cs * n * 8);
  }
  const size_t fR = take(E * nn * 8), fB = take(E * nn * 8), fY = take(E * nn * 8), fc = take(E * n * 8);
  void* mem = dev_alloc(b, o, "device memory for the mixed path's FP64 levels");
  if (!mem) return nullptr;
  char* w = (char*)mem;
  ukhip_batch::TopStore ts;
  ts.mem = mem;
  for (int par = 0; par < 2; ++par) {
    ts.ch[par].B = (double*)(w + ch[par][0]);
    ts.ch[par].D = (double*)(w + ch[par][1]);
    ts.ch[par].c = (double*)(w + ch[par][2]);
    ts.ch[par].C = (double*)(w + ch[par][3]);
    ts.ch[par].o = (double*)(w + ch[par][4]);
  }
  ts.fac.R = (double*)(w + fR);
  ts.fac.Bt = (double*)(w + fB);
  ts.fac.Y = (double*)(w + fY);
  ts.fac.cf = (double*)(w + fc);
  return &(b->tops[lsw] = ts);
}

/* first FP64 level of a mixed smoothing with fp32_levels FP32 levels (level 0
 * stays FP32); nl when every level is FP32 */
static int mixed_lsw(const ukhip_batch* b, int fp32_levels) {
  const int nl = (int)b->levels.size();
  return std::max(1, std::min(fp32_levels, nl));
}

/* Everything a mixed smoothing allocates, before any capture. */
static bool mixed_prepare(ukhip_batch* b, int fp32_levels, bool cov) {
  if (!b->parts.empty()) {
    for (auto* q : b->parts)
      if (!mixed_prepare(q, fp32_levels, cov)) return false;
    return true;
  }
  if (!mixed_workspace(b)) return false;
  const int lsw = mixed_lsw(b, fp32_levels);
  if (lsw < (int)b->levels.size() && !ensure_top(b, lsw)) return false;
  return !cov || ensure_elim(b);
}

static uk_factors32_t fac32_at(const ukhip_batch* b, int l) {
  uk_factors32_t f = b->fac32;
  const int64_t fb = b->levels[l].fac_base;
  const size_t nn = (size_t)b->n * b->n;
  f.R += fb * nn;
  f.Bt += fb * nn;
  f.Y += fb * nn;
  f.cf += fb * b->n;
  return f;
}

/* level l's FP64 factor slots: in the full factor storage (base 0), or in a
 * TopStore whose first level is `from` */
static uk_factors_t fac64_at(const ukhip_batch* b, int l, const uk_factors_t& all, int from = 0) {
  uk_factors_t f = all;
  const int64_t fb = b->levels[l].fac_base - b->levels[from].fac_base;
  const size_t nn = (size_t)b->n * b->n;
  f.R += fb * nn;
  f.Bt += fb * nn;
  f.Y += fb * nn;
  f.cf += fb * b->n;
  return f;
}

/* Levels 0 .. fp32_levels - 1 of the elimination run in FP32, the levels
 * above in FP64 (where the reduced blocks' conditioning compounds); the
 * estimates, in the original step order of xout, come from the FP32/FP64
 * transformed right-hand sides and `refinements` FP64 refinement steps over
 * the factors of both precisions. With cov, the covariances are then formed
 * from the same factors by the FP64 selected-inversion back substitution
 * (the FP32 factors widened into the FP64 factor slots), into Sout. */
static bool run_mixed(ukhip_batch* b, int refinements, int fp32_levels, bool cov, SlotTimer* tm, hipStream_t st) {
  if (!mixed_prepare(b, fp32_levels, cov)) return false;
  auto tb = [&]() {
    if (tm) tm->begin();
  };
  auto te = [&](int slot, int launches) {
    if (tm) tm->end(slot, launches);
  };
  const int n = b->n, nl = (int)b->levels.size();
  const int lsw = mixed_lsw(b, fp32_levels); /* first FP64 level (level 0 stays FP32) */
  const ukhip_batch::TopStore* ts = lsw < nl ? &b->tops.at(lsw) : nullptr;
  auto fac64 = [&](int l) { return fac64_at(b, l, ts->fac, lsw); };
  const size_t nn = (size_t)n * n, S = (size_t)b->steps;
  /* the level-0 chain in FP32 (the FP64 chain stays for the residuals):
   * written beside the FP64 one when the whitening is diagonal, else converted */
  const uk_chain_t& c64 = b->chain[0];
  uk_chain32_t c32 = b->chain32[0];
  const uk_chain_t w0 = level0_chain(b, b->chain[0]); /* its flags imply the fused path below */
  /* With every K and C given as 'w' the FP32 elimination forms the level-0 C =
   * diag(w) G on load (phases 1 and 3), so the whitening does not store it. */
  const bool csynth = w0.b_synth != 0;
  c32.c_synth = csynth ? 1 : 0;
  c32.rG = b->raw.G;
  c32.rC = b->raw.C;
  c32.rm = b->m0;
  /* With diagonal whitening and n, m <= 32 the refinement's k_normal_w forms the
   * FP64 level-0 chain from the raw inputs, so the whitening writes only the
   * FP32 copy. */
  const bool from_raw = w0.d_diag && w0.b_synth && uk_bucket_qr(std::max(n, b->mr0)) <= 32;
  tb();
  const hipError_t fused =
      b->compress ? hipErrorNotSupported : uk_launch_weigh_f32(&b->raw, &w0, &c32, from_raw ? 0 : 1, st);
  if (fused != hipErrorNotSupported) {
    if (!hip_ok(fused, "k_weigh launch")) return false;
    te(UKHIP_OE_WEIGH, 1);
  } else {
    if (!weigh_level0(b, st)) return false;
    te(UKHIP_OE_WEIGH, b->compress ? 2 : 1);
    tb();
    if (!hip_ok(uk_launch_to32(c32.B, c64.B, S * nn, st), "to32") ||
        !hip_ok(uk_launch_to32(c32.D, c64.D, S * nn, st), "to32") ||
        !hip_ok(uk_launch_to32(c32.c, c64.c, S * n, st), "to32") ||
        !hip_ok(uk_launch_to32(c32.C, c64.C, S * nn, st), "to32") ||
        !hip_ok(uk_launch_to32(c32.o, c64.o, S * n, st), "to32"))
      return false;
    te(UKHIP_OE_CONVERT, 5);
  }
  uk_chain_t top[2] = {ts ? ts->ch[0] : uk_chain_t{}, ts ? ts->ch[1] : uk_chain_t{}};
  for (int l = 0; l < nl; ++l) {
    uk_level_t lv = b->levels[l].dev;
    if (l < lsw) {
      uk_chain32_t in = l == 0 ? level0_chain(b, c32) : b->chain32[l & 1];
      if (l > 0) {
        in.mr = n;
        in.c_tri = 1;
        in.c_synth = 0;
      }
      uk_chain32_t out = b->chain32[(l + 1) & 1]; /* its flags stay 0: a dense chain */
      out.mr = n;
      const uk_factors32_t f = fac32_at(b, l);
      if (!hip_ok(uk_launch_eliminate32(&lv, &in, &out, &f, b->fix32, n, st, tm ? tm->phases() : nullptr),
                  "FP32 elimination launch"))
        return false;
      if (tm) tm->phases_done();
      if (l + 1 == lsw && lsw < nl) { /* the next level's chain enters FP64 */
        const size_t ns = (size_t)b->levels[lsw].steps;
        const uk_chain_t& t = top[lsw & 1];
        tb();
        if (!hip_ok(uk_launch_to64(t.B, out.B, ns * nn, st), "to64") ||
            !hip_ok(uk_launch_to64(t.D, out.D, ns * nn, st), "to64") ||
            !hip_ok(uk_launch_to64(t.c, out.c, ns * n, st), "to64") ||
            !hip_ok(uk_launch_to64(t.C, out.C, ns * nn, st), "to64") ||
            !hip_ok(uk_launch_to64(t.o, out.o, ns * n, st), "to64"))
          return false;
        te(UKHIP_OE_CONVERT, 5);
      }
    } else {
      uk_chain_t in = top[l & 1];
      in.mr = n;
      in.c_tri = 1;
      uk_chain_t out = top[(l + 1) & 1];
      out.mr = n;
      const uk_factors_t f = fac64(l);
      if (!hip_ok(uk_launch_eliminate(&lv, &in, &out, &f, b->fix, n, st, tm ? tm->phases() : nullptr),
                  "FP64 elimination launch"))
        return false;
      if (tm) tm->phases_done();
    }
  }
  /* x0 from the transformed right-hand sides */
  for (int l = nl - 1; l >= 0; --l) {
    uk_level_t lv = b->levels[l].dev;
    tb();
    if (l < lsw) {
      const uk_factors32_t f = fac32_at(b, l);
      if (!hip_ok(uk_launch_solve_est(&lv, &f, f.cf, nullptr, b->xout, n, st), "solve launch")) return false;
    } else {
      const uk_factors_t f = fac64(l);
      if (!hip_ok(uk_launch_solve_est64(&lv, &f, f.cf, b->xout, n, st), "solve launch")) return false;
    }
    te(UKHIP_OE_SOLVE, 1);
  }
  uk_chain_t r64 = level0_chain(b, b->chain[0]);
  r64.from_raw = from_raw ? 1 : 0;
  r64.rG = b->raw.G;
  r64.rC = b->raw.C;
  r64.ro = b->raw.o;
  r64.rc = b->raw.c;
  for (int it = 0; it < refinements; ++it) {
    tb();
    if (!hip_ok(uk_launch_normal(&r64, b->raw.first, b->m0, b->xout, b->gA, b->gB, b->gbuf[0], (int64_t)S, n, st),
                "residual launch"))
      return false;
    te(UKHIP_OE_NORMAL, 2);
    for (int l = 0; l < nl; ++l) {
      uk_level_t lv = b->levels[l].dev;
      double* y = b->ys + b->levels[l].fac_base * n;
      hipError_t e;
      tb();
      if (l < lsw) {
        const uk_factors32_t f = fac32_at(b, l);
        e = uk_launch_fwd(&lv, &f, b->gbuf[l & 1], y, b->gbuf[(l + 1) & 1], n, st);
      } else {
        const uk_factors_t f = fac64(l);
        e = uk_launch_fwd64(&lv, &f, b->gbuf[l & 1], y, b->gbuf[(l + 1) & 1], n, st);
      }
      if (!hip_ok(e, "forward sweep launch")) return false;
      te(UKHIP_OE_FWD, 2);
    }
    for (int l = nl - 1; l >= 0; --l) {
      uk_level_t lv = b->levels[l].dev;
      const double* y = b->ys + b->levels[l].fac_base * n;
      hipError_t e;
      tb();
      if (l < lsw) {
        const uk_factors32_t f = fac32_at(b, l);
        e = uk_launch_solve_est(&lv, &f, nullptr, y, b->dx, n, st);
      } else {
        const uk_factors_t f = fac64(l);
        e = uk_launch_solve_est64(&lv, &f, y, b->dx, n, st);
      }
      if (!hip_ok(e, "backward sweep launch")) return false;
      te(UKHIP_OE_SOLVE, 1);
    }
    tb();
    if (!hip_ok(uk_launch_add(b->xout, b->dx, (int64_t)(S * n), st), "update launch")) return false;
    te(UKHIP_OE_NORMAL, 1);
  }
  if (!cov) return true;
  /* covariances: the FP32 levels' factors widened into their FP64 slots (the
   * slots of levels 0 .. lsw-1 are contiguous), then the FP64 back
   * substitution over those and the FP64 levels' own factors; its estimates
   * go to the scratch dx, xout keeps the refined ones */
  const size_t E32 = (size_t)(lsw < nl ? b->levels[lsw].fac_base : b->total_evens);
  tb();
  if (!hip_ok(uk_launch_to64(b->fac.R, b->fac32.R, E32 * nn, st), "to64") ||
      !hip_ok(uk_launch_to64(b->fac.Bt, b->fac32.Bt, E32 * nn, st), "to64") ||
      !hip_ok(uk_launch_to64(b->fac.Y, b->fac32.Y, E32 * nn, st), "to64") ||
      !hip_ok(uk_launch_to64(b->fac.cf, b->fac32.cf, E32 * n, st), "to64"))
    return false;
  te(UKHIP_OE_CONVERT, 4);
  for (int l = nl - 1; l >= 0; --l) {
    const uk_factors_t f = l < lsw ? fac64_at(b, l, b->fac) : fac64(l);
    uk_level_t lv = b->levels[l].dev;
    double* Xc = l == 0 ? nullptr : b->X[l & 1];
    double* XTc = l == 0 ? nullptr : b->XT[l & 1];
    tb();
    if (!hip_ok(uk_launch_backsub(&lv, &f, b->dx, b->Sout, b->X[(l + 1) & 1], b->XT[(l + 1) & 1], Xc, XTc, n, 1, st),
                "k_backsub launch"))
      return false;
    te(UKHIP_OE_BACKSUB, 1);
  }
  return true;
}

/* Default precision switch of the mixed path: levels 0..7 in FP32, the levels
 * above (at most 1/256 of a problem's steps each) in FP64. On MI355X the FP64
 * levels cost 0.4% of the n = 64, k = 200000 smoothing; they did not change
 * the error after one refinement step (1.74e-9 either way), which the level-0
 * FP32 factors set. */
constexpr int MIXED_FP32_LEVELS = 8;
static int mixed_fp32_levels(const ukhip_mixed_options_t* o) {
  return o->fp32_levels <= 0 ? MIXED_FP32_LEVELS : std::min(o->fp32_levels, 64);
}

/* Enqueues a whole mixed-precision smoothing on the library stream (the parts
 * of a composite on forked streams, as enqueue_oddeven). */
static bool enqueue_mixed(ukhip_batch* b, int refinements, int fp32_levels, bool cov) {
  if (b->parts.empty()) return run_mixed(b, refinements, fp32_levels, cov, nullptr, g_stream);
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> joins(b->parts.size(), nullptr);
  bool ok = hip_ok(hipEventCreateWithFlags(&fork, hipEventDisableTiming), "hipEventCreate") &&
            hip_ok(hipEventRecord(fork, g_stream), "hipEventRecord");
  for (size_t i = 0; ok && i < b->parts.size(); ++i) {
    ok = hip_ok(hipStreamWaitEvent(b->part_streams[i], fork, 0), "hipStreamWaitEvent") &&
         run_mixed(b->parts[i], refinements, fp32_levels, cov, nullptr, b->part_streams[i]) &&
         hip_ok(hipEventCreateWithFlags(&joins[i], hipEventDisableTiming), "hipEventCreate") &&
         hip_ok(hipEventRecord(joins[i], b->part_streams[i]), "hipEventRecord") &&
         hip_ok(hipStreamWaitEvent(g_stream, joins[i], 0), "hipStreamWaitEvent");
  }
  if (fork) (void)hipEventDestroy(fork);
  for (auto e : joins)
    if (e) (void)hipEventDestroy(e);
  return ok;
}

/* Captures enqueue_mixed once per (refinements, FP32 levels, covariances),
 * replays it afterwards. The workspaces are allocated before the capture. */
static bool graph_mixed(ukhip_batch* b, int refinements, int fp32_levels, bool cov) {
  if (!mixed_prepare(b, fp32_levels, cov)) return false; /* no allocation inside the capture */
  const int64_t key = ((int64_t)refinements << 20) | ((int64_t)(fp32_levels & 0xFFFF) << 1) | (cov ? 1 : 0);
  hipGraphExec_t& ge = b->mixed_graph[key];
  if (!ge) {
    if (!hip_ok(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"))
      return false;
    const bool ok = enqueue_mixed(b, refinements, fp32_levels, cov);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(g_stream, &g);
    if (!ok || !hip_ok(e, "hipStreamEndCapture")) {
      if (g) (void)hipGraphDestroy(g);
      b->mixed_graph.erase(key);
      return false;
    }
    const hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (!hip_ok(ei, "hipGraphInstantiate")) {
      b->mixed_graph.erase(key);
      return false;
    }
  }
  return hip_ok(hipGraphLaunch(ge, g_stream), "hipGraphLaunch");
}

extern "C" int ukhip_batch_smooth_oddeven_mixed_opt(ukhip_batch_t* b, const ukhip_mixed_options_t* opt) {
  if (!b || !opt) { uk_set_error("NULL batch or options"); return -1; }
  if (!ensure_stream()) return -1;
  if (opt->refinements < 0 || opt->refinements > 64) { uk_set_error("refinements must be in 0..64"); return -1; }
  const int fl = mixed_fp32_levels(opt);
  const bool cov = opt->covariances != 0;
  if (g_use_graphs) return graph_mixed(b, opt->refinements, fl, cov) ? 0 : -1;
  return enqueue_mixed(b, opt->refinements, fl, cov) ? 0 : -1;
}

extern "C" int ukhip_batch_smooth_oddeven_mixed(ukhip_batch_t* b, int32_t refinements) {
  const ukhip_mixed_options_t o{refinements, 0, 0};
  return ukhip_batch_smooth_oddeven_mixed_opt(b, &o);
}

extern "C" int ukhip_batch_download(ukhip_batch_t* b, int32_t p, double* states, double* covs) {
  if (!b || p < 0 || p >= b->nprob) { uk_set_error("bad batch or problem index"); return -1; }
  if (!b->parts.empty()) return ukhip_batch_download(b->parts[b->where[p].first], b->where[p].second, states, covs);
  const size_t n = b->n, nn = n * n;
  const int64_t o = b->orig_start[p], k = b->k[p];
  if (states && !hip_ok(hipMemcpyAsync(states, b->xout + o * n, k * n * 8, hipMemcpyDeviceToHost, g_stream), "download"))
    return -1;
  if (covs && !b->Sout) {
    uk_set_error("no covariances were computed for this batch (only mixed-precision estimates)");
    return -1;
  }
  if (covs && !hip_ok(hipMemcpyAsync(covs, b->Sout + o * nn, k * nn * 8, hipMemcpyDeviceToHost, g_stream), "download"))
    return -1;
  return hip_ok(hipStreamSynchronize(g_stream), "download") ? 0 : -1;
}

extern "C" int ukhip_batch_download_cov(ukhip_batch_t* b, int32_t p, char format, double* out) {
  if (!b || p < 0 || p >= b->nprob || !out) { uk_set_error("bad batch, problem index or output"); return -1; }
  if (!b->parts.empty()) return ukhip_batch_download_cov(b->parts[b->where[p].first], b->where[p].second, format, out);
  if (format == 'C') return ukhip_batch_download(b, p, nullptr, out);
  if (format != 'D' && format != 'W') { uk_set_error("covariance format '%c': use 'C', 'D' or 'W'", format); return -1; }
  if (!b->Sout) {
    uk_set_error("no covariances were computed for this batch (only mixed-precision estimates)");
    return -1;
  }
  const size_t n = b->n, nn = n * n;
  const int64_t o = b->orig_start[p], k = b->k[p];
  const size_t bytes = (size_t)k * (format == 'D' ? n : nn) * 8;
  double* d = nullptr;
  if (!hip_ok(hipMalloc((void**)&d, bytes), "hipMalloc(covariance format)")) return -1;
  bool ok = hip_ok(uk_launch_cov_format(b->Sout + o * nn, k, (int)n, format == 'D' ? 1 : 2, d, g_stream),
                   "covariance format launch") &&
            hip_ok(hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, g_stream), "download") &&
            hip_ok(hipStreamSynchronize(g_stream), "download");
  (void)hipFree(d);
  return ok ? 0 : -1;
}

/* Device pointers for sibling translation units (associative path). */
/* The covariance output is allocated here when the batch has none yet (*Sout
 * is NULL, with the error recorded, when that fails). */
extern "C" void uk_batch_view(ukhip_batch* b, const uk_raw_t** raw, double** xout, double** Sout, int64_t* steps,
                              int* nprob, const int64_t** k, void** assoc_ws, size_t* assoc_bytes) {
  *raw = &b->raw;
  *xout = b->xout;
  *Sout = ensure_elim(b) ? b->Sout : nullptr;
  *steps = b->steps;
  *nprob = b->nprob;
  *k = b->k.data();
  *assoc_ws = b->d_assoc;
  *assoc_bytes = b->assoc_bytes;
}
extern "C" void uk_batch_set_assoc_ws(ukhip_batch* b, void* p, size_t bytes) {
  b->d_assoc = p;
  b->assoc_bytes = bytes;
}
/* the level-0 chain buffers, which hold the whitened steps of the batch (B
 * and D stored in full), with the level-0 observation slots: ld mr0 and m0[s]
 * rows per step */
extern "C" void uk_batch_chain0(ukhip_batch* b, uk_chain_t* ch, const int32_t** m0) {
  *ch = b->chain[0];
  ch->mr = b->mr0;
  *m0 = b->m0;
}
/* whitens the batch into that chain (every block stored in full) */
extern "C" hipError_t uk_batch_weigh_dense0(ukhip_batch* b, hipStream_t st) {
  if (b->compress) return uk_launch_weigh_obs(&b->raw, &b->chain[0], b->wobs, st);
  return uk_launch_weigh(&b->raw, &b->chain[0], st);
}

/* ------------------------------------------------------------------------ */
/* the reference's smoother entry points (c/kalman.h:148-149)                */
/* ------------------------------------------------------------------------ */

static bool smooth_equations(kalman_options_t options, kalman_step_equations_t** equations, kalman_step_index_t length,
                             bool associative) {
  if (!ensure_stream()) return false;
  if (!equations || length < 1) {
    uk_set_error("kalman_smooth: need length >= 1 and an equations array");
    return false;
  }
  Staging st;
  if (!pack_equations(equations, length, st)) {
    staging_free(st);
    return false;
  }
  if (associative) {
    /* the associative elements need square, invertible H_i and one state
     * dimension (c/kalman_associative_smoother.c:309-505 never reads H) */
    for (kalman_step_index_t i = 0; i < length; ++i) {
      const kalman_step_equations_t* e = equations[i];
      if (e->dimension != st.n || (i > 0 && e->H->row_dim != e->H->col_dim)) {
        uk_set_error("kalman_smooth_associative: step %lld has state dimension %d of %d or a %d x %d H; the "
                     "associative smoother needs one dimension and square H (use kalman_smooth_oddeven)",
                     (long long)i, e->dimension, st.n, i > 0 ? e->H->row_dim : 0, i > 0 ? e->H->col_dim : 0);
        staging_free(st);
        return false;
      }
    }
  }
  std::vector<int64_t> lengths{(int64_t)length};
  ukhip_batch* b = batch_from_staging(st, lengths);
  staging_free(st);
  if (!b) return false;
  const int cov = associative ? 1 : ((options & KALMAN_NO_COVARIANCE) == 0);
  int rc = associative ? uk_assoc_smooth(b, options) : (run_oddeven(b, cov, nullptr, g_stream) ? 0 : -1);
  if (rc != 0) {
    batch_release(b);
    return false;
  }
  const int n = b->n;
  const bool wform = cov && (options & UKHIP_COVARIANCE_W) != 0;
  std::vector<double> xs((size_t)length * n), Ss(cov ? (size_t)length * n * n : 0);
  if (ukhip_batch_download(b, 0, xs.data(), (cov && !wform) ? Ss.data() : nullptr) != 0 ||
      (wform && ukhip_batch_download_cov(b, 0, 'W', Ss.data()) != 0)) {
    batch_release(b);
    return false;
  }
  batch_release(b);
  for (kalman_step_index_t i = 0; i < length; ++i) {
    kalman_step_equations_t* e = equations[i];
    if (associative) { /* c/kalman_associative_smoother.c:1084-1086 frees earlier results */
      matrix_free(e->state);
      matrix_free(e->covariance);
    }
    /* the step's own n_i components: the leading part of the padded state
     * (and the leading block of its covariance, or of W = L^{-1}: both
     * Cholesky factor and inverse of a lower-triangular matrix keep their
     * leading blocks) */
    const int ni = e->dimension;
    e->state = matrix_create(ni, 1);
    std::memcpy(e->state->elements, &xs[(size_t)i * n], (size_t)ni * 8);
    if (cov) {
      e->covariance = matrix_create(ni, ni);
      for (int c = 0; c < ni; ++c)
        std::memcpy(e->covariance->elements + (size_t)c * ni, &Ss[(size_t)i * n * n + (size_t)c * n], (size_t)ni * 8);
    } else {
      e->covariance = NULL;
    }
    e->covariance_type = wform ? 'W' : 'C';
  }
  return true;
}

extern "C" void kalman_smooth_oddeven(kalman_options_t options, kalman_step_equations_t** equations,
                                      kalman_step_index_t length) {
  smooth_equations(options, equations, length, false);
}

extern "C" void kalman_smooth_associative(kalman_options_t options, kalman_step_equations_t** equations,
                                          kalman_step_index_t length) {
  smooth_equations(options, equations, length, true);
}

/* ------------------------------------------------------------------------ */
/* serving: continuous batching of smoothing requests                        */
/* ------------------------------------------------------------------------ */

namespace {

using Clock = std::chrono::steady_clock;

/* The pinned download buffers of one batch. They stay with the batch's
 * requests until each has been waited for (the results are copied out of
 * them once, by ukhip_server_wait). */
struct ResultBuf {
  double* x = nullptr;
  double* S = nullptr;
  size_t xb = 0, Sb = 0;
  int refs = 0;
};

/* One request: its own copy of the inputs (packed as ukhip_problem_t), where
 * its results are, and timestamps. */
struct Request {
  int64_t id = 0;
  ukhip_problem_t prob{};
  char* in = nullptr; /* the inputs prob points into: one block from the server's input pool */
  size_t in_bytes = 0;
  ResultBuf* res = nullptr;
  const double* rx = nullptr; /* k*n estimates in res */
  const double* rS = nullptr; /* k*n*n covariances in res, or NULL */
  Clock::time_point t_submit, t_done;
  int status = 0; /* 0 queued, 1 running, 2 done, -1 failed */
  std::string error;
};

/* A batch on the device: its requests, stream slot, staging and results. */
struct InFlight {
  ukhip_batch* b = nullptr;
  Staging st;
  std::vector<Request*> reqs;
  int slot = 0;
  hipEvent_t done = nullptr;
  double* host_x = nullptr; /* pinned: steps * n */
  double* host_S = nullptr; /* pinned: steps * n * n (FP64 mode) */
  size_t bytes = 0;
  int64_t b_steps = 0;
  Clock::time_point t_start;
};

/* Device bytes of a batch of requests of state dimension n (the layout of
 * batch_from_staging and, in mixed mode, mixed_workspace), with slack. */
size_t batch_device_bytes(int n, int mr, int64_t steps, bool mixed) {
  const double nn = (double)n * n;
  double per = 8.0 * (14.0 * nn + 10.0 * n + 4.0 * mr + mr * (double)n + (double)mr * mr);
  if (mr > n) per += 8.0 * mr * (n + 1.0) + 4.0; /* whitened observations staged for k_obs_compress, m0 */
  if (mixed) per += 4.0 * (7.5 * nn + 4 * n) + 8.0 * (6.0 * n) + 8.0 * (0.8 * nn + 0.5 * n);
  return (size_t)(per * (double)steps * 1.05) + (size_t)(4u << 20);
}

}  // namespace

struct ukhip_server {
  ukhip_server_options_t opt{};
  DevArena arena;
  PinnedPool pinned;
  PinnedPool inputs; /* pageable: requests' copies of their inputs */
  std::vector<hipStream_t> streams;
  std::vector<bool> slot_busy;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;
  std::map<int64_t, Request*> all;
  std::vector<InFlight*> flight;
  std::atomic<bool> stop{false};
  std::thread worker;
  int64_t next_id = 0;
  size_t in_flight_bytes = 0;
  bool log = false; /* UKHIP_SERVER_LOG set: one stderr line per batch */
  Clock::time_point t0 = Clock::now();
};

static double ms_since(Clock::time_point a, Clock::time_point b = Clock::now()) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

/* Drops one request's reference to its batch's result buffers (lock held). */
static void server_unref(ukhip_server* s, ResultBuf* rb) {
  if (!rb || --rb->refs > 0) return;
  s->pinned.put(rb->x, rb->xb);
  s->pinned.put(rb->S, rb->Sb);
  delete rb;
}

/* Returns a request's input block to the pool (its batch has been packed). */
static void release_inputs(ukhip_server* s, Request* r) {
  if (r->in) s->inputs.put(r->in, r->in_bytes);
  r->in = nullptr;
}

static void server_fail(ukhip_server* s, Request* r, const std::string& why) {
  release_inputs(s, r);
  r->status = -1;
  r->error = why;
  r->t_done = Clock::now();
}

/* Starts one batch from the queue. Called with the lock held; the lock is
 * released while the batch is packed, uploaded and launched (only this
 * scheduler thread allocates from the arena and starts batches, and the
 * stream slot is reserved first), so submissions and waits do not stall
 * behind it. Returns false when nothing could be admitted. */
static bool server_admit(ukhip_server* s, std::unique_lock<std::mutex>& lk) {
  if (s->queue.empty()) return false;
  int slot = -1;
  for (size_t i = 0; i < s->slot_busy.size(); ++i)
    if (!s->slot_busy[i]) { slot = (int)i; break; }
  if (slot < 0) return false;
  const bool mixed = s->opt.mixed_refinements >= 0;
  /* the oldest request fixes the state dimension; then every queued request
   * of that dimension in arrival order, while the batch fits the budget */
  const int n = s->queue.front()->prob.n;
  std::vector<Request*> take;
  int64_t steps = 0;
  int mr = 1;
  for (auto it = s->queue.begin(); it != s->queue.end() && (int)take.size() < s->opt.max_batch;) {
    Request* r = *it;
    if (r->prob.n != n) { ++it; continue; }
    const int mr2 = std::max(mr, r->prob.m);
    const size_t need = batch_device_bytes(n, mr2, steps + r->prob.k, mixed);
    if (s->in_flight_bytes + need > s->arena.size) {
      if (take.empty() && s->flight.empty()) { /* alone and still too large: reject it */
        server_fail(s, r, "request exceeds the device memory budget of the server");
        it = s->queue.erase(it);
        s->cv_done.notify_all();
        continue;
      }
      break;
    }
    take.push_back(r);
    steps += r->prob.k;
    mr = mr2;
    it = s->queue.erase(it);
  }
  if (take.empty()) return false;
  InFlight* f = new InFlight();
  f->reqs = take;
  f->slot = slot;
  hipStream_t st = s->streams[slot];
  std::vector<ukhip_problem_t> probs;
  for (auto* r : take) {
    probs.push_back(r->prob);
    r->status = 1;
  }
  s->slot_busy[slot] = true;
  const size_t queued = s->queue.size();
  lk.unlock();
  const Clock::time_point ta = Clock::now();
  std::vector<int64_t> lengths;
  f->st.pool = &s->pinned;
  bool ok = pack_problems((int32_t)probs.size(), probs.data(), f->st, lengths);
  const Clock::time_point tp = Clock::now();
  if (ok) {
    f->bytes = batch_device_bytes(n, f->st.mr, f->st.steps, mixed);
    f->b = batch_from_staging(f->st, lengths, st, &s->arena, false);
    ok = f->b != nullptr;
  }
  if (ok) ok = mixed ? run_mixed(f->b, s->opt.mixed_refinements, MIXED_FP32_LEVELS, false, nullptr, st) : run_oddeven(f->b, 1, nullptr, st);
  const size_t S = (size_t)f->st.steps, nn = (size_t)n * n;
  if (ok) ok = (f->host_x = (double*)s->pinned.get(S * n * 8)) != nullptr;
  if (ok && !mixed) ok = (f->host_S = (double*)s->pinned.get(S * nn * 8)) != nullptr;
  if (ok) ok = hip_ok(hipMemcpyAsync(f->host_x, f->b->xout, S * n * 8, hipMemcpyDeviceToHost, st), "download");
  if (ok && !mixed)
    ok = hip_ok(hipMemcpyAsync(f->host_S, f->b->Sout, S * nn * 8, hipMemcpyDeviceToHost, st), "download");
  if (ok) ok = hip_ok(hipEventCreateWithFlags(&f->done, hipEventDisableTiming), "hipEventCreate") &&
               hip_ok(hipEventRecord(f->done, st), "hipEventRecord");
  if (!ok) {
    const std::string why = ukhip_last_error();
    const bool arena_full = why.find("device arena") != std::string::npos;
    (void)hipStreamSynchronize(st);
    if (f->b) batch_release(f->b);
    staging_free(f->st);
    s->pinned.put(f->host_x, S * n * 8);
    s->pinned.put(f->host_S, S * nn * 8);
    if (f->done) (void)hipEventDestroy(f->done);
    delete f;
    lk.lock();
    s->slot_busy[slot] = false;
    if (arena_full && !s->flight.empty()) { /* fragmented arena: retry after a batch retires */
      for (auto it = take.rbegin(); it != take.rend(); ++it) {
        (*it)->status = 0;
        s->queue.push_front(*it);
      }
      return false;
    }
    for (auto* r : take) server_fail(s, r, why);
    s->cv_done.notify_all();
    return true;
  }
  f->b_steps = f->st.steps;
  if (s->log)
    std::fprintf(stderr, "[ukhip server] %.3f ms: batch n=%d requests=%zu steps=%lld pack %.3f ms enqueue %.3f ms, "
                 "%zu still queued, slot %d\n", ms_since(s->t0, ta), n, take.size(), (long long)f->b_steps,
                 ms_since(ta, tp), ms_since(tp), queued, slot);
  f->t_start = ta;
  for (auto* r : take) release_inputs(s, r); /* packed and uploading from staging */
  lk.lock();
  s->in_flight_bytes += f->bytes;
  s->flight.push_back(f);
  return true;
}

/* Retires finished batches (lock held by the caller). */
static bool server_reap(ukhip_server* s) {
  bool any = false;
  for (size_t i = 0; i < s->flight.size();) {
    InFlight* f = s->flight[i];
    const hipError_t q = hipEventQuery(f->done);
    if (q == hipErrorNotReady) { ++i; continue; }
    const bool ok = q == hipSuccess;
    const int n = f->b->n;
    const size_t nn = (size_t)n * n;
    const size_t S = (size_t)f->b_steps;
    ResultBuf* rb = nullptr;
    if (ok) {
      rb = new ResultBuf();
      rb->x = f->host_x;
      rb->S = f->host_S;
      rb->xb = S * n * 8;
      rb->Sb = S * nn * 8;
      rb->refs = (int)f->reqs.size();
    }
    for (size_t p = 0; p < f->reqs.size(); ++p) {
      Request* r = f->reqs[p];
      if (!ok) { server_fail(s, r, hipGetErrorString(q)); continue; }
      const int64_t o = f->b->orig_start[p];
      r->res = rb;
      r->rx = f->host_x + o * n;
      r->rS = f->host_S ? f->host_S + o * nn : nullptr;
      r->t_done = Clock::now();
      r->status = 2;
    }
    if (s->log)
      std::fprintf(stderr, "[ukhip server] %.3f ms: slot %d done, %lld steps in %.3f ms\n", ms_since(s->t0), f->slot,
                   (long long)f->b_steps, ms_since(f->t_start));
    batch_release(f->b);
    staging_free(f->st);
    if (!ok) {
      s->pinned.put(f->host_x, S * n * 8);
      s->pinned.put(f->host_S, S * nn * 8);
    }
    (void)hipEventDestroy(f->done);
    s->slot_busy[f->slot] = false;
    s->in_flight_bytes -= f->bytes;
    s->flight.erase(s->flight.begin() + i);
    delete f;
    any = true;
  }
  return any;
}

static void server_loop(ukhip_server* s) {
  std::unique_lock<std::mutex> lk(s->mu);
  while (true) {
    const bool reaped = server_reap(s);
    if (reaped) s->cv_done.notify_all();
    while (server_admit(s, lk)) {
    }
    if (s->stop && s->queue.empty() && s->flight.empty()) break;
    /* in flight: poll the events at a short interval; idle: sleep until a submission */
    if (!s->flight.empty()) s->cv_work.wait_for(lk, std::chrono::microseconds(50));
    else s->cv_work.wait_for(lk, std::chrono::milliseconds(20));
  }
}

extern "C" ukhip_server_t* ukhip_server_create(const ukhip_server_options_t* options) {
  if (!ensure_stream()) return nullptr;
  ukhip_server* s = new ukhip_server();
  if (options) s->opt = *options;
  if (s->opt.max_batch <= 0) s->opt.max_batch = 256;
  if (s->opt.max_inflight <= 0) s->opt.max_inflight = 4;
  size_t budget = s->opt.hbm_budget_bytes;
  if (budget == 0) {
    size_t fr = 0, tot = 0;
    if (!hip_ok(hipMemGetInfo(&fr, &tot), "hipMemGetInfo")) { delete s; return nullptr; }
    budget = fr / 5 * 2;
  }
  if (!hip_ok(hipMalloc((void**)&s->arena.base, budget), "hipMalloc(server arena)")) { delete s; return nullptr; }
  s->arena.size = budget;
  s->arena.free_blocks[0] = budget;
  for (int i = 0; i < s->opt.max_inflight; ++i) {
    hipStream_t st = nullptr;
    if (!hip_ok(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate")) break;
    s->streams.push_back(st);
  }
  if ((int)s->streams.size() != s->opt.max_inflight) {
    for (auto st : s->streams) (void)hipStreamDestroy(st);
    (void)hipFree(s->arena.base);
    delete s;
    return nullptr;
  }
  s->slot_busy.assign(s->streams.size(), false);
  s->inputs.pageable = true;
  s->log = std::getenv("UKHIP_SERVER_LOG") != nullptr;
  s->worker = std::thread(server_loop, s);
  return s;
}

extern "C" int64_t ukhip_server_submit(ukhip_server_t* s, const ukhip_problem_t* p) {
  if (!s || !p) { uk_set_error("NULL server or problem"); return -1; }
  /* the whole request is checked here, so that a malformed one never joins
   * (and fails) a batch of valid requests */
  if (!problem_ok(*p, 0)) {
    const std::string why = ukhip_last_error();
    uk_set_error("ukhip_server_submit: %s", why.c_str());
    return -1;
  }
  /* the request's own copy of its inputs, in one block reused from the
   * server's pool (once warm, the copy runs without page faults) */
  const size_t n = p->n, m = p->m, k = p->k;
  const void* src[8] = {p->F, p->H, p->c, p->K, p->G, p->o, p->C, p->m_i};
  const size_t bytes[8] = {(p->F_shared ? 1 : k) * n * n * 8, k * n * n * 8, k * n * 8,
                           k * (p->K_type == 'w' ? n : n * n) * 8, k * m * n * 8, k * m * 8,
                           k * (p->C_type == 'w' ? m : m * m) * 8, k * 4};
  size_t off[8], total = 0;
  for (int f = 0; f < 8; ++f) {
    off[f] = total;
    if (src[f]) total += (bytes[f] + 63) & ~(size_t)63;
  }
  Request* r = new Request();
  r->in_bytes = std::max<size_t>(total, 64);
  r->in = (char*)s->inputs.get(r->in_bytes);
  if (!r->in) {
    delete r;
    return -1;
  }
  const void* dst[8];
  for (int f = 0; f < 8; ++f) {
    dst[f] = src[f] ? r->in + off[f] : nullptr;
    if (src[f]) std::memcpy(r->in + off[f], src[f], bytes[f]);
  }
  r->prob = *p;
  r->prob.F = (const double*)dst[0];
  r->prob.H = (const double*)dst[1];
  r->prob.c = (const double*)dst[2];
  r->prob.K = (const double*)dst[3];
  r->prob.G = (const double*)dst[4];
  r->prob.o = (const double*)dst[5];
  r->prob.C = (const double*)dst[6];
  r->prob.m_i = (const int32_t*)dst[7];
  {
    std::lock_guard<std::mutex> g(s->mu);
    r->id = s->next_id++;
    r->t_submit = Clock::now();
    s->all[r->id] = r;
    s->queue.push_back(r);
  }
  s->cv_work.notify_one();
  return r->id;
}

extern "C" int ukhip_server_poll(ukhip_server_t* s, int64_t id) {
  if (!s) return -1;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) return -1;
  return it->second->status == 2 ? 1 : (it->second->status < 0 ? -1 : 0);
}

extern "C" int ukhip_server_wait(ukhip_server_t* s, int64_t id, double* states, double* covs, double* latency_ms) {
  if (!s) { uk_set_error("NULL server"); return -1; }
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) { uk_set_error("unknown request %lld", (long long)id); return -1; }
  Request* r = it->second;
  s->cv_done.wait(lk, [&] { return r->status == 2 || r->status < 0; });
  s->all.erase(it);
  lk.unlock();
  int rc = 0;
  if (r->status < 0) {
    uk_set_error("request %lld: %s", (long long)id, r->error.c_str());
    rc = -1;
  } else { /* r->res stays allocated while this request holds a reference */
    const size_t kn = (size_t)r->prob.k * r->prob.n, knn = kn * r->prob.n;
    if (states) std::memcpy(states, r->rx, kn * 8);
    if (covs) {
      if (!r->rS) {
        for (size_t i = 0; i < knn; ++i) covs[i] = kalman_nan;
      } else {
        std::memcpy(covs, r->rS, knn * 8);
      }
    }
    lk.lock();
    server_unref(s, r->res);
    lk.unlock();
  }
  if (latency_ms) latency_ms[0] = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  delete r;
  return rc;
}

extern "C" void ukhip_server_destroy(ukhip_server_t* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (auto& kv : s->all) {
    server_unref(s, kv.second->res);
    release_inputs(s, kv.second);
    delete kv.second;
  }
  for (auto st : s->streams) (void)hipStreamDestroy(st);
  (void)hipFree(s->arena.base);
  s->pinned.clear();
  s->inputs.clear();
  delete s;
}

/* ------------------------------------------------------------------------ */
/* batched sequential filter                                                  */
/* ------------------------------------------------------------------------ */

struct ukhip_filter {
  int nf = 0, n = 0, mr = 0;
  int64_t steps = 0;
  uk_filter_t st{};
  void* d_state = nullptr;
  char* d_in = nullptr;
  char* h_in = nullptr; /* pinned staging of one step's inputs */
  size_t in_bytes = 0;
  /* full-representation K or C: the step whitened by k_weigh into wch, with
   * per-filter tables for its uk_raw_t view (F block indices own / shared,
   * first-step flags 1 / 0, m_i = mr, a zero vector) */
  void* d_aux = nullptr;
  const int64_t *fidx_own = nullptr, *fidx_zero = nullptr;
  const int8_t *first1 = nullptr, *first0 = nullptr;
  const int32_t* mfull = nullptr;
  const double* zeros = nullptr;
  void* d_wch = nullptr;
  uk_chain_t wch{};
  /* history (ukhip_filter_keep_history) and the smoother's outputs */
  void* d_hist = nullptr;
  int64_t hist_cap = 0; /* slots: steps 0 .. hist_cap - 1 */
  void* d_smooth = nullptr;
  int64_t smooth_cap = 0;
};

extern "C" ukhip_filter_t* ukhip_filter_create(int32_t nfilters, int32_t n, int32_t m) {
  if (!ensure_stream()) return nullptr;
  if (nfilters < 1 || n < 1 || n > UKHIP_MAX_N || m < 0 || m > n) {
    uk_set_error("ukhip_filter_create: need nfilters >= 1, 1 <= n <= %d, 0 <= m <= n", UKHIP_MAX_N);
    return nullptr;
  }
  ukhip_filter* f = new ukhip_filter();
  f->nf = nfilters;
  f->n = n;
  f->mr = std::max(m, 1);
  const size_t P = (size_t)nfilters, nn = (size_t)n * n, mr = (size_t)f->mr;
  const size_t state_bytes = P * (nn + 2 * (size_t)n) * 8 + P * 4;
  /* one step's inputs: F, H, c, K (n*n at most), G, o, C (m*m at most), m_i,
   * the K and C type codes; each array 256-byte aligned */
  f->in_bytes = P * (3 * nn + (size_t)n + mr * n + mr + mr * mr) * 8 + P * 6 + 12 * 256;
  bool ok = hip_ok(hipMalloc(&f->d_state, state_bytes), "hipMalloc(filter state)") &&
            hip_ok(hipMemsetAsync(f->d_state, 0, state_bytes, g_stream), "hipMemset(filter state)") &&
            hip_ok(hipMalloc((void**)&f->d_in, f->in_bytes), "hipMalloc(filter inputs)") &&
            hip_ok(hipHostMalloc((void**)&f->h_in, f->in_bytes, hipHostMallocDefault), "hipHostMalloc(filter inputs)");
  if (!ok) {
    ukhip_filter_free(f);
    return nullptr;
  }
  char* d = (char*)f->d_state;
  f->st.R = (double*)d;
  f->st.y = (double*)(d + P * nn * 8);
  f->st.x = (double*)(d + P * (nn + n) * 8);
  f->st.z = (int32_t*)(d + P * (nn + 2 * (size_t)n) * 8);
  f->st.nf = nfilters;
  return f;
}

/* The tables and whitened-step buffers of the full-representation path,
 * allocated on its first use. */
static bool filter_aux(ukhip_filter* f) {
  if (f->d_aux) return true;
  const size_t P = (size_t)f->nf, n = (size_t)f->n, nn = n * n;
  std::vector<char> h;
  auto put = [&](const void* src, size_t bytes) {
    const size_t at = h.size();
    h.resize(at + ((bytes + 255) & ~(size_t)255), 0);
    std::memcpy(h.data() + at, src, bytes);
    return at;
  };
  std::vector<int64_t> own(P), zero(P, 0);
  for (size_t p = 0; p < P; ++p) own[p] = (int64_t)p;
  std::vector<int8_t> one8(P, 1), zero8(P, 0);
  std::vector<int32_t> mf(P, f->mr);
  std::vector<double> z(P * n, 0.0);
  const size_t o_own = put(own.data(), P * 8), o_zero = put(zero.data(), P * 8), o_f1 = put(one8.data(), P),
               o_f0 = put(zero8.data(), P), o_m = put(mf.data(), P * 4), o_z = put(z.data(), P * n * 8);
  if (!hip_ok(hipMalloc(&f->d_aux, h.size()), "hipMalloc(filter tables)")) return false;
  if (!hip_ok(hipMemcpy(f->d_aux, h.data(), h.size(), hipMemcpyHostToDevice), "upload filter tables")) {
    (void)hipFree(f->d_aux);
    f->d_aux = nullptr;
    return false;
  }
  char* a = (char*)f->d_aux;
  f->fidx_own = (const int64_t*)(a + o_own);
  f->fidx_zero = (const int64_t*)(a + o_zero);
  f->first1 = (const int8_t*)(a + o_f1);
  f->first0 = (const int8_t*)(a + o_f0);
  f->mfull = (const int32_t*)(a + o_m);
  f->zeros = (const double*)(a + o_z);
  const size_t wbytes = P * (3 * nn + 2 * n) * 8;
  if (!hip_ok(hipMalloc(&f->d_wch, wbytes), "hipMalloc(filter whitened step)")) return false;
  double* w = (double*)f->d_wch;
  f->wch.B = w;
  f->wch.D = w + P * nn;
  f->wch.C = w + 2 * P * nn;
  f->wch.c = w + 3 * P * nn;
  f->wch.o = w + 3 * P * nn + P * n;
  f->wch.mr = f->mr;
  return true;
}

static bool full_rep(char t) { return t == 'W' || t == 'C' || t == 'U' || t == 'F'; }

/* One filter step; dev: the arrays of `in` are device pointers, read in place. */
static int filter_step(ukhip_filter_t* f, const ukhip_filter_inputs_t* in, bool dev) {
  if (!f || !in) { uk_set_error("NULL filter or inputs"); return -1; }
  const bool first = f->steps == 0;
  const size_t P = (size_t)f->nf, n = (size_t)f->n, nn = n * n, mr = (size_t)f->mr;
  const char kt = in->K_type ? in->K_type : 'w', ct = in->C_type ? in->C_type : 'w';
  if ((kt != 'w' && !full_rep(kt)) || (ct != 'w' && !full_rep(ct))) {
    uk_set_error("ukhip_filter_step: K_type and C_type must be 'w', 'W', 'C', 'U' or 'F'");
    return -1;
  }
  const bool full = full_rep(kt) || full_rep(ct);
  if (f->d_hist && !first && f->steps - 1 >= f->hist_cap) {
    uk_set_error("ukhip_filter_step: the history holds %lld steps", (long long)(f->hist_cap + 1));
    return -1;
  }
  if (!in->K && !first) { uk_set_error("ukhip_filter_step: K is required after the first step"); return -1; }
  if (!first && !in->F) { uk_set_error("ukhip_filter_step: F is required after the first step"); return -1; }
  bool observed = in->m_i == nullptr || dev; /* device counts are clamped by the kernel, not read here */
  for (size_t p = 0; !dev && in->m_i && p < P; ++p) {
    if (in->m_i[p] < 0 || in->m_i[p] > f->mr) { uk_set_error("ukhip_filter_step: m_i outside 0..m"); return -1; }
    observed = observed || in->m_i[p] > 0;
  }
  if (observed && (!in->G || !in->o || !in->C)) { uk_set_error("ukhip_filter_step: G, o and C are required"); return -1; }
  /* the previous step's upload reads the staging buffer (device inputs use it
   * only for the representation codes of a whitened step) */
  if ((!dev || full) && !hip_ok(hipStreamSynchronize(g_stream), "ukhip_filter_step")) return -1;
  size_t off = 0;
  uk_filter_in_t din{};
  auto stage = [&](const void* src, size_t bytes) -> const void* {
    if (!src) return nullptr;
    std::memcpy(f->h_in + off, src, bytes);
    const void* dp = f->d_in + off;
    off += (bytes + 255) & ~(size_t)255;
    return dp;
  };
  auto put = [&](const void* src, size_t bytes) -> const void* { return dev ? src : stage(src, bytes); };
  const size_t mcols = (size_t)f->mr; /* rows of G, o, C per filter as given: ld m */
  din.F = (const double*)put(first ? nullptr : in->F, (in->F_shared ? 1 : P) * nn * 8);
  din.F_shared = in->F_shared;
  din.H = (const double*)put(first ? nullptr : in->H, P * nn * 8);
  din.c = (const double*)put(first ? nullptr : in->c, P * n * 8);
  din.K = (const double*)put(in->K, P * (full_rep(kt) ? nn : n) * 8);
  din.G = (const double*)put(in->G, P * mcols * n * 8);
  din.o = (const double*)put(in->o, P * mcols * 8);
  din.C = (const double*)put(in->C, P * mcols * (full_rep(ct) ? mcols : 1) * 8);
  din.m_i = (const int32_t*)put(in->m_i, P * 4);
  din.mr = (int32_t)mr;
  std::vector<int8_t> codes;
  const int8_t* dKt = nullptr;
  const int8_t* dCt = nullptr;
  if (full) {
    codes.assign(2 * P, 0);
    std::fill(codes.begin(), codes.begin() + P, (int8_t)cov_code(kt));
    std::fill(codes.begin() + P, codes.end(), (int8_t)cov_code(ct));
    dKt = (const int8_t*)stage(codes.data(), P);
    dCt = (const int8_t*)stage(codes.data() + P, P);
  }
  if (!din.K) din.K = f->st.y; /* first step: K is never read */
  if (off > f->in_bytes) { uk_set_error("ukhip_filter_step: staging overflow"); return -1; }
  if (off > 0 && !hip_ok(hipMemcpyAsync(f->d_in, f->h_in, off, hipMemcpyHostToDevice, g_stream), "filter upload"))
    return -1;
  if (full) { /* whiten the step on the device (k_weigh), then the filter step reads it */
    if (!filter_aux(f)) return -1;
    uk_raw_t raw{};
    raw.n = f->n;
    raw.mr = f->mr;
    raw.steps = f->nf;
    raw.H_ident = din.H == nullptr;
    raw.K_full = full_rep(kt);
    raw.C_full = full_rep(ct);
    raw.F = din.F ? din.F : f->zeros;
    raw.Fidx = in->F_shared ? f->fidx_zero : f->fidx_own;
    raw.H = din.H;
    raw.c = din.c ? din.c : f->zeros;
    raw.K = din.K;
    raw.G = din.G ? din.G : f->zeros;
    raw.o = din.o ? din.o : f->zeros;
    raw.C = din.C ? din.C : f->zeros;
    raw.Kt = dKt;
    raw.Ct = dCt;
    raw.m = din.m_i ? din.m_i : f->mfull;
    raw.first = first ? f->first1 : f->first0;
    if (!hip_ok(uk_launch_weigh(&raw, &f->wch, g_stream), "filter whitening launch")) return -1;
    din.pre = 1;
    din.wB = f->wch.B;
    din.wD = f->wch.D;
    din.wc = f->wch.c;
    din.wC = f->wch.C;
    din.wo = f->wch.o;
  }
  f->st.hist = (f->d_hist && !first) ? 1 : 0;
  f->st.hslot = first ? 0 : f->steps - 1;
  if (!hip_ok(uk_launch_filter_step(&f->st, &din, f->nf, f->n, first ? 1 : 0, g_stream), "filter step launch"))
    return -1;
  ++f->steps;
  return 0;
}

extern "C" int ukhip_filter_step(ukhip_filter_t* f, const ukhip_filter_inputs_t* in) {
  return filter_step(f, in, false);
}

extern "C" int ukhip_filter_step_device(ukhip_filter_t* f, const ukhip_filter_inputs_t* in) {
  return filter_step(f, in, true);
}

extern "C" int ukhip_filter_keep_history(ukhip_filter_t* f, int64_t max_steps) {
  if (!f) { uk_set_error("NULL filter"); return -1; }
  if (f->steps > 0) { uk_set_error("ukhip_filter_keep_history: call it before the first step"); return -1; }
  if (max_steps < 1) { uk_set_error("ukhip_filter_keep_history: max_steps must be >= 1"); return -1; }
  if (f->d_hist) (void)hipFree(f->d_hist);
  f->d_hist = nullptr;
  f->hist_cap = 0;
  const size_t P = (size_t)f->nf, n = (size_t)f->n, nn = n * n, slots = (size_t)std::max<int64_t>(max_steps - 1, 1);
  if (!hip_ok(hipMalloc(&f->d_hist, slots * P * (2 * nn + n) * 8), "hipMalloc(filter history)")) return -1;
  double* h = (double*)f->d_hist;
  f->st.hRd = h;
  f->st.hRs = h + slots * P * nn;
  f->st.hy = h + 2 * slots * P * nn;
  f->hist_cap = max_steps - 1;
  return 0;
}

extern "C" int ukhip_filter_smooth(ukhip_filter_t* f, double* states, double* W) {
  if (!f) { uk_set_error("NULL filter"); return -1; }
  if (!f->d_hist) { uk_set_error("ukhip_filter_smooth: needs ukhip_filter_keep_history before the first step"); return -1; }
  if (f->steps < 1) { uk_set_error("ukhip_filter_smooth: no steps yet"); return -1; }
  const size_t P = (size_t)f->nf, n = (size_t)f->n, nn = n * n, K = (size_t)f->steps;
  if (f->smooth_cap < f->steps) {
    if (f->d_smooth) (void)hipFree(f->d_smooth);
    f->d_smooth = nullptr;
    f->smooth_cap = 0;
    if (!hip_ok(hipMalloc(&f->d_smooth, K * P * (n + nn) * 8), "hipMalloc(filter smoother outputs)")) return -1;
    f->smooth_cap = f->steps;
  }
  double* xs = (double*)f->d_smooth;
  double* Ws = xs + (size_t)f->smooth_cap * P * n;
  if (!hip_ok(uk_launch_filter_smooth(&f->st, f->n, f->steps, xs, Ws, g_stream), "filter smoother launch")) return -1;
  if (states && !hip_ok(hipMemcpyAsync(states, xs, K * P * n * 8, hipMemcpyDeviceToHost, g_stream), "download"))
    return -1;
  if (W && !hip_ok(hipMemcpyAsync(W, Ws, K * P * nn * 8, hipMemcpyDeviceToHost, g_stream), "download")) return -1;
  return hip_ok(hipStreamSynchronize(g_stream), "filter smoother") ? 0 : -1;
}

extern "C" int ukhip_filter_estimates(ukhip_filter_t* f, double* states, double* R) {
  if (!f) { uk_set_error("NULL filter"); return -1; }
  const size_t P = (size_t)f->nf, n = (size_t)f->n;
  if (states && !hip_ok(hipMemcpyAsync(states, f->st.x, P * n * 8, hipMemcpyDeviceToHost, g_stream), "download"))
    return -1;
  if (R && !hip_ok(hipMemcpyAsync(R, f->st.R, P * n * n * 8, hipMemcpyDeviceToHost, g_stream), "download"))
    return -1;
  return hip_ok(hipStreamSynchronize(g_stream), "download") ? 0 : -1;
}

extern "C" void ukhip_filter_free(ukhip_filter_t* f) {
  if (!f) return;
  (void)hipStreamSynchronize(g_stream);
  if (f->d_state) (void)hipFree(f->d_state);
  if (f->d_in) (void)hipFree(f->d_in);
  if (f->h_in) (void)hipHostFree(f->h_in);
  if (f->d_aux) (void)hipFree(f->d_aux);
  if (f->d_wch) (void)hipFree(f->d_wch);
  if (f->d_hist) (void)hipFree(f->d_hist);
  if (f->d_smooth) (void)hipFree(f->d_smooth);
  delete f;
}

/* ------------------------------------------------------------------------ */
/* kalman_t ingest: steps appended straight into structure-of-arrays fields   */
/* ------------------------------------------------------------------------ */

/* The steps of one kalman_t (c/kalman_explicit_representation.c:164-252) as
 * the device layout's fields, slot after slot (slot j of a field at j*stride):
 * kalman_evolve / kalman_observe write a step's matrices into its slot once,
 * kalman_smooth uploads every field's live range as it is (no per-matrix
 * allocations, no repacking) and keeps the results compactly. Live steps are
 * slots [start, start + count); slot start + count is the step between
 * kalman_evolve and kalman_observe. A field's stride changes only when a wider
 * representation first appears (a non-'w' K or C, more observation rows, a
 * non-identity H); then that field is re-laid once. */
struct uk_ingest {
  int n = -1, mr = 1;
  int64_t start = 0, count = 0, step0 = 0;
  bool current = false;
  /* General mode: once a step has another state dimension, an evolution
   * equation of other than n rows, or more observation rows than states, the
   * live steps (and the pending one, at the back) are kept as deep copies of
   * their equations, as the reference's explicit representation keeps them
   * (c/kalman_explicit_representation.c:164-244), and kalman_smooth hands
   * them to the packing of kalman_smooth_oddeven, which pads every step to
   * one dimension. eqs[q] is step step0 + q. */
  bool general = false;
  std::deque<kalman_step_equations_t*> eqs;
  bool H_any = false, K_full = false, C_full = false;
  std::vector<double> F, H, c, K, G, o, C, x, S;
  std::vector<int8_t> Kt, Ct, hasx, hasS;
  std::vector<int32_t> m;
  char Stype = 'C';
  int64_t cap = 0;

  size_t nn() const { return (size_t)n * n; }
  size_t Kw() const { return K_full ? nn() : (size_t)n; }
  size_t Cw() const { return C_full ? (size_t)mr * mr : (size_t)mr; }
  void grow(int64_t slots) {
    if (slots <= cap) return;
    int64_t nc = std::max<int64_t>(1024, cap);
    while (nc < slots) nc *= 2;
    if (start > 0 && count + 1 <= nc / 2) compact();
    if (start + count + 1 <= cap) return;
    cap = nc;
    F.resize(cap * nn());
    if (H_any) H.resize(cap * nn());
    c.resize(cap * n);
    K.resize(cap * Kw());
    G.resize(cap * (size_t)mr * n);
    o.resize(cap * (size_t)mr);
    C.resize(cap * Cw());
    x.resize(cap * n);
    S.resize(cap * nn());
    Kt.resize(cap);
    Ct.resize(cap);
    hasx.resize(cap);
    hasS.resize(cap);
    m.resize(cap);
  }
  /* moves the live slots (and the pending one) to slot 0 */
  void compact() {
    const int64_t k = count + (current ? 1 : 0);
    auto mv = [&](auto& v, size_t stride) {
      if (!v.empty()) std::memmove(v.data(), v.data() + start * stride, k * stride * sizeof(v[0]));
    };
    mv(F, nn());
    if (H_any) mv(H, nn());
    mv(c, n); mv(K, Kw()); mv(G, (size_t)mr * n); mv(o, mr); mv(C, Cw()); mv(x, n); mv(S, nn());
    mv(Kt, 1); mv(Ct, 1); mv(hasx, 1); mv(hasS, 1); mv(m, 1);
    start = 0;
  }
  ~uk_ingest() {
    for (auto* e : eqs) eq_free(e);
  }
  static kalman_matrix_t* mat_from(const double* src, int rows, int cols, int ld) {
    kalman_matrix_t* A = matrix_create(rows, cols);
    for (int c = 0; c < cols; ++c) std::memcpy(A->elements + (size_t)c * rows, src + (size_t)c * ld, (size_t)rows * 8);
    return A;
  }
  static void eq_clear_observation(kalman_step_equations_t* e) {
    matrix_free(e->G);
    matrix_free(e->o);
    matrix_free(e->C);
    matrix_free(e->state);
    matrix_free(e->covariance);
    e->G = e->o = e->C = e->state = e->covariance = nullptr;
  }
  static void eq_free(kalman_step_equations_t* e) {
    if (!e) return;
    eq_clear_observation(e);
    matrix_free(e->H);
    matrix_free(e->F);
    matrix_free(e->c);
    matrix_free(e->K);
    std::free(e);
  }
  static char cov_char(int code) {
    return code == UK_COV_W ? 'W' : code == UK_COV_w ? 'w' : code == UK_COV_C ? 'C' : 'F';
  }
  /* moves every live slot (and the pending one) into eqs */
  void to_general() {
    if (general) return;
    const int64_t total = count + (current ? 1 : 0);
    const size_t nn_ = nn();
    for (int64_t q = 0; q < total; ++q) {
      const int64_t j = start + q;
      auto* e = (kalman_step_equations_t*)std::calloc(1, sizeof(kalman_step_equations_t));
      e->step = (kalman_step_index_t)(step0 + q);
      e->dimension = n;
      e->covariance_type = 'C';
      if (q > 0) { /* the earliest live step's evolution equation is never used */
        e->F = mat_from(&F[j * nn_], n, n, n);
        if (H_any) {
          e->H = mat_from(&H[j * nn_], n, n, n);
        } else {
          e->H = matrix_create(n, n);
          for (size_t t = 0; t < nn_; ++t) e->H->elements[t] = (t % (n + 1) == 0) ? 1.0 : 0.0;
        }
        e->c = mat_from(&c[j * n], n, 1, n);
        e->K_type = cov_char(Kt[j]);
        e->K = Kt[j] == UK_COV_w ? mat_from(&K[j * Kw()], n, 1, n) : mat_from(&K[j * Kw()], n, n, n);
      }
      if (m[j] > 0) {
        const int mi = m[j];
        e->G = mat_from(&G[j * (size_t)mr * n], mi, n, mr);
        e->o = mat_from(&o[j * (size_t)mr], mi, 1, mr);
        e->C_type = cov_char(Ct[j]);
        e->C = Ct[j] == UK_COV_w ? mat_from(&C[j * Cw()], mi, 1, mr) : mat_from(&C[j * Cw()], mi, mi, mr);
      }
      if (hasx[j]) e->state = mat_from(&x[j * n], n, 1, n);
      if (hasS[j]) {
        e->covariance = mat_from(&S[j * nn_], n, n, n);
        e->covariance_type = Stype;
      }
      eqs.push_back(e);
    }
    general = true;
    for (auto* v : {&F, &H, &c, &K, &G, &o, &C, &x, &S}) std::vector<double>().swap(*v);
    start = 0;
    cap = 0;
  }
  /* re-lays a field from stride a to stride b (b >= a), the first `keep` entries of each slot */
  static void restride(std::vector<double>& v, int64_t slots, size_t a, size_t b, size_t keep) {
    std::vector<double> w(slots * b, 0.0);
    for (int64_t j = 0; j < slots; ++j) std::memcpy(&w[j * b], &v[j * a], keep * 8);
    v.swap(w);
  }
};

extern "C" uk_ingest* uk_ingest_create(void) { return new uk_ingest(); }
extern "C" void uk_ingest_free(uk_ingest* g) { delete g; }
extern "C" int64_t uk_ingest_count(const uk_ingest* g) { return g ? g->count : 0; }
extern "C" int64_t uk_ingest_first(const uk_ingest* g) { return g ? g->step0 : 0; }
extern "C" int32_t uk_ingest_dimension(const uk_ingest* g) { return g ? g->n : -1; }

/* general mode: the step's equations as deep copies (checked by the packing at smoothing time) */
static int ingest_evolve_general(uk_ingest* g, int32_t n, const kalman_matrix_t* Hm, const kalman_matrix_t* Fm,
                                 const kalman_matrix_t* cm, const kalman_matrix_t* Km, char K_type) {
  if (n < 1 || n > UKHIP_MAX_N) { uk_set_error("kalman_evolve: state dimension %d outside 1..%d", n, UKHIP_MAX_N); return -1; }
  if (g->count > 0 && (!Hm || !Fm || !cm || !Km)) {
    uk_set_error("kalman_evolve: H, F, c and K are required after the first step");
    return -1;
  }
  if (g->current) { /* a second evolve replaces the pending step */
    uk_ingest::eq_free(g->eqs.back());
    g->eqs.pop_back();
  }
  auto* e = (kalman_step_equations_t*)std::calloc(1, sizeof(kalman_step_equations_t));
  e->step = (kalman_step_index_t)(g->step0 + g->count);
  e->dimension = n;
  e->covariance_type = 'C';
  if (g->count > 0) {
    e->H = matrix_create_copy((kalman_matrix_t*)Hm);
    e->F = matrix_create_copy((kalman_matrix_t*)Fm);
    e->c = matrix_create_copy((kalman_matrix_t*)cm);
    e->K = matrix_create_copy((kalman_matrix_t*)Km);
    e->K_type = K_type;
  }
  g->eqs.push_back(e);
  g->current = true;
  return 0;
}

extern "C" int uk_ingest_evolve(uk_ingest* g, int32_t n, const kalman_matrix_t* Hm, const kalman_matrix_t* Fm,
                                const kalman_matrix_t* cm, const kalman_matrix_t* Km, char K_type) {
  if (!g) return -1;
  if (!g->general && g->n >= 0 &&
      (n != g->n || (g->count > 0 && Hm && Fm && cm &&
                     (Hm->row_dim != n || Hm->col_dim != n || Fm->row_dim != n || Fm->col_dim != n || cm->row_dim != n))))
    g->to_general(); /* another dimension or a rectangular evolution equation */
  if (g->general) return ingest_evolve_general(g, n, Hm, Fm, cm, Km, K_type);
  if (g->n < 0) {
    if (n < 1 || n > UKHIP_MAX_N) { uk_set_error("kalman_evolve: state dimension %d outside 1..%d", n, UKHIP_MAX_N); return -1; }
    g->n = n;
  } else if (n != g->n) {
    uk_set_error("kalman_evolve: dimension %d differs from %d (libukhip smooths constant dimensions)", n, g->n);
    return -1;
  }
  g->grow(g->start + g->count + 1);
  const int64_t j = g->start + g->count;
  g->current = true;
  g->hasx[j] = g->hasS[j] = 0;
  g->m[j] = 0;
  if (g->count == 0) return 0; /* the first step has no evolution equation */
  const size_t nn = g->nn();
  if (!Hm || !Fm || !cm) { uk_set_error("kalman_evolve: H, F and c are required after the first step"); g->current = false; return -1; }
  if (Fm->row_dim != n || Fm->col_dim != n || Hm->row_dim != n || Hm->col_dim != n || cm->row_dim != n ||
      cm->col_dim != 1 || cov_code(K_type) < 0 || !cov_shape_ok(Km, K_type, n)) {
    uk_set_error("kalman_evolve: H, F, c or K (type '%c') has the wrong shape or type", K_type);
    g->current = false;
    return -1;
  }
  if (K_type != 'w' && !g->K_full) {
    uk_ingest::restride(g->K, g->cap, n, nn, n);
    g->K_full = true;
  }
  if (!g->H_any && !is_identity(Hm)) {
    g->H.assign(g->cap * nn, 0.0);
    for (int64_t q = 0; q < g->cap; ++q)
      for (int d = 0; d < n; ++d) g->H[q * nn + (size_t)d * n + d] = 1.0;
    g->H_any = true;
  }
  copy_cols(&g->F[j * nn], n, Fm, n, n);
  if (g->H_any) copy_cols(&g->H[j * nn], n, Hm, n, n);
  copy_cols(&g->c[j * n], n, cm, n, 1);
  copy_cols(&g->K[j * g->Kw()], n, Km, Km->row_dim, Km->col_dim);
  g->Kt[j] = (int8_t)cov_code(K_type);
  return 0;
}

extern "C" int uk_ingest_observe(uk_ingest* g, const kalman_matrix_t* Gm, const kalman_matrix_t* om,
                                 const kalman_matrix_t* Cm, char C_type) {
  if (!g || !g->current) { uk_set_error("kalman_observe: call kalman_evolve first"); return -1; }
  if (!g->general && om && om->row_dim > g->n) g->to_general(); /* more observation rows than states */
  if (g->general) {
    kalman_step_equations_t* e = g->eqs.back();
    if (om) {
      if (!Gm || !Cm) { uk_set_error("kalman_observe: G and C are required with o"); return -1; }
      e->G = matrix_create_copy((kalman_matrix_t*)Gm);
      e->o = matrix_create_copy((kalman_matrix_t*)om);
      e->C = matrix_create_copy((kalman_matrix_t*)Cm);
      e->C_type = C_type;
    }
    g->current = false;
    g->count++;
    return 0;
  }
  const int n = g->n;
  const int64_t j = g->start + g->count;
  if (om) {
    const int mi = om->row_dim;
    if (mi > n || om->col_dim != 1 || !Gm || Gm->row_dim != mi || Gm->col_dim != n || cov_code(C_type) < 0 ||
        !cov_shape_ok(Cm, C_type, mi)) {
      uk_set_error("kalman_observe: G, o or C (type '%c') has the wrong shape, or more rows than states", C_type);
      return -1;
    }
    if (mi > g->mr || (C_type != 'w' && !g->C_full)) {
      const int nm = std::max(mi, g->mr);
      const bool cf = g->C_full || C_type != 'w';
      std::vector<double> G2(g->cap * (size_t)nm * n, 0.0), o2(g->cap * (size_t)nm, 0.0),
          C2(g->cap * (cf ? (size_t)nm * nm : (size_t)nm), 0.0);
      for (int64_t q = 0; q < g->cap; ++q) {
        for (int col = 0; col < n; ++col)
          std::memcpy(&G2[q * nm * n + (size_t)col * nm], &g->G[q * g->mr * n + (size_t)col * g->mr], g->mr * 8);
        std::memcpy(&o2[q * nm], &g->o[q * g->mr], g->mr * 8);
        const size_t cw2 = cf ? (size_t)nm * nm : (size_t)nm;
        if (g->C_full)
          for (int col = 0; col < g->mr; ++col)
            std::memcpy(&C2[q * cw2 + (size_t)col * nm], &g->C[q * g->Cw() + (size_t)col * g->mr], g->mr * 8);
        else
          std::memcpy(&C2[q * cw2], &g->C[q * g->Cw()], g->mr * 8);
      }
      g->G.swap(G2);
      g->o.swap(o2);
      g->C.swap(C2);
      g->mr = nm;
      g->C_full = cf;
    }
    const int mr = g->mr;
    copy_cols(&g->G[j * (size_t)mr * n], mr, Gm, mi, n);
    copy_cols(&g->o[j * (size_t)mr], mr, om, mi, 1);
    if (C_type == 'w') copy_cols(&g->C[j * g->Cw()], mr, Cm, mi, 1);
    else copy_cols(&g->C[j * g->Cw()], mr, Cm, mi, mi);
    g->Ct[j] = (int8_t)cov_code(C_type);
    g->m[j] = mi;
  } else {
    g->m[j] = 0;
  }
  g->current = false;
  g->count++;
  return 0;
}

extern "C" int uk_ingest_smooth(uk_ingest* g, uint32_t options) {
  if (!g || g->count == 0) return 0;
  if (!ensure_stream()) return -1;
  const bool assoc = (options & KALMAN_ALGORITHM_ASSOCIATIVE) != 0;
  if (g->general) {
    std::vector<kalman_step_equations_t*> live(g->eqs.begin(), g->eqs.begin() + g->count);
    for (auto* e : live) { /* earlier results are replaced */
      matrix_free(e->state);
      matrix_free(e->covariance);
      e->state = e->covariance = nullptr;
    }
    return smooth_equations((kalman_options_t)options, live.data(), (kalman_step_index_t)g->count, assoc) ? 0 : -1;
  }
  const int n = g->n;
  const size_t nn = g->nn(), S = (size_t)g->count, j0 = (size_t)g->start;
  Staging st;
  st.n = n;
  st.mr = g->mr;
  st.steps = g->count;
  st.nF = g->count;
  st.H_ident = g->H_any ? 0 : 1;
  st.K_full = g->K_full ? 1 : 0;
  st.C_full = g->C_full ? 1 : 0;
  st.layout();
  std::vector<int64_t> lengths{g->count};
  ukhip_batch* b = batch_from_staging(st, lengths, g_stream, nullptr, true);
  if (!b) return -1;
  std::vector<int64_t> fidx(S);
  std::vector<int8_t> first(S, 0);
  for (size_t i = 0; i < S; ++i) fidx[i] = (int64_t)i;
  first[0] = 1; /* the earliest live step: its evolution equation refers to a forgotten state */
  char* d = (char*)b->d_raw;
  auto up = [&](size_t off, const void* src, size_t bytes) {
    return bytes == 0 || hip_ok(hipMemcpyAsync(d + off, src, bytes, hipMemcpyHostToDevice, g_stream), "upload");
  };
  bool ok = up(st.offFidx, fidx.data(), S * 8) && up(st.offF, &g->F[j0 * nn], S * nn * 8) &&
            (!g->H_any || up(st.offH, &g->H[j0 * nn], S * nn * 8)) && up(st.offc, &g->c[j0 * n], S * n * 8) &&
            up(st.offK, &g->K[j0 * g->Kw()], S * g->Kw() * 8) && up(st.offG, &g->G[j0 * g->mr * n], S * g->mr * n * 8) &&
            up(st.offo, &g->o[j0 * g->mr], S * g->mr * 8) && up(st.offC, &g->C[j0 * g->Cw()], S * g->Cw() * 8) &&
            up(st.offKt, &g->Kt[j0], S) && up(st.offCt, &g->Ct[j0], S) && up(st.offm, &g->m[j0], S * 4) &&
            up(st.offfirst, first.data(), S);
  const int cov = assoc ? 1 : ((options & KALMAN_NO_COVARIANCE) == 0);
  if (ok) ok = assoc ? uk_assoc_smooth(b, options) == 0 : run_oddeven(b, cov, nullptr, g_stream);
  const bool wform = cov && (options & UKHIP_COVARIANCE_W) != 0;
  if (ok) ok = ukhip_batch_download(b, 0, &g->x[j0 * n], (cov && !wform) ? &g->S[j0 * nn] : nullptr) == 0;
  if (ok && wform) ok = ukhip_batch_download_cov(b, 0, 'W', &g->S[j0 * nn]) == 0;
  batch_release(b);
  if (!ok) return -1;
  for (size_t i = 0; i < S; ++i) {
    g->hasx[j0 + i] = 1;
    g->hasS[j0 + i] = cov ? 1 : 0;
  }
  g->Stype = wform ? 'W' : 'C';
  return 0;
}

static int64_t ingest_slot(const uk_ingest* g, int64_t si) {
  if (!g || g->count == 0) return -1;
  if (si < 0) si = g->step0 + g->count - 1;
  if (si < g->step0 || si >= g->step0 + g->count) {
    uk_set_error("step %lld is not in memory", (long long)si);
    return -1;
  }
  return (g->general ? 0 : g->start) + (si - g->step0);
}

extern "C" const double* uk_ingest_state(const uk_ingest* g, int64_t si) {
  const int64_t j = ingest_slot(g, si);
  if (j >= 0 && g->general) {
    const kalman_step_equations_t* e = g->eqs[j];
    return e->state ? e->state->elements : nullptr;
  }
  return (j >= 0 && g->hasx[j]) ? &g->x[j * g->n] : nullptr;
}

extern "C" const double* uk_ingest_covariance(const uk_ingest* g, int64_t si, char* type) {
  const int64_t j = ingest_slot(g, si);
  if (j >= 0 && g->general) {
    const kalman_step_equations_t* e = g->eqs[j];
    if (type) *type = e->covariance ? e->covariance_type : 'C';
    return e->covariance ? e->covariance->elements : nullptr;
  }
  if (type) *type = g ? g->Stype : 'C';
  return (j >= 0 && g->hasS[j]) ? &g->S[j * g->nn()] : nullptr;
}

/* state dimension of step si (-1: not in memory) */
extern "C" int32_t uk_ingest_step_dimension(const uk_ingest* g, int64_t si) {
  const int64_t j = ingest_slot(g, si);
  if (j < 0) return -1;
  return g->general ? g->eqs[j]->dimension : g->n;
}

extern "C" int uk_ingest_valid(const uk_ingest* g, int64_t si) { return ingest_slot(g, si) >= 0; }

/* c/kalman_base.c kalman_forget: drops the steps up to si (default: all but the latest), never the latest. */
extern "C" void uk_ingest_forget(uk_ingest* g, int64_t si) {
  if (!g || g->count == 0) return;
  const int64_t latest = g->step0 + g->count - 1;
  if (si < 0) si = latest - 1;
  if (si > latest - 1) return;
  const int64_t d = si - g->step0 + 1;
  if (d <= 0) return;
  if (g->general) {
    for (int64_t q = 0; q < d; ++q) {
      uk_ingest::eq_free(g->eqs.front());
      g->eqs.pop_front();
    }
  } else {
    g->start += d;
  }
  g->count -= d;
  g->step0 += d;
}

/* c/kalman_base.c kalman_rollback: drops the steps after si; step si returns
 * to the state just after its kalman_evolve (observation and results cleared). */
extern "C" void uk_ingest_rollback(uk_ingest* g, int64_t si) {
  if (!g || g->count == 0) return;
  if (si > g->step0 + g->count - 1 || si < g->step0) return;
  g->count = si - g->step0;
  if (g->general) {
    while ((int64_t)g->eqs.size() > g->count + 1) {
      uk_ingest::eq_free(g->eqs.back());
      g->eqs.pop_back();
    }
    uk_ingest::eq_clear_observation(g->eqs.back());
    g->current = true;
    return;
  }
  const int64_t j = g->start + g->count;
  g->m[j] = 0;
  g->hasx[j] = g->hasS[j] = 0;
  g->current = true;
}

/* ------------------------------------------------------------------------ */
/* run-time autotuning of the QR launch table                                 */
/* ------------------------------------------------------------------------ */



/* For every bucket (24, 32, 48, 64) up to n_max: a synthetic problem (n = bucket,
 * m = n/2, k steps, F = 0.98 times a Householder reflection, G uniform with
 * variance 1/n, 'w' covariances) is smoothed with each phase kernel in each
 * of its launch variants (1, 2, 4 workgroups per CU asked of the register
 * allocator, two columns per lane group, and for phase 2 the blocked-WY form),
 * timed with HIP events; the fastest variant of each phase goes into the
 * launch table. The register variants execute the same instructions, so the
 * choice among them changes timing only; the blocked-WY form applies the
 * same reflectors in another order (results agree to rounding). Batches whose
 * smoothing was captured into a graph earlier keep their captured launches. */
extern "C" int ukhip_autotune(int32_t n_max, int32_t k) {
  if (!ensure_stream()) return -1;
  if (k < 64) k = 4096;
  uint64_t st = 0x9E3779B97F4A7C15ull;
  auto uni = [&]() { /* xorshift64*, uniform in (-1, 1) */
    st ^= st >> 12; st ^= st << 25; st ^= st >> 27;
    return (double)((st * 0x2545F4914F6CDD1Dull) >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  };
  for (int nb : {24, 32, 48, 64}) {
    if (nb > std::max(32, (int)n_max)) break;
    const int n = nb, m = nb / 2;
    const size_t nn = (size_t)n * n;
    std::vector<double> v(n), F(nn), K(n, 1.0), G((size_t)k * m * n), o((size_t)k * m), C((size_t)m, 1.0 / std::sqrt(0.1));
    double vv = 0.0;
    for (int i = 0; i < n; ++i) { v[i] = uni(); vv += v[i] * v[i]; }
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) F[i + (size_t)j * n] = 0.98 * ((i == j ? 1.0 : 0.0) - 2.0 * v[i] * v[j] / vv);
    const double gs = std::sqrt(3.0 / n);
    for (auto& g : G) g = gs * uni();
    for (auto& x : o) x = uni();
    std::vector<double> Kall((size_t)k * n, 1.0), Call((size_t)k * m);
    for (size_t i = 0; i < Call.size(); ++i) Call[i] = C[i % m];
    ukhip_problem_t p{};
    p.n = n; p.m = m; p.k = k;
    p.F = F.data(); p.F_shared = 1;
    p.K = Kall.data(); p.K_type = 'w';
    p.G = G.data(); p.o = o.data(); p.C = Call.data(); p.C_type = 'w';
    ukhip_batch_t* b = ukhip_batch_create(1, &p);
    if (!b) return -1;
    /* {cpl, wg}; cpl 0 (blocked WY) exists for phase 2 only */
    static const int kVariants[9][2] = {{1, 1}, {1, 2}, {1, 4}, {2, 2}, {2, 3}, {2, 4}, {3, 2}, {3, 3}, {0, 1}};
    /* The phase's HIP events of an eager smoothing: FP64 odd-even, or the
     * mixed-precision path without a refinement step for the FP32
     * elimination. One phase varied at a time. */
    for (int fp32 = 0; fp32 <= 1; ++fp32) {
      for (int phase = 1; phase <= 3; ++phase) {
        int best_c = 1;
        int best_w = ukhip_qr_launch_get(fp32, phase, n, &best_c);
        double best_ms = 1e30;
        for (const auto& var : kVariants) {
          if (ukhip_qr_launch_set(fp32, phase, n, var[0], var[1]) != 0) continue; /* not a variant of this phase */
          double t = 0.0;
          for (int rep = 0; rep < 3; ++rep) {
            double ms[UKHIP_OE_SLOTS];
            const ukhip_mixed_options_t mo{0, 0, 0};
            const int rc = ukhip_batch_timed_parts(b, 0, fp32 ? &mo : nullptr, ms, nullptr, nullptr, 1) < 0;
            if (rc != 0) {
              ukhip_batch_free(b);
              return -1;
            }
            if (rep == 0) continue;
            t += ms[UKHIP_OE_PHASE1 + phase - 1];
          }
          if (t < best_ms) {
            best_ms = t;
            best_c = var[0];
            best_w = var[1];
          }
        }
        ukhip_qr_launch_set(fp32, phase, n, best_c, best_w);
      }
    }
    ukhip_batch_free(b);
  }
  return 0;
}
