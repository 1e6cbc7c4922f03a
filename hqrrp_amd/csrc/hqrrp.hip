// hqrrp.hip -- host orchestration of HQRRP on one MI355X and the C ABI.
//
// The block loop restates NoFLA_HQRRP_WY_blk_var4 (reference
// lapack_compatible_sources/NoFLA_HQRRP_WY_blk_var4.c:303-571): sketch
// Y = Omega*A, then per block of nb columns
//   sketch pivots (k_sketch) -> gather of A/Y/jpvt columns (k_misc)
//   -> panel QRCP + T (k_panel) -> trailing dlarfb (k_gemm)
//   -> sketch downdate (k_gemm).
// Everything is stream-ordered; pivots never leave the device and the host
// never waits inside the loop (all shapes depend only on j).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/hqrrp_amd.h"
#include "common.h"
#include "kernels.h"

#define HQ_TRY(expr)                                                  \
  do {                                                                \
    hipError_t e_ = (expr);                                           \
    if (e_ != hipSuccess) {                                           \
      fprintf(stderr, "hqrrp: %s failed: %s (%s:%d)\n", #expr,        \
              hipGetErrorString(e_), __FILE__, __LINE__);             \
      return HQRRP_ERR_HIP;                                           \
    }                                                                 \
  } while (0)

#define HQ_RET(expr)            \
  do {                          \
    int rc_ = (expr);           \
    if (rc_ != HQRRP_OK) return rc_; \
  } while (0)

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

int ensure(DevBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (b.bytes >= bytes) return HQRRP_OK;
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
  if (hipMalloc(&b.p, bytes) != hipSuccess) {
    fprintf(stderr, "hqrrp: device allocation of %zu bytes failed\n", bytes);
    return HQRRP_ERR_ALLOC;
  }
  b.bytes = bytes;
  return HQRRP_OK;
}

// Reentrant restatement of glibc rand(): additive lagged-Fibonacci generator
// r[i] = r[i-3] + r[i-31] (mod 2^32) seeded by the minimal-standard LCG,
// first 310 outputs discarded, outputs shifted right by one bit.
struct RefRand {
  uint32_t r[31];
  int f = 3, b = 0;
  explicit RefRand(unsigned seed) {
    int32_t word = (int32_t)(seed == 0 ? 1 : seed);
    r[0] = (uint32_t)word;
    for (int i = 1; i < 31; ++i) {
      int32_t hi = word / 127773, lo = word % 127773;
      word = 16807 * lo - 2836 * hi;
      if (word < 0) word += 2147483647;
      r[i] = (uint32_t)word;
    }
    for (int i = 0; i < 310; ++i) next();
  }
  int next() {
    r[f] += r[b];
    int out = (int)((r[f] >> 1) & 0x7fffffff);
    f = (f + 1) % 31;
    b = (b + 1) % 31;
    return out;
  }
};

// Exchange buffers of a persistent kernel (common.h): two buffers used by
// alternate launches.  used[i] is the prefix of buffer i written by the last
// launch that ran on it; the next launch (on the other buffer) restores the
// sentinel over it, so a launch always starts on an all-sentinel buffer.
struct Xchg {
  DevBuf buf[2];
  long used[2] = {0, 0};
  int par = 0;
  bool dirty = false;  // a launch aborted: both buffers are refilled
};

// Workspace of the allocation-free classic QRCP (hqrrp_geqp3_noalloc_dev):
// a file-scope device array, one copy per GPU, held until the process ends.
// Host code never uses its name as an address: hqrrp_create looks up this
// device's copy with hipGetSymbolAddress and the handle keeps that pointer.
constexpr size_t kScratchBytes = HQRRP_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

std::mutex g_scratch_mu;
void* g_scratch_ptr[64];  // per device, from hipGetSymbolAddress

// Device address of this device's scratch (the device must be current).
int scratch_address(int device, void** out) {
  if (device < 0 || device >= 64) return HQRRP_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "hqrrp: hipGetSymbolAddress(scratch) failed: %s\n",
              hipGetErrorString(e));
      return HQRRP_ERR_HIP;
    }
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return HQRRP_OK;
}

}  // namespace

namespace {

// Marsaglia polar method exactly as NoFLA_Normal_random_number
// (NoFLA_HQRRP_WY_blk_var4.c:593-614): pairs of uniforms on (-1,1) with
// 0 < a < 1, returning c1*f and then the cached c2*f, column-major fill.
// (have, cached) is the generator's cached second variate on entry and on
// exit: the reference keeps it in function statics that survive srand().
void polar_fill(long total, double* G, unsigned seed, bool& have,
                double& cached) {
  RefRand rng(seed);
  const double rmax = 2147483647.0;
  for (long idx = 0; idx < total; ++idx) {
    if (have) {
      have = false;
      G[idx] = 0.0 + 1.0 * cached;
      continue;
    }
    double c1, c2, a;
    do {
      c1 = -1.0 + 2.0 * ((double)rng.next() / rmax);
      c2 = -1.0 + 2.0 * ((double)rng.next() / rmax);
      a = c1 * c1 + c2 * c2;
    } while ((a == 0) || (a >= 1));
    double f = sqrt((-2 * log(a)) / a);
    cached = c2 * f;
    have = true;
    G[idx] = 0.0 + 1.0 * (c1 * f);
  }
}

// The process-wide cached variate of the reference generator.  The
// reference seeds rand() with srand(12) at the start of every HQRRP call
// (NoFLA_HQRRP_WY_blk_var4.c:380) but its normal generator's statics
// (lines 594-595) carry over: a call that draws an odd number r*m of
// normals leaves the second variate of its last pair cached, and the next
// call in the same process returns that value first.  Every factorisation
// that draws the reference's Omega reads and advances this state the same
// way, so a sequence of calls sees the sequence of Omegas the reference's
// would.
std::mutex g_carry_mu;
bool g_carry_have = false;
double g_carry_value = 0.0;

}  // namespace

extern "C" void hqrrp_reference_omega(int r, int m, double* G, unsigned seed) {
  bool have = false;
  double cached = 0.0;
  polar_fill((long)r * m, G, seed, have, cached);
}

extern "C" void hqrrp_reference_omega_carry(int r, int m, double* G,
                                            unsigned seed, int* have,
                                            double* cached) {
  bool hv = have && *have;
  double cv = (hv && cached) ? *cached : 0.0;
  polar_fill((long)r * m, G, seed, hv, cv);
  if (have) *have = hv ? 1 : 0;
  if (cached) *cached = cv;
}

extern "C" void hqrrp_omega_carry_get(int* have, double* value) {
  std::lock_guard<std::mutex> lk(g_carry_mu);
  if (have) *have = g_carry_have ? 1 : 0;
  if (value) *value = g_carry_value;
}

extern "C" void hqrrp_omega_carry_set(int have, double value) {
  std::lock_guard<std::mutex> lk(g_carry_mu);
  g_carry_have = have != 0;
  g_carry_value = have ? value : 0.0;
}

// Phases timed by the optional profiler (hqrrp_profile_*).
enum {
  PH_SKETCH_GEMM = 0,  // Y = Omega*A
  PH_SKETCH_QRCP,      // pivot selection on the sketch
  PH_GATHER,           // A/Y/jpvt column gathers, look-ahead column copies
  PH_PANEL,            // panel QRCP kernel
  PH_PANEL_POST,       // T factor + permutation of the rows above the panel
  PH_TRAIL_VTC,        // trailing M = V'C (split-K partials)
  PH_TRAIL_TW,         // W = T'M; in the driver also R12, B' and Y2
  PH_TRAIL_UPD,        // trailing C -= V W
  PH_DOWNDATE,         // sketch downdate (component entry only)
  PH_SELECT,           // (unused: the panel kernel selects and updates)
  HQ_PHASES
};
static const char* const kPhaseNames[HQ_PHASES] = {
    "sketch_gemm", "sketch_qrcp", "gather", "panel", "panel_post",
    "trail_vtc", "trail_tw", "trail_update", "downdate", "select_update"};

struct hqrrp_handle_s {
  int device = 0;
  hipStream_t own = nullptr, stream = nullptr;
  hipStream_t cur = nullptr;  // stream the internal helpers launch on
  // Look-ahead streams: the next block's panel runs on s_chain (restricted
  // to chain_cus CUs, so its persistent workgroups are co-resident at once)
  // while the trailing update runs on s_upd (the remaining CUs).
  hipStream_t s_chain = nullptr, s_upd = nullptr;
  int chain_cus = 0;
  hipEvent_t ev_fork = nullptr, ev_upd = nullptr, ev_chain = nullptr;
  hipEvent_t ev_perm = nullptr;  // look-ahead: the panel copied into A
  hipEvent_t ev_sketch = nullptr, ev_t = nullptr;  // look-ahead: moves, T
  // Joins the caller's stream to the handle's own stream when the caller's
  // is the null stream (whose implicit synchronisation with every blocking
  // stream would slow the loop's many small launches).
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  int cu_count = 0;
  size_t lds_max = 0;
  int sketch_max_wg = 0, panel_max_wg = 0;
  DevBuf words;  // ints [8]: error word of the persistent kernels;
                 // u64 [8, 14): the prescaling state (launch_scale_pre)
  DevBuf Y, G, Gcache, moves, piv, gram, T, perm;
  Xchg xs, xp;  // exchange buffers of the sketch and panel kernels
  DevBuf tmpA, tmpY, tmpJ, split, Mx, W2, Pp, Qm;
  // Sketch precision of the running factorisation: Y and G hold floats
  // (sk32) and the sketch GEMM rounds its operands to bf16 (skbf).
  int sk32 = 0, skbf = 0;
  DevBuf Pn;  // look-ahead: the next panel (two, used in turn)
  // Cholesky-QR panel (k_cqr.hip): R1 (64 x 64) and the fallback flag.
  // panel_alg 0 selects it (with the per-step kernel as its fallback), 1
  // keeps every panel on the per-step kernel.
  DevBuf cqr;
  int panel_alg = 0;
  // The last launch on xp was the gated per-step kernel after a
  // Cholesky-QR panel, whose flag (in cqr) says whether it wrote xp.
  bool xp_after_cqr = false;
  DevBuf gen, genT, genM;  // generic block step: sketch copy / norms, T
                           // products, trailing products
  DevBuf Xb;             // look-ahead: second trailing-matrix buffer (m x n)
  DevBuf adjAP, adjTau, adjK, adjX;  // fixed-pivot QR and its adjoint
  DevBuf hA, hJ, hT, hW;  // device copies used by the host entry points
  DevBuf vn;              // classic QRCP: partial and reference norms
  void* scratch = nullptr;  // this device's allocation-free workspace
  int om_r = -1, om_m = -1;
  unsigned om_seed = 0;
  // Cached-variate state the cached Omega was drawn with, and after it.
  bool om_have = false, om_out_have = false;
  double om_in = 0.0, om_out = 0.0;
  // Batched path: block size per 64 x 64 shape bucket chosen by
  // hqrrp_batch_tune for (bt_pp, bt_piv); 0 = not tuned.
  int bt_nb[8][8] = {};
  int bt_pp = -1, bt_piv = -1;
  std::mutex mu;
  // Optional per-phase timing with HIP events recorded on `stream`.
  bool prof = false;
  unsigned prof_mask = ~0u;  // phases recorded when prof is on
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  struct Span {
    int phase;
    hipEvent_t a, b;
  };
  std::vector<Span> spans;
  double ph_flops[HQ_PHASES] = {}, ph_bytes[HQ_PHASES] = {};
  long long ph_launch[HQ_PHASES] = {};

  int* err() const { return words.as<int>() + 8; }
  // Prescaling state: slot 0 for a factorisation, slot 1 for a host entry
  // that wraps one (dgeqp4 with fixed columns).
  unsigned long long* scale_words(int slot = 0) const {
    return words.as<unsigned long long>() + 8 + 8 * slot;
  }
};

namespace {

int check_err(hqrrp_handle_t h) {
  int e = 0;
  HQ_TRY(hipMemcpyAsync(&e, h->err(), sizeof(int), hipMemcpyDeviceToHost,
                        h->stream));
  HQ_TRY(hipStreamSynchronize(h->stream));
  if (e != 0) {
    fprintf(stderr, "hqrrp: kernel reported error %d (grid barrier timeout)\n",
            e);
    HQ_TRY(hipMemsetAsync(h->words.p, 0, h->words.bytes, h->stream));
    HQ_TRY(hipStreamSynchronize(h->stream));
    h->xs.dirty = h->xp.dirty = true;
    return HQRRP_ERR_KERNEL;
  }
  return HQRRP_OK;
}

hipEvent_t next_event(hqrrp_handle_t h) {
  if (h->ev_used == h->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    h->ev_pool.push_back(e);
  }
  return h->ev_pool[h->ev_used++];
}

// Brackets the launches of one phase with events when profiling is on and
// accumulates the phase's algorithmic flops/bytes and launch count.
struct PhaseScope {
  hqrrp_handle_t h;
  int ph;
  hipEvent_t a = nullptr;
  PhaseScope(hqrrp_handle_t h_, int ph_, double flops, double bytes,
             int launches)
      : h(h_), ph(ph_) {
    if (!h->prof || !((h->prof_mask >> ph) & 1u)) return;
    h->ph_flops[ph] += flops;
    h->ph_bytes[ph] += bytes;
    h->ph_launch[ph] += launches;
    a = next_event(h);
    if (a) (void)hipEventRecord(a, h->cur);
  }
  ~PhaseScope() {
    if (!a) return;
    hipEvent_t b = next_event(h);
    if (!b) return;
    (void)hipEventRecord(b, h->cur);
    h->spans.push_back({ph, a, b});
  }
};

// Sets h->cur for an entry point: the caller's stream, or -- when that is
// the null stream -- the handle's own stream joined to it by events at
// entry and exit (the null stream's implicit synchronisation with every
// blocking stream would slow a sequence of small launches).
struct StreamScope {
  hqrrp_handle_t h;
  bool joined = false;
  explicit StreamScope(hqrrp_handle_t h_) : h(h_) {
    h->cur = h->stream;
    if (!h->stream && hipEventRecord(h->ev_in, nullptr) == hipSuccess &&
        hipStreamWaitEvent(h->own, h->ev_in, 0) == hipSuccess) {
      h->cur = h->own;
      joined = true;
    }
  }
  ~StreamScope() {
    if (!joined) return;
    (void)hipEventRecord(h->ev_out, h->own);
    (void)hipStreamWaitEvent(nullptr, h->ev_out, 0);
    h->cur = h->stream;
  }
};

// Split-K factor for a product with `tiles` output tiles and reduction
// length K: aim for >= 2 workgroups per CU, slices of >= 256.
int choose_splits(hqrrp_handle_t h, long tiles, long K, long* kchunk) {
  long want = std::max(1L, (2L * h->cu_count + tiles - 1) / tiles);
  long maxs = std::max(1L, K / 256);
  long s = std::min(want, maxs);
  long kc = (K + s - 1) / s;
  kc = (kc + 15) / 16 * 16;
  if (kc <= 0) kc = 16;
  *kchunk = kc;
  return (int)std::max(1L, (K + kc - 1) / kc);
}

// C := alpha op(A) op(B) + beta C with automatic split-K.
int gemm(hqrrp_handle_t h, bool ta, bool tb, int M, int N, int K, double alpha,
         const double* A, long lda, int maskA, const double* B, long ldb,
         int maskB, double beta, double* C, long ldc, bool allow_split = true,
         int tag = 0) {
  if (M <= 0 || N <= 0) return HQRRP_OK;
  GemmArgs g{};
  g.M = M;
  g.N = N;
  g.K = K;
  g.alpha = alpha;
  g.beta = beta;
  g.A = A;
  g.lda = lda;
  g.B = B;
  g.ldb = ldb;
  g.C = C;
  g.ldc = ldc;
  g.maskA = maskA;
  g.maskB = maskB;
  long tiles = (long)hq_cdiv(M, 64) * hq_cdiv(N, 64);
  long kc = K > 0 ? K : 16;
  int S = 1;
  if (allow_split && K > 0) S = choose_splits(h, tiles, K, &kc);
  g.kchunk = kc;
  if (S > 1) {
    HQ_RET(ensure(h->split, sizeof(double) * (size_t)S * M * N));
    g.part = h->split.as<double>();
    HQ_TRY(launch_gemm(g, ta, tb, S, h->cur, tag));
    HQ_TRY(launch_reduce_splitk(M, N, S, g.part, beta, C, ldc, h->cur));
  } else {
    g.part = nullptr;
    HQ_TRY(launch_gemm(g, ta, tb, 1, h->cur, tag));
  }
  return HQRRP_OK;
}

// Buffers for a launch that writes `need` doubles of exchange slots.
int xchg_begin(hqrrp_handle_t h, Xchg& x, size_t need, double** cur,
               double** prev, long* prev_len) {
  for (int i = 0; i < 2; ++i) {
    const bool grow = x.buf[i].bytes < need * sizeof(double);
    if (grow) HQ_RET(ensure(x.buf[i], need * sizeof(double)));
    if (grow || x.dirty) {
      HQ_TRY(launch_fill_sentinel(x.buf[i].as<double>(),
                                  x.buf[i].bytes / sizeof(double), h->cur));
      x.used[i] = 0;
    }
  }
  x.dirty = false;
  *cur = x.buf[x.par].as<double>();
  *prev = x.buf[1 - x.par].as<double>();
  *prev_len = x.used[1 - x.par];
  return HQRRP_OK;
}

// After a successful launch that wrote `used` doubles of its buffer.
void xchg_end(Xchg& x, size_t used) {
  x.used[1 - x.par] = 0;
  x.used[x.par] = (long)used;
  x.par ^= 1;
}

// C (float) := alpha op(A) op(B) + beta C in FP32 arithmetic; each operand
// FP64 (x64 = 1) or FP32; automatic split-K (k_gemm32.hip).
int gemm32(hqrrp_handle_t h, bool ta, bool tb, int M, int N, int K,
           float alpha, const void* A, long lda, int a64, int maskA,
           const void* B, long ldb, int b64, int maskB, float beta, float* C,
           long ldc, int bf16, bool allow_split = true) {
  if (M <= 0 || N <= 0) return HQRRP_OK;
  Gemm32Args g{};
  g.M = M;
  g.N = N;
  g.K = K;
  g.alpha = alpha;
  g.beta = beta;
  g.A = A;
  g.lda = lda;
  g.a64 = a64;
  g.B = B;
  g.ldb = ldb;
  g.b64 = b64;
  g.C = C;
  g.ldc = ldc;
  g.maskA = maskA;
  g.maskB = maskB;
  g.bf16 = bf16;
  long tiles = (long)hq_cdiv(M, 64) * hq_cdiv(N, 64);
  long kc = K > 0 ? K : 16;
  int S = 1;
  if (allow_split && K > 0) S = choose_splits(h, tiles, K, &kc);
  g.kchunk = kc;
  if (S > 1) {
    HQ_RET(ensure(h->split, sizeof(float) * (size_t)S * M * N));
    g.part = h->split.as<float>();
    HQ_TRY(launch_gemm32(g, ta, tb, S, h->cur));
    HQ_TRY(launch_reduce32(M, N, S, g.part, alpha, beta, C, ldc, h->cur));
  } else {
    g.part = nullptr;
    HQ_TRY(launch_gemm32(g, ta, tb, 1, h->cur));
  }
  return HQRRP_OK;
}

// The mixed-precision sketch Y = Omega A (r x n, FP32) from G' = Omega'
// (m x r, FP64, already rounded to the sketch precision's values) and A
// (m x n, ld lda, FP64): on the matrix cores (k_mixed.hip), split K so the
// grid fills the CUs twice; beyond the kernel's r, an FP32 VALU product
// of the same (rounded) operator.
int sketch_gemm_mixed(hqrrp_handle_t h, int r, int m, int n, const double* Gt,
                      const double* dA, long lda, float* Y) {
  if (r > mixed_sketch_rmax())
    return gemm32(h, true, false, r, n, m, 1.0f, Gt, m, 1, 0, dA, lda, 1, 0,
                  0.0f, Y, r, 0);
  MixedSketchArgs a{};
  a.G = Gt;
  a.ldg = m;
  a.A = dA;
  a.lda = lda;
  a.m = m;
  a.n = n;
  a.r = r;
  a.Y = Y;
  a.ldy = r;
  const long tiles = hq_cdiv(n, mixed_sketch_cols_per_wg());
  const long kb = mixed_sketch_kbk();
  long S = std::max(1L, std::min((2L * h->cu_count + tiles - 1) / tiles,
                                 (long)m / (4 * kb)));
  long kc = (m + S - 1) / S;
  kc = (kc + kb - 1) / kb * kb;
  S = std::max(1L, (m + kc - 1) / kc);
  a.kchunk = kc;
  const int prec = h->skbf ? 2 : 1;
  if (S == 1) {
    HQ_TRY(launch_sketch_mixed(a, prec, 1, h->cur));
    return HQRRP_OK;
  }
  HQ_RET(ensure(h->split, sizeof(float) * (size_t)S * r * n));
  a.part = h->split.as<float>();
  HQ_TRY(launch_sketch_mixed(a, prec, (int)S, h->cur));
  HQ_TRY(launch_reduce32(r, n, (int)S, a.part, 1.0f, 0.0f, Y, r, h->cur));
  return HQRRP_OK;
}

// T (b x b, ld ldt) of b reflectors from their inner products (gram:
// strictly upper part, ld ldg) and tau, for any b: larft on 64-column
// diagonal blocks, then the off-diagonal blocks by doubling,
//   T(B1, B2) = -T(B1, B1) G(B1, B2) T(B2, B2)  (dlarft's recurrence in
// block form; the same matrix in exact arithmetic).
int larft_any(hqrrp_handle_t h, int b, const double* gram, long ldg,
              const double* tau, double* T, long ldt) {
  if (b <= 64) {
    HQ_TRY(launch_larft(b, gram, ldg, tau, T, ldt, h->cur));
    return HQRRP_OK;
  }
  HQ_TRY(launch_fill_eye(T, ldt, b, b, 0.0, h->cur));
  for (int q0 = 0; q0 < b; q0 += 64)
    HQ_TRY(launch_larft(std::min(64, b - q0), gram + q0 + (long)q0 * ldg, ldg,
                        tau + q0, T + q0 + (long)q0 * ldt, ldt, h->cur));
  HQ_RET(ensure(h->genT, sizeof(double) * (size_t)b * b));
  double* X = h->genT.as<double>();
  for (int sz = 64; sz < b; sz *= 2) {
    for (int c0 = 0; c0 + sz < b; c0 += 2 * sz) {
      const int c1 = c0 + sz, s2 = std::min(sz, b - c1);
      HQ_RET(gemm(h, false, false, sz, s2, s2, 1.0, gram + c0 + (long)c1 * ldg,
                  ldg, 0, T + c1 + (long)c1 * ldt, ldt, 0, 0.0, X, sz, false));
      HQ_RET(gemm(h, false, false, sz, s2, sz, -1.0, T + c0 + (long)c0 * ldt,
                  ldt, 0, X, sz, 0, 0.0, T + c0 + (long)c1 * ldt, ldt, false));
    }
  }
  return HQRRP_OK;
}

// Greedy QRCP for `stages` steps on a copy of the sketch (r x ncols)
// without the persistent kernel's limits (r, columns, stages): the
// reference's NoFLA_QRPmod_WY_unb_var4 on V_R = Y_R (NoFLA_HQRRP_WY_blk_var4
// .c:500-506), two launches per step, then the composed moves.
int run_sketch_generic(hqrrp_handle_t h, int r, int ncols, const void* dY,
                       long ldy, int stages, int* dpiv, int* dmoves, int f32) {
  const size_t rc = (size_t)r * ncols;
  HQ_RET(ensure(h->gen, sizeof(double) * (rc + 2 * (size_t)ncols + stages) +
                            sizeof(int) * (size_t)ncols));
  double* V = h->gen.as<double>();
  double* vn1 = V + rc;
  double* vn2 = vn1 + ncols;
  double* tw = vn2 + ncols;
  int* pos = reinterpret_cast<int*>(tw + stages);
  if (f32)
    HQ_TRY(launch_widen_block(V, r, static_cast<const float*>(dY), ldy, r,
                              ncols, h->cur));
  else
    HQ_TRY(launch_copy_block(V, r, static_cast<const double*>(dY), ldy, r,
                             ncols, h->cur));
  HQ_TRY(launch_iota(pos, ncols, h->cur));
  HQ_TRY(launch_classic_norms(V, r, r, ncols, 0, vn1, vn2, h->cur));
  for (int k = 0; k < stages; ++k)
    HQ_TRY(launch_qrpmod_step(V, r, r, ncols, k, 1, pos, dpiv, tw, vn1, vn2,
                              h->cur));
  HQ_TRY(launch_perm_moves(pos, ncols, dmoves, h->cur));
  return HQRRP_OK;
}

// Householder QR with optional pivoting of an mrows x b panel without the
// persistent kernel's limits (b, rows): NoFLA_QRPmod_WY_unb_var4 with
// num_stages = -1 and build_T (NoFLA_HQRRP_WY_blk_var4.c:520-524), two
// launches per step; the reflectors' inner products by one product, T by
// larft_any.
int run_panel_generic(hqrrp_handle_t h, int mrows, int b, double* dP, long ldp,
                      int pivoting, double* dtau, int* dperm, double* dT,
                      long ldt) {
  HQ_RET(ensure(h->gen, sizeof(double) * 2 * (size_t)b));
  HQ_RET(ensure(h->gram, sizeof(double) * (size_t)b * b));
  double* vn1 = h->gen.as<double>();
  double* vn2 = vn1 + b;
  HQ_TRY(launch_iota(dperm, b, h->cur));
  if (pivoting) HQ_TRY(launch_classic_norms(dP, ldp, mrows, b, 0, vn1, vn2,
                                            h->cur));
  {
    PhaseScope ps(h, PH_PANEL, 2.0 * mrows * b * b - 2.0 / 3.0 * b * b * b,
                  16.0 * mrows * b, 2 * b);
    for (int k = 0; k < std::min(b, mrows); ++k)
      HQ_TRY(launch_qrpmod_step(dP, ldp, mrows, b, k, pivoting, dperm, nullptr,
                                dtau, vn1, vn2, h->cur));
  }
  double* gram = h->gram.as<double>();
  HQ_RET(gemm(h, true, false, b, b, mrows, 1.0, dP, ldp, 1, dP, ldp, 1, 0.0,
              gram, b));
  if (!dT) return HQRRP_OK;
  PhaseScope ps(h, PH_PANEL_POST, 0.0, 0.0, 1);
  return larft_any(h, b, gram, b, dtau, dT, ldt);
}

int panel_geometry(hqrrp_handle_t h, int mrows, int* nw);

int panel_geometry(hqrrp_handle_t h, int mrows, int* nw) {
  int n = hq_cdiv(mrows, panel_rows_per_wg());
  if (n > h->panel_max_wg) return HQRRP_ERR_UNSUPPORTED;
  *nw = n;
  return HQRRP_OK;
}

int run_sketch(hqrrp_handle_t h, int r, int ncols, const void* dY, long ldy,
               int stages, int* dpiv, int* dmoves, int f32 = 0) {
  int rmax = sketch_rmax(r);
  if (rmax < 0 || stages > sketch_max_stages() ||
      hq_cdiv(ncols, sketch_cols_per_wg(ncols)) > h->sketch_max_wg)
    return run_sketch_generic(h, r, ncols, dY, ldy, stages, dpiv, dmoves, f32);
  SketchArgs a{};
  a.Y = dY;
  a.ldy = ldy;
  a.f32 = f32;
  a.r = r;
  a.ncols = ncols;
  a.stages = stages;
  a.cols = sketch_cols_per_wg(ncols);
  a.nw = hq_cdiv(ncols, a.cols);
  if (a.nw > h->sketch_max_wg) return HQRRP_ERR_UNSUPPORTED;
  const size_t need = sketch_xchg_doubles(a.nw, rmax, stages);
  HQ_RET(xchg_begin(h, h->xs, need, &a.xcur, &a.xprev, &a.xprev_len));
  a.piv = dpiv;
  a.moves = dmoves;
  a.err = h->err();
  HQ_TRY(launch_sketch_qrcp(a, rmax, h->cur));
  xchg_end(h->xs, need);
  return HQRRP_OK;
}

// The panel from two Gram passes (k_cqr.hip): Cholesky QR2, the reference's
// QRCP of the b x b factor R, Householder reconstruction; T is written
// directly.  A panel the kernels flag (nearly dependent columns, a column
// with nothing below its diagonal) is factored by the per-step kernel,
// whose workgroup 0 then forms T as larft would: one launch, gated on the
// flag, so the host never waits.  sel: as in run_panel (the selected columns are
// formed and stored into dP by the first pass).
int run_panel_cqr(hqrrp_handle_t h, int mrows, int b, double* dP, long ldp,
                  int pivoting, double* dtau, int* dperm, double* dT, long ldt,
                  const PanelArgs* sel) {
  const int nwc = hq_cdiv(mrows, cqr_rows_per_wg());
  HQ_RET(ensure(h->cqr, sizeof(double) * (4096 + 64)));
  CqrArgs c{};
  c.R1 = h->cqr.as<double>();
  c.flag = reinterpret_cast<int*>(c.R1 + 4096);
  c.P = dP;
  c.ldp = ldp;
  c.mrows = mrows;
  c.b = b;
  c.pivoting = pivoting;
  c.nw = nwc;
  c.tau = dtau;
  c.perm = dperm;
  c.T = dT;
  c.ldt = ldt;
  c.err = h->err();
  if (sel) {
    c.sel_moves = sel->sel_moves;
    c.sel_b2 = sel->sel_b2;
    c.sel_r0 = sel->sel_r0;
    c.sel_b = sel->sel_b;
    c.sel_C = sel->sel_C;
    c.sel_ldc = sel->sel_ldc;
    c.sel_V = sel->sel_V;
    c.sel_ldv = sel->sel_ldv;
    c.sel_W = sel->sel_W;
    c.sel_ldw = sel->sel_ldw;
  }
  {
    const size_t need = cqr_xchg_doubles(c.nw);
    HQ_RET(xchg_begin(h, h->xp, need, &c.xcur, &c.xprev, &c.xprev_len));
    c.xprev_if = h->xp_after_cqr ? c.flag : nullptr;
    PhaseScope ps(h, PH_PANEL, 2.0 * mrows * b * b - 2.0 / 3.0 * b * b * b,
                  16.0 * mrows * b, 1);
    HQ_TRY(launch_cqr_panel(c, h->cur));
    // The kernel restores its partials' sentinels itself.
    xchg_end(h->xp, cqr_xchg_dirty());
  }
  // Fallback, gated on the flag: the per-step kernel on the panel as stored.
  PanelArgs a{};
  HQ_RET(panel_geometry(h, mrows, &a.nw));
  a.P = dP;
  a.ldp = ldp;
  a.mrows = mrows;
  a.b = b;
  a.pivoting = pivoting;
  a.tau = dtau;
  a.perm = dperm;
  a.run_if = c.flag;
  HQ_RET(ensure(h->gram, sizeof(double) * 64 * 64));
  const size_t xneed = panel_xchg_doubles(a.nw, b);
  HQ_RET(xchg_begin(h, h->xp, xneed, &a.xcur, &a.xprev, &a.xprev_len));
  a.gram = h->gram.as<double>();
  a.err = h->err();
  a.T = dT;
  a.ldt = ldt;
  {
    PhaseScope ps(h, PH_PANEL_POST, 0.0, 0.0, 1);
    HQ_TRY(launch_panel_qrcp(a, h->cur));
    xchg_end(h->xp, xneed);
    h->xp_after_cqr = true;
  }
  return HQRRP_OK;
}

// dT == nullptr: T is left to the caller (larft from h->gram).  sel: the
// look-ahead selection and update of the panel's columns, done by the panel
// kernel as it loads them (its sel_* fields; null: the panel is read from
// dP).
int run_panel(hqrrp_handle_t h, int mrows, int b, double* dP, long ldp,
              int pivoting, double* dtau, int* dperm, double* dT, long ldt,
              const PanelArgs* sel = nullptr) {
  if (b < 1 || mrows < b) return HQRRP_ERR_UNSUPPORTED;
  int nw_fit = 0;
  if (b > 64 || panel_geometry(h, mrows, &nw_fit) != HQRRP_OK) {
    // The look-ahead only runs where the persistent kernels fit.
    if (sel) return HQRRP_ERR_UNSUPPORTED;
    return run_panel_generic(h, mrows, b, dP, ldp, pivoting, dtau, dperm, dT,
                             ldt);
  }
  if (h->panel_alg == 0 && dT)
    return run_panel_cqr(h, mrows, b, dP, ldp, pivoting, dtau, dperm, dT, ldt,
                         sel);
  PanelArgs a{};
  if (sel) {
    a.sel_moves = sel->sel_moves;
    a.sel_b2 = sel->sel_b2;
    a.sel_r0 = sel->sel_r0;
    a.sel_b = sel->sel_b;
    a.sel_C = sel->sel_C;
    a.sel_ldc = sel->sel_ldc;
    a.sel_V = sel->sel_V;
    a.sel_ldv = sel->sel_ldv;
    a.sel_W = sel->sel_W;
    a.sel_ldw = sel->sel_ldw;
  }
  HQ_RET(panel_geometry(h, mrows, &a.nw));
  a.P = dP;
  a.ldp = ldp;
  a.mrows = mrows;
  a.b = b;
  a.pivoting = pivoting;
  a.tau = dtau;
  a.perm = dperm;
  HQ_RET(ensure(h->gram, sizeof(double) * 64 * 64));
  const size_t need = panel_xchg_doubles(a.nw, b);
  HQ_RET(xchg_begin(h, h->xp, need, &a.xcur, &a.xprev, &a.xprev_len));
  a.gram = h->gram.as<double>();
  a.err = h->err();
  {
    // Algorithmic traffic: the panel is read and written once.
    PhaseScope ps(h, PH_PANEL, 2.0 * mrows * b * b - 2.0 / 3.0 * b * b * b,
                  16.0 * mrows * b, 1);
    HQ_TRY(launch_panel_qrcp(a, h->cur));
  }
  xchg_end(h->xp, need);
  h->xp_after_cqr = false;
  if (!dT) return HQRRP_OK;
  PhaseScope ps(h, PH_PANEL_POST, 0.0, 0.0, 1);
  HQ_TRY(launch_larft(b, a.gram, b, dtau, dT, ldt, h->cur));
  return HQRRP_OK;
}

// [A12; A22] := (I - V T V')' [A12; A22]  (dlarfb Left/Transpose):
// M = V'C (split-K), W = T'M, C -= V W  (k_trail.hip).  trail_prepare sizes
// the split and the buffers, trail_w computes W' (nc x b, in h->W2),
// trail_apply the rank-b update; the look-ahead driver runs the last on a
// stream of its own.
int trail_prepare(hqrrp_handle_t h, int mrows, int nc, int b, const double* V,
                  long ldv, const double* T, long ldt, double* C, long ldc,
                  TrailArgs* out, double* C2 = nullptr, long ldc2 = 0,
                  int nc2 = 0) {
  TrailArgs a{};
  a.mrows = mrows;
  a.nc = nc;
  a.b = b;
  // Optional second column range (updated in place), from virtual column
  // nca on: tiles of every kernel (64, 128, 256 columns) never straddle it.
  a.C2 = (C2 && nc2 > 0) ? C2 : nullptr;
  a.ldc2 = ldc2;
  a.nc2 = a.C2 ? nc2 : 0;
  a.nca = a.C2 ? (int)(hq_cdiv(nc, 256) * 256L) : nc;
  a.ncv = a.C2 ? a.nca + a.nc2 : nc;
  nc = a.ncv;  // tiles, partials and W' below count virtual columns
  a.vmask = 1;
  a.V = V;
  a.ldv = ldv;
  a.C = C;
  a.ldc = ldc;
  a.T = T;
  a.ldt = ldt;
  // V'C: the stream-K partition over all wavefronts of one workgroup per
  // CU when every wavefront gets >= 128 rows of a 64-column group; else
  // 64-column tiles (two workgroups per CU) with K split uniformly, in
  // slices of at least 64 rows, so that the grid fills but does not exceed
  // those slots (a second, partial round of workgroups would double the
  // time; a slice walks its rows serially, so few long slices leave a small
  // product latency-bound).
  const long nch = hq_cdiv(mrows, kVtcChunk);
  const long ngroups = hq_cdiv(a.nc, 64) + hq_cdiv(a.nc2, 64);
  const long waves = 4L * h->cu_count;
  long S;
  if (nch * ngroups >= 8 * waves) {
    a.sk = 1;
    a.sk_waves = (int)waves;
    a.sk_nch = nch;
    a.sk_q = nch * ngroups / waves;
    a.sk_r = nch * ngroups % waves;
    const VtcPart vp = vtc_part(a);
    S = 1;
    for (long g = 0; g < ngroups; ++g) S = std::max(S, (long)vp.count(g));
  } else {
    const long maxs = std::max(1L, (long)mrows / 64);
    const long tiles = hq_cdiv(nc, 64);
    const long slots = 2L * h->cu_count;
    const int kbk = trail_kbk();
    S = std::max(1L, slots / tiles);
    S = std::min(S, maxs);
    long kc = (mrows + S - 1) / S;
    kc = (kc + kbk - 1) / kbk * kbk;
    S = (mrows + kc - 1) / kc;
    a.kchunk = kc;
  }
  a.splits = (int)S;
  HQ_RET(ensure(h->split, sizeof(double) * (size_t)S * b * nc));
  HQ_RET(ensure(h->W2, sizeof(double) * (size_t)b * nc));
  a.part = h->split.as<double>();
  a.W = h->W2.as<double>();
  a.ldw = nc;
  *out = a;
  return HQRRP_OK;
}

int trail_vtc_only(hqrrp_handle_t h, const TrailArgs& a) {
  const double mn = (double)a.mrows * (a.nc + (a.C2 ? a.nc2 : 0)), b = a.b;
  PhaseScope ps(h, PH_TRAIL_VTC, 2.0 * b * mn, 8.0 * mn + 8.0 * a.mrows * b,
                1);
  HQ_TRY(launch_trail_vtc(a, h->cur));
  return HQRRP_OK;
}

int trail_w(hqrrp_handle_t h, const TrailArgs& a) {
  const double b = a.b, S = a.splits;
  HQ_RET(trail_vtc_only(h, a));
  PhaseScope ps(h, PH_TRAIL_TW, 2.0 * b * b * a.nc, 8.0 * (S + 1.0) * b * a.nc,
                1);
  HQ_TRY(launch_trail_tw(a, h->cur));
  return HQRRP_OK;
}

int trail_apply(hqrrp_handle_t h, const TrailArgs& a) {
  if (a.mrows <= 0 || a.nc <= 0) return HQRRP_OK;
  // C (and C2) are read and written once; V and W are small.
  const double mn = (double)a.mrows * (a.nc + (a.C2 ? a.nc2 : 0));
  PhaseScope ps(h, PH_TRAIL_UPD, 2.0 * a.b * mn, 16.0 * mn, 1);
  // The CUs this stream may use (the look-ahead update stream holds all
  // but the chain's); the persistent update kernel needs 160 KB of LDS.
  TrailArgs u = a;
  u.ncu = h->lds_max >= 160 * 1024
              ? (h->cur == h->s_upd && h->s_upd ? h->cu_count - h->chain_cus
                                                : h->cu_count)
              : 0;
  HQ_TRY(launch_trail_update(u, h->cur));
  return HQRRP_OK;
}

int apply_qt_left(hqrrp_handle_t h, int mrows, int nc, int b, const double* V,
                  long ldv, const double* T, long ldt, double* C, long ldc) {
  if (nc <= 0 || mrows <= 0) return HQRRP_OK;
  TrailArgs a;
  HQ_RET(trail_prepare(h, mrows, nc, b, V, ldv, T, ldt, C, ldc, &a));
  HQ_RET(trail_w(h, a));
  return trail_apply(h, a);
}

// Sketch downdate with the old G_R = [G1 G2]:  P = G_R V, Q = P T,
// G_R -= Q V' (whose first b columns are the reference's B matrix), then
// Y2 -= G1_new R12 (four launches, k_downdate.hip).
int downdate_y(hqrrp_handle_t h, int r, int b, int mU, int nY2, const double* U,
               long ldu, const double* R12, long ldr, const double* T, long ldt,
               double* Y2, long ldy, double* G, long ldg) {
  HQ_RET(ensure(h->Pp, sizeof(double) * (size_t)r * b));
  HQ_RET(ensure(h->Qm, sizeof(double) * (size_t)r * b));
  if (r > 80 || b > 64) {  // beyond the fused kernels: generic products
    double* P = h->Pp.as<double>();
    double* Q = h->Qm.as<double>();
    HQ_RET(gemm(h, false, false, r, b, mU, 1.0, G, ldg, 0, U, ldu, 1, 0.0, P,
                r));
    HQ_RET(gemm(h, false, false, r, b, b, 1.0, P, r, 0, T, ldt, 0, 0.0, Q, r,
                false));
    HQ_RET(gemm(h, false, true, r, mU, b, -1.0, Q, r, 0, U, ldu, 1, 1.0, G,
                ldg, false));
    if (nY2 > 0)
      HQ_RET(gemm(h, false, false, r, nY2, b, -1.0, G, ldg, 0, R12, ldr, 0,
                  1.0, Y2, ldy, false));
    return HQRRP_OK;
  }
  DowndateArgs d{};
  d.S = downdate_slices(mU, &d.kc);
  HQ_RET(ensure(h->split, sizeof(double) * (size_t)d.S * r * b));
  d.r = r;
  d.b = b;
  d.mU = mU;
  d.nY2 = std::max(nY2, 0);
  d.U = U;
  d.ldu = ldu;
  d.T = T;
  d.ldt = ldt;
  d.R12 = R12;
  d.ldr = ldr;
  d.G = G;
  d.ldg = ldg;
  d.Y2 = Y2;
  d.ldy = ldy;
  d.part = h->split.as<double>();
  d.P = h->Pp.as<double>();
  d.Q = h->Qm.as<double>();
  HQ_TRY(launch_downdate(d, h->cur));
  return HQRRP_OK;
}


int load_omega(hqrrp_handle_t h, int r, int m, const hqrrp_options& o) {
  // The driver keeps Omega' (m x r, ld m, FP64) for both sketch precisions:
  // it is updated as extra columns of the FP64 trailing matrix.
  HQ_RET(ensure(h->G, sizeof(double) * (size_t)r * m));
  if (o.omega_kind == HQRRP_OMEGA_USER) {
    if (!o.omega) return HQRRP_ERR_ARG;
    HQ_TRY(launch_transpose(o.omega, r, r, m, h->G.as<double>(), m, h->cur));
    return HQRRP_OK;
  }
  // The reference's Omega depends on the generator's cached variate as
  // this call finds it (process-wide, advanced by every call).
  std::lock_guard<std::mutex> lk(g_carry_mu);
  const bool in_have = g_carry_have;
  const double in_value = g_carry_value;
  if (h->om_r != r || h->om_m != m || h->om_seed != o.seed ||
      h->om_have != in_have ||
      (in_have && __builtin_memcmp(&h->om_in, &in_value, sizeof(double)))) {
    std::vector<double> host((size_t)r * m);
    bool hv = in_have;
    double cv = in_value;
    polar_fill((long)r * m, host.data(), o.seed, hv, cv);
    HQ_RET(ensure(h->Gcache, sizeof(double) * (size_t)r * m));
    HQ_TRY(hipMemcpy(h->Gcache.p, host.data(), sizeof(double) * host.size(),
                     hipMemcpyHostToDevice));
    h->om_r = r;
    h->om_m = m;
    h->om_seed = o.seed;
    h->om_have = in_have;
    h->om_in = in_value;
    h->om_out_have = hv;
    h->om_out = cv;
  }
  g_carry_have = h->om_out_have;
  g_carry_value = h->om_out;
  HQ_TRY(launch_transpose(h->Gcache.as<double>(), r, r, m, h->G.as<double>(),
                          m, h->cur));
  return HQRRP_OK;
}

// Sketch pivots of the block starting at column j (b stages over the n - j
// remaining sketch columns) into h->piv / h->moves.
// Element j*r of the sketch (FP64 or FP32 per h->sk32).
void* sk_at(hqrrp_handle_t h, void* base, long off) {
  return static_cast<char*>(base) + off * (h->sk32 ? 4 : 8);
}

int block_sketch(hqrrp_handle_t h, int r, int j, int n, int b, void* Y) {
  PhaseScope ps(h, PH_SKETCH_QRCP, 4.0 * r * (double)(n - j) * b,
                (h->sk32 ? 4.0 : 8.0) * r * (double)(n - j), 1);
  return run_sketch(h, r, n - j, sk_at(h, Y, (long)j * r), r, b,
                    h->piv.as<int>(), h->moves.as<int>(), h->sk32);
}

// Applies the composed moves of the last sketch to columns j.. of A (all m
// rows), of Y and of jpvt.
int block_gather(hqrrp_handle_t h, int r, int j, int m, int b, double* dA,
                 long lda, void* Y, int* djpvt) {
  PhaseScope ps(h, PH_GATHER, 0.0, 0.0, 2);
  HQ_TRY(launch_gather_cols(h->moves.as<int>(), 2 * b, dA + (long)j * lda, lda,
                            m, sk_at(h, Y, (long)j * r), r, r,
                            h->sk32 ? 4 : 8, djpvt + j, h->tmpA.as<double>(),
                            h->tmpY.as<double>(), h->tmpJ.as<int>(), h->cur));
  return HQRRP_OK;
}

// Rows 0..j-1 of the panel columns and jpvt follow the panel's pivots (for
// panels beyond 64 columns: as the permutation's moves, through the column
// gathers).
int block_permute(hqrrp_handle_t h, int j, int b, double* dA, long lda,
                  int* djpvt) {
  PhaseScope ps(h, PH_PANEL_POST, 0.0, 16.0 * j * b, 1);
  if (b <= 64) {
    HQ_TRY(launch_panel_permute(dA + (long)j * lda, lda, j, b,
                                h->perm.as<int>(), djpvt + j, h->cur));
    return HQRRP_OK;
  }
  int* mv = h->moves.as<int>();  // free between the gathers and the next sketch
  HQ_TRY(launch_perm_moves(h->perm.as<int>(), b, mv, h->cur));
  HQ_TRY(launch_gather_cols(mv, b, dA + (long)j * lda, lda, j, nullptr, 0, 0,
                            8, djpvt + j, h->tmpA.as<double>(),
                            h->tmpY.as<double>(), h->tmpJ.as<int>(), h->cur));
  return HQRRP_OK;
}

// Block tail and trailing update with generic FP64 products, for blocks
// beyond the fused kernels (b > 64 reflectors or r > 80 sketch rows):
// dlarfb (NoFLA_HQRRP_WY_blk_var4.c:529-541, 711-740) on the trailing
// matrix C (mrows x nc, rows 0..b-1 become R12) and, with Gt, the sketch
// downdate as in block_tail: G' -= V (T'(V'G')) on G' rows j.., whose first
// b rows are then the reference's B (617-686), and Y2 -= B R12 (691-693).
int block_trail_generic(hqrrp_handle_t h, int j, int jn, int mrows, int nc,
                        int b, const double* V, long ldv, const double* T,
                        long ldt, double* C, long ldc, double* Gt, int m,
                        int r, void* Y) {
  const int ncg = Gt ? r : 0;
  HQ_RET(ensure(h->genM, sizeof(double) * 2 * (size_t)b * (nc + ncg)));
  double* M = h->genM.as<double>();          // b x (nc + ncg): V'[C G']
  double* W = M + (size_t)b * (nc + ncg);    // b x (nc + ncg): T'M
  PhaseScope ps(h, PH_TRAIL_VTC, 4.0 * b * (double)mrows * (nc + ncg),
                16.0 * (double)mrows * (nc + ncg), 6);
  HQ_RET(gemm(h, true, false, b, nc, mrows, 1.0, V, ldv, 1, C, ldc, 0, 0.0, M,
              b));
  if (Gt)
    HQ_RET(gemm(h, true, false, b, r, mrows, 1.0, V, ldv, 1, Gt + j, m, 0, 0.0,
                M + (size_t)b * nc, b));
  HQ_RET(gemm(h, true, false, b, nc + ncg, b, 1.0, T, ldt, 0, M, b, 0, 0.0, W,
              b, false));
  HQ_RET(gemm(h, false, false, mrows, nc, b, -1.0, V, ldv, 1, W, b, 0, 1.0, C,
              ldc));
  if (!Gt) return HQRRP_OK;
  HQ_RET(gemm(h, false, false, mrows, r, b, -1.0, V, ldv, 1,
              W + (size_t)b * nc, b, 0, 1.0, Gt + j, m));
  const double* R12 = C;  // rows 0..b-1 of the updated trailing matrix
  if (h->sk32)
    return gemm32(h, true, false, r, nc, b, -1.0f, Gt + j, m, 1, 0, R12, ldc,
                  1, 0, 1.0f, static_cast<float*>(Y) + (long)jn * r, r, 0,
                  false);
  return gemm(h, true, false, r, nc, b, -1.0, Gt + j, m, 0, R12, ldc, 0, 1.0,
              static_cast<double*>(Y) + (long)jn * r, r, false);
}


// FP64 sketch downdate (NoFLA_Downdate_Y, NoFLA_HQRRP_WY_blk_var4.c:617-708,
// and the G update 743-771) folded into the trailing update: G' = G_R'
// (rows j..m-1 of the transposed sketch operator) rides along as r extra
// columns of the trailing matrix, so V'[C G'] gives P' = (G_R V)' next to
// M, the tail forms Q' = T'P' next to W and B' = G'(0:b, :) - V1 Q' (the
// reference's B, stored over G'(0:b, :)) next to R12, and the trailing
// update applies G' -= V Q' (G_R := G_R (I - V T V')) next to C -= V W.
// block_tail launches the tail and then Y2 -= B R12.
int block_tail(hqrrp_handle_t h, const TrailArgs& ta, int j, int jn,
               double* dA, long lda, void* Y, int r) {
  const double b = ta.b, nc = ta.ncv;
  PhaseScope ps(h, PH_TRAIL_TW,
                4.0 * b * b * nc + (ta.C2 ? 2.0 * r * b * ta.nc : 0.0),
                8.0 * nc * (b * ta.splits + 2.0 * b) +
                    (ta.C2 ? 16.0 * r * ta.nc : 0.0),
                ta.C2 ? 2 : 1);
  TailArgs t{};
  t.b = ta.b;
  t.nc = ta.nc;
  t.S = ta.splits;
  t.sk = ta.sk;
  t.vp = vtc_part(ta);
  t.nc2 = ta.nc2;
  t.nca = ta.nca;
  t.ncv = ta.ncv;
  t.part = ta.part;
  t.T = ta.T;
  t.ldt = ta.ldt;
  t.V = ta.V;
  t.ldv = ta.ldv;
  t.C = ta.C;
  t.ldc = ta.ldc;
  t.R = dA + j + (long)jn * lda;
  t.ldr = lda;
  t.C2 = ta.C2;
  t.ldc2 = ta.ldc2;
  t.W = ta.W;
  t.ldw = ta.ldw;
  // Y2 in the sketch's precision (FP32: computed in FP64, stored rounded).
  if (h->sk32)
    t.Y2f = static_cast<float*>(Y) + (long)jn * r;
  else
    t.Y2 = static_cast<double*>(Y) + (long)jn * r;
  t.ldy = r;
  HQ_TRY(launch_trail_tail(t, h->cur));
  if (ta.C2) HQ_TRY(launch_trail_y2(t, h->cur));
  return HQRRP_OK;
}

// The trailing update below row block j (R12 is already in A).
TrailArgs below_r12(const TrailArgs& ta) {
  TrailArgs u = ta;
  u.mrows -= ta.b;
  u.V += ta.b;
  u.vmask = 0;
  u.C += ta.b;
  if (u.Cout) u.Cout += ta.b;
  if (u.C2) u.C2 += ta.b;
  return u;
}

// The block loop of NoFLA_HQRRP_WY_blk_var4.c:415-562.
//
// Per block: V'[C G'] (split-K partials), then the block tail (W', R12
// into A, B' over G's top rows, Y2 -= B R12), then the update of the rows
// below R12 (C -= V W, G' -= V Q').
//
// Look-ahead: once block j's tail is done, the pivot chain of block j+1 runs
// on stream s_chain (a reserved quarter of the CUs, so its persistent
// kernels are co-resident) while block j's trailing update runs on s_upd
// (the other CUs):
//   sketch pivots of block j+1;
//   the b2 selected columns (rows below R12) read from the not yet updated
//             C and updated -- the same products in the same order as the
//             full update's -- into a panel buffer;
//   panel QRCP and T of block j+1 on that buffer.
// To let the chain read C while the update rewrites it, the update runs out
// of place: the trailing matrix alternates between A and an internal buffer
// of the same shape (X[0] = A, X[1]); every final value (R rows, panels) is
// stored into A.  After both streams join, the column gather, the panel
// write-back and the row permutation follow on the caller's stream.  Every
// value is produced by the same kernel with the same operands as without
// look-ahead, so the results are bitwise identical.
int factor_loop(hqrrp_handle_t h, int m, int n, int kmax, double* dA,
                long lda, int* djpvt, double* dtau, const hqrrp_options& o);

// Columns a factorisation truncated at kmax leaves factored: blocks of nb
// up to kmax, the block that would cross kmax cut there, except the
// reference's final block (j + nb >= m or n), which is factored whole.
int factored_columns(int m, int n, int nb, int kmax) {
  const int mn = std::min(m, n);
  for (int j = 0; j < mn; j += nb) {
    if (j + nb >= m || j + nb >= n) break;
    if (j + nb >= kmax) return kmax;
  }
  return mn;
}

int factor_locked(hqrrp_handle_t h, int m, int n, double* dA, long lda,
                  int* djpvt, double* dtau, const hqrrp_options& o) {
  const int mn = std::min(m, n);
  h->cur = h->stream;
  if (mn == 0) return HQRRP_OK;
  const int kmax = (o.max_rank > 0) ? std::min(mn, o.max_rank) : mn;
  if (o.sketch_prec < HQRRP_SKETCH_FP64 || o.sketch_prec > HQRRP_SKETCH_BF16)
    return HQRRP_ERR_ARG;

  HQ_TRY(hipSetDevice(h->device));
  h->sk32 = o.sketch_prec != HQRRP_SKETCH_FP64;
  h->skbf = o.sketch_prec == HQRRP_SKETCH_BF16;
  StreamScope ss(h);
  // A matrix of extreme magnitude is factored as 2^p A (exact: the pivots,
  // reflectors and tau do not change, R is scaled back): max |a| then lies
  // in [0.5, 1), so squared norms neither overflow nor underflow, and a
  // mixed sketch stays inside FP32's range.
  const int lim = h->sk32 ? 60 : 256;
  HQ_TRY(launch_scale_pre(dA, lda, m, n, h->scale_words(), -lim, lim, h->cur));
  HQ_RET(factor_loop(h, m, n, kmax, dA, lda, djpvt, dtau, o));
  HQ_TRY(launch_scale_post(dA, lda, m, n, factored_columns(m, n, o.nb, kmax),
                           h->scale_words(), h->cur));
  return HQRRP_OK;
}

int factor_loop(hqrrp_handle_t h, int m, int n, int kmax, double* dA,
                long lda, int* djpvt, double* dtau, const hqrrp_options& o) {
  const int nb = o.nb, pp = o.pp, r = nb + pp;
  HQ_RET(load_omega(h, r, m, o));
  HQ_RET(ensure(h->Y, sizeof(double) * (size_t)r * n));
  HQ_RET(ensure(h->moves, sizeof(int) * (size_t)(1 + 4 * nb)));
  HQ_RET(ensure(h->piv, sizeof(int) * (size_t)nb));
  HQ_RET(ensure(h->perm, sizeof(int) * (size_t)std::max(64, nb)));
  HQ_RET(ensure(h->gram, sizeof(double) * (size_t)std::max(64, nb) *
                             std::max(64, nb)));
  HQ_RET(ensure(h->T, sizeof(double) * (size_t)nb * nb));
  HQ_RET(ensure(h->tmpA, sizeof(double) * (size_t)2 * nb * m));
  HQ_RET(ensure(h->tmpY, sizeof(double) * (size_t)2 * nb * r));
  HQ_RET(ensure(h->tmpJ, sizeof(int) * (size_t)2 * nb));
  // Blocks beyond the fused kernels (more than 64 reflectors or 80 sketch
  // rows) run the generic block step, which has no look-ahead.
  const bool gen_trail = nb > 64 || r > 80;
  const bool use_la = o.lookahead != 0 && h->s_chain && h->s_upd &&
                      kmax > nb && n > nb && !gen_trail;
  if (use_la) {
    HQ_RET(ensure(h->Pn, sizeof(double) * (size_t)2 * nb * m));
    HQ_RET(ensure(h->Xb, sizeof(double) * (size_t)m * n));
  }
  void* Y = h->Y.p;  // FP64 or FP32 (h->sk32)
  void* G = h->G.p;
  double* T = h->T.as<double>();
  hipStream_t main_s = h->cur;
  // Trailing-matrix buffers; the trailing matrix lives in X[xc].
  double* X[2] = {dA, use_la ? h->Xb.as<double>() : nullptr};
  const long ldx[2] = {lda, (long)m};
  int xc = 0;

  // The mixed sketch's operator: Omega' rounded to the sketch precision's
  // values once, so that the FP64 downdate below is that of exactly the
  // operator the mixed product uses.
  if (h->sk32)
    HQ_TRY(launch_round_omega(static_cast<double*>(G), (long)m * r,
                              h->skbf ? 2 : 1, h->cur));
  // Y = Omega * A (NoFLA_HQRRP_WY_blk_var4.c:407-409).
  {
    PhaseScope ps(h, PH_SKETCH_GEMM, 2.0 * r * (double)m * n,
                  8.0 * ((double)m * n + (double)r * (m + n)), 1);
    // G holds Omega' (m x r, FP64); the mixed sketch forms Y in FP32
    // (optionally bf16-rounded operands) with FP32 accumulation.
    if (h->sk32) {
      HQ_RET(sketch_gemm_mixed(h, r, m, n, static_cast<const double*>(G), dA,
                               lda, static_cast<float*>(Y)));
    } else {
      // Y = G'' A is a V'C product with G' (unmasked) in the role of V: up
      // to 64 sketch rows per pass, each pass one streaming read of A with
      // the operands going straight into the matrix cores (split-K
      // partials, then a fixed-order sum into Y's rows).
      for (int p0 = 0; p0 < r; p0 += 64) {
        const int pb = std::min(64, r - p0);
        TrailArgs ta;
        HQ_RET(trail_prepare(h, m, n, pb, static_cast<double*>(G) + (long)p0 * m,
                             m, nullptr, 0, dA, lda, &ta));
        ta.vmask = 0;
        HQ_TRY(launch_trail_vtc(ta, h->cur));
        HQ_TRY(launch_trail_sum(ta, static_cast<double*>(Y) + p0, r,
                                h->cur));
      }
    }
  }

  const hqrrp_trace* tr = o.trace;
  const size_t ysz = h->sk32 ? sizeof(float) : sizeof(double);
  if (tr && tr->y0)
    HQ_TRY(hipMemcpyAsync(tr->y0, Y, ysz * (size_t)r * n,
                          hipMemcpyDeviceToDevice, h->cur));
  // Block j's sketch choices, in the sketch's order: jpvt(j:j+b) right after
  // the block's column gather, before the panel reorders them.
  auto trace_sketch = [&](int j, int b) -> int {
    if (tr && tr->sketch_cols)
      HQ_TRY(hipMemcpyAsync(tr->sketch_cols + j, djpvt + j, sizeof(int) * b,
                            hipMemcpyDeviceToDevice, h->cur));
    return HQRRP_OK;
  };

  // G' (m x r, ld m, FP64) is updated as extra columns of the trailing
  // matrix and the downdate of Y follows the block tail (block_tail); Y is
  // FP64 or, for the mixed sketch, FP32.
  double* Gt = static_cast<double*>(G);
  bool have_panel = false;  // block j's panel already factored by look-ahead
  bool t_pending = false;   // its T still being formed on s_chain (ev_t)
  // Width of the block starting at j: nb columns, cut to end at kmax unless
  // it is the reference's final block (the panel of all that remains).
  auto block_width = [&](int j) -> int {
    const int w = std::min(nb, std::min(n - j, m - j));
    const bool fin = (j + nb >= m) || (j + nb >= n);
    return fin ? w : std::min(w, kmax - j);
  };
  // A look-ahead panel's reflectors are read from its buffer in Pn (ld its
  // row count): its copy into A runs on s_upd beside the next V'C.
  double* Vp = nullptr;
  long ldvp = 0;
  int pn_par = 0;
  bool perm_pending = false;  // a panel copy on s_upd not yet joined (ev_perm)
  for (int j = 0; j < kmax; j += nb) {
    const int b = block_width(j);
    const bool last = (j + nb >= m) || (j + nb >= n);
    const int mrows = m - j, nc = n - j - b;
    double* Ajj = dA + j + (long)j * lda;
    double* Vj = have_panel ? Vp : Ajj;  // the block's reflectors
    const long ldvj = have_panel ? ldvp : lda;
    if (!have_panel) {
      if (!last) {
        HQ_RET(block_sketch(h, r, j, n, b, Y));
        HQ_RET(block_gather(h, r, j, m, b, dA, lda, Y, djpvt));
        HQ_RET(trace_sketch(j, b));
      }
      HQ_RET(run_panel(h, mrows, b, Ajj, lda, o.panel_pivoting, dtau + j,
                       h->perm.as<int>(), T, nb));
      HQ_RET(block_permute(h, j, b, dA, lda, djpvt));
    }
    have_panel = false;
    if (nc <= 0) continue;

    const int jn = j + b;
    int nw_panel = 0;
    const bool la =
        use_la && jn < kmax &&
        panel_geometry(h, m - jn, &nw_panel) == HQRRP_OK &&
        nw_panel <= h->chain_cus &&
        hq_cdiv(n - jn, sketch_cols_per_wg(n - jn)) <= h->chain_cus;
    if (!la && xc != 0) {
      // The in-place update below needs the trailing matrix back in A.
      PhaseScope ps(h, PH_PANEL_POST, 0.0, 16.0 * (double)mrows * nc, 1);
      HQ_TRY(launch_copy_block(dA + j + (long)jn * lda, lda,
                               X[1] + j + (long)jn * ldx[1], ldx[1], mrows, nc,
                               h->cur));
      xc = 0;
    }
    double* C = X[xc] + j + (long)jn * ldx[xc];
    // With a next block to pivot, G' rows j.. ride along.
    const bool gcols = !last;
    if (gen_trail) {
      HQ_RET(block_trail_generic(h, j, jn, mrows, nc, b, Ajj, lda, T, nb, C,
                                 ldx[xc], gcols ? Gt : nullptr, m, r, Y));
      continue;
    }
    TrailArgs ta;
    HQ_RET(trail_prepare(h, mrows, nc, b, Vj, ldvj, T, nb, C, ldx[xc], &ta,
                         gcols ? Gt + j : nullptr, m, gcols ? r : 0));
    HQ_RET(trail_vtc_only(h, ta));
    if (t_pending) HQ_TRY(hipStreamWaitEvent(main_s, h->ev_t, 0));
    t_pending = false;
    HQ_RET(block_tail(h, ta, j, jn, dA, lda, Y, r));
    if (!la) {
      HQ_RET(trail_apply(h, below_r12(ta)));
      continue;
    }

    // jn < kmax implies !last and b == nb.
    const int xn = xc ^ 1;
    double* Pn = h->Pn.as<double>() + (size_t)pn_par * nb * m;
    const int mnext = m - jn;
    const int b2 = block_width(jn);
    const bool last_next = (jn + nb >= m) || (jn + nb >= n);
    ta.Cout = X[xn] + j + (long)jn * ldx[xn];
    ta.ldco = ldx[xn];
    // R12 is already in A: the update covers the rows below it.
    const TrailArgs tu = below_r12(ta);
    // Fork: block j's update || block jn's pivot chain.  s_upd: the update,
    // then (once the chain's sketch has named them) block jn's column
    // gathers, which touch the updated trailing matrix, Y and jpvt.
    // s_chain: sketch, column selection and update, panel; its T follows
    // on s_chain after the join point, needed only by block jn's tail.
    HQ_TRY(hipEventRecord(h->ev_fork, main_s));
    HQ_TRY(hipStreamWaitEvent(h->s_upd, h->ev_fork, 0));
    HQ_TRY(hipStreamWaitEvent(h->s_chain, h->ev_fork, 0));
    h->cur = h->s_upd;
    int rc = trail_apply(h, tu);
    h->cur = h->s_chain;
    if (rc == HQRRP_OK && !last_next) {
      rc = block_sketch(h, r, jn, n, b2, Y);
      if (rc == HQRRP_OK && hipEventRecord(h->ev_sketch, h->s_chain) != hipSuccess)
        rc = HQRRP_ERR_HIP;
    }
    auto gathers = [&]() -> int {
      if (last_next) return HQRRP_OK;
      HQ_TRY(hipStreamWaitEvent(h->s_upd, h->ev_sketch, 0));
      if (xn == 0) {
        HQ_RET(block_gather(h, r, jn, m, b2, dA, lda, Y, djpvt));
      } else {
        HQ_RET(block_gather(h, r, jn, jn, b2, dA, lda, Y, djpvt));
        PhaseScope ps(h, PH_GATHER, 0.0, 0.0, 2);
        HQ_TRY(launch_gather_cols(h->moves.as<int>(), 2 * b2,
                                  X[1] + jn + (long)jn * ldx[1], ldx[1], mnext,
                                  nullptr, 0, 0, 8, nullptr, h->tmpA.as<double>(),
                                  h->tmpY.as<double>(), h->tmpJ.as<int>(),
                                  h->cur));
      }
      return trace_sketch(jn, b2);
    };
    if (rc == HQRRP_OK) {
      h->cur = h->s_upd;
      rc = gathers();
      if (rc == HQRRP_OK && hipEventRecord(h->ev_upd, h->s_upd) != hipSuccess)
        rc = HQRRP_ERR_HIP;
      h->cur = h->s_chain;
    }
    auto chain = [&]() -> int {
      // The next panel's columns are selected and updated (the same
      // products, in the same order, as the full update's) by the panel
      // kernel as it loads them.
      PanelArgs sel{};
      sel.sel_moves = last_next ? nullptr : h->moves.as<int>();
      sel.sel_b2 = b2;
      sel.sel_C = C;
      sel.sel_ldc = ldx[xc];
      sel.sel_r0 = b;
      sel.sel_V = Vj + b;
      sel.sel_ldv = ldvj;
      sel.sel_W = ta.W;
      sel.sel_ldw = ta.ldw;
      sel.sel_b = b;
      // The Cholesky-QR panel writes T itself; the per-step kernel leaves
      // it to larft, which then runs beside the next block's V'C.
      const bool direct_t = h->panel_alg == 0;
      HQ_RET(run_panel(h, mnext, b2, Pn, mnext, o.panel_pivoting, dtau + jn,
                       h->perm.as<int>(), direct_t ? T : nullptr, nb, &sel));
      HQ_TRY(hipEventRecord(h->ev_chain, h->s_chain));
      if (!direct_t) {
        PhaseScope ps(h, PH_PANEL_POST, 0.0, 0.0, 1);
        HQ_TRY(launch_larft(b2, h->gram.as<double>(), b2, dtau + jn, T, nb,
                            h->cur));
        HQ_TRY(hipEventRecord(h->ev_t, h->s_chain));
      }
      return HQRRP_OK;
    };
    if (rc == HQRRP_OK) rc = chain();
    h->cur = main_s;
    HQ_RET(rc);
    HQ_TRY(hipStreamWaitEvent(main_s, h->ev_upd, 0));
    HQ_TRY(hipStreamWaitEvent(main_s, h->ev_chain, 0));
    // The Cholesky-QR panel (and its gated fallback) wrote T before
    // ev_chain; the per-step panel's T follows on s_chain (ev_t).
    t_pending = h->panel_alg != 0;
    // Block jn's factored panel over its gathered columns and the rows
    // above it permuted, in one launch on s_upd (after the gathers, which
    // move the rows above, and the panel): beside block jn's V'C, which
    // reads the reflectors from Pn.
    h->cur = h->s_upd;
    rc = hipStreamWaitEvent(h->s_upd, h->ev_chain, 0) == hipSuccess
             ? HQRRP_OK
             : HQRRP_ERR_HIP;
    if (rc == HQRRP_OK) {
      PhaseScope ps(h, PH_PANEL_POST, 0.0,
                    16.0 * (double)mnext * b2 + 16.0 * jn * b2, 1);
      if (launch_panel_permute(dA + (long)jn * lda, lda, jn, b2,
                               h->perm.as<int>(), djpvt + jn, h->cur, Pn,
                               mnext, mnext) != hipSuccess ||
          hipEventRecord(h->ev_perm, h->s_upd) != hipSuccess)
        rc = HQRRP_ERR_HIP;
    }
    h->cur = main_s;
    HQ_RET(rc);
    perm_pending = true;
    Vp = Pn;
    ldvp = mnext;
    pn_par ^= 1;
    xc = xn;
    have_panel = true;
  }
  if (perm_pending) HQ_TRY(hipStreamWaitEvent(main_s, h->ev_perm, 0));
  // A last T formed on s_chain but never consumed (no trailing columns)
  // still reads h->gram: order the caller's next work after it.
  if (t_pending) HQ_TRY(hipStreamWaitEvent(main_s, h->ev_t, 0));
  if (tr && tr->y)
    HQ_TRY(hipMemcpyAsync(tr->y, Y, ysz * (size_t)r * n,
                          hipMemcpyDeviceToDevice, h->cur));
  if (tr && tr->g)
    HQ_TRY(hipMemcpyAsync(tr->g, Gt, sizeof(double) * (size_t)m * r,
                          hipMemcpyDeviceToDevice, h->cur));
  HQ_TRY(hipGetLastError());
  return HQRRP_OK;
}

// Householder QR without pivoting of the first kmax columns of A (m x n),
// each block's reflectors applied to all columns to its right (LAPACK
// dgeqrf; the reference calls it for dgeqp4's fixed columns, WY.c:221-225):
// panel (pivoting off) + T, then the trailing dlarfb.
int run_geqrf(hqrrp_handle_t h, int m, int n, int kmax, double* dA, long lda,
              double* dtau) {
  const int nb = 64;
  HQ_RET(ensure(h->T, sizeof(double) * nb * nb));
  HQ_RET(ensure(h->perm, sizeof(int) * 64));
  for (int j = 0; j < kmax; j += nb) {
    const int b = std::min(nb, std::min(kmax - j, m - j));
    double* Ajj = dA + j + (long)j * lda;
    HQ_RET(run_panel(h, m - j, b, Ajj, lda, 0, dtau + j, h->perm.as<int>(),
                     h->T.as<double>(), nb));
    if (j + b < n)
      HQ_RET(apply_qt_left(h, m - j, n - j - b, b, Ajj, lda, h->T.as<double>(),
                           nb, Ajj + (long)b * lda, lda));
  }
  return HQRRP_OK;
}

// Classic greedy QRCP of A(k0:m, k0:n) with LAPACK dgeqpf / dgeqp3 semantics
// (the small-size branches of NoFLA_HQRRP_WY_blk_var4.c:189-199): partial
// norms of rows k0.., then one pivot + update per column.  Swaps move whole
// columns (all m rows) and jpvt entries, as LAPACK's dswap does, so rows
// above k0 follow the pivots with no re-pivot pass.
// Workspace bytes of run_classic_qrcp for an m x n matrix.
size_t classic_ws_bytes(int m, int n) {
  if (classic_fused_ok(m, n))
    return sizeof(double) * (2 * (size_t)m * n + 4 * (size_t)n) +
           sizeof(int) * 2 * (size_t)n;
  return sizeof(double) * 2 * (size_t)n;
}

// ws: caller-provided workspace of classic_ws_bytes(m, n) bytes, or null
// for the handle's own (grown on demand).
int run_classic_qrcp(hqrrp_handle_t h, int m, int n, int k0, double* dA,
                     long lda, int* djpvt, double* dtau,
                     double* ws = nullptr) {
  const int mn = std::min(m, n);
  if (k0 >= mn) return HQRRP_OK;
  if (!ws) {
    HQ_RET(ensure(h->vn, classic_ws_bytes(m, n)));
    ws = h->vn.as<double>();
  }
  if (classic_fused_ok(m, n)) {
    // One launch per step, ping-ponging X <-> Y (matrix, jpvt, norms).
    const size_t mn2 = (size_t)m * n;
    double* X = ws;
    double* Y = X + mn2;
    double* vX = Y + mn2;
    double* vY = vX + 2 * (size_t)n;
    int* jX = reinterpret_cast<int*>(vY + 2 * (size_t)n);
    int* jY = jX + n;
    HQ_TRY(launch_copy_block(X, m, dA, lda, m, n, h->cur));
    HQ_TRY(hipMemcpyAsync(jX, djpvt, sizeof(int) * n,
                          hipMemcpyDeviceToDevice, h->cur));
    HQ_TRY(launch_classic_norms(X, m, m, n, k0, vX, vX + n, h->cur));
    for (int i = k0; i < mn; ++i) {
      HQ_TRY(launch_classic_fused(X, Y, m, m, n, i, jX, jY, vX, vY, dA, lda,
                                  djpvt, dtau, h->cur));
      std::swap(X, Y);
      std::swap(jX, jY);
      std::swap(vX, vY);
    }
    if (n > mn) {  // m < n: columns m.. are finished in X
      HQ_TRY(launch_copy_block(dA + (long)mn * lda, lda, X + (size_t)mn * m,
                               m, m, n - mn, h->cur));
      HQ_TRY(hipMemcpyAsync(djpvt + mn, jX + mn, sizeof(int) * (n - mn),
                            hipMemcpyDeviceToDevice, h->cur));
    }
    return HQRRP_OK;
  }
  double* vn1 = ws;
  double* vn2 = vn1 + n;
  HQ_TRY(launch_classic_norms(dA, lda, m, n, k0, vn1, vn2, h->cur));
  for (int i = k0; i < mn; ++i)
    HQ_TRY(launch_classic_step(dA, lda, m, n, i, djpvt, dtau, vn1, vn2,
                               h->cur));
  return HQRRP_OK;
}

// The classic QRCP of a whole matrix with the prescaling of factor_locked
// (an exact power of two when max |a| is extreme).
int run_classic_scaled(hqrrp_handle_t h, int m, int n, double* dA, long lda,
                       int* djpvt, double* dtau, double* ws) {
  HQ_TRY(launch_scale_pre(dA, lda, m, n, h->scale_words(), -256, 256, h->cur));
  HQ_RET(run_classic_qrcp(h, m, n, 0, dA, lda, djpvt, dtau, ws));
  HQ_TRY(launch_scale_post(dA, lda, m, n, std::min(m, n), h->scale_words(),
                           h->cur));
  return HQRRP_OK;
}

// Q (m x n) = the first n columns of H_1 ... H_k, the reflectors held in the
// first k columns of A below the diagonal (LAPACK dorgqr, written to Q):
// Q = [I; 0], then block reflectors from the last to the first applied to
// Q(j:m, j:n) -- (I - V T V') C with T rebuilt from V'V (same larft kernel).
int run_orgqr(hqrrp_handle_t h, int m, int n, int k, const double* dA,
              long lda, const double* dtau, double* dQ, long ldq) {
  const int nb = 64;
  HQ_RET(ensure(h->T, sizeof(double) * nb * nb));
  HQ_RET(ensure(h->gram, sizeof(double) * 64 * 64));
  HQ_TRY(launch_fill_eye(dQ, ldq, m, n, 1.0, h->cur));
  if (k <= 0) return HQRRP_OK;
  for (int j = ((k - 1) / nb) * nb; j >= 0; j -= nb) {
    const int b = std::min(nb, k - j), mrows = m - j;
    const double* V = dA + j + (long)j * lda;
    double* gram = h->gram.as<double>();
    HQ_RET(gemm(h, true, false, b, b, mrows, 1.0, V, lda, 1, V, lda, 1, 0.0,
                gram, b));
    HQ_TRY(launch_larft(b, gram, b, dtau + j, h->T.as<double>(), nb, h->cur));
    TrailArgs ta;
    HQ_RET(trail_prepare(h, mrows, n - j, b, V, lda, h->T.as<double>(), nb,
                         dQ + j + (long)j * ldq, ldq, &ta));
    ta.tnotrans = 1;
    HQ_RET(trail_w(h, ta));
    HQ_RET(trail_apply(h, ta));
  }
  return HQRRP_OK;
}

// op(Q) C (left) or C op(Q) (right), Q = H_1 ... H_k from the reflectors in
// the first k columns of A below the diagonal and tau (LAPACK dormqr), by
// block reflectors of 64: Q = Q_1 Q_2 ..., Q_b = I - V T V' with T rebuilt
// from V'V.  Left: Q'C applies Q_1' first (W = T'V'C), QC the last block
// first (W = T V'C), both through the trailing-update kernels on the rows
// the block touches; right: C Q applies Q_1 first and C Q' the last block
// first, C(:, j:) -= ((C(:, j:) V) T or T') V' by generic products.
int run_ormqr(hqrrp_handle_t h, bool left, bool trans, int m, int n, int k,
              const double* dA, long lda, const double* dtau, double* dC,
              long ldc) {
  const int nb = 64;
  HQ_RET(ensure(h->T, sizeof(double) * nb * nb));
  HQ_RET(ensure(h->gram, sizeof(double) * 64 * 64));
  const int nq = left ? m : n;
  const int nblk = (k + nb - 1) / nb;
  // Left Q' and right Q take the blocks first to last.
  const bool forward = left == trans;
  for (int q = 0; q < nblk; ++q) {
    const int j = (forward ? q : nblk - 1 - q) * nb;
    const int b = std::min(nb, k - j), rows = nq - j;
    const double* V = dA + j + (long)j * lda;
    double* gram = h->gram.as<double>();
    HQ_RET(gemm(h, true, false, b, b, rows, 1.0, V, lda, 1, V, lda, 1, 0.0,
                gram, b));
    HQ_TRY(launch_larft(b, gram, b, dtau + j, h->T.as<double>(), nb, h->cur));
    if (left) {
      TrailArgs ta;
      HQ_RET(trail_prepare(h, rows, n, b, V, lda, h->T.as<double>(), nb,
                           dC + j, ldc, &ta));
      ta.tnotrans = trans ? 0 : 1;
      HQ_RET(trail_w(h, ta));
      HQ_RET(trail_apply(h, ta));
    } else {
      HQ_RET(ensure(h->genM, sizeof(double) * 2 * (size_t)m * b));
      double* W = h->genM.as<double>();  // m x b: C(:, j:) V
      double* W2 = W + (size_t)m * b;    // m x b: W T (or W T')
      double* Cj = dC + (long)j * ldc;
      HQ_RET(gemm(h, false, false, m, b, rows, 1.0, Cj, ldc, 0, V, lda, 1, 0.0,
                  W, m));
      HQ_RET(gemm(h, false, trans, m, b, b, 1.0, W, m, 0, h->T.as<double>(),
                  nb, 0, 0.0, W2, m, false));
      HQ_RET(gemm(h, false, true, m, rows, b, -1.0, W2, m, 0, V, lda, 1, 1.0,
                  Cj, ldc, false));
    }
  }
  return HQRRP_OK;
}

// QR of AP = A(:, jpvt-1) truncated to k columns of Q: AP(:, 0:k) = Q R11,
// R = [R11 R12] = Q' AP(:, 0:n) restricted to its first k rows.
int run_qr_fixed(hqrrp_handle_t h, int m, int n, int k, const double* dA,
                 long lda, const int* djpvt, double* dQ, long ldq, double* dR,
                 long ldr) {
  HQ_RET(ensure(h->adjAP, sizeof(double) * (size_t)m * n));
  HQ_RET(ensure(h->adjTau, sizeof(double) * (size_t)std::max(1, k)));
  double* AP = h->adjAP.as<double>();
  HQ_TRY(launch_permute_cols(dA, lda, m, n, djpvt, 0, AP, m, h->cur));
  HQ_RET(run_geqrf(h, m, n, k, AP, m, h->adjTau.as<double>()));
  HQ_TRY(launch_copy_r(AP, m, k, n, dR, ldr, h->cur));
  return run_orgqr(h, m, k, k, AP, m, h->adjTau.as<double>(), dQ, ldq);
}

// Reverse-mode derivative of run_qr_fixed (pivots held fixed).  With
// X = AP(:, :k), Y = AP(:, k:), dQ' = dQ + Y dR12', M = R11 dR11' - dQ'' Q:
//   dX = (dQ' + Q copyltu(M)) R11^-T,   dY = Q dR12,   dA(:, p) = [dX dY].
int run_qr_fixed_adjoint(hqrrp_handle_t h, int m, int n, int k,
                         const double* dA, long lda, const int* djpvt,
                         const double* dQ, long ldq, const double* dR,
                         long ldr, const double* gQ, long ldgq,
                         const double* gR, long ldgr, double* gA, long ldga) {
  HQ_RET(ensure(h->adjAP, sizeof(double) * (size_t)m * n));
  HQ_RET(ensure(h->adjK, sizeof(double) * (size_t)3 * k * k));
  double* G = h->adjAP.as<double>();  // m x n: [Z | Y], later [dX | dY]
  double* Mk = h->adjK.as<double>();  // k x k each: M, copyltu(M), R11^-1
  double* Bk = Mk + (size_t)k * k;
  double* Ri = Bk + (size_t)k * k;
  double* Z = G;
  double* Y = G + (size_t)k * m;
  const int nr = n - k;
  // Z = dQ, Y = AP(:, k:)
  HQ_TRY(launch_copy_block(Z, m, gQ, ldgq, m, k, h->cur));
  if (nr > 0) {
    HQ_TRY(launch_permute_cols(dA, lda, m, nr, djpvt + k, 0, Y, m, h->cur));
    // Z += Y dR12'
    HQ_RET(gemm(h, false, true, m, k, nr, 1.0, Y, m, 0, gR + (long)k * ldgr,
                ldgr, 0, 1.0, Z, m));
  }
  // M = R11 dR11' - Z' Q
  HQ_RET(gemm(h, false, true, k, k, k, 1.0, dR, ldr, 0, gR, ldgr, 0, 0.0, Mk,
              k));
  HQ_RET(gemm(h, true, false, k, k, m, -1.0, Z, m, 0, dQ, ldq, 0, 1.0, Mk, k));
  HQ_TRY(launch_copyltu(k, Mk, k, Bk, k, h->cur));
  // Z += Q copyltu(M)
  HQ_RET(gemm(h, false, false, m, k, k, 1.0, dQ, ldq, 0, Bk, k, 0, 1.0, Z, m));
  // R11^-1: diagonal 64-blocks in LDS, then doubling:
  //   Ri(B1,B2) = -Ri(B1,B1) R(B1,B2) Ri(B2,B2)
  HQ_TRY(launch_fill_eye(Ri, k, k, k, 0.0, h->cur));
  HQ_TRY(launch_tri_inv_diag(k, dR, ldr, Ri, k, h->cur));
  for (int s = tri_inv_block(); s < k; s *= 2) {
    for (int i0 = 0; i0 + s < k; i0 += 2 * s) {
      const int i1 = i0 + s, s2 = std::min(s, k - i1);
      // Bk(0:s, 0:s2) = Ri(B1,B1) R(B1,B2); Ri(B1,B2) = -Bk Ri(B2,B2)
      HQ_RET(gemm(h, false, false, s, s2, s, 1.0, Ri + i0 + (long)i0 * k, k, 0,
                  dR + i0 + (long)i1 * ldr, ldr, 0, 0.0, Bk, k));
      HQ_RET(gemm(h, false, false, s, s2, s2, -1.0, Bk, k, 0,
                  Ri + i1 + (long)i1 * k, k, 0, 0.0, Ri + i0 + (long)i1 * k, k));
    }
  }
  // dY = Q dR12 (over Y), dX = Z R11^-T (then over Z)
  if (nr > 0)
    HQ_RET(gemm(h, false, false, m, nr, k, 1.0, dQ, ldq, 0,
                gR + (long)k * ldgr, ldgr, 0, 0.0, Y, m));
  HQ_RET(ensure(h->adjX, sizeof(double) * (size_t)m * k));
  HQ_RET(gemm(h, false, true, m, k, k, 1.0, Z, m, 0, Ri, k, 0, 0.0,
              h->adjX.as<double>(), m));
  HQ_TRY(launch_copy_block(Z, m, h->adjX.as<double>(), m, m, k, h->cur));
  // dA(:, p[c]-1) = [dX dY](:, c)
  HQ_TRY(launch_permute_cols(G, m, m, n, djpvt, 1, gA, ldga, h->cur));
  return HQRRP_OK;
}

std::mutex g_default_mu;
hqrrp_handle_t g_default[64];

int default_handle(hqrrp_handle_t* out) {
  int dev = 0;
  HQ_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return HQRRP_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_default_mu);
  if (!g_default[dev]) HQ_RET(hqrrp_create(&g_default[dev], dev));
  *out = g_default[dev];
  return HQRRP_OK;
}

}  // namespace

// ------------------------------------------------------------------ C ABI

extern "C" void hqrrp_default_options(hqrrp_options* o) {
  memset(o, 0, sizeof(*o));
  o->nb = 64;
  o->pp = 10;
  o->panel_pivoting = 1;
  o->sketch_prec = HQRRP_SKETCH_FP64;
  o->omega_kind = HQRRP_OMEGA_REFERENCE;
  o->seed = 12;
  o->omega = nullptr;
  o->max_rank = 0;
  o->lookahead = 1;
}

extern "C" int hqrrp_create(hqrrp_handle_t* out, int device) {
  if (!out) return HQRRP_ERR_ARG;
  hqrrp_handle_t h = new (std::nothrow) hqrrp_handle_s();
  if (!h) return HQRRP_ERR_ALLOC;
  h->device = device;
  HQ_TRY(hipSetDevice(device));
  HQ_TRY(hipStreamCreateWithFlags(&h->own, hipStreamNonBlocking));
  h->stream = h->own;
  HQ_TRY(hipDeviceGetAttribute(&h->cu_count, hipDeviceAttributeMultiprocessorCount,
                               device));
  int lds = 0;
  HQ_TRY(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                               device));
  h->lds_max = (size_t)lds;
  // Persistent kernels need all their workgroups co-resident; both kernels
  // are launched at most one workgroup per CU (sketch: register-bound,
  // panel: LDS-bound).
  h->sketch_max_wg = std::min(h->cu_count, 256);
  h->panel_max_wg = h->cu_count;
  // Look-ahead streams: a quarter of the CUs for the panel chain, the rest
  // for the trailing update.  Mask bits are taken in runs of 8 (one run in
  // four), so that whether the bits map to the 8 XCDs interleaved or in
  // contiguous ranges, every XCD contributes the same number of CUs to each
  // stream (workgroups are dealt to the XCDs round-robin).  A panel needing
  // more workgroups than the chain holds runs without look-ahead.
  {
    const int words = (h->cu_count + 31) / 32;
    std::vector<uint32_t> mc(words, 0u), mu(words, 0u);
    int nchain = 0;
    for (int c = 0; c < h->cu_count; ++c) {
      if ((c / 8) % 4 == 0) {
        mc[c / 32] |= 1u << (c % 32);
        ++nchain;
      } else {
        mu[c / 32] |= 1u << (c % 32);
      }
    }
    if (nchain > 0 &&
        hipExtStreamCreateWithCUMask(&h->s_chain, words, mc.data()) ==
            hipSuccess &&
        hipExtStreamCreateWithCUMask(&h->s_upd, words, mu.data()) ==
            hipSuccess) {
      h->chain_cus = nchain;
    } else {
      (void)hipGetLastError();
      if (h->s_chain) (void)hipStreamDestroy(h->s_chain);
      if (h->s_upd) (void)hipStreamDestroy(h->s_upd);
      h->s_chain = h->s_upd = nullptr;
      h->chain_cus = 0;
    }
    HQ_TRY(hipEventCreateWithFlags(&h->ev_in, hipEventDisableTiming));
    HQ_TRY(hipEventCreateWithFlags(&h->ev_out, hipEventDisableTiming));
    if (h->s_chain) {
      // The look-ahead's events only order device work between this
      // device's queues: a device-scope release is enough (no system-scope
      // cache writeback and invalidation at each fork and join).
      const unsigned f = hipEventDisableTiming | hipEventDisableSystemFence;
      HQ_TRY(hipEventCreateWithFlags(&h->ev_fork, f));
      HQ_TRY(hipEventCreateWithFlags(&h->ev_perm, f));
      HQ_TRY(hipEventCreateWithFlags(&h->ev_upd, f));
      HQ_TRY(hipEventCreateWithFlags(&h->ev_chain, f));
      HQ_TRY(hipEventCreateWithFlags(&h->ev_sketch, f));
      HQ_TRY(hipEventCreateWithFlags(&h->ev_t, f));
    }
  }
  int rc = ensure(h->words, 256);
  if (rc) return rc;
  HQ_TRY(hipMemset(h->words.p, 0, h->words.bytes));
  // The allocation-free entry's workspace on this device (looked up once
  // per device, with the device current).
  rc = scratch_address(device, &h->scratch);
  if (rc) return rc;
  *out = h;
  return HQRRP_OK;
}

extern "C" int hqrrp_destroy(hqrrp_handle_t h) {
  if (!h) return HQRRP_OK;
  (void)hipSetDevice(h->device);
  (void)hipStreamSynchronize(h->stream);
  DevBuf* bufs[] = {&h->words,     &h->Y,         &h->G,     &h->Gcache,
                    &h->xs.buf[0], &h->xs.buf[1], &h->xp.buf[0], &h->xp.buf[1],
                    &h->moves,     &h->piv,       &h->gram,
                    &h->T,     &h->perm,  &h->tmpA,  &h->tmpY,   &h->tmpJ,
                    &h->split, &h->Mx,    &h->W2,    &h->Pp,     &h->Qm,
                    &h->hA,    &h->hJ,    &h->hT,     &h->hW,    &h->vn,
                    &h->Pn,    &h->Xb,     &h->adjAP, &h->gen,  &h->cqr,
                    &h->genT,  &h->genM,
                    &h->adjTau, &h->adjK, &h->adjX};
  for (DevBuf* b : bufs)
    if (b->p) (void)hipFree(b->p);
  for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
  for (hipEvent_t e : {h->ev_fork, h->ev_perm, h->ev_upd, h->ev_chain, h->ev_sketch,
                       h->ev_t, h->ev_in, h->ev_out})
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t st : {h->s_chain, h->s_upd})
    if (st) (void)hipStreamDestroy(st);
  if (h->own) (void)hipStreamDestroy(h->own);
  delete h;
  return HQRRP_OK;
}

extern "C" int hqrrp_set_panel_alg(hqrrp_handle_t h, int alg) {
  if (!h || alg < 0 || alg > 1) return HQRRP_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  h->panel_alg = alg;
  return HQRRP_OK;
}

extern "C" int hqrrp_phase_count(void) { return HQ_PHASES; }

extern "C" const char* hqrrp_phase_name(int phase) {
  return (phase >= 0 && phase < HQ_PHASES) ? kPhaseNames[phase] : "";
}

extern "C" int hqrrp_profile_enable(hqrrp_handle_t h, int on) {
  if (!h) return HQRRP_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  HQ_TRY(hipStreamSynchronize(h->stream));
  h->prof = on != 0;
  h->spans.clear();
  h->ev_used = 0;
  for (int p = 0; p < HQ_PHASES; ++p) {
    h->ph_flops[p] = h->ph_bytes[p] = 0.0;
    h->ph_launch[p] = 0;
  }
  return HQRRP_OK;
}

extern "C" int hqrrp_profile_phases(hqrrp_handle_t h, unsigned mask) {
  if (!h) return HQRRP_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  h->prof_mask = mask;
  return HQRRP_OK;
}

extern "C" int hqrrp_profile_read(hqrrp_handle_t h, int nphase, double* ms,
                                  long long* launches, double* flops,
                                  double* bytes) {
  if (!h || nphase < HQ_PHASES) return HQRRP_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  HQ_TRY(hipStreamSynchronize(h->stream));
  for (int p = 0; p < HQ_PHASES; ++p) ms[p] = 0.0;
  for (const auto& sp : h->spans) {
    float t = 0.f;
    HQ_TRY(hipEventElapsedTime(&t, sp.a, sp.b));
    ms[sp.phase] += t;
  }
  for (int p = 0; p < HQ_PHASES; ++p) {
    launches[p] = h->ph_launch[p];
    flops[p] = h->ph_flops[p];
    bytes[p] = h->ph_bytes[p];
  }
  return HQRRP_OK;
}

extern "C" int hqrrp_set_stream(hqrrp_handle_t h, void* stream) {
  if (!h) return HQRRP_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  h->stream = (hipStream_t)stream;
  return HQRRP_OK;
}

extern "C" int hqrrp_reset_stream(hqrrp_handle_t h) {
  if (!h) return HQRRP_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  h->stream = h->own;
  return HQRRP_OK;
}

extern "C" int hqrrp_synchronize(hqrrp_handle_t h) {
  if (!h) return HQRRP_ERR_ARG;
  return check_err(h);
}

extern "C" int hqrrp_factor_dev(hqrrp_handle_t h, int m, int n, double* dA,
                                int lda, int* djpvt, double* dtau,
                                const hqrrp_options* opt) {
  if (!h || m < 0 || n < 0 || lda < std::max(1, m)) return HQRRP_ERR_ARG;
  hqrrp_options o;
  if (opt)
    o = *opt;
  else
    hqrrp_default_options(&o);
  if (o.nb < 1 || o.pp < 0) return HQRRP_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  return factor_locked(h, m, n, dA, lda, djpvt, dtau, o);
}

// The libflame UT-variant front end (FLA_HQRRP_UT_blk_var2,
// libflame_sources/FLA_HQRRP_UT_blk_var2.c:124-401): the same algorithm,
// its Householder transforms in libflame's UT form.  The factorisation runs
// as HQRRP (WY form) and is then rewritten: s = 1/tau, and each step where
// nothing remained below the diagonal (tau = 0 here; libflame's
// FLA_Househ2_UT then negates alpha and sets tau = 1/2, the reflection
// I - 2 e e') has its row of R negated.  Same pivots, same Householder
// vectors below the diagonal.
int ut_rewrite(hqrrp_handle_t h, int m, int n, double* dA, long lda,
               double* ds, const hqrrp_options& o) {
  const int mn = std::min(m, n);
  if (mn == 0) return HQRRP_OK;
  const int kmax = (o.max_rank > 0) ? std::min(mn, o.max_rank) : mn;
  HQ_TRY(launch_ut_convert(dA, lda, m, n, factored_columns(m, n, o.nb, kmax),
                           ds, h->stream));
  return HQRRP_OK;
}

extern "C" int hqrrp_ut_factor_dev(hqrrp_handle_t h, int m, int n, double* dA,
                                   int lda, int* dp, double* ds,
                                   const hqrrp_options* opt) {
  if (!h || m < 0 || n < 0 || lda < std::max(1, m)) return HQRRP_ERR_ARG;
  hqrrp_options o;
  if (opt)
    o = *opt;
  else
    hqrrp_default_options(&o);
  if (o.nb < 1 || o.pp < 0) return HQRRP_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  HQ_RET(factor_locked(h, m, n, dA, lda, dp, ds, o));
  return ut_rewrite(h, m, n, dA, lda, ds, o);
}

extern "C" int hqrrp_sketch_qrcp_dev(hqrrp_handle_t h, int r, int ncols,
                                     const double* dY, int ldy, int stages,
                                     int* dpiv, int* dmoves) {
  if (!h || r < 1 || ncols < 1 || ldy < r || stages < 1 ||
      stages > std::min(r, ncols))
    return HQRRP_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  h->cur = h->stream;
  return run_sketch(h, r, ncols, dY, ldy, stages, dpiv, dmoves);
}

extern "C" int hqrrp_panel_qrcp_dev(hqrrp_handle_t h, int mrows, int b,
                                    double* dP, int ldp, int pivoting,
                                    double* dtau, int* dperm, double* dT,
                                    int ldt) {
  if (!h || b < 1 || mrows < b || ldp < mrows || ldt < b) return HQRRP_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  h->cur = h->stream;
  return run_panel(h, mrows, b, dP, ldp, pivoting, dtau, dperm, dT, ldt);
}

extern "C" int hqrrp_apply_qt_left_dev(hqrrp_handle_t h, int mrows, int ncols,
                                       int b, const double* dV, int ldv,
                                       const double* dT, int ldt, double* dC,
                                       int ldc) {
  if (!h || b < 1 || mrows < b || ncols < 0 || ldv < mrows || ldc < mrows ||
      ldt < b)
    return HQRRP_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  h->cur = h->stream;
  return apply_qt_left(h, mrows, ncols, b, dV, ldv, dT, ldt, dC, ldc);
}

extern "C" int hqrrp_downdate_y_dev(hqrrp_handle_t h, int r, int b, int mU,
                                    int nY2, const double* dU, int ldu,
                                    const double* dR12, int ldr,
                                    const double* dT, int ldt, double* dY2,
                                    int ldy, double* dG, int ldg) {
  if (!h || r < 1 || b < 1 || mU < b || nY2 < 0 || ldu < mU || ldr < b ||
      ldt < b || ldy < r || ldg < r)
    return HQRRP_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  h->cur = h->stream;
  return downdate_y(h, r, b, mU, nY2, dU, ldu, dR12, ldr, dT, ldt, dY2, ldy,
                    dG, ldg);
}

extern "C" int hqrrp_dgemm_dev(hqrrp_handle_t h, int ta, int tb, int M, int N,
                               int K, double alpha, const double* dA, int lda,
                               const double* dB, int ldb, double beta,
                               double* dC, int ldc) {
  if (!h || M < 0 || N < 0 || K < 0 || ldc < std::max(1, M)) return HQRRP_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  h->cur = h->stream;
  return gemm(h, ta != 0, tb != 0, M, N, K, alpha, dA, lda, 0, dB, ldb, 0,
              beta, dC, ldc);
}

extern "C" int hqrrp_geqrf_dev(hqrrp_handle_t h, int m, int n, double* dA,
                               int lda, double* dtau) {
  if (!h || m < 0 || n < 0 || lda < std::max(1, m)) return HQRRP_ERR_ARG;
  if (std::min(m, n) == 0) return HQRRP_OK;
  std::lock_guard<std::mutex> lk(h->mu);
  StreamScope ss(h);
  return run_geqrf(h, m, n, std::min(m, n), dA, lda, dtau);
}

extern "C" int hqrrp_geqp3_dev(hqrrp_handle_t h, int m, int n, double* dA,
                               int lda, int* djpvt, double* dtau) {
  if (!h || m < 0 || n < 0 || lda < std::max(1, m)) return HQRRP_ERR_ARG;
  if (std::min(m, n) == 0) return HQRRP_OK;
  std::lock_guard<std::mutex> lk(h->mu);
  StreamScope ss(h);
  return run_classic_scaled(h, m, n, dA, lda, djpvt, dtau, nullptr);
}

extern "C" int hqrrp_geqp3_noalloc_dev(hqrrp_handle_t h, int m, int n,
                                       double* dA, int lda, int* djpvt,
                                       double* dtau) {
  if (!h || m < 0 || n < 0 || lda < std::max(1, m)) return HQRRP_ERR_ARG;
  if (std::min(m, n) == 0) return HQRRP_OK;
  const size_t need = classic_ws_bytes(m, n);
  if (need > kScratchBytes) {
    fprintf(stderr,
            "hqrrp_geqp3_noalloc_dev: a %d x %d matrix needs %zu bytes of "
            "workspace, more than the %zu of the static scratch\n",
            m, n, need, kScratchBytes);
    return HQRRP_ERR_UNSUPPORTED;
  }
  if (!h->scratch) return HQRRP_ERR_HIP;
  std::lock_guard<std::mutex> lk(h->mu);
  StreamScope ss(h);
  return run_classic_scaled(h, m, n, dA, lda, djpvt, dtau,
                            static_cast<double*>(h->scratch));
}

extern "C" int hqrrp_orgqr_dev(hqrrp_handle_t h, int m, int n, int k,
                               const double* dA, int lda, const double* dtau,
                               double* dQ, int ldq) {
  if (!h || m < 0 || n < 0 || n > m || k < 0 || k > n ||
      lda < std::max(1, m) || ldq < std::max(1, m))
    return HQRRP_ERR_ARG;
  if (n == 0) return HQRRP_OK;
  std::lock_guard<std::mutex> lk(h->mu);
  StreamScope ss(h);
  return run_orgqr(h, m, n, k, dA, lda, dtau, dQ, ldq);
}

extern "C" int hqrrp_ormqr_dev(hqrrp_handle_t h, char side, char trans,
                               int m, int n, int k, const double* dA, int lda,
                               const double* dtau, double* dC, int ldc) {
  const bool left = side == 'L' || side == 'l';
  const bool tr = trans == 'T' || trans == 't';
  const int nq = left ? m : n;
  if (!h || !(left || side == 'R' || side == 'r') ||
      !(tr || trans == 'N' || trans == 'n') || m < 0 || n < 0 || k < 0 ||
      k > nq || lda < std::max(1, nq) || ldc < std::max(1, m))
    return HQRRP_ERR_ARG;
  if (m == 0 || n == 0 || k == 0) return HQRRP_OK;
  std::lock_guard<std::mutex> lk(h->mu);
  StreamScope ss(h);
  return run_ormqr(h, left, tr, m, n, k, dA, lda, dtau, dC, ldc);
}

extern "C" int hqrrp_qr_fixed_dev(hqrrp_handle_t h, int m, int n, int k,
                                  const double* dA, int lda, const int* djpvt,
                                  double* dQ, int ldq, double* dR, int ldr) {
  if (!h || k < 1 || n < k || m < k || lda < m || ldq < m || ldr < k)
    return HQRRP_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  StreamScope ss(h);
  return run_qr_fixed(h, m, n, k, dA, lda, djpvt, dQ, ldq, dR, ldr);
}

extern "C" int hqrrp_qr_fixed_adjoint_dev(
    hqrrp_handle_t h, int m, int n, int k, const double* dA, int lda,
    const int* djpvt, const double* dQ, int ldq, const double* dR, int ldr,
    const double* gQ, int ldgq, const double* gR, int ldgr, double* gA,
    int ldga) {
  if (!h || k < 1 || n < k || m < k || lda < m || ldq < m || ldr < k ||
      ldgq < m || ldgr < k || ldga < m)
    return HQRRP_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  StreamScope ss(h);
  return run_qr_fixed_adjoint(h, m, n, k, dA, lda, djpvt, dQ, ldq, dR, ldr, gQ,
                              ldgq, gR, ldgr, gA, ldga);
}

// ------------------------------------------------------------ batched path

// A prepared batch of small matrices (shapes, block sizes, queue order,
// shared Omega, per-workgroup workspace).  Factorisations reuse it.
struct hqrrp_batch_s {
  int device = 0;
  int count = 0, pp = 8, pivoting = 1;
  // Items are grouped by kernel variant (block sizes up to 16, 32, 64):
  // class c holds items [first[c], first[c] + cnt[c]), launched with wgs[c]
  // workgroups on queue counter c.
  int first[3] = {0, 0, 0}, cnt[3] = {0, 0, 0}, wgs[3] = {0, 0, 0};
  std::vector<int> m, n, lda, nb;
  DevBuf items, omega, work, counter;
  DevBuf prof;  // per-phase cycle sums when profiling is on
  bool prof_on = false;
};

namespace {

// Block size of the batched path for an m x n matrix when the caller gives
// none: the handle's tuned table (hqrrp_batch_tune) if it holds an entry for
// the shape's 64 x 64 bucket, else 16 below 64 rows or columns and 32 above.
int batch_pick_nb(hqrrp_handle_t h, int pp, int pivoting, int m, int n) {
  const int bm = std::max(0, std::min(7, (m - 1) / 64));
  const int bn = std::max(0, std::min(7, (n - 1) / 64));
  if (h->bt_pp == pp && h->bt_piv == pivoting && h->bt_nb[bm][bn] > 0)
    return h->bt_nb[bm][bn];
  return std::min(m, n) < 64 ? 16 : 32;
}

// Relative cost of one matrix for the queue order (largest first): the
// factorisation's flops plus a latency term per pivot stage.
double batch_cost(int m, int n) {
  const double a = std::max(m, n), b = std::min(m, n);
  return 2.0 * a * b * b - 2.0 / 3.0 * b * b * b + 4.0e5 * b;
}

}  // namespace

extern "C" int hqrrp_batch_create(hqrrp_handle_t h, int count, const int* m,
                                  const int* n, const int* lda, const int* nb,
                                  int pp, int panel_pivoting,
                                  hqrrp_batch_t* out) {
  if (!h || !out || count < 0 || pp < 0 || pp > HQB_PP_MAX) return HQRRP_ERR_ARG;
  if (count > 0 && (!m || !n)) return HQRRP_ERR_ARG;
  *out = nullptr;
  hqrrp_batch_t b = new (std::nothrow) hqrrp_batch_s();
  if (!b) return HQRRP_ERR_ALLOC;
  std::unique_ptr<hqrrp_batch_s> guard(b);
  b->device = h->device;
  b->count = count;
  b->pp = pp;
  b->pivoting = panel_pivoting ? 1 : 0;
  b->m.assign(m, m + count);
  b->n.assign(n, n + count);
  b->lda.resize(count);
  b->nb.resize(count);
  long omega_len = 1;
  for (int i = 0; i < count; ++i) {
    const int mi = m[i], ni = n[i];
    if (mi < 0 || ni < 0 || mi > HQB_DIM_MAX || ni > HQB_DIM_MAX)
      return HQRRP_ERR_UNSUPPORTED;
    b->lda[i] = lda ? lda[i] : std::max(1, mi);
    if (b->lda[i] < std::max(1, mi)) return HQRRP_ERR_ARG;
    b->nb[i] = nb ? nb[i] : batch_pick_nb(h, pp, b->pivoting, mi, ni);
    if (b->nb[i] < 1 || b->nb[i] > 64) return HQRRP_ERR_UNSUPPORTED;
    omega_len = std::max(omega_len, (long)(b->nb[i] + pp) * std::max(1, mi));
  }
  auto cls = [&](int i) {
    const int c = batch_class_nb(b->nb[i]);
    return c == 16 ? 0 : (c == 32 ? 1 : 2);
  };
  std::vector<int> order(count);
  for (int i = 0; i < count; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    if (cls(x) != cls(y)) return cls(x) < cls(y);
    return batch_cost(b->m[x], b->n[x]) > batch_cost(b->m[y], b->n[y]);
  });
  for (int q = 0; q < count; ++q) ++b->cnt[cls(order[q])];
  b->first[1] = b->cnt[0];
  b->first[2] = b->cnt[0] + b->cnt[1];
  int max_wgs = 1;
  for (int c = 0; c < 3; ++c) {
    b->wgs[c] = std::min(b->cnt[c], batch_wgs_per_cu(16 << c) * h->cu_count);
    max_wgs = std::max(max_wgs, b->wgs[c]);
  }
  std::vector<BatchItem> items(std::max(1, count));
  for (int q = 0; q < count; ++q) {
    const int i = order[q];
    BatchItem& it = items[q];
    memset(&it, 0, sizeof(it));
    it.m = b->m[i];
    it.n = b->n[i];
    it.lda = b->lda[i];
    it.nb = b->nb[i];
    it.idx = i;
  }
  HQ_TRY(hipSetDevice(h->device));
  HQ_RET(ensure(b->items, sizeof(BatchItem) * items.size()));
  HQ_TRY(hipMemcpy(b->items.p, items.data(), sizeof(BatchItem) * items.size(),
                   hipMemcpyHostToDevice));
  // Omega of an r x m matrix is the first r*m normals of the reference's
  // sequence (seed 12) in column order: one sequence serves every shape.
  std::vector<double> om(omega_len);
  hqrrp_reference_omega(1, (int)omega_len, om.data(), 12u);
  HQ_RET(ensure(b->omega, sizeof(double) * om.size()));
  HQ_TRY(hipMemcpy(b->omega.p, om.data(), sizeof(double) * om.size(),
                   hipMemcpyHostToDevice));
  HQ_RET(ensure(b->work, sizeof(double) * (size_t)batch_work_doubles() * max_wgs));
  HQ_TRY(hipMemset(b->work.p, 0, b->work.bytes));
  HQ_RET(ensure(b->counter, 256));
  HQ_TRY(hipMemset(b->counter.p, 0, b->counter.bytes));
  *out = guard.release();
  return HQRRP_OK;
}

extern "C" int hqrrp_batch_get_nb(hqrrp_batch_t b, int* nb) {
  if (!b || (!nb && b->count > 0)) return HQRRP_ERR_ARG;
  for (int i = 0; i < b->count; ++i) nb[i] = b->nb[i];
  return HQRRP_OK;
}

extern "C" int hqrrp_batch_factor_dev(hqrrp_handle_t h, hqrrp_batch_t b,
                                      double* const* dA, int* const* djpvt,
                                      double* const* dtau) {
  if (!h || !b || b->device != h->device) return HQRRP_ERR_ARG;
  if (b->count == 0) return HQRRP_OK;
  if (!dA || !djpvt || !dtau) return HQRRP_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  StreamScope ss(h);
  HQ_TRY(launch_batch_reset(b->counter.as<int>(), 3, h->cur));
  for (int c = 0; c < 3; ++c) {
    if (b->cnt[c] == 0) continue;
    BatchArgs a{};
    a.items = b->items.as<BatchItem>() + b->first[c];
    a.count = b->cnt[c];
    a.pp = b->pp;
    a.pivoting = b->pivoting;
    a.A = dA;
    a.jpvt = djpvt;
    a.tau = dtau;
    a.omega = b->omega.as<double>();
    a.work = b->work.as<double>();
    a.counter = b->counter.as<int>() + c;
    a.prof = b->prof_on ? b->prof.as<unsigned long long>() : nullptr;
    HQ_TRY(launch_batch_hqrrp(a, 16 << c, b->wgs[c], h->cur));
  }
  return HQRRP_OK;
}

// Autotuner of the batched path's block size: for every 64 x 64 shape
// bucket, a batch of its midpoint shape (hash-filled) is factored with each
// candidate block size and the fastest kept in the handle's table.
extern "C" int hqrrp_batch_tune(hqrrp_handle_t h, int pp, int panel_pivoting) {
  if (!h || pp < 0 || pp > HQB_PP_MAX) return HQRRP_ERR_ARG;
  HQ_TRY(hipSetDevice(h->device));
  const int count = std::max(256, 4 * h->cu_count);
  const int cand[3] = {16, 32, 64};
  const long maxsz = (long)HQB_DIM_MAX * HQB_DIM_MAX;
  DevBuf A, J, T, P;
  HQ_RET(ensure(A, sizeof(double) * (size_t)count * maxsz));
  HQ_RET(ensure(J, sizeof(int) * (size_t)count * HQB_DIM_MAX));
  HQ_RET(ensure(T, sizeof(double) * (size_t)count * HQB_DIM_MAX));
  HQ_RET(ensure(P, sizeof(void*) * 3 * (size_t)count));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HQ_TRY(hipEventCreate(&e0));
  HQ_TRY(hipEventCreate(&e1));
  hipStream_t s = h->own;
  int rc = HQRRP_OK;
  int table[8][8] = {};
  std::vector<void*> ptrs(3 * (size_t)count);
  std::vector<int> jinit((size_t)count * HQB_DIM_MAX);
  for (int bm = 0; bm < 8 && rc == HQRRP_OK; ++bm) {
    for (int bn = 0; bn < 8 && rc == HQRRP_OK; ++bn) {
      const int m = 64 * bm + 32, n = 64 * bn + 32;
      for (int i = 0; i < count; ++i) {
        ptrs[i] = A.as<double>() + (size_t)i * m * n;
        ptrs[count + i] = J.as<int>() + (size_t)i * n;
        ptrs[2 * count + i] = T.as<double>() + (size_t)i * std::min(m, n);
      }
      if (hipMemcpy(P.p, ptrs.data(), sizeof(void*) * ptrs.size(),
                    hipMemcpyHostToDevice) != hipSuccess) {
        rc = HQRRP_ERR_HIP;
        break;
      }
      for (int i = 0; i < count; ++i)
        for (int c = 0; c < n; ++c) jinit[(size_t)i * n + c] = c + 1;
      std::vector<int> vm(count, m), vn(count, n);
      float best = 0.0f;
      int best_nb = 0;
      for (int ci = 0; ci < 3 && rc == HQRRP_OK; ++ci) {
        std::vector<int> vnb(count, cand[ci]);
        hqrrp_batch_t plan = nullptr;
        rc = hqrrp_batch_create(h, count, vm.data(), vn.data(), nullptr,
                                vnb.data(), pp, panel_pivoting, &plan);
        if (rc != HQRRP_OK) break;
        float ms_min = 0.0f;
        for (int rep = 0; rep < 3 && rc == HQRRP_OK; ++rep) {
          if (launch_batch_fill(A.as<double>(), (long)count * m * n,
                                1234u + rep, s) != hipSuccess ||
              hipMemcpyAsync(J.p, jinit.data(), sizeof(int) * (size_t)count * n,
                             hipMemcpyHostToDevice, s) != hipSuccess ||
              hipEventRecord(e0, s) != hipSuccess) {
            rc = HQRRP_ERR_HIP;
            break;
          }
          {
            std::lock_guard<std::mutex> lk(h->mu);
            h->cur = s;
            if (launch_batch_reset(plan->counter.as<int>(), 3, s) != hipSuccess)
              rc = HQRRP_ERR_HIP;
            for (int c = 0; c < 3 && rc == HQRRP_OK; ++c) {
              if (plan->cnt[c] == 0) continue;
              BatchArgs a{};
              a.items = plan->items.as<BatchItem>() + plan->first[c];
              a.count = plan->cnt[c];
              a.pp = plan->pp;
              a.pivoting = plan->pivoting;
              a.A = P.as<double* const>();
              a.jpvt = reinterpret_cast<int* const*>(P.as<void*>() + count);
              a.tau = reinterpret_cast<double* const*>(P.as<void*>() + 2 * count);
              a.omega = plan->omega.as<double>();
              a.work = plan->work.as<double>();
              a.counter = plan->counter.as<int>() + c;
              if (launch_batch_hqrrp(a, 16 << c, plan->wgs[c], s) != hipSuccess)
                rc = HQRRP_ERR_HIP;
            }
            h->cur = h->stream;
          }
          float ms = 0.0f;
          if (rc != HQRRP_OK || hipEventRecord(e1, s) != hipSuccess ||
              hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = rc ? rc : HQRRP_ERR_HIP;
            break;
          }
          if (rep > 0 && (ms_min == 0.0f || ms < ms_min)) ms_min = ms;
        }
        hqrrp_batch_destroy(plan);
        if (rc == HQRRP_OK && (best_nb == 0 || ms_min < best)) {
          best = ms_min;
          best_nb = cand[ci];
        }
      }
      table[bm][bn] = best_nb;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (DevBuf* d : {&A, &J, &T, &P})
    if (d->p) (void)hipFree(d->p);
  if (rc != HQRRP_OK) return rc;
  memcpy(h->bt_nb, table, sizeof(table));
  h->bt_pp = pp;
  h->bt_piv = panel_pivoting ? 1 : 0;
  return HQRRP_OK;
}

// The tuned block sizes (64 ints, [m bucket][n bucket], 0 = not tuned).
extern "C" int hqrrp_batch_tune_table(hqrrp_handle_t h, int* table) {
  if (!h || !table) return HQRRP_ERR_ARG;
  memcpy(table, h->bt_nb, sizeof(h->bt_nb));
  return HQRRP_OK;
}

static const char* const kBatchPhaseNames[HQB_PHASES] = {
    "idle", "sketch_gemm", "sketch_qrcp", "panel", "t_factor",
    "trailing_update", "downdate", "finish_permutation"};

extern "C" int hqrrp_batch_phase_count(void) { return HQB_PHASES; }

extern "C" const char* hqrrp_batch_phase_name(int phase) {
  return (phase >= 0 && phase < HQB_PHASES) ? kBatchPhaseNames[phase] : "";
}

extern "C" int hqrrp_batch_profile(hqrrp_batch_t b, int on) {
  if (!b) return HQRRP_ERR_ARG;
  HQ_TRY(hipSetDevice(b->device));
  if (on) {
    HQ_RET(ensure(b->prof, sizeof(unsigned long long) * HQB_PHASES));
    HQ_TRY(hipMemset(b->prof.p, 0, b->prof.bytes));
  }
  b->prof_on = on != 0;
  return HQRRP_OK;
}

extern "C" int hqrrp_batch_profile_read(hqrrp_batch_t b, int nphase,
                                        double* cycles) {
  if (!b || nphase < HQB_PHASES || !cycles) return HQRRP_ERR_ARG;
  unsigned long long v[HQB_PHASES] = {};
  if (b->prof.p) {
    HQ_TRY(hipSetDevice(b->device));
    HQ_TRY(hipDeviceSynchronize());
    HQ_TRY(hipMemcpy(v, b->prof.p, sizeof(v), hipMemcpyDeviceToHost));
  }
  for (int i = 0; i < HQB_PHASES; ++i) cycles[i] = (double)v[i];
  return HQRRP_OK;
}

extern "C" int hqrrp_batch_destroy(hqrrp_batch_t b) {
  if (!b) return HQRRP_OK;
  (void)hipSetDevice(b->device);
  for (DevBuf* d : {&b->items, &b->omega, &b->work, &b->counter, &b->prof})
    if (d->p) (void)hipFree(d->p);
  delete b;
  return HQRRP_OK;
}

// --------------------------------------------------- drop-in host entries

namespace {

// Host-pointer HQRRP on the default handle of the current device.
// ut: libflame's UT form of the result (hqrrp_ut_blk_var2).
int host_hqrrp(int m, int n, double* A, int lda, int* jpvt, double* tau,
               const hqrrp_options& o, bool ut = false) {
  hqrrp_handle_t h = nullptr;
  HQ_RET(default_handle(&h));
  std::lock_guard<std::mutex> lk(h->mu);
  const int mn = std::min(m, n);
  HQ_RET(ensure(h->hA, sizeof(double) * (size_t)std::max(1, m) * n));
  HQ_RET(ensure(h->hJ, sizeof(int) * (size_t)std::max(1, n)));
  HQ_RET(ensure(h->hT, sizeof(double) * (size_t)std::max(1, mn)));
  double* dA = h->hA.as<double>();
  HQ_TRY(hipMemcpy2DAsync(dA, sizeof(double) * m, A, sizeof(double) * lda,
                          sizeof(double) * m, n, hipMemcpyHostToDevice,
                          h->stream));
  HQ_TRY(hipMemcpyAsync(h->hJ.p, jpvt, sizeof(int) * n, hipMemcpyHostToDevice,
                        h->stream));
  HQ_RET(factor_locked(h, m, n, dA, m, h->hJ.as<int>(), h->hT.as<double>(), o));
  if (ut) HQ_RET(ut_rewrite(h, m, n, dA, m, h->hT.as<double>(), o));
  HQ_RET(check_err(h));
  HQ_TRY(hipMemcpy2DAsync(A, sizeof(double) * lda, dA, sizeof(double) * m,
                          sizeof(double) * m, n, hipMemcpyDeviceToHost,
                          h->stream));
  HQ_TRY(hipMemcpyAsync(jpvt, h->hJ.p, sizeof(int) * n, hipMemcpyDeviceToHost,
                        h->stream));
  HQ_TRY(hipMemcpyAsync(tau, h->hT.p, sizeof(double) * mn,
                        hipMemcpyDeviceToHost, h->stream));
  HQ_TRY(hipStreamSynchronize(h->stream));
  return HQRRP_OK;
}

// dgeqp4 with nf > 0 fixed columns already moved to the front
// (NoFLA_HQRRP_WY_blk_var4.c:220-292): QR of the fixed columns with the
// reflectors applied to the rest (the reference's dgeqrf + dormqr), HQRRP of
// the free block A(nf:, nf:) permuting jpvt(nf:), then the rows above the
// free block follow its pivots -- one gather by the inverse of the previous
// jpvt instead of the reference's O(n^2) search (same columns end up in the
// same places).
int host_dgeqp4_fixed(int m, int n, double* A, int lda, int* jpvt, double* tau,
                      int nf, const hqrrp_options& o) {
  hqrrp_handle_t h = nullptr;
  HQ_RET(default_handle(&h));
  std::lock_guard<std::mutex> lk(h->mu);
  const int mn = std::min(m, n);
  HQ_RET(ensure(h->hA, sizeof(double) * (size_t)std::max(1, m) * n));
  HQ_RET(ensure(h->hJ, sizeof(int) * (size_t)std::max(1, n)));
  HQ_RET(ensure(h->hT, sizeof(double) * (size_t)std::max(1, mn)));
  double* dA = h->hA.as<double>();
  int* dJ = h->hJ.as<int>();
  double* dT = h->hT.as<double>();
  h->cur = h->stream;
  HQ_TRY(hipMemcpy2DAsync(dA, sizeof(double) * m, A, sizeof(double) * lda,
                          sizeof(double) * m, n, hipMemcpyHostToDevice,
                          h->stream));
  HQ_TRY(hipMemcpyAsync(dJ, jpvt, sizeof(int) * n, hipMemcpyHostToDevice,
                        h->stream));
  // The fixed columns' QR on the prescaled matrix (the free block's own
  // prescaling inside factor_locked then finds nothing to do); R scaled back
  // at the end.
  HQ_TRY(launch_scale_pre(dA, m, m, n, h->scale_words(1), -256, 256,
                          h->cur));
  HQ_RET(run_geqrf(h, m, n, nf, dA, m, dT));
  if (nf < mn) {
    std::vector<int> prev(jpvt, jpvt + n);
    HQ_RET(factor_locked(h, m - nf, n - nf, dA + nf + (long)nf * m, m, dJ + nf,
                         dT + nf, o));
    HQ_RET(check_err(h));
    HQ_TRY(hipMemcpyAsync(jpvt, dJ, sizeof(int) * n, hipMemcpyDeviceToHost,
                          h->stream));
    HQ_TRY(hipStreamSynchronize(h->stream));
    // Column j >= nf of the top block comes from the position that held
    // column id jpvt[j] before the HQRRP.
    std::vector<int> pos(n + 1, -1), src(n - nf);
    for (int k = nf; k < n; ++k) pos[prev[k]] = k;
    for (int j = nf; j < n; ++j) {
      const int v = jpvt[j];
      if (v < 1 || v > n || pos[v] < nf) return HQRRP_ERR_KERNEL;
      src[j - nf] = pos[v] - nf + 1;
    }
    HQ_RET(ensure(h->hW, sizeof(double) * (size_t)nf * (n - nf) +
                             sizeof(int) * (size_t)(n - nf)));
    double* top = h->hW.as<double>();
    int* dsrc = reinterpret_cast<int*>(top + (size_t)nf * (n - nf));
    HQ_TRY(hipMemcpyAsync(dsrc, src.data(), sizeof(int) * (n - nf),
                          hipMemcpyHostToDevice, h->stream));
    HQ_TRY(launch_permute_cols(dA + (long)nf * m, m, nf, n - nf, dsrc, 0, top,
                               nf, h->stream));
    HQ_TRY(launch_copy_block(dA + (long)nf * m, m, top, nf, nf, n - nf,
                             h->stream));
  }
  HQ_TRY(launch_scale_post(dA, m, m, n, mn, h->scale_words(1), h->stream));
  HQ_RET(check_err(h));
  HQ_TRY(hipMemcpy2DAsync(A, sizeof(double) * lda, dA, sizeof(double) * m,
                          sizeof(double) * m, n, hipMemcpyDeviceToHost,
                          h->stream));
  HQ_TRY(hipMemcpyAsync(tau, dT, sizeof(double) * mn, hipMemcpyDeviceToHost,
                        h->stream));
  HQ_TRY(hipStreamSynchronize(h->stream));
  return HQRRP_OK;
}

// dgeqp4 for min(m,n) < 1000, where the reference calls LAPACK dgeqpf or
// dgeqp3 (NoFLA_HQRRP_WY_blk_var4.c:189-199) with the fixed columns already
// at the front: their QR with the reflectors applied to the rest (dgeqpf's
// dgeqr2 + dorm2r), then the classic QRCP of the free columns.
int host_dgeqp4_classic(int m, int n, double* A, int lda, int* jpvt,
                        double* tau, int nf) {
  hqrrp_handle_t h = nullptr;
  HQ_RET(default_handle(&h));
  std::lock_guard<std::mutex> lk(h->mu);
  const int mn = std::min(m, n);
  HQ_RET(ensure(h->hA, sizeof(double) * (size_t)std::max(1, m) * n));
  HQ_RET(ensure(h->hJ, sizeof(int) * (size_t)std::max(1, n)));
  HQ_RET(ensure(h->hT, sizeof(double) * (size_t)std::max(1, mn)));
  double* dA = h->hA.as<double>();
  int* dJ = h->hJ.as<int>();
  double* dT = h->hT.as<double>();
  h->cur = h->stream;
  HQ_TRY(hipMemcpy2DAsync(dA, sizeof(double) * m, A, sizeof(double) * lda,
                          sizeof(double) * m, n, hipMemcpyHostToDevice,
                          h->stream));
  HQ_TRY(hipMemcpyAsync(dJ, jpvt, sizeof(int) * n, hipMemcpyHostToDevice,
                        h->stream));
  HQ_TRY(launch_scale_pre(dA, m, m, n, h->scale_words(), -256, 256, h->cur));
  if (nf > 0) HQ_RET(run_geqrf(h, m, n, nf, dA, m, dT));
  HQ_RET(run_classic_qrcp(h, m, n, nf, dA, m, dJ, dT));
  HQ_TRY(launch_scale_post(dA, m, m, n, mn, h->scale_words(), h->cur));
  HQ_RET(check_err(h));
  HQ_TRY(hipMemcpy2DAsync(A, sizeof(double) * lda, dA, sizeof(double) * m,
                          sizeof(double) * m, n, hipMemcpyDeviceToHost,
                          h->stream));
  HQ_TRY(hipMemcpyAsync(jpvt, dJ, sizeof(int) * n, hipMemcpyDeviceToHost,
                        h->stream));
  HQ_TRY(hipMemcpyAsync(tau, dT, sizeof(double) * mn, hipMemcpyDeviceToHost,
                        h->stream));
  HQ_TRY(hipStreamSynchronize(h->stream));
  return HQRRP_OK;
}

}  // namespace

extern "C" int NoFLA_HQRRP_WY_blk_var4(int m_A, int n_A, double* buff_A,
                                       int ldim_A, int* buff_jpvt,
                                       double* buff_tau, int nb_alg, int pp,
                                       int panel_pivoting) {
  // Argument messages as in NoFLA_HQRRP_WY_blk_var4.c:358-367; unlike the
  // reference, an illegal argument returns an error instead of continuing.
  int bad = 0;
  if (m_A < 0) {
    fprintf(stderr, "ERROR in NoFLA_HQRRP_WY_blk_var4: m_A is < 0.\n");
    bad = 1;
  }
  if (n_A < 0) {
    fprintf(stderr, "ERROR in NoFLA_HQRRP_WY_blk_var4: n_A is < 0.\n");
    bad = 1;
  }
  if (ldim_A < std::max(1, m_A)) {
    fprintf(stderr,
            "ERROR in NoFLA_HQRRP_WY_blk_var4: ldim_A is < max( 1, m_A ).\n");
    bad = 1;
  }
  if (bad) return HQRRP_ERR_ARG;
  if (std::min(m_A, n_A) == 0) return 0;
  if (nb_alg < 1 || pp < 0) return HQRRP_ERR_ARG;
  hqrrp_options o;
  hqrrp_default_options(&o);
  o.nb = nb_alg;
  o.pp = pp;
  o.panel_pivoting = panel_pivoting;
  return host_hqrrp(m_A, n_A, buff_A, ldim_A, buff_jpvt, buff_tau, o);
}

extern "C" int hqrrp_ut_blk_var2(int m_A, int n_A, double* buff_A,
                                 int ldim_A, int* buff_p, double* buff_s,
                                 int nb_alg, int pp, int panel_pivoting) {
  // FLA_HQRRP_UT_blk_var2 (FLA_HQRRP_UT_blk_var2.c:124-401) with its FLA_Obj
  // arguments unpacked (buffer, length, width, column stride).  It has no
  // argument checks of its own (FLA_Obj dimensions cannot be illegal);
  // these reject what a caller could not have passed in an FLA_Obj.
  if (m_A < 0 || n_A < 0 || ldim_A < std::max(1, m_A)) return HQRRP_ERR_ARG;
  if (std::min(m_A, n_A) == 0) return 0;  // quick return, UT.c:177-180
  if (nb_alg < 1 || pp < 0) return HQRRP_ERR_ARG;
  hqrrp_options o;
  hqrrp_default_options(&o);
  o.nb = nb_alg;
  o.pp = pp;
  o.panel_pivoting = panel_pivoting;
  return host_hqrrp(m_A, n_A, buff_A, ldim_A, buff_p, buff_s, o, true);
}

extern "C" void dgeqp4(int* m, int* n, double* A, int* lda, int* jpvt,
                       double* tau, double* work, int* lwork, int* info) {
  // Argument checking and workspace query as NoFLA_HQRRP_WY_blk_var4.c:
  // 148-182 (the reference asks LAPACK's ilaenv for DGEQRF's block size;
  // this library has no LAPACK and reports with nb = 32, netlib's value).
  const int mA = *m, nA = *n, ld = *lda;
  const int mn = std::min(mA, nA);
  const bool lquery = (*lwork == -1);
  *info = 0;
  if (mA < 0)
    *info = -1;
  else if (nA < 0)
    *info = -2;
  else if (ld < std::max(1, mA))
    *info = -4;
  int iws = 1, lwkopt = 1;
  if (*info == 0) {
    if (mn != 0) {
      iws = 3 * nA + 1;
      lwkopt = 2 * nA + (nA + 1) * 32;
    }
    work[0] = (double)lwkopt;
    if (*lwork < iws && !lquery) *info = -8;
  }
  if (*info != 0) {
    fprintf(stderr,
            " ** On entry to DGEQP3 parameter number %d had an illegal value\n",
            -*info);
    return;
  }
  if (lquery || mn == 0) return;

  // Fixed columns (jpvt != 0) to the front, as WY.c:201-218.
  int nfixed = 0;
  std::vector<double> colbuf;
  for (int c = 0; c < nA; ++c) {
    if (jpvt[c] != 0) {
      if (c != nfixed) {
        colbuf.assign(A + (long)c * ld, A + (long)c * ld + mA);
        memcpy(A + (long)c * ld, A + (long)nfixed * ld, sizeof(double) * mA);
        memcpy(A + (long)nfixed * ld, colbuf.data(), sizeof(double) * mA);
        jpvt[c] = jpvt[nfixed];
        jpvt[nfixed] = c + 1;
      } else {
        jpvt[c] = c + 1;
      }
      ++nfixed;
    } else {
      jpvt[c] = c + 1;
    }
  }
  const int nf = std::min(mA, nfixed);
  hqrrp_options o;
  hqrrp_default_options(&o);
  // min(m,n) < 250 goes to dgeqpf and < 1000 to dgeqp3 in the reference
  // (WY.c:189-199); both are the classic greedy QRCP here.
  int rc = mn < 1000 ? host_dgeqp4_classic(mA, nA, A, ld, jpvt, tau, nf)
           : nf > 0  ? host_dgeqp4_fixed(mA, nA, A, ld, jpvt, tau, nf, o)
                     : host_hqrrp(mA, nA, A, ld, jpvt, tau, o);
  if (rc != HQRRP_OK) {
    fprintf(stderr, "dgeqp4: device factorisation failed (%d)\n", rc);
    *info = HQRRP_INFO_DEVICE - rc;
    return;
  }
  work[0] = iws;
}
