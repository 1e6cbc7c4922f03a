// common.h -- device helpers shared by the HQRRP kernels (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HQ_NT 256  // threads per workgroup for every kernel in this library

typedef double hq_d4 __attribute__((ext_vector_type(4)));
typedef double hq_d2 __attribute__((ext_vector_type(2)));

// Error flags written by kernels (slot 0 of the handle's error word).
enum {
  HQ_ERR_NONE = 0,
  HQ_ERR_BARRIER_TIMEOUT = 1,  // a persistent kernel's grid barrier timed out
};

// Upper bound on polls of a cross-workgroup exchange before the kernel gives
// up and reports HQ_ERR_BARRIER_TIMEOUT (a poll is a coherent load plus a
// short sleep, about 1 us: a few seconds in total).
#define HQ_SPIN_LIMIT (1u << 22)

// Cross-workgroup data of the persistent kernels is exchanged with
// agent-scope (device-coherent) atomic stores and loads, which write through
// to and read from the coherence point instead of the per-XCD L2.
__device__ __forceinline__ void hq_publish(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double hq_fetch(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// Barrier-free exchange.  Every exchange of a launch has slots of its own in
// an exchange buffer that holds only the sentinel below when the launch
// starts.  A producer stores its values with hq_publish; a consumer re-reads
// each value until it no longer holds the sentinel.  Every value is one
// single-copy-atomic 64-bit word that validates itself, so no fence, counter
// or barrier orders anything, and an exchange costs one memory round trip
// once the data is there.  The sentinel is a signalling NaN whose payload no
// arithmetic produces (arithmetic NaNs are quiet).  Two buffers alternate
// between consecutive launches: each launch restores the sentinel over the
// region the previous launch used in the other buffer.
#define HQ_XSENT 0x7FF4DEAD7FF4DEADll

__device__ __forceinline__ double hq_sentinel() {
  return __longlong_as_double(HQ_XSENT);
}
__device__ __forceinline__ bool hq_pending(double v) {
  return __double_as_longlong(v) == HQ_XSENT;
}

// v[i] = *addr(i) for every i < N with addr(i) != nullptr (0.0 otherwise),
// all loads in flight together, re-reading those still pending.  Returns
// false after HQ_SPIN_LIMIT polls (recording HQ_ERR_BARRIER_TIMEOUT) or when
// another workgroup has recorded an error.
template <int N, class Addr>
__device__ __forceinline__ bool hq_await(double (&v)[N], Addr addr, int* err) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double* p = addr(i);
    v[i] = p ? hq_fetch(p) : 0.0;
  }
  for (unsigned spins = 0;; ++spins) {
    bool pend = false;
#pragma unroll
    for (int i = 0; i < N; ++i) pend |= hq_pending(v[i]);
    if (!pend) return true;
    if (spins > HQ_SPIN_LIMIT) {
      __hip_atomic_store(err, HQ_ERR_BARRIER_TIMEOUT, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    if ((spins & 255) == 255 &&
        __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      return false;
    __builtin_amdgcn_s_sleep(1);
#pragma unroll
    for (int i = 0; i < N; ++i)
      if (hq_pending(v[i])) v[i] = hq_fetch(addr(i));
  }
}

// Restores the sentinel over x[0, len), the share of workgroup w of nw.
__device__ __forceinline__ void hq_xreset(double* x, long len, int w, int nw,
                                          int nthreads) {
  const double s = hq_sentinel();
  for (long i = (long)w * nthreads + threadIdx.x; i < len;
       i += (long)nw * nthreads)
    x[i] = s;
}

// Wavefront (64-lane) sum; every lane receives the total.  The butterfly
// order is fixed, so every workgroup computing the same sum gets the same
// bits.
__device__ __forceinline__ double hq_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Argmax over aligned groups of `width` lanes (a power of two <= 64): the
// largest key, and among equal keys the smallest pos (idamax's first-index
// rule); every lane of a group receives its group's (key, pos).  A max-only
// butterfly on the key, then the position from the one lane holding the
// maximum -- or, only if some group holds it in several lanes, a min
// butterfly on the positions.  All 64 lanes must be active.
__device__ __forceinline__ void hq_wave_argmax(double& key, int& pos,
                                               int width) {
  double m = key;
  for (int off = width >> 1; off > 0; off >>= 1)
    m = fmax(m, __shfl_xor(m, off, 64));
  const int lane = threadIdx.x & 63;
  const unsigned long long grp =
      (width == 64) ? ~0ull
                    : (((1ull << width) - 1ull) << (lane & ~(width - 1)));
  const unsigned long long cand = __ballot(key == m) & grp;
  if (__any(__popcll(cand) > 1)) {
    int pp = (key == m) ? pos : 0x7fffffff;
    for (int off = width >> 1; off > 0; off >>= 1)
      pp = min(pp, __shfl_xor(pp, off, 64));
    pos = pp;
  } else {
    pos = __shfl(pos, __builtin_ctzll(cand), 64);
  }
  key = m;
}

// Value of v in lane `lane` (wave-uniform) for every lane: two scalar reads,
// no LDS round trip.
__device__ __forceinline__ double hq_readlane(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ int hq_readlane(int v, int lane) {
  return __builtin_amdgcn_readlane(v, lane);
}

// Value of v in lane l ^ 32 (the other half of the wavefront) by gfx950's
// v_permlane32_swap (a VALU swap of the lower half of one register with the
// upper half of another; no LDS).  With both operands holding v, each lane
// ends up with its own value in one result and its partner's in the other.
__device__ __forceinline__ unsigned hq_xor32_u32(unsigned v) {
  const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
  return (threadIdx.x & 32) ? r[0] : r[1];
}
__device__ __forceinline__ float hq_xor32(float v) {
  return __uint_as_float(hq_xor32_u32(__float_as_uint(v)));
}
__device__ __forceinline__ double hq_xor32(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = hq_xor32_u32((unsigned)b);
  const unsigned hi = hq_xor32_u32((unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// One argmax step: (key, pos) becomes the better of itself and (ko, po) --
// the larger key, the smaller position among equal keys.
__device__ __forceinline__ void hq_argmax_take(double& key, int& pos, double ko,
                                               int po) {
  const bool take = (ko > key) | ((ko == key) & (po < pos));  // branch-free
  key = take ? ko : key;
  pos = take ? po : pos;
}
// Combines each lane with the lane that DPP control CTRL selects inside its
// row of 16 (quad permutations and mirrors: a VALU move, no LDS).
template <int CTRL>
__device__ __forceinline__ void hq_argmax_dpp(double& key, int& pos) {
  const long long b = __double_as_longlong(key);
  const unsigned lo =
      (unsigned)__builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(
      0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  const int po = __builtin_amdgcn_update_dpp(0, pos, CTRL, 0xF, 0xF, false);
  hq_argmax_take(
      key, pos,
      __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo)),
      po);
}
// Maximum of v over the wavefront, in every lane: within each row of 16
// lanes a prefix maximum by DPP row shifts (1, 2, 4, 8; lanes shifted in from
// outside the row read 0), then the row maxima combined by row broadcasts of
// lanes 15 and 31, so lane 63 holds the maximum.  All 64 lanes active.
__device__ __forceinline__ unsigned hq_wave_max_u32(unsigned v) {
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, true));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, true));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xE, true));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xC, true));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, true));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, true));
  return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// Order-preserving map of a non-NaN double to an unsigned 64-bit key (both
// zeros to the same key): sign bit set -> all bits inverted, else the sign
// bit set.
__device__ __forceinline__ unsigned long long hq_order_key(double key) {
  const unsigned long long b =
      (unsigned long long)__double_as_longlong(key == 0.0 ? 0.0 : key);
  return (b >> 63) ? ~b : (b | (1ull << 63));
}

// Argmax over the 64 lanes of a wavefront with the rule of hq_argmax_take
// (idamax's: largest key, first position among equals); every lane receives
// the result.  Keys are compared through hq_order_key: the maximum of the
// upper 32 bits, then -- only when several lanes share it -- of the lower 32
// bits among those lanes, then -- only for an exact tie -- the smallest
// position among the tied lanes.  Keys must not be NaN (hq_pivot_key maps
// NaN norms to a finite key).  All 64 lanes must be active.
__device__ __forceinline__ void hq_wave_argmax64(double& key, int& pos) {
  const unsigned long long k = hq_order_key(key);
  const unsigned hi = (unsigned)(k >> 32), lo = (unsigned)k;
  const unsigned mhi = hq_wave_max_u32(hi);
  unsigned long long cand = __ballot(hi == mhi);
  if (__popcll(cand) > 1) {
    const unsigned mlo = hq_wave_max_u32(hi == mhi ? lo : 0u);
    cand = __ballot(hi == mhi && lo == mlo);
    if (__popcll(cand) > 1) {
      const bool mine = (cand >> (threadIdx.x & 63)) & 1ull;
      int pp = mine ? pos : 0x7fffffff;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) pp = min(pp, __shfl_xor(pp, off, 64));
      const int lane = __builtin_ctzll(__ballot(mine && pos == pp));
      key = hq_readlane(key, lane);
      pos = pp;
      return;
    }
  }
  const int lane = __builtin_ctzll(cand);
  key = hq_readlane(key, lane);
  pos = hq_readlane(pos, lane);
}

// Pivot key of a candidate's partial norm: a NaN norm (NaN or Inf entries
// in its column) orders after every finite candidate and before the columns
// already taken (key -1), so every lane's argmax agrees and such a column is
// chosen last among the candidates, in position order.
__device__ __forceinline__ double hq_pivot_key(double d) {
  return d == d ? d : -0.5;
}

// Largest finite value of T (LAPACK's overflow threshold in dlapy2/slapy2).
template <class T>
__device__ __forceinline__ T hq_huge();
template <>
__device__ __forceinline__ double hq_huge<double>() {
  return 1.79769313486231571e308;
}
template <>
__device__ __forceinline__ float hq_huge<float>() {
  return 3.40282347e38f;
}

// LAPACK dlapy2 / slapy2: sqrt(x^2 + y^2) without destructive overflow.
template <class T>
__device__ __forceinline__ T hq_dlapy2(T x, T y) {
  T xa = fabs(x), ya = fabs(y);
  T w = fmax(xa, ya), z = fmin(xa, ya);
  if (z == T(0) || w > hq_huge<T>()) return w;
  T q = z / w;
  return w * sqrt(T(1) + q * q);
}

// Householder generation with LAPACK's conventions (dlarfg): given alpha and
// the squared 2-norm of x, returns beta = -sign(alpha)*||(alpha,x)||,
// tau = (beta-alpha)/beta and the scale 1/(alpha-beta) applied to x.  When
// x is zero tau = 0, beta = alpha and x is left unscaled (scale = 1).
template <class T>
struct hq_reflector_t {
  T beta, tau, scale;
};
typedef hq_reflector_t<double> hq_reflector;
// In the arithmetic of T (dlarfg for double, slarfg for float).
template <class T>
__device__ __forceinline__ hq_reflector_t<T> hq_larfg(T alpha, T xnorm2) {
  hq_reflector_t<T> h;
  T xnorm = sqrt(xnorm2);
  if (xnorm == T(0)) {
    h.beta = alpha;
    h.tau = T(0);
    h.scale = T(1);
    return h;
  }
  T beta = -copysign(hq_dlapy2<T>(alpha, xnorm), alpha);
  h.beta = beta;
  h.tau = (beta - alpha) / beta;
  h.scale = T(1) / (alpha - beta);
  return h;
}
// The same with the norm sqrt(alpha^2 + ||x||^2) taken directly (one square
// root instead of dlapy2's): for the kernels of a factorisation, which runs
// on a matrix prescaled to magnitudes within 2^+-256 (2^+-60 with the float
// sketch), where these squares can neither overflow nor underflow.
template <class T>
__device__ __forceinline__ hq_reflector_t<T> hq_larfg_direct(T alpha, T xnorm2) {
  hq_reflector_t<T> h;
  if (xnorm2 == T(0)) {
    h.beta = alpha;
    h.tau = T(0);
    h.scale = T(1);
    return h;
  }
  T beta = -copysign(sqrt(__builtin_fma(alpha, alpha, xnorm2)), alpha);
  h.beta = beta;
  h.tau = (beta - alpha) / beta;
  h.scale = T(1) / (alpha - beta);
  return h;
}
__device__ __forceinline__ hq_reflector hq_dlarfg(double alpha, double xnorm2) {
  return hq_larfg<double>(alpha, xnorm2);
}

// sqrt(eps) with eps = 2^-53 (LAPACK dlamch('E')): Drmac's recompute bound;
// for float, sqrt(slamch('E')) = sqrt(2^-24).
#define HQ_TOL3Z 1.0536712127723509e-08
#define HQ_TOL3Z_F32 2.44140625e-04f
template <class T>
__device__ __forceinline__ T hq_tol3z();
template <>
__device__ __forceinline__ double hq_tol3z<double>() {
  return HQ_TOL3Z;
}
template <>
__device__ __forceinline__ float hq_tol3z<float>() {
  return HQ_TOL3Z_F32;
}

// LAWN 176 partial-norm downdate for one column.  `wt` is the entry of the
// column in the row just eliminated, `below2` the squared norm of the part of
// the column under that row (used when the downdated value is not trusted),
// `mbelow` its length.
template <class T>
__device__ __forceinline__ void hq_downdate_norm(T& d, T& e, T wt, T below2,
                                                 int mbelow) {
  if (d != T(0)) {
    T t = fabs(wt) / d;
    T u = (T(1) + t) * (T(1) - t);
    t = (T(0) > u) ? T(0) : u;
    T ratio = d / e;
    T t2 = t * ratio * ratio;
    if (t2 <= hq_tol3z<T>()) {
      if (mbelow > 0) {
        d = sqrt(below2);
        e = d;
      } else {
        d = T(0);
        e = T(0);
      }
    } else {
      d = d * sqrt(t);
    }
  }
}

__host__ __device__ inline int hq_cdiv(long a, long b) {
  return (int)((a + b - 1) / b);
}
