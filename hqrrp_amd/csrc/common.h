// common.h -- device helpers shared by the HQRRP kernels (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HQ_NT 256  // threads per workgroup for every kernel in this library

typedef double hq_d4 __attribute__((ext_vector_type(4)));

// Error flags written by kernels (slot 0 of the handle's error word).
enum {
  HQ_ERR_NONE = 0,
  HQ_ERR_BARRIER_TIMEOUT = 1,  // a persistent kernel's grid barrier timed out
};

// Upper bound on spins of a grid barrier before the kernel gives up and
// reports HQ_ERR_BARRIER_TIMEOUT (a spin is a coherent load plus a short
// sleep, about 1 us: a few seconds in total).
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

// Grid-wide barrier for persistent kernels whose grid the host has sized to
// be co-resident.  bar[0] counts arrivals and only grows within a launch
// (barrier number `phase` completes at phase*nblocks arrivals); bar[1] counts
// workgroups leaving the kernel so the last one out can zero both words for
// the next launch on the stream.  Returns false (in every thread) if the wait
// timed out or another workgroup reported a timeout.
//
// Precondition: every global store that another workgroup reads after this
// barrier was issued by wavefront 0 (the wavefront of thread 0), so thread
// 0's release (which waits for its wavefront's stores) orders all of them
// before the arrival; thread 0's acquire, followed by the workgroup barrier,
// orders every thread's later loads after all arrivals.  Only thread 0 pays
// for the release/acquire cache maintenance.
__device__ __forceinline__ bool hq_grid_sync(unsigned* bar, unsigned nblocks,
                                             unsigned& phase, int* err) {
  __shared__ int s_ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    phase += 1;
    const unsigned target = phase * nblocks;
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    int ok = 1;
    unsigned spins = 0;
    // Poll with relaxed loads (an acquire load would invalidate the caches
    // on every poll); one acquire fence once the count is reached.
    while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
           target) {
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        ok = 0;
        break;
      }
      if (++spins > HQ_SPIN_LIMIT) {
        __hip_atomic_store(err, HQ_ERR_BARRIER_TIMEOUT, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

// Called once by every workgroup after its last hq_grid_sync: the last
// workgroup to leave resets the barrier words.
__device__ __forceinline__ void hq_grid_exit(unsigned* bar, unsigned nblocks) {
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned old = __hip_atomic_fetch_add(bar + 1, 1u, __ATOMIC_ACQ_REL,
                                          __HIP_MEMORY_SCOPE_AGENT);
    if (old == nblocks - 1) {
      __hip_atomic_store(bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(bar + 1, 0u, __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Wavefront (64-lane) sum; every lane receives the total.  The butterfly
// order is fixed, so every workgroup computing the same sum gets the same
// bits.
__device__ __forceinline__ double hq_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// LAPACK dlapy2: sqrt(x^2 + y^2) without destructive overflow.
__device__ __forceinline__ double hq_dlapy2(double x, double y) {
  double xa = fabs(x), ya = fabs(y);
  double w = fmax(xa, ya), z = fmin(xa, ya);
  if (z == 0.0 || w > 1.79769313486231571e308) return w;
  double q = z / w;
  return w * sqrt(1.0 + q * q);
}

// Householder generation with LAPACK's conventions (dlarfg): given alpha and
// the squared 2-norm of x, returns beta = -sign(alpha)*||(alpha,x)||,
// tau = (beta-alpha)/beta and the scale 1/(alpha-beta) applied to x.  When
// x is zero tau = 0, beta = alpha and x is left unscaled (scale = 1).
struct hq_reflector {
  double beta, tau, scale;
};
__device__ __forceinline__ hq_reflector hq_dlarfg(double alpha, double xnorm2) {
  hq_reflector h;
  double xnorm = sqrt(xnorm2);
  if (xnorm == 0.0) {
    h.beta = alpha;
    h.tau = 0.0;
    h.scale = 1.0;
    return h;
  }
  double beta = -copysign(hq_dlapy2(alpha, xnorm), alpha);
  h.beta = beta;
  h.tau = (beta - alpha) / beta;
  h.scale = 1.0 / (alpha - beta);
  return h;
}

// sqrt(eps) with eps = 2^-53 (LAPACK dlamch('E')): Drmac's recompute bound.
#define HQ_TOL3Z 1.0536712127723509e-08

// LAWN 176 partial-norm downdate for one column.  `wt` is the entry of the
// column in the row just eliminated, `below2` the squared norm of the part of
// the column under that row (used when the downdated value is not trusted),
// `mbelow` its length.
__device__ __forceinline__ void hq_downdate_norm(double& d, double& e,
                                                 double wt, double below2,
                                                 int mbelow) {
  if (d != 0.0) {
    double t = fabs(wt) / d;
    double u = (1.0 + t) * (1.0 - t);
    t = (0.0 > u) ? 0.0 : u;
    double ratio = d / e;
    double t2 = t * ratio * ratio;
    if (t2 <= HQ_TOL3Z) {
      if (mbelow > 0) {
        d = sqrt(below2);
        e = d;
      } else {
        d = 0.0;
        e = 0.0;
      }
    } else {
      d = d * sqrt(t);
    }
  }
}

__host__ __device__ inline int hq_cdiv(long a, long b) {
  return (int)((a + b - 1) / b);
}
