// k_misc.hip -- column gathers applying a block's composed pivot permutation.
//
// The sketch step of block j swaps up to b pairs of columns of A (all m
// rows), of the sketch Y and of jpvt (reference NoFLA_QRP_pivot_G_B_C,
// NoFLA_HQRRP_WY_blk_var4.c:990-1036, called b times per block).  Here the b
// swaps are composed on the device into at most 2b (dst, src) moves and
// applied in two coalesced passes: every source column is copied out to a
// staging buffer, then every staging column is written to its destination.
#include <algorithm>

#include "common.h"
#include "kernels.h"

namespace {

constexpr int kRowsPerBlock = 4 * HQ_NT;

// Y's columns move as r*yw 32-bit words (yw = 1 for float, 2 for double).
__global__ __launch_bounds__(HQ_NT) void gather_out_kernel(
    const int* moves, const double* A, long lda, int m, const unsigned* Y,
    long ldy, int r, const int* jpvt, double* tmpA, unsigned* tmpY,
    int* tmpJ) {
  const int e = blockIdx.y;
  if (e >= moves[0]) return;
  const long src = moves[2 + 2 * e];
  const long i0 = (long)blockIdx.x * kRowsPerBlock;
  const double* a = A + src * lda;
  double* ta = tmpA + (long)e * m;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    long i = i0 + q * HQ_NT + threadIdx.x;
    if (i < m) ta[i] = a[i];
  }
  if (blockIdx.x == 0) {
    for (int i = threadIdx.x; i < r; i += HQ_NT)
      tmpY[(long)e * r + i] = Y[src * ldy + i];
    if (threadIdx.x == 0 && jpvt) tmpJ[e] = jpvt[src];
  }
}

__global__ __launch_bounds__(HQ_NT) void gather_in_kernel(
    const int* moves, double* A, long lda, int m, unsigned* Y, long ldy, int r,
    int* jpvt, const double* tmpA, const unsigned* tmpY, const int* tmpJ) {
  const int e = blockIdx.y;
  if (e >= moves[0]) return;
  const long dst = moves[1 + 2 * e];
  const long i0 = (long)blockIdx.x * kRowsPerBlock;
  double* a = A + dst * lda;
  const double* ta = tmpA + (long)e * m;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    long i = i0 + q * HQ_NT + threadIdx.x;
    if (i < m) a[i] = ta[i];
  }
  if (blockIdx.x == 0) {
    for (int i = threadIdx.x; i < r; i += HQ_NT)
      Y[dst * ldy + i] = tmpY[(long)e * r + i];
    if (threadIdx.x == 0 && jpvt) jpvt[dst] = tmpJ[e];
  }
}

// dst(rows x cols, ld ldd) = src(rows x cols, ld lds).
__global__ __launch_bounds__(HQ_NT) void copy_block_kernel(double* dst, long ldd,
                                                           const double* src,
                                                           long lds, int rows,
                                                           int cols) {
  const long i0 = (long)blockIdx.x * kRowsPerBlock;
  for (long c = blockIdx.y; c < cols; c += gridDim.y) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      long i = i0 + q * HQ_NT + threadIdx.x;
      if (i < rows) dst[i + c * ldd] = src[i + c * lds];
    }
  }
}

__global__ __launch_bounds__(HQ_NT) void fill_sentinel_kernel(double* p,
                                                              size_t n) {
  const double v = hq_sentinel();
  for (size_t i = (size_t)blockIdx.x * HQ_NT + threadIdx.x; i < n;
       i += (size_t)gridDim.x * HQ_NT)
    p[i] = v;
}

// dst(c, i) = src(i, c) for a rows x cols column-major src, through a
// 32 x 32 LDS tile (both sides coalesced).
__global__ __launch_bounds__(HQ_NT) void transpose_kernel(
    const double* src, long lds, int rows, int cols, double* dst, long ldd) {
  __shared__ double tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const long i0 = (long)blockIdx.x * 32, c0 = (long)blockIdx.y * 32;
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const long i = i0 + tx, c = c0 + ty + k;
    tile[ty + k][tx] = (i < rows && c < cols) ? src[i + c * lds] : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const long c = c0 + tx, i = i0 + ty + k;
    if (i < rows && c < cols) dst[c + i * ldd] = tile[tx][ty + k];
  }
}

__global__ __launch_bounds__(HQ_NT) void iota_kernel(int* p, int n) {
  const int i = blockIdx.x * HQ_NT + threadIdx.x;
  if (i < n) p[i] = i;
}

// One workgroup walks the positions in chunks of 1024, compacting the moved
// ones in order: a ballot per wavefront, the wavefront offsets through LDS.
constexpr int kPM = 1024;
__global__ __launch_bounds__(kPM) void perm_moves_kernel(const int* perm,
                                                         int n, int* moves) {
  __shared__ int s_cnt[kPM / 64];
  __shared__ int s_base;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (t == 0) s_base = 0;
  __syncthreads();
  for (int c0 = 0; c0 < n; c0 += kPM) {
    const int pos = c0 + t;
    const int src = pos < n ? perm[pos] : pos;
    const bool mv = pos < n && src != pos;
    const unsigned long long b = __ballot(mv);
    if (lane == 0) s_cnt[w] = __popcll(b);
    __syncthreads();
    int off = s_base;
    for (int q = 0; q < w; ++q) off += s_cnt[q];
    if (mv) {
      const int idx = off + __popcll(b & ((1ull << lane) - 1ull));
      moves[1 + 2 * idx] = pos;
      moves[2 + 2 * idx] = src;
    }
    __syncthreads();
    if (t == 0) {
      int tot = 0;
      for (int q = 0; q < kPM / 64; ++q) tot += s_cnt[q];
      s_base += tot;
    }
    __syncthreads();
  }
  if (t == 0) moves[0] = s_base;
}

__global__ __launch_bounds__(HQ_NT) void widen_block_kernel(
    double* dst, long ldd, const float* src, long lds, int rows, int cols) {
  const long i0 = (long)blockIdx.x * kRowsPerBlock;
  for (long c = blockIdx.y; c < cols; c += gridDim.y) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      long i = i0 + q * HQ_NT + threadIdx.x;
      if (i < rows) dst[i + c * ldd] = (double)src[i + c * lds];
    }
  }
}

// Power-of-two prescaling of a matrix whose magnitude is extreme (words:
// [0] the bits of max |a_ij| over the finite entries, [1] 1 if some entry
// is NaN or Inf, [2] s, [3] 1/s, [4] 1 when the forward pass must run, [5]
// 1 when the backward pass must).
__global__ __launch_bounds__(64) void scale_init_kernel(
    unsigned long long* w) {
  if (threadIdx.x < 6) w[threadIdx.x] = 0ull;
}

__global__ __launch_bounds__(HQ_NT) void amax_kernel(const double* A, long lda,
                                                     int m, int n,
                                                     unsigned long long* w) {
  const long i0 = (long)blockIdx.x * kRowsPerBlock;
  double mx = 0.0;
  bool bad = false;
  for (long c = blockIdx.y; c < n; c += gridDim.y) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long i = i0 + q * HQ_NT + threadIdx.x;
      if (i < m) {
        const double v = fabs(A[i + c * lda]);
        if (v <= 1.79769313486231571e308)
          mx = fmax(mx, v);
        else
          bad = true;  // Inf or NaN
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  // One atomic per workgroup (the four wavefront maxima combined in LDS):
  // thousands of workgroups updating one word is what this pass waits on.
  __shared__ double s_mx[HQ_NT / 64];
  __shared__ int s_bad[HQ_NT / 64];
  const bool any_bad = __any(bad);
  if ((threadIdx.x & 63) == 0) {
    s_mx[threadIdx.x >> 6] = mx;
    s_bad[threadIdx.x >> 6] = any_bad ? 1 : 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m4 = s_mx[0];
    int b4 = s_bad[0];
#pragma unroll
    for (int q = 1; q < HQ_NT / 64; ++q) {
      m4 = fmax(m4, s_mx[q]);
      b4 |= s_bad[q];
    }
    // Non-negative doubles order like their bit patterns.
    atomicMax(&w[0], (unsigned long long)__double_as_longlong(m4));
    if (b4) atomicOr(&w[1], 1ull);
  }
}

// s = 2^-(e+1) with e = exponent(max |a|) when that exponent lies outside
// [lo, hi] (s max|a| in [0.5, 1)); else 1.  The forward pass also runs when
// there are non-finite entries: multiplying by s (even 1) turns signalling
// NaNs into quiet ones.
__global__ __launch_bounds__(64) void scale_decide_kernel(unsigned long long* w,
                                                          int lo, int hi) {
  if (threadIdx.x != 0) return;
  const double amax = __longlong_as_double((long long)w[0]);
  double s = 1.0;
  if (amax > 0.0) {
    const int e = ilogb(amax);
    if (e < lo || e > hi) s = ldexp(1.0, -e - 1);
  }
  w[2] = (unsigned long long)__double_as_longlong(s);
  w[3] = (unsigned long long)__double_as_longlong(1.0 / s);
  w[4] = (s != 1.0 || w[1]) ? 1ull : 0ull;
  w[5] = (s != 1.0) ? 1ull : 0ull;
}

// Forward (fwd = 1): A *= s everywhere.  Backward: A *= 1/s on R and the
// trailing matrix -- entries (i, j) with i <= j and i < k, or i >= k and
// j >= k -- leaving the reflectors, which do not scale.
__global__ __launch_bounds__(HQ_NT) void scale_apply_kernel(
    double* A, long lda, int m, int n, int k, const unsigned long long* w,
    int fwd) {
  if (!w[fwd ? 4 : 5]) return;
  const double s = __longlong_as_double((long long)w[fwd ? 2 : 3]);
  const long i0 = (long)blockIdx.x * kRowsPerBlock;
  for (long c = blockIdx.y; c < n; c += gridDim.y) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long i = i0 + q * HQ_NT + threadIdx.x;
      if (i >= m) continue;
      if (fwd || (i <= c && i < k) || (i >= k && c >= k)) A[i + c * lda] *= s;
    }
  }
}

// Row k of R negated (columns >= k) for every reflector k < nk with
// tau = 0; one thread per column.
__global__ __launch_bounds__(HQ_NT) void ut_rows_kernel(double* A, long lda,
                                                        int n, int nk,
                                                        const double* tau) {
  const long c = (long)blockIdx.x * HQ_NT + threadIdx.x;
  if (c >= n) return;
  const int last = (int)min((long)nk - 1, c);
  for (int k = 0; k <= last; ++k)
    if (tau[k] == 0.0) A[k + c * lda] = -A[k + c * lda];
}

// tau -> the UT transform's s (1/tau; 1/2 for tau = 0).
__global__ __launch_bounds__(HQ_NT) void ut_tau_kernel(double* tau, int nk) {
  const int k = blockIdx.x * HQ_NT + threadIdx.x;
  if (k < nk) tau[k] = (tau[k] == 0.0) ? 0.5 : 1.0 / tau[k];
}

}  // namespace

hipError_t launch_ut_convert(double* A, long lda, int m, int n, int k,
                             double* tau, hipStream_t s) {
  const int nk = std::min(k, std::min(m, n));
  if (nk <= 0) return hipSuccess;
  hipLaunchKernelGGL(ut_rows_kernel, dim3(hq_cdiv(n, HQ_NT)), dim3(HQ_NT), 0,
                     s, A, lda, n, nk, tau);
  hipLaunchKernelGGL(ut_tau_kernel, dim3(hq_cdiv(nk, HQ_NT)), dim3(HQ_NT), 0,
                     s, tau, nk);
  return hipGetLastError();
}

hipError_t launch_scale_pre(double* A, long lda, int m, int n,
                            unsigned long long* w, int lo, int hi,
                            hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(scale_init_kernel, dim3(1), dim3(64), 0, s, w);
  dim3 grid(hq_cdiv(m, kRowsPerBlock), n < 2048 ? n : 2048);
  // The max pass loops over more columns per workgroup (fewer atomics).
  dim3 mgrid(grid.x, n < 256 ? n : 256);
  hipLaunchKernelGGL(amax_kernel, mgrid, dim3(HQ_NT), 0, s, A, lda, m, n, w);
  hipLaunchKernelGGL(scale_decide_kernel, dim3(1), dim3(64), 0, s, w, lo, hi);
  hipLaunchKernelGGL(scale_apply_kernel, grid, dim3(HQ_NT), 0, s, A, lda, m, n,
                     0, w, 1);
  return hipGetLastError();
}

hipError_t launch_scale_post(double* A, long lda, int m, int n, int k,
                             const unsigned long long* w, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  dim3 grid(hq_cdiv(m, kRowsPerBlock), n < 2048 ? n : 2048);
  hipLaunchKernelGGL(scale_apply_kernel, grid, dim3(HQ_NT), 0, s, A, lda, m, n,
                     k, w, 0);
  return hipGetLastError();
}

hipError_t launch_iota(int* p, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(iota_kernel, dim3(hq_cdiv(n, HQ_NT)), dim3(HQ_NT), 0, s,
                     p, n);
  return hipGetLastError();
}

hipError_t launch_perm_moves(const int* perm, int n, int* moves,
                             hipStream_t s) {
  hipLaunchKernelGGL(perm_moves_kernel, dim3(1), dim3(kPM), 0, s, perm, n,
                     moves);
  return hipGetLastError();
}

hipError_t launch_widen_block(double* dst, long ldd, const float* src,
                              long lds, int rows, int cols, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  dim3 grid(hq_cdiv(rows, kRowsPerBlock), cols < 32768 ? cols : 32768);
  hipLaunchKernelGGL(widen_block_kernel, grid, dim3(HQ_NT), 0, s, dst, ldd,
                     src, lds, rows, cols);
  return hipGetLastError();
}

hipError_t launch_fill_sentinel(double* p, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  size_t blocks = (n + HQ_NT - 1) / HQ_NT;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(fill_sentinel_kernel, dim3((unsigned)blocks), dim3(HQ_NT),
                     0, s, p, n);
  return hipGetLastError();
}

hipError_t launch_copy_block(double* dst, long ldd, const double* src,
                             long lds, int rows, int cols, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  dim3 grid(hq_cdiv(rows, kRowsPerBlock), cols < 32768 ? cols : 32768);
  hipLaunchKernelGGL(copy_block_kernel, grid, dim3(HQ_NT), 0, s, dst, ldd, src,
                     lds, rows, cols);
  return hipGetLastError();
}

hipError_t launch_gather_cols(const int* moves, int max_moves, double* A,
                              long lda, int m, void* Y, long ldy, int r,
                              int ysz, int* jpvt, double* tmpA, double* tmpY,
                              int* tmpJ, hipStream_t s) {
  if (max_moves <= 0) return hipSuccess;
  if (r > 0 && ysz != 4 && ysz != 8) return hipErrorInvalidValue;
  const int yw = ysz / 4;  // 32-bit words per Y element
  unsigned* Yw = static_cast<unsigned*>(Y);
  unsigned* tY = reinterpret_cast<unsigned*>(tmpY);
  dim3 grid(m > 0 ? hq_cdiv(m, kRowsPerBlock) : 1, max_moves);
  hipLaunchKernelGGL(gather_out_kernel, grid, dim3(HQ_NT), 0, s, moves, A, lda,
                     m, Yw, ldy * yw, r * yw, jpvt, tmpA, tY, tmpJ);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(gather_in_kernel, grid, dim3(HQ_NT), 0, s, moves, A, lda,
                     m, Yw, ldy * yw, r * yw, jpvt, tmpA, tY, tmpJ);
  return hipGetLastError();
}

hipError_t launch_transpose(const double* src, long lds, int rows, int cols,
                            double* dst, long ldd, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  hipLaunchKernelGGL(transpose_kernel, dim3(hq_cdiv(rows, 32), hq_cdiv(cols, 32)),
                     dim3(HQ_NT), 0, s, src, lds, rows, cols, dst, ldd);
  return hipGetLastError();
}
