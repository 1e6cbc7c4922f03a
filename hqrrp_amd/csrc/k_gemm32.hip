// k_gemm32.hip -- GEMM in FP32 arithmetic for the mixed-precision sketch:
// the sketch Y = Omega A (NoFLA_HQRRP_WY_blk_var4.c:407-409) and the sketch
// downdate products (617-708, 743-771) when the sketch is held in FP32.
//
// C (M x N, float) = alpha op(A) op(B) + beta C, where each operand is
// stored as FP64 or FP32 (converted to FP32 when staged) and may be read as
// unit lower trapezoidal (the reflector block V).  With bf16 set, both
// operands are rounded to bfloat16 (round to nearest even) as they are
// staged: products of bf16 values are exact in FP32 and accumulate in FP32,
// as a bf16 matrix-core product would.  64 x 64 tiles, 256 threads with 4x4
// outputs each, K staged 16 at a time through LDS; optional split-K into
// float slices reduced in a fixed order.
#include "common.h"
#include "kernels.h"

namespace {

constexpr int kT = 64;
constexpr int kK = 16;

__device__ __forceinline__ float to_bf16(float v) {
  unsigned b = __float_as_uint(v);
  b += 0x7FFFu + ((b >> 16) & 1u);
  return __uint_as_float(b & 0xFFFF0000u);
}

__device__ __forceinline__ float load_el(const void* p, long idx, int is64) {
  return is64 ? (float)static_cast<const double*>(p)[idx]
              : static_cast<const float*>(p)[idx];
}

__device__ __forceinline__ float unit_lower32(float v, long row, long col,
                                              int mask) {
  if (!mask) return v;
  return (row < col) ? 0.0f : ((row == col) ? 1.0f : v);
}

template <bool TA, bool TB>
__global__ __launch_bounds__(HQ_NT) void gemm32_kernel(Gemm32Args g) {
  __shared__ float As[kK][kT + 4];  // As[k][m] = op(A)(m0+m, k0+k)
  __shared__ float Bs[kK][kT + 4];  // Bs[k][n] = op(B)(k0+k, n0+n)
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const long m0 = (long)blockIdx.x * kT, n0 = (long)blockIdx.y * kT;
  const long kbeg = (long)blockIdx.z * g.kchunk;
  const long kend = min((long)g.K, kbeg + g.kchunk);
  float acc[4][4] = {};
  for (long k0 = kbeg; k0 < kend; k0 += kK) {
#pragma unroll
    for (int q = 0; q < kT * kK / HQ_NT; ++q) {
      const int idx = t + HQ_NT * q;
      // A tile: coalesce along the stored leading dimension.
      int mi, kk;
      if (!TA) {
        mi = idx & (kT - 1);
        kk = idx / kT;
      } else {
        kk = idx & (kK - 1);
        mi = idx / kK;
      }
      const long gm = m0 + mi, gk = k0 + kk;
      float v = 0.0f;
      if (gm < g.M && gk < kend) {
        if (!TA)
          v = unit_lower32(load_el(g.A, gm + gk * g.lda, g.a64), gm, gk,
                           g.maskA);
        else
          v = unit_lower32(load_el(g.A, gk + gm * g.lda, g.a64), gk, gm,
                           g.maskA);
      }
      As[kk][mi] = g.bf16 ? to_bf16(v) : v;
      int nj, kb;
      if (!TB) {
        kb = idx & (kK - 1);
        nj = idx / kK;
      } else {
        nj = idx & (kT - 1);
        kb = idx / kT;
      }
      const long gn = n0 + nj, gkb = k0 + kb;
      float w = 0.0f;
      if (gn < g.N && gkb < kend) {
        if (!TB)
          w = unit_lower32(load_el(g.B, gkb + gn * g.ldb, g.b64), gkb, gn,
                           g.maskB);
        else
          w = unit_lower32(load_el(g.B, gn + gkb * g.ldb, g.b64), gn, gkb,
                           g.maskB);
      }
      Bs[kb][nj] = g.bf16 ? to_bf16(w) : w;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kK; ++kk) {
      float av[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) av[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = Bs[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(av[i], bv[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long gm = m0 + ty * 4 + i, gn = n0 + tx * 4 + j;
      if (gm >= g.M || gn >= g.N) continue;
      if (g.part) {
        g.part[(long)blockIdx.z * g.M * g.N + gm + gn * g.M] = acc[i][j];
      } else {
        float* c = g.C + gm + gn * g.ldc;
        *c = (g.beta == 0.0f) ? g.alpha * acc[i][j]
                              : g.alpha * acc[i][j] + g.beta * *c;
      }
    }
}

__global__ __launch_bounds__(HQ_NT) void reduce32_kernel(int M, int N, int S,
                                                         const float* part,
                                                         float alpha,
                                                         float beta, float* C,
                                                         long ldc) {
  const long total = (long)M * N;
  for (long e = (long)blockIdx.x * HQ_NT + threadIdx.x; e < total;
       e += (long)gridDim.x * HQ_NT) {
    float s = 0.0f;
    for (int q = 0; q < S; ++q) s += part[(long)q * total + e];
    const long i = e % M, j = e / M;
    float* c = C + i + j * ldc;
    *c = (beta == 0.0f) ? alpha * s : alpha * s + beta * *c;
  }
}

__global__ __launch_bounds__(HQ_NT) void cast_f32_kernel(const double* src,
                                                         float* dst, long n) {
  for (long i = (long)blockIdx.x * HQ_NT + threadIdx.x; i < n;
       i += (long)gridDim.x * HQ_NT)
    dst[i] = (float)src[i];
}

}  // namespace

hipError_t launch_cast_f32(const double* src, float* dst, long n,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  long blocks = (n + HQ_NT - 1) / HQ_NT;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(cast_f32_kernel, dim3((unsigned)blocks), dim3(HQ_NT), 0, s,
                     src, dst, n);
  return hipGetLastError();
}

hipError_t launch_gemm32(const Gemm32Args& g, bool ta, bool tb, int splits,
                         hipStream_t s) {
  dim3 grid(hq_cdiv(g.M, kT), hq_cdiv(g.N, kT), splits);
  if (!ta && !tb)
    hipLaunchKernelGGL((gemm32_kernel<false, false>), grid, dim3(HQ_NT), 0, s, g);
  else if (!ta && tb)
    hipLaunchKernelGGL((gemm32_kernel<false, true>), grid, dim3(HQ_NT), 0, s, g);
  else if (ta && !tb)
    hipLaunchKernelGGL((gemm32_kernel<true, false>), grid, dim3(HQ_NT), 0, s, g);
  else
    hipLaunchKernelGGL((gemm32_kernel<true, true>), grid, dim3(HQ_NT), 0, s, g);
  return hipGetLastError();
}

hipError_t launch_reduce32(int M, int N, int S, const float* part, float alpha,
                           float beta, float* C, long ldc, hipStream_t s) {
  long total = (long)M * N;
  long blocks = (total + HQ_NT - 1) / HQ_NT;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(reduce32_kernel, dim3((unsigned)blocks), dim3(HQ_NT), 0, s,
                     M, N, S, part, alpha, beta, C, ldc);
  return hipGetLastError();
}
