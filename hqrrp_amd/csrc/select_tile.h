// select_tile.h -- one wavefront's tile of the look-ahead column update.
//
// The b2 <= 64 columns that the next block's panel will hold (position q <-
// source column src[q], src in LDS) are read from rows r0 + m of the not
// yet updated trailing matrix C and updated with the block's reflectors:
// P(m, q) = C(r0 + m, src[q]) - sum_k V(m, k) W'(src[q], k) (V's rows
// aligned with P's), for the 32 rows m in [mw, mw + 32) and all 64
// positions.  Computed transposed (P' =
// C' + W'(-V')) on the matrix cores with the operands loaded straight into
// MFMA registers, in the product order of trail_update_kernel (accumulator
// initialised from C, reflectors in steps of 4, lane l supplying reflector
// k0 + l/16), so the result is bitwise what the full trailing update writes
// for those columns.  acc[i][j] element e: P(m = mw + 16j + l%16,
// q = 16i + l/16 + 4e).  Rows m >= rows and positions q >= b2 read as 0.
//
// The selected rows of W' are gathered once per workgroup into `ws` (LDS,
// at least 64 x kSelLd doubles): every element of W'(src[q], k) lies in a
// cache line of its own, so wavefronts gathering them each for themselves
// multiply that traffic by the wavefront count.  Every thread of the
// workgroup calls select_tile; it ends with a barrier, after which `ws` may
// be reused.
#pragma once
#include "common.h"

constexpr int kSelLd = 65;  // LDS leading dimension of the gathered W' rows

__device__ __forceinline__ void select_tile(hq_d4 (&acc)[4][2],
                                            const int* src, int b2,
                                            const double* C, long ldc, int r0,
                                            int rows, const double* V,
                                            long ldv, const double* Wt,
                                            long ldw, int b, long mw,
                                            double* ws) {
  const int lane = threadIdx.x & 63;
  const int l15 = lane & 15, l4 = lane >> 4;
  // ws[q * kSelLd + k] = W'(src[q], k), zero for q >= b2 or k >= b.
  for (int idx = threadIdx.x; idx < 64 * 64; idx += blockDim.x) {
    const int q = idx & 63, k = idx >> 6;
    ws[q * kSelLd + k] = (q < b2 && k < b) ? Wt[src[q] + (long)k * ldw] : 0.0;
  }
  __syncthreads();
  if (b == 64 && b2 == 64 && mw + 32 <= rows) {
    // Every operand exists: C and the first operand steps are loaded
    // without predicates, in flight together, and a four-step register ring
    // keeps later steps' loads in flight (the same products in the same
    // order as below).
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const double* cc = C + r0 + (long)src[16 * i + l4 + 4 * e] * ldc;
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j][e] = cc[mw + 16 * j + l15];
      }
    auto load = [&](double (&af)[4], double (&bf)[2], int st) {
      const int k = 4 * st + l4;
#pragma unroll
      for (int i = 0; i < 4; ++i) af[i] = ws[(16 * i + l15) * kSelLd + k];
#pragma unroll
      for (int j = 0; j < 2; ++j) bf[j] = -V[mw + 16 * j + l15 + (long)k * ldv];
    };
    auto mma = [&](const double (&af)[4], const double (&bf)[2]) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j],
                                                            acc[i][j], 0, 0, 0);
    };
    double ra[4][4], rb[4][2];
    load(ra[0], rb[0], 0);
    load(ra[1], rb[1], 1);
    load(ra[2], rb[2], 2);
#pragma unroll
    for (int st = 0; st < 16; st += 4) {
      load(ra[3], rb[3], st + 3);
      mma(ra[0], rb[0]);
      if (st + 4 < 16) load(ra[0], rb[0], st + 4);
      mma(ra[1], rb[1]);
      if (st + 5 < 16) load(ra[1], rb[1], st + 5);
      mma(ra[2], rb[2]);
      if (st + 6 < 16) load(ra[2], rb[2], st + 6);
      mma(ra[3], rb[3]);
    }
    __syncthreads();
    return;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int n = 16 * i + l4 + 4 * e;
      const double* cc = C + r0 + (long)(n < b2 ? src[n] : 0) * ldc;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const long m = mw + 16 * j + l15;
        acc[i][j][e] = (n < b2 && m < rows) ? cc[m] : 0.0;
      }
    }
#pragma unroll
  for (int kk = 0; kk < 64; kk += 4) {
    const int k = kk + l4;
    double af[4], bf[2];
#pragma unroll
    for (int i = 0; i < 4; ++i) af[i] = ws[(16 * i + l15) * kSelLd + k];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const long m = mw + 16 * j + l15;
      bf[j] = (k < b && m < rows) ? -V[m + (long)k * ldv] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j],
                                                          acc[i][j], 0, 0, 0);
  }
  __syncthreads();
}
