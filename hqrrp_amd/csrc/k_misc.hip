// k_misc.hip -- column gathers applying a block's composed pivot permutation.
//
// The sketch step of block j swaps up to b pairs of columns of A (all m
// rows), of the sketch Y and of jpvt (reference NoFLA_QRP_pivot_G_B_C,
// NoFLA_HQRRP_WY_blk_var4.c:990-1036, called b times per block).  Here the b
// swaps are composed on the device into at most 2b (dst, src) moves and
// applied in two coalesced passes: every source column is copied out to a
// staging buffer, then every staging column is written to its destination.
#include "common.h"
#include "kernels.h"

namespace {

constexpr int kRowsPerBlock = 4 * HQ_NT;

__global__ __launch_bounds__(HQ_NT) void gather_out_kernel(
    const int* moves, const double* A, long lda, int m, const double* Y,
    long ldy, int r, const int* jpvt, double* tmpA, double* tmpY, int* tmpJ) {
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
    if (threadIdx.x == 0) tmpJ[e] = jpvt[src];
  }
}

__global__ __launch_bounds__(HQ_NT) void gather_in_kernel(
    const int* moves, double* A, long lda, int m, double* Y, long ldy, int r,
    int* jpvt, const double* tmpA, const double* tmpY, const int* tmpJ) {
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
    if (threadIdx.x == 0) jpvt[dst] = tmpJ[e];
  }
}

}  // namespace

hipError_t launch_gather_cols(const int* moves, int max_moves, double* A,
                              long lda, int m, double* Y, long ldy, int r,
                              int* jpvt, double* tmpA, double* tmpY, int* tmpJ,
                              hipStream_t s) {
  if (max_moves <= 0) return hipSuccess;
  dim3 grid(m > 0 ? hq_cdiv(m, kRowsPerBlock) : 1, max_moves);
  hipLaunchKernelGGL(gather_out_kernel, grid, dim3(HQ_NT), 0, s, moves, A, lda,
                     m, Y, ldy, r, jpvt, tmpA, tmpY, tmpJ);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(gather_in_kernel, grid, dim3(HQ_NT), 0, s, moves, A, lda,
                     m, Y, ldy, r, jpvt, tmpA, tmpY, tmpJ);
  return hipGetLastError();
}
