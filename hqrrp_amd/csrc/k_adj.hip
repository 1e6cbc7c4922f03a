// k_adj.hip -- kernels for explicit Q and the adjoint of the QR at fixed
// pivots (no counterpart in the reference; the reference's test driver forms
// Q with LAPACK dorgqr, WYtest.c:89-93).
//
//   fill_eye:      Q = [dI; 0] (m x n): [I; 0] starts dorgqr's backward
//                  sweep, d = 0 clears a matrix
//   copy_r:        R (k x n) = upper trapezoid of the factored matrix
//   tri_inv_diag:  inverses of the 64 x 64 diagonal blocks of an upper
//                  triangular matrix (back substitution in LDS, one column
//                  per thread); the off-diagonal blocks of R^-1 follow by
//                  doubling with GEMMs on the host side
//   copyltu:       B(i,j) = M(max(i,j), min(i,j))
//   permute_cols:  column gather dst(:, c) = src(:, p(c)) or scatter
//                  dst(:, p(c)) = src(:, c), p 1-based (jpvt)
#include "common.h"
#include "kernels.h"

namespace {

constexpr int kTB = 64;  // diagonal block of the triangular inverse

__global__ __launch_bounds__(HQ_NT) void fill_eye_kernel(double* Q, long ldq,
                                                         int m, int n,
                                                         double diag) {
  const int c = blockIdx.y;
  for (long i = (long)blockIdx.x * HQ_NT + threadIdx.x; i < m;
       i += (long)gridDim.x * HQ_NT)
    Q[i + c * ldq] = (i == c) ? diag : 0.0;
}

__global__ __launch_bounds__(HQ_NT) void copy_r_kernel(const double* A,
                                                       long lda, int k, int n,
                                                       double* R, long ldr) {
  const int c = blockIdx.x;
  for (int i = threadIdx.x; i < k; i += HQ_NT)
    R[i + (long)c * ldr] = (i <= c) ? A[i + (long)c * lda] : 0.0;
}

// Block d of the diagonal: X = R(B,B)^-1 with B = [64d, 64d+bs).  Column c
// of X solves R(B,B) x = e_c by back substitution; thread c owns it.  R's
// block is staged in LDS (upper part only, the rest read as zero).
__global__ __launch_bounds__(kTB) void tri_inv_diag_kernel(int k,
                                                           const double* R,
                                                           long ldr, double* X,
                                                           long ldx) {
  __shared__ double Rs[kTB][kTB + 1];
  const int d0 = blockIdx.x * kTB;
  const int bs = min(kTB, k - d0);
  const int c = threadIdx.x;
  for (int j = 0; j < bs; ++j) {
    const int i = c;
    Rs[i][j] = (i < bs && i <= j) ? R[(d0 + i) + (long)(d0 + j) * ldr] : 0.0;
  }
  __syncthreads();
  if (c >= bs) return;
  double x[kTB];
#pragma unroll
  for (int i = 0; i < kTB; ++i) x[i] = 0.0;
  // x[i] for i = c down to 0; x[i] = 0 for i > c.
#pragma unroll
  for (int ii = kTB - 1; ii >= 0; --ii) {
    if (ii <= c) {
      double s = (ii == c) ? 1.0 : 0.0;
#pragma unroll
      for (int j = ii + 1; j < kTB; ++j)
        if (j <= c) s -= Rs[ii][j] * x[j];
      x[ii] = s / Rs[ii][ii];
    }
  }
  for (int i = 0; i < bs; ++i) X[(d0 + i) + (long)(d0 + c) * ldx] = x[i];
}

__global__ __launch_bounds__(HQ_NT) void copyltu_kernel(int k, const double* M,
                                                        long ldm, double* B,
                                                        long ldb) {
  const int j = blockIdx.x;
  for (int i = threadIdx.x; i < k; i += HQ_NT)
    B[i + (long)j * ldb] = (i >= j) ? M[i + (long)j * ldm] : M[j + (long)i * ldm];
}

__global__ __launch_bounds__(HQ_NT) void permute_cols_kernel(
    const double* src, long lds, int m, const int* p, int scatter,
    double* dst, long ldd) {
  const int c = blockIdx.y;
  const long pc = p[c] - 1;
  const double* s = src + (scatter ? c : pc) * lds;
  double* d = dst + (scatter ? pc : c) * ldd;
  for (long i = (long)blockIdx.x * HQ_NT + threadIdx.x; i < m;
       i += (long)gridDim.x * HQ_NT)
    d[i] = s[i];
}

unsigned row_blocks(long m) {
  long b = (m + HQ_NT - 1) / HQ_NT;
  return (unsigned)(b < 1 ? 1 : (b > 64 ? 64 : b));
}

}  // namespace

hipError_t launch_fill_eye(double* Q, long ldq, int m, int n, double diag,
                           hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_eye_kernel, dim3(row_blocks(m), n), dim3(HQ_NT), 0, s,
                     Q, ldq, m, n, diag);
  return hipGetLastError();
}

hipError_t launch_copy_r(const double* A, long lda, int k, int n, double* R,
                         long ldr, hipStream_t s) {
  if (k <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(copy_r_kernel, dim3(n), dim3(HQ_NT), 0, s, A, lda, k, n, R,
                     ldr);
  return hipGetLastError();
}

hipError_t launch_tri_inv_diag(int k, const double* R, long ldr, double* X,
                               long ldx, hipStream_t s) {
  if (k <= 0) return hipSuccess;
  hipLaunchKernelGGL(tri_inv_diag_kernel, dim3((k + kTB - 1) / kTB), dim3(kTB),
                     0, s, k, R, ldr, X, ldx);
  return hipGetLastError();
}

int tri_inv_block() { return kTB; }

hipError_t launch_copyltu(int k, const double* M, long ldm, double* B, long ldb,
                          hipStream_t s) {
  if (k <= 0) return hipSuccess;
  hipLaunchKernelGGL(copyltu_kernel, dim3(k), dim3(HQ_NT), 0, s, k, M, ldm, B,
                     ldb);
  return hipGetLastError();
}

hipError_t launch_permute_cols(const double* src, long lds, int m, int ncols,
                               const int* p, int scatter, double* dst,
                               long ldd, hipStream_t s) {
  if (m <= 0 || ncols <= 0) return hipSuccess;
  hipLaunchKernelGGL(permute_cols_kernel, dim3(row_blocks(m), ncols),
                     dim3(HQ_NT), 0, s, src, lds, m, p, scatter, dst, ldd);
  return hipGetLastError();
}
