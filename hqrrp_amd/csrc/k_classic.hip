// k_classic.hip -- classic greedy column-pivoted QR (LAPACK dgeqpf / dgeqp3
// semantics) for the sizes that dgeqp4 does not hand to HQRRP.
//
// NoFLA_HQRRP_WY_blk_var4.c:189-199 sends min(m,n) < 250 to LAPACK dgeqpf and
// min(m,n) < 1000 to dgeqp3.  Both choose at step i the first column of
// largest partial norm, swap it to position i (all m rows, jpvt and both norm
// arrays), reflect it with dlarfg, apply the reflector to the columns on its
// right and downdate their partial norms with LAWN 176's test: a norm whose
// downdate loses more than half the digits (temp * (vn1/vn2)^2 <= sqrt(eps))
// is recomputed from the updated column.  dgeqp3 blocks the updates (dlaqps)
// without changing which column is chosen in exact arithmetic.
//
// Small shapes (m <= 1024, n <= 2048: every square size below 1000) take one
// launch per step, classic_fused_kernel: the matrix, jpvt and the norms
// ping-pong between two workspaces (read X, write Y), so every workgroup can
// recompute the step's pivot and reflector from X (the same data and the same
// reduction order give the same bits everywhere) and update its four columns
// into Y while workgroup 0 writes the finished column i to A; no workgroup
// writes what another reads.  The latency of one step is then about three
// dependent L2 round trips (norms, pivot column, own column) instead of two
// kernels' worth.
//
// Other shapes take two launches per step:
//   classic_pivot_kernel   one workgroup: argmax, swap, dlarfg, tau(i).
//   classic_update_kernel  one wave per column j > i: w = v'A(i:m, j),
//                          A(i:m, j) += (-tau w) v, the norm downdate (the
//                          recomputed norm comes from the same pass).
// The update reads each element of A(i:m, i+1:n) twice and writes it once
// (v is read by every wave from L2): 16 bytes of HBM per element per step when
// the trailing columns do not fit in L2, less when they do.
#include "common.h"
#include "kernels.h"

namespace {

constexpr int kWaves = HQ_NT / 64;

// vn1 = vn2 = ||A(k0:m, j)|| for j in [k0, n), one wave per column.
__global__ __launch_bounds__(HQ_NT) void classic_norms_kernel(
    const double* A, long lda, int m, int n, int k0, double* vn1,
    double* vn2) {
  const int lane = threadIdx.x & 63;
  const long j = (long)k0 + (long)blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (j >= n) return;
  const double* c = A + j * lda;
  double s = 0.0;
  for (int r = k0 + lane; r < m; r += 64) s = fma(c[r], c[r], s);
  s = sqrt(hq_wave_sum(s));
  if (lane == 0) {
    vn1[j] = s;
    vn2[j] = s;
  }
}

// Step i, part 1 (dgeqpf's DO 40 body up to DLARFG): pivot = first index of
// max vn1(i:n); swap A(:, i) <-> A(:, pvt), jpvt, vn1(pvt) = vn1(i),
// vn2(pvt) = vn2(i); then the reflector of A(i:m, i): beta into A(i, i), the
// scaled vector into A(i+1:m, i), tau(i).  Each thread reads back only the
// rows it wrote itself; A(i, i) and the sum of squares go through LDS.
// pivoting = 0: no choice (p = i).  piv (optional): piv[i] = p.
__global__ __launch_bounds__(HQ_NT) void classic_pivot_kernel(
    double* A, long lda, int m, int n, int i, int* jpvt, double* tau,
    double* vn1, double* vn2, int pivoting, int* piv) {
  __shared__ double s_key[kWaves], s_sum[kWaves], s_alpha;
  __shared__ int s_pos[kWaves];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  double key = -1.0;
  int pos = 0x7fffffff;
  if (pivoting) {
    for (int j = i + t; j < n; j += HQ_NT) {
      const double v = vn1[j];
      if (v > key) {  // increasing j: the first of equal maxima stays
        key = v;
        pos = j;
      }
    }
  }
  hq_wave_argmax64(key, pos);
  if (lane == 0) {
    s_key[w] = key;
    s_pos[w] = pos;
  }
  __syncthreads();
  key = s_key[0];
  pos = s_pos[0];
#pragma unroll
  for (int q = 1; q < kWaves; ++q) hq_argmax_take(key, pos, s_key[q], s_pos[q]);
  // In [i, n) since vn1 >= 0 > the initial key; a NaN norm (NaN in A) is
  // never taken, and an all-NaN range keeps column i.
  const int p = (pos < n) ? pos : i;
  if (piv && t == 0) piv[i] = p;
  double* ci = A + (long)i * lda;
  if (p != i) {
    double* cp = A + (long)p * lda;
    for (int r = t; r < m; r += HQ_NT) {
      const double x = cp[r];
      cp[r] = ci[r];
      ci[r] = x;
    }
    if (t == 0) {
      const int jt = jpvt[p];
      jpvt[p] = jpvt[i];
      jpvt[i] = jt;
      vn1[p] = vn1[i];
      vn2[p] = vn2[i];
    }
  }
  if (t == i % HQ_NT) s_alpha = ci[i];
  double ss = 0.0;
  for (int r = t; r < m; r += HQ_NT)
    if (r > i) ss = fma(ci[r], ci[r], ss);
  ss = hq_wave_sum(ss);
  if (lane == 0) s_sum[w] = ss;
  __syncthreads();
  double x2 = s_sum[0];
#pragma unroll
  for (int q = 1; q < kWaves; ++q) x2 += s_sum[q];
  // dlarfg: tau = 0 (H = I) when the part below the diagonal is zero, which
  // includes i = m - 1.
  const hq_reflector h = hq_dlarfg(s_alpha, x2);
  if (h.tau != 0.0)
    for (int r = t; r < m; r += HQ_NT)
      if (r > i) ci[r] *= h.scale;
  if (t == 0) {
    ci[i] = h.beta;
    tau[i] = h.tau;
  }
}

// Step i, part 2 (dgeqpf's DLARF and DO 30): columns j in (i, n), one wave
// each.  With tau = 0 the column is left untouched, as dlarf does.  REF:
// the downdate of the reference's NoFLA_QRP_downdate_partial_norms
// (NoFLA_HQRRP_WY_blk_var4.c:925-951: t = max(0, (1 + t)(1 - t))) instead
// of LAPACK dlaqp2's 1 - t^2; the two differ in rounding only.  pivoting = 0
// skips the norms.
template <bool REF>
__global__ __launch_bounds__(HQ_NT) void classic_update_kernel(
    double* A, long lda, int m, int n, int i, const double* tau, double* vn1,
    double* vn2, int pivoting) {
  const int lane = threadIdx.x & 63;
  const long j = (long)i + 1 + (long)blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (j >= n) return;  // wave-uniform
  const double* v = A + (long)i * lda;  // v(i) = 1, v(r > i) below A(i, i)
  double* c = A + j * lda;
  const double tu = tau[i];
  const double cij = c[i];
  double aij = cij, s2 = 0.0;
  if (tu != 0.0) {
    double d = 0.0;
    for (int r = i + 1 + lane; r < m; r += 64) d = fma(v[r], c[r], d);
    const double f = -tu * (hq_wave_sum(d) + cij);
    for (int r = i + 1 + lane; r < m; r += 64) {
      const double x = fma(v[r], f, c[r]);
      c[r] = x;
      s2 = fma(x, x, s2);
    }
    aij = cij + f;
    if (lane == 0) c[i] = aij;
  } else {
    for (int r = i + 1 + lane; r < m; r += 64) s2 = fma(c[r], c[r], s2);
  }
  s2 = hq_wave_sum(s2);
  if (lane != 0 || !pivoting) return;
  const double d = vn1[j];
  if (d == 0.0) return;
  const double q = fabs(aij) / d;
  double temp = REF ? (1.0 + q) * (1.0 - q) : 1.0 - q * q;
  temp = temp > 0.0 ? temp : 0.0;
  const double ratio = d / vn2[j];
  const double temp2 = temp * (ratio * ratio);
  if (temp2 <= HQ_TOL3Z) {
    const double e = (m - i - 1 > 0) ? sqrt(s2) : 0.0;
    vn1[j] = e;
    vn2[j] = e;
  } else {
    vn1[j] = d * sqrt(temp);
  }
}

// One step of the small-shape path.  X, Y: m x n (ld ldx), jX/jY: jpvt,
// vX/vY: vn1 at [0, n), vn2 at [n, 2n).  Positions >= i of X are current;
// this step writes positions > i of Y and column i of A (ld lda), tau(i),
// jpvt(i).
__global__ __launch_bounds__(HQ_NT) void classic_fused_kernel(
    const double* X, double* Y, long ldx, int m, int n, int i, const int* jX,
    int* jY, const double* vX, double* vY, double* A, long lda, int* jpvt,
    double* tau) {
  constexpr int kR = kClassicMaxRows / HQ_NT;  // pivot-column rows per thread
  constexpr int kL = kClassicMaxRows / 64;     // own-column rows per lane
  __shared__ double s_v[kClassicMaxRows];
  __shared__ double s_key[kWaves], s_sum[kWaves], s_alpha;
  __shared__ int s_pos[kWaves];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  // This wave's position j; its column, norms and jpvt entry are loaded
  // before the pivot is known (only the wave with j == pivot reloads).
  const int j = i + 1 + blockIdx.x * kWaves + w;
  const bool own = j < n;  // wave-uniform
  double c[kL];
  double dn = 0.0, en = 0.0;
  int jv = 0;
  if (own) {
    const double* xc = X + (long)j * ldx;
#pragma unroll
    for (int k = 0; k < kL; ++k) {
      const int r = lane + 64 * k;
      c[k] = (r < m) ? xc[r] : 0.0;
    }
    dn = vX[j];
    en = vX[n + j];
    jv = jX[j];
  }
  // Pivot: first index of the largest vn1(i:n).
  double key = -1.0;
  int pos = 0x7fffffff;
  for (int q = i + t; q < n; q += HQ_NT) {
    const double v = vX[q];
    if (v > key) {
      key = v;
      pos = q;
    }
  }
  hq_wave_argmax64(key, pos);
  if (lane == 0) {
    s_key[w] = key;
    s_pos[w] = pos;
  }
  __syncthreads();
  key = s_key[0];
  pos = s_pos[0];
#pragma unroll
  for (int q = 1; q < kWaves; ++q) hq_argmax_take(key, pos, s_key[q], s_pos[q]);
  const int p = (pos < n) ? pos : i;
  // Reflector of the pivot column's rows i..m-1.
  const double* xp = X + (long)p * ldx;
  double xr[kR];
#pragma unroll
  for (int q = 0; q < kR; ++q) {
    const int r = t + q * HQ_NT;
    xr[q] = (r < m) ? xp[r] : 0.0;
  }
  double ss = 0.0;
#pragma unroll
  for (int q = 0; q < kR; ++q) {
    const int r = t + q * HQ_NT;
    if (r > i) ss = fma(xr[q], xr[q], ss);
    if (r == i) s_alpha = xr[q];
  }
  ss = hq_wave_sum(ss);
  if (lane == 0) s_sum[w] = ss;
  __syncthreads();
  double x2 = s_sum[0];
#pragma unroll
  for (int q = 1; q < kWaves; ++q) x2 += s_sum[q];
  const hq_reflector h = hq_dlarfg(s_alpha, x2);
#pragma unroll
  for (int q = 0; q < kR; ++q) {
    const int r = t + q * HQ_NT;
    const double vr = (r > i) ? xr[q] * h.scale : (r == i ? 1.0 : 0.0);
    s_v[r] = vr;
    if (blockIdx.x == 0 && r < m) {
      A[(long)i * lda + r] = (r < i) ? xr[q] : (r == i ? h.beta : vr);
    }
  }
  if (blockIdx.x == 0 && t == 0) {
    tau[i] = h.tau;
    jpvt[i] = jX[p];
  }
  __syncthreads();
  if (!own) return;
  if (j == p) {  // position p receives X's column i
    const double* xc = X + (long)i * ldx;
#pragma unroll
    for (int k = 0; k < kL; ++k) {
      const int r = lane + 64 * k;
      c[k] = (r < m) ? xc[r] : 0.0;
    }
    dn = vX[i];
    en = vX[n + i];
    jv = jX[i];
  }
  double aij = 0.0, s2 = 0.0;
  if (h.tau != 0.0) {
    double d = 0.0;
#pragma unroll
    for (int k = 0; k < kL; ++k) d = fma(s_v[lane + 64 * k], c[k], d);
    const double f = -h.tau * hq_wave_sum(d);
#pragma unroll
    for (int k = 0; k < kL; ++k) c[k] = fma(s_v[lane + 64 * k], f, c[k]);
  }
  double* yc = Y + (long)j * ldx;
#pragma unroll
  for (int k = 0; k < kL; ++k) {
    const int r = lane + 64 * k;
    if (r < m) yc[r] = c[k];
    if (r > i) s2 = fma(c[k], c[k], s2);
    if (r == i) aij = c[k];
  }
  s2 = hq_wave_sum(s2);
  aij = hq_readlane(aij, i & 63);
  if (lane != 0) return;
  jY[j] = jv;
  double d = dn, e = en;
  if (d != 0.0) {
    const double q = fabs(aij) / d;
    double temp = 1.0 - q * q;
    temp = temp > 0.0 ? temp : 0.0;
    const double ratio = d / e;
    const double temp2 = temp * (ratio * ratio);
    if (temp2 <= HQ_TOL3Z) {
      d = (m - i - 1 > 0) ? sqrt(s2) : 0.0;
      e = d;
    } else {
      d = d * sqrt(temp);
    }
  }
  vY[j] = d;
  vY[n + j] = e;
}

}  // namespace

hipError_t launch_classic_norms(const double* A, long lda, int m, int n,
                                int k0, double* vn1, double* vn2,
                                hipStream_t s) {
  if (k0 >= n) return hipSuccess;
  hipLaunchKernelGGL(classic_norms_kernel, dim3(hq_cdiv(n - k0, kWaves)),
                     dim3(HQ_NT), 0, s, A, lda, m, n, k0, vn1, vn2);
  return hipGetLastError();
}

hipError_t launch_classic_step(double* A, long lda, int m, int n, int i,
                               int* jpvt, double* tau, double* vn1,
                               double* vn2, hipStream_t s) {
  if (i < 0 || i >= m || i >= n) return hipErrorInvalidValue;
  hipLaunchKernelGGL(classic_pivot_kernel, dim3(1), dim3(HQ_NT), 0, s, A, lda,
                     m, n, i, jpvt, tau, vn1, vn2, 1, nullptr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || i + 1 >= n) return e;
  hipLaunchKernelGGL(classic_update_kernel<false>,
                     dim3(hq_cdiv(n - i - 1, kWaves)), dim3(HQ_NT), 0, s, A,
                     lda, m, n, i, tau, vn1, vn2, 1);
  return hipGetLastError();
}

hipError_t launch_qrpmod_step(double* A, long lda, int m, int n, int i,
                              int pivoting, int* perm, int* piv, double* tau,
                              double* vn1, double* vn2, hipStream_t s) {
  if (i < 0 || i >= m || i >= n) return hipErrorInvalidValue;
  hipLaunchKernelGGL(classic_pivot_kernel, dim3(1), dim3(HQ_NT), 0, s, A, lda,
                     m, n, i, perm, tau, vn1, vn2, pivoting, piv);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || i + 1 >= n) return e;
  hipLaunchKernelGGL(classic_update_kernel<true>,
                     dim3(hq_cdiv(n - i - 1, kWaves)), dim3(HQ_NT), 0, s, A,
                     lda, m, n, i, tau, vn1, vn2, pivoting);
  return hipGetLastError();
}

int classic_fused_ok(int m, int n) {
  return m <= kClassicMaxRows && n <= kClassicMaxCols;
}

hipError_t launch_classic_fused(const double* X, double* Y, long ldx, int m,
                                int n, int i, const int* jX, int* jY,
                                const double* vX, double* vY, double* A,
                                long lda, int* jpvt, double* tau,
                                hipStream_t s) {
  if (i < 0 || i >= m || i >= n || !classic_fused_ok(m, n))
    return hipErrorInvalidValue;
  const int grid = (i + 1 < n) ? hq_cdiv(n - i - 1, kWaves) : 1;
  hipLaunchKernelGGL(classic_fused_kernel, dim3(grid), dim3(HQ_NT), 0, s, X,
                     Y, ldx, m, n, i, jX, jY, vX, vY, A, lda, jpvt, tau);
  return hipGetLastError();
}
