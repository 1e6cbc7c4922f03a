// k_downdate.hip -- the FP64 sketch downdate of one block (reference
// NoFLA_Downdate_Y, NoFLA_HQRRP_WY_blk_var4.c:617-708, with the G update of
// NoFLA_Apply_Q_WY_rnfc_blk_var4, 743-771):
//   P  = G U                 (r x b; U = the block's reflectors, m' x b,
//                             unit lower trapezoidal)
//   Q  = P T                 (T upper triangular)
//   G := G - Q U'            (so G(:, 0:b) becomes the reference's
//                             B = G1 - (G1 U11 + G2 U21) T U11')
//   Y2 := Y2 - G(:, 0:b) R12
// in four short launches: split-K partials of P over row slices of U (MFMA),
// a fixed-order sum of the slices (deterministic), Q and the first b columns
// of the new G on one workgroup per 16 rows, then one launch applying
// G2 -= Q U2' and Y2 -= B R12 on 64-column tiles.  Every operand is small
// (r <= 80, b <= 64): the products are latency-bound, so the design goal is
// few dependent launches with every load of a launch in flight together.
//
// v_mfma_f64_16x16x4f64 fragments: A lane l = A[l%16][l/16], B lane l =
// B[l/16][l%16]; accumulator element q of lane l = C[l/16 + 4q][l%16].
#include <algorithm>

#include "common.h"
#include "kernels.h"

namespace {

constexpr int kRB = 5;        // 16-row blocks of r (r <= 80)
constexpr int kRP = 16 * kRB + 4;  // LDS pitch over rows (doubles)
constexpr int kCP = 64 + 4;        // LDS pitch over 64 columns (doubles)
constexpr int kK = 64;             // rows of U per LDS stage (partials)

__device__ __forceinline__ double unit_lower_dd(double v, long i, long c) {
  return (i < c) ? 0.0 : ((i == c) ? 1.0 : v);
}

// part[s] (r x b, ld r) = G(:, Ks) U(Ks, :) over the rows
// Ks = [s*kc, min(mU, (s+1)*kc)).  Wavefront w owns columns 16w..16w+15 of
// the slice and every 16-row block of r.
__global__ __launch_bounds__(HQ_NT) void dd_partial_kernel(DowndateArgs a) {
  __shared__ double Gs[kK][kRP];  // Gs[k][i] = G(i, k0+k)
  __shared__ double Us[kK][kCP];  // Us[k][p] = U(k0+k, p)
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int r = a.r, b = a.b;
  const long kb = (long)blockIdx.x * a.kc;
  const long ke = min((long)a.mU, kb + a.kc);
  hq_d4 acc[kRB];
#pragma unroll
  for (int q = 0; q < kRB; ++q) acc[q] = hq_d4{0.0, 0.0, 0.0, 0.0};
  for (long k0 = kb; k0 < ke; k0 += kK) {
#pragma unroll
    for (int u = 0; u < kK * 16 * kRB / HQ_NT; ++u) {
      const int idx = t + u * HQ_NT;
      const int k = idx / (16 * kRB), i = idx % (16 * kRB);
      const long gk = k0 + k;
      Gs[k][i] = (i < r && gk < ke) ? a.G[i + gk * a.ldg] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kK * 64 / HQ_NT; ++u) {
      const int idx = t + u * HQ_NT;
      const int k = idx & (kK - 1), p = idx / kK;
      const long gk = k0 + k;
      double v = 0.0;
      if (p < b && gk < ke) v = unit_lower_dd(a.U[gk + (long)p * a.ldu], gk, p);
      Us[k][p] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kK; kk += 4) {
      const double bf = Us[kk + (lane >> 4)][wave * 16 + (lane & 15)];
#pragma unroll
      for (int q = 0; q < kRB; ++q) {
        const double af = Gs[kk + (lane >> 4)][q * 16 + (lane & 15)];
        acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(af, bf, acc[q], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  double* out = a.part + (long)blockIdx.x * r * b;
  const int p = wave * 16 + (lane & 15);
#pragma unroll
  for (int q = 0; q < kRB; ++q)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = q * 16 + (lane >> 4) + 4 * e;
      if (i < r && p < b) out[i + (long)p * r] = acc[q][e];
    }
}

// P = sum_s part[s], the slices added in a fixed order: wavefront w of a
// workgroup sums slices w, w+4, ... of the workgroup's 64 entries, then the
// four sums are added in wavefront order.
__global__ __launch_bounds__(HQ_NT) void dd_reduce_kernel(DowndateArgs a) {
  __shared__ double s4[4][64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const long total = (long)a.r * a.b;
  const long o = (long)blockIdx.x * 64 + lane;
  double s = 0.0;
  if (o < total) {
    constexpr int U = 16;
    int q = wave;
    for (; q + 4 * (U - 1) < a.S; q += 4 * U) {
      double v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = a.part[(long)(q + 4 * u) * total + o];
#pragma unroll
      for (int u = 0; u < U; ++u) s += v[u];
    }
    for (; q < a.S; q += 4) s += a.part[(long)q * total + o];
  }
  s4[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && o < total)
    a.P[o] = ((s4[0][lane] + s4[1][lane]) + s4[2][lane]) + s4[3][lane];
}

// Rows 16*blockIdx.x .. +15 of Q = P T and of B = G1 - Q U1' (U1 = the top
// b x b block of U, unit lower); B overwrites G(:, 0:b).
__global__ __launch_bounds__(HQ_NT) void dd_qb_kernel(DowndateArgs a) {
  __shared__ double Ps[16][kCP];
  __shared__ double Qs[16][kCP];
  __shared__ double Ts[64][kCP];  // Ts[p][q] = T(p, q), p <= q
  __shared__ double Us[64][kCP];  // Us[c][p] = U1(c, p)
  const int t = threadIdx.x, r = a.r, b = a.b;
  const int i0 = blockIdx.x * 16;
  #pragma unroll
  for (int u_ = 0; u_ < (16 * 64) / HQ_NT; ++u_) {
    const int idx = t + u_ * HQ_NT;
    const int i = idx & 15, p = idx >> 4;
    Ps[i][p] = (i0 + i < r && p < b) ? a.P[i0 + i + (long)p * r] : 0.0;
  }
  #pragma unroll
  for (int u_ = 0; u_ < (64 * 64) / HQ_NT; ++u_) {
    const int idx = t + u_ * HQ_NT;
    const int p = idx & 63, q = idx >> 6;
    Ts[p][q] = (p <= q && q < b) ? a.T[p + (long)q * a.ldt] : 0.0;
    Us[p][q] = (q < b && p < b) ? unit_lower_dd(a.U[p + (long)q * a.ldu], p, q)
                                : 0.0;
  }
  __syncthreads();
  #pragma unroll
  for (int u_ = 0; u_ < (16 * 64) / HQ_NT; ++u_) {
    const int idx = t + u_ * HQ_NT;
    const int i = idx & 15, q = idx >> 4;
    double s = 0.0;
    for (int p = 0; p <= q; ++p) s += Ps[i][p] * Ts[p][q];
    Qs[i][q] = s;
    if (i0 + i < r && q < b) a.Q[i0 + i + (long)q * r] = s;
  }
  __syncthreads();
  #pragma unroll
  for (int u_ = 0; u_ < (16 * 64) / HQ_NT; ++u_) {
    const int idx = t + u_ * HQ_NT;
    const int i = idx & 15, c = idx >> 4;
    if (i0 + i >= r || c >= b) continue;
    double s = 0.0;
    for (int p = 0; p <= c; ++p) s += Qs[i][p] * Us[c][p];
    double* g = a.G + i0 + i + (long)c * a.ldg;
    *g = *g - s;
  }
}

// Workgroups [0, ng): G(:, b+64g .. +63) -= Q U(b+64g .., :)'.
// Workgroups [ng, ng+ny): Y2(:, 64y .. +63) -= G(:, 0:b) R12(:, 64y ..).
// Wavefront w owns columns 16w..16w+15 of the tile and every 16-row block;
// the tile is loaded straight into the accumulators.
__global__ __launch_bounds__(HQ_NT) void dd_apply_kernel(DowndateArgs a) {
  __shared__ double As[64][kRP];  // As[p][i] = -Q(i, p) or -B(i, p)
  __shared__ double Bs[64][kCP];  // Bs[p][c] = U(c0+c, p) or R12(p, c0+c)
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int r = a.r, b = a.b;
  const bool isg = (int)blockIdx.x < a.ng;
  const double* Asrc = isg ? a.Q : a.G;
  const long ldas = isg ? r : a.ldg;
  long c0, ncol;
  double* out;
  long ldo;
  if (isg) {
    c0 = b + 64L * blockIdx.x;
    ncol = a.mU;
    out = a.G;
    ldo = a.ldg;
  } else {
    c0 = 64L * (blockIdx.x - a.ng);
    ncol = a.nY2;
    out = a.Y2;
    ldo = a.ldy;
  }
  #pragma unroll
  for (int u_ = 0; u_ < (64 * 16 * kRB) / HQ_NT; ++u_) {
    const int idx = t + u_ * HQ_NT;
    const int i = idx % (16 * kRB), p = idx / (16 * kRB);
    As[p][i] = (i < r && p < b) ? -Asrc[i + (long)p * ldas] : 0.0;
  }
  if (isg) {
    #pragma unroll
    for (int u_ = 0; u_ < (64 * 64) / HQ_NT; ++u_) {
      const int idx = t + u_ * HQ_NT;
      const int c = idx & 63, p = idx >> 6;
      const long gc = c0 + c;
      Bs[p][c] = (gc < ncol && p < b) ? a.U[gc + (long)p * a.ldu] : 0.0;
    }
  } else {
    #pragma unroll
    for (int u_ = 0; u_ < (64 * 64) / HQ_NT; ++u_) {
      const int idx = t + u_ * HQ_NT;
      const int p = idx & 63, c = idx >> 6;
      const long gc = c0 + c;
      Bs[p][c] = (gc < ncol && p < b) ? a.R12[p + gc * a.ldr] : 0.0;
    }
  }
  const long gc = c0 + wave * 16 + (lane & 15);
  const bool cok = gc < ncol;
  hq_d4 acc[kRB];
#pragma unroll
  for (int q = 0; q < kRB; ++q)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = q * 16 + (lane >> 4) + 4 * e;
      acc[q][e] = (cok && i < r) ? out[i + gc * ldo] : 0.0;
    }
  __syncthreads();
#pragma unroll
  for (int kk = 0; kk < 64; kk += 4) {
    const double bf = Bs[kk + (lane >> 4)][wave * 16 + (lane & 15)];
#pragma unroll
    for (int q = 0; q < kRB; ++q) {
      const double af = As[kk + (lane >> 4)][q * 16 + (lane & 15)];
      acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(af, bf, acc[q], 0, 0, 0);
    }
  }
#pragma unroll
  for (int q = 0; q < kRB; ++q)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = q * 16 + (lane >> 4) + 4 * e;
      if (cok && i < r) out[i + gc * ldo] = acc[q][e];
    }
}

}  // namespace

int downdate_slices(long mU, long* kc) {
  long s = std::min((mU + kK - 1) / kK, 128L);
  if (s < 1) s = 1;
  long c = (mU + s - 1) / s;
  c = (c + kK - 1) / kK * kK;
  *kc = c;
  return (int)std::max(1L, (mU + c - 1) / c);
}

hipError_t launch_downdate(DowndateArgs a, hipStream_t s) {
  if (a.r < 1 || a.r > 16 * kRB || a.b < 1 || a.b > 64 || a.mU < a.b ||
      a.S < 1)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(dd_partial_kernel, dim3(a.S), dim3(HQ_NT), 0, s, a);
  const long total = (long)a.r * a.b;
  hipLaunchKernelGGL(dd_reduce_kernel, dim3(hq_cdiv(total, 64)), dim3(HQ_NT),
                     0, s, a);
  hipLaunchKernelGGL(dd_qb_kernel, dim3(hq_cdiv(a.r, 16)), dim3(HQ_NT), 0, s,
                     a);
  a.ng = hq_cdiv(a.mU - a.b, 64);
  const int ny = hq_cdiv(a.nY2, 64);
  if (a.ng + ny > 0)
    hipLaunchKernelGGL(dd_apply_kernel, dim3(a.ng + ny), dim3(HQ_NT), 0, s, a);
  return hipGetLastError();
}
