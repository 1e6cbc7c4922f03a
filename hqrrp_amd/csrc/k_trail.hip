// k_trail.hip -- the trailing update of a block, the FP64 hot spot:
//   [A12; A22] := (I - V T V')' [A12; A22]
// (dlarfb Left/Transpose/Forward/Columnwise, reference
// NoFLA_HQRRP_WY_blk_var4.c:529-541 and 711-740), split in three kernels:
//   trail_vtc:    M_s = V(Ks,:)' C(Ks,:)   split-K partial products (MFMA)
//   trail_tw:     W = T' * sum_s M_s       fixed-order reduction, then T'
//   trail_update: C -= V W                 rank-b update (MFMA)
// V is read straight out of the factored panel as a unit lower trapezoidal
// matrix (R above its diagonal is masked off).
//
// Rooflines (b = 64): trail_vtc reads C once, 8 B per 2b = 128 flops
// (16 flop/B, at the MI355X ridge of ~9.8 flop/B it is MFMA-bound);
// trail_update reads and writes C once, 16 B per 128 flops (8 flop/B,
// HBM-bound).  trail_update therefore loads its C tile directly into the
// MFMA accumulators and multiplies by -V, so C makes exactly one round trip.
//
// v_mfma_f64_16x16x4f64 fragments: A lane l = A[l%16][l/16], B lane l =
// B[l/16][l%16]; accumulator element r of lane l = C[l/16 + 4r][l%16].
#include "common.h"
#include "kernels.h"

namespace {

constexpr int kTile = 64;  // output tile edge (rows and columns)
constexpr int kPitch = 72; // LDS row pitch in doubles
constexpr int kBK = 32;    // K rows staged per step in trail_vtc

__device__ __forceinline__ double vmask(double v, long i, long c) {
  return (i < c) ? 0.0 : ((i == c) ? 1.0 : v);
}

// C(mrows x nc) -= V(mrows x b) * W(b x nc).
__global__ __launch_bounds__(HQ_NT, 2) void trail_update_kernel(TrailArgs a) {
  __shared__ double Vs[kTile][kPitch];  // Vs[k][m] = -V(m0+m, k)
  __shared__ double Ws[kTile][kPitch];  // Ws[k][n] = W(k, n0+n)
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const long m0 = (long)blockIdx.x * kTile, n0 = (long)blockIdx.y * kTile;
  const int b = a.b, kpad = (b + 3) & ~3;

  hq_d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        long gm = m0 + wm * 32 + i * 16 + (lane >> 4) + 4 * r;
        long gn = n0 + wn * 32 + j * 16 + (lane & 15);
        acc[i][j][r] =
            (gm < a.mrows && gn < a.nc) ? a.C[gm + gn * a.ldc] : 0.0;
      }
  for (int idx = t; idx < kTile * kpad; idx += HQ_NT) {
    int m = idx & 63, k = idx >> 6;
    long gm = m0 + m;
    double v = 0.0;
    if (gm < a.mrows && k < b) v = vmask(a.V[gm + (long)k * a.ldv], gm, k);
    Vs[k][m] = -v;
  }
  for (int idx = t; idx < kTile * kpad; idx += HQ_NT) {
    int k = idx % kpad, n = idx / kpad;
    long gn = n0 + n;
    Ws[k][n] = (k < b && gn < a.nc) ? a.W[k + gn * a.ldw] : 0.0;
  }
  __syncthreads();
  for (int kk = 0; kk < kpad; kk += 4) {
    double af[2], bf[2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
      af[i] = Vs[kk + (lane >> 4)][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
    for (int j = 0; j < 2; ++j)
      bf[j] = Ws[kk + (lane >> 4)][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] =
            __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        long gm = m0 + wm * 32 + i * 16 + (lane >> 4) + 4 * r;
        long gn = n0 + wn * 32 + j * 16 + (lane & 15);
        if (gm < a.mrows && gn < a.nc) a.C[gm + gn * a.ldc] = acc[i][j][r];
      }
}

// part[s](b x nc) = V(Ks, :)' * C(Ks, :) over the K slice Ks of split s.
// K rows are staged kBK at a time; the next slice of V and C is loaded into
// registers while the matrix cores work on the current one.
__global__ __launch_bounds__(HQ_NT, 2) void trail_vtc_kernel(TrailArgs a) {
  __shared__ double Vs[kBK][kPitch];  // Vs[k][p] = V(k0+k, p)
  __shared__ double Cs[kBK][kPitch];  // Cs[k][n] = C(k0+k, n0+n)
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const long n0 = (long)blockIdx.x * kTile;
  const int s = blockIdx.y;
  const long kbeg = (long)s * a.kchunk;
  const long kend = min((long)a.mrows, kbeg + a.kchunk);
  const int b = a.b;
  constexpr int kPer = kBK * kTile / HQ_NT;  // 8 elements per thread

  hq_d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = hq_d4{0.0, 0.0, 0.0, 0.0};

  double rv[kPer], rc[kPer];
  auto load = [&](long k0) {
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      int idx = t + HQ_NT * q;
      int k = idx & (kBK - 1), p = idx / kBK;
      long gk = k0 + k;
      rv[q] = (gk < kend && p < b) ? vmask(a.V[gk + (long)p * a.ldv], gk, p)
                                   : 0.0;
      long gn = n0 + p;
      rc[q] = (gk < kend && gn < a.nc) ? a.C[gk + gn * a.ldc] : 0.0;
    }
  };
  if (kbeg < kend) load(kbeg);
  for (long k0 = kbeg; k0 < kend; k0 += kBK) {
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      int idx = t + HQ_NT * q;
      int k = idx & (kBK - 1), p = idx / kBK;
      Vs[k][p] = rv[q];
      Cs[k][p] = rc[q];
    }
    __syncthreads();
    if (k0 + kBK < kend) load(k0 + kBK);
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 4) {
      double af[2], bf[2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
        af[i] = Vs[kk + (lane >> 4)][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j)
        bf[j] = Cs[kk + (lane >> 4)][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j],
                                                            acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  double* out = a.part + (long)s * b * a.nc;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int p = wm * 32 + i * 16 + (lane >> 4) + 4 * r;
        long gn = n0 + wn * 32 + j * 16 + (lane & 15);
        if (p < b && gn < a.nc) out[p + gn * b] = acc[i][j][r];
      }
}

// W(:, n0:n0+64) = T' * sum_s part[s](:, n0:n0+64); slices added in order.
__global__ __launch_bounds__(HQ_NT) void trail_tw_kernel(TrailArgs a) {
  __shared__ double Ms[kTile][kTile + 1];  // Ms[p][n]
  __shared__ double Ts[kTile][kTile + 1];  // Ts[p][q] = T(p, q)
  const int t = threadIdx.x, b = a.b;
  const long n0 = (long)blockIdx.x * kTile;
  const long slice = (long)b * a.nc;
  for (int idx = t; idx < b * kTile; idx += HQ_NT) {
    int p = idx % b, n = idx / b;
    long gn = n0 + n;
    double s = 0.0;
    if (gn < a.nc)
      for (int q = 0; q < a.splits; ++q) s += a.part[q * slice + p + gn * b];
    Ms[p][n] = s;
  }
  for (int idx = t; idx < b * b; idx += HQ_NT) {
    int p = idx % b, q = idx / b;
    Ts[p][q] = (p <= q) ? a.T[p + (long)q * a.ldt] : 0.0;
  }
  __syncthreads();
  for (int idx = t; idx < b * kTile; idx += HQ_NT) {
    int q = idx % b, n = idx / b;
    long gn = n0 + n;
    if (gn >= a.nc) continue;
    double s = 0.0;
    for (int p = 0; p <= q; ++p) s += Ts[p][q] * Ms[p][n];
    a.W[q + gn * a.ldw] = s;
  }
}

}  // namespace

hipError_t launch_trail_vtc(const TrailArgs& a, hipStream_t s) {
  dim3 grid(hq_cdiv(a.nc, kTile), a.splits);
  hipLaunchKernelGGL(trail_vtc_kernel, grid, dim3(HQ_NT), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_trail_tw(const TrailArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(trail_tw_kernel, dim3(hq_cdiv(a.nc, kTile)), dim3(HQ_NT),
                     0, s, a);
  return hipGetLastError();
}

hipError_t launch_trail_update(const TrailArgs& a, hipStream_t s) {
  dim3 grid(hq_cdiv(a.mrows, kTile), hq_cdiv(a.nc, kTile));
  hipLaunchKernelGGL(trail_update_kernel, grid, dim3(HQ_NT), 0, s, a);
  return hipGetLastError();
}

int trail_kbk() { return kBK; }
