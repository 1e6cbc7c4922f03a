// k_mixed.hip -- the mixed-precision sketch Y = Omega A on the matrix cores
// (the sketch product of NoFLA_HQRRP_WY_blk_var4.c:407-409; the FP64
// factorisation is unchanged, the sketch only steers the pivot choice).
//
//   FP32: v_mfma_f32_32x32x2_f32, A and Omega rounded to FP32;
//   BF16: v_mfma_f32_32x32x16_bf16, Omega rounded to bf16 and A split into
//         two bf16 terms (hi = bf16(a), lo = bf16(a - hi): 16 significant
//         bits), both products accumulated in FP32.
//
// The driver rounds its own FP64 copy of Omega' to the same FP32 / bf16
// values before the product (round_omega), so the downdate it applies later
// (in FP64) is that of exactly the operator used here, and the only error
// of the sketch is A's rounding.  A single bf16 term would leave an
// absolute error of ~2^-9 of the initial sketch that no downdate removes;
// once the trailing matrix is 2^-9 smaller than A that error is as large as
// the sketch itself and the pivots become noise.  Two terms push that point
// to ~2^-16.  bf16 and FP32 share FP32's exponent range, and the split is
// relative to each entry, so no per-column scaling is needed (entries of A
// beyond FP32's range, ~1e+-38, are not supported by the mixed sketch).
//
// Layout: Y' (n x r) = A' (n x m) G' (m x r).  A workgroup of four
// wavefronts takes 128 columns of A and one K slice of rows; wavefront w
// owns columns 32w..32w+31 and all r sketch rows (three 32-row tiles).  For
// each 16 rows of K, lane l loads rows 8(l/32)..8(l/32)+7 of column l%32 of
// its tile (64 contiguous bytes, the next 16 rows in flight meanwhile) and
// the same K rows of the sketch operator from LDS, where each 64-row chunk
// of G' is staged (converted, [column][row]) by the whole workgroup.  The
// FP32 instruction takes two K values per lane pair, so its 16-row step is
// eight instructions over K permuted as (k, k + 8) -- the same permutation
// for both operands.  Accumulator element v of lane l holds
// D(8(v/4) + 4(l/32) + v%4, l%32): Y row 32t + l%32, column 8(v/4) +
// 4(l/32) + v%4 of the tile -- 32 lanes store 128 contiguous bytes.
#include "common.h"
#include "kernels.h"

namespace {

typedef float hq_f16v __attribute__((ext_vector_type(16)));
typedef __bf16 hq_bf8 __attribute__((ext_vector_type(8)));

constexpr int kMT = 4;          // wavefronts per workgroup
constexpr int kMC = 32 * kMT;   // columns of A per workgroup
constexpr int kMK = 64;         // K rows per LDS chunk
constexpr int kNT = 3;          // 32-row tiles of the sketch: r <= 96
constexpr int kRP = 32 * kNT;   // sketch rows covered

__device__ __forceinline__ __bf16 bf16_rne(float v) {
  unsigned b = __float_as_uint(v);
  b += 0x7FFFu + ((b >> 16) & 1u);
  const unsigned short h = (unsigned short)(b >> 16);
  return __builtin_bit_cast(__bf16, h);
}
__device__ __forceinline__ float bf16_to_f32(__bf16 v) {
  const unsigned short h = __builtin_bit_cast(unsigned short, v);
  return __uint_as_float((unsigned)h << 16);
}

template <int PREC>
struct StageT;
template <>
struct StageT<1> {
  typedef float T;
};
template <>
struct StageT<2> {
  typedef __bf16 T;
};

// PREC 1: FP32, 2: BF16 (two terms).
template <int PREC>
__global__ __launch_bounds__(HQ_NT) void sketch_mixed_kernel(MixedSketchArgs a) {
  typedef typename StageT<PREC>::T ST;
  constexpr int kPad = (PREC == 1) ? 4 : 8;  // keeps 16-byte row alignment
  __shared__ __align__(16) ST Gs[kRP][kMK + kPad];  // Gs[c][k] = G'(k0+k, c)
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int l32 = lane & 31, h = lane >> 5;
  const long c0 = (long)blockIdx.x * kMC + 32 * w;  // this wavefront's tile
  const long col = c0 + l32;                        // column this lane loads
  const bool cok = col < a.n;
  const long kbeg = (long)blockIdx.y * a.kchunk;
  const long kend = min((long)a.m, kbeg + a.kchunk);
  const double* acol = a.A + (cok ? col : 0) * a.lda;

  hq_f16v acc[kNT];
#pragma unroll
  for (int q = 0; q < kNT; ++q)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[q][v] = 0.0f;

  // Rows kk + 8h .. kk + 8h + 7 of my column (zero beyond the slice).
  auto load = [&](double (&x)[8], long kk) {
    const long r0 = kk + 8 * h;
#pragma unroll
    for (int e = 0; e < 8; ++e) x[e] = (cok && r0 + e < kend) ? acol[r0 + e] : 0.0;
  };
  double xa[8], xb[8];
  load(xa, kbeg);
  for (long k0 = kbeg; k0 < kend; k0 += kMK) {
    __syncthreads();  // the previous chunk's reads are done
    for (int idx = t; idx < kRP * kMK; idx += HQ_NT) {
      const int k = idx & (kMK - 1), c = idx / kMK;
      const long gk = k0 + k;
      const double g = (c < a.r && gk < kend) ? a.G[gk + (long)c * a.ldg] : 0.0;
      if (PREC == 1)
        Gs[c][k] = (float)g;
      else
        Gs[c][k] = bf16_rne((float)g);
    }
    __syncthreads();
#pragma unroll
    for (int sub = 0; sub < kMK / 16; ++sub) {
      const long kk = k0 + 16 * sub;
      if (kk >= kend) break;
      if (kk + 16 < kend) load(xb, kk + 16);
      const int kl = 16 * sub + 8 * h;  // this lane's first K row in the chunk
      if (PREC == 2) {
        hq_bf8 hi, lo;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const float f = (float)xa[e];
          hi[e] = bf16_rne(f);
          lo[e] = bf16_rne(f - bf16_to_f32(hi[e]));
        }
#pragma unroll
        for (int q = 0; q < kNT; ++q) {
          const hq_bf8 b =
              *reinterpret_cast<const hq_bf8*>(&Gs[32 * q + l32][kl]);
          acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi, b, acc[q], 0, 0,
                                                           0);
          acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lo, b, acc[q], 0, 0,
                                                           0);
        }
      } else {
#pragma unroll
        for (int q = 0; q < kNT; ++q) {
          const float* gq = reinterpret_cast<const float*>(&Gs[32 * q + l32][kl]);
#pragma unroll
          for (int e = 0; e < 8; ++e)
            acc[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(
                (float)xa[e], gq[e], acc[q], 0, 0, 0);
        }
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) xa[e] = xb[e];
    }
  }
  // Y row 32q + l32, tile column 8(v/4) + 4h + v%4.
  float* out = a.part ? a.part + (long)blockIdx.y * a.r * a.n : a.Y;
  const long ldo = a.part ? a.r : a.ldy;
#pragma unroll
  for (int q = 0; q < kNT; ++q) {
    const int row = 32 * q + l32;
    if (row >= a.r) continue;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const long c = c0 + 8 * (v / 4) + 4 * h + (v % 4);
      if (c < a.n) out[row + c * ldo] = acc[q][v];
    }
  }
}

// Rounds x[0, n) to the nearest FP32 (prec 1) or bf16 (prec 2) value, kept
// in FP64.
__global__ __launch_bounds__(HQ_NT) void round_omega_kernel(double* x, long n,
                                                            int prec) {
  for (long i = (long)blockIdx.x * HQ_NT + threadIdx.x; i < n;
       i += (long)gridDim.x * HQ_NT) {
    const float f = (float)x[i];
    x[i] = (prec == 2) ? (double)bf16_to_f32(bf16_rne(f)) : (double)f;
  }
}

}  // namespace

int mixed_sketch_rmax() { return kRP; }
int mixed_sketch_cols_per_wg() { return kMC; }
int mixed_sketch_kbk() { return kMK; }

hipError_t launch_sketch_mixed(const MixedSketchArgs& a, int prec, int splits,
                               hipStream_t s) {
  if (a.r < 1 || a.r > kRP || (prec != 1 && prec != 2) || splits < 1)
    return hipErrorInvalidValue;
  dim3 grid(hq_cdiv(a.n, kMC), splits);
  if (prec == 1)
    hipLaunchKernelGGL(sketch_mixed_kernel<1>, grid, dim3(HQ_NT), 0, s, a);
  else
    hipLaunchKernelGGL(sketch_mixed_kernel<2>, grid, dim3(HQ_NT), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_round_omega(double* x, long n, int prec, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  long blocks = (n + HQ_NT - 1) / HQ_NT;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(round_omega_kernel, dim3((unsigned)blocks), dim3(HQ_NT),
                     0, s, x, n, prec);
  return hipGetLastError();
}
