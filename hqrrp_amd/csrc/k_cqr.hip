// k_cqr.hip -- panel QRCP from two Gram passes: pivoted Cholesky QR with the
// reference's pivot rule, a second Cholesky QR pass, and Householder
// reconstruction.
//
// Restates NoFLA_QRPmod_WY_unb_var4(panel_pivoting, -1, AB1, ..., build_T=1)
// (reference NoFLA_HQRRP_WY_blk_var4.c:520-524, 774-886; its dlarft at
// 875-878) for an mrows x b panel P, b <= 64, with a different order of
// operations that makes the same pivot decisions on the same quantities:
//
//   1. G = P'P (one pass over P).  Pivoted Cholesky of G: at step k the
//      Schur complement's row of the pivot p gives row k of R, R(k, c) =
//      S(p, c) / sqrt(S(p, p)), and its diagonal the squared norms of the
//      columns' parts below row k -- exactly the quantities the reference's
//      Householder steps produce.  The pivot is chosen as the reference
//      chooses it: idamax (first position among equals, positions following
//      the reference's swaps) over partial norms kept by LAWN 176's downdate
//      from |R(k, c)|, recomputed from the Schur diagonal when the downdate
//      loses accuracy (NoFLA_QRP_downdate_partial_norms, 912-986).  This
//      gives P Pi = Q1 R1 with Q1 = P Pi R1^-1.
//   2. G2 = Q1'Q1 (one pass), R2 = chol(G2): P Pi = Q R with Q = Q1 R2^-1
//      orthonormal to working precision and R = R2 R1 (Cholesky QR2,
//      Yamamoto, Nakatsukasa, Yanagisawa, Fukaya 2015).
//   3. The panel's Householder vectors and T from Q (Ballard, Demmel,
//      Grigori, Jacquelin, Knight, Nguyen 2015): the LU factorisation
//      without pivoting of [S; 0] - Q = L U, with S_k = -sign of the k-th
//      Schur pivot of Q's top block, gives V = L, T = U S L1^-T, tau_k =
//      U_kk S_k (in [1, 2]: LAPACK's dlarfg convention) and R_H = S R.
//
// Panels the rearrangement would not factor to working accuracy set a flag
// and are factored by the per-step panel kernel (k_panel.hip) instead, in
// the same stream order: a Cholesky pivot below 1e-6 of its column's squared
// norm (nearly dependent or zero columns; also bounds the error of every
// partial norm a decision compares to ~1e-13 of the column norm), a second
// Gram matrix far from I, or a column whose part below the diagonal
// vanishes (LAPACK then takes tau = 0; the last column of a square panel).
//
// One persistent launch per panel (all workgroups co-resident, one per CU):
// each workgroup holds 256 rows of P and then of Q1 in registers; the
// workgroups' Gram partials are summed in two levels through the
// barrier-free exchange of common.h (fixed order, deterministic), workgroup
// 0 does the 64 x 64 work and publishes the matrices the others apply.
#include <climits>

#include "common.h"
#include "kernels.h"
#include "select_tile.h"

namespace {

constexpr int kT = 512;       // threads: 8 wavefronts
constexpr int kRows = 256;    // panel rows per workgroup: 32 per wavefront
constexpr int kTiles = 10;    // upper 16 x 16 tiles of a 64 x 64 Gram matrix
constexpr int kPart = kTiles * 256;  // doubles of one Gram partial
constexpr int kGroup = 8;     // partials summed by one workgroup
static_assert(kPart % kT == 0, "whole Gram entries per thread");
constexpr int kXl = 33;       // LDS transposition: 64 columns x 32 rows
constexpr int LD = 65;        // leading dimension of 64 x 64 LDS matrices
constexpr int kMat = 64 * LD;
constexpr int kSmem = 4 * kMat + 32 * 33;  // >= 8 * 64 * kXl
static_assert(kSmem >= 8 * 64 * kXl, "transposition regions");
constexpr double kChol1Thr = 1e-6;   // pivot / squared column norm, pass 1

__device__ __forceinline__ int tile_i(int t) {
  return t < 4 ? 0 : t < 7 ? 1 : t < 9 ? 2 : 3;
}
__device__ __forceinline__ int tile_j(int t) {
  constexpr int tj[kTiles] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};
  return tj[t];
}

__device__ __forceinline__ hq_d4 mfma4(double a, double b, hq_d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------- row slabs
//
// A wavefront's 32 rows x 64 columns of a panel-shaped matrix X live in the
// "A layout": xa[i][j][e] = X(mw + 16 j + l%16, 16 i + l/16 + 4 e) -- the
// layout select_tile produces, and the operand layout of a product X Z on
// the matrix cores (X(m, 4 s + l/16) is xa[s/4][j][s%4]).

__device__ __forceinline__ void load_alayout(hq_d4 (&xa)[4][2], const double* P,
                                             long ldp, int mrows, int b,
                                             long mw) {
  const int l = threadIdx.x & 63, l15 = l & 15, l4 = l >> 4;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int q = 16 * i + l4 + 4 * e;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const long m = mw + 16 * j + l15;
        xa[i][j][e] = (q < b && m < mrows) ? P[m + (long)q * ldp] : 0.0;
      }
    }
}

// Y = X M for the wavefront's rows, M 64 x 64 in LDS (ld LD), computed
// transposed (Y' = M' X') so that Y is again in the A layout.
__device__ __forceinline__ void mul_alayout(hq_d4 (&y)[4][2],
                                            const hq_d4 (&xa)[4][2],
                                            const double* M) {
  const int l = threadIdx.x & 63, l15 = l & 15, l4 = l >> 4;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
#pragma unroll
    for (int j = 0; j < 2; ++j) y[c][j] = hq_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const double mv = M[(4 * s + l4) + LD * (16 * c + l15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) y[c][j] = mfma4(mv, xa[s >> 2][j][s & 3], y[c][j]);
    }
  }
}

// Adds the wavefront's part of the upper tiles of X'X to g (X in the A
// layout; sw: this wavefront's LDS region of 64 x kXl doubles).  The rows
// reach the K position of the products through LDS: fragment (ks, c) of lane
// l is X(8 (l/16) + ks, 16 c + l%16).
__device__ __forceinline__ void gram_acc(hq_d4 (&g)[kTiles],
                                         const hq_d4 (&xa)[4][2], double* sw) {
  const int l = threadIdx.x & 63, l15 = l & 15, l4 = l >> 4;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        sw[(16 * i + l4 + 4 * e) * kXl + 16 * j + l15] = xa[i][j][e];
  wave_sync();
#pragma unroll
  for (int ks = 0; ks < 8; ++ks) {
    double f[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) f[c] = sw[(16 * c + l15) * kXl + 8 * l4 + ks];
#pragma unroll
    for (int t = 0; t < kTiles; ++t)
      g[t] = mfma4(f[tile_i(t)], f[tile_j(t)], g[t]);
  }
}

// s[u] += the sum over q < nq (nq <= kGroup, in order) of X[q * kPart +
// t + kT u], u < kPer, every value polled until it no longer holds the
// exchange sentinel (common.h); all of a thread's values are in flight
// together.  The addresses are uniform bases plus one per-thread offset.
// Returns false after the spin limit or another workgroup's error.
constexpr int kPer = kPart / kT;
template <int U0, int NU>
__device__ __forceinline__ bool poll_sum_part(const double* X, int nq,
                                              double (&s)[kPer], int* err) {
  const int t = threadIdx.x;
  const double* x0 = X + t + kT * U0;
  double v[kGroup][NU];
#pragma unroll
  for (int q = 0; q < kGroup; ++q)
#pragma unroll
    for (int u = 0; u < NU; ++u)
      v[q][u] = (q < nq) ? hq_fetch(x0 + q * kPart + kT * u) : 0.0;
  for (unsigned spins = 0;; ++spins) {
    bool pend = false;
#pragma unroll
    for (int q = 0; q < kGroup; ++q)
#pragma unroll
      for (int u = 0; u < NU; ++u) pend |= hq_pending(v[q][u]);
    if (!pend) break;
    if (spins > HQ_SPIN_LIMIT) {
      __hip_atomic_store(err, HQ_ERR_BARRIER_TIMEOUT, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    if ((spins & 255) == 255 &&
        __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      return false;
    __builtin_amdgcn_s_sleep(1);
#pragma unroll
    for (int q = 0; q < kGroup; ++q)
#pragma unroll
      for (int u = 0; u < NU; ++u)
        if (hq_pending(v[q][u])) v[q][u] = hq_fetch(x0 + q * kPart + kT * u);
  }
#pragma unroll
  for (int u = 0; u < NU; ++u)
#pragma unroll
    for (int q = 0; q < kGroup; ++q)
      if (q < nq) s[U0 + u] += v[q][u];
  return true;
}
// Two rounds (entries 0-2, then 3-4) keep the registers in flight bounded.
__device__ __forceinline__ bool poll_sum(const double* X, int nq,
                                         double (&s)[kPer], int* err) {
  return poll_sum_part<0, 3>(X, nq, s, err) && poll_sum_part<3, 2>(X, nq, s, err);
}

// Sums the eight wavefronts' tiles (fixed tree) and publishes the
// workgroup's partial to xA; the first workgroup of each group of kGroup
// sums its group's partials in order into xB; workgroup 0 sums the group
// sums in order into G (LDS, ld LD, full symmetric 64 x 64).  Returns false
// (in every thread) when an exchange timed out.
__device__ __forceinline__ bool gram_exchange(hq_d4 (&g)[kTiles], double* red, double* xA,
                              double* xB, int nw, int* err, double* G,
                              int* s_abort) {
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63, w = blockIdx.x;
  __syncthreads();  // `red` overlaps the transposition regions
#pragma unroll
  for (int half = 4; half >= 1; half >>= 1) {
    if (wave >= half && wave < 2 * half) {
#pragma unroll
      for (int tt = 0; tt < kTiles; ++tt)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          red[(((wave - half) * kTiles + tt) * 64 + lane) * 4 + e] = g[tt][e];
    }
    __syncthreads();
    if (wave < half) {
#pragma unroll
      for (int tt = 0; tt < kTiles; ++tt)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          g[tt][e] += red[((wave * kTiles + tt) * 64 + lane) * 4 + e];
    }
    __syncthreads();
  }
  if (wave == 0) {
    double* mine = xA + (long)w * kPart;
#pragma unroll
    for (int tt = 0; tt < kTiles; ++tt)
#pragma unroll
      for (int e = 0; e < 4; ++e) hq_publish(mine + (tt * 64 + lane) * 4 + e, g[tt][e]);
  }
  bool ok = true;
  if (w % kGroup == 0) {
    double s[kPer] = {0.0, 0.0, 0.0, 0.0, 0.0};
    ok = poll_sum(xA + (long)w * kPart, min(kGroup, nw - w), s, err);
#pragma unroll
    for (int u = 0; u < kPer; ++u)
      hq_publish(xB + (long)(w / kGroup) * kPart + t + kT * u, s[u]);
  }
  if (w == 0) {
    const int ng = hq_cdiv(nw, kGroup);
    double s[kPer] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int q0 = 0; ok && q0 < ng; q0 += kGroup)
      ok = poll_sum(xB + (long)q0 * kPart, min(kGroup, ng - q0), s, err);
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int idx = t + kT * u;
      const int tt = idx >> 8, ln = (idx >> 2) & 63, e = idx & 3;
      const int row = 16 * tile_i(tt) + (ln >> 4) + 4 * e;
      const int col = 16 * tile_j(tt) + (ln & 15);
      G[row + LD * col] = s[u];
      G[col + LD * row] = s[u];
    }
  }
  if (!ok) *s_abort = 1;
  __syncthreads();
  return *s_abort == 0;
}

// Waits for the matrix workgroup 0 publishes at xC (4096 values, column-major
// 64 x 64, then a flag at xC[4096]).  Returns the flag (1: the per-step
// kernel takes the panel) in every thread, or -1 after a timeout; with flag
// 0 the matrix is in M (LDS, ld LD).  Thread 0 polls the flag alone, so the
// waiting workgroups keep the polling traffic low.
__device__ __forceinline__ int await_matrix(const double* xC, double* M, int* err, int* s_flag) {
  const int t = threadIdx.x;
  if (t == 0) {
    double f[1];
    const bool ok = hq_await(f, [&](int) -> const double* { return xC + 4096; },
                             err);
    *s_flag = ok ? (f[0] != 0.0 ? 1 : 0) : -1;
  }
  __syncthreads();
  const int flag = *s_flag;
  if (flag != 0) return flag;
  double v[8];
  const bool ok = hq_await(v, [&](int i) -> const double* {
    return xC + t + kT * i;
  }, err);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int idx = t + kT * i;
    M[(idx & 63) + LD * (idx >> 6)] = v[i];
  }
  if (!ok) *s_flag = -1;
  __syncthreads();
  return *s_flag;
}

// ------------------------------------------------------ 64 x 64 work (WG 0)
//
// 64 x 64 matrices in LDS, column-major with leading dimension LD.  b <= 64
// is the panel width; rows and columns from b on hold the identity (or
// zeros where stated), so every operation acts on the leading b x b block.

__device__ __forceinline__ hq_d4 tile_mm(const double* A, const double* B,
                                         int r0, int c0) {
  const int l = threadIdx.x & 63, l15 = l & 15, l4 = l >> 4;
  hq_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int k = 0; k < 64; k += 4)
    acc = mfma4(A[(r0 + l15) + LD * (k + l4)], B[(k + l4) + LD * (c0 + l15)], acc);
  return acc;
}

// C = alpha A B (each wavefront two tiles), handed to put(row, col, value).
// Callers synchronise around it.
template <class Put>
__device__ __forceinline__ void mm64(const double* A, const double* B,
                                     double alpha, Put put) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int tt = wave + 8 * u;
    const int r0 = 16 * (tt & 3), c0 = 16 * (tt >> 2);
    const hq_d4 acc = tile_mm(A, B, r0, c0);
#pragma unroll
    for (int e = 0; e < 4; ++e)
      put(r0 + (l >> 4) + 4 * e, c0 + (l & 15), alpha * acc[e]);
  }
}

// X = U^-1 for an upper triangular 64 x 64 U with U(i, j) = u(i, j) (i < j)
// and reciprocal diagonal dinv(i), by doubling: for adjacent diagonal
// blocks B1, B2 of size s = 1, 2, ..., 32, X(B1, B2) = -X(B1, B1) U(B1, B2)
// X(B2, B2).  X: LDS matrix, upper part written, strictly lower part zeroed;
// tmp: 32 x 33 doubles.  Blocks of 16 and 32 on the matrix cores.  Callers
// synchronise before; the state is synchronised on return.
template <class UF, class DF>
__device__ __forceinline__ void tri_inv_upper(double* X, UF u, DF dinv, double* tmp) {
  const int t = threadIdx.x, wave = t >> 6, l = t & 63;
  const int l15 = l & 15, l4 = l >> 4;
  for (int idx = t; idx < 64 * 64; idx += kT) {
    const int i = idx & 63, j = idx >> 6;
    X[i + LD * j] = (i == j) ? dinv(i) : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int lv = 0; lv < 4; ++lv) {  // s = 1, 2, 4, 8: one element a thread
    const int s = 1 << lv, per = s * s;
    const int e = t;
    const bool act = e < 32 * s;
    const int p = e >> (2 * lv), r = (e & (per - 1)) >> lv, c = e & (s - 1);
    const int c0 = 2 * s * p, row = c0 + r, col = c0 + s + c;
    if (act) {
      double y = 0.0;
      for (int q = 0; q <= c; ++q) y += u(row, c0 + s + q) * X[(c0 + s + q) + LD * col];
      tmp[e] = y;
    }
    __syncthreads();
    if (act) {
      double x = 0.0;
      for (int q = r; q < s; ++q) x += X[row + LD * (c0 + q)] * tmp[p * per + q * s + c];
      X[row + LD * col] = -x;
    }
    __syncthreads();
  }
  // s = 16: two pairs, one tile each (wavefronts 0, 1).
  if (wave < 2) {
    const int c0 = 32 * wave;
    hq_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 16; k += 4)
      acc = mfma4(u(c0 + l15, c0 + 16 + k + l4),
                  X[(c0 + 16 + k + l4) + LD * (c0 + 16 + l15)], acc);
#pragma unroll
    for (int e = 0; e < 4; ++e) tmp[(l4 + 4 * e) + 33 * (16 * wave + l15)] = acc[e];
  }
  __syncthreads();
  if (wave < 2) {
    const int c0 = 32 * wave;
    hq_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 16; k += 4)
      acc = mfma4(X[(c0 + l15) + LD * (c0 + k + l4)],
                  tmp[(k + l4) + 33 * (16 * wave + l15)], acc);
#pragma unroll
    for (int e = 0; e < 4; ++e) X[(c0 + l4 + 4 * e) + LD * (c0 + 16 + l15)] = -acc[e];
  }
  __syncthreads();
  // s = 32: one pair, four tiles (wavefronts 0..3).
  const int r0 = 16 * (wave & 1), cc0 = 16 * ((wave >> 1) & 1);
  if (wave < 4) {
    hq_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 32; k += 4)
      acc = mfma4(u(r0 + l15, 32 + k + l4),
                  X[(32 + k + l4) + LD * (32 + cc0 + l15)], acc);
#pragma unroll
    for (int e = 0; e < 4; ++e) tmp[(r0 + l4 + 4 * e) + 33 * (cc0 + l15)] = acc[e];
  }
  __syncthreads();
  if (wave < 4) {
    hq_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 32; k += 4)
      acc = mfma4(X[(r0 + l15) + LD * (k + l4)], tmp[(k + l4) + 33 * (cc0 + l15)], acc);
#pragma unroll
    for (int e = 0; e < 4; ++e) X[(r0 + l4 + 4 * e) + LD * (32 + cc0 + l15)] = -acc[e];
  }
  __syncthreads();
}

// Upper Cholesky factor of a matrix close to the identity, G = I + E (A:
// leading b x b block, the identity beyond; // caller: max |E| <= kNearI), as R = I + U by the fixed-point iteration U =
// up(E - U'U), whose strictly upper part is that of E - U'U and whose
// diagonal is half its diagonal (from R'R = I + U + U' + U'U = I + E).  The
// error contracts by about ||E|| <= 64 kNearI per iteration; the iteration
// stops when an update changes nothing (at most kCholIters).  Every product is a 64 x 64 tile product on
// the matrix cores.  A receives R (zero below the diagonal); X is scratch
// (U'U), s_flag one int.  Returns false if it did not converge.
constexpr int kCholIters = 12;
constexpr double kNearI = 1e-4;
__device__ __forceinline__ bool chol_near_identity(double* A, double* U,
                                                   double* X, int b,
                                                   int* s_flag) {
  const int t = threadIdx.x;
  // E in A (upper part used); U0 = up(E).
  for (int idx = t; idx < 64 * 64; idx += kT) {
    const int i = idx & 63, j = idx >> 6;
    const double e = (i < b && j < b) ? A[i + LD * j] - (i == j ? 1.0 : 0.0) : 0.0;
    U[i + LD * j] = (i < j) ? e : (i == j ? 0.5 * e : 0.0);
  }
  __syncthreads();
  bool conv = false;
  for (int it = 0; it < kCholIters && !conv; ++it) {
    if (t == 0) *s_flag = 0;
    // X = U'U (U upper: column tiles need rows up to their own).
    {
      const int wave = t >> 6, l = t & 63;
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int tt = wave + 8 * u;
        const int r0 = 16 * (tt & 3), c0 = 16 * (tt >> 2);
        hq_d4 acc = {0.0, 0.0, 0.0, 0.0};
        const int kend = min(r0, c0) + 16;  // U(k, r) = 0 for k > r
        for (int k = 0; k < kend; k += 4)
          acc = mfma4(U[(k + (l >> 4)) + LD * (r0 + (l & 15))],
                      U[(k + (l >> 4)) + LD * (c0 + (l & 15))], acc);
#pragma unroll
        for (int e = 0; e < 4; ++e) X[(r0 + (l >> 4) + 4 * e) + LD * (c0 + (l & 15))] = acc[e];
      }
    }
    __syncthreads();
    bool changed = false;
    for (int idx = t; idx < 64 * 64; idx += kT) {
      const int i = idx & 63, j = idx >> 6;
      if (i <= j && i < b && j < b) {
        const double e = A[i + LD * j] - (i == j ? 1.0 : 0.0);
        const double f = e - X[i + LD * j];
        const double un = (i == j) ? 0.5 * f : f;
        changed |= un != U[i + LD * j];
        U[i + LD * j] = un;
      }
    }
    if (changed) *s_flag = 1;
    __syncthreads();
    conv = *s_flag == 0;
    __syncthreads();
  }
  for (int idx = t; idx < 64 * 64; idx += kT) {
    const int i = idx & 63, j = idx >> 6;
    A[i + LD * j] = (i == j ? 1.0 : 0.0) + ((i <= j) ? U[i + LD * j] : 0.0);
  }
  __syncthreads();
  return conv;
}

// LU without pivoting of W + diag(S) (W in A, leading b x b block, zero
// beyond), S_k the sign of the k-th Schur pivot w (sign 0 = +), so that
// U(k, k) = S_k + w has |U(k, k)| >= 1; U on and above the diagonal of A,
// L's multipliers below (the identity beyond b), S_k, U(k, k) and 1 /
// U(k, k) in s_S[k], s_v[k], s_v[64 + k].  |w| = 1 means the column's part
// below the diagonal vanishes (LAPACK then takes tau = 0): returns false.
// By blocks of 16 columns: for each block,
// wavefront 0 eliminates the block's columns in registers (lane i: row i, 16
// steps without barriers, row k read from lane k), solves for the block row
// to the right with the block's unit lower triangle (lane j: column j), and
// all wavefronts apply the block's rank-16 update to the trailing matrix on
// the matrix cores.  The arithmetic of every entry is that of the unblocked
// elimination regrouped (W's Schur complements, the same multipliers).
__device__ __forceinline__ bool lu_blocked(double* A, int b, double* s_S,
                                           double* s_v, int* s_fail) {
  const int t = threadIdx.x, wave = t >> 6, l = t & 63;
  if (t == 0) *s_fail = 0;
  __syncthreads();
  for (int c0 = 0; c0 < 64; c0 += 16) {
    if (c0 >= b) break;
    // Block panel: rows c0..63 of columns c0..c0+15 (wavefront 0).
    if (wave == 0) {
      const int i = l;
      double x[16];
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) x[jj] = A[i + LD * (c0 + jj)];
      bool fail = false;
#pragma unroll
      for (int kk = 0; kk < 16; ++kk) {
        const int k = c0 + kk;
        if (k < b && !fail) {
          const double wkk = hq_readlane(x[kk], k);
          if (!(fabs(wkk) < 1.0 - 1e-12)) {
            fail = true;
          } else {
            const double sk = (wkk >= 0.0) ? 1.0 : -1.0;
            const double ukk = sk + wkk, uinv = 1.0 / ukk;
            if (l == 0) {
              s_S[k] = sk;
              s_v[k] = ukk;
              s_v[64 + k] = uinv;
            }
            const double lik = (i > k) ? x[kk] * uinv : 0.0;
#pragma unroll
            for (int jj = kk + 1; jj < 16; ++jj)
              x[jj] = __builtin_fma(-lik, hq_readlane(x[jj], k), x[jj]);
            if (i > k) x[kk] = lik;
            if (i == k) x[kk] = ukk;
          }
        }
      }
#pragma unroll
      for (int jj = 0; jj < 16; ++jj)
        if (i >= c0) A[i + LD * (c0 + jj)] = x[jj];
      if (l == 0 && fail) *s_fail = 1;
    }
    __syncthreads();
    if (*s_fail) break;
    const int c1 = c0 + 16;
    if (c1 >= 64) break;
    // Block row: U12 = L11^-1 A12 (lane j: column j >= c1; wavefront 0).
    if (wave == 0 && l >= c1) {
      const int j = l;
      double y[16];
#pragma unroll
      for (int ii = 0; ii < 16; ++ii) y[ii] = A[(c0 + ii) + LD * j];
#pragma unroll
      for (int ii = 0; ii < 16; ++ii)
#pragma unroll
        for (int i2 = ii + 1; i2 < 16; ++i2)
          y[i2] = __builtin_fma(-A[(c0 + i2) + LD * (c0 + ii)], y[ii], y[i2]);
#pragma unroll
      for (int ii = 0; ii < 16; ++ii) A[(c0 + ii) + LD * j] = y[ii];
    }
    __syncthreads();
    // Trailing update A22 -= L21 U12 (rows and columns c1..63, K = 16).
    const int nt = (64 - c1) / 16;
    for (int tt = wave; tt < nt * nt; tt += kT / 64) {
      const int r0 = c1 + 16 * (tt % nt), q0 = c1 + 16 * (tt / nt);
      hq_d4 acc;
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[e] = A[(r0 + (l >> 4) + 4 * e) + LD * (q0 + (l & 15))];
#pragma unroll
      for (int k = 0; k < 16; k += 4)
        acc = mfma4(-A[(r0 + (l & 15)) + LD * (c0 + k + (l >> 4))],
                    A[(c0 + k + (l >> 4)) + LD * (q0 + (l & 15))], acc);
#pragma unroll
      for (int e = 0; e < 4; ++e) A[(r0 + (l >> 4) + 4 * e) + LD * (q0 + (l & 15))] = acc[e];
    }
    __syncthreads();
  }
  __syncthreads();
  const bool ok = *s_fail == 0;
  if (ok) {
    // The identity beyond b (W was zero there: S = 1, U = 1, L = 0).
    for (int idx = t; idx < 64 * 64; idx += kT) {
      const int i = idx & 63, j = idx >> 6;
      if (i >= b || j >= b) A[i + LD * j] = (i == j) ? 1.0 : 0.0;
    }
  }
  __syncthreads();
  return ok;
}

// Step 1 on workgroup 0: pivoted Cholesky of G (in S: LDS, full symmetric,
// leading b x b block) with the reference's pivot rule, by blocks of 16
// steps.  Within a block wavefront 0 works left-looking and without
// barriers: lane c keeps column c's Schur diagonal, its LAWN 176 partial
// norm (vn1, vn2), position and state in registers, and row k of R
// (physical columns, stored to Rk, LD, zero on entry) is S(p, c) minus the
// block's earlier rows' contribution, over R(k, k).  After each block all
// wavefronts subtract the block's rank-16 contribution from S on the matrix
// cores.  On success s_perm[pos] = column.  Returns the outcome in every
// thread.
__device__ __forceinline__ bool chol_pivoted(double* S, double* Rk, int b,
                                             int mrows, int pivoting,
                                             int* s_perm, int* s_fail) {
  const int t = threadIdx.x, wave = t >> 6, c = t & 63;
  const double gc = S[c + LD * c];
  double d = gc;  // Schur diagonal: squared norm of the part below row k
  double vn1 = (c < b) ? sqrt(gc) : 0.0, vn2 = vn1;
  // The pivot key: vn1^2 as the downdate forms it (the same order as vn1;
  // the square root then stays off the argmax's path).
  double q2 = (c < b) ? gc : 0.0;
  int pos = c;
  bool done = c >= b;
  double rprev = 0.0;
  if (t == 0) *s_fail = 0;
  __syncthreads();
  for (int k0 = 0; k0 < b; k0 += 16) {
    const int k1 = min(k0 + 16, b);
    if (wave == 0) {
      bool ok = true;
      // This lane's rows k0.. of R (registers; the block's steps unrolled
      // so that the index is static).
      double rc[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) rc[i] = 0.0;
#pragma unroll
      for (int kk = 0; kk < 16; ++kk) {
        const int k = k0 + kk;
        if (k >= k1 || !ok) break;
        // LAWN 176 downdate with row k-1 of R (NoFLA_HQRRP_WY_blk_var4.c:
        // 941-986), in the same quantities without divisions: with t = |r|
        // / vn1, (1 + t)(1 - t) vn1^2 = (vn1 - |r|)(vn1 + |r|) and the
        // recompute test t (vn1 / vn2)^2 <= tol3z reads (vn1 - |r|)(vn1 +
        // |r|) <= tol3z vn2^2; the recompute takes the Schur diagonal.
        if (k > 0 && pivoting && !done && vn1 != 0.0) {
          const double ar = fabs(rprev);
          double u2 = (vn1 - ar) * (vn1 + ar);
          u2 = (0.0 > u2) ? 0.0 : u2;
          if (u2 <= HQ_TOL3Z * (vn2 * vn2)) {
            q2 = (mrows - (k - 1) - 1 > 0) ? (d > 0.0 ? d : 0.0) : 0.0;
            vn1 = sqrt(q2);
            vn2 = vn1;
          } else {
            q2 = u2;
            vn1 = sqrt(u2);
          }
        }
        // idamax over positions k..b-1 (largest norm, first position).
        int pk = k;
        if (pivoting) {
          double key = done ? -1.0 : hq_pivot_key(q2);
          int q = done ? INT_MAX : pos;
          hq_wave_argmax64(key, q);
          pk = q;
        }
        const int p = __builtin_ctzll(__ballot(!done && pos == pk));
        const bool pend = !done && c != p;
        if (c == p) {
          pos = k;
          done = true;
        } else if (pend && pos == k) {
          pos = pk;
        }
        // The block's earlier rows: R(k0:k, p)' R(k0:k, c), over all 16 rows
        // of the block so that every load is in flight at once: rows k.. of
        // Rk are still zero and add exact zeros.
        const double spc = S[p + LD * c];
        double rp[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) rp[i] = Rk[k0 + i + LD * p];
        const double dp = hq_readlane(d, p), gp = hq_readlane(gc, p);
        if (!(dp > kChol1Thr * gp)) {
          ok = false;
          break;
        }
        // 1 / sqrt(dp): the hardware estimate and two Newton steps.
        double rinv = __builtin_amdgcn_rsq(dp);
#pragma unroll
        for (int it = 0; it < 2; ++it)
          rinv = rinv * __builtin_fma(-0.5 * dp * rinv, rinv, 1.5);
        const double rkk = dp * rinv;
        double s4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int i = 0; i < 16; i += 4)
#pragma unroll
          for (int u = 0; u < 4; ++u) s4[u] = __builtin_fma(rp[i + u], rc[i + u], s4[u]);
        const double dot = (s4[0] + s4[1]) + (s4[2] + s4[3]);
        const double r = (c == p) ? rkk : (pend ? (spc - dot) * rinv : 0.0);
        if (pend) d = __builtin_fma(-r, r, d);
        rprev = r;
        rc[kk] = r;
        Rk[k + LD * c] = r;  // row k is not read before the next step
        wave_sync();
      }
      if (!ok && c == 0) *s_fail = 1;
    }
    __syncthreads();
    if (*s_fail) break;
    if (k1 < b) {
      // S -= R(k0:k0+16, :)' R(k0:k0+16, :) (16 tiles of 16 x 16, K = 16;
      // rows of Rk beyond k1 are zero).
      const int l = c;
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int tt = wave + 8 * u;
        const int r0 = 16 * (tt & 3), q0 = 16 * (tt >> 2);
        hq_d4 acc;
#pragma unroll
        for (int e = 0; e < 4; ++e) acc[e] = S[(r0 + (l >> 4) + 4 * e) + LD * (q0 + (l & 15))];
#pragma unroll
        for (int kk = 0; kk < 16; kk += 4)
          acc = mfma4(-Rk[(k0 + kk + (l >> 4)) + LD * (r0 + (l & 15))],
                      Rk[(k0 + kk + (l >> 4)) + LD * (q0 + (l & 15))], acc);
#pragma unroll
        for (int e = 0; e < 4; ++e) S[(r0 + (l >> 4) + 4 * e) + LD * (q0 + (l & 15))] = acc[e];
      }
    }
    __syncthreads();
  }
  const bool ok = *s_fail == 0;
  if (ok && wave == 0) s_perm[c < b ? pos : c] = c;
  __syncthreads();
  return ok;
}

__global__ __launch_bounds__(kT) void cqr_panel_kernel(CqrArgs a) {
  __shared__ double smem[kSmem];
  __shared__ double s_S[64], s_v2[128];
  __shared__ int s_src[64], s_perm[64];
  __shared__ int s_abort, s_flag, s_fail;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int l15 = lane & 15, l4 = lane >> 4;
  const int w = blockIdx.x, nw = a.nw, b = a.b;
  const int ng = hq_cdiv(nw, kGroup);
  const long mw = (long)w * kRows + 32 * wave;
  // Exchange layout: the two published matrices first (left for the next
  // launch to restore), then the two passes' partials, which every
  // workgroup restores as soon as it knows they have been consumed.
  const long npart = (long)(nw + ng) * kPart;
  double* xC2 = a.xcur;
  double* xC1 = xC2 + 4097;
  double* xA1 = xC1 + 4097;
  double* xB1 = xA1 + (long)nw * kPart;
  double* xA2 = xA1 + npart;
  double* xB2 = xA2 + (long)nw * kPart;
  // LDS: pass phases use [0, 8 * 64 * kXl) (the M staging at [0, kMat)
  // first); workgroup 0's 64 x 64 work uses four matrices and tmp.
  double* mA = smem;
  double* mB = smem + kMat;
  double* mC = smem + 2 * kMat;
  double* mD = smem + 3 * kMat;
  double* tmp = smem + 4 * kMat;
  if (t == 0) s_abort = 0;
  // The previous launch's exchange region (the other buffer) gets its
  // sentinel back now, overlapping the loads below.
  if (!a.xprev_if || *a.xprev_if != 0) hq_xreset(a.xprev, a.xprev_len, w, nw, kT);

  // ---- the panel's rows in registers (formed from the trailing matrix
  // under look-ahead, and stored: the fallback reads them)
  hq_d4 xa[4][2];
  if (a.sel_C) {
    if (t < 64) s_src[t] = t;
    __syncthreads();
    if (a.sel_moves && t < a.sel_moves[0]) {  // at most 2 * 80 moves < kT
      const int dst = a.sel_moves[1 + 2 * t];
      if (dst < a.sel_b2) s_src[dst] = a.sel_moves[2 + 2 * t];
    }
    __syncthreads();
    static_assert(kSmem >= 64 * kSelLd, "W' staging fits in smem");
    select_tile(xa, s_src, a.sel_b2, a.sel_C, a.sel_ldc, a.sel_r0, a.mrows,
                a.sel_V, a.sel_ldv, a.sel_W, a.sel_ldw, a.sel_b, mw, smem);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int q = 16 * i + l4 + 4 * e;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const long m = mw + 16 * j + l15;
          if (q < b && m < a.mrows) a.P[m + (long)q * a.ldp] = xa[i][j][e];
        }
      }
  } else {
    load_alayout(xa, a.P, a.ldp, a.mrows, b, mw);
  }

  // ---- pass 1: G = P'P; workgroup 0: pivoted Cholesky, M1 = Pi R1^-1.
  hq_d4 g[kTiles];
#pragma unroll
  for (int tt = 0; tt < kTiles; ++tt) g[tt] = hq_d4{0.0, 0.0, 0.0, 0.0};
  gram_acc(g, xa, smem + wave * 64 * kXl);
  if (!gram_exchange(g, smem, xA1, xB1, nw, a.err, mA, &s_abort)) return;
  if (w == 0) {
    // mA holds G; the identity beyond b.
    for (int idx = t; idx < 64 * 64; idx += kT) {
      const int i = idx & 63, j = idx >> 6;
      if (i >= b || j >= b) mA[i + LD * j] = (i == j) ? 1.0 : 0.0;
    }
    for (int idx = t; idx < kMat; idx += kT) mC[idx] = 0.0;
    __syncthreads();
    const bool ok = chol_pivoted(mA, mC, b, a.mrows, a.pivoting, s_perm, &s_fail);
    if (ok) {
      // R1 in pivoted order (mB and the workspace): R1(i, j) = Rk(i, perm j).
      for (int idx = t; idx < 64 * 64; idx += kT) {
        const int i = idx & 63, j = idx >> 6;
        const double v = (i < b && j < b) ? (i <= j ? mC[i + LD * s_perm[j]] : 0.0)
                                          : (i == j ? 1.0 : 0.0);
        mB[i + LD * j] = v;
        a.R1[idx] = v;
      }
      __syncthreads();
      // M1(perm[k], :) = R1^-1(k, :), published column-major.
      tri_inv_upper(
          mD, [&](int i, int j) { return mB[i + LD * j]; },
          [&](int i) { return 1.0 / mB[i + LD * i]; }, tmp);
      for (int idx = t; idx < 64 * 64; idx += kT) {
        const int k = idx & 63, j = idx >> 6;
        hq_publish(xC1 + s_perm[k] + 64 * j, mD[k + LD * j]);
      }
    } else if (t == 0) {
      *a.flag = 1;
    }
    if (t == 0) hq_publish(xC1 + 4096, ok ? 0.0 : 1.0);
    __syncthreads();
  }
  int f = await_matrix(xC1, smem, a.err, &s_flag);
  // Pass 1's partials are consumed once workgroup 0 has published (leaders
  // read xA1 before publishing xB1, workgroup 0 read xB1 before xC1).
  if (f >= 0) hq_xreset(xA1, npart, w, nw, kT);
  if (f != 0) return;

  // ---- pass 2: Q1 = P M1, G2 = Q1'Q1; workgroup 0: R2, the reconstruction.
  // The panel's rows are read again (each thread reads what it loaded or
  // stored above) rather than held in registers through pass 1.
  load_alayout(xa, a.P, a.ldp, a.mrows, b, mw);
  hq_d4 q1[4][2];
  mul_alayout(q1, xa, smem);
#pragma unroll
  for (int tt = 0; tt < kTiles; ++tt) g[tt] = hq_d4{0.0, 0.0, 0.0, 0.0};
  __syncthreads();  // the M1 staging is read before the regions are reused
  gram_acc(g, q1, smem + wave * 64 * kXl);
  if (!gram_exchange(g, smem, xA2, xB2, nw, a.err, mA, &s_abort)) return;
  if (w == 0) {
    for (int idx = t; idx < 64 * 64; idx += kT) {
      const int i = idx & 63, j = idx >> 6;
      if (i >= b || j >= b) mA[i + LD * j] = (i == j) ? 1.0 : 0.0;
    }
    if (t == 0) s_fail = 0;
    __syncthreads();
    for (int idx = t; idx < b * b; idx += kT) {
      const int i = idx % b, j = idx / b;
      if (i <= j && !(fabs(mA[i + LD * j] - (i == j ? 1.0 : 0.0)) <= kNearI)) s_fail = 1;
    }
    __syncthreads();
    bool ok = s_fail == 0 && chol_near_identity(mA, mB, mC, b, &s_fail);
    if (ok) {
      // mB = R2^-1; mD = R = R2 R1 (R1 from pass 1); mC = Q1's first 64
      // rows (this workgroup's wavefronts 0, 1 hold them).
      tri_inv_upper(
          mB, [&](int i, int j) { return mA[i + LD * j]; },
          [&](int i) { return 1.0 / mA[i + LD * i]; }, tmp);
      for (int idx = t; idx < 64 * 64; idx += kT) {
        const int i = idx & 63, j = idx >> 6;
        mC[i + LD * j] = a.R1[idx];
      }
      __syncthreads();
      mm64(mA, mC, 1.0, [&](int i, int j, double v) { mD[i + LD * j] = v; });
      __syncthreads();
      if (wave < 2) {
#pragma unroll
        for (int cq = 0; cq < 4; ++cq)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e)
              mC[(32 * wave + 16 * j + l15) + LD * (16 * cq + l4 + 4 * e)] = q1[cq][j][e];
      }
      __syncthreads();
      // W = -(Q's top block) = -Q1(0:64, :) R2^-1, zero outside b x b.
      mm64(mC, mB, -1.0, [&](int i, int j, double v) {
        mA[i + LD * j] = (i < b && j < b) ? v : 0.0;
      });
      __syncthreads();
      ok = lu_blocked(mA, b, s_S, s_v2, &s_fail);
    }
    if (ok) {
      // The identity beyond b; mC = U^-1; N2 = -R2^-1 U^-1 (published).
      for (int idx = t; idx < 64 * 64; idx += kT) {
        const int i = idx & 63, j = idx >> 6;
        if (i >= b || j >= b) mA[i + LD * j] = (i == j) ? 1.0 : 0.0;
      }
      if (t >= b && t < 64) s_S[t] = 1.0;
      __syncthreads();
      tri_inv_upper(
          mC, [&](int i, int j) { return mA[i + LD * j]; },
          [&](int i) { return 1.0 / mA[i + LD * i]; }, tmp);
      mm64(mB, mC, -1.0, [&](int i, int j, double v) { hq_publish(xC2 + i + 64 * j, v); });
      __syncthreads();
      if (t == 0) {
        *a.flag = 0;
        hq_publish(xC2 + 4096, 0.0);
      }
      // T = (U S) L1^-T: mC = (L1')^-1 (unit upper), mB = U S.
      tri_inv_upper(
          mC, [&](int i, int j) { return mA[j + LD * i]; },
          [&](int) { return 1.0; }, tmp);
      for (int idx = t; idx < 64 * 64; idx += kT) {
        const int i = idx & 63, j = idx >> 6;
        mB[i + LD * j] = (i <= j) ? mA[i + LD * j] * s_S[j] : 0.0;
      }
      __syncthreads();
      mm64(mB, mC, 1.0, [&](int i, int j, double v) {
        if (i < b && j < b) a.T[i + a.ldt * j] = (i <= j) ? v : 0.0;
      });
      // The panel's first b rows: R_H = S R on and above the diagonal, L
      // below; tau; the permutation.
      for (int idx = t; idx < b * b; idx += kT) {
        const int i = idx % b, j = idx / b;
        a.P[i + a.ldp * j] = (i <= j) ? s_S[i] * mD[i + LD * j] : mA[i + LD * j];
      }
      if (t < b) {
        a.tau[t] = mA[t + LD * t] * s_S[t];
        a.perm[t] = s_perm[t];
      }
    } else {
      if (t == 0) {
        *a.flag = 1;
        hq_publish(xC2 + 4096, 1.0);
      }
    }
    __syncthreads();
  }
  f = await_matrix(xC2, smem, a.err, &s_flag);
  if (f >= 0) hq_xreset(xA2, npart, w, nw, kT);
  if (f != 0) return;

  // ---- V = Q1 N2 below the panel's first b rows (V' = N2' Q1', so that 16
  // consecutive rows of a column are stored together).
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    hq_d4 v[2] = {hq_d4{0.0, 0.0, 0.0, 0.0}, hq_d4{0.0, 0.0, 0.0, 0.0}};
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const double nv = smem[(4 * s + l4) + LD * (16 * c + l15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) v[j] = mfma4(nv, q1[s >> 2][j][s & 3], v[j]);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int q = 16 * c + l4 + 4 * e;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const long m = mw + 16 * j + l15;
        if (q < b && m >= b && m < a.mrows) a.P[m + (long)q * a.ldp] = v[j][e];
      }
    }
  }
}

}  // namespace

int cqr_rows_per_wg() { return kRows; }

size_t cqr_xchg_dirty() { return 2 * 4097; }

size_t cqr_xchg_doubles(int nw) {
  const int ng = hq_cdiv(nw, kGroup);
  return 2 * ((size_t)(nw + ng) * kPart + 4097);
}

hipError_t launch_cqr_panel(const CqrArgs& a, hipStream_t stream) {
  if (a.b < 1 || a.b > 64 || a.mrows < a.b ||
      a.nw != hq_cdiv(a.mrows, kRows))
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(cqr_panel_kernel, dim3(a.nw), dim3(kT), 0, stream, a);
  return hipGetLastError();
}
