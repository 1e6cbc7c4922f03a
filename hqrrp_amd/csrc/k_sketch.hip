// k_sketch.hip -- pivot selection on the sketch (greedy QRCP, pivots only).
//
// Restates NoFLA_QRPmod_WY_unb_var4(pivoting=1, num_stages=b) applied to the
// working copy V_R of the sketch Y_R (reference NoFLA_HQRRP_WY_blk_var4.c:
// 500-506, 774-886, norms 889-909, Drmac downdate 912-986, swaps 990-1036).
//
// Design (MI355X): one persistent launch per block of b pivots.  Two lanes of
// a wavefront own one sketch column and keep it in registers for all b
// steps; a 512-thread workgroup holds 256 columns, so n = 16384 needs 64
// workgroups.  A step is:
//   1. per-workgroup argmax of the partial norms (first position on ties),
//      the winning column published to global memory;
//   2. one grid barrier;
//   3. every workgroup reduces the candidates identically, builds the same
//      Householder reflector from the winning column and applies it, the
//      LAWN-176 norm downdate, and the bookkeeping of positions, to its own
//      columns.
// Columns never move: a column's *position* is tracked instead, which is what
// the reference's swap-not-shift semantics and idamax's first-index tie rule
// depend on.  At the end each moved column appends (dst, src) to a list, the
// composed permutation that the gather kernel applies to A, Y and jpvt.
#include <climits>

#include "common.h"
#include "kernels.h"

namespace {

constexpr int kSlotHead = 8;  // candidate slot: [key, pos, ...pad, column]
constexpr int kThreads = 512;  // 8 wavefronts, 32 columns each
constexpr int kCols = 256;     // sketch columns per workgroup
constexpr int kMaxNw = 64;     // workgroups whose candidates are prefetched

// Column c of the workgroup lives in wavefront c/32; lane c%32 holds the
// first RMAX/2 of the column's *active* rows and lane c%32+32 the next
// RMAX/2.  Active rows are rows k..r-1 at step k: after every step the
// registers shift up by one row (the boundary row crosses from the upper to
// the lower lane by one shuffle), so row k is always x[0] of the lower lane
// and every per-row operation uses static register indices with no tests
// against k.  Rows at or beyond r are zero.  A column's dot product is one
// cross-lane add (a+b == b+a, both lanes get the same bits).
template <int RMAX>
__global__ __launch_bounds__(kThreads) void sketch_qrcp_kernel(SketchArgs a) {
  constexpr int H = RMAX / 2;
  const int t = threadIdx.x, w = blockIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int half = lane >> 5;
  const int r = a.r;
  const int c = w * kCols + wave * 32 + (lane & 31);
  const bool own = c < a.ncols;
  const int ibase = half * H;  // first active row (relative) of this lane
  const long slot = kSlotHead + RMAX;

  __shared__ double s_v[RMAX];    // reflector on the active rows (relative)
  __shared__ double s_col[RMAX];  // this workgroup's candidate column
  __shared__ double s_all[kMaxNw][RMAX];  // every workgroup's candidate
  __shared__ double s_key[8];
  __shared__ int s_pos[8];
  __shared__ int s_win_pos, s_win_wg;
  __shared__ double s_tau;

  double x[H];
#pragma unroll
  for (int q = 0; q < H; ++q) {
    const int i = ibase + q;
    x[q] = (own && i < r) ? a.Y[(long)c * a.ldy + i] : 0.0;
  }
  double ss = 0.0;
#pragma unroll
  for (int q = 0; q < H; ++q) ss += x[q] * x[q];
  ss += __shfl_xor(ss, 32, 64);
  double d = sqrt(ss), e = d;
  int pos = own ? c : -1;  // current position; -1 once chosen as a pivot
  int done_at = -1;

  if (w == 0 && t == 0) a.moves[0] = 0;
  unsigned phase = 0;

  for (int k = 0; k < a.stages; ++k) {
    double* cbuf = a.cand + (long)(k & 1) * a.nw * slot;
    const int span = r - k;  // active rows
    // 1. local argmax: largest norm, smallest position among equals.
    double key = (pos >= 0) ? d : -1.0;
    int kp = (pos >= 0) ? pos : INT_MAX;
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) {
      double ok = __shfl_xor(key, off, 64);
      int op = __shfl_xor(kp, off, 64);
      if (ok > key || (ok == key && op < kp)) {
        key = ok;
        kp = op;
      }
    }
    if (lane == 0) {
      s_key[wave] = key;
      s_pos[wave] = kp;
    }
    __syncthreads();
    if (t == 0) {
      double bk = s_key[0];
      int bp = s_pos[0];
      for (int q = 1; q < 8; ++q)
        if (s_key[q] > bk || (s_key[q] == bk && s_pos[q] < bp)) {
          bk = s_key[q];
          bp = s_pos[q];
        }
      hq_publish(cbuf + (long)w * slot + 0, bk);
      hq_publish(cbuf + (long)w * slot + 1, (double)bp);
      s_win_pos = bp;
    }
    __syncthreads();
    if (pos >= 0 && pos == s_win_pos) {
#pragma unroll
      for (int q = 0; q < H; ++q) s_col[ibase + q] = x[q];
    }
    __syncthreads();
    // Wavefront 0 publishes the candidate's active rows (hq_grid_sync
    // precondition).
    if (wave == 0 && s_win_pos != INT_MAX) {
      double* dst = cbuf + (long)w * slot + kSlotHead;
      if (lane < span) hq_publish(dst + lane, s_col[lane]);
      if (lane + 64 < span) hq_publish(dst + lane + 64, s_col[lane + 64]);
    }
    // 2. grid barrier
    if (!hq_grid_sync(a.bar, a.nw, phase, a.err)) return;
    // 3. global argmax over the workgroup candidates (wavefront 0) while the
    //    other wavefronts fetch every candidate column into LDS in the same
    //    memory round trip (when nw <= kMaxNw); the winner's column is then
    //    already local.
    const bool prefetch = a.nw <= kMaxNw;
    if (wave == 0) {
      double bk = -2.0;
      int bp = INT_MAX, bw = -1;
      for (int q = lane; q < a.nw; q += 64) {
        double kq = hq_fetch(cbuf + (long)q * slot + 0);
        int pq = (int)hq_fetch(cbuf + (long)q * slot + 1);
        if (kq > bk || (kq == bk && pq < bp)) {
          bk = kq;
          bp = pq;
          bw = q;
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        double ok = __shfl_xor(bk, off, 64);
        int op = __shfl_xor(bp, off, 64);
        int ow = __shfl_xor(bw, off, 64);
        if (ok > bk || (ok == bk && op < bp)) {
          bk = ok;
          bp = op;
          bw = ow;
        }
      }
      if (lane == 0) {
        s_win_pos = bp;
        s_win_wg = bw;
      }
    } else if (prefetch) {
      constexpr int kPF = (kMaxNw * RMAX + kThreads - 65) / (kThreads - 64);
      const int total = a.nw * span;
      double vbuf[kPF];
#pragma unroll
      for (int u = 0; u < kPF; ++u) {
        const int e2 = (t - 64) + u * (kThreads - 64);
        vbuf[u] = 0.0;
        if (e2 < total) {
          const int q = e2 / span, i = e2 % span;
          vbuf[u] = hq_fetch(cbuf + (long)q * slot + kSlotHead + i);
        }
      }
#pragma unroll
      for (int u = 0; u < kPF; ++u) {
        const int e2 = (t - 64) + u * (kThreads - 64);
        if (e2 < total) s_all[e2 / span][e2 % span] = vbuf[u];
      }
    }
    __syncthreads();
    const int p = s_win_pos;
    // 4. winning column (active rows) into wavefront 0's registers;
    // 5. reflector from a fixed-order wavefront reduction, identical in every
    //    workgroup; v (relative rows, zero beyond span) into LDS.
    if (wave == 0) {
      const int wg = s_win_wg;
      const double* src = cbuf + (long)wg * slot + kSlotHead;
      const int i0 = lane, i1 = lane + 64;
      double x0 = 0.0, x1 = 0.0;
      if (prefetch) {
        if (i0 < span) x0 = s_all[wg][i0];
        if (i1 < span) x1 = s_all[wg][i1];
      } else {
        if (i0 < span) x0 = hq_fetch(src + i0);
        if (i1 < span) x1 = hq_fetch(src + i1);
      }
      const double alpha = __shfl(x0, 0, 64);
      double sq = ((i0 > 0) ? x0 * x0 : 0.0) + x1 * x1;
      sq = hq_wave_sum(sq);
      hq_reflector h = hq_dlarfg(alpha, sq);
      if (i0 < RMAX) s_v[i0] = (i0 == 0) ? 1.0 : x0 * h.scale;
      if (i1 < RMAX) s_v[i1] = x1 * h.scale;
      if (lane == 0) {
        s_tau = h.tau;
        if (w == 0) a.piv[k] = p;
      }
    }
    __syncthreads();
    // 6. bookkeeping of positions (swap of positions k and p).
    if (pos >= 0) {
      if (pos == p) {
        pos = -1;
        done_at = k;
      } else if (pos == k) {
        pos = p;
      }
    }
    // 7. apply the reflector to my column, downdate its norm, and shift the
    //    active rows up by one (row k leaves).  Both lanes of a column take
    //    the same branches (pos is per column).
    if (pos >= 0) {
      const double tau = s_tau;
      if (tau != 0.0) {
        double wd = 0.0;
#pragma unroll
        for (int q = 0; q < H; ++q) wd += s_v[ibase + q] * x[q];
        wd += __shfl_xor(wd, 32, 64);
        const double f = -tau * wd;
#pragma unroll
        for (int q = 0; q < H; ++q) x[q] += f * s_v[ibase + q];
      }
      double b2 = 0.0;
#pragma unroll
      for (int q = 1; q < H; ++q) b2 += x[q] * x[q];
      const double x0 = x[0];
      const double up = __shfl_xor(x0, 32, 64);  // partner's first row
      b2 += half ? x0 * x0 : 0.0;
      b2 += __shfl_xor(b2, 32, 64);
      const double wt = half ? up : x0;  // row k
      hq_downdate_norm(d, e, wt, b2, r - k - 1);
#pragma unroll
      for (int q = 0; q < H - 1; ++q) x[q] = x[q + 1];
      x[H - 1] = half ? 0.0 : up;
    }
  }
  // Composed permutation: final position of every column that moved.
  if (own && half == 0) {
    const int fp = (done_at >= 0) ? done_at : pos;
    if (fp != c) {
      int idx = atomicAdd(a.moves, 1);
      a.moves[1 + 2 * idx] = fp;
      a.moves[2 + 2 * idx] = c;
    }
  }
  hq_grid_exit(a.bar, a.nw);
}

}  // namespace

size_t sketch_cand_doubles(int nw, int rmax) {
  return (size_t)2 * nw * (kSlotHead + rmax);
}

int sketch_cols_per_wg() { return kCols; }

int sketch_rmax(int r) {
  if (r <= 32) return 32;
  if (r <= 80) return 80;
  return -1;
}

hipError_t launch_sketch_qrcp(const SketchArgs& a, int rmax,
                              hipStream_t stream) {
  dim3 grid(a.nw), block(kThreads);
  if (rmax == 32)
    hipLaunchKernelGGL(sketch_qrcp_kernel<32>, grid, block, 0, stream, a);
  else if (rmax == 80)
    hipLaunchKernelGGL(sketch_qrcp_kernel<80>, grid, block, 0, stream, a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}
