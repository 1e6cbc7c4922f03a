// k_sketch.hip -- pivot selection on the sketch (greedy QRCP, pivots only).
//
// Restates NoFLA_QRPmod_WY_unb_var4(pivoting=1, num_stages=b) applied to the
// working copy V_R of the sketch Y_R (reference NoFLA_HQRRP_WY_blk_var4.c:
// 500-506, 774-886, norms 889-909, Drmac downdate 912-986, swaps 990-1036).
//
// Design (MI355X): one persistent launch per block of b pivots.  Two lanes of
// a wavefront own one sketch column and keep it in registers for all b
// steps; a 512-thread workgroup holds 256 columns, so n = 16384 needs 64
// workgroups; up to 8192 columns, 256-thread workgroups of 128 columns (one
// wavefront per SIMD instead of two: a step's reflector application, which
// is bound by the SIMDs' FP64 issue, takes half as long).  A step is:
//   1. per-workgroup argmax of the partial norms (first position on ties),
//      the winning column published to this step's exchange slot;
//   2. every workgroup reads every slot header as soon as it is written (no
//      barrier: the exchange in common.h), reduces the candidates
//      identically, builds the same Householder reflector from the winner's
//      header and rows and applies it, the LAWN-176 norm downdate, and the
//      bookkeeping of positions, to its own columns.  Reading only the
//      winner's rows (one more round trip) is cheaper than polling every
//      candidate's rows: the polling traffic of 64 workgroups x 64 columns
//      slows the exchange more than the round trip costs.
// Columns never move: a column's *position* is tracked instead, which is what
// the reference's swap-not-shift semantics and idamax's first-index tie rule
// depend on.  At the end workgroup 0 composes the b swaps into the list of
// (dst, src) moves that the gather kernel applies to A, Y and jpvt.
#include <climits>

#include "common.h"
#include "kernels.h"

namespace {

constexpr int kSlotHead = 8;  // slot: [key, pos, alpha, ||x||^2, pad, rows]
// Slots start 1 KB apart, so no cache line holds parts of two workgroups'
// slots (lines written by several workgroups slow the polling).
constexpr long slot_stride(long used) { return (used + 127) / 128 * 128; }
// A workgroup of NT threads (NT / 64 wavefronts, 32 columns each) holds
// NT / 2 columns: 512 threads (256 columns), or 256 (128 columns) when every
// column fits in kNarrowWg workgroups.
constexpr int kNarrowWg = 64;
constexpr int kMaxWg = 256;    // workgroups of one launch at most
constexpr int kMaxStages = 80;

// Column c of the workgroup lives in wavefront c/32; lane c%32 holds the
// first RMAX/2 of the column's *active* rows and lane c%32+32 the next
// RMAX/2.  Active rows are rows k..r-1 at step k: after every step the
// registers shift up by one row (the boundary row crosses from the upper to
// the lower lane by one shuffle), so row k is always x[0] of the lower lane
// and every per-row operation uses static register indices with no tests
// against k.  Rows at or beyond r are zero.  A column's dot product is one
// cross-lane add (a+b == b+a, both lanes get the same bits).
//
// Exchange slot of a workgroup at step k: [0] key (partial norm of its best
// column, -1 if none), [1] its position, [2] its row k (alpha of the next
// reflector), [3] the squared norm of its rows below k, [kSlotHead + i] its
// row k + i.  Every workgroup polls the headers, picks the same winner,
// builds the reflector from the winner's alpha and norm, and reads the
// winner's rows.
//
// T is the sketch's arithmetic: double, or float for the mixed-precision
// sketch (values travel through the exchange as doubles, exactly).
template <class T, int RMAX, int NT>
__global__ __launch_bounds__(NT) void sketch_qrcp_kernel(SketchArgs a) {
  constexpr int NW = NT / 64;  // wavefronts
  constexpr int H = RMAX / 2;
  const int t = threadIdx.x, w = blockIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int half = lane >> 5;
  const int r = a.r;
  const int c = w * (NT / 2) + wave * 32 + (lane & 31);
  const bool own = c < a.ncols;
  const int ibase = half * H;  // first active row (relative) of this lane
  const long slot = kSlotHead + RMAX;
  const long stride = slot_stride(slot);

  __shared__ T s_v[RMAX];  // reflector on the active rows (relative)
  __shared__ double s_slot[kSlotHead + RMAX];  // this workgroup's slot
  __shared__ double s_key[8];
  __shared__ int s_pos[8];
  __shared__ int s_win_pos;
  __shared__ T s_tau;
  __shared__ int s_piv[kMaxStages];
  __shared__ int s_abort;

  const T* Yp = reinterpret_cast<const T*>(a.Y);
  T x[H];
#pragma unroll
  for (int q = 0; q < H; ++q) {
    const int i = ibase + q;
    x[q] = (own && i < r) ? Yp[(long)c * a.ldy + i] : T(0);
  }
  T ss = 0;
#pragma unroll
  for (int q = 0; q < H; ++q) ss += x[q] * x[q];
  ss += hq_xor32(ss);
  T d = sqrt(ss), e = d;
  // Squared norm of this column's rows below the first active one (the
  // ||x||^2 of its reflector, or LAWN 176's recompute): lane pair sum.
  // Four interleaved partial sums keep the dependent chain short.
  auto below2 = [&]() -> T {
    T s4[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int q = 1; q < H; ++q) s4[q & 3] += x[q] * x[q];
    s4[0] += half ? x[0] * x[0] : T(0);
    const T s1 = (s4[0] + s4[1]) + (s4[2] + s4[3]);
    return s1 + hq_xor32(s1);
  };
  int pos = own ? c : -1;  // current position; -1 once chosen as a pivot

  if (t == 0) s_abort = 0;
  __syncthreads();

  for (int k = 0; k < a.stages; ++k) {
    const double* cbuf = a.xcur + (long)k * a.nw * stride;
    double* mine = a.xcur + ((long)k * a.nw + w) * stride;
    const int span = r - k;  // active rows
    // 1. local argmax: largest norm, smallest position among equals; every
    //    thread combines the 8 wavefront winners identically and the
    //    winning column's lanes publish the slot.
    {
      // Lanes l and l+32 hold the same column: the upper half stands aside
      // (a duplicate would only make the argmax resolve a tie).
      double key = (pos >= 0 && half == 0) ? hq_pivot_key((double)d) : -1.0;
      int kp = (pos >= 0 && half == 0) ? pos : INT_MAX;
      hq_wave_argmax64(key, kp);
      if (lane == 0) {
        s_key[wave] = key;
        s_pos[wave] = kp;
      }
    }
    __syncthreads();
    {
      // Every wavefront combines the NW wavefront winners (lane l < NW
      // holds winner l; identical result in every lane); bq: the winner's
      // wavefront.
      const int wq = (lane < NW) ? s_pos[lane] : INT_MAX;
      double bk = (lane < NW) ? s_key[lane] : -1.0;
      int bp = wq;
      hq_wave_argmax64(bk, bp);
      const int bq =
          __builtin_ctzll(__ballot(wq == bp) & ((1ull << NW) - 1ull));
      if (bp == INT_MAX) {  // no candidate here
        if (t == 0) {
          hq_publish(mine + 0, -1.0);
          hq_publish(mine + 1, (double)INT_MAX);
          hq_publish(mine + 2, 0.0);
          hq_publish(mine + 3, 0.0);
        }
      } else if (wave == bq) {
        // The winner's two lanes stage the slot in LDS and their wavefront
        // publishes it with two coalesced stores (the LDS traffic stays
        // inside this wavefront, ordered by wavefront-scope fences).  Only
        // this wavefront computes the ||x||^2 that the slot carries.
        const T sqn = below2();
        if (pos == bp) {
#pragma unroll
          for (int q = 0; q < H; ++q) s_slot[kSlotHead + ibase + q] = x[q];
          if (half == 0) {
            s_slot[0] = bk;
            s_slot[1] = (double)bp;
            s_slot[2] = (double)x[0];
            s_slot[3] = (double)sqn;
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < 4 || (lane >= kSlotHead && lane < slot))
          hq_publish(mine + lane, s_slot[lane]);
        if (lane + 64 < slot) hq_publish(mine + lane + 64, s_slot[lane + 64]);
      }
    }
    // 2. wavefront 0: every header, the global winner, the reflector's
    //    scalars (dlarfg needs only alpha and ||x||^2) and the winner's rows
    //    below k, scaled, as the reflector (relative rows, 1 at row k, zero
    //    beyond span) -- the rows are fetched as soon as the winner is
    //    known and dlarfg runs while they are in flight; one workgroup
    //    barrier then hands the reflector to the other wavefronts.
    bool ok = true;
    if (wave == 0) {
      double bk = -2.0, ba = 0.0, bs = 0.0;
      int bp = INT_MAX, bw = -1;
      for (int q = lane; ok && q - lane < a.nw; q += 64) {
        double hv[4];
        ok = hq_await(hv, [&](int i) -> const double* {
          return q < a.nw ? cbuf + (long)q * stride + i : nullptr;
        }, a.err);
        const int pq = (int)hv[1];
        if (q < a.nw && (hv[0] > bk || (hv[0] == bk && pq < bp))) {
          bk = hv[0];
          bp = pq;
          bw = q;
          ba = hv[2];
          bs = hv[3];
        }
      }
      // The lane holding the winner (positions are unique) supplies its
      // workgroup, alpha and norm.
      const double lk = bk;
      const int lp = bp;
      hq_wave_argmax64(bk, bp);
      const int bl = __builtin_ctzll(__ballot(lk == bk && lp == bp));
      bw = hq_readlane(bw, bl);
      ba = hq_readlane(ba, bl);
      bs = hq_readlane(bs, bl);
      // Lane l: relative rows l and l + 64 of the winner.
      const double* src = cbuf + (long)bw * stride + kSlotHead;
      const bool f0 = lane > 0 && lane < span;
      const bool f1 = lane + 64 < span;
      double cv[2];
      cv[0] = f0 ? hq_fetch(src + lane) : 0.0;
      cv[1] = f1 ? hq_fetch(src + lane + 64) : 0.0;
      const hq_reflector_t<T> h = hq_larfg_direct<T>((T)ba, (T)bs);
      if (ok && (hq_pending(cv[0]) || hq_pending(cv[1])))
        ok = hq_await(cv, [&](int i) -> const double* {
          return (i == 0 ? f0 : f1) ? src + lane + 64 * i : nullptr;
        }, a.err);
      if (lane < RMAX) s_v[lane] = (lane == 0) ? T(1) : (T)cv[0] * h.scale;
      if (lane + 64 < RMAX) s_v[lane + 64] = (T)cv[1] * h.scale;
      if (lane == 0) {
        s_tau = h.tau;
        s_win_pos = bp;
        s_piv[k] = bp;
        if (w == 0) a.piv[k] = bp;
      }
      if (!ok) s_abort = 1;
    }
    __syncthreads();
    if (s_abort) return;
    const int p = s_win_pos;
    // 4. bookkeeping of positions (swap of positions k and p).
    if (pos >= 0) {
      if (pos == p) {
        pos = -1;
      } else if (pos == k) {
        pos = p;
      }
    }
    // 5. apply the reflector to my column, downdate its norm, and shift the
    //    active rows up by one (row k leaves).  Both lanes of a column take
    //    the same branches (pos is per column).
    if (pos >= 0) {
      const T tau = s_tau;
      if (tau != T(0)) {
        T w4[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int q = 0; q < H; ++q) w4[q & 3] += s_v[ibase + q] * x[q];
        T wd = (w4[0] + w4[1]) + (w4[2] + w4[3]);
        wd += hq_xor32(wd);
        const T f = -tau * wd;
#pragma unroll
        for (int q = 0; q < H; ++q) x[q] += f * s_v[ibase + q];
      }
      // LAWN-176 downdate with row k; the rows below k are summed only
      // when some column of the wavefront needs Drmac's recompute.
      const T x0 = x[0];
      const T up = hq_xor32(x0);  // partner's first row
      const T wt = half ? up : x0;  // row k
      // The reference's t = |w| / d, (1 + t)(1 - t), test t (d / e)^2 <=
      // tol3z and d sqrt((1 + t)(1 - t)), without divisions: (d - |w|)(d +
      // |w|) is d^2 (1 + t)(1 - t), compared with tol3z e^2.
      bool need = false;
      if (d != T(0)) {
        const T aw = fabs(wt);
        T u2 = (d - aw) * (d + aw);
        u2 = (T(0) > u2) ? T(0) : u2;
        need = u2 <= hq_tol3z<T>() * (e * e);
        if (!need) d = sqrt(u2);
      }
      if (__ballot(need)) {
        const T b2 = below2();
        if (need) {
          if (r - k - 1 > 0) {
            d = sqrt(b2);
            e = d;
          } else {
            d = T(0);
            e = T(0);
          }
        }
      }
#pragma unroll
      for (int q = 0; q < H - 1; ++q) x[q] = x[q + 1];
      x[H - 1] = half ? T(0) : up;
    }
  }
  // Composed permutation (workgroup 0, wavefront 0): replay the swaps of
  // positions k <-> piv[k] on a map position -> column held in registers
  // (lane l owns map entries l, l+64, l+128; at most 2*stages entries), then
  // emit (position, column) for every entry that moved.
  if (w == 0 && wave == 0) {
    int key[3] = {-1, -1, -1}, val[3] = {0, 0, 0};
    int cnt = 0;
    auto entry = [&](int ps) -> int {
#pragma unroll
      for (int u = 0; u < 3; ++u) {
        const unsigned long long m = __ballot(key[u] == ps);
        if (m) return 64 * u + __builtin_ctzll(m);
      }
      const int sidx = cnt++;
#pragma unroll
      for (int u = 0; u < 3; ++u)
        if (sidx >> 6 == u && lane == (sidx & 63)) {
          key[u] = ps;
          val[u] = ps;
        }
      return sidx;
    };
    auto get = [&](int sidx) -> int {
      int v = 0;
#pragma unroll
      for (int u = 0; u < 3; ++u)
        if (sidx >> 6 == u) v = __shfl(val[u], sidx & 63, 64);
      return v;
    };
    auto put = [&](int sidx, int v) {
#pragma unroll
      for (int u = 0; u < 3; ++u)
        if (sidx >> 6 == u && lane == (sidx & 63)) val[u] = v;
    };
    for (int k = 0; k < a.stages; ++k) {
      const int p = s_piv[k];
      if (p == k) continue;
      const int ia = entry(k), ib = entry(p);
      const int va = get(ia), vb = get(ib);
      put(ia, vb);
      put(ib, va);
    }
    int base = 0;
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      const bool mv = key[u] >= 0 && val[u] != key[u];
      const unsigned long long m = __ballot(mv);
      if (mv) {
        const int idx =
            base + __popcll(m & ((1ull << lane) - 1ull));
        a.moves[1 + 2 * idx] = key[u];
        a.moves[2 + 2 * idx] = val[u];
      }
      base += __popcll(m);
    }
    if (lane == 0) a.moves[0] = base;
  }
  hq_xreset(a.xprev, a.xprev_len, w, a.nw, NT);
}

}  // namespace

size_t sketch_xchg_doubles(int nw, int rmax, int stages) {
  return (size_t)stages * nw * slot_stride(kSlotHead + rmax);
}

int sketch_cols_per_wg(int ncols) {
  return ncols <= kNarrowWg * 128 ? 128 : 256;
}

int sketch_max_stages() { return kMaxStages; }

int sketch_rmax(int r) {
  if (r <= 32) return 32;
  if (r <= 80) return 80;
  return -1;
}

template <int NT>
hipError_t launch_sketch_nt(const SketchArgs& a, int rmax, hipStream_t stream) {
  dim3 grid(a.nw), block(NT);
  if (a.f32) {
    if (rmax == 32)
      hipLaunchKernelGGL((sketch_qrcp_kernel<float, 32, NT>), grid, block, 0,
                         stream, a);
    else if (rmax == 80)
      hipLaunchKernelGGL((sketch_qrcp_kernel<float, 80, NT>), grid, block, 0,
                         stream, a);
    else
      return hipErrorInvalidValue;
  } else if (rmax == 32) {
    hipLaunchKernelGGL((sketch_qrcp_kernel<double, 32, NT>), grid, block, 0,
                       stream, a);
  } else if (rmax == 80) {
    hipLaunchKernelGGL((sketch_qrcp_kernel<double, 80, NT>), grid, block, 0,
                       stream, a);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_sketch_qrcp(const SketchArgs& a, int rmax,
                              hipStream_t stream) {
  if (a.nw < 1 || a.nw > kMaxWg || a.stages < 1 || a.stages > kMaxStages ||
      (a.cols != 128 && a.cols != 256) || (long)a.nw * a.cols < a.ncols)
    return hipErrorInvalidValue;
  return a.cols == 128 ? launch_sketch_nt<256>(a, rmax, stream)
                       : launch_sketch_nt<512>(a, rmax, stream);
}
