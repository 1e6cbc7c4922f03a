// k_panel.hip -- panel factorisation: Householder QR with column pivoting
// inside an (m-j) x b panel, b <= 64, plus the compact-WY T factor.
//
// Restates NoFLA_QRPmod_WY_unb_var4(panel_pivoting, -1, AB1, ..., build_T=1)
// (reference NoFLA_HQRRP_WY_blk_var4.c:520-524, 774-886): per step an
// idamax over the partial norms, dlarfg on the pivot column, dlarf on the
// remaining columns, the LAWN-176 norm downdate (912-986), and at the end
// dlarft('Forward','Columnwise') (875-878).
//
// Design (MI355X): one persistent launch per panel.  Workgroup w owns 256
// rows of all b columns, resident in registers for the whole factorisation
// (thread = one column x 32 consecutive rows), so the panel is read from and
// written to HBM exactly once.  A step needs ONE grid-wide reduction (a
// barrier-free exchange, common.h): with p the pivot column and
// v = [1; scale * x_p(k+1:)] its reflector,
//   v' x_c = x_c(k) + scale * G_c,   G_c = x_p(k+1:)' x_c(k+1:),
// so every workgroup publishes its partial G_c for all columns (G_p is the
// ||x||^2 that dlarfg needs) and the owner of row k publishes row k.  From
// those every wavefront computes, redundantly and with identical bits, the
// reflector, the dots (for the pending columns: dlarf's w; for the factored
// columns: the inner products dlarft needs), row k of R (x_c(k) - tau * dot),
// the norm downdate and the next pivot -- no workgroup barrier beyond the
// exchange itself.  Only when a downdate loses accuracy (LAWN 176's
// recompute test) is a second exchange run, for the fresh column norms.
// The reflector entries reach the other columns' lanes through LDS from the
// pivot column's lane (each wavefront holds all b columns of its rows).
// Reductions are done in a fixed order everywhere.  Columns stay with their
// threads; a position -> column map records the pivoting and the slab is
// written back in pivoted order.
#include "common.h"
#include "kernels.h"
#include "select_tile.h"

namespace {

constexpr int kMaxB = 64;

constexpr int kPT = 512;          // threads of the panel kernel
constexpr int kPG = kPT / 64;     // row groups (one per wavefront)
constexpr int kRT = 32;           // rows per thread (per row group)
constexpr int kPR = kRT * kPG;    // panel rows held by one workgroup
constexpr int kXS = 2 * kMaxB;    // exchange slot of one workgroup: [0,64)
                                  // partial sums per column, [64,128) row k
                                  // (owner only)
constexpr int kChunk = 8;         // workgroups' partials awaited together

// dlarft('Forward','Columnwise') from the reflector inner products
// G(l,q) = v_l'v_q (l < q, gram[l + q ldg]): T is the compact-WY factor of
// H_0 H_1 ... with T(q,q) = tau_q and, for adjacent column blocks B1, B2,
//   T(B1,B2) = -T(B1,B1) G(B1,B2) T(B2,B2),
// built by doubling the block size (1, 2, 4, ..., 32): six levels of small
// triangular products instead of LAPACK's b-step recurrence (same matrix in
// exact arithmetic; a zero tau gives a zero column, as LAPACK's).  Run by
// every thread of one workgroup of NT threads; P: kMaxB x (kMaxB + 1)
// doubles of LDS holding T's upper triangle (with the diagonal) and G's
// strict upper part transposed below the diagonal.  A level's X(B1, B2) =
// G(B1, B2) T(B2, B2) goes to T(B1, B2)'s place, and T(B1, B2) is formed in
// registers and written after a barrier.  Every entry is the same sum in
// the same order whatever NT is.
template <int NT>
__device__ __forceinline__ void larft_lds(int b, const double* gram, long ldg,
                                          const double* tau, double* T,
                                          long ldt, double (*P)[kMaxB + 1]) {
  constexpr int kE = (kMaxB * kMaxB / 4 + NT - 1) / NT;  // a level's blocks
  const int t = threadIdx.x;
#pragma unroll
  for (int u = 0; u < kMaxB * kMaxB / NT; ++u) {
    const int idx = t + u * NT, i = idx & (kMaxB - 1), j = idx >> 6;
    const bool in = i < b && j < b;
    P[i][j] = !in ? 0.0 : (i == j ? tau[i] : (i < j ? 0.0 : gram[j + i * ldg]));
  }
  __syncthreads();
  // Level s: for the pairs of column blocks B1 = [c0, c0+s), B2 = [c0+s,
  // c0+2s), element (r, cc) of the s x s block (row c0+r, column c0+s+cc).
  // s is a compile-time power of two after unrolling, so the index split is
  // shifts and masks.
#pragma unroll
  for (int lv = 0; lv < 6; ++lv) {
    const int s2 = 1 << lv;
    if (s2 >= b) break;
    const int per = s2 * s2, npairs = (b + 2 * s2 - 1) / (2 * s2);
    // X(B1, B2) = G(B1, B2) T(B2, B2).
    for (int e = t; e < npairs * per; e += NT) {
      const int pr = e >> (2 * lv), rem = e & (per - 1);
      const int r = rem >> lv, cc = rem & (s2 - 1);
      const int c0 = pr * 2 * s2, c1 = c0 + s2, col = c1 + cc;
      if (col >= b) continue;
      double acc = 0.0;
#pragma unroll
      for (int u = 0; u < s2; ++u)  // u <= cc: T(B2, B2) is upper
        acc += (u <= cc) ? P[c1 + u][c0 + r] * P[c1 + u][col] : 0.0;
      P[c0 + r][col] = acc;
    }
    __syncthreads();
    // T(B1, B2) = -T(B1, B1) X(B1, B2).
    double nv[kE];
#pragma unroll
    for (int q = 0; q < kE; ++q) {
      const int e = t + q * NT;
      const int pr = e >> (2 * lv), rem = e & (per - 1);
      const int r = rem >> lv, cc = rem & (s2 - 1);
      const int c0 = pr * 2 * s2, col = c0 + s2 + cc;
      double acc = 0.0;
      if (e < npairs * per && col < b) {
#pragma unroll
        for (int u = 0; u < s2; ++u)  // u >= r: T(B1, B1) is upper
          acc += (u >= r) ? P[c0 + r][c0 + u] * P[c0 + u][col] : 0.0;
      }
      nv[q] = -acc;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kE; ++q) {
      const int e = t + q * NT;
      const int pr = e >> (2 * lv), rem = e & (per - 1);
      const int r = rem >> lv, cc = rem & (s2 - 1);
      const int c0 = pr * 2 * s2, col = c0 + s2 + cc;
      if (e < npairs * per && col < b) P[c0 + r][col] = nv[q];
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < kMaxB * kMaxB / NT; ++u) {
    const int idx = t + u * NT, i = idx & (kMaxB - 1), j = idx >> 6;
    if (i < b && j < b) T[i + (long)j * ldt] = (i <= j) ? P[i][j] : 0.0;
  }
}

// Thread (c, g) = (t % 64, t / 64) holds column c of the 32 consecutive rows
// row0 + 32g + i, i < 32, in registers x[i]; rows beyond mrows are zero.
// Relative to the current step k, a wavefront's rows are either all above
// k (finished: skipped), all below k (branch-free fast path), or contain row
// k (masked path, one wavefront of one workgroup per step).
__global__ __launch_bounds__(kPT) void panel_qrcp_kernel(PanelArgs a) {
  __shared__ double stage[kMaxB][129];  // staging of 128 rows x 64 columns
  __shared__ double s_red[kPG][kMaxB];  // this workgroup's per-group partials
  __shared__ double s_acc[kPG][kMaxB];  // per-group sums of published ones
  __shared__ double s_rowk[kMaxB];      // row k, staged by its owner group
  __shared__ double s_rowin[kMaxB];     // row k, as received
  __shared__ double s_pv[kPG][kRT];     // each group's rows of the pivot column
  __shared__ int s_abort;
  __shared__ int s_src[kMaxB];          // look-ahead: source column per position

  const int t = threadIdx.x, w = blockIdx.x;
  const int b = a.b, nw = a.nw;
  if (a.run_if && *a.run_if == 0) {
    hq_xreset(a.xprev, a.xprev_len, w, nw, kPT);
    return;
  }
  const long row0 = (long)w * kPR;
  const int cc = t & 63, g = t >> 6;
  const bool col_ok = cc < b;
  const long glo = row0 + kRT * g;  // first global row of my group
  // The wavefront holding row k keeps it in x[0]: after each step it stashes
  // its finished row in `stage` and shifts its registers up by one, so x[i]
  // holds row glo + sh + i and no per-row test against k is needed.  Only
  // rows k < b <= 64 are ever finished, so only wavefronts 0 and 1 of
  // workgroup 0 shift, and `stage` (idle during the steps) holds them.
  int sh = 0;
  // Row kk relative to my first unfinished row: -1 when all my rows are
  // below kk, 0 when kk is my first row, kRT when all my rows are done.
  auto rel = [&](long kk) -> int {
    if (sh >= kRT) return kRT;
    return kk < glo + sh ? -1 : 0;
  };
  if (t == 0) s_abort = 0;

  double x[kRT];
  if (a.sel_C) {
    // ---- look-ahead load: every wavefront forms its 32 rows of the
    // selected, updated columns on the matrix cores (select_tile.h), then
    // the tiles reach the column-per-thread layout through `stage`, 128 rows
    // per round.
    if (t < kMaxB) s_src[t] = t;
    __syncthreads();
    // The composed moves have distinct destinations: one thread per move.
    if (a.sel_moves && t < a.sel_moves[0]) {  // at most 2 * 80 moves < kPT
      const int dst = a.sel_moves[1 + 2 * t];
      if (dst < a.sel_b2) s_src[dst] = a.sel_moves[2 + 2 * t];
    }
    __syncthreads();
    hq_d4 acc[4][2];
    static_assert(kMaxB * 129 >= 64 * kSelLd, "W' staging fits in stage");
    select_tile(acc, s_src, a.sel_b2, a.sel_C, a.sel_ldc, a.sel_r0, a.mrows,
                a.sel_V, a.sel_ldv, a.sel_W, a.sel_ldw, a.sel_b, glo,
                &stage[0][0]);
    const int l15 = cc & 15, l4 = cc >> 4;
#pragma unroll
    for (int rd = 0; rd < kPR / 128; ++rd) {
      if (g / (128 / kRT) == rd) {
        const int rb = (g % (128 / kRT)) * kRT;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e)
              stage[16 * i + l4 + 4 * e][rb + 16 * j + l15] = acc[i][j][e];
      }
      __syncthreads();
      if (g / (128 / kRT) == rd) {
#pragma unroll
        for (int i = 0; i < kRT; ++i) x[i] = stage[cc][(g % (128 / kRT)) * kRT + i];
      }
      __syncthreads();
    }
  } else {
  // ---- load, 128 rows per round, coalesced along the rows of a column.
#pragma unroll
  for (int rd = 0; rd < kPR / 128; ++rd) {
#pragma unroll
    for (int q = 0; q < 128 * kMaxB / kPT; ++q) {
      const int idx = t + kPT * q, rl = idx & 127, c = idx >> 7;
      const long gr = row0 + rd * 128 + rl;
      stage[c][rl] = (c < b && gr < a.mrows) ? a.P[gr + (long)c * a.ldp] : 0.0;
    }
    __syncthreads();
    if (g / (128 / kRT) == rd) {
#pragma unroll
      for (int i = 0; i < kRT; ++i) x[i] = stage[cc][(g % (128 / kRT)) * kRT + i];
    }
    __syncthreads();
  }
  }

  // Exchange e's slot of workgroup q.  e = 0: initial norms; step k: 1 + 2k
  // (G and row k), 2 + 2k (recomputed norms, when needed).
  auto xslot = [&](int e, int q) -> double* {
    return a.xcur + ((long)e * nw + q) * kXS;
  };
  const int per = (nw + kPG - 1) / kPG;
  const int q0 = min(nw, g * per), q1 = min(nw, q0 + per);
  // Precondition: s_red holds this workgroup's per-group partials and, if
  // this workgroup is `rowner`, s_rowk its row; all threads have synced.
  // Wavefront 0 publishes; row group g sums the partials of workgroups
  // [q0, q1) into s_acc[g]; group 0 also fetches the row of `rowner` (if
  // >= 0) into s_rowin.  Returns false, in every thread, on failure.
  auto exchange = [&](int e, int rowner) -> bool {
    if (nw == 1) {
      // One workgroup: the same sums through LDS only (bitwise what the
      // published/polled path gives: group 0 adds 0 + the sum, the other
      // groups have no workgroups to add).  s_red was completed before the
      // barrier preceding this call; s_acc is not read again before the
      // barrier below.
      if (col_ok) {
        double a1 = 0.0;
#pragma unroll
        for (int q = 0; q < kPG; ++q) a1 += s_red[q][cc];
        s_acc[g][cc] = (g == 0) ? 0.0 + a1 : 0.0;
        if (g == 0 && rowner == 0) s_rowin[cc] = s_rowk[cc];
      }
      __syncthreads();
      return true;
    }
    if (t < b) {
      double a1 = 0.0;
#pragma unroll
      for (int q = 0; q < kPG; ++q) a1 += s_red[q][t];
      double* my = xslot(e, w);
      hq_publish(my + t, a1);
      if (rowner == w) hq_publish(my + kMaxB + t, s_rowk[t]);
    }
    bool ok = true;
    if (col_ok) {
      double acc = 0.0;
      for (int base = q0; ok && base < q1; base += kChunk) {
        const bool row = g == 0 && base == q0 && rowner >= 0;
        double v[kChunk + 1];
        ok = hq_await(v, [&](int i) -> const double* {
          if (i < kChunk) {
            const int q = base + i;
            return q < q1 ? xslot(e, q) + cc : nullptr;
          }
          return row ? xslot(e, rowner) + kMaxB + cc : nullptr;
        }, a.err);
#pragma unroll
        for (int i = 0; i < kChunk; ++i) acc += v[i];
        if (row) s_rowin[cc] = v[kChunk];
      }
      s_acc[g][cc] = acc;
    }
    if (!ok) s_abort = 1;
    __syncthreads();
    return s_abort == 0;
  };
  // Full sum for column c (every thread, the same order).
  auto full = [&](int c) -> double {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < kPG; ++q) s += s_acc[q][c];
    return s;
  };

  // Per-column state, one copy per wavefront (lane c: column c): partial
  // norm dn, its reference value en (LAWN 176's vn2), done flag, and the
  // column's current position pos (the pivoting swaps positions, columns
  // stay in their lanes).
  {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < kRT; ++i) s += x[i] * x[i];
    s_red[g][cc] = s;
  }
  __syncthreads();
  if (!exchange(0, -1)) return;
  double dn = col_ok ? sqrt(full(cc)) : 0.0, en = dn;
  bool done = !col_ok;
  int pos = cc;

  // Pivot of step k (idamax over positions k..b-1: largest partial norm,
  // first position on ties): the argmax runs over the columns at positions
  // >= k keyed by (norm, position), so no gather by position is needed.
  // Swaps positions k and p and returns the column now at position k.
  auto choose = [&](int k) -> int {
    int p = k;
    if (a.pivoting) {
      const bool cand = col_ok && pos >= k;
      double key = cand ? hq_pivot_key(dn) : -1.0;
      int q = cand ? pos : 1 << 30;
      hq_wave_argmax64(key, q);
      p = q;
    }
    const int col = __builtin_ctzll(__ballot(col_ok && pos == p));
    if (col_ok) {
      if (pos == p)
        pos = k;
      else if (pos == k)
        pos = p;
    }
    return col;
  };

  int pc = choose(0);
  for (int k = 0; k < b; ++k) {
    const int r = rel(k);
    const double* pv = s_pv[g];
    // 1. The pivot column's lane stages its rows for its wavefront (read
    //    back as LDS broadcasts: the LDS traffic stays inside the wavefront,
    //    ordered by wavefront-scope fences); partial G_c over my rows below
    //    k (four independent partial sums); the group holding row k stages
    //    it.
    {
      double s = 0.0;
      if (r < kRT) {
        if (cc == pc) {
#pragma unroll
          for (int i = 0; i < kRT; ++i)
            s_pv[g][i] = (i == 0 && r == 0) ? 0.0 : x[i];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double s4[4] = {0.0, 0.0, 0.0, 0.0};
        if (r < 0) {
#pragma unroll
          for (int i = 0; i < kRT; ++i) s4[i & 3] += pv[i] * x[i];
        } else {  // row k is x[0]
#pragma unroll
          for (int i = 1; i < kRT; ++i) s4[i & 3] += pv[i] * x[i];
          s_rowk[cc] = x[0];
        }
        s = (s4[0] + s4[1]) + (s4[2] + s4[3]);
      }
      s_red[g][cc] = s;
    }
    __syncthreads();
    if (!exchange(1 + 2 * k, (int)(k / kPR))) return;

    // 2. Reflector, dots and gram entries (every thread, identical bits):
    //    the pivot's row-k entry and G_p are read from the pivot's lane.
    const double gsum = full(cc), rin = s_rowin[cc];
    const hq_reflector h =
        hq_larfg_direct<double>(hq_readlane(rin, pc), hq_readlane(gsum, pc));
    if (w == 0 && t == 0) a.tau[k] = h.tau;
    const double xk = col_ok ? rin : 0.0;
    const double dot = xk + h.scale * (col_ok ? gsum : 0.0);
    // Gram entry (position of my column, k) for the factored columns.
    if (w == 0 && g == 0 && col_ok && pos < k) a.gram[pos + (long)k * b] = dot;
    // 3. dlarf on my rows >= k: pending columns x += v * (-tau * dot), the
    //    pivot column becomes (beta; v).  Branch-free per row: x becomes
    //    fma(v, fq, x * keep) with (keep, fq) = (1, f) for a pending
    //    column, (0, 1) for the pivot column and (1, 0) for the others.
    const double f = -h.tau * dot;
    const bool upd = col_ok && !done && cc != pc && h.tau != 0.0;
    const double keep = (cc == pc) ? 0.0 : 1.0;
    const double fq = upd ? f : ((cc == pc) ? 1.0 : 0.0);
    if (r < 0) {
#pragma unroll
      for (int i = 0; i < kRT; ++i)
        x[i] = __builtin_fma(pv[i] * h.scale, fq, x[i] * keep);
    } else if (r < kRT) {  // row k is x[0]
#pragma unroll
      for (int i = 1; i < kRT; ++i)
        x[i] = __builtin_fma(pv[i] * h.scale, fq, x[i] * keep);
      if (upd)
        x[0] += f;
      else if (cc == pc)
        x[0] = h.beta;
      // Row k is final: stash it (stage[column][row]) and shift up.
      if (col_ok) stage[cc][k - row0] = x[0];
#pragma unroll
      for (int i = 0; i < kRT - 1; ++i) x[i] = x[i + 1];
      x[kRT - 1] = 0.0;
      ++sh;
    }
    if (cc == pc) done = true;
    if (k + 1 == b) break;

    // 4. LAWN-176 downdate with row k of R (= the owner's updated x[r]).
    bool need = false;
    // The reference's t = |r| / dn, (1 + t)(1 - t), test t (dn / en)^2 <=
    // tol3z and dn sqrt((1 + t)(1 - t)), without divisions: (dn - |r|)(dn +
    // |r|) is dn^2 (1 + t)(1 - t), compared with tol3z en^2.
    if (a.pivoting && !done && dn != 0.0) {
      const double rk = (h.tau != 0.0) ? xk + f : xk;
      const double ar = fabs(rk);
      double u2 = (dn - ar) * (dn + ar);
      u2 = (0.0 > u2) ? 0.0 : u2;
      if (u2 <= HQ_TOL3Z * (en * en))
        need = true;
      else
        dn = sqrt(u2);
    }
    if (__ballot(need)) {  // identical in every wavefront of every workgroup
      // Every row still held is below k (the row-k wavefront has shifted).
      double s = 0.0;
      if (r < kRT) {
#pragma unroll
        for (int i = 0; i < kRT; ++i) s += x[i] * x[i];
      }
      s_red[g][cc] = s;
      __syncthreads();
      if (!exchange(2 + 2 * k, -1)) return;
      if (need) {
        dn = (a.mrows - k - 1 > 0) ? sqrt(full(cc)) : 0.0;
        en = dn;
      }
    }
    // 5. Next pivot.
    pc = choose(k + 1);
  }
  // ---- write back in pivoted order (my column goes to position pos).
  // A shifted wavefront first adds its unfinished rows to the stash (rows
  // glo+sh.., from x[0..]), then reloads all its rows from it.
  const int lg = (int)(glo - row0);
  if (sh > 0 && col_ok) {
#pragma unroll
    for (int i = 0; i < kRT; ++i)
      if (i < kRT - sh) stage[cc][lg + sh + i] = x[i];
  }
  __syncthreads();
  if (sh > 0 && col_ok) {
#pragma unroll
    for (int i = 0; i < kRT; ++i) x[i] = stage[cc][lg + i];
  }
  __syncthreads();
  const int myq = col_ok ? pos : 0;
#pragma unroll
  for (int rd = 0; rd < kPR / 128; ++rd) {
    if (col_ok && g / (128 / kRT) == rd) {
#pragma unroll
      for (int i = 0; i < kRT; ++i) stage[myq][(g % (128 / kRT)) * kRT + i] = x[i];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 128 * kMaxB / kPT; ++q) {
      const int idx = t + kPT * q, rl = idx & 127, c = idx >> 7;
      const long gr = row0 + rd * 128 + rl;
      if (c < b && gr < a.mrows) a.P[gr + (long)c * a.ldp] = stage[c][rl];
    }
    __syncthreads();
  }
  if (w == 0 && g == 0 && col_ok) a.perm[pos] = cc;
  hq_xreset(a.xprev, a.xprev_len, w, nw, kPT);
  // T from this workgroup's gram and tau (written above; the barrier orders
  // them before they are read back), in `stage`, free again.
  if (a.T && w == 0)
    larft_lds<kPT>(b, a.gram, b, a.tau, a.T, a.ldt,
                   reinterpret_cast<double (*)[kMaxB + 1]>(&stage[0][0]));
}

// dlarft('Forward','Columnwise'): T of the b reflectors, formed by one
// workgroup (larft_lds); see larft_lds for the method.
__global__ __launch_bounds__(HQ_NT) void larft_kernel(int b, const double* gram,
                                                      long ldg,
                                                      const double* tau,
                                                      double* T, long ldt,
                                                      const int* run_if) {
  if (run_if && *run_if == 0) return;
  __shared__ double sP[kMaxB][kMaxB + 1];
  larft_lds<HQ_NT>(b, gram, ldg, tau, T, ldt, sP);
}

// Applies the panel's column permutation to rows [0, nrows) of the b panel
// columns (the rows above the panel, A01 in the reference) and to jpvt.
// With P: blocks past the permutation's also copy the factored panel
// (prows x b, ld ldp, already in pivoted order) into A's rows nrows.. of the
// same columns -- the look-ahead's write-back, in the same launch.
__global__ __launch_bounds__(HQ_NT) void panel_permute_kernel(
    double* A, long lda, int nrows, int b, const int* perm, int* jpvt,
    const double* P, long ldp, int prows, int nperm) {
  __shared__ double buf[kMaxB][65];
  __shared__ int sj[kMaxB], sp[kMaxB];
  const int t = threadIdx.x;
  if ((int)blockIdx.x >= nperm) {
    const long i0 = (long)(blockIdx.x - nperm) * 64;
    constexpr int kPer = 64 * kMaxB / HQ_NT;
    double v[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int idx = t + u * HQ_NT, i = idx & 63, c = idx >> 6;
      v[u] = (c < b && i0 + i < prows) ? P[(i0 + i) + (long)c * ldp] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int idx = t + u * HQ_NT, i = idx & 63, c = idx >> 6;
      if (c < b && i0 + i < prows) A[(nrows + i0 + i) + (long)c * lda] = v[u];
    }
    return;
  }
  const long r0 = (long)blockIdx.x * 64;
  const bool jp = blockIdx.x == 0 && jpvt;
  if (t < b) {
    const int pt = perm[t];
    sp[t] = pt;
    if (jp) sj[t] = jpvt[pt];
  }
  // Every load of the block's 64 rows in flight together (compile-time
  // trip count), then the stores in the new column order.
  constexpr int kPer = 64 * kMaxB / HQ_NT;
  double v[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int idx = t + u * HQ_NT, i = idx & 63, c = idx >> 6;
    v[u] = (c < b && r0 + i < nrows) ? A[(r0 + i) + (long)c * lda] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int idx = t + u * HQ_NT;
    buf[idx >> 6][idx & 63] = v[u];
  }
  __syncthreads();
  if (jp && t < b) jpvt[t] = sj[t];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int idx = t + u * HQ_NT, i = idx & 63, q = idx >> 6;
    if (q < b && r0 + i < nrows) A[(r0 + i) + (long)q * lda] = buf[sp[q]][i];
  }
}

}  // namespace

hipError_t launch_panel_qrcp(const PanelArgs& a, hipStream_t stream) {
  hipLaunchKernelGGL(panel_qrcp_kernel, dim3(a.nw), dim3(kPT), 0, stream, a);
  return hipGetLastError();
}

int panel_rows_per_wg() { return kPR; }

size_t panel_xchg_doubles(int nw, int b) {
  return (size_t)(2 * b + 1) * nw * kXS;
}

hipError_t launch_larft(int b, const double* gram, long ldg, const double* tau,
                        double* T, long ldt, hipStream_t stream,
                        const int* run_if) {
  if (b < 1 || b > kMaxB) return hipErrorInvalidValue;
  hipLaunchKernelGGL(larft_kernel, dim3(1), dim3(HQ_NT), 0, stream, b, gram,
                     ldg, tau, T, ldt, run_if);
  return hipGetLastError();
}

hipError_t launch_panel_permute(double* A, long lda, int nrows, int b,
                                const int* perm, int* jpvt,
                                hipStream_t stream, const double* P, long ldp,
                                int prows) {
  const int nperm = nrows > 0 ? hq_cdiv(nrows, 64) : 1;
  const int ncopy = (P && prows > 0) ? hq_cdiv(prows, 64) : 0;
  hipLaunchKernelGGL(panel_permute_kernel, dim3(nperm + ncopy), dim3(HQ_NT), 0,
                     stream, A, lda, nrows, b, perm, jpvt, P, ldp, prows,
                     nperm);
  return hipGetLastError();
}
