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
// written to HBM exactly once and every per-step loop runs on registers; LDS
// only stages the coalesced load/store and broadcasts the reflector.  At
// 16384 rows that is 64 workgroups; below 256 rows a single workgroup needs
// no cross-workgroup exchange at all.  Each step needs two grid-wide
// reductions, each a barrier-free exchange (common.h) through slots of its
// own:
//   A: the dot products of the reflector with every column (the dlarf
//      vector w for the pending columns and, for the columns already
//      factored, the inner products dlarft needs -- same pass, same barrier);
//   B: after the rank-1 update, the squared norms of every pending column
//      below rows k and k+1 (Drmac's recompute and the next dlarfg's ||x||)
//      together with rows k and k+1 themselves.
// Reductions are done in a fixed order by every workgroup, so all of them
// hold bit-identical pivot/reflector state without further communication;
// rows k and k+1 come from the slots of the workgroups that own them, in the
// same memory round trip as the partial sums.
// Columns stay with their threads; a position->column map records the
// pivoting and the slab is written back in pivoted order.
#include "common.h"
#include "kernels.h"

namespace {

constexpr int kMaxB = 64;

constexpr int kPT = 512;          // threads of the panel kernel
constexpr int kPG = kPT / 64;     // row groups (one per wavefront)
constexpr int kRT = 32;           // rows per thread (per row group)
constexpr int kPR = kRT * kPG;    // panel rows held by one workgroup
constexpr int kXS = 4 * kMaxB;    // exchange slot of one workgroup: A uses
                                  // [0,64) (dots), B [0,128) (S1, S2 of
                                  // column c at 2c, 2c+1), [128,192) row kk,
                                  // [192,256) row kk+1
constexpr int kChunk = 8;         // workgroups' partials awaited together

// Thread (c, g) = (t % 64, t / 64) holds column c of the 64 consecutive rows
// row0 + 64g + i, i < 64, in registers x[i]; rows beyond mrows are zero.
// Relative to the current step k, a wavefront's rows are either all above
// k (finished: skipped), all at or below k+2 (branch-free fast path), or they
// straddle k, k+1 (masked path, at most two wavefronts per step).
__global__ __launch_bounds__(kPT) void panel_qrcp_kernel(PanelArgs a) {
  __shared__ double stage[kMaxB][129];  // staging of 128 rows x 64 columns
  __shared__ double s_v[kPR];           // reflector on this WG's rows
  // s_red*: this workgroup's per-group partials (combined by wavefront 0
  // and published); s_acc*: per-group sums of the published partials.  Two
  // sets, since wavefront 0 may still be combining the former when the other
  // wavefronts already hold the exchange's results.
  __shared__ double s_red[kPG][kMaxB], s_red2[kPG][kMaxB];
  __shared__ double s_acc[kPG][kMaxB], s_acc2[kPG][kMaxB];
  __shared__ double s_d[kMaxB], s_e[kMaxB], s_S2[kMaxB], s_rowk[kMaxB];
  __shared__ double s_full[kMaxB], s_rk[kMaxB], s_rk1[kMaxB];
  __shared__ double s_rkin[kMaxB], s_rk1in[kMaxB];  // rows kk, kk+1 (reduced)
  __shared__ int s_phys[kMaxB], s_done[kMaxB], s_pos[kMaxB];
  __shared__ double s_tau, s_beta, s_scale;
  __shared__ int s_pc, s_abort;

  const int t = threadIdx.x, w = blockIdx.x;
  const int b = a.b, nw = a.nw;
  const long row0 = (long)w * kPR;
  const int cc = t & 63, g = t >> 6;
  const bool col_ok = cc < b;
  const long glo = row0 + kRT * g;  // first global row of my group
  // Position of global row kk inside my group, clamped to [-1, 64]:
  // -1: kk is above my rows, 64: kk is below them.
  auto rel = [&](long kk) -> int {
    long v = kk - glo;
    return (int)(v < -1 ? -1 : (v > kRT ? kRT : v));
  };

  // ---- load, 128 rows per round, coalesced along the rows of a column.
  double x[kRT];
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
  if (t < kMaxB) {
    s_phys[t] = t;
    s_done[t] = (t < b) ? 0 : 1;
  }
  if (t == 0) s_abort = 0;
  __syncthreads();

  // Exchange e's slot of workgroup q; e = 0: initial norms, then per step
  // k: A at 1 + 2k, B at 2 + 2k.  Row group g reduces the partials of
  // workgroups [q0, q1).
  auto xslot = [&](int e, int q) -> double* {
    return a.xcur + ((long)e * nw + q) * kXS;
  };
  const int per = (nw + kPG - 1) / kPG;
  const int q0 = min(nw, g * per), q1 = min(nw, q0 + per);

  // Squared norms of my column over rows >= kk+1 (S1) and >= kk+2 (S2), and
  // rows kk, kk+1 (to s_rk / s_rk1); wavefront 0 combines and publishes.
  auto norm_pass = [&](long kk, int e) {
    double s1 = 0.0, s2 = 0.0;
    const int r = rel(kk);
    if (r < 0) {  // every row of mine is >= kk+1; S2 differs if kk+1 is mine
      if (rel(kk + 1) < 0) {
#pragma unroll
        for (int i = 0; i < kRT; ++i) s1 += x[i] * x[i];
        s2 = s1;
      } else {  // kk+1 is my first row
#pragma unroll
        for (int i = 1; i < kRT; ++i) s2 += x[i] * x[i];
        s1 = s2 + x[0] * x[0];
        s_rk1[cc] = x[0];
      }
    } else if (r < kRT) {  // straddle: masked path
      double rk = 0.0, rk1 = 0.0;
#pragma unroll
      for (int i = 0; i < kRT; ++i) {
        const double v2 = x[i] * x[i];
        s1 += (i > r) ? v2 : 0.0;
        s2 += (i > r + 1) ? v2 : 0.0;
        rk = (i == r) ? x[i] : rk;
        rk1 = (i == r + 1) ? x[i] : rk1;
      }
      s_rk[cc] = rk;
      if (r + 1 < kRT) s_rk1[cc] = rk1;
    }
    s_red[g][cc] = s1;
    s_red2[g][cc] = s2;
    __syncthreads();
    if (t < b) {
      double a1 = 0.0, a2 = 0.0;
#pragma unroll
      for (int q = 0; q < kPG; ++q) {
        a1 += s_red[q][t];
        a2 += s_red2[q][t];
      }
      double* my = xslot(e, w);
      hq_publish(my + 2 * t, a1);
      hq_publish(my + 2 * t + 1, a2);
      const long lk = kk - row0, lk1 = kk + 1 - row0;
      if (lk >= 0 && lk < kPR) hq_publish(my + 128 + t, s_rk[t]);
      if (lk1 >= 0 && lk1 < kPR) hq_publish(my + 192 + t, s_rk1[t]);
    }
  };
  // Full S1 / S2 of every column in a fixed order (s_full, s_S2) and rows
  // kk, kk+1 (s_rkin, s_rk1in; row group 0 fetches them with its first
  // chunk).  Returns false, in every thread, if the exchange failed.
  auto reduce_B = [&](long kk, int e) -> bool {
    bool ok = true;
    if (col_ok) {
      double a1 = 0.0, a2 = 0.0;
      for (int base = q0; ok && base < q1; base += kChunk) {
        const bool rows = g == 0 && base == q0;
        double v[2 * kChunk + 2];
        ok = hq_await(v, [&](int i) -> const double* {
          if (i < 2 * kChunk) {
            const int q = base + i / 2;
            return q < q1 ? xslot(e, q) + 2 * cc + (i & 1) : nullptr;
          }
          if (!rows) return nullptr;
          if (i == 2 * kChunk)
            return kk >= 0 ? xslot(e, (int)(kk / kPR)) + 128 + cc : nullptr;
          return kk + 1 < a.mrows ? xslot(e, (int)((kk + 1) / kPR)) + 192 + cc
                                  : nullptr;
        }, a.err);
#pragma unroll
        for (int i = 0; i < kChunk; ++i) {
          a1 += v[2 * i];
          a2 += v[2 * i + 1];
        }
        if (rows) {
          s_rkin[cc] = v[2 * kChunk];
          s_rk1in[cc] = v[2 * kChunk + 1];
        }
      }
      s_acc[g][cc] = a1;
      s_acc2[g][cc] = a2;
    }
    if (!ok) s_abort = 1;
    __syncthreads();
    if (s_abort) return false;
    if (t < b) {
      double a1 = 0.0, a2 = 0.0;
#pragma unroll
      for (int q = 0; q < kPG; ++q) {
        a1 += s_acc[q][t];
        a2 += s_acc2[q][t];
      }
      s_full[t] = a1;
      s_S2[t] = a2;
    }
    __syncthreads();
    return true;
  };

  // Initial column norms (kk = -1): S1 = full squared norms, S2 = rows >= 1.
  norm_pass(-1, 0);
  if (!reduce_B(-1, 0)) return;
  if (t < b) {
    s_d[t] = sqrt(s_full[t]);
    s_e[t] = s_d[t];
    s_rowk[t] = s_rk1in[t];
  }
  __syncthreads();

  for (int k = 0; k < b; ++k) {
    // Phase 1: pivot (wavefront 0: largest partial norm, first position on
    // ties, as idamax) and reflector; identical in every workgroup.
    if (t < 64) {
      int p = k;
      if (a.pivoting) {
        double v = (t >= k && t < b) ? s_d[s_phys[t]] : -1.0;
        int q = (t >= k && t < b) ? t : 1 << 30;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
          double ov = __shfl_xor(v, off, 64);
          int oq = __shfl_xor(q, off, 64);
          if (ov > v || (ov == v && oq < q)) {
            v = ov;
            q = oq;
          }
        }
        p = q;
      }
      if (t == 0) {
        int tmp = s_phys[k];
        s_phys[k] = s_phys[p];
        s_phys[p] = tmp;
        const int pc = s_phys[k];
        s_pc = pc;
        s_done[pc] = 1;
        hq_reflector h =
            hq_dlarfg(s_rowk[pc], (a.mrows - k > 1) ? s_S2[pc] : 0.0);
        s_tau = h.tau;
        s_beta = h.beta;
        s_scale = h.scale;
        if (w == 0) a.tau[k] = h.tau;
      }
    }
    __syncthreads();
    const int pc = s_pc;
    const double tau = s_tau;
    const int r = rel(k);  // k relative to my group (-1: above, kRT: below)
    const bool active = r < kRT;  // some of my rows are >= k
    double* vg = s_v + kRT * g;
    // Phase 2: the pivot column becomes the reflector; its values on my rows
    // go to LDS (1 at row k, 0 above).
    if (cc == pc && active) {
      const double sc = s_scale, be = s_beta;
      if (r < 0) {
#pragma unroll
        for (int i = 0; i < kRT; ++i) {
          x[i] *= sc;
          vg[i] = x[i];
        }
      } else {
#pragma unroll
        for (int i = 0; i < kRT; ++i) {
          const bool below = i > r, diag = i == r;
          x[i] = below ? x[i] * sc : (diag ? be : x[i]);
          vg[i] = below ? x[i] : (diag ? 1.0 : 0.0);
        }
      }
    }
    __syncthreads();
    // Phase 3: partial dots v'*col for every other column (rows >= k).
    {
      double s = 0.0;
      if (active && col_ok && cc != pc) {
#pragma unroll
        for (int i = 0; i < kRT; ++i) s += vg[i] * x[i];
      }
      s_red[g][cc] = s;
    }
    __syncthreads();
    const int eA = 1 + 2 * k;
    if (t < b) {
      double a1 = 0.0;
#pragma unroll
      for (int q = 0; q < kPG; ++q) a1 += s_red[q][t];
      hq_publish(xslot(eA, w) + t, a1);
    }
    // Phase 4: full dots in a fixed order.
    {
      bool ok = true;
      if (col_ok) {
        double a1 = 0.0;
        for (int base = q0; ok && base < q1; base += kChunk) {
          double v[kChunk];
          ok = hq_await(v, [&](int i) -> const double* {
            const int q = base + i;
            return q < q1 ? xslot(eA, q) + cc : nullptr;
          }, a.err);
#pragma unroll
          for (int i = 0; i < kChunk; ++i) a1 += v[i];
        }
        s_acc[g][cc] = a1;
      }
      if (!ok) s_abort = 1;
    }
    __syncthreads();
    if (s_abort) return;
    if (t < b) {
      double a1 = 0.0;
#pragma unroll
      for (int q = 0; q < kPG; ++q) a1 += s_acc[q][t];
      s_full[t] = a1;
    }
    __syncthreads();
    if (w == 0 && t < k) a.gram[t + (long)k * b] = s_full[s_phys[t]];
    // Phase 5: rank-1 update of the pending columns over rows >= k.
    if (active && tau != 0.0 && col_ok && !s_done[cc]) {
      const double f = -tau * s_full[cc];
#pragma unroll
      for (int i = 0; i < kRT; ++i) x[i] += vg[i] * f;
    }
    if (k + 1 < b) {
      __syncthreads();
      norm_pass(k, 2 + 2 * k);
      if (!reduce_B(k, 2 + 2 * k)) return;
      // Phase 6: norm downdate and the next step's alpha / ||x||^2.
      if (t < b && !s_done[t]) {
        const double wt = s_rkin[t];
        if (a.pivoting)
          hq_downdate_norm(s_d[t], s_e[t], wt, s_full[t], a.mrows - k - 1);
        if (k + 1 < a.mrows) s_rowk[t] = s_rk1in[t];
      }
      __syncthreads();
    }
  }
  // ---- write back in pivoted order (position q holds column s_phys[q]).
  if (t < b) s_pos[s_phys[t]] = t;
  __syncthreads();
  const int myq = col_ok ? s_pos[cc] : 0;
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
  if (w == 0 && t < b) a.perm[t] = s_phys[t];
  hq_xreset(a.xprev, a.xprev_len, w, nw, kPT);
}

// dlarft('Forward','Columnwise') from the reflector inner products
// G(l,q) = v_l'v_q (l < q): T is the compact-WY factor of H_0 H_1 ... with
// T(q,q) = tau_q and, for adjacent column blocks B1, B2,
//   T(B1,B2) = -T(B1,B1) G(B1,B2) T(B2,B2),
// built by doubling the block size (1, 2, 4, ..., 32): six levels of small
// triangular products instead of LAPACK's b-step recurrence (same matrix in
// exact arithmetic; a zero tau gives a zero column, as LAPACK's).
__global__ __launch_bounds__(HQ_NT) void larft_kernel(int b, const double* gram,
                                                      const double* tau,
                                                      double* T, long ldt) {
  __shared__ double sT[kMaxB][kMaxB + 1];
  __shared__ double sG[kMaxB][kMaxB + 1];
  __shared__ double sX[kMaxB][kMaxB + 1];
  const int t = threadIdx.x;
  for (int idx = t; idx < kMaxB * (kMaxB + 1); idx += HQ_NT) {
    (&sT[0][0])[idx] = 0.0;
    (&sX[0][0])[idx] = 0.0;
  }
  __syncthreads();
  for (int idx = t; idx < b * b; idx += HQ_NT) {
    int i = idx % b, j = idx / b;
    sG[i][j] = (i < j) ? gram[idx] : 0.0;
    if (i == j) sT[i][i] = tau[i];
  }
  __syncthreads();
  for (int s = 1; s < b; s *= 2) {
    // X(B1,B2) = G(B1,B2) T(B2,B2) for every pair (B1,B2) at this level.
    for (int e = t; e < b * s; e += HQ_NT) {
      const int col = e / s, r = e % s;           // col: global column in B2
      const int c0 = (col / (2 * s)) * 2 * s;     // start of B1
      const int c1 = c0 + s;                      // start of B2
      if (col < c1 || col >= min(b, c0 + 2 * s)) continue;
      double acc = 0.0;
      for (int u = c1; u <= col; ++u) acc += sG[c0 + r][u] * sT[u][col];
      sX[c0 + r][col] = acc;
    }
    __syncthreads();
    // T(B1,B2) = -T(B1,B1) X(B1,B2).
    for (int e = t; e < b * s; e += HQ_NT) {
      const int col = e / s, r = e % s;
      const int c0 = (col / (2 * s)) * 2 * s, c1 = c0 + s;
      if (col < c1 || col >= min(b, c0 + 2 * s)) continue;
      double acc = 0.0;
      for (int u = c0 + r; u < c1; ++u) acc += sT[c0 + r][u] * sX[u][col];
      sT[c0 + r][col] = -acc;
    }
    __syncthreads();
  }
  for (int idx = t; idx < b * b; idx += HQ_NT) {
    int i = idx % b, j = idx / b;
    T[i + (long)j * ldt] = (i <= j) ? sT[i][j] : 0.0;
  }
}

// Applies the panel's column permutation to rows [0, nrows) of the b panel
// columns (the rows above the panel, A01 in the reference) and to jpvt.
__global__ __launch_bounds__(HQ_NT) void panel_permute_kernel(
    double* A, long lda, int nrows, int b, const int* perm, int* jpvt) {
  __shared__ double buf[kMaxB][65];
  __shared__ int sj[kMaxB];
  const int t = threadIdx.x;
  const long r0 = (long)blockIdx.x * 64;
  if (blockIdx.x == 0 && jpvt) {
    if (t < b) sj[t] = jpvt[perm[t]];
    __syncthreads();
    if (t < b) jpvt[t] = sj[t];
  }
  if (r0 >= nrows) return;
  for (int idx = t; idx < 64 * b; idx += HQ_NT) {
    int i = idx & 63, c = idx >> 6;
    if (r0 + i < nrows) buf[c][i] = A[(r0 + i) + (long)c * lda];
  }
  __syncthreads();
  for (int idx = t; idx < 64 * b; idx += HQ_NT) {
    int i = idx & 63, q = idx >> 6;
    if (r0 + i < nrows) A[(r0 + i) + (long)q * lda] = buf[perm[q]][i];
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

hipError_t launch_larft(int b, const double* gram, const double* tau,
                        double* T, long ldt, hipStream_t stream) {
  hipLaunchKernelGGL(larft_kernel, dim3(1), dim3(HQ_NT), 0, stream, b, gram,
                     tau, T, ldt);
  return hipGetLastError();
}

hipError_t launch_panel_permute(double* A, long lda, int nrows, int b,
                                const int* perm, int* jpvt,
                                hipStream_t stream) {
  int blocks = nrows > 0 ? hq_cdiv(nrows, 64) : 1;
  hipLaunchKernelGGL(panel_permute_kernel, dim3(blocks), dim3(HQ_NT), 0,
                     stream, A, lda, nrows, b, perm, jpvt);
  return hipGetLastError();
}
