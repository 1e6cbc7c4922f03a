// k_panel.hip -- panel factorisation: Householder QR with column pivoting
// inside an (m-j) x b panel, b <= 64, plus the compact-WY T factor.
//
// Restates NoFLA_QRPmod_WY_unb_var4(panel_pivoting, -1, AB1, ..., build_T=1)
// (reference NoFLA_HQRRP_WY_blk_var4.c:520-524, 774-886): per step an
// idamax over the partial norms, dlarfg on the pivot column, dlarf on the
// remaining columns, the LAWN-176 norm downdate (912-986), and at the end
// dlarft('Forward','Columnwise') (875-878).
//
// Design (MI355X): one persistent launch per panel.  Workgroup w owns a slab
// of R consecutive rows of all b columns, resident in LDS for the whole
// factorisation (R <= 240 rows x 64 columns = 120 KB), so the panel is read
// from and written to HBM exactly once.  Each step needs two grid-wide
// reductions:
//   A: the dot products of the reflector with every column (the dlarf
//      vector w for the pending columns and, for the columns already
//      factored, the inner products dlarft needs -- same pass, same barrier);
//   B: after the rank-1 update, the squared norms of every pending column
//      below rows k and k+1 (Drmac's recompute and the next dlarfg's ||x||)
//      together with rows k and k+1 themselves.
// Reductions are done in a fixed order by every workgroup, so all of them
// hold bit-identical pivot/reflector state without further communication.
// Columns stay where they are in LDS; a position->column map records the
// pivoting and the slab is written back in pivoted order.
#include "common.h"
#include "kernels.h"

namespace {

constexpr int kMaxB = 64;

__global__ __launch_bounds__(HQ_NT) void panel_qrcp_kernel(PanelArgs a) {
  extern __shared__ double slab[];  // kMaxB columns x ldR doubles
  __shared__ double s_d[kMaxB], s_e[kMaxB], s_S2[kMaxB], s_rowk[kMaxB];
  __shared__ double s_full[kMaxB];
  __shared__ double s_red[4][kMaxB], s_red2[4][kMaxB];
  __shared__ int s_phys[kMaxB], s_done[kMaxB];
  __shared__ double s_tau, s_beta, s_scale;
  __shared__ int s_pc;

  const int t = threadIdx.x, w = blockIdx.x;
  const int b = a.b, R = a.R, ldR = a.ldR, nw = a.nw;
  const long row0 = (long)w * R;
  const int nrows = (int)min((long)R, (long)a.mrows - row0);
  const int cc = t & 63, g = t >> 6;  // column / row-group of this thread

  for (int idx = t; idx < R * b; idx += HQ_NT) {
    int i = idx % R, c = idx / R;
    slab[c * ldR + i] = (i < nrows) ? a.P[(row0 + i) + (long)c * a.ldp] : 0.0;
  }
  if (t < kMaxB) {
    s_phys[t] = t;
    s_done[t] = (t < b) ? 0 : 1;
  }
  __syncthreads();

  unsigned phase = 0;
  double* partA = a.part;                 // nw x 64
  double* partB = a.part + (long)nw * 64;  // nw x 128

  // Sums of squares of pending columns over rows >= kk+1 (S1) and >= kk+2
  // (S2), plus rows kk and kk+1 published by their owners.
  auto norm_pass = [&](int kk) {
    if (cc < b && !s_done[cc]) {
      double s1 = 0.0, s2 = 0.0;
      const double* col = slab + cc * ldR;
      for (int i = g; i < nrows; i += 4) {
        long gi = row0 + i;
        double v = col[i];
        if (gi >= kk + 1) s1 += v * v;
        if (gi >= kk + 2) s2 += v * v;
      }
      s_red[g][cc] = s1;
      s_red2[g][cc] = s2;
    } else if (cc < kMaxB) {
      s_red[g][cc] = 0.0;
      s_red2[g][cc] = 0.0;
    }
    if (t < b && !s_done[t]) {
      long lk = kk - row0, lk1 = kk + 1 - row0;
      if (lk >= 0 && lk < nrows) a.rowbuf[t] = slab[t * ldR + lk];
      if (lk1 >= 0 && lk1 < nrows) a.rowbuf[kMaxB + t] = slab[t * ldR + lk1];
    }
    __syncthreads();
    if (t < b) {
      partB[(long)w * 128 + t] =
          ((s_red[0][t] + s_red[1][t]) + s_red[2][t]) + s_red[3][t];
      partB[(long)w * 128 + 64 + t] =
          ((s_red2[0][t] + s_red2[1][t]) + s_red2[2][t]) + s_red2[3][t];
    }
  };
  // Reduce partB in a fixed order; returns S1 in s_full, S2 in s_red2[0].
  auto reduce_B = [&]() {
    if (cc < b) {
      double s1 = 0.0, s2 = 0.0;
      for (int q = g; q < nw; q += 4) {
        s1 += partB[(long)q * 128 + cc];
        s2 += partB[(long)q * 128 + 64 + cc];
      }
      s_red[g][cc] = s1;
      s_red2[g][cc] = s2;
    }
    __syncthreads();
    if (t < b) {
      s_full[t] = ((s_red[0][t] + s_red[1][t]) + s_red[2][t]) + s_red[3][t];
      s_S2[t] = ((s_red2[0][t] + s_red2[1][t]) + s_red2[2][t]) + s_red2[3][t];
    }
    __syncthreads();
  };

  // Initial column norms (k = -1): S1 = full squared norms, S2 = rows >= 1.
  norm_pass(-1);
  if (!hq_grid_sync(a.bar, nw, phase, a.err)) return;
  reduce_B();
  if (t < b) {
    s_d[t] = sqrt(s_full[t]);
    s_e[t] = s_d[t];
    s_rowk[t] = a.rowbuf[kMaxB + t];
  }
  __syncthreads();

  for (int k = 0; k < b; ++k) {
    // Phase 1: pivot choice and reflector (serial, identical in every WG).
    if (t == 0) {
      int p = k;
      if (a.pivoting) {
        double best = s_d[s_phys[k]];
        for (int q = k + 1; q < b; ++q) {
          double v = s_d[s_phys[q]];
          if (v > best) {
            best = v;
            p = q;
          }
        }
      }
      int tmp = s_phys[k];
      s_phys[k] = s_phys[p];
      s_phys[p] = tmp;
      const int pc = s_phys[k];
      s_pc = pc;
      s_done[pc] = 1;
      hq_reflector h = hq_dlarfg(s_rowk[pc], (a.mrows - k > 1) ? s_S2[pc] : 0.0);
      s_tau = h.tau;
      s_beta = h.beta;
      s_scale = h.scale;
      if (w == 0) a.tau[k] = h.tau;
    }
    __syncthreads();
    const int pc = s_pc;
    const double tau = s_tau;
    // Phase 2: the pivot column becomes the reflector (unit entry implied).
    {
      double* col = slab + pc * ldR;
      const double sc = s_scale, be = s_beta;
      for (int i = t; i < nrows; i += HQ_NT) {
        long gi = row0 + i;
        if (gi > k)
          col[i] *= sc;
        else if (gi == k)
          col[i] = be;
      }
    }
    __syncthreads();
    // Phase 3: partial dots v'*col for every other column.
    if (cc < b) {
      double s = 0.0;
      if (cc != pc) {
        const double* v = slab + pc * ldR;
        const double* col = slab + cc * ldR;
        for (int i = g; i < nrows; i += 4) {
          long gi = row0 + i;
          if (gi >= k) s += ((gi == k) ? 1.0 : v[i]) * col[i];
        }
      }
      s_red[g][cc] = s;
    }
    __syncthreads();
    if (t < b)
      partA[(long)w * 64 + t] =
          ((s_red[0][t] + s_red[1][t]) + s_red[2][t]) + s_red[3][t];
    if (!hq_grid_sync(a.bar, nw, phase, a.err)) return;
    // Phase 4: full dots in a fixed order.
    if (cc < b) {
      double s = 0.0;
      for (int q = g; q < nw; q += 4) s += partA[(long)q * 64 + cc];
      s_red[g][cc] = s;
    }
    __syncthreads();
    if (t < b)
      s_full[t] = ((s_red[0][t] + s_red[1][t]) + s_red[2][t]) + s_red[3][t];
    __syncthreads();
    if (w == 0 && t < k) a.gram[t + (long)k * b] = s_full[s_phys[t]];
    // Phase 5: rank-1 update of the pending columns over rows >= k.
    if (tau != 0.0 && cc < b && !s_done[cc]) {
      const double f = -tau * s_full[cc];
      const double* v = slab + pc * ldR;
      double* col = slab + cc * ldR;
      for (int i = g; i < nrows; i += 4) {
        long gi = row0 + i;
        if (gi >= k) col[i] += ((gi == k) ? 1.0 : v[i]) * f;
      }
    }
    __syncthreads();
    if (k + 1 < b) {
      norm_pass(k);
      if (!hq_grid_sync(a.bar, nw, phase, a.err)) return;
      reduce_B();
      // Phase 6: norm downdate and the next step's alpha / ||x||^2.
      if (t < b && !s_done[t]) {
        double wt = a.rowbuf[t];
        if (a.pivoting)
          hq_downdate_norm(s_d[t], s_e[t], wt, s_full[t], a.mrows - k - 1);
        if (k + 1 < a.mrows) s_rowk[t] = a.rowbuf[kMaxB + t];
      }
      __syncthreads();
    }
  }
  // Write the slab back in pivoted order.
  for (int idx = t; idx < R * b; idx += HQ_NT) {
    int i = idx % R, q = idx / R;
    if (i < nrows) a.P[(row0 + i) + (long)q * a.ldp] = slab[s_phys[q] * ldR + i];
  }
  if (w == 0 && t < b) a.perm[t] = s_phys[t];
  hq_grid_exit(a.bar, nw);
}

// dlarft('Forward','Columnwise') from the reflector inner products:
// T(0:k,k) = -tau_k * T(0:k,0:k) * gram(0:k,k), T(k,k) = tau_k; a zero tau
// gives a zero column.  One workgroup; thread l owns row l of T.
__global__ __launch_bounds__(HQ_NT) void larft_kernel(int b, const double* gram,
                                                      const double* tau,
                                                      double* T, long ldt) {
  __shared__ double sT[kMaxB][kMaxB + 1];
  __shared__ double sg[kMaxB];
  const int t = threadIdx.x;
  for (int idx = t; idx < kMaxB * (kMaxB + 1); idx += HQ_NT)
    (&sT[0][0])[idx] = 0.0;
  __syncthreads();
  for (int k = 0; k < b; ++k) {
    const double tk = tau[k];
    if (t < k) sg[t] = -tk * gram[t + (long)k * b];
    __syncthreads();
    if (t < k) {
      double s = 0.0;
      if (tk != 0.0)
        for (int q = t; q < k; ++q) s += sT[t][q] * sg[q];
      sT[t][k] = s;
    }
    if (t == 0) sT[k][k] = tk;
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
  size_t lds = sizeof(double) * (size_t)kMaxB * a.ldR;
  hipLaunchKernelGGL(panel_qrcp_kernel, dim3(a.nw), dim3(HQ_NT), lds, stream,
                     a);
  return hipGetLastError();
}

hipError_t panel_set_lds_limit(size_t bytes) {
  return hipFuncSetAttribute((const void*)panel_qrcp_kernel,
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes);
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
