// k_batch.hip -- batched HQRRP of many small matrices in one persistent
// launch (no counterpart in the reference, which factors one matrix per
// call).  Each workgroup takes matrices from a work queue (sorted largest
// first on the host) and runs the whole of NoFLA_HQRRP_WY_blk_var4
// (reference NoFLA_HQRRP_WY_blk_var4.c:303-571) on one matrix at a time:
//
//   Y = Omega A                                 MFMA tiles    (WY.c:407-409)
//   per block of nb columns (b = min(nb, n-j, m-j)):
//     sketch QRCP on Y(:, j:n), b stages        columns in lanes
//                                                (WY.c:458-470, 774-886)
//     panel QRCP of A(j:m, j:j+b) and its T     rows in lanes
//                                                (WY.c:490-506, 774-886)
//     A(j:m, j+b:n) -= (V T') (V' A(j:m, j+b:n))   MFMA strips (WY.c:529-541)
//     G(:, j:m) := G(:, j:m) (I - V T V')       the same strips on G'
//     Y2 -= G(:, j:j+b) R12                     MFMA tiles    (WY.c:617-708)
//
// The last line is the reference's sketch downdate: its
// B = G1 - (G1 U11 + G2 U21) T U11' is exactly the first b columns of the
// updated G(I - V T V'), so the downdate is the G update followed by one
// product.
//
// Column swaps never move data: a logical -> physical column map in LDS
// follows them (the sketch, the panel, R12 and Y all address columns through
// it), and A's columns and jpvt are put in logical order once at the end.
// Intermediates live in registers, LDS and a private workspace per
// workgroup (Y, G, V, V T'), small enough to stay in L2.
//
// Shapes: m, n <= 512; nb <= 64; pp <= 10.  512 threads: one sketch column
// and one panel row per thread.
// Every cross-thread dependency is a workgroup barrier; workgroups share
// nothing but the queue counter, so no wait depends on another workgroup.
//
// v_mfma_f64_16x16x4f64 fragments: A lane l = A[l%16][l/16], B lane l =
// B[l/16][l%16]; accumulator element e of lane l = D[l/16 + 4e][l%16].
#include "common.h"
#include "kernels.h"

namespace {

constexpr int kPPMax = HQB_PP_MAX;
constexpr int kDim = HQB_DIM_MAX;
constexpr int kBT = 512;            // threads per workgroup (one row/column each)
constexpr int kWaves = kBT / 64;
constexpr int kLdv = kDim;          // leading dimension of V T' (workspace)
constexpr int kPermRows = 4;        // rows per pass of the final permutation
constexpr int kPermPitch = kDim + 1;

// Workspace of one workgroup (doubles): Y (r x n), G (r x m), V' (nb x R,
// row-major V), V T' (R x nb, ld kLdv).
constexpr long kYW = (64L + kPPMax) * kDim;
constexpr long kGW = (64L + kPPMax) * kDim;
constexpr long kVW = 64L * kDim;
constexpr long kTW = (long)kLdv * 64;
constexpr long kWork = kYW + kGW + kVW + kTW;

template <int NB>
struct Cfg {
  static constexpr int RMAX = NB + kPPMax;  // sketch rows held per column
  static constexpr int CPT = 1;                // sketch columns per thread
  static constexpr int RPT = 1;                // panel rows per thread
  static constexpr int IT = NB / 16;           // 16-row tiles of V'
  static constexpr int SB = 1;  // strips per round (more spill at 128 VGPRs)
  static constexpr int RT = (RMAX + 15) / 16;  // 16-row tiles of Y
  static constexpr int SH = NB == 64 ? 0 : (NB == 32 ? 1 : 2);
};

template <int NB>
struct Lds {
  static constexpr int RMAX = NB + kPPMax;
  static constexpr int IT = NB / 16;
  int map[kDim];       // logical column -> physical column of A (and Y)
  int slot_col[NB];    // panel slot -> physical column
  int slot_step[NB];   // panel slot -> elimination step
  int item;
  unsigned long long clk[HQB_PHASES + 1];  // optional phase cycle counts
  double tau[NB];
  double onehot[192];  // onehot[64] = 1, zeros elsewhere
  double step_gt[192]; // step_gt[i] = (i > 64)
  double pn_scale[NB]; // panel: dlarfg scale of each slot's reflector
  double sk_hdr[2][kWaves][4];       // sketch: wave's (key, pos, tau, alpha)
  double sk_x[2][kWaves][RMAX];      // sketch: wave's candidate reflector
  double pn_part[2][kWaves][NB];     // panel: per-wave partial sums per slot
  double pn_rowk[2][NB];             // panel: row k of every slot
  double pn_bc[kWaves][2][NB];       // panel: per-wave broadcast (f, row k)
  union {
    double T[NB][NB + 1];  // T upper triangle; S(l,k) = v_l'v_k at T[k][l]
    double W[kWaves * Cfg<NB>::SB * IT * 4 * 64];  // [wave][SB][IT][4][64]
    double perm[kPermRows * kPermPitch];
  } u;
};

__device__ __forceinline__ hq_d4 mfma4(double a, double b, hq_d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Operands outside a matrix are read from here (a load from a selected
// address, never a load under a branch, so loads stay in flight together).
__device__ const double kZeros[4] = {0.0, 0.0, 0.0, 0.0};

// Optional per-phase cycle counts (BatchArgs::prof): thread 0 of each
// workgroup times the phases of its matrices (shader clock, after the
// barrier that ends each phase) into LDS and adds them up at exit.
struct PhaseClock {
  unsigned long long* acc;  // LDS, HQB_PHASES + 1 (the last: phase start)
  bool on;
  __device__ void start(unsigned long long* lds, bool enable) {
    acc = lds;
    on = enable;
    if (on && threadIdx.x == 0) {
      for (int i = 0; i < HQB_PHASES; ++i) acc[i] = 0;
      acc[HQB_PHASES] = clock64();
    }
  }
  __device__ __forceinline__ void mark(int ph) {
    if (on && threadIdx.x == 0) {
      const unsigned long long t = clock64();
      acc[ph] += t - acc[HQB_PHASES];
      acc[HQB_PHASES] = t;
    }
  }
  __device__ void flush(unsigned long long* out) {
    if (on && threadIdx.x == 0)
      for (int i = 0; i < HQB_PHASES; ++i) atomicAdd(out + i, acc[i]);
  }
};

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double readlane_d(double v, int l) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)u, l);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ int readlane_i(int v, int l) {
  return __builtin_amdgcn_readlane(v, l);
}

// Reduce-scatter over the 64 lanes of a wavefront: lane l returns the wave
// sum of value (l >> SH) (NB = 64 >> SH values per lane, the first level
// taking them from val(c)).  Halving exchanges (each lane keeps half of what
// it holds and adds its partner's copy of that half), then a butterfly over
// the low lane bits; the order is fixed, so every wave summing the same
// values gets the same bits.
template <int H, int O, int N>
__device__ __forceinline__ void rs_levels(double (&v)[N], int lane) {
  if constexpr (H >= 1) {
    const bool up = (lane & O) != 0;
#pragma unroll
    for (int i = 0; i < H; ++i) {
      const double send = up ? v[i] : v[i + H];
      const double keep = up ? v[i + H] : v[i];
      v[i] = keep + __shfl_xor(send, O, 64);
    }
    rs_levels<H / 2, O / 2, N>(v, lane);
  }
}

template <int NB, class Val>
__device__ __forceinline__ double wave_reduce_scatter(Val val) {
  const int lane = threadIdx.x & 63;
  constexpr int H = NB / 2;
  double v[H];
  const bool up = (lane & 32) != 0;
#pragma unroll
  for (int i = 0; i < H; ++i) {
    const double lo = val(i), hi = val(i + H);
    const double send = up ? lo : hi;
    const double keep = up ? hi : lo;
    v[i] = keep + __shfl_xor(send, 32, 64);
    if ((i & 7) == 7) __builtin_amdgcn_sched_barrier(0);
  }
  rs_levels<H / 2, 16, H>(v, lane);
  double s = v[0];
  constexpr int SH = Cfg<NB>::SH;
#pragma unroll
  for (int o = (1 << SH) >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s;
}

// Sums val(c) over the workgroup: lane c (< NB) of every wave returns the
// total for slot c.  Slots are reduced in chunks of at most 32 (bounding the
// registers of the reduce-scatter); one barrier; buffers alternate with
// `buf`.
template <int NB, class Val>
__device__ __forceinline__ double pn_exchange(Lds<NB>& L, Val val, int buf) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int CK = NB < 32 ? NB : 32;
  constexpr int SH = Cfg<CK>::SH;
#pragma unroll
  for (int ch = 0; ch < NB / CK; ++ch) {
    const double s =
        wave_reduce_scatter<CK>([&](int c) { return val(ch * CK + c); });
    if ((lane & ((1 << SH) - 1)) == 0)
      L.pn_part[buf][wave][ch * CK + (lane >> SH)] = s;
  }
  __syncthreads();
  double tot = 0.0;
  if (lane < NB) {
#pragma unroll
    for (int w = 0; w < kWaves; ++w) tot += L.pn_part[buf][w][lane];
  }
  return tot;
}

// ---------------------------------------------------------------- phases

// Y (r x n, ld r) = G (r x m, ld r) * A (m x n).  Work item: one 16-column
// tile of Y with all its row tiles.
template <int NB>
__device__ __forceinline__ void sketch_gemm(const double* __restrict__ G, int r, int m, int n,
                            const double* __restrict__ A, long lda,
                            double* __restrict__ Y) {
  constexpr int RT = Cfg<NB>::RT;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l15 = lane & 15, l4 = lane >> 4;
  const int rtn = (r + 15) / 16;
  const int ctiles = (n + 15) / 16;
  for (int ct = wave; ct < ctiles; ct += kWaves) {
    const int col = ct * 16 + l15;
    const bool cok = col < n;
    const double* Ac = A + (long)(cok ? col : 0) * lda;
    hq_d4 acc[RT];
#pragma unroll
    for (int i = 0; i < RT; ++i) acc[i] = hq_d4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < m; k0 += 4) {
      const int k = k0 + l4;
      const bool kok = k < m;
      // k >= m reads zero for the A operand; rows of Y >= r are not stored
      const double bv = *(kok ? Ac + k : kZeros);
      const int kc = kok ? k : m - 1;
#pragma unroll
      for (int i = 0; i < RT; ++i) {
        if (i < rtn) {
          const int row = i * 16 + l15;
          const double av = G[min(row, r - 1) + kc * r];
          acc[i] = mfma4(av, bv, acc[i]);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < RT; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = i * 16 + l4 + 4 * e;
        if (i < rtn && row < r && cok) Y[row + (long)col * r] = acc[i][e];
      }
  }
}

// Pivot selection on the sketch Y(:, j:j+N) (NoFLA_QRPmod_WY_unb_var4 with
// pivoting, WY.c:774-886, on a copy of the sketch): b stages of column
// pivoting by the partial norms (LAWN 176 downdate, idamax's first-index
// rule), Householder reflections of the remaining columns.  Thread t holds
// logical column t (all r rows) in registers.  Each stage is
// one barrier: the best candidate of every wave stages its reflector u
// (u_k = 1, zero outside rows k..r-1) in LDS, every thread picks the best of
// the four and applies it.  Row k of a column is read as its dot product with
// a one-hot vector, so no register is ever indexed by k.  Only the pivots are
// kept: the logical -> physical map takes each swap.
template <int NB>
__device__ __forceinline__ void sketch_qrcp(Lds<NB>& L, const double* __restrict__ Y, int r, int j,
                            int N, int b) {
  constexpr int RMAX = Cfg<NB>::RMAX, CPT = Cfg<NB>::CPT;
  constexpr int CH = 8, NCH = (RMAX + CH - 1) / CH;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  double v[CPT][RMAX];
  double d[CPT], e[CPT];
  int pos[CPT];
  bool live[CPT];
#pragma unroll
  for (int c = 0; c < CPT; ++c) {
    pos[c] = t + c * kBT;
    live[c] = pos[c] < N;
    const double* y = Y + (long)L.map[j + (live[c] ? pos[c] : 0)] * r;
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < RMAX; ++i) {
      v[c][i] = (live[c] && i < r) ? y[i] : 0.0;
      s += v[c][i] * v[c][i];
    }
    d[c] = sqrt(s);
    e[c] = d[c];
  }
  for (int k = 0; k < b; ++k) {
    const int buf = k & 1;
    // best live column of this thread, then of the wave
    double key = -1.0;
    int kp = 0x7fffffff, kc = -1;
#pragma unroll
    for (int c = 0; c < CPT; ++c)
      if (live[c] && (d[c] > key || (d[c] == key && pos[c] < kp))) {
        key = d[c];
        kp = pos[c];
        kc = c;
      }
    double wk = key;
    int wp = kp;
    hq_wave_argmax(wk, wp, 64);
    double* xw = L.sk_x[buf][wave];
    double* hd = L.sk_hdr[buf][wave];
    const double* oh = L.onehot + 64 - k;   // oh[i] = (i == k)
    const double* gt = L.step_gt + 64 - k;  // gt[i] = (i > k)
    if (kc >= 0 && kp == wp) {
      // the wave's candidate forms its reflector (dlarfg) in LDS; rows past
      // r hold zeros, so u_i = (i > k) scale x_i + (i == k) needs no test
      double alpha = 0.0, xn = 0.0;
#pragma unroll
      for (int i = 0; i < RMAX; ++i) {
        const double xv = (CPT == 1 || kc == 0) ? v[0][i] : v[CPT - 1][i];
        alpha = fma(oh[i], xv, alpha);
        const double xg = gt[i] * xv;
        xn = fma(xg, xg, xn);
      }
      const hq_reflector h = hq_dlarfg(alpha, xn);
#pragma unroll
      for (int i = 0; i < RMAX; ++i) {
        const double xv = (CPT == 1 || kc == 0) ? v[0][i] : v[CPT - 1][i];
        xw[i] = fma(gt[i] * h.scale, xv, oh[i]);
      }
      hd[0] = wk;
      hd[1] = (double)wp;
      hd[2] = h.tau;
    } else if (lane == 0 && wk < 0.0) {
      hd[0] = -1.0;
      hd[1] = 2147483647.0;
    }
    __syncthreads();
    int ww = 0;
    double bk = L.sk_hdr[buf][0][0];
    int bp = (int)L.sk_hdr[buf][0][1];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) {
      const double kk = L.sk_hdr[buf][w][0];
      const int pw = (int)L.sk_hdr[buf][w][1];
      if (kk > bk || (kk == bk && pw < bp)) {
        bk = kk;
        bp = pw;
        ww = w;
      }
    }
    if (bp < k || bp >= N) bp = k;  // only reachable with non-finite norms
    const double tauk = L.sk_hdr[buf][ww][2];
    const double* us = L.sk_x[buf][ww];
    // w = u'v and the old row k of every column held
    double w[CPT], vk[CPT];
#pragma unroll
    for (int c = 0; c < CPT; ++c) w[c] = vk[c] = 0.0;
#pragma unroll
    for (int q = 0; q < NCH; ++q) {
      if (q * CH + CH > k && q * CH < r) {
#pragma unroll
        for (int ii = 0; ii < CH; ++ii) {
          const int i = q * CH + ii;
          if (i < RMAX) {
            const double ui = us[i], si = oh[i];
#pragma unroll
            for (int c = 0; c < CPT; ++c) {
              w[c] = fma(ui, v[c][i], w[c]);
              vk[c] = fma(si, v[c][i], vk[c]);
            }
          }
        }
      }
    }
    asm volatile("" ::: "memory");  // reload u rather than keep it live
    double f[CPT];
#pragma unroll
    for (int c = 0; c < CPT; ++c) f[c] = tauk * w[c];
#pragma unroll
    for (int q = 0; q < NCH; ++q) {
      if (q * CH + CH > k && q * CH < r) {
#pragma unroll
        for (int ii = 0; ii < CH; ++ii) {
          const int i = q * CH + ii;
          if (i < RMAX) {
            const double ui = us[i];
#pragma unroll
            for (int c = 0; c < CPT; ++c) v[c][i] = fma(-f[c], ui, v[c][i]);
          }
        }
      }
    }
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      if (live[c] && pos[c] != bp && d[c] != 0.0) {
        // LAWN 176 downdate (WY.c:912-986) with the new row-k entry; the
        // recompute is thread-local
        const double vkn = vk[c] - f[c];
        double tt = fabs(vkn) / d[c];
        const double uu = (1.0 + tt) * (1.0 - tt);
        tt = (0.0 > uu) ? 0.0 : uu;
        const double ratio = d[c] / e[c];
        if (tt * ratio * ratio <= HQ_TOL3Z) {
          double s = 0.0;
#pragma unroll
          for (int i = 0; i < RMAX; ++i) {
            const double pv = gt[i] * v[c][i];
            s = fma(pv, pv, s);
          }
          d[c] = (r - k - 1 > 0) ? sqrt(s) : 0.0;
          e[c] = d[c];
        } else {
          d[c] = d[c] * sqrt(tt);
        }
      }
      // swap semantics (WY.c:990-1036): the chosen column goes to k, the
      // column at k to the chosen one's place, norms travel with columns
      if (live[c]) {
        if (pos[c] == bp)
          live[c] = false;
        else if (pos[c] == k)
          pos[c] = bp;
      }
    }
    if (t == 0) {
      const int a0 = L.map[j + k];
      L.map[j + k] = L.map[j + bp];
      L.map[j + bp] = a0;
    }
  }
}

// Panel QRCP of A(j:m, j:j+b) (R = m - j rows) with optional column
// pivoting (NoFLA_QRPmod_WY_unb_var4, WY.c:774-886), writing the factored
// panel back to A, tau, V' (unit lower V, row-major, ld NB) and
// S(l, k) = v_l'v_k for the T factor.  Thread t holds row t of
// every slot (slot c = the panel column at logical position j + c on entry);
// lane c of every wave holds slot c's norms and position.  One exchange per
// step: with x the pivot column and k the step,
//   G_c = x(k+1:)' A(k+1:, c) and row k of every slot, summed over the
//   workgroup, give ||x(k+1:)||^2 (c = pivot), v'A(:, c) = A(k, c) +
//   scale G_c for the live columns and v_l'v_k for the eliminated ones.
// An eliminated slot keeps x unscaled below its diagonal (its dlarfg scale is
// applied on the way out).  A second exchange recomputes partial norms when
// the LAWN test asks for it.
template <int NB>
__device__ __forceinline__ void panel_qrcp(Lds<NB>& L, double* __restrict__ A, long lda, int R,
                           int j, int b, int pivoting,
                           double* __restrict__ tau, double* __restrict__ Vt) {
  constexpr int RPT = Cfg<NB>::RPT;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  double P[RPT][NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    const long pcol = (c < b) ? L.map[j + c] : 0;
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
      const int row = t + s * kBT;
      P[s][c] = (c < b && row < R) ? A[(j + row) + pcol * lda] : 0.0;
    }
  }
  if (wave == 0 && lane < b) L.slot_col[lane] = L.map[j + lane];
  bool sdone = lane >= b;
  int spos = lane;
  double dn = 0.0, en = 0.0, ssc = 1.0;
  int xi = 0;
  if (pivoting) {
    dn = sqrt(pn_exchange<NB>(
        L,
        [&](int c) {
          double s = 0.0;
#pragma unroll
          for (int q = 0; q < RPT; ++q) s = fma(P[q][c], P[q][c], s);
          return s;
        },
        xi & 1));
    en = dn;
    ++xi;
  }
  double* bc = &L.pn_bc[wave][0][0];
  for (int k = 0; k < b; ++k) {
    // 1. pivot slot p (uniform): largest partial norm, first position on ties
    int ppos;
    if (pivoting) {
      double key = sdone ? -1.0 : hq_pivot_key(dn);
      int kp = sdone ? 0x7fffffff : spos;
      hq_wave_argmax(key, kp, 64);
      ppos = kp;
    } else {
      ppos = k;
    }
    unsigned long long pm = __ballot(!sdone && spos == ppos);
    if (pm == 0ull) {  // only reachable with non-finite norms
      pm = __ballot(!sdone);
      ppos = readlane_i(spos, __builtin_ctzll(pm));
    }
    const int p = __builtin_ctzll(pm);
    // 2. pivot column on my rows (one-hot product), rows <= k masked
    const double* oh = L.onehot + 64 - p;  // oh[c] = (c == p)
    double x[RPT], xm[RPT];
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
      double a0 = 0.0, a1 = 0.0;
#pragma unroll
      for (int c = 0; c < NB; c += 2) {
        a0 = fma(oh[c], P[s][c], a0);
        a1 = fma(oh[c + 1], P[s][c + 1], a1);
        if ((c & 7) == 6) __builtin_amdgcn_sched_barrier(0);
      }
      x[s] = a0 + a1;
      const int row = t + s * kBT;
      xm[s] = (row > k && row < R) ? x[s] : 0.0;
    }
    // 3. exchange: G_c and row k
    const int buf = xi & 1;
    ++xi;
    const bool owner = t == (k % kBT);
    const int sk = k / kBT;
    if (owner) {
#pragma unroll
      for (int c = 0; c < NB; ++c)
        L.pn_rowk[buf][c] = (RPT == 1 || sk == 0) ? P[0][c] : P[RPT - 1][c];
    }
    const double G = pn_exchange<NB>(
        L,
        [&](int c) {
          double s = 0.0;
#pragma unroll
          for (int q = 0; q < RPT; ++q) s = fma(xm[q], P[q][c], s);
          return s;
        },
        buf);
    const double ak = (lane < NB) ? L.pn_rowk[buf][lane] : 0.0;
    // 4. reflector (dlarfg) from alpha = A(k, p), ||x(k+1:)||^2 = G_p
    const hq_reflector h = hq_dlarfg(readlane_d(ak, p), readlane_d(G, p));
    // 5. per slot: multiplier of x below row k, new row-k value; for an
    // eliminated slot l, v_l'v_k = scale_l (A(k, l) + scale G_l)
    const bool act = !sdone && lane != p;
    const double w = ak + h.scale * G;
    const double f = act ? h.tau * h.scale * w : 0.0;
    const double rk = act ? ak - h.tau * w : (lane == p ? h.beta : ak);
    if (wave == 0 && sdone && lane < b) L.u.T[k][L.slot_step[lane]] = ssc * w;
    if (lane < NB) {
      bc[lane] = f;
      bc[NB + lane] = rk;
    }
    wave_lds_sync();
    // 6. update my rows of every slot (f = 0 but for the live ones); the
    // owner of row k stores the new row
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      const double fc = bc[c];
#pragma unroll
      for (int s = 0; s < RPT; ++s) P[s][c] = fma(-fc, xm[s], P[s][c]);
      if ((c & 7) == 7) __builtin_amdgcn_sched_barrier(0);
    }
    if (owner) {
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        const double rc = bc[NB + c];
        if (RPT == 1 || sk == 0)
          P[0][c] = rc;
        else
          P[RPT - 1][c] = rc;
      }
    }
    // 7. partial-norm downdate (LAWN 176) of the live slots
    if (pivoting) {
      bool need = false;
      if (act && dn != 0.0) {
        double tt = fabs(rk) / dn;
        const double uu = (1.0 + tt) * (1.0 - tt);
        tt = (0.0 > uu) ? 0.0 : uu;
        const double ratio = dn / en;
        if (tt * ratio * ratio <= HQ_TOL3Z)
          need = true;
        else
          dn = dn * sqrt(tt);
      }
      // every wave holds the same flags, so all take the branch together
      if (__any(need)) {
        const double bel = pn_exchange<NB>(
            L,
            [&](int c) {
              double s = 0.0;
#pragma unroll
              for (int q = 0; q < RPT; ++q) {
                const int row = t + q * kBT;
                const double pv = (row > k && row < R) ? P[q][c] : 0.0;
                s = fma(pv, pv, s);
              }
              return s;
            },
            xi & 1);
        ++xi;
        if (need) {
          dn = (R - k - 1 > 0) ? sqrt(bel) : 0.0;
          en = dn;
        }
      }
    }
    // 8. bookkeeping
    if (lane == p) {
      sdone = true;
      ssc = h.scale;
    } else if (!sdone && spos == k) {
      spos = ppos;
    }
    if (t == 0) {
      tau[j + k] = h.tau;
      L.tau[k] = h.tau;
      L.slot_step[p] = k;
      const int a0 = L.map[j + k];
      L.map[j + k] = L.map[j + ppos];
      L.map[j + ppos] = a0;
    }
  }
  if (wave == 0 && lane < b) L.pn_scale[lane] = ssc;
  __syncthreads();
  // V' rows (steps) b..NB-1 and columns (rows of the panel) R..16*ceil(R/16)
  // hold zeros, so that the products read V' unmasked
  {
    const int row = t;
    if (row < ((R + 15) & ~15)) {
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c >= b || row >= R) Vt[c + (long)row * NB] = 0.0;
    }
  }
  // write the panel back (slot c -> its physical column) and V'
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    if (c < b) {
      const long pcol = L.slot_col[c];
      const int st = L.slot_step[c];
      const double sc = L.pn_scale[c];
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        const int row = t + s * kBT;
        if (row < R) {
          const double val = (row > st) ? sc * P[s][c] : P[s][c];
          A[(j + row) + pcol * lda] = val;
          Vt[st + (long)row * NB] = (row < st) ? 0.0 : ((row == st) ? 1.0 : val);
        }
      }
    }
  }
}

// T (dlarft Forward/Columnwise, WY.c:504-506) from S(l, k) = v_l'v_k and tau:
// T(k,k) = tau_k, T(0:k,k) = -tau_k T(0:k,0:k) S(0:k,k).  Wave 0, lane i owns
// row i of T.
template <int NB>
__device__ __forceinline__ void build_t(Lds<NB>& L, int b) {
  const int lane = threadIdx.x & 63;
  if ((threadIdx.x >> 6) != 0) return;
  const int i = lane;
  for (int k = 0; k < b; ++k) {
    const double tk = L.tau[k];
    if (i < k) {
      double s = 0.0;
#pragma unroll 8
      for (int l = 0; l < k; ++l)
        if (l >= i) s += L.u.T[i][l] * L.u.T[k][l];
      L.u.T[i][k] = -tk * s;
    } else if (i == k) {
      L.u.T[k][k] = tk;
    }
  }
}

// -V T' (R x b, ld kLdv) from V' and T, stored negated for the update.
template <int NB>
__device__ __forceinline__ void form_vt(Lds<NB>& L, const double* __restrict__ Vt,
                        double* __restrict__ VTc, int R, int b) {
  constexpr int IT = Cfg<NB>::IT;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l15 = lane & 15, l4 = lane >> 4;
  const int groups = (R + 15) / 16;
  for (int g = wave; g < groups; g += kWaves) {
    hq_d4 acc[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) acc[it] = hq_d4{0.0, 0.0, 0.0, 0.0};
    const int row = g * 16 + l15;
#pragma unroll
    for (int k0 = 0; k0 < NB; k0 += 4) {
      if (k0 < b) {
        const int k = k0 + l4;
        const double av = Vt[k + row * NB];  // zero-padded
#pragma unroll
        for (int it = 0; it < IT; ++it) {
          const int i = it * 16 + l15;
          const double bl = L.u.T[i][k];
          const double bv = (i < b && k < b && i <= k) ? bl : 0.0;
          acc[it] = mfma4(av, bv, acc[it]);
        }
      }
    }
#pragma unroll
    for (int it = 0; it < IT; ++it)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int rr = g * 16 + l4 + 4 * e, i = it * 16 + l15;
        VTc[rr + (long)i * kLdv] = -acc[it][e];  // zeros outside R x b
      }
  }
}

// C (R x ncols) := (I - V T V')' C = C - (V T') (V' C), C(row, col) read
// at cin(col)[row * rs] and written to cout(col)[row * rs].  Rounds of SB
// strips of 16 columns: the waves split the rows (16-row groups dealt
// round-robin), form their part of V'C for every strip of the round, sum the
// parts through LDS in a fixed order (two barriers per round), then re-read
// their C tiles into the MFMA accumulators and apply the rank-b update (the
// second read hits L2).  SB strips keep SB times the loads in flight.
template <int NB, class CIn, class COut>
__device__ __forceinline__ void apply_qt(Lds<NB>& L, const double* __restrict__ Vt,
                         const double* __restrict__ VTc, int R, int b,
                         int ncols, long rs, CIn cin, COut cout) {
  constexpr int IT = Cfg<NB>::IT, SB = Cfg<NB>::SB;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l15 = lane & 15, l4 = lane >> 4;
  const int groups = (R + 15) / 16;
  double* Wl = L.u.W;
  constexpr int WS = SB * IT * 4 * 64;  // doubles per wave in Wl
  for (int c0 = 0; c0 < ncols; c0 += 16 * SB) {
    const double* cp[SB];
    bool cok[SB];
#pragma unroll
    for (int sb = 0; sb < SB; ++sb) {
      const int col = c0 + 16 * sb + l15;
      cok[sb] = col < ncols;
      cp[sb] = cin(cok[sb] ? col : 0);
    }
    hq_d4 wacc[SB][IT];
#pragma unroll
    for (int sb = 0; sb < SB; ++sb)
#pragma unroll
      for (int it = 0; it < IT; ++it) wacc[sb][it] = hq_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
    for (int g = wave; g < groups; g += kWaves) {
      hq_d4 cg[SB];
#pragma unroll
      for (int sb = 0; sb < SB; ++sb)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          // rows past R repeat row R-1; V' is zero there
          const int row = g * 16 + l4 + 4 * e;
          cg[sb][e] = cp[sb][(long)min(row, R - 1) * rs];
        }
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        const int row = g * 16 + 4 * s4 + l4;
#pragma unroll
        for (int it = 0; it < IT; ++it) {
          const double av = Vt[it * 16 + l15 + row * NB];  // zero-padded
#pragma unroll
          for (int sb = 0; sb < SB; ++sb)
            wacc[sb][it] = mfma4(av, cg[sb][s4], wacc[sb][it]);
        }
      }
    }
#pragma unroll
    for (int sb = 0; sb < SB; ++sb)
#pragma unroll
      for (int it = 0; it < IT; ++it)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          Wl[wave * WS + ((sb * IT + it) * 4 + e) * 64 + lane] = wacc[sb][it][e];
    __syncthreads();
    hq_d4 ws[SB][IT];
#pragma unroll
    for (int sb = 0; sb < SB; ++sb)
#pragma unroll
      for (int it = 0; it < IT; ++it) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int o = ((sb * IT + it) * 4 + e) * 64 + lane;
          double sum = 0.0;
#pragma unroll
          for (int w = 0; w < kWaves; ++w) sum += Wl[w * WS + o];
          ws[sb][it][e] = sum;
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    __syncthreads();
#pragma unroll 1
    for (int g = wave; g < groups; g += kWaves) {
      hq_d4 cg[SB];
#pragma unroll
      for (int sb = 0; sb < SB; ++sb)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = g * 16 + l4 + 4 * e;
          cg[sb][e] = cp[sb][(long)min(row, R - 1) * rs];  // rows past R unused
        }
      const int row = g * 16 + l15;
#pragma unroll
      for (int ks = 0; ks < NB / 4; ++ks) {
        if (ks * 4 < b) {
          const double av = VTc[row + (ks * 4 + l4) * kLdv];  // zero-padded
#pragma unroll
          for (int sb = 0; sb < SB; ++sb)
            cg[sb] = mfma4(av, ws[sb][ks / 4][ks % 4], cg[sb]);
        }
      }
#pragma unroll
      for (int sb = 0; sb < SB; ++sb) {
        if (cok[sb]) {
          double* op = cout(c0 + 16 * sb + l15);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int rr = g * 16 + l4 + 4 * e;
            if (rr < R) op[(long)rr * rs] = cg[sb][e];
          }
        }
      }
    }
  }
}

// Y(:, c) -= G(:, j:j+b) R12(:, c) for the logical columns c >= j + b
// (R12 = A(j:j+b, c)).  Work item: a 16-column tile with all row tiles.
template <int NB>
__device__ __forceinline__ void y2_update(Lds<NB>& L, double* __restrict__ Y,
                          const double* __restrict__ G, int r,
                          const double* __restrict__ A, long lda, int j, int b,
                          int nrest) {
  constexpr int RT = Cfg<NB>::RT;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l15 = lane & 15, l4 = lane >> 4;
  const int rtn = (r + 15) / 16;
  for (int ct = wave; ct * 16 < nrest; ct += kWaves) {
    const int col = ct * 16 + l15;
    const bool cok = col < nrest;
    const long pcol = cok ? L.map[j + b + col] : 0;
    double* Yc = Y + pcol * r;
    const double* Ac = A + pcol * lda + j;
    hq_d4 acc[RT];
#pragma unroll
    for (int i = 0; i < RT; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = i * 16 + l4 + 4 * e;
        acc[i][e] = Yc[min(row, r - 1)];  // rows past r are not stored
      }
#pragma unroll
    for (int k0 = 0; k0 < NB; k0 += 4) {
      if (k0 < b) {
        const int k = k0 + l4;
        const int kc = min(k, b - 1);
        const double bv = *(k < b ? Ac + k : kZeros);  // R12, zero past b
#pragma unroll
        for (int i = 0; i < RT; ++i) {
          if (i < rtn) {
            const int row = i * 16 + l15;
            const double av = -G[min(row, r - 1) + (j + kc) * r];
            acc[i] = mfma4(av, bv, acc[i]);
          }
        }
      }
    }
#pragma unroll
    for (int i = 0; i < RT; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = i * 16 + l4 + 4 * e;
        if (i < rtn && row < r && cok) Yc[row] = acc[i][e];
      }
  }
}

// A(:, pos) = A(:, map[pos]) for every pos (a few rows at a time through
// LDS) and jpvt(pos) = jpvt(map[pos]).
template <int NB>
__device__ __forceinline__ void finish_perm(Lds<NB>& L, double* __restrict__ A, long lda, int m,
                            int n, int* __restrict__ jp) {
  const int t = threadIdx.x;
  int jv[kDim / kBT];
#pragma unroll
  for (int s = 0; s < kDim / kBT; ++s) {
    const int pos = t + s * kBT;
    jv[s] = (pos < n) ? jp[L.map[pos]] : 0;
  }
  for (int i0 = 0; i0 < m; i0 += kPermRows) {
    __syncthreads();
    for (int idx = t; idx < kPermRows * n; idx += kBT) {
      const int rr = idx % kPermRows, c = idx / kPermRows;
      if (i0 + rr < m) L.u.perm[rr * kPermPitch + c] = A[(i0 + rr) + (long)c * lda];
    }
    __syncthreads();
    for (int idx = t; idx < kPermRows * n; idx += kBT) {
      const int rr = idx % kPermRows, pos = idx / kPermRows;
      if (i0 + rr < m)
        A[(i0 + rr) + (long)pos * lda] = L.u.perm[rr * kPermPitch + L.map[pos]];
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < kDim / kBT; ++s) {
    const int pos = t + s * kBT;
    if (pos < n) jp[pos] = jv[s];
  }
}

template <int NB>
__device__ __forceinline__ void factor(Lds<NB>& L, const BatchArgs& a, const BatchItem& it,
                       double* __restrict__ work, PhaseClock& pc) {
  const int t = threadIdx.x;
  const int m = it.m, n = it.n, nb = it.nb, pp = a.pp, r = nb + pp;
  const long lda = it.lda;
  double* A = a.A[it.idx];
  int* jp = a.jpvt[it.idx];
  double* tau = a.tau[it.idx];
  const int mn = min(m, n);
  if (mn <= 0) return;
  double* Yw = work;
  double* Gw = Yw + kYW;
  double* Vt = Gw + kGW;
  double* VTc = Vt + kVW;
  for (int c = t; c < n; c += kBT) L.map[c] = c;
  pc.mark(HQB_PH_IDLE);
  sketch_gemm<NB>(a.omega, r, m, n, A, lda, Yw);
  __syncthreads();
  pc.mark(HQB_PH_SKETCH_GEMM);
  for (int j = 0; j < mn; j += nb) {
    const int b = min(nb, min(n - j, m - j));
    const bool last = (j + nb >= m) || (j + nb >= n);
    const int R = m - j, nrest = n - j - b;
    if (!last) {
      sketch_qrcp<NB>(L, Yw, r, j, n - j, b);
      __syncthreads();
      pc.mark(HQB_PH_SKETCH_QRCP);
    }
    panel_qrcp<NB>(L, A, lda, R, j, b, a.pivoting, tau, Vt);
    __syncthreads();
    pc.mark(HQB_PH_PANEL);
    if (nrest <= 0 && last) continue;
    build_t<NB>(L, b);
    __syncthreads();
    form_vt<NB>(L, Vt, VTc, R, b);
    __syncthreads();
    pc.mark(HQB_PH_T);
    if (nrest > 0) {
      auto acol = [&](int c) -> double* {
        return A + (long)L.map[j + b + c] * lda + j;
      };
      apply_qt<NB>(L, Vt, VTc, R, b, nrest, 1, acol, acol);
      __syncthreads();
      pc.mark(HQB_PH_TRAIL);
    }
    if (!last) {
      const double* Gin = (j == 0) ? a.omega : Gw;
      auto gin = [&](int c) -> const double* { return Gin + c + (long)j * r; };
      auto gout = [&](int c) -> double* { return Gw + c + (long)j * r; };
      apply_qt<NB>(L, Vt, VTc, R, b, r, r, gin, gout);
      __syncthreads();
      if (nrest > 0) {
        y2_update<NB>(L, Yw, Gw, r, A, lda, j, b, nrest);
        __syncthreads();
      }
      pc.mark(HQB_PH_DOWNDATE);
    }
  }
  finish_perm<NB>(L, A, lda, m, n, jp);
  pc.mark(HQB_PH_FINISH);
}

// Workgroups per CU of each variant (registers: 128 VGPRs per thread allow
// two 512-thread workgroups per CU).
template <int NB>
struct Occ {
  static constexpr int value = NB <= 16 ? 2 : 1;
};

// (__launch_bounds__' second argument: waves per SIMD.)
template <int NB>
__global__ __launch_bounds__(kBT, Occ<NB>::value * kWaves / 4) void
batch_hqrrp_kernel(
    BatchArgs a) {
  __shared__ Lds<NB> L;
  double* work = a.work + (long)blockIdx.x * kWork;
  PhaseClock pc;
  pc.start(L.clk, a.prof != nullptr);
  for (int i = threadIdx.x; i < 192; i += kBT) {
    L.onehot[i] = (i == 64) ? 1.0 : 0.0;
    L.step_gt[i] = (i > 64) ? 1.0 : 0.0;
  }
  for (;;) {
    if (threadIdx.x == 0) L.item = atomicAdd(a.counter, 1);
    __syncthreads();
    const int q = L.item;
    __syncthreads();
    if (q >= a.count) break;
    const BatchItem it = a.items[q];
    factor<NB>(L, a, it, work, pc);
    __syncthreads();
  }
  pc.mark(HQB_PH_IDLE);
  pc.flush(a.prof);
}

// Uniform values in (-1, 1) from a hash of (seed, index): input for timing
// runs (the autotuner), not a statistical generator.
__global__ void batch_fill_kernel(double* p, long n, unsigned seed) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    unsigned long long z = (unsigned long long)i * 0x9E3779B97F4A7C15ull +
                           ((unsigned long long)seed << 32);
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    p[i] = ((double)(z >> 11) + 0.5) * (2.0 / 9007199254740992.0) - 1.0;
  }
}

__global__ void batch_reset_kernel(int* counter, int n) {
  if ((int)threadIdx.x < n) counter[threadIdx.x] = 0;
}

}  // namespace

long batch_work_doubles() { return kWork; }

int batch_class_nb(int nb) { return nb > 32 ? 64 : (nb > 16 ? 32 : 16); }

int batch_wgs_per_cu(int class_nb) {
  return class_nb == 64 ? Occ<64>::value
                        : (class_nb == 32 ? Occ<32>::value : Occ<16>::value);
}

hipError_t launch_batch_fill(double* p, long n, unsigned seed, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  long blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(batch_fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                     p, n, seed);
  return hipGetLastError();
}

hipError_t launch_batch_reset(int* counters, int n, hipStream_t s) {
  hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, s, counters, n);
  return hipGetLastError();
}

hipError_t launch_batch_hqrrp(const BatchArgs& a, int class_nb, int wgs,
                              hipStream_t s) {
  if (a.count <= 0 || wgs <= 0) return hipSuccess;
  if (class_nb == 64)
    hipLaunchKernelGGL(batch_hqrrp_kernel<64>, dim3(wgs), dim3(kBT), 0, s, a);
  else if (class_nb == 32)
    hipLaunchKernelGGL(batch_hqrrp_kernel<32>, dim3(wgs), dim3(kBT), 0, s, a);
  else
    hipLaunchKernelGGL(batch_hqrrp_kernel<16>, dim3(wgs), dim3(kBT), 0, s, a);
  return hipGetLastError();
}
