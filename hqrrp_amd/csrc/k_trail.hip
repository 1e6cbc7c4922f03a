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
#include <algorithm>
#include <type_traits>

#include "common.h"
#include "kernels.h"

namespace {

constexpr int kTile = 64;  // output tile edge (rows and columns)
constexpr int kBK64 = 32;  // K rows staged per step in trail_vtc64

__device__ __forceinline__ double vmask(double v, long i, long c) {
  return (i < c) ? 0.0 : ((i == c) ? 1.0 : v);
}

// C(mrows x nc) -= V(mrows x b) * W(b x nc) on 256 x 64 tiles (written to
// Cout when given: the update then runs out of place), computed
// transposed as C' := C' + W' (-V') so that the row-strided f64 accumulator
// layout maps to 16 consecutive rows of one column of C per lane group:
// every C load and store moves whole 128-byte lines.  The C tile is loaded
// straight into the accumulators, so C makes exactly one round trip through
// HBM.  Each wavefront owns a 64x64 quadrant of the tile -- the four stacked
// down the rows, so a tile is 2 KB contiguous in each of its columns (C
// streams ~15% faster through HBM that way than as 128 x 128 tiles) -- and
// runs on its own: no LDS and no barrier.  Both MFMA operands go from memory (W' and V stay
// in L2: every tile of a row or column range re-reads the same panel) into
// operand registers -- lane l supplies rows/columns l % 16 of reflector
// k0 + l / 16, 16 lanes covering a 128-byte line -- through a four-step
// register ring that keeps three K steps of loads in flight while the
// matrix cores work.  K runs over all 16 steps of 4 (operands beyond b are
// zero, adding exact zeros), so the ring has no exit branch.  Row tiles vary
// fastest across workgroups so concurrently running tiles cover different
// rows (different HBM channels) of the power-of-two-strided columns.
// Base of virtual column v (C's columns, then C2's from a.nca), or null
// for the gap and beyond.
__device__ __forceinline__ const double* vcol(const TrailArgs& a, long v) {
  if (v < a.nc) return a.C + v * a.ldc;
  if (a.C2 && v >= a.nca && v < a.nca + a.nc2)
    return a.C2 + (v - a.nca) * a.ldc2;
  return nullptr;
}

constexpr int kUR = 256;  // update tile rows (four wavefronts of 64)
constexpr int kUC = 64;   // update tile columns

// One wavefront's 64x64 quadrant of a tile.  FAST: an interior tile with
// all 64 reflectors and no unit-lower mask, so no load or store needs a
// predicate and nothing separates the C and operand loads from their uses.
template <bool FAST>
__device__ __forceinline__ void update_quadrant(const TrailArgs& a, long m0,
                                                long nv, long n0, long nc,
                                                const double* Cb, long ldc,
                                                double* Ob, long ldo,
                                                bool masked, int b, int wave) {
  const int lane = threadIdx.x & 63;
  const int wm = wave, wn = 0;  // 64-row quadrant of the 256 x 64 tile
  const int l15 = lane & 15, l4 = lane >> 4;
  // acc[i][j] element r: C'(n, m), n = n0+64wn+16i+l/16+4r, m = m0+64wm+16j+l%16,
  // i.e. cbase[16j + (16i+4r)*ldc] with the lane's base below.
  const bool interior = FAST || (m0 + kUR <= a.mrows && n0 + kUC <= nc);
  const long lm = m0 + wm * 64 + l15, ln = n0 + wn * 64 + l4;
  const double* cbase = Cb + lm + ln * ldc;
  double* obase = Ob + lm + ln * ldo;
  hq_d4 acc[4][4];
  if (interior) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double* cc = cbase + (long)(16 * i + 4 * r) * ldc;
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j][r] = __builtin_nontemporal_load(cc + 16 * j);
      }
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const bool nok = ln + 16 * i + 4 * r < nc;
        const double* cc = cbase + (long)(16 * i + 4 * r) * ldc;
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j][r] = (nok && lm + 16 * j < a.mrows) ? cc[16 * j] : 0.0;
      }
  }

  // Operands of K step s (reflectors 4s..4s+3): af[i] = W'(n_i, k), n_i =
  // n0+64wn+16i+l%16; bf[j] = -V(m_j, k) (unit lower on the top tile),
  // m_j = m0+64wm+16j+l%16; k = 4s + l/16.  Every load is issued
  // unconditionally (an out-of-range operand reads the tile's first row or
  // column, which exists, and is then replaced by zero), so the loads stay
  // in flight across the ring instead of behind branches.
  const long wcol = n0 + wn * 64 + l15, vrow = m0 + wm * 64 + l15;
  auto load = [&](double (&af)[4], double (&bf)[4], int st) {
    const int k = 4 * st + l4;
    if (FAST) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
        af[i] = a.W[nv + wn * 64 + 16 * i + l15 + (long)k * a.ldw];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        bf[j] = -a.V[m0 + wm * 64 + 16 * j + l15 + (long)k * a.ldv];
      return;
    }
    const bool kin = k < b;
    const long kw = (kin ? k : 0) * a.ldw, kv = (kin ? k : 0) * a.ldv;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const bool ok = kin && wcol + 16 * i < nc;
      const double w = a.W[nv + (ok ? wn * 64 + 16 * i + l15 : 0) + kw];
      af[i] = ok ? w : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const bool ok = kin && vrow + 16 * j < a.mrows;
      double v = a.V[m0 + (ok ? wm * 64 + 16 * j + l15 : 0) + kv];
      if (masked) v = vmask(v, vrow + 16 * j, k);
      bf[j] = ok ? -v : 0.0;
    }
  };
  auto mma = [&](const double (&af)[4], const double (&bf)[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j],
                                                          acc[i][j], 0, 0, 0);
  };
  constexpr int kSteps = 16;  // 64 / 4
  double ra[4][4], rb[4][4];
  load(ra[0], rb[0], 0);
  load(ra[1], rb[1], 1);
  load(ra[2], rb[2], 2);
#pragma unroll
  for (int st = 0; st < kSteps; st += 4) {
    load(ra[3], rb[3], st + 3);
    mma(ra[0], rb[0]);
    if (st + 4 < kSteps) load(ra[0], rb[0], st + 4);
    mma(ra[1], rb[1]);
    if (st + 5 < kSteps) load(ra[1], rb[1], st + 5);
    mma(ra[2], rb[2]);
    if (st + 6 < kSteps) load(ra[2], rb[2], st + 6);
    mma(ra[3], rb[3]);
  }
  if (interior) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double* cc = obase + (long)(16 * i + 4 * r) * ldo;
#pragma unroll
        for (int j = 0; j < 4; ++j)
          __builtin_nontemporal_store(acc[i][j][r], cc + 16 * j);
      }
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const bool nok = ln + 16 * i + 4 * r < nc;
        double* cc = obase + (long)(16 * i + 4 * r) * ldo;
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (nok && lm + 16 * j < a.mrows) cc[16 * j] = acc[i][j][r];
      }
  }
}

// The 256 x 64 tile at row m0 and virtual column nv, by four wavefronts
// (wave = 0..3, one 64-row quadrant each).
__device__ __forceinline__ void update_tile(const TrailArgs& a, long m0,
                                            long nv, int wave) {
  const int b = a.b;
  const bool masked = a.vmask && m0 < b;  // only the top tile holds R / the diagonal
  // The tile's column range: C's, C2's (updated in place) or the gap.
  long n0, nc, ldc, ldo;
  double *Cb, *Ob;
  if (nv < a.nc) {
    n0 = nv;
    nc = a.nc;
    Cb = a.C;
    ldc = a.ldc;
    Ob = a.Cout ? a.Cout : a.C;
    ldo = a.Cout ? a.ldco : a.ldc;
  } else if (a.C2 && nv >= a.nca && nv < a.nca + a.nc2) {
    n0 = nv - a.nca;
    nc = a.nc2;
    Cb = Ob = a.C2;
    ldc = ldo = a.ldc2;
  } else {
    return;
  }

  const bool fast = m0 + kUR <= a.mrows && n0 + kUC <= nc && b == 64 && !masked;
  if (fast)
    update_quadrant<true>(a, m0, nv, n0, nc, Cb, ldc, Ob, ldo, false, 64, wave);
  else
    update_quadrant<false>(a, m0, nv, n0, nc, Cb, ldc, Ob, ldo, masked, b,
                           wave);
}

__global__ __launch_bounds__(HQ_NT, 2) void trail_update_kernel(TrailArgs a) {
  update_tile(a, (long)blockIdx.x * kUR, (long)blockIdx.y * kUC,
              threadIdx.x >> 6);
}

// ---------------------------------------------------------------------------
// trail_band_kernel: the same update (every element of C and C2 receives the
// same MFMA products in the same order as in trail_update_kernel, so the
// results are bitwise equal) as a persistent launch of one 512-thread
// workgroup per CU, for launches large enough to give every workgroup
// several tiles.
//
// The tiles of full columns (256 rows x 64 virtual columns, the last band
// possibly short) are numbered band-major (a band = 256 rows) and dealt to
// the workgroups in contiguous runs, so that a workgroup keeps one band's V
// for many tiles:
//   * -V of the band (256 rows x 64 reflectors) sits in LDS, staged once per
//     band; W' of the tile (64 columns x 64 reflectors) sits in LDS, the next
//     tile's loaded into registers while the current tile's MFMAs run and
//     written between two workgroup barriers at the tile's end;
//   * the eight wavefronts (two per SIMD) each own a 32-row slice of the
//     tile and hold two accumulator sets: while the matrix cores work on one
//     tile in one set, the previous tile's results are stored out of the
//     other set and the next tile's C is loaded into it (four stores and
//     four loads per K step, the stores ahead of the loads), so the
//     wavefront never has more than a few dozen memory operations in flight
//     and never waits for C;
//   * C is addressed as a buffer resource based at the tile's origin (its
//     size ends after row 255 of the tile's last column); lanes whose rows lie
//     beyond the matrix get a lane offset past that size, so their loads
//     return zero and their stores are dropped;
//   * the operand LDS layouts give every lane 16 contiguous bytes per
//     operand (-V: [wave][k][row % 16][row / 16 % 2]; W': [k][col % 16]
//     [col / 16]), whole wavefronts reading 1-2 KB runs.
// Tiles with fewer than 64 columns (C's and C2's last ones) are left to
// trail_edge_kernel (the per-tile code of trail_update_kernel).  V beyond b reflectors or below mrows is staged as zero, so
// every tile runs all 16 K steps (exact zeros added, as in
// trail_update_kernel); the unit-lower mask of the top rows is applied while
// staging.
namespace {

constexpr int kBT = 512;        // threads of a band workgroup
constexpr int kBW = kBT / 64;   // wavefronts (two per SIMD)
constexpr int kBR = kUR / kBW;  // rows of a wavefront's slice (32)
typedef unsigned int hq_u2 __attribute__((ext_vector_type(2)));

// Wavefront-uniform copies (the compiler then keeps them, and the buffer
// resources built from them, in scalar registers).
__device__ __forceinline__ long band_u(long v) {
  const unsigned long long u = (unsigned long long)v;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  return (long)(((unsigned long long)hi << 32) | lo);
}
template <class T>
__device__ __forceinline__ T* band_up(T* p) {
  return reinterpret_cast<T*>(band_u(reinterpret_cast<long>(p)));
}

// Tile counts per band: full C tiles, full C2 tiles, partial ones.
struct BandGeom {
  long nfc, nf2, full, npart, nbands;
};

__device__ __forceinline__ BandGeom band_geom(const TrailArgs& a) {
  BandGeom g;
  g.nfc = a.nc / kUC;
  g.nf2 = a.C2 ? a.nc2 / kUC : 0;
  g.full = g.nfc + g.nf2;
  g.npart = (a.nc % kUC ? 1 : 0) + (a.C2 && a.nc2 % kUC ? 1 : 0);
  g.nbands = hq_cdiv(a.mrows, kUR);
  return g;
}

// Full-column tile t (band-major) as a scalar description.
struct BandTile {
  const double* in;  // origin of the tile in the input
  double* out;       // ... and in the output
  long ldi, ldo;
  long nv;    // first virtual column (row of W')
  long band;
  int nrows;  // valid rows (<= 256)
};

__device__ __forceinline__ BandTile band_tile(const TrailArgs& a,
                                              const BandGeom& g, long t) {
  BandTile bt;
  bt.band = t / g.full;
  const long k = t - bt.band * g.full;
  const long m0 = bt.band * kUR;
  bt.nrows = (int)((long)a.mrows - m0 < kUR ? (long)a.mrows - m0 : kUR);
  if (k < g.nfc) {
    const long n0 = k * kUC;
    bt.ldi = a.ldc;
    bt.in = a.C + m0 + n0 * a.ldc;
    bt.ldo = a.Cout ? a.ldco : a.ldc;
    bt.out = (a.Cout ? a.Cout : a.C) + m0 + n0 * bt.ldo;
    bt.nv = n0;
  } else {
    const long n0 = (k - g.nfc) * kUC;
    bt.ldi = bt.ldo = a.ldc2;
    bt.in = bt.out = a.C2 + m0 + n0 * a.ldc2;
    bt.nv = a.nca + n0;
  }
  bt.in = band_up(bt.in);
  bt.out = band_up(bt.out);
  bt.ldi = band_u(bt.ldi);
  bt.ldo = band_u(bt.ldo);
  bt.nv = band_u(bt.nv);
  bt.band = band_u(bt.band);
  bt.nrows = __builtin_amdgcn_readfirstlane(bt.nrows);
  return bt;
}

// Buffer resource over a full-column tile (needs ld >= 256, so its size,
// which ends after row 255 of column 63, lies below the next column).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t band_rsrc(const double* p,
                                                           long ld) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(p), 0,
                                           (int)((63 * ld + kUR) * 8),
                                           0x00020000);
}

// Byte offset of entry e's column group (16 i + 4 r) within a tile.
__device__ __forceinline__ int band_soff(int e, long ld) {
  return (int)((16 * (e >> 3) + 4 * ((e >> 1) & 3)) * ld * 8);
}

}  // namespace

__global__ __launch_bounds__(kBT, 1) void trail_band_kernel(TrailArgs a) {
  __shared__ __attribute__((aligned(16))) double vl[kBW * 64 * kBR];  // 128 KB
  __shared__ __attribute__((aligned(16))) double wl[64 * 64];          // 32 KB
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int l15 = lane & 15, l4 = lane >> 4;
  const BandGeom geo = band_geom(a);
  const long total = geo.nbands * geo.full;
  const long G = gridDim.x, g = blockIdx.x;
  const long t0 = total * g / G, t1 = total * (g + 1) / G;
  const int b = a.b;

  if (t0 < t1) {
    // -V(band rows, k) into LDS: thread t stages row m = t % 256 for the
    // reflectors of parity t / 256.  Rows beyond mrows and reflectors beyond
    // b are zero; the top rows carry the unit-lower mask.
    auto stage_v = [&](long band) {
      const int m = t & (kUR - 1), kh = t >> 8;
      const long row = band * kUR + m;
      const bool rok = row < a.mrows;
      const double* vr = a.V + (rok ? row : 0);
      const int base = (m / kBR) * (64 * kBR) + (m & 15) * 2 + ((m >> 4) & 1);
#pragma unroll 1  // few registers: this runs with both accumulator sets live
      for (int c = 0; c < 4; ++c) {
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int k = 2 * (8 * c + q) + kh;
          v[q] = vr[(long)(k < b ? k : 0) * a.ldv];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int k = 2 * (8 * c + q) + kh;
          const double x = a.vmask ? vmask(v[q], row, k) : v[q];
          vl[base + k * kBR] = (rok && k < b) ? -x : 0.0;
        }
      }
    };
    // W' of a tile: thread t holds columns 16 i + (p % 16) of reflector
    // p / 16 for p = t and t + 512; reflectors beyond b are zero.
    double ws[2][4];
    auto load_w = [&](const BandTile& bt) {
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int p = t + kBT * u, k = p >> 4, nn = p & 15;
        const double* wr = a.W + bt.nv + nn + (long)(k < b ? k : 0) * a.ldw;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const double w = wr[16 * i];
          ws[u][i] = k < b ? w : 0.0;
        }
      }
    };
    auto store_w = [&]() {
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int p = t + kBT * u;
        hq_d2* d = reinterpret_cast<hq_d2*>(&wl[p * 4]);
        d[0] = hq_d2{ws[u][0], ws[u][1]};
        d[1] = hq_d2{ws[u][2], ws[u][3]};
      }
    };

    // Accumulator set s, entry e (0..31): acc[s][i][j][r] = C'(16 i + 4 r +
    // l4, 32 wave + 16 j + l15) of its tile, i = e / 8, j = e % 2,
    // r = e / 2 % 4.
    hq_d4 acc[2][4][2];
#define HQ_ENT(s, e) acc[s][(e) >> 3][(e) & 1][((e) >> 1) & 3]
    // Lane byte offsets of the slice's rows j = 0, 1 (past the buffer for
    // rows beyond the tile's valid ones).
    auto lane_off = [&](const BandTile& bt, long ld, int (&vo)[2]) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int row = wave * kBR + 16 * j + l15;
        vo[j] = row < bt.nrows ? (int)((row + l4 * ld) * 8) : 0x7ffffff0;
      }
    };
    auto ld_entry = [&](__amdgpu_buffer_rsrc_t r, const int (&vo)[2], long ld,
                        int e) -> double {
      return __builtin_bit_cast(
          double, __builtin_amdgcn_raw_buffer_load_b64(r, vo[e & 1],
                                                       band_soff(e, ld), 2));
    };
    auto st_entry = [&](double v, __amdgpu_buffer_rsrc_t r, const int (&vo)[2],
                        long ld, int e) {
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(hq_u2, v), r,
                                            vo[e & 1], band_soff(e, ld), 2);
    };

    long tt = t0;
    BandTile cur = band_tile(a, geo, tt);
    long band = cur.band;
    stage_v(band);
    {
      int vo[2];
      lane_off(cur, cur.ldi, vo);
      const auto r = band_rsrc(cur.in, cur.ldi);
#pragma unroll
      for (int e = 0; e < 32; ++e) HQ_ENT(0, e) = ld_entry(r, vo, cur.ldi, e);
    }
    load_w(cur);
    store_w();
    // Nothing of the prologue stays in flight into the tile loop.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    bool prev = false;  // the other set holds the previous tile's results
    BandTile pre = cur;
    const double* wb = &wl[4 * lane];
    const double* vb = &vl[wave * (64 * kBR) + 2 * lane];

    auto step = [&](auto set) {
      constexpr int S = decltype(set)::value, O = 1 - S;
      const bool nxt = tt + 1 < t1;
      // Without a next tile the loads below re-read this tile (harmless:
      // the other set is not used again).
      const BandTile nx = nxt ? band_tile(a, geo, tt + 1) : cur;
      const bool chg = nx.band != band;
      int von[2], vop[2];
      lane_off(nx, nx.ldi, von);
      lane_off(pre, pre.ldo, vop);
      const auto rn = band_rsrc(nx.in, nx.ldi);
      const auto rp = band_rsrc(pre.out, pre.ldo);
      load_w(nx);
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        const hq_d2 a01 = *reinterpret_cast<const hq_d2*>(wb + 256 * s);
        const hq_d2 a23 = *reinterpret_cast<const hq_d2*>(wb + 256 * s + 2);
        const hq_d2 b01 = *reinterpret_cast<const hq_d2*>(vb + 2 * 64 * s);
        const double af[4] = {a01[0], a01[1], a23[0], a23[1]};
        const double bf[2] = {b01[0], b01[1]};
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acc[S][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                af[i], bf[j], acc[S][i][j], 0, 0, 0);
        if (s < 8 && prev) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int e = 4 * s + q;
            st_entry(HQ_ENT(O, e), rp, vop, pre.ldo, e);
          }
        }
        if (s >= 4 && s < 12) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int e = 4 * (s - 4) + q;
            HQ_ENT(O, e) = ld_entry(rn, von, nx.ldi, e);
          }
        }
      }
      if (!nxt || chg) {
        // The end of a band (or of the run): this tile's results go out now.
        int vo[2];
        lane_off(cur, cur.ldo, vo);
        const auto r = band_rsrc(cur.out, cur.ldo);
#pragma unroll
        for (int e = 0; e < 32; ++e) st_entry(HQ_ENT(S, e), r, vo, cur.ldo, e);
        prev = false;
      } else {
        prev = true;
      }
      pre = cur;
      cur = nx;
      __syncthreads();  // every wavefront is done with this tile's W' and V
      if (nxt) {
        if (chg) {
          band = nx.band;
          stage_v(band);
        }
        store_w();
        __syncthreads();
      }
      ++tt;
    };
    while (tt < t1) {
      step(std::integral_constant<int, 0>{});
      if (tt >= t1) break;
      step(std::integral_constant<int, 1>{});
    }
#undef HQ_ENT
  }

}

// The tiles trail_band_kernel leaves: the partial-column tiles (C's last and
// C2's last, at most two per band), one per workgroup.
__global__ __launch_bounds__(HQ_NT, 2) void trail_edge_kernel(TrailArgs a) {
  const BandGeom geo = band_geom(a);
  const long e = blockIdx.x;
  const long band = e / geo.npart, w = e - band * geo.npart;
  const bool cpart = (a.nc % kUC) && w == 0;
  const long nv = cpart ? geo.nfc * kUC : a.nca + geo.nf2 * kUC;
  update_tile(a, band * kUR, nv, threadIdx.x >> 6);
}

// The partials of M = V'[C C2] on the stream-K partition (VtcPart): each
// wavefront walks its run of (64-column group, 16-row chunk) units against
// all 64 reflectors (4 x 4 MFMA blocks, accumulators in AGPRs), computed
// transposed (C' V) so the accumulator stores whole columns of a partial,
// and writes its accumulators to the group's partial slice whenever its run
// leaves a group.  No LDS and no barrier: both operands go from memory
// straight into MFMA operand registers.  K is permuted inside each 16-row
// chunk: in MFMA substep e (0..3) lane l supplies row 4*(l/16) + e of the
// chunk for both operands, so every lane reads 4 consecutive doubles (32
// bytes) of one column of C and of V per chunk, and 16 lanes cover a
// column's 128-byte line.  A four-chunk register ring keeps the next three
// chunks' loads in flight (across group boundaries) while the matrix cores
// work on the current one.  Every wavefront runs q or q + 1 chunks, so all
// CUs finish together whatever the column count.  VEC: the 32-byte groups
// are 32-byte aligned (vector loads); otherwise the same values are loaded
// one double at a time, so both forms compute identical bits.
constexpr int kCK = kVtcChunk;  // K rows per chunk
constexpr int kRing = 4;  // chunks of the register ring (kRing - 1 in flight)
template <bool VEC, int RING = kRing>
__global__ __launch_bounds__(HQ_NT, 1) void trail_vtc_kernel(TrailArgs a) {
  const int lane = threadIdx.x & 63;
  const int l15 = lane & 15, g4 = lane >> 4;
  const VtcPart vp = vtc_part(a);
  const long w = (long)blockIdx.x * (HQ_NT / 64) +
                 __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long u0 = vp.start(w);
  const long nu = vp.start(w + 1) - u0;
  if (nu <= 0) return;
  const long nch = vp.nch, kend = a.mrows;
  const int b = a.b;

  const double* vptr[4];
  bool vok[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int p = 16 * i + l15;
    vok[i] = p < b;
    vptr[i] = a.V + (long)(vok[i] ? p : 0) * a.ldv + 4 * g4;
  }
  // Load cursor: group lg, chunk lc of the next unit to load; cp/cok are
  // that group's columns.
  long lg = u0 / nch, lc = u0 - lg * nch, ld_left = nu;
  const double* cp[4];
  bool cok[4];
  auto set_cols = [&](long g) {
    const long v0 = vp.vcol0(g);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const double* col = vcol(a, v0 + 16 * i + l15);
      cok[i] = col != nullptr;
      cp[i] = (cok[i] ? col : a.C) + 4 * g4;
    }
  };
  set_cols(lg);

  hq_d4 acc[4][4];  // [n block][p block]
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = hq_d4{0.0, 0.0, 0.0, 0.0};

  hq_d4 cb[RING][4], vb[RING][4];
  auto load = [&](hq_d4 (&cr)[4], hq_d4 (&vr)[4]) {
    if (ld_left <= 0) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        cr[i] = hq_d4{0.0, 0.0, 0.0, 0.0};
        vr[i] = hq_d4{0.0, 0.0, 0.0, 0.0};
      }
      return;
    }
    const long k0 = lc * kCK;
    const long r0 = k0 + 4 * g4;  // this lane's first row
    if (VEC && k0 + kCK <= kend) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        cr[i] = cok[i] ? __builtin_nontemporal_load(
                             reinterpret_cast<const hq_d4*>(cp[i] + k0))
                       : hq_d4{0.0, 0.0, 0.0, 0.0};
        vr[i] = vok[i] ? *reinterpret_cast<const hq_d4*>(vptr[i] + k0)
                       : hq_d4{0.0, 0.0, 0.0, 0.0};
      }
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const bool rok = r0 + e < kend;
          cr[i][e] = (cok[i] && rok) ? cp[i][k0 + e] : 0.0;
          vr[i][e] = (vok[i] && rok) ? vptr[i][k0 + e] : 0.0;
        }
    }
    if (a.vmask && k0 < b) {  // the top b x b block of V is unit lower
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e) vr[i][e] = vmask(vr[i][e], r0 + e, 16 * i + l15);
    }
    --ld_left;
    if (++lc == nch) {
      lc = 0;
      ++lg;
      if (ld_left > 0) set_cols(lg);
    }
  };
  auto compute = [&](const hq_d4 (&cr)[4], const hq_d4 (&vr)[4]) {
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(cr[i][e], vr[j][e],
                                                            acc[i][j], 0, 0, 0);
  };
  // Group cg's partial: acc[i][j] element r of lane l is partial(p, n) with
  // p = 16j + l%16 and n = vcol0(cg) + 16i + l/16 + 4r.
  long cg = u0 / nch;
  long cnext = (cg + 1) * nch - u0;  // first unit (relative) of group cg + 1
  auto flush = [&]() {
    double* out = a.part + (w - vp.first(cg)) * b * a.ncv;
    const long v0 = vp.vcol0(cg);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int p = j * 16 + l15;
          const long gn = v0 + i * 16 + g4 + 4 * r;
          if (p < b && gn < a.ncv) out[p + gn * b] = acc[i][j][r];
        }
  };
  // Before unit u is multiplied, unit u + RING - 1 is loaded into the slot
  // unit u - 1 left; ring indices are compile-time constants.
#pragma unroll
  for (int q = 0; q + 1 < RING; ++q) load(cb[q], vb[q]);
  for (long it = 0; it < nu; it += RING) {
#pragma unroll
    for (int q = 0; q < RING; ++q) {
      load(cb[(q + RING - 1) % RING], vb[(q + RING - 1) % RING]);
      if (it + q < nu) {
        if (it + q == cnext) {
          flush();
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = hq_d4{0.0, 0.0, 0.0, 0.0};
          ++cg;
          cnext += nch;
        }
        compute(cb[q], vb[q]);
      }
    }
  }
  flush();
}

// The same product on 64-column tiles (4 wavefronts of 32 x 32, two
// workgroups per CU): more workgroups for the small trailing matrices of the
// last blocks, where 256-column tiles leave CUs idle.  K rows are staged
// kBK64 at a time, k-contiguous in LDS (pitch kBK64+4: conflict-free stores,
// at most 2-way conflicts on fragment reads); the next slice is loaded into
// registers while the matrix cores work.
__global__ __launch_bounds__(HQ_NT, 2) void trail_vtc64_kernel(TrailArgs a) {
  constexpr int P = kBK64 + 4;
  __shared__ double Vs[kTile][P];  // Vs[p][k] = V(k0+k, p)
  __shared__ double Cs[kTile][P];  // Cs[n][k] = C(k0+k, n0+n)
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wn = wave & 1, wp = wave >> 1;  // 32-column / 32-p quadrant
  const long n0 = (long)blockIdx.x * kTile;
  const int s = blockIdx.y;
  const long kbeg = (long)s * a.kchunk;
  const long kend = min((long)a.mrows, kbeg + a.kchunk);
  const int b = a.b;
  constexpr int kPer = kBK64 * kTile / HQ_NT;  // 8 elements per thread
  const int l15 = lane & 15, l4 = lane >> 4;

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
      int k = idx & (kBK64 - 1), p = idx / kBK64;
      long gk = k0 + k;
      double vv = (gk < kend && p < b) ? a.V[gk + (long)p * a.ldv] : 0.0;
      rv[q] = a.vmask ? vmask(vv, gk, p) : vv;
      const double* col = vcol(a, n0 + p);
      rc[q] = (gk < kend && col) ? col[gk] : 0.0;
    }
  };
  // The K slice is walked starting at a per-column-tile rotation, so that
  // concurrently running workgroups read different rows (spreading the
  // power-of-two-strided columns over HBM channels).
  const long nsteps = (kend - kbeg + kBK64 - 1) / kBK64;
  const long rot = nsteps > 0 ? ((long)blockIdx.x * 7) % nsteps : 0;
  auto kstart = [&](long it) { return kbeg + ((it + rot) % nsteps) * kBK64; };
  if (nsteps > 0) load(kstart(0));
  for (long it = 0; it < nsteps; ++it) {
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      int idx = t + HQ_NT * q;
      int k = idx & (kBK64 - 1), p = idx / kBK64;
      Vs[p][k] = rv[q];
      Cs[p][k] = rc[q];
    }
    __syncthreads();
    if (it + 1 < nsteps) load(kstart(it + 1));
#pragma unroll
    for (int kk = 0; kk < kBK64; kk += 4) {
      double af[2], bf[2];  // A = C' (n x k), B = V (k x p)
#pragma unroll
      for (int i = 0; i < 2; ++i) af[i] = Cs[wn * 32 + i * 16 + l15][kk + l4];
#pragma unroll
      for (int j = 0; j < 2; ++j) bf[j] = Vs[wp * 32 + j * 16 + l15][kk + l4];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j],
                                                            acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // acc element r of lane l: partial(p, n) with p = 32wp+16j+l%16 and
  // n = n0+32wn+16i+l/16+4r.
  double* out = a.part + (long)s * b * a.ncv;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int p = wp * 32 + j * 16 + l15;
        long gn = n0 + wn * 32 + i * 16 + l4 + 4 * r;
        if (p < b && gn < a.ncv) out[p + gn * b] = acc[i][j][r];
      }
}


// W(:, n0:n0+64) = T' * sum_s part[s](:, n0:n0+64) (or T * ..., to apply Q
// instead of Q'), slices added in order, stored transposed (W' is nc x b) for
// coalesced staging in trail_update.
__global__ __launch_bounds__(HQ_NT) void trail_tw_kernel(TrailArgs a) {
  __shared__ double Ms[kTile][kTile + 1];  // Ms[p][n]
  __shared__ double Ts[kTile][kTile + 1];  // Ts[p][q] = T(p, q)
  const int t = threadIdx.x, b = a.b;
  const long n0 = (long)blockIdx.x * kTile;
  const long slice = (long)b * a.ncv;
  const int S = a.sk ? vtc_part(a).count(n0 / 64) : a.splits;
  for (int idx = t; idx < b * kTile; idx += HQ_NT) {
    int p = idx % b, n = idx / b;
    long gn = n0 + n;
    double s = 0.0;
    if (gn < a.nc)
      for (int q = 0; q < S; ++q) s += a.part[q * slice + p + gn * b];
    Ms[p][n] = s;
  }
  for (int idx = t; idx < b * b; idx += HQ_NT) {
    int p = idx % b, q = idx / b;
    Ts[p][q] = (p <= q) ? a.T[p + (long)q * a.ldt] : 0.0;
  }
  __syncthreads();
  for (int idx = t; idx < b * kTile; idx += HQ_NT) {
    int n = idx & (kTile - 1), q = idx / kTile;
    long gn = n0 + n;
    if (gn >= a.nc) continue;
    double s = 0.0;
    if (a.tnotrans) {
      for (int p = q; p < b; ++p) s += Ts[q][p] * Ms[p][n];  // (T M)(q, n)
    } else {
      for (int p = 0; p <= q; ++p) s += Ts[p][q] * Ms[p][n];  // (T' M)(q, n)
    }
    a.W[gn + (long)q * a.ldw] = s;  // W' (nc x b), ld ldw
  }
}

// Y(p, n) = sum over the partial slices of column n, in slice order (the
// product V'C alone, for callers without T).
__global__ __launch_bounds__(HQ_NT) void trail_sum_kernel(TrailArgs a,
                                                          double* Y, long ldy) {
  const long idx = (long)blockIdx.x * HQ_NT + threadIdx.x;
  const int b = a.b;
  if (idx >= (long)b * a.nc) return;
  const long p = idx % b, n = idx / b;
  const int S = a.sk ? vtc_part(a).count(n / 64) : a.splits;
  const long slice = (long)b * a.ncv;
  double s = 0.0;
  for (int q = 0; q < S; ++q) s += a.part[q * slice + idx];
  Y[p + n * ldy] = s;
}

}  // namespace

// The block tail of the HQRRP driver, per 32 virtual columns (two
// workgroups for each 64-column group of the partials, so that their loads
// spread over twice as many CUs), after the split-K partials of M = V'[C C2]:
//   W   = T' sum_s part[s]       (written as W' for the trailing update)
//   R12 = C(0:b, :) - V1 W        (V1: the top b x b block of V, unit lower;
//                                  R12 is row block j of R, written to R)
// and for C2 = G' (the sketch operator, transposed) the same two lines give
// Q' = T' P' and B' = G'(0:b, :) - V1 Q', written over G'(0:b, :): the
// reference's B = G1 - (G1 U11 + G2 U21) T U11' (NoFLA_HQRRP_WY_blk_var4.c:
// 646-686).  A second launch (trail_y2_kernel) then applies
// Y2 -= B R12 (691-693).  Each product is 64 x 64 x 32 on the matrix cores
// with both operands in LDS (67 KB: two workgroups per CU).  Eight
// wavefronts per workgroup, each one 16 x 16 tile of each product; the
// slices are added in slice order and every tile's products in the same
// order as for any other split, so the results do not depend on it.
namespace {
constexpr int kTP = 65;  // LDS pitch of the 64-wide tiles
constexpr int kTB = 81;  // LDS pitch of B' (r <= 80 rows)
constexpr int kTT = 512;  // threads of a tail workgroup
constexpr int kTH = 33;   // LDS pitch of the 32-wide tiles
__global__ __launch_bounds__(kTT) void trail_tail_kernel(TailArgs a) {
  // T and V1 share one array: T(p, q) on and above the diagonal, V1(i, p)
  // below it (its unit diagonal is implied where it is read).
  __shared__ double TVs[64][kTP];
  __shared__ double Ms[64][kTH];  // M(p, c)
  __shared__ double Ws[64][kTH];  // W(q, c)
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int l15 = lane & 15, l4 = lane >> 4;
  const int b = a.b;
  // Two workgroups per 64 virtual columns (one group of the partials), each
  // 32 of them.
  const long v0 = (long)(blockIdx.x >> 1) * 64;  // the group's first
  const int hc = 32 * (blockIdx.x & 1);          // this half's offset
  long c0, nc, ldc, ldr;
  const double* Cb;
  double* Rb;
  if (v0 < a.nc) {
    c0 = v0;
    nc = a.nc;
    Cb = a.C;
    ldc = a.ldc;
    Rb = a.R;
    ldr = a.ldr;
  } else if (a.C2 && v0 >= a.nca && v0 < a.nca + a.nc2) {
    c0 = v0 - a.nca;
    nc = a.nc2;
    Cb = Rb = a.C2;
    ldc = ldr = a.ldc2;
  } else {
    return;
  }
  const long h0 = c0 + hc;  // first column of this workgroup
  const long slice = (long)b * a.ncv;
  const int S = a.sk ? a.vp.count(a.vp.group(v0)) : a.S;  // partial slices
  // Wavefront w: columns 16 (w % 2) .. + 15 of the half and row tile w / 2
  // of both products.
  const int cw = (wave & 1) * 16 + l15;  // this lane's column of the half
  const int qt = wave >> 1;
  const bool cok = h0 + cw < nc;
  const long gc = h0 + cw;
  // Every load up front (T, V1, the top rows for R12, then the partial
  // slices summed in slice order, four slices' loads in flight at a time),
  // then the LDS stores.  An interior half with all 64 reflectors (fast)
  // issues every load without a predicate.
  const bool fast = b == 64 && h0 + 32 <= nc;
  constexpr int kPer = 64 * 32 / kTT;  // partial entries per thread
  constexpr int kPtv = 64 * 64 / kTT;  // T and V1 entries per thread
  hq_d4 top;
  double msum[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) msum[u] = 0.0;
  const long off0 = (long)(t & 63) + (v0 + hc + (t >> 6)) * b;  // u = 0
  if (fast) {
    double tv[kPtv], vv[kPtv];
#pragma unroll
    for (int u = 0; u < kPtv; ++u) {
      const int idx = t + u * kTT, p = idx & 63, c = idx >> 6;
      tv[u] = a.T[p + (long)c * a.ldt];
      vv[u] = a.V[p + (long)c * a.ldv];
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) top[e] = Cb[qt * 16 + l4 + 4 * e + gc * ldc];
    const int S4 = S & ~3;
    for (int q0 = 0; q0 < S4; q0 += 4) {
      double x[4][kPer];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double* src = a.part + (q0 + k) * slice + off0;
#pragma unroll
        for (int u = 0; u < kPer; ++u) x[k][u] = src[(long)u * (kTT / 64) * b];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int u = 0; u < kPer; ++u) msum[u] += x[k][u];
    }
    for (int q = S4; q < S; ++q) {
      const double* src = a.part + q * slice + off0;
      double x[kPer];
#pragma unroll
      for (int u = 0; u < kPer; ++u) x[u] = src[(long)u * (kTT / 64) * b];
#pragma unroll
      for (int u = 0; u < kPer; ++u) msum[u] += x[u];
    }
#pragma unroll
    for (int u = 0; u < kPtv; ++u) {
      const int idx = t + u * kTT, p = idx & 63, c = idx >> 6;
      TVs[p][c] = (p <= c) ? tv[u] : vv[u];
    }
  } else {
#pragma unroll
    for (int u = 0; u < kPtv; ++u) {
      const int idx = t + u * kTT, p = idx & 63, c = idx >> 6;
      TVs[p][c] = (p <= c) ? ((c < b) ? a.T[p + (long)c * a.ldt] : 0.0)
                           : ((p < b && c < b) ? a.V[p + (long)c * a.ldv] : 0.0);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = qt * 16 + l4 + 4 * e;
      top[e] = (cok && i < b) ? Cb[i + gc * ldc] : 0.0;
    }
    // The slices four at a time: 4 x kPer independent loads in flight, then
    // added in slice order.
    const bool pok = (t & 63) < b;
    for (int q0 = 0; q0 < S; q0 += 4) {
      double x[4][kPer];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const bool sok = q0 + k < S;
        const double* src = a.part + (q0 + k) * slice + off0;
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
          const int c = (t >> 6) + u * (kTT / 64);
          x[k][u] = (sok && pok && h0 + c < nc)
                        ? src[(long)u * (kTT / 64) * b]
                        : 0.0;
        }
      }
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int u = 0; u < kPer; ++u) msum[u] += x[k][u];
    }
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int idx = t + u * kTT;
    Ms[idx & 63][idx >> 6] = msum[u];
  }
  __syncthreads();
  const int qr = qt * 16 + l15;
  // W = T' M: A = T' (q x p) from T(p, q), B = M.
  {
    hq_d4 acc = hq_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < 64; kk += 4)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(
          (kk + l4 <= qr) ? TVs[kk + l4][qr] : 0.0, Ms[kk + l4][cw], acc, 0,
          0, 0);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int qq = qt * 16 + l4 + 4 * e;
      Ws[qq][cw] = acc[e];
      if (cok && qq < b) a.W[v0 + hc + cw + (long)qq * a.ldw] = acc[e];
    }
  }
  __syncthreads();  // W complete
  // R12 (or B') = top rows - V1 W: A = -V1 (i x p), B = W; the top rows
  // straight into the accumulators.
  {
    hq_d4 acc = top;
#pragma unroll
    for (int kk = 0; kk < 64; kk += 4)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(
          -vmask(TVs[qr][kk + l4], qr, kk + l4), Ws[kk + l4][cw], acc, 0, 0,
          0);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = qt * 16 + l4 + 4 * e;
      if (cok && i < b) Rb[i + gc * ldr] = acc[e];
    }
  }
}

// Y2(:, c) -= B R12(:, c) on 64-column tiles of C, B = (C2(0:b, :))' (r x b,
// r = nc2 <= 80) and R12 = R(0:b, :); an FP32 Y2 (the mixed sketch) is read
// into and computed in FP64 and stored rounded.  Wavefront w owns columns
// 16w..16w+15 and every 16-row block of r; Y2 goes straight into the
// accumulators.
__global__ __launch_bounds__(HQ_NT) void trail_y2_kernel(TailArgs a) {
  __shared__ double As[64][kTB];  // As[p][i] = -B(i, p) = -C2(p, i)
  __shared__ double Bs[64][kTP];  // Bs[p][c] = R12(p, c0 + c)
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int l15 = lane & 15, l4 = lane >> 4;
  const int b = a.b, r = a.nc2;
  const long c0 = (long)blockIdx.x * 64;
  // Every load in flight together (B', R12, the Y2 tile), then the LDS
  // stores.  An interior tile with all 64 reflectors (fast) loads without
  // predicates: rows i >= r of B' and Y2 only feed accumulator rows that are
  // never stored, so they read row r - 1 instead of being zeroed.
  constexpr int kPa = 64 * 80 / HQ_NT, kPb = 64 * 64 / HQ_NT;
  double va[kPa], vb[kPb];
  const int cw = wave * 16 + l15;
  const long gc = c0 + cw;
  const bool cok = gc < a.nc;
  hq_d4 acc[5];
  if (b == 64 && c0 + 64 <= a.nc) {
#pragma unroll
    for (int u = 0; u < kPa; ++u) {
      const int idx = t + u * HQ_NT, p = idx & 63, i = idx >> 6;
      va[u] = -a.C2[p + (long)min(i, r - 1) * a.ldc2];
    }
#pragma unroll
    for (int u = 0; u < kPb; ++u) {
      const int idx = t + u * HQ_NT, p = idx & 63, c = idx >> 6;
      vb[u] = a.R[p + (c0 + c) * a.ldr];
    }
    if (a.Y2f) {
#pragma unroll
      for (int q = 0; q < 5; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          acc[q][e] = (double)a.Y2f[min(q * 16 + l4 + 4 * e, r - 1) + gc * a.ldy];
    } else {
#pragma unroll
      for (int q = 0; q < 5; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          acc[q][e] = a.Y2[min(q * 16 + l4 + 4 * e, r - 1) + gc * a.ldy];
    }
  } else {
#pragma unroll
    for (int u = 0; u < kPa; ++u) {
      const int idx = t + u * HQ_NT, p = idx & 63, i = idx >> 6;
      va[u] = (p < b && i < r) ? -a.C2[p + (long)i * a.ldc2] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kPb; ++u) {
      const int idx = t + u * HQ_NT, p = idx & 63, c = idx >> 6;
      vb[u] = (p < b && c0 + c < a.nc) ? a.R[p + (c0 + c) * a.ldr] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 5; ++q)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = q * 16 + l4 + 4 * e;
        const long at = i + gc * a.ldy;
        acc[q][e] = (cok && i < r) ? (a.Y2f ? (double)a.Y2f[at] : a.Y2[at]) : 0.0;
      }
  }
#pragma unroll
  for (int u = 0; u < kPa; ++u) {
    const int idx = t + u * HQ_NT;
    As[idx & 63][idx >> 6] = va[u];
  }
#pragma unroll
  for (int u = 0; u < kPb; ++u) {
    const int idx = t + u * HQ_NT;
    Bs[idx & 63][idx >> 6] = vb[u];
  }
  __syncthreads();
#pragma unroll
  for (int kk = 0; kk < 64; kk += 4) {
    const double bf = Bs[kk + l4][cw];
#pragma unroll
    for (int q = 0; q < 5; ++q)
      acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(As[kk + l4][q * 16 + l15],
                                                    bf, acc[q], 0, 0, 0);
  }
#pragma unroll
  for (int q = 0; q < 5; ++q)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = q * 16 + l4 + 4 * e;
      const long at = i + gc * a.ldy;
      if (cok && i < r) {
        if (a.Y2f)
          a.Y2f[at] = (float)acc[q][e];
        else
          a.Y2[at] = acc[q][e];
      }
    }
}
}  // namespace

hipError_t launch_trail_tail(const TailArgs& a, hipStream_t s) {
  if (a.b < 1 || a.b > 64 || a.nc < 1 || a.S < 1 || (a.C2 && a.nc2 > 80))
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(trail_tail_kernel, dim3(2 * hq_cdiv(a.ncv, 64)),
                     dim3(kTT), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_trail_y2(const TailArgs& a, hipStream_t s) {
  if (a.b < 1 || a.b > 64 || a.nc < 1 || !a.C2 || a.nc2 < 1 || a.nc2 > 80)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(trail_y2_kernel, dim3(hq_cdiv(a.nc, 64)), dim3(HQ_NT), 0,
                     s, a);
  return hipGetLastError();
}

hipError_t launch_trail_vtc(const TrailArgs& a, hipStream_t s) {
  if (a.sk) {
    if (a.sk_waves < 1 || a.sk_waves % (HQ_NT / 64) || a.b > 64)
      return hipErrorInvalidValue;
    dim3 grid(a.sk_waves / (HQ_NT / 64));
    // 32-byte groups of rows 16c + 4g are 32-byte aligned when both bases
    // and leading dimensions are.
    const bool vec = (reinterpret_cast<uintptr_t>(a.C) % 32 == 0) &&
                     (reinterpret_cast<uintptr_t>(a.V) % 32 == 0) &&
                     a.ldc % 4 == 0 && a.ldv % 4 == 0 &&
                     (!a.C2 || (reinterpret_cast<uintptr_t>(a.C2) % 32 == 0 &&
                                a.ldc2 % 4 == 0));
    if (vec)
      hipLaunchKernelGGL(trail_vtc_kernel<true>, grid, dim3(HQ_NT), 0, s, a);
    else
      hipLaunchKernelGGL(trail_vtc_kernel<false>, grid, dim3(HQ_NT), 0, s, a);
  } else {
    dim3 grid(hq_cdiv(a.ncv, kTile), a.splits);
    hipLaunchKernelGGL(trail_vtc64_kernel, grid, dim3(HQ_NT), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_trail_sum(const TrailArgs& a, double* Y, long ldy,
                            hipStream_t s) {
  if (a.C2) return hipErrorInvalidValue;
  hipLaunchKernelGGL(trail_sum_kernel, dim3(hq_cdiv((long)a.b * a.nc, HQ_NT)),
                     dim3(HQ_NT), 0, s, a, Y, ldy);
  return hipGetLastError();
}

hipError_t launch_trail_tw(const TrailArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(trail_tw_kernel, dim3(hq_cdiv(a.nc, kTile)), dim3(HQ_NT),
                     0, s, a);
  return hipGetLastError();
}

hipError_t launch_trail_update(const TrailArgs& a, hipStream_t s) {
  // The persistent band kernel when every one of a.ncu workgroups (one per
  // CU the stream may use) gets at least 16 tiles (below that its per-band
  // staging costs more than it saves); its buffer resources need leading
  // dimensions of at least 256 and tile extents below 2^30 bytes.
  const long ntiles = hq_cdiv(a.mrows, kUR) *
                      (a.nc / kUC + (a.C2 ? a.nc2 / kUC : 0));
  const long ldmax = std::max(std::max(a.ldc, a.Cout ? a.ldco : a.ldc),
                              a.C2 ? a.ldc2 : a.ldc);
  const long ldmin = std::min(std::min(a.ldc, a.Cout ? a.ldco : a.ldc),
                              a.C2 ? a.ldc2 : a.ldc);
  if (a.ncu > 0 && ntiles >= 16L * a.ncu && ldmin >= kUR &&
      ldmax * 8 * kUC < (1L << 30)) {
    const long nedge = hq_cdiv(a.mrows, kUR) *
                       ((a.nc % kUC ? 1 : 0) + (a.C2 && a.nc2 % kUC ? 1 : 0));
    if (nedge > 0)
      hipLaunchKernelGGL(trail_edge_kernel, dim3(nedge), dim3(HQ_NT), 0, s, a);
    hipLaunchKernelGGL(trail_band_kernel, dim3(a.ncu), dim3(kBT), 0, s, a);
    return hipGetLastError();
  }
  dim3 grid(hq_cdiv(a.mrows, kUR), hq_cdiv(a.C2 ? a.ncv : a.nc, kUC));
  hipLaunchKernelGGL(trail_update_kernel, grid, dim3(HQ_NT), 0, s, a);
  return hipGetLastError();
}

int trail_kbk() { return kBK64; }  // multiple of kBK too

