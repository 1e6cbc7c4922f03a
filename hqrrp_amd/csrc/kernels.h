// kernels.h -- argument blocks and launchers of the device kernels
// (host-side declarations shared by the driver and the kernel files).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

struct SketchArgs {
  const void* Y;  // sketch columns to pivot: r x ncols, ld ldy (elements)
  long ldy;
  int f32;        // Y holds floats (the pivoting then runs in float)
  int r, ncols, stages, nw;
  int cols;  // columns per workgroup: sketch_cols_per_wg(ncols)
  int* piv;    // stages: position swapped with position k at step k
  int* moves;  // [0] = count, then (dst, src) position pairs
  // Exchange buffers (common.h): xcur holds only the sentinel on entry
  // (sketch_xchg_doubles(nw, rmax, stages) of it are used); the sentinel is
  // restored over xprev[0, xprev_len).
  double* xcur;
  double* xprev;
  long xprev_len;
  int* err;
};

struct PanelArgs {
  double* P;  // panel: mrows x b, ld ldp (factored in place)
  long ldp;
  int mrows, b, pivoting;
  int nw;          // workgroups (panel_rows_per_wg() rows each)
  double* tau;     // b
  int* perm;       // b: original column at each position (0-based)
  double* gram;    // b x b, strictly upper part: v_l' v_k
  // Exchange buffers as in SketchArgs; panel_xchg_doubles(nw, b) used.
  double* xcur;
  double* xprev;
  long xprev_len;
  int* err;
  // Look-ahead (sel_C set): the panel's columns are formed on load as
  // P(m, q) = C(sel_r0 + m, src(q)) - V(m, :) W'(src(q), :)' (select_tile.h;
  // src from the composed moves, the identity when sel_moves is null)
  // instead of read from P; P then only receives the factored panel.
  const int* sel_moves;
  int sel_b2, sel_r0, sel_b;
  const double* sel_C;
  long sel_ldc;
  const double* sel_V;
  long sel_ldv;
  const double* sel_W;
  long sel_ldw;
  // Run only when null or *run_if != 0 (the fallback of the Cholesky-QR
  // panel); otherwise the launch only restores its exchange sentinel.
  const int* run_if;
  // Non-null: workgroup 0 also forms T (b x b, ld ldt) from gram and tau at
  // the end, as larft would (null: T is left to a larft launch).
  double* T;
  long ldt;
};

// k_cqr.hip: panel QRCP from two Gram passes (pivoted Cholesky QR with the
// reference's pivot rule, a second Cholesky QR pass) and Householder
// reconstruction, b <= 64, one persistent launch; the per-step panel kernel
// takes the panels it flags.
struct CqrArgs {
  double* P;  // panel: mrows x b, ld ldp (factored in place)
  long ldp;
  int mrows, b, pivoting;
  int nw;          // workgroups: cqr_rows_per_wg() rows each, co-resident
  double* tau;     // b
  int* perm;       // b: original column at each position (0-based)
  double* T;       // b x b, ld ldt
  long ldt;
  double* R1;      // workspace: 64 x 64
  int* flag;       // set to 1 when the per-step kernel must take the panel
  // Exchange buffers as in SketchArgs: cqr_xchg_doubles(nw) used, of which
  // the kernel restores the sentinel itself over all but the first
  // cqr_xchg_dirty().  xprev_if: when set, xprev is restored only if
  // *xprev_if != 0 (the previous launch was the gated per-step kernel after
  // a Cholesky-QR panel, which wrote its buffer only if that flag was set).
  double* xcur;
  double* xprev;
  long xprev_len;
  const int* xprev_if;
  int* err;
  // Look-ahead column selection as in PanelArgs (sel_C set); the formed
  // panel is stored into P.
  const int* sel_moves;
  int sel_b2, sel_r0, sel_b;
  const double* sel_C;
  long sel_ldc;
  const double* sel_V;
  long sel_ldv;
  const double* sel_W;
  long sel_ldw;
};
int cqr_rows_per_wg();
size_t cqr_xchg_doubles(int nw);
size_t cqr_xchg_dirty();
hipError_t launch_cqr_panel(const CqrArgs& a, hipStream_t s);

struct GemmArgs {
  int M, N, K;
  double alpha, beta;
  const double* A;
  long lda;
  const double* B;
  long ldb;
  double* C;
  long ldc;
  int maskA, maskB;  // read the stored operand as unit lower trapezoidal
  long kchunk;       // K range of one split (blockIdx.z)
  double* part;      // split-K output slices (M x N each), or null
};

struct TrailArgs {
  int mrows, nc, b, splits;
  int vmask;        // 1: V's top b x b block is unit lower (masked)
  int tnotrans;     // trail_tw: W = T M (apply Q) instead of T' M (apply Q')
  const double* V;  // mrows x b reflectors (unit lower trapezoidal), ld ldv
  long ldv;
  double* C;  // mrows x nc trailing matrix, ld ldc
  double* Cout;  // trail_update's output (null: C in place), ld ldco
  long ldco;
  long ldc;
  double* W;  // W' = (T' V' C)': nc x b, ld ldw
  long ldw;
  const double* T;  // b x b upper triangular, ld ldt
  long ldt;
  double* part;  // splits slices of b x ncv
  long kchunk;   // rows of C per split
  // Optional second column range, updated with C (in place): virtual
  // columns [0, nc) are C's, [nca, nca + nc2) are C2's (mrows x nc2, ld
  // ldc2, rows aligned with C's; nca a multiple of 256 >= nc, the gap
  // empty).  ncv = nca + nc2 with C2, else nc: the column count of the
  // partials and of W'.
  double* C2;
  long ldc2;
  int nc2, nca, ncv;
  // trail_update: CUs the launch stream may use (0: the per-tile kernel
  // only); the persistent band kernel takes one workgroup per CU.
  int ncu;
  // V'C partition (sk = 1; else `splits` uniform K slices of kchunk rows):
  // see VtcPart.
  int sk, sk_waves;
  long sk_nch, sk_q, sk_r;
};

// The stream-K partition of V'[C C2] (trail_vtc with sk = 1).  A unit is
// (group g, chunk c): 64 virtual columns by kVtcChunk rows, numbered u = g *
// nch + c over C's ceil(nc / 64) groups, then C2's.  The units are dealt in
// consecutive runs to the wavefronts: wavefront w takes [start(w),
// start(w + 1)), q units each and one more for the first r.  A wavefront
// writes one partial per group it touches, into slice w - first(g) of the
// partials, so group g's sum is over count(g) slices, added in slice order.
constexpr int kVtcChunk = 16;
struct VtcPart {
  long nch, q, r;
  int nc, nca;
  __host__ __device__ long start(long w) const {
    return w * q + (w < r ? w : r);
  }
  __host__ __device__ long wave_of(long u) const {
    const long big = r * (q + 1);
    return u < big ? u / (q + 1) : r + (u - big) / q;
  }
  __host__ __device__ long first(long g) const { return wave_of(g * nch); }
  __host__ __device__ int count(long g) const {
    return (int)(wave_of((g + 1) * nch - 1) - wave_of(g * nch) + 1);
  }
  __host__ __device__ int ngc() const { return (nc + 63) / 64; }
  // First virtual column of group g; the group of virtual column v (a
  // multiple of 64 in C's or C2's range).
  __host__ __device__ long vcol0(long g) const {
    return g < ngc() ? 64 * g : nca + 64 * (g - ngc());
  }
  __host__ __device__ long group(long v) const {
    return v < nca ? v / 64 : ngc() + (v - nca) / 64;
  }
};
__host__ __device__ inline VtcPart vtc_part(const TrailArgs& a) {
  return VtcPart{a.sk_nch, a.sk_q, a.sk_r, a.nc, a.nca};
}

// k_downdate.hip: FP64 sketch downdate of one block (r <= 80, b <= 64).
struct DowndateArgs {
  int r, b, mU, nY2;
  const double* U;  // mU x b reflectors (unit lower trapezoidal), ld ldu
  long ldu;
  const double* T;  // b x b upper triangular, ld ldt
  long ldt;
  const double* R12;  // b x nY2, ld ldr
  long ldr;
  double* G;  // r x mU, ld ldg (updated)
  long ldg;
  double* Y2;  // r x nY2, ld ldy (updated)
  long ldy;
  double* part;  // S slices of r x b
  double* P;     // r x b
  double* Q;     // r x b
  int S;         // row slices of U (downdate_slices)
  long kc;       // rows per slice
  int ng;        // set by launch_downdate
};
// Slices and rows per slice for mU rows of U.
int downdate_slices(long mU, long* kc);
hipError_t launch_downdate(DowndateArgs a, hipStream_t s);

// k_batch.hip: batched HQRRP of small matrices (m, n <= HQB_DIM_MAX;
// nb <= 64; pp <= HQB_PP_MAX).
#define HQB_PP_MAX 10
#define HQB_DIM_MAX 512
struct BatchItem {
  int m, n, lda, nb;
  int idx;  // index into the pointer arrays
  int pad[3];
};
struct BatchArgs {
  const BatchItem* items;  // queue order (largest first)
  int count;
  int pp, pivoting;
  double* const* A;  // per-matrix device pointers, indexed by BatchItem::idx
  int* const* jpvt;
  double* const* tau;
  const double* omega;  // normals: Omega of r x m is omega[0, r*m), ld r
  double* work;         // batch_work_doubles() per workgroup
  int* counter;         // queue head, reset by launch_batch_hqrrp
  unsigned long long* prof;  // HQB_PHASES cycle sums, or null
};
// Phases of the batched kernel's optional cycle counts.
enum {
  HQB_PH_IDLE = 0,  // queue, setup, waiting for the last matrices
  HQB_PH_SKETCH_GEMM,
  HQB_PH_SKETCH_QRCP,
  HQB_PH_PANEL,
  HQB_PH_T,          // T and V T'
  HQB_PH_TRAIL,      // trailing update of A
  HQB_PH_DOWNDATE,   // G update and Y2
  HQB_PH_FINISH,     // column permutation of A and jpvt
  HQB_PHASES
};
long batch_work_doubles();
// Kernel variant (16, 32 or 64) serving block size nb; its workgroups per CU.
int batch_class_nb(int nb);
int batch_wgs_per_cu(int class_nb);
// Zeroes n queue counters; launches the variant's persistent kernel on the
// items of one class (a.items .. a.items + a.count, all of that class).
hipError_t launch_batch_reset(int* counters, int n, hipStream_t s);
hipError_t launch_batch_fill(double* p, long n, unsigned seed, hipStream_t s);
hipError_t launch_batch_hqrrp(const BatchArgs& a, int class_nb, int wgs,
                              hipStream_t s);

// k_trail.hip
// Block tail of the HQRRP driver (see trail_tail_kernel).
struct TailArgs {
  int b, nc, S;
  int nc2, nca, ncv;   // C2's columns, as in TrailArgs
  const double* part;  // S slices of b x ncv (trail_vtc's partials)
  int sk;              // the partials follow the stream-K partition vp
  VtcPart vp;
  const double* T;     // b x b upper triangular, ld ldt
  long ldt;
  const double* V;  // the block's reflectors (top b x b read unit lower)
  long ldv;
  const double* C;  // rows 0..b-1 of the trailing matrix, ld ldc
  long ldc;
  double* R;  // R12 out (b x nc), ld ldr (may equal C)
  long ldr;
  double* C2;  // rows 0..b-1 of the second range (updated in place), ld ldc2
  long ldc2;
  double* W;  // W' out (ncv x b), ld ldw
  long ldw;
  double* Y2;  // r x nc, ld ldy: Y2 -= C2(0:b, :)' R12 (r = nc2)
  float* Y2f;  // the same for an FP32 sketch (computed in FP64, stored
               // rounded); exactly one of Y2, Y2f is set
  long ldy;
};
// W', R12 (and C2's top rows); then, with Y2, the sketch downdate.
hipError_t launch_trail_tail(const TailArgs& a, hipStream_t s);
hipError_t launch_trail_y2(const TailArgs& a, hipStream_t s);
hipError_t launch_trail_vtc(const TrailArgs& a, hipStream_t s);
hipError_t launch_trail_tw(const TrailArgs& a, hipStream_t s);
// Y(p, n) = sum of the partials of V'C (b x nc, no C2), ld ldy.
hipError_t launch_trail_sum(const TrailArgs& a, double* Y, long ldy,
                            hipStream_t s);
hipError_t launch_trail_update(const TrailArgs& a, hipStream_t s);
int trail_kbk();

// k_sketch.hip
size_t sketch_xchg_doubles(int nw, int rmax, int stages);
int sketch_rmax(int r);
int sketch_cols_per_wg(int ncols);  // 128 or 256
int sketch_max_stages();
hipError_t launch_sketch_qrcp(const SketchArgs& a, int rmax, hipStream_t s);

// k_panel.hip
hipError_t launch_panel_qrcp(const PanelArgs& a, hipStream_t s);
int panel_rows_per_wg();
size_t panel_xchg_doubles(int nw, int b);
// T (b x b, ld ldt) from the reflector inner products gram (strictly upper
// part, ld ldg) and tau; b <= 64.  With run_if, only when *run_if != 0.
hipError_t launch_larft(int b, const double* gram, long ldg, const double* tau,
                        double* T, long ldt, hipStream_t s,
                        const int* run_if = nullptr);
// Rows [0, nrows) of the b panel columns follow the panel's pivots (and
// jpvt); with P, the same launch also copies the factored panel (prows x b,
// ld ldp) into rows nrows.. of those columns.
hipError_t launch_panel_permute(double* A, long lda, int nrows, int b,
                                const int* perm, int* jpvt, hipStream_t s,
                                const double* P = nullptr, long ldp = 0,
                                int prows = 0);

// k_gemm.hip
hipError_t launch_gemm(const GemmArgs& g, bool ta, bool tb, int splits,
                       hipStream_t s, int tag = 0);
hipError_t launch_reduce_splitk(int M, int N, int S, const double* part,
                                double beta, double* C, long ldc,
                                hipStream_t s);

// k_misc.hip
hipError_t launch_fill_sentinel(double* p, size_t n, hipStream_t s);
// dst (cols x rows, ld ldd) = src' (src: rows x cols, ld lds).
hipError_t launch_transpose(const double* src, long lds, int rows, int cols,
                            double* dst, long ldd, hipStream_t s);
// Exact power-of-two prescaling of an m x n matrix whose largest entry is
// extreme (exponent outside [lo, hi]): pre records max |A| and s in w (6
// words of device memory) and multiplies A by s (also when A holds NaN or
// Inf, which quiets signalling NaNs); post multiplies R and the trailing
// matrix of a factorisation truncated at k by 1/s.  Products, quotients and
// square roots of power-of-two multiples are exact, so pivots, reflectors
// and tau are unchanged and R scales exactly.
hipError_t launch_scale_pre(double* A, long lda, int m, int n,
                            unsigned long long* w, int lo, int hi,
                            hipStream_t s);
hipError_t launch_scale_post(double* A, long lda, int m, int n, int k,
                             const unsigned long long* w, hipStream_t s);
// The libflame UT-transform form of a factorisation's first k reflectors
// (LAPACK's H = I - tau v v' becomes H = I - u u' / s with u = v): s =
// 1 / tau, and where LAPACK's dlarfg found nothing below the diagonal
// (tau = 0, H = I) the UT Householder is the reflection I - 2 e e' (s =
// 1/2), so row k of R (columns k..n-1) changes sign.  In place on A and
// on tau (which becomes s).
hipError_t launch_ut_convert(double* A, long lda, int m, int n, int k,
                             double* tau, hipStream_t s);
// p[i] = i for i < n.
hipError_t launch_iota(int* p, int n, hipStream_t s);
// The moves of a permutation (perm[pos] = the position whose column ends at
// pos): moves[0] = count, then (dst = pos, src = perm[pos]) for every pos
// with perm[pos] != pos, in increasing pos (one workgroup).
hipError_t launch_perm_moves(const int* perm, int n, int* moves,
                             hipStream_t s);
// dst (rows x cols, ld ldd, FP64) = src (FP32, ld lds).
hipError_t launch_widen_block(double* dst, long ldd, const float* src,
                              long lds, int rows, int cols, hipStream_t s);
hipError_t launch_copy_block(double* dst, long ldd, const double* src,
                             long lds, int rows, int cols, hipStream_t s);
// Moves m rows of A's columns (A may point at any row), r rows of Y's
// columns (r = 0: none; Y's elements are ysz = 4 or 8 bytes, ldy in
// elements) and jpvt (null: none).
hipError_t launch_gather_cols(const int* moves, int max_moves, double* A,
                              long lda, int m, void* Y, long ldy, int r,
                              int ysz, int* jpvt, double* tmpA, double* tmpY,
                              int* tmpJ, hipStream_t s);

// k_adj.hip
hipError_t launch_fill_eye(double* Q, long ldq, int m, int n, double diag,
                           hipStream_t s);
hipError_t launch_copy_r(const double* A, long lda, int k, int n, double* R,
                         long ldr, hipStream_t s);
hipError_t launch_tri_inv_diag(int k, const double* R, long ldr, double* X,
                               long ldx, hipStream_t s);
int tri_inv_block();
hipError_t launch_copyltu(int k, const double* M, long ldm, double* B, long ldb,
                          hipStream_t s);
// scatter = 0: dst(:, c) = src(:, p[c]-1); 1: dst(:, p[c]-1) = src(:, c).
hipError_t launch_permute_cols(const double* src, long lds, int m, int ncols,
                               const int* p, int scatter, double* dst,
                               long ldd, hipStream_t s);

// k_classic.hip: classic greedy QRCP (LAPACK dgeqpf / dgeqp3), one step per
// launch_classic_step; vn1/vn2: partial and reference norms (n each).
hipError_t launch_classic_norms(const double* A, long lda, int m, int n,
                                int k0, double* vn1, double* vn2,
                                hipStream_t s);
hipError_t launch_classic_step(double* A, long lda, int m, int n, int i,
                               int* jpvt, double* tau, double* vn1,
                               double* vn2, hipStream_t s);
// One step of NoFLA_QRPmod_WY_unb_var4 (NoFLA_HQRRP_WY_blk_var4.c:820-872)
// on A (m x n, ld lda) with the same two kernels: the first position of the
// largest vn1 (pivoting = 1; else position i), a whole-column swap (perm,
// vn1, vn2 follow; piv[i] = the chosen position when piv is set), dlarfg,
// dlarf on the columns to the right and the reference's LAWN 176 downdate.
hipError_t launch_qrpmod_step(double* A, long lda, int m, int n, int i,
                              int pivoting, int* perm, int* piv, double* tau,
                              double* vn1, double* vn2, hipStream_t s);
// Small shapes: one launch per step, X -> Y ping-pong (m x n, ld ldx each;
// jpvt jX -> jY; norms vX -> vY with vn1 at [0, n), vn2 at [n, 2n)).
constexpr int kClassicMaxRows = 1024;
constexpr int kClassicMaxCols = 2048;
int classic_fused_ok(int m, int n);
hipError_t launch_classic_fused(const double* X, double* Y, long ldx, int m,
                                int n, int i, const int* jX, int* jY,
                                const double* vX, double* vY, double* A,
                                long lda, int* jpvt, double* tau,
                                hipStream_t s);

// k_mixed.hip: the mixed-precision sketch Y (r x n, FP32, ld ldy) =
// G'' A with G' (m x r, ld ldg) and A (m x n, ld lda) in FP64, on the
// matrix cores (prec 1: FP32, 2: BF16 with A split in two terms).  With
// part set, split s of K (rows [s kchunk, (s+1) kchunk), kchunk a multiple
// of mixed_sketch_kbk()) writes slice s of part (r x n each, ld r).
struct MixedSketchArgs {
  const double* G;
  long ldg;
  const double* A;
  long lda;
  int m, n, r;
  long kchunk;
  float* part;
  float* Y;
  long ldy;
};
int mixed_sketch_rmax();
int mixed_sketch_cols_per_wg();
int mixed_sketch_kbk();
hipError_t launch_sketch_mixed(const MixedSketchArgs& a, int prec, int splits,
                               hipStream_t s);
// x[0, n) rounded to the nearest FP32 (prec 1) or bf16 (prec 2) value.
hipError_t launch_round_omega(double* x, long n, int prec, hipStream_t s);

// k_gemm32.hip
struct Gemm32Args {
  int M, N, K;
  float alpha, beta;
  const void* A;  // FP64 (a64 = 1) or FP32 elements
  long lda;
  int a64;
  const void* B;
  long ldb;
  int b64;
  float* C;
  long ldc;
  int maskA, maskB;  // read the stored operand as unit lower trapezoidal
  int bf16;          // round both operands to bfloat16 when staged
  long kchunk;       // K range of one split (blockIdx.z)
  float* part;       // split-K output slices (M x N each), or null
};
hipError_t launch_gemm32(const Gemm32Args& g, bool ta, bool tb, int splits,
                         hipStream_t s);
hipError_t launch_reduce32(int M, int N, int S, const float* part, float alpha,
                           float beta, float* C, long ldc, hipStream_t s);
hipError_t launch_cast_f32(const double* src, float* dst, long n,
                           hipStream_t s);
