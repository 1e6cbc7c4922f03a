/*
 * hqrrp_amd.h -- C ABI of the MI355X-native HQRRP library (libhqrrp_amd.so).
 *
 * Two groups of entry points:
 *
 *  1. Drop-in replacements for the reference's API (host pointers, synchronous,
 *     same signatures and argument meaning).  A program that links the
 *     reference's NoFLA_HQRRP_WY_blk_var4.o can link this library instead.
 *
 *  2. Device-pointer, stream-ordered entry points used by the drop-in entries,
 *     by the Python layer and by the tests.  All matrices are column-major
 *     FP64 in device memory; pivot vectors are 1-based int32 like LAPACK's.
 *
 * Every function returns (or stores in *info) 0 on success.  Device entries
 * return a negative hqrrp_status on failure; they never fall back to the CPU.
 */
#ifndef HQRRP_AMD_H
#define HQRRP_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
typedef enum {
  HQRRP_OK = 0,
  HQRRP_ERR_ARG = -1,      /* illegal argument                            */
  HQRRP_ERR_ALLOC = -2,    /* device or host allocation failed            */
  HQRRP_ERR_HIP = -3,      /* a HIP runtime call failed                   */
  HQRRP_ERR_KERNEL = -4,   /* a kernel reported failure (barrier timeout) */
  HQRRP_ERR_UNSUPPORTED = -5 /* parameters outside the implemented range  */
} hqrrp_status;

/* ------------------------------------------------- 1. drop-in entry points */

/* LAPACK dgeqp3-compatible QR with column pivoting (HQRRP for the free
 * columns).  Replaces the reference's
 *   void dgeqp4(int*, int*, double*, int*, int*, double*, double*, int*, int*)
 * (lapack_compatible_sources/NoFLA_HQRRP_WY_blk_var4.h:2-3,
 *  NoFLA_HQRRP_WY_blk_var4.c:127-300).  Host pointers; A is overwritten with
 * R and the Householder vectors (dgeqrf storage), jpvt/tau as in dgeqp3.
 * info: 0, -1 (m<0), -2 (n<0), -4 (lda<max(1,m)), -8 (lwork too small);
 * lwork == -1 is a workspace query answered in work[0].  A failure of the
 * device (no counterpart in LAPACK, whose negative codes name arguments)
 * is reported as info = HQRRP_INFO_DEVICE - status (> 0), e.g. 1002 for an
 * allocation failure; A, jpvt and tau are then unspecified. */
#define HQRRP_INFO_DEVICE 1000
void dgeqp4(int *m, int *n, double *A, int *lda, int *jpvt, double *tau,
            double *work, int *lwork, int *info);

/* Blocked HQRRP driver.  Replaces the reference's
 *   int NoFLA_HQRRP_WY_blk_var4(int m_A, int n_A, double *buff_A, int ldim_A,
 *        int *buff_jpvt, double *buff_tau, int nb_alg, int pp,
 *        int panel_pivoting)
 * (NoFLA_HQRRP_WY_blk_var4.h:5-7, NoFLA_HQRRP_WY_blk_var4.c:303-571).
 * Host pointers.  jpvt is permuted in place (the caller initialises it, as
 * with the reference).  Omega is the reference's stream (seed 12).  Returns
 * 0, or a negative hqrrp_status (the reference prints and continues). */
int NoFLA_HQRRP_WY_blk_var4(int m_A, int n_A, double *buff_A, int ldim_A,
                            int *buff_jpvt, double *buff_tau, int nb_alg,
                            int pp, int panel_pivoting);

/* libflame UT-variant front end.  Replaces the reference's
 *   int FLA_HQRRP_UT_blk_var2(FLA_Obj A, FLA_Obj p, FLA_Obj s, int nb_alg,
 *        int pp, int panel_pivoting)
 * (libflame_sources/FLA_HQRRP_UT_blk_var2.h:3-4, FLA_HQRRP_UT_blk_var2.c:
 * 124-401) with each FLA_Obj unpacked into its buffer and sizes: A (m_A x
 * n_A, column stride ldim_A), p (n_A pivots, permuted in place; the caller
 * initialises them), s (min(m_A, n_A) values).  Host pointers.  The same
 * factorisation as NoFLA_HQRRP_WY_blk_var4 with libflame's UT Householder
 * convention: H_k = I - u_k u_k' / s_k (u_k: 1 on the diagonal, the stored
 * vector below), s_k = 1 / tau_k; where nothing remains below the diagonal
 * (tau_k = 0 in LAPACK's convention, H_k = I) libflame reflects instead,
 * H_k = I - 2 e_k e_k', s_k = 1/2, with row k of R negated.  Returns 0 or a
 * negative hqrrp_status. */
int hqrrp_ut_blk_var2(int m_A, int n_A, double *buff_A, int ldim_A,
                      int *buff_p, double *buff_s, int nb_alg, int pp,
                      int panel_pivoting);

/* -------------------------------------------- 2. device-pointer interface */

typedef struct hqrrp_handle_s *hqrrp_handle_t;

/* Sketch precision: the sketch Y = Omega*A and its downdate only steer the
 * pivot choice.  FP64 reproduces the reference. */
typedef enum {
  HQRRP_SKETCH_FP64 = 0,
  HQRRP_SKETCH_FP32 = 1,
  HQRRP_SKETCH_BF16 = 2
} hqrrp_sketch_prec;

/* Where Omega comes from. */
typedef enum {
  HQRRP_OMEGA_REFERENCE = 0, /* the reference's srand(seed)/polar stream    */
  HQRRP_OMEGA_USER = 1       /* caller-supplied device matrix (r x m, ld r) */
} hqrrp_omega_kind;

/* Optional read-back of the driver's internal state, for tests of the
 * pieces the reference checks only in debug builds (its
 * CHECK_DOWNDATING_OF_Y block, NoFLA_HQRRP_WY_blk_var4.c:457-490) and of
 * the pivot decisions.  Every pointer is a device pointer or NULL (not
 * written); r = nb + pp. */
typedef struct {
  int *sketch_cols; /* min(m,n) ints: entry j+s = the column (jpvt value)
                       that block j's sketch chose at step s, before the
                       panel reorders the block                          */
  void *y0;         /* r x n, ld r: the sketch Y = Omega*A right after it is
                       formed (FP64, or FP32 for a mixed sketch)         */
  void *y;          /* r x n, ld r: the sketch when the factorisation
                       stops (columns kmax.. hold the downdated Y_R)     */
  double *g;        /* m x r, ld m: Omega' when it stops (rows kmax.. hold
                       the downdated G_R')                               */
} hqrrp_trace;

typedef struct {
  int nb;              /* block size nb_alg (reference default 64)           */
  int pp;              /* oversampling (reference default 10)                */
  int panel_pivoting;  /* 1: pivot inside each panel (reference default)     */
  int sketch_prec;     /* hqrrp_sketch_prec                                  */
  int omega_kind;      /* hqrrp_omega_kind                                   */
  unsigned seed;       /* seed of the reference stream (reference: 12)       */
  const double *omega; /* device pointer when omega_kind == HQRRP_OMEGA_USER */
  /* Truncation (no counterpart in the reference): stop after k = max_rank
   * columns (<= 0: min(m,n)).  Blocks of nb columns run as in the full
   * factorisation; the block that reaches k is cut to end at column k (its
   * sketch takes k - j pivots, its panel has k - j columns), except the
   * reference's final block (j + nb >= m or j + nb >= n), which is always
   * factored whole.  On exit columns 0..k-1 hold R's first k rows and the
   * reflectors, and A(k:m, k:n) holds the trailing matrix Q' A P with Q of
   * those k reflectors. */
  int max_rank;
  int lookahead;       /* 1: overlap the next block's panel with the trailing
                          update (results bitwise equal to 0)                */
  const hqrrp_trace *trace; /* NULL: no read-back                            */
} hqrrp_options;

/* Fills *opt with the reference defaults (64, 10, 1, FP64, seed 12). */
void hqrrp_default_options(hqrrp_options *opt);

/* Handle: owns device workspace, the grid-barrier words and a stream. */
int hqrrp_create(hqrrp_handle_t *h, int device);
int hqrrp_destroy(hqrrp_handle_t h);
/* Use `stream` (a hipStream_t) for all work of this handle; NULL is the
 * default (null) stream, as in hipBLAS/cuBLAS.  A new handle uses a stream
 * of its own; reset_stream returns to it. */
int hqrrp_set_stream(hqrrp_handle_t h, void *stream);
int hqrrp_reset_stream(hqrrp_handle_t h);
int hqrrp_synchronize(hqrrp_handle_t h);

/* Panel factorisation used by every block of a handle's factorisations (no
 * counterpart in the reference, whose panel is NoFLA_QRPmod_WY_unb_var4,
 * NoFLA_HQRRP_WY_blk_var4.c:520-524, 774-886): 0 (default) = two Gram
 * passes (pivoted Cholesky QR with the reference's pivot rule, a second
 * Cholesky QR pass) and Householder reconstruction, falling back on the
 * device to the per-step kernel for panels it cannot factor to working
 * accuracy; 1 = the per-step kernel (one grid-wide reduction per pivot) for
 * every panel.  Both give the reference's pivots up to near-ties and values
 * equal to rounding. */
int hqrrp_set_panel_alg(hqrrp_handle_t h, int alg);

/* Optional per-phase timing (no counterpart in the reference, which only has
 * compile-time wall-clock timers in its libflame variant): HIP events are
 * recorded around the launches of each phase of the block loop, on the
 * stream those launches go to.  enable resets the counters; phases limits
 * the recording to the phases whose bit is set in mask (default: all; each
 * event pair costs a few microseconds of GPU time); read synchronises the
 * stream and returns, per phase, the summed event time (ms), the number of
 * kernel launches, and the algorithmic flops and bytes of those launches. */
int hqrrp_phase_count(void);
const char *hqrrp_phase_name(int phase);
int hqrrp_profile_enable(hqrrp_handle_t h, int on);
int hqrrp_profile_phases(hqrrp_handle_t h, unsigned mask);
int hqrrp_profile_read(hqrrp_handle_t h, int nphase, double *ms,
                       long long *launches, double *flops, double *bytes);

/* Device-pointer HQRRP, stream-ordered, no host synchronisation inside.
 * Same semantics as NoFLA_HQRRP_WY_blk_var4 on device arrays:
 * dA (m x n, ld lda), djpvt (n, permuted in place), dtau (min(m,n)). */
int hqrrp_factor_dev(hqrrp_handle_t h, int m, int n, double *dA, int lda,
                     int *djpvt, double *dtau, const hqrrp_options *opt);

/* Device form of hqrrp_ut_blk_var2 (the libflame UT-variant front end,
 * FLA_HQRRP_UT_blk_var2.c:124-401): hqrrp_factor_dev followed by the
 * rewrite into libflame's UT convention; ds (min(m,n)) receives s. */
int hqrrp_ut_factor_dev(hqrrp_handle_t h, int m, int n, double *dA, int lda,
                        int *dp, double *ds, const hqrrp_options *opt);

/* Host generation of the reference's Omega (r x m, column-major, ld r) with
 * the stream srand(seed) + polar method, without touching the caller's
 * rand() state.  (NoFLA_HQRRP_WY_blk_var4.c:380, 575-614) */
void hqrrp_reference_omega(int r, int m, double *G, unsigned seed);

/* The same generator started from a given cached second variate (*have,
 * *cached) and returning the one it leaves: the reference's generator keeps
 * it in function statics (NoFLA_HQRRP_WY_blk_var4.c:594-602) that srand()
 * does not reset, so a call drawing an odd number r*m of normals hands its
 * last pair's second value to the next call. */
void hqrrp_reference_omega_carry(int r, int m, double *G, unsigned seed,
                                 int *have, double *cached);

/* The library's process-wide copy of that state, read and advanced by every
 * factorisation that draws the reference Omega (HQRRP_OMEGA_REFERENCE:
 * hqrrp_factor_dev, NoFLA_HQRRP_WY_blk_var4, dgeqp4's HQRRP branch), as the
 * reference's statics are by its calls.  A new process starts with none
 * (have = 0); _set restores any state, e.g. that of a fresh process. */
void hqrrp_omega_carry_get(int *have, double *value);
void hqrrp_omega_carry_set(int have, double value);

/* ------------------------------- Q, fixed-pivot QR and its adjoint (device)
 * Device pointers, stream-ordered on the handle's stream. */

/* Householder QR without pivoting, LAPACK dgeqrf semantics: A (m x n, ld
 * lda) factored in place, tau (min(m,n)).  Replaces the dgeqrf_ call the
 * reference makes for dgeqp4's fixed columns (NoFLA_HQRRP_WY_blk_var4.c:
 * 221-225). */
int hqrrp_geqrf_dev(hqrrp_handle_t h, int m, int n, double *dA, int lda,
                    double *dtau);

/* Classic greedy column-pivoted QR, LAPACK dgeqp3 semantics (the first
 * column of largest partial norm at every step, LAWN 176 norm downdating
 * with recomputation): A (m x n, ld lda) factored in place, djpvt (n)
 * permuted in place like hqrrp_factor_dev's, dtau (min(m,n)).  Replaces the
 * dgeqpf_ / dgeqp3_ calls dgeqp4 makes for min(m,n) < 1000
 * (NoFLA_HQRRP_WY_blk_var4.c:189-199); dgeqp4 itself routes those sizes
 * here after moving its fixed columns to the front. */
int hqrrp_geqp3_dev(hqrrp_handle_t h, int m, int n, double *dA, int lda,
                    int *djpvt, double *dtau);

/* Allocation-free variant of hqrrp_geqp3_dev (no counterpart in the
 * reference): for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid it).  Same
 * arguments, kernels and results (bitwise) as hqrrp_geqp3_dev, but the
 * workspace is a static device array of HQRRP_NOALLOC_SCRATCH_BYTES per GPU
 * (held from the first use of the library on that GPU until the process
 * ends) instead of handle-owned memory; the call performs no allocation.
 * Shapes whose workspace exceeds it -- 16 (m n + 2 n) + 8 n bytes for
 * m <= 1024 and n <= 2048, else 16 n bytes -- return HQRRP_ERR_UNSUPPORTED
 * with a message on stderr.  Every call on one GPU shares that array: calls
 * that overlap on one GPU (other streams, other handles, other threads)
 * race; calls on one stream are ordered. */
#define HQRRP_NOALLOC_SCRATCH_BYTES ((size_t)34 << 20)
int hqrrp_geqp3_noalloc_dev(hqrrp_handle_t h, int m, int n, double *dA,
                            int lda, int *djpvt, double *dtau);

/* Q (m x n) = the first n columns of H_1 ... H_k, with the reflectors in the
 * first k columns of A below the diagonal and tau (LAPACK dorgqr semantics,
 * written out of place to Q); k <= n <= m.  Replaces the dorgqr_ call of
 * the reference's test driver (WYtest.c:89-93). */
int hqrrp_orgqr_dev(hqrrp_handle_t h, int m, int n, int k, const double *dA,
                    int lda, const double *dtau, double *dQ, int ldq);

/* op(Q) C (side 'L') or C op(Q) (side 'R'), op = transpose for trans 'T',
 * with Q = H_1 ... H_k from the reflectors in the first k columns of A below
 * the diagonal and tau (LAPACK dormqr semantics, in place on C: m x n, ld
 * ldc; A: nq x k, ld lda >= nq, nq = m for 'L', n for 'R').  Replaces the
 * dormqr_ call dgeqp4 makes to apply the fixed columns' reflectors
 * (NoFLA_HQRRP_WY_blk_var4.c:231-235) and the Q products callers form
 * after a factorisation (the test driver's dorgqr, WYtest.c:89-93). */
int hqrrp_ormqr_dev(hqrrp_handle_t h, char side, char trans, int m, int n,
                    int k, const double *dA, int lda, const double *dtau,
                    double *dC, int ldc);

/* QR at fixed pivots, truncated to k (no counterpart in the reference):
 * with AP = A(:, jpvt-1) (jpvt 1-based, n entries), AP(:, 0:k) = Q R11 and
 * R = [R11 R12] = the first k rows of the R factor of AP.  Q: m x k (ld
 * ldq), R: k x n (ld ldr, zero below the diagonal).  1 <= k <= min(m,n). */
int hqrrp_qr_fixed_dev(hqrrp_handle_t h, int m, int n, int k, const double *dA,
                       int lda, const int *djpvt, double *dQ, int ldq,
                       double *dR, int ldr);

/* Reverse-mode derivative of hqrrp_qr_fixed_dev with the pivots held fixed
 * (no counterpart in the reference): given A, jpvt, the forward's Q and R,
 * and the cotangents gQ (m x k), gR (k x n), writes gA (m x n):
 *   X = AP(:, :k), Y = AP(:, k:), gQ' = gQ + Y gR12',
 *   M = R11 gR11' - gQ''Q, gX = (gQ' + Q copyltu(M)) R11^-T, gY = Q gR12,
 *   gA(:, jpvt-1) = [gX gY]. */
int hqrrp_qr_fixed_adjoint_dev(hqrrp_handle_t h, int m, int n, int k,
                               const double *dA, int lda, const int *djpvt,
                               const double *dQ, int ldq, const double *dR,
                               int ldr, const double *gQ, int ldgq,
                               const double *gR, int ldgr, double *gA,
                               int ldga);

/* ------------------------------------------- batched small matrices */

/* Many independent small QRCPs in one persistent launch (no counterpart in
 * the reference, whose NoFLA_HQRRP_WY_blk_var4 factors one matrix per call,
 * NoFLA_HQRRP_WY_blk_var4.c:303-571).  Each matrix i (m[i] x n[i], ld
 * lda[i], m, n <= 512) is factored exactly as NoFLA_HQRRP_WY_blk_var4(m[i],
 * n[i], A_i, lda[i], jpvt_i, tau_i, nb[i], pp, panel_pivoting) would: Omega
 * from the reference's generator with seed 12, jpvt_i permuted in place
 * (callers pass 1..n[i] for the reference's usual call), tau_i of length
 * min(m[i], n[i]).
 *
 * A plan holds the shapes, the block sizes (nb == NULL: chosen per shape,
 * from the table of hqrrp_batch_tune when it has run on this handle), the
 * queue order and a workspace; nb[i] in 1..64, pp in 0..10.  Returns
 * HQRRP_ERR_UNSUPPORTED for shapes or block sizes outside that range. */
typedef struct hqrrp_batch_s *hqrrp_batch_t;
int hqrrp_batch_create(hqrrp_handle_t h, int count, const int *m, const int *n,
                       const int *lda, const int *nb, int pp,
                       int panel_pivoting, hqrrp_batch_t *plan);
/* The block size used for each matrix (count ints). */
int hqrrp_batch_get_nb(hqrrp_batch_t plan, int *nb);
/* Factors the plan's matrices: dA, djpvt, dtau are DEVICE arrays of count
 * device pointers (matrix i at dA[i], ...).  Stream-ordered on the handle's
 * stream, one launch, no allocation and no host synchronisation (so it can
 * be captured in a HIP graph).  Calls sharing a plan must not overlap. */
int hqrrp_batch_factor_dev(hqrrp_handle_t h, hqrrp_batch_t plan,
                           double *const *dA, int *const *djpvt,
                           double *const *dtau);
int hqrrp_batch_destroy(hqrrp_batch_t plan);
/* Optional cycle counts of the batched kernel's phases (no counterpart in
 * the reference): when on, every later factorisation adds, per phase, the
 * shader-clock cycles its workgroups spent there (summed over workgroups);
 * profile_read waits for the device and returns the sums. */
int hqrrp_batch_profile(hqrrp_batch_t plan, int on);
/* Autotuner of the batched block size (no counterpart in the reference):
 * times each of nb = 16, 32, 64 on a batch of the midpoint shape of every
 * 64 x 64 bucket of (m, n) and keeps the fastest per bucket in the handle;
 * later plans created with nb == NULL (same pp and panel_pivoting) use it.
 * Takes about a second and a few GB of scratch device memory. */
int hqrrp_batch_tune(hqrrp_handle_t h, int pp, int panel_pivoting);
/* The tuned table: 64 ints, table[8*bm + bn] for m in (64 bm, 64 bm + 64],
 * n in (64 bn, 64 bn + 64]; 0 where not tuned. */
int hqrrp_batch_tune_table(hqrrp_handle_t h, int *table);
int hqrrp_batch_profile_read(hqrrp_batch_t plan, int nphase, double *cycles);
int hqrrp_batch_phase_count(void);
const char *hqrrp_batch_phase_name(int phase);

/* ------------------------------------- component entry points (device) */

/* Greedy QRCP of the sketch for `stages` steps, choosing pivots only
 * (NoFLA_QRPmod_WY_unb_var4 with pivoting=1, num_stages=stages, on V_R,
 * NoFLA_HQRRP_WY_blk_var4.c:500-506 and 774-886).  dY: r x ncols (ld ldy) is
 * read, not modified.  dpiv[k] receives the position (0-based, relative) that
 * is swapped with position k at step k.  dmoves (1 + 4*stages ints)
 * receives the composed permutation: dmoves[0] = count (<= 2*stages), then
 * (dst, src) position pairs.  r <= 80, stages <= min(r, ncols). */
int hqrrp_sketch_qrcp_dev(hqrrp_handle_t h, int r, int ncols,
                          const double *dY, int ldy, int stages, int *dpiv,
                          int *dmoves);

/* Householder QR with optional column pivoting of an mrows x b panel
 * (b <= 64), in place, with LAPACK dgeqrf storage; outputs tau (b), the
 * position -> original-column permutation dperm (b, 0-based) and the
 * compact-WY factor T (b x b, ld ldt, zero below the diagonal)
 * (NoFLA_QRPmod_WY_unb_var4 with num_stages=-1 and build_T=1,
 * NoFLA_HQRRP_WY_blk_var4.c:520-524, 774-886, dlarft at 875-878). */
int hqrrp_panel_qrcp_dev(hqrrp_handle_t h, int mrows, int b, double *dP,
                         int ldp, int pivoting, double *dtau, int *dperm,
                         double *dT, int ldt);

/* Trailing update C := (I - V T V')' C, V unit lower trapezoidal
 * (mrows x b) and T upper triangular (dlarfb Left/Transpose/Forward/
 * Columnwise, NoFLA_HQRRP_WY_blk_var4.c:711-740). */
int hqrrp_apply_qt_left_dev(hqrrp_handle_t h, int mrows, int ncols, int b,
                            const double *dV, int ldv, const double *dT,
                            int ldt, double *dC, int ldc);

/* Sketch downdate (NoFLA_Downdate_Y, NoFLA_HQRRP_WY_blk_var4.c:617-708):
 * Y2 -= (G1 - (G1 U11 + G2 U21) T U11') R12 and [G1 G2] := [G1 G2](I-UTU').
 * dU: mU x b unit lower (mU = b + rows of U21), dR12: b x nY2 (ld ldr),
 * dY2: r x nY2 (ld ldy), dG: r x mU (ld ldg). */
int hqrrp_downdate_y_dev(hqrrp_handle_t h, int r, int b, int mU, int nY2,
                         const double *dU, int ldu, const double *dR12,
                         int ldr, const double *dT, int ldt, double *dY2,
                         int ldy, double *dG, int ldg);

/* General FP64 matrix product on the matrix cores:
 * C := alpha*op(A)*op(B) + beta*C, op = transpose when ta/tb != 0. */
int hqrrp_dgemm_dev(hqrrp_handle_t h, int ta, int tb, int M, int N, int K,
                    double alpha, const double *dA, int lda, const double *dB,
                    int ldb, double beta, double *dC, int ldc);

#ifdef __cplusplus
}
#endif
#endif
