/*
 * hqrrp_oracle.c -- CPU restatement of the reference HQRRP (WY variant).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product library (hqrrp_amd/) never links or calls it.
 *
 * It restates lapack_compatible_sources/NoFLA_HQRRP_WY_blk_var4.c of the
 * reference: it issues the same LAPACK calls, with the same arguments and in
 * the same order, so that against the same LAPACK library it reproduces the
 * reference's A, tau and jpvt bit for bit (checked against the compiled
 * reference in tests/test_oracle.py and against committed fixtures).
 *
 * LAPACK is reached through function pointers resolved with dlopen() so the
 * same object works with scipy's bundled OpenBLAS (symbols prefixed
 * "scipy_") or a plain LAPACK.  Character arguments are passed with the
 * gfortran hidden-length convention.
 *
 * Additions that do not change the arithmetic:
 *   - a reentrant generator for Omega (the reference uses srand(12)/rand());
 *   - an optional trace of the top-two partial norms at every pivot choice,
 *     used by tests to classify near-ties;
 *   - an optional block limit so the CPU baseline can time a bounded sample.
 *
 * Parity status of each function (what pins it):
 *   orc_rand_*         glibc rand() known answers after srand(12)   (pinned)
 *   orc_omega          reference Omega known answers                (pinned)
 *   orc_omega_carry    compiled reference, two calls in one process  (pinned)
 *   orc_qrpmod         compiled reference helper, bitwise           (pinned)
 *   orc_downdate_y     compiled reference helper, bitwise;          (pinned)
 *                      invariant Y_R = G_R * A_R
 *   orc_hqrrp          compiled reference driver, bitwise;          (pinned)
 *                      ||AP-QR||, ||Q'Q-I|| closed-form checks
 *   orc_dgeqp4         compiled reference front end, bitwise         (pinned)
 */
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hqrrp_oracle.h"

#define OMAX(a, b) ((a) > (b) ? (a) : (b))
#define OMIN(a, b) ((a) > (b) ? (b) : (a))

/* ------------------------------------------------------------------------- */
/* LAPACK / BLAS entry points (Fortran ABI, LP64, gfortran hidden lengths).   */

typedef size_t flen;
typedef void (*dgemm_t)(const char *, const char *, const int *, const int *,
                        const int *, const double *, const double *,
                        const int *, const double *, const int *,
                        const double *, double *, const int *, flen, flen);
typedef void (*dlacpy_t)(const char *, const int *, const int *,
                         const double *, const int *, double *, const int *,
                         flen);
typedef void (*dlarfg_t)(const int *, double *, double *, const int *,
                         double *);
typedef void (*dlarf_t)(const char *, const int *, const int *,
                        const double *, const int *, const double *, double *,
                        const int *, double *, flen);
typedef void (*dlarft_t)(const char *, const char *, const int *, const int *,
                         const double *, const int *, const double *,
                         double *, const int *, flen, flen);
typedef void (*dlarfb_t)(const char *, const char *, const char *,
                         const char *, const int *, const int *, const int *,
                         const double *, const int *, const double *,
                         const int *, double *, const int *, double *,
                         const int *, flen, flen, flen, flen);
typedef double (*dnrm2_t)(const int *, const double *, const int *);
typedef int (*idamax_t)(const int *, const double *, const int *);
typedef void (*dswap_t)(const int *, double *, const int *, double *,
                        const int *);
typedef void (*dtrmm_t)(const char *, const char *, const char *, const char *,
                        const int *, const int *, const double *,
                        const double *, const int *, double *, const int *,
                        flen, flen, flen, flen);
typedef double (*dlamch_t)(const char *, flen);
typedef void (*dgeqrf_t)(const int *, const int *, double *, const int *,
                         double *, double *, const int *, int *);
typedef void (*dormqr_t)(const char *, const char *, const int *, const int *,
                         const int *, const double *, const int *,
                         const double *, double *, const int *, double *,
                         const int *, int *, flen, flen);
typedef void (*dgeqpf_t)(const int *, const int *, double *, const int *,
                         int *, double *, double *, int *);
typedef void (*dgeqp3_t)(const int *, const int *, double *, const int *,
                         int *, double *, double *, const int *, int *);
typedef int (*ilaenv_t)(const int *, const char *, const char *, const int *,
                        const int *, const int *, const int *, flen, flen);
typedef void (*xerbla_t)(const char *, const int *, flen);
typedef void (*dorgqr_t)(const int *, const int *, const int *, double *,
                         const int *, const double *, double *, const int *,
                         int *);
typedef void (*setthreads_t)(int);
typedef void (*dtrsm_t)(const char *, const char *, const char *, const char *,
                        const int *, const int *, const double *,
                        const double *, const int *, double *, const int *,
                        flen, flen, flen, flen);
typedef void (*dgemv_t)(const char *, const int *, const int *,
                        const double *, const double *, const int *,
                        const double *, const int *, const double *, double *,
                        const int *, flen);
typedef void (*dger_t)(const int *, const int *, const double *,
                       const double *, const int *, const double *,
                       const int *, double *, const int *);
typedef void (*dcopy_t)(const int *, const double *, const int *, double *,
                        const int *);
typedef void (*daxpy_t)(const int *, const double *, const double *,
                        const int *, double *, const int *);
typedef void (*dscal_t)(const int *, const double *, double *, const int *);

static struct {
  void *h;
  char prefix[16];
  dgemm_t dgemm;
  dlacpy_t dlacpy;
  dlarfg_t dlarfg;
  dlarf_t dlarf;
  dlarft_t dlarft;
  dlarfb_t dlarfb;
  dnrm2_t dnrm2;
  idamax_t idamax;
  dswap_t dswap;
  dtrmm_t dtrmm;
  dlamch_t dlamch;
  dgeqrf_t dgeqrf;
  dormqr_t dormqr;
  dgeqpf_t dgeqpf;
  dgeqp3_t dgeqp3;
  ilaenv_t ilaenv;
  xerbla_t xerbla;
  dorgqr_t dorgqr;
  dtrsm_t dtrsm;
  dgemv_t dgemv;
  dger_t dger;
  dcopy_t dcopy;
  daxpy_t daxpy;
  dscal_t dscal;
  setthreads_t set_threads;
} LP;

static void *orc_sym(const char *name) {
  char buf[96];
  snprintf(buf, sizeof buf, "%s%s", LP.prefix, name);
  return dlsym(LP.h, buf);
}

int orc_init(const char *libpath) {
  if (LP.h) return 0;
  void *h = dlopen(libpath, RTLD_NOW | RTLD_LOCAL);
  if (!h) {
    fprintf(stderr, "oracle: dlopen(%s) failed: %s\n", libpath, dlerror());
    return -1;
  }
  LP.h = h;
  LP.prefix[0] = 0;
  if (!dlsym(h, "dgemm_") && dlsym(h, "scipy_dgemm_"))
    strcpy(LP.prefix, "scipy_");
#define GET(field, name)                                                    \
  do {                                                                      \
    *(void **)&LP.field = orc_sym(name);                                    \
    if (!LP.field) {                                                        \
      fprintf(stderr, "oracle: missing %s%s in %s\n", LP.prefix, name,      \
              libpath);                                                     \
      return -2;                                                            \
    }                                                                       \
  } while (0)
  GET(dgemm, "dgemm_");
  GET(dlacpy, "dlacpy_");
  GET(dlarfg, "dlarfg_");
  GET(dlarf, "dlarf_");
  GET(dlarft, "dlarft_");
  GET(dlarfb, "dlarfb_");
  GET(dnrm2, "dnrm2_");
  GET(idamax, "idamax_");
  GET(dswap, "dswap_");
  GET(dtrmm, "dtrmm_");
  GET(dlamch, "dlamch_");
  GET(dgeqrf, "dgeqrf_");
  GET(dormqr, "dormqr_");
  GET(dgeqpf, "dgeqpf_");
  GET(dgeqp3, "dgeqp3_");
  GET(ilaenv, "ilaenv_");
  GET(xerbla, "xerbla_");
  GET(dorgqr, "dorgqr_");
  GET(dtrsm, "dtrsm_");
  GET(dgemv, "dgemv_");
  GET(dger, "dger_");
  GET(dcopy, "dcopy_");
  GET(daxpy, "daxpy_");
  GET(dscal, "dscal_");
#undef GET
  *(void **)&LP.set_threads = orc_sym("openblas_set_num_threads");
  return 0;
}

int orc_set_threads(int n) {
  if (!LP.set_threads) return -1;
  LP.set_threads(n);
  return 0;
}

const char *orc_lapack_prefix(void) { return LP.prefix; }

/* ------------------------------------------------------------------------- */
/* Omega: the reference seeds glibc rand() with 12 (WY.c:380) and draws      */
/* normals with Marsaglia's polar method (WY.c:593-614), filling G column by  */
/* column (WY.c:575-590).  glibc's rand() is the additive lagged-Fibonacci    */
/* generator r[i] = r[i-3] + r[i-31] (mod 2^32) seeded by the Park-Miller     */
/* minimal standard LCG, with the first 310 outputs discarded and each output */
/* shifted right by one bit.  Restated here with caller-owned state so the    */
/* caller's own rand() stream is left alone.                                  */

void orc_rand_seed(orc_rand_state *s, unsigned int seed) {
  int32_t word = (int32_t)(seed == 0 ? 1 : seed);
  s->r[0] = (uint32_t)word;
  for (int i = 1; i < 31; ++i) {
    /* 16807 * word mod (2^31 - 1) without overflow (Schrage). */
    int32_t hi = word / 127773, lo = word % 127773;
    word = 16807 * lo - 2836 * hi;
    if (word < 0) word += 2147483647;
    s->r[i] = (uint32_t)word;
  }
  s->f = 3;
  s->b = 0;
  for (int i = 0; i < 310; ++i) (void)orc_rand_next(s);
}

int orc_rand_next(orc_rand_state *s) {
  s->r[s->f] += s->r[s->b];
  int out = (int)((s->r[s->f] >> 1) & 0x7fffffff);
  s->f = (s->f + 1) % 31;
  s->b = (s->b + 1) % 31;
  return out;
}

#define ORC_RAND_MAX 2147483647

/* One call of the reference's polar normal generator; `cache`/`have` hold the
 * second variate between calls as the reference's static locals do. */
static double orc_polar(orc_rand_state *s, double *cache, int *have) {
  if (*have) {
    *have = 0;
    return 0.0 + 1.0 * *cache;
  }
  double c1, c2, a;
  do {
    c1 = -1.0 + 2.0 * ((double)orc_rand_next(s) / ORC_RAND_MAX);
    c2 = -1.0 + 2.0 * ((double)orc_rand_next(s) / ORC_RAND_MAX);
    a = c1 * c1 + c2 * c2;
  } while ((a == 0) || (a >= 1));
  double factor = sqrt((-2 * log(a)) / a);
  *cache = c2 * factor;
  *have = 1;
  return 0.0 + 1.0 * (c1 * factor);
}

/* Omega as a call finds the generator's statics (*have, *cache) and leaves
 * them (WY.c:594-595: they survive the srand(12) of every call, WY.c:380);
 * NULL pointers: a fresh process. */
void orc_omega_carry(int r, int m, double *G, int ldg, unsigned int seed,
                     int *have, double *cache) {
  orc_rand_state s;
  double c = (have && *have && cache) ? *cache : 0.0;
  int h = (have && *have) ? 1 : 0;
  orc_rand_seed(&s, seed);
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < r; ++i)
      G[i + (size_t)j * ldg] = orc_polar(&s, &c, &h);
  if (have) *have = h;
  if (cache) *cache = c;
}

void orc_omega(int r, int m, double *G, int ldg, unsigned int seed) {
  orc_omega_carry(r, m, G, ldg, seed, NULL, NULL);
}

/* ------------------------------------------------------------------------- */
/* Norm helpers (WY.c:889-909 and WY.c:912-986).                              */

static void orc_colnorms(int m, int n, const double *A, int lda, double *d,
                         double *e) {
  static const int one = 1;
  for (int c = 0; c < n; ++c) {
    d[c] = LP.dnrm2(&m, A + (size_t)c * lda, &one);
    e[c] = d[c];
  }
}

/* LAWN 176 (Drmac & Bujanovic) partial-norm downdate.  `w` is the row just
 * produced by the reflector (stride ldw), `A` the part of each column below
 * it (m rows) used when the downdated value is no longer trusted. */
static void orc_downdate_norms(int m, int n, double *d, double *e,
                               const double *w, int ldw, const double *A,
                               int lda) {
  static const int one = 1;
  const double tol3z = sqrt(LP.dlamch("Epsilon", 7));
  for (int c = 0; c < n; ++c) {
    if (d[c] != 0.0) {
      double t = fabs(w[(size_t)c * ldw]) / d[c];
      t = OMAX(0.0, (1.0 + t) * (1 - t));
      double ratio = d[c] / e[c];
      double t2 = t * ratio * ratio;
      if (t2 <= tol3z) {
        if (m > 0) {
          d[c] = LP.dnrm2(&m, A + (size_t)c * lda, &one);
          e[c] = d[c];
        } else {
          d[c] = 0.0;
          e[c] = 0.0;
        }
      } else {
        d[c] = d[c] * sqrt(t);
      }
    }
  }
}

/* Swap semantics of WY.c:990-1036: column 0 <-> column k of G, B, C; pivot
 * entries swapped; the norms of column 0 copied (not swapped) to slot k. */
static void orc_pivot(int k, int mG, double *G, int ldg, int doB, int mB,
                      double *B, int ldb, int doC, int mC, double *C, int ldc,
                      int *p, double *d, double *e) {
  static const int one = 1;
  if (k == 0) return;
  LP.dswap(&mG, G, &one, G + (size_t)k * ldg, &one);
  if (doB) LP.dswap(&mB, B, &one, B + (size_t)k * ldb, &one);
  if (doC) LP.dswap(&mC, C, &one, C + (size_t)k * ldc, &one);
  int tmp = p[k];
  p[k] = p[0];
  p[0] = tmp;
  d[k] = d[0];
  e[k] = e[0];
}

/* ------------------------------------------------------------------------- */
/* Unblocked QR with optional column pivoting (WY.c:774-886).                 */

int orc_qrpmod(int pivoting, int num_stages, int m, int n, double *A, int lda,
               int *p, double *t, int pivot_B, int mB, double *B, int ldb,
               int pivot_C, int mC, double *C, int ldc, int build_T, double *T,
               int ldt, double *trace) {
  static const int one = 1;
  int mn = OMIN(m, n);
  if (num_stages < 0) num_stages = mn;
  double *d = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
  double *e = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
  double *wk = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
  if (!d || !e || !wk) {
    free(d);
    free(e);
    free(wk);
    return -1;
  }
  if (pivoting == 1) orc_colnorms(m, n, A, lda, d, e);

  for (int j = 0; j < num_stages; ++j) {
    int nleft = n - j;
    if (pivoting == 1) {
      int idx = LP.idamax(&nleft, d + j, &one) - 1;
      if (trace) {
        /* Largest and second largest candidate partial norms. */
        double best = d[j + idx], second = 0.0;
        for (int c = j; c < n; ++c)
          if (c != j + idx && fabs(d[c]) > second) second = fabs(d[c]);
        trace[2 * j] = best;
        trace[2 * j + 1] = second;
      }
      orc_pivot(idx, m, A + (size_t)j * lda, lda, pivot_B, mB,
                B + (size_t)j * ldb, ldb, pivot_C, mC, C + (size_t)j * ldc,
                ldc, p + j, d + j, e + j);
    }
    int hlen = m - j;
    double *ajj = A + j + (size_t)j * lda;
    LP.dlarfg(&hlen, ajj, A + OMIN(m - 1, j + 1) + (size_t)j * lda, &one,
              t + j);
    double saved = *ajj;
    *ajj = 1.0;
    int mrest = m - j, nrest = n - j - 1;
    LP.dlarf("Left", &mrest, &nrest, ajj, &one, t + j,
             A + j + (size_t)(j + 1) * lda, &lda, wk, 4);
    *ajj = saved;
    if (pivoting == 1)
      orc_downdate_norms(m - j - 1, n - j - 1, d + j + 1, e + j + 1,
                         A + j + (size_t)(j + 1) * lda, lda,
                         A + (j + 1) + (size_t)OMIN(n - 1, j + 1) * lda, lda);
  }
  if (build_T)
    LP.dlarft("Forward", "Columnwise", &m, &num_stages, A, &lda, t, T, &ldt, 7,
              10);
  free(d);
  free(e);
  free(wk);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Block reflector wrappers (WY.c:711-771).                                   */

int orc_apply_q_left_t(int mU, int nU, const double *U, int ldu,
                       const double *T, int ldt, int mB, int nB, double *B,
                       int ldb) {
  int ldw = OMAX(1, nB);
  double *W = (double *)malloc(sizeof(double) * (size_t)OMAX(1, nB) *
                               OMAX(1, nU));
  if (!W) return -1;
  (void)mU;
  LP.dlarfb("Left", "Transpose", "Forward", "Columnwise", &mB, &nB, &nU, U,
            &ldu, T, &ldt, B, &ldb, W, &ldw, 4, 9, 7, 10);
  free(W);
  return 0;
}

int orc_apply_q_right_n(int mU, int nU, const double *U, int ldu,
                        const double *T, int ldt, int mB, int nB, double *B,
                        int ldb) {
  int ldw = OMAX(1, mB);
  double *W = (double *)malloc(sizeof(double) * (size_t)OMAX(1, mB) *
                               OMAX(1, nU));
  if (!W) return -1;
  (void)mU;
  LP.dlarfb("Right", "No transpose", "Forward", "Columnwise", &mB, &nB, &nU,
            U, &ldu, T, &ldt, B, &ldb, W, &ldw, 5, 12, 7, 10);
  free(W);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Sketch downdate (WY.c:617-708):                                            */
/*   Y2 := Y2 - (G1 - (G1*U11 + G2*U21) * T * U11') * R12,                    */
/*   [G1 G2] := [G1 G2] * (I - U*T*U'),  U = [U11; U21] unit lower.           */

int orc_downdate_y(int b, int mU21, const double *U11, int ldu,
                   const double *R12, int ldr, int nY2, double *Y2, int ldy,
                   const double *T, int ldt, int r, double *G1, int nG2,
                   double *G2, int ldg) {
  static const double one = 1.0, mone = -1.0;
  double *Bm = (double *)malloc(sizeof(double) * (size_t)OMAX(1, r) *
                                OMAX(1, b));
  if (!Bm) return -1;
  int ldb = r;
  LP.dlacpy("All", &r, &b, G1, &ldg, Bm, &ldb, 3);
  LP.dtrmm("Right", "Lower", "No transpose", "Unit", &r, &b, &one, U11, &ldu,
           Bm, &ldb, 5, 5, 12, 4);
  LP.dgemm("No transpose", "No transpose", &r, &b, &mU21, &one, G2, &ldg,
           U11 + b, &ldu, &one, Bm, &ldb, 12, 12);
  LP.dtrmm("Right", "Upper", "No transpose", "Non-unit", &r, &b, &one, T,
           &ldt, Bm, &ldb, 5, 5, 12, 8);
  LP.dtrmm("Right", "Lower", "Conjugate transpose", "Unit", &r, &b, &mone, U11,
           &ldu, Bm, &ldb, 5, 5, 19, 4);
  for (int c = 0; c < b; ++c)
    for (int i = 0; i < r; ++i) Bm[i + (size_t)c * ldb] += G1[i + (size_t)c * ldg];
  LP.dgemm("No transpose", "No transpose", &r, &nY2, &b, &mone, Bm, &ldb, R12,
           &ldr, &one, Y2, &ldy, 12, 12);
  int nG = b + nG2;
  orc_apply_q_right_n(b + mU21, b, U11, ldu, T, ldt, r, nG, G1, ldg);
  free(Bm);
  (void)G2;
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Blocked HQRRP driver (WY.c:303-571).                                       */

/* Beyond the reference (test instrumentation and truncation): max_blocks
 * stops after that many blocks; max_cols > 0 stops after max_cols columns,
 * the block that reaches max_cols cut to end there unless it is the
 * reference's final block (last_iter), which is factored whole;
 * sketch_cols[j + s] receives the column block j's sketch chose at step s
 * (jpvt after the sketch QRCP, before the panel reorders the block). */
int orc_hqrrp_ex(int m, int n, double *A, int lda, int *jpvt, double *tau,
                 int nb, int pp, int panel_pivoting, const double *omega,
                 int max_blocks, double *trace_sketch, double *trace_panel,
                 double *omega_out, double *y_out, int max_cols,
                 int *sketch_cols) {
  static const double one = 1.0, zero = 0.0;
  if (m < 0) fprintf(stderr, "ERROR in oracle HQRRP: m_A is < 0.\n");
  if (n < 0) fprintf(stderr, "ERROR in oracle HQRRP: n_A is < 0.\n");
  if (lda < OMAX(1, m))
    fprintf(stderr, "ERROR in oracle HQRRP: ldim_A is < max( 1, m_A ).\n");
  int mn = OMIN(m, n);
  if (mn == 0) return 0;

  int r = nb + pp;
  double *Y = (double *)malloc(sizeof(double) * (size_t)r * n);
  double *V = (double *)malloc(sizeof(double) * (size_t)r * n);
  double *W = (double *)malloc(sizeof(double) * (size_t)nb * n);
  double *G = (double *)malloc(sizeof(double) * (size_t)r * m);
  if (!Y || !V || !W || !G) {
    free(Y);
    free(V);
    free(W);
    free(G);
    return -1;
  }
  if (omega)
    memcpy(G, omega, sizeof(double) * (size_t)r * m);
  else
    orc_omega(r, m, G, r, 12u);
  if (omega_out) memcpy(omega_out, G, sizeof(double) * (size_t)r * m);

  LP.dgemm("No transpose", "No transpose", &r, &n, &m, &one, G, &r, A, &lda,
           &zero, Y, &r, 12, 12);
  if (y_out) memcpy(y_out, Y, sizeof(double) * (size_t)r * n);

  int blocks = 0;
  for (int j = 0; j < mn; j += nb) {
    if (max_blocks > 0 && blocks >= max_blocks) break;
    if (max_cols > 0 && j >= max_cols) break;
    ++blocks;
    int b = OMIN(nb, OMIN(n - j, m - j));
    int last = (j + nb >= m) || (j + nb >= n);
    if (max_cols > 0 && !last) b = OMIN(b, max_cols - j);
    double *A_R = A + (size_t)j * lda;
    double *Y_R = Y + (size_t)j * r;
    double *V_R = V + (size_t)j * r;
    double *A11 = A + j + (size_t)j * lda;
    double *T1 = W + (size_t)j * nb;
    int mU21 = OMAX(0, m - j - b);
    int nA12 = OMAX(0, n - j - b);
    double *A12 = A + j + (size_t)OMIN(n - 1, j + b) * lda;
    double *Y2 = Y + (size_t)OMIN(n - 1, j + b) * r;
    double *G1 = G + (size_t)j * r;
    double *G2 = G + (size_t)OMIN(m - 1, j + b) * r;

    if (!last) {
      int nVR = n - j;
      LP.dlacpy("All", &r, &nVR, Y_R, &r, V_R, &r, 3);
      orc_qrpmod(1, b, r, nVR, V_R, r, jpvt + j, tau + j, 1, m, A_R, lda, 1, r,
                 Y_R, r, 0, Y, r, trace_sketch ? trace_sketch + 2 * j : NULL);
      if (sketch_cols) memcpy(sketch_cols + j, jpvt + j, sizeof(int) * b);
    }
    orc_qrpmod(panel_pivoting, -1, m - j, b, A11, lda, jpvt + j, tau + j, 1, j,
               A_R, lda, 1, r, Y_R, r, 1, T1, nb,
               trace_panel ? trace_panel + 2 * j : NULL);
    if (j + b < n)
      orc_apply_q_left_t(m - j, b, A11, lda, T1, nb, m - j, nA12, A12, lda);
    if (!last)
      orc_downdate_y(b, mU21, A11, lda, A12, lda, OMAX(0, n - j - b), Y2, r,
                     T1, nb, r, G1, OMAX(0, m - j - b), G2, r);
  }
  free(G);
  free(Y);
  free(V);
  free(W);
  return 0;
}

int orc_hqrrp(int m, int n, double *A, int lda, int *jpvt, double *tau, int nb,
              int pp, int panel_pivoting) {
  return orc_hqrrp_ex(m, n, A, lda, jpvt, tau, nb, pp, panel_pivoting, NULL, 0,
                      NULL, NULL, NULL, NULL, 0, NULL);
}

/* ------------------------------------------------------------------------- */
/* dgeqp3-compatible front end (WY.c:127-300).                                */

void orc_dgeqp4(int *m, int *n, double *A, int *lda, int *jpvt, double *tau,
                double *work, int *lwork, int *info) {
  static const int one = 1, ispec = 1, mone = -1;
  int mA = *m, nA = *n, ld = *lda;
  int mn = OMIN(mA, nA);
  int iws = 1, lwkopt = 1;
  int lquery = (*lwork == -1);
  *info = 0;
  if (mA < 0)
    *info = -1;
  else if (nA < 0)
    *info = -2;
  else if (ld < OMAX(1, mA))
    *info = -4;
  if (*info == 0) {
    if (mn != 0) {
      iws = 3 * nA + 1;
      int nbq = LP.ilaenv(&ispec, "DGEQRF", " ", &mA, &nA, &mone, &mone, 6, 1);
      lwkopt = 2 * nA + (nA + 1) * nbq;
    }
    work[0] = (double)lwkopt;
    if (*lwork < iws && !lquery) *info = -8;
  }
  if (*info != 0) {
    int arg = -*info;
    LP.xerbla("DGEQP3", &arg, 6);
    return;
  }
  if (lquery || mn == 0) return;
  if (mn < 250) {
    LP.dgeqpf(m, n, A, lda, jpvt, tau, work, info);
    return;
  }
  if (mn < 1000) {
    LP.dgeqp3(m, n, A, lda, jpvt, tau, work, lwork, info);
    return;
  }
  /* Fixed (jpvt != 0) columns to the front. */
  int nfixed = 0;
  for (int c = 0; c < nA; ++c) {
    if (jpvt[c] != 0) {
      if (c != nfixed) {
        LP.dswap(&mA, A + (size_t)c * ld, &one, A + (size_t)nfixed * ld, &one);
        jpvt[c] = jpvt[nfixed];
        jpvt[nfixed] = c + 1;
      } else {
        jpvt[c] = c + 1;
      }
      ++nfixed;
    } else {
      jpvt[c] = c + 1;
    }
  }
  int nf = OMIN(mA, nfixed);
  if (nf > 0) {
    LP.dgeqrf(&mA, &nf, A, &ld, tau, work, lwork, info);
    if (*info != 0) fprintf(stderr, "ERROR in dgeqrf: Info: %d \n", *info);
    iws = OMAX(iws, (int)work[0]);
    if (nf < nA) {
      int nrest = nA - nf;
      LP.dormqr("Left", "Transpose", &mA, &nrest, &nf, A, &ld, tau,
                A + (size_t)nf * ld, &ld, work, lwork, info, 4, 9);
      if (*info != 0) fprintf(stderr, "ERROR in dormqr: Info: %d \n", *info);
      iws = OMAX(iws, (int)work[0]);
    }
  }
  int *prev = (int *)malloc(sizeof(int) * (size_t)nA);
  if (nf > 0)
    for (int c = 0; c < nA; ++c) prev[c] = jpvt[c];
  if (nf < mn)
    *info = orc_hqrrp(mA - nf, nA - nf, A + nf + (size_t)nf * ld, ld,
                      jpvt + nf, tau + nf, 64, 10, 1);
  if (nf > 0) {
    /* Re-pivot the rows above the HQRRP block to the new column order. */
    for (int c = nf; c < nA; ++c) {
      int k = c;
      for (; k < nA; ++k)
        if (jpvt[c] == prev[k]) break;
      if (k != c) {
        int tmp = prev[c];
        prev[c] = prev[k];
        prev[k] = tmp;
        LP.dswap(&nf, A + (size_t)c * ld, &one, A + (size_t)k * ld, &one);
      }
    }
  }
  free(prev);
  work[0] = iws;
}

/* ------------------------------------------------------------------------- */
/* Plain LAPACK entry points used by tests and the CPU baseline.              */

void orc_dgeqp3(int *m, int *n, double *A, int *lda, int *jpvt, double *tau,
                double *work, int *lwork, int *info) {
  LP.dgeqp3(m, n, A, lda, jpvt, tau, work, lwork, info);
}

void orc_dorgqr(int *m, int *n, int *k, double *A, int *lda, const double *tau,
                double *work, int *lwork, int *info) {
  LP.dorgqr(m, n, k, A, lda, tau, work, lwork, info);
}

void orc_dormqr(const char *side, const char *trans, int *m, int *n, int *k,
                double *A, int *lda, const double *tau, double *C, int *ldc,
                double *work, int *lwork, int *info) {
  LP.dormqr(side, trans, m, n, k, A, lda, tau, C, ldc, work, lwork, info, 1,
            1);
}

void orc_dlarft(int n, int k, const double *V, int ldv, const double *tau,
                double *T, int ldt) {
  LP.dlarft("Forward", "Columnwise", &n, &k, V, &ldv, tau, T, &ldt, 7, 10);
}

/* ------------------------------------------------------------------------- */
/* The libflame UT variant, FLA_HQRRP_UT_blk_var2                             */
/* (libflame_sources/FLA_HQRRP_UT_blk_var2.c, below UT.c), restated on plain  */
/* column-major arrays.  Its FLA_Obj views become pointers and sizes; its     */
/* libflame calls (FLA_Gemm, FLA_Trmm, FLA_Trsm, FLA_Copy, FLA_Axpy) become   */
/* the BLAS calls they perform.  Its Householder vectors come from libflame's */
/* FLA_Househ2_UT_l_opd, which is not in the reference tree (libflame is a    */
/* third-party library; the reference pins no version); its published        */
/* algorithm is restated in orc_househ2_ut.  libflame itself being absent,    */
/* this driver is pinned by the mathematics of the UT transform and by the    */
/* WY driver above (tests/test_oracle.py), not by a compiled reference.       */

/* FLA_Househ2_UT_l_opd: chi_1 := alpha = -sign(chi_1) ||[chi_1; x2]||,
 * x2 := u2 = x2 / (chi_1 - alpha), tau := (1 + u2'u2) / 2, so that
 * (I - [1; u2][1; u2]' / tau) [chi_1; x2] = [alpha; 0]; when x2 = 0 it
 * returns chi_1 := -chi_1, tau := 1/2 (the reflection I - 2 e1 e1'). */
static void orc_househ2_ut(int m_x2, double *chi_1, double *x2, double *tau) {
  static const int one = 1;
  const double norm_x_2 = m_x2 > 0 ? LP.dnrm2(&m_x2, x2, &one) : 0.0;
  if (norm_x_2 == 0.0) {
    *chi_1 = -(*chi_1);
    *tau = 0.5;
    return;
  }
  const int two = 2;
  double y[2] = {fabs(*chi_1), norm_x_2};
  const double norm_x = LP.dnrm2(&two, y, &one);
  const double alpha = (*chi_1 < 0.0 ? 1.0 : -1.0) * norm_x;
  const double cma = *chi_1 - alpha;
  const double inv = 1.0 / cma;
  LP.dscal(&m_x2, &inv, x2, &one);
  const double r = norm_x_2 / fabs(cma);
  *tau = 0.5 + 0.5 * (r * r);
  *chi_1 = alpha;
}

/* MyFLA_Apply_H2_UT_l_opd_var2 (UT.c:826-928): [a1t; A2] := H [a1t; A2]. */
static void orc_apply_h2_ut(int m_u2, int n_a1t, const double *tau,
                            const double *u2, double *a1t, int inc_a1t,
                            double *A2, int lda2, double *w) {
  static const int one = 1;
  static const double d1 = 1.0, dm1 = -1.0;
  if (n_a1t == 0) return;
  LP.dcopy(&n_a1t, a1t, &inc_a1t, w, &one);
  if (m_u2 > 0)
    LP.dgemv("Transpose", &m_u2, &n_a1t, &d1, A2, &lda2, u2, &one, &d1, w, &one,
             1);
  const double rtau = 1.0 / *tau; /* tau >= 1/2 in this convention */
  LP.dscal(&n_a1t, &rtau, w, &one);
  LP.daxpy(&n_a1t, &dm1, w, &one, a1t, &inc_a1t);
  if (m_u2 > 0) LP.dger(&m_u2, &n_a1t, &dm1, u2, &one, w, &one, A2, &lda2);
}

/* NoFLA_my_idamax (UT.c:1123-1166): first index of the largest |d|. */
static int orc_my_idamax(int n, const double *d) {
  if (n < 1) return -1;
  int res = 0;
  double dmax = fabs(d[0]);
  for (int i = 1; i < n; ++i)
    if (dmax < fabs(d[i])) {
      res = i;
      dmax = fabs(d[i]);
    }
  return res;
}

/* MyFLA_QRPmod_UT_unb_var2 (UT.c:633-823), with NoFLA_QRP_pivot_B_C
 * (UT.c:986-1040), MyFLA_QRP_compute_norms (UT.c:931-983) and
 * NoFLA_QRP_downdate_partial_norms (UT.c:1043-1120). */
static void orc_qrpmod_ut(int pivoting, int num_stages, int m, int n,
                          double *A, int lda, int *p, double *t, int pivot_B,
                          int mB, double *B, int ldb, int pivot_C, int mC,
                          double *Cm, int ldc, int build_T, double *T,
                          int ldt) {
  static const int one = 1;
  static const double d1 = 1.0;
  const int mn = OMIN(m, n);
  if (num_stages < 0) num_stages = mn;
  double *d = (double *)malloc(sizeof(double) * OMAX(1, n));
  double *e = (double *)malloc(sizeof(double) * OMAX(1, n));
  double *w = (double *)malloc(sizeof(double) * OMAX(1, n));
  if (pivoting == 1)
    for (int j = 0; j < n; ++j) {
      d[j] = LP.dnrm2(&m, A + (size_t)j * lda, &one);
      e[j] = d[j];
    }
  const double tol3z = sqrt(LP.dlamch("Epsilon", 7));
  for (int j = 0; j < num_stages; ++j) {
    const int n_dB = n - j, m_a21 = m - j - 1, n_A22 = n - j - 1;
    if (pivoting == 1) {
      const int q = orc_my_idamax(n_dB, d + j);
      if (q != 0) {
        LP.dswap(&m, A + (size_t)j * lda, &one, A + (size_t)(j + q) * lda,
                 &one);
        if (pivot_B)
          LP.dswap(&mB, B + (size_t)j * ldb, &one, B + (size_t)(j + q) * ldb,
                   &one);
        if (pivot_C)
          LP.dswap(&mC, Cm + (size_t)j * ldc, &one, Cm + (size_t)(j + q) * ldc,
                   &one);
        const int iv = p[j + q];
        p[j + q] = p[j];
        p[j] = iv;
        d[j + q] = d[j];
        e[j + q] = e[j];
      }
    }
    double *a11 = A + j + (size_t)j * lda;
    orc_househ2_ut(m_a21, a11, a11 + 1, t + j);
    orc_apply_h2_ut(m_a21, n_A22, t + j, a11 + 1, A + j + (size_t)(j + 1) * lda,
                    lda, A + (j + 1) + (size_t)(j + 1) * lda, lda, w);
    if (build_T) {
      T[j + (size_t)j * ldt] = t[j];
      LP.dcopy(&j, A + j, &lda, T + (size_t)j * ldt, &one);
      if (m_a21 > 0 && j > 0)
        LP.dgemv("Transpose", &m_a21, &j, &d1, A + (j + 1), &lda, a11 + 1,
                 &one, &d1, T + (size_t)j * ldt, &one, 1);
    }
    if (pivoting == 1) {
      for (int c = 0; c < n_A22; ++c) {
        double *dc = d + j + 1 + c, *ec = e + j + 1 + c;
        if (*dc != 0.0) {
          double temp = fabs(A[j + (size_t)(j + 1 + c) * lda]) / *dc;
          temp = OMAX(0.0, (1.0 + temp) * (1 - temp));
          const double temp5 = *dc / *ec;
          const double temp2 = temp * temp5 * temp5;
          if (temp2 <= tol3z) {
            if (m_a21 > 0) {
              *dc = LP.dnrm2(&m_a21, A + (j + 1) + (size_t)(j + 1 + c) * lda,
                             &one);
              *ec = *dc;
            } else {
              *dc = 0.0;
              *ec = 0.0;
            }
          } else {
            *dc = *dc * sqrt(temp);
          }
        }
      }
    }
  }
  free(d);
  free(e);
  free(w);
}

/* MyFLA_Apply_Q_UT_lhfc_blk_var2 (UT.c:521-573): [B1; B2] := Q' [B1; B2]
 * with Q = I - U inv(T)' U'... as written there: W = inv(triu(T))' (U11'B1 +
 * U21'B2), B2 -= U21 W, B1 -= U11 W.  U11 (b x b unit lower), U21 (mU x b),
 * B1 (b x nB), B2 (mU x nB). */
static void orc_apply_q_ut_lhfc(int b, int mU, int nB, const double *U11,
                                int ldu, const double *U21, const double *T,
                                int ldt, double *B1, double *B2, int ldb) {
  static const double d1 = 1.0, dm1 = -1.0;
  if (nB == 0 || b == 0) return;
  double *W = (double *)malloc(sizeof(double) * (size_t)b * nB);
  LP.dlacpy("All", &b, &nB, B1, &ldb, W, &b, 3);
  LP.dtrmm("Left", "Lower", "Conjugate transpose", "Unit", &b, &nB, &d1, U11,
           &ldu, W, &b, 4, 5, 19, 4);
  if (mU > 0)
    LP.dgemm("Conjugate transpose", "No transpose", &b, &nB, &mU, &d1, U21,
             &ldu, B2, &ldb, &d1, W, &b, 19, 12);
  LP.dtrsm("Left", "Upper", "Conjugate transpose", "Non-unit", &b, &nB, &d1, T,
           &ldt, W, &b, 4, 5, 19, 8);
  if (mU > 0)
    LP.dgemm("No transpose", "No transpose", &mU, &nB, &b, &dm1, U21, &ldu, W,
             &b, &d1, B2, &ldb, 12, 12);
  LP.dtrmm("Left", "Lower", "No transpose", "Unit", &b, &nB, &dm1, U11, &ldu, W,
           &b, 4, 5, 12, 4);
  for (int c = 0; c < nB; ++c)
    for (int i = 0; i < b; ++i) B1[i + (size_t)c * ldb] += W[i + (size_t)c * b];
  free(W);
}

/* MyFLA_Apply_Q_UT_rnfc_blk_var2 (UT.c:576-630): [B1 B2] := [B1 B2] Q with
 * W = inv(triu(T))' (U11'B1' + U21'B2'), B2 -= W'U21', B1 -= (U11 W)'.
 * B1 (r x b), B2 (r x mU). */
static void orc_apply_q_ut_rnfc(int b, int mU, int r, const double *U11,
                                int ldu, const double *U21, const double *T,
                                int ldt, double *B1, double *B2, int ldg) {
  static const double d1 = 1.0, dm1 = -1.0;
  double *W = (double *)malloc(sizeof(double) * (size_t)b * OMAX(1, r));
  for (int i = 0; i < r; ++i)
    for (int c = 0; c < b; ++c) W[c + (size_t)i * b] = B1[i + (size_t)c * ldg];
  LP.dtrmm("Left", "Lower", "Transpose", "Unit", &b, &r, &d1, U11, &ldu, W, &b,
           4, 5, 9, 4);
  if (mU > 0)
    LP.dgemm("Transpose", "Transpose", &b, &r, &mU, &d1, U21, &ldu, B2, &ldg,
             &d1, W, &b, 9, 9);
  LP.dtrsm("Left", "Upper", "Transpose", "Non-unit", &b, &r, &d1, T, &ldt, W, &b,
           4, 5, 9, 8);
  if (mU > 0)
    LP.dgemm("Transpose", "Conjugate transpose", &r, &mU, &b, &dm1, W, &b, U21,
             &ldu, &d1, B2, &ldg, 9, 19);
  LP.dtrmm("Left", "Lower", "No transpose", "Unit", &b, &r, &dm1, U11, &ldu, W,
           &b, 4, 5, 12, 4);
  for (int i = 0; i < r; ++i)
    for (int c = 0; c < b; ++c) B1[i + (size_t)c * ldg] += W[c + (size_t)i * b];
  free(W);
}

/* MyFLA_Downdate_Y (UT.c:465-518): B = G1 - (G1 U11 + G2 U21) inv(T) U11',
 * Y2 -= B R12, then [G1 G2] := [G1 G2] Q. */
static void orc_downdate_y_ut(int b, int mU, int r, int nY2, const double *U11,
                              int ldu, const double *U21, const double *R12,
                              int ldr, const double *T, int ldt, double *Y2,
                              int ldy, double *G1, double *G2, int ldg) {
  static const double d1 = 1.0, dm1 = -1.0;
  double *Bm = (double *)malloc(sizeof(double) * (size_t)r * b);
  LP.dlacpy("All", &r, &b, G1, &ldg, Bm, &r, 3);
  LP.dtrmm("Right", "Lower", "No transpose", "Unit", &r, &b, &d1, U11, &ldu, Bm,
           &r, 5, 5, 12, 4);
  if (mU > 0)
    LP.dgemm("No transpose", "No transpose", &r, &b, &mU, &d1, G2, &ldg, U21,
             &ldu, &d1, Bm, &r, 12, 12);
  LP.dtrsm("Right", "Upper", "No transpose", "Non-unit", &r, &b, &d1, T, &ldt,
           Bm, &r, 5, 5, 12, 8);
  LP.dtrmm("Right", "Lower", "Conjugate transpose", "Unit", &r, &b, &dm1, U11,
           &ldu, Bm, &r, 5, 5, 19, 4);
  for (int c = 0; c < b; ++c)
    for (int i = 0; i < r; ++i) Bm[i + (size_t)c * r] += G1[i + (size_t)c * ldg];
  if (nY2 > 0)
    LP.dgemm("No transpose", "No transpose", &r, &nY2, &b, &dm1, Bm, &r, R12,
             &ldr, &d1, Y2, &ldy, 12, 12);
  orc_apply_q_ut_rnfc(b, mU, r, U11, ldu, U21, T, ldt, G1, G2, ldg);
  free(Bm);
}

/* FLA_HQRRP_UT_blk_var2 (UT.c:124-401).  carry: the generator's statics
 * as the call finds them (NULL: a fresh process). */
int orc_hqrrp_ut(int m, int n, double *A, int lda, int *p, double *s, int nb,
                 int pp, int panel_pivoting, int *carry_have,
                 double *carry_value) {
  static const double d1 = 1.0, d0 = 0.0;
  const int mn = OMIN(m, n);
  if (mn == 0) return 0;
  const int r = nb + pp;
  double *G = (double *)malloc(sizeof(double) * (size_t)r * m);
  double *Y = (double *)malloc(sizeof(double) * (size_t)r * n);
  double *V = (double *)malloc(sizeof(double) * (size_t)r * n);
  double *W = (double *)malloc(sizeof(double) * (size_t)nb * n);
  orc_omega_carry(r, m, G, r, 12u, carry_have, carry_value);
  LP.dgemm("No transpose", "No transpose", &r, &n, &m, &d1, G, &r, A, &lda, &d0,
           Y, &r, 12, 12);
  for (int j = 0; j < mn;) {
    const int b = OMIN(OMIN(m - j, n - j), nb);
    const int m22 = m - j - b, n22 = n - j - b;
    const int last = (OMIN(m22, n22) <= 0);
    int *pB = p + j;
    double *sB = s + j;
    if (!last) {
      const int nR = n - j;
      LP.dlacpy("All", &r, &nR, Y + (size_t)j * r, &r, V + (size_t)j * r, &r, 3);
      orc_qrpmod_ut(1, b, r, nR, V + (size_t)j * r, r, pB, sB, 1, m,
                    A + (size_t)j * lda, lda, 1, r, Y + (size_t)j * r, r, 0,
                    NULL, 0);
    }
    double *T1 = W + (size_t)j * nb; /* W1's first b rows, ld nb */
    double *A11 = A + j + (size_t)j * lda;
    orc_qrpmod_ut(panel_pivoting, -1, m - j, b, A11, lda, pB, sB, 1, j,
                  A + (size_t)j * lda, lda, 1, r, Y + (size_t)j * r, r, 1, T1,
                  nb);
    const int nA12 = n - j - b;
    if (nA12 > 0) {
      orc_apply_q_ut_lhfc(b, m - j - b, nA12, A11, lda, A11 + b, T1, nb,
                          A + j + (size_t)(j + b) * lda,
                          A + (j + b) + (size_t)(j + b) * lda, lda);
    }
    const int nY2 = n - j - b;
    if (nY2 > 0)
      orc_downdate_y_ut(b, m - j - b, r, nY2, A11, lda, A11 + b,
                        A + j + (size_t)(j + b) * lda, lda, T1, nb,
                        Y + (size_t)(j + b) * r, r, G + (size_t)j * r,
                        G + (size_t)(j + b) * r, r);
    j += b;
  }
  free(G);
  free(Y);
  free(V);
  free(W);
  return 0;
}
