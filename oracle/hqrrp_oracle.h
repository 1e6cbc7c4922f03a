/* hqrrp_oracle.h -- CPU restatement of the reference HQRRP.  TEST
 * INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, never by the product library. */
#ifndef HQRRP_ORACLE_H
#define HQRRP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  uint32_t r[31];
  int f, b;
} orc_rand_state;

int orc_init(const char *lapack_path);
int orc_set_threads(int n);
const char *orc_lapack_prefix(void);

void orc_rand_seed(orc_rand_state *s, unsigned int seed);
int orc_rand_next(orc_rand_state *s);
void orc_omega(int r, int m, double *G, int ldg, unsigned int seed);
void orc_omega_carry(int r, int m, double *G, int ldg, unsigned int seed,
                     int *have, double *cache);
int orc_hqrrp_ut(int m, int n, double *A, int lda, int *p, double *s, int nb,
                 int pp, int panel_pivoting, int *carry_have,
                 double *carry_value);

int orc_qrpmod(int pivoting, int num_stages, int m, int n, double *A, int lda,
               int *p, double *t, int pivot_B, int mB, double *B, int ldb,
               int pivot_C, int mC, double *C, int ldc, int build_T, double *T,
               int ldt, double *trace);
int orc_apply_q_left_t(int mU, int nU, const double *U, int ldu,
                       const double *T, int ldt, int mB, int nB, double *B,
                       int ldb);
int orc_apply_q_right_n(int mU, int nU, const double *U, int ldu,
                        const double *T, int ldt, int mB, int nB, double *B,
                        int ldb);
int orc_downdate_y(int b, int mU21, const double *U11, int ldu,
                   const double *R12, int ldr, int nY2, double *Y2, int ldy,
                   const double *T, int ldt, int r, double *G1, int nG2,
                   double *G2, int ldg);
int orc_hqrrp_ex(int m, int n, double *A, int lda, int *jpvt, double *tau,
                 int nb, int pp, int panel_pivoting, const double *omega,
                 int max_blocks, double *trace_sketch, double *trace_panel,
                 double *omega_out, double *y_out, int max_cols,
                 int *sketch_cols);
int orc_hqrrp(int m, int n, double *A, int lda, int *jpvt, double *tau, int nb,
              int pp, int panel_pivoting);
void orc_dgeqp4(int *m, int *n, double *A, int *lda, int *jpvt, double *tau,
                double *work, int *lwork, int *info);
void orc_dgeqp3(int *m, int *n, double *A, int *lda, int *jpvt, double *tau,
                double *work, int *lwork, int *info);
void orc_dorgqr(int *m, int *n, int *k, double *A, int *lda, const double *tau,
                double *work, int *lwork, int *info);
void orc_dormqr(const char *side, const char *trans, int *m, int *n, int *k,
                double *A, int *lda, const double *tau, double *C, int *ldc,
                double *work, int *lwork, int *info);
void orc_dlarft(int n, int k, const double *V, int ldv, const double *tau,
                double *T, int ldt);

#ifdef __cplusplus
}
#endif
#endif
