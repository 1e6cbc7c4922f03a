/*
 * ref_expose.c -- build recipe helper for oracle/_ref (TEST INFRASTRUCTURE).
 *
 * Compiles the reference's own lapack_compatible_sources/NoFLA_HQRRP_WY_blk_var4.c
 * (included by path from the reference tree, never copied) and exports thin
 * wrappers around its file-static helpers so tests can compare the oracle's
 * restatement with the reference function by function.  The reference's LAPACK
 * calls are bound to scipy's bundled OpenBLAS by -D renames in oracle/Makefile.
 */
#include REF_SOURCE

int ref_qrpmod(int pivoting, int num_stages, int m, int n, double *A, int lda,
               int *p, double *t, int pivot_B, int mB, double *B, int ldb,
               int pivot_C, int mC, double *C, int ldc, int build_T, double *T,
               int ldt) {
  return NoFLA_QRPmod_WY_unb_var4(pivoting, num_stages, m, n, A, lda, p, t,
                                  pivot_B, mB, B, ldb, pivot_C, mC, C, ldc,
                                  build_T, T, ldt);
}

int ref_downdate_y(int b, int mU21, double *U11, int ldu, double *R12, int ldr,
                   int nY2, double *Y2, int ldy, double *T, int ldt, int r,
                   double *G1, int nG2, double *G2, int ldg) {
  return NoFLA_Downdate_Y(b, b, U11, ldu, mU21, b, U11 + b, ldu, b, nY2, R12,
                          ldr, b, b, T, ldt, r, nY2, Y2, ldy, r, b, G1, ldg, r,
                          nG2, G2, ldg);
}

/* Omega exactly as the reference draws it: srand(12) then the polar method. */
int ref_omega(int r, int m, double *G, int ldg) {
  srand(12);
  return NoFLA_Normal_random_matrix(r, m, G, ldg);
}
