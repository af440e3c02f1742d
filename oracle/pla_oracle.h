/*
 * pla_oracle.h -- CPU restatement of PLAPACK R3.0's dense factor-and-solve
 * path, used ONLY as a checker.
 *
 * This is test infrastructure.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only to check or time the
 * HIP library against it.  The product library (libpla_mi355x.so) never
 * links, loads or calls anything here.
 *
 * All matrices are column-major double with a leading dimension, exactly as
 * a PLA_Obj's local buffer on a 1x1 template (reference OBJ/PLA_Obj.c:303).
 * Pivots use the reference's encoding: piv[j] is the 0-based offset of the
 * pivot row relative to row j (0 = no swap), SOLVERS/LU/PLA_seq_lu.c:33-35.
 *
 * Pinning (what checks each function against something independent of this
 * file; the checks live in tests/test_oracle.py):
 *   ora_srand48/ora_drand48/ora_fill_uniform : glibc srand48/drand48 via libc
 *   ora_lcg_skip                              : stepping the generator k times
 *   ora_idamax                                : numpy argmax(|x|) (first max)
 *   ora_lu_linpack                            : ora_lu_blocked's U, pivots and
 *       appended column bit for bit; its L after the later interchanges
 *   ora_lu_unblocked / ora_lu_blocked         : scipy LAPACK getrf pivots,
 *       exact rational elimination (tiny n, incl. exact ties), and the
 *       backward-error bound ||PA-LU|| / (n ||A|| eps) < 16
 *   ora_chol_unblocked / ora_chol_blocked     : numpy.linalg.cholesky, and
 *       the failure index against exact leading-minor determinants
 *   ora_laswp                                 : an explicit permutation
 *   ora_trsv / ora_trsm_left / ora_trsm_right : scipy solve_triangular
 *   ora_gemm / ora_syrk                       : numpy matmul (ora_syrk also:
 *       the opposite triangle is exactly beta C)
 *   ora_zlu_blocked / ora_zlu_solve           : scipy LAPACK zgetrf pivots and
 *       factors; solve residuals
 *   ora_zchol_blocked                         : numpy.linalg.cholesky (complex),
 *       failure index against exact leading-minor signs
 *   ora_api_*                                 : numpy slice arithmetic,
 *       bit-exact (one rounded product, one rounded add per element)
 * No function here is "parity unpinned".
 */
#ifndef PLA_ORACLE_H
#define PLA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* POSIX drand48: X <- (0x5DEECE66D X + 0xB) mod 2^48; srand48(s): X = s<<16 | 0x330E */
uint64_t ora_srand48(long seed);
double ora_drand48(uint64_t *state);
uint64_t ora_lcg_skip(uint64_t state, uint64_t k);
/* A(i,j) = drand48()*2-1 in column-major draw order after `skip` draws from srand48(seed)
   (TEST/create_problem.c:31-52). */
void ora_fill_uniform(int m, int n, double *a, int lda, long seed, int64_t skip);

int ora_idamax(int n, const double *x, int incx);

/* PLA_dlu: unblocked right-looking LU (SOLVERS/LU/PLA_seq_lu.c:21-46). */
void ora_lu_unblocked(int m, int n, double *a, int lda, int *piv);
/* PLA_LU_right: blocked right-looking LU (SOLVERS/LU/PLA_LU.c:37-212). */
void ora_lu_blocked(int m, int n, double *a, int lda, int nb, int *piv);
/* PLA_LU_linpack (SOLVERS/LINPACK/PLA_LU_linpack.c:27-171): as ora_lu_blocked,
   but each panel's interchanges reach only the columns from the panel on. */
void ora_lu_linpack(int m, int n, double *a, int lda, int nb, int *piv);

/* PLA_Apply_pivots_to_rows: forward sequential interchanges of rows
   r0+i <-> r0+i+piv[i], i = 0..npiv-1, over ncols columns
   (SOLVERS/LU/PLA_Apply_pivots_new.c:104-164). */
void ora_laswp(int ncols, double *a, int lda, int r0, int npiv, const int *piv);

/* PLA_dchol: unblocked lower Cholesky; returns 0 or k+1 at the first
   non-positive pivot (SOLVERS/Cholesky/PLA_Local_chol.c:103-143). */
int ora_chol_unblocked(int n, double *a, int lda);
/* PLA_Chol_lower_large (SOLVERS/Cholesky/PLA_Chol.c:83-234); returns 0 or the
   global 1-based failure index k + i + 1. */
int ora_chol_blocked(int n, double *a, int lda, int nb);

/* Triangular solve, one right-hand side: uplo 'L'/'U', trans 'N'/'T', diag 'U'/'N'. */
void ora_trsv(char uplo, char trans, char diag, int n, const double *a, int lda, double *x);
/* B <- alpha * op(A)^{-1} B, A n x n triangular, B n x nrhs. */
void ora_trsm_left(char uplo, char trans, char diag, int n, int nrhs, double alpha,
                   const double *a, int lda, double *b, int ldb);
/* B <- alpha B op(A)^{-1}, A n x n triangular, B m x n. */
void ora_trsm_right(char uplo, char trans, char diag, int m, int n, double alpha, const double *a, int lda,
                    double *b, int ldb);
/* all of C <- beta C, then the uplo triangle += alpha op(A) op(A)^T (PLA_Syrk) */
void ora_syrk(char uplo, char trans, int n, int k, double alpha, const double *a, int lda, double beta, double *c,
              int ldc);
/* C <- alpha op(A) op(B) + beta C */
void ora_gemm(char ta, char tb, int m, int n, int k, double alpha, const double *a, int lda,
              const double *b, int ldb, double beta, double *c, int ldc);

/* Application interface on one process (API/PLA_API.c:390-911):
   A(dr:, dc:) += alpha buf, buf += alpha A(dr:, dc:), and the multi-block
   forms (to_global != 0: buffer into A). */
void ora_api_to_global(int m, int n, double alpha, const double *buf, int ldb, double *a, int lda, int dr, int dc);
void ora_api_from_global(int m, int n, double alpha, const double *a, int lda, int dr, int dc, double *buf,
                         int ldb);
void ora_api_multi(int to_global, int nr, int nc, const int *rsz, const int *csz, double alpha, double *buf,
                   int ldb, double *a, int lda, const int *rd, const int *cd);

/* MPI_DOUBLE_COMPLEX PLA_LU (interleaved re/im, column-major): izamax
   pivots (|re| + |im|), zgeru updates; and PLA_General_solve on its factors. */
void ora_zlu_blocked(int m, int n, void *a, int lda, int nb, int *piv);
void ora_zlu_solve(int n, int nrhs, const void *lu, int lda, const int *piv, void *b, int ldb);
/* MPI_DOUBLE_COMPLEX PLA_Chol (lower, A = L L^H): 0 or the failure index. */
int ora_zchol_blocked(int n, void *a, int lda, int nb);

/* PLA_General_solve (SOLVERS/LU/PLA_General.c:67-119) on already-factored A:
   apply pivots to B, then L (unit) and U solves. */
void ora_lu_solve(int n, int nrhs, const double *lu, int lda, const int *piv, double *b, int ldb);
/* PLA_Pos_def_solve (SOLVERS/Cholesky/PLA_Pos_def.c:44-99) on factored L. */
void ora_chol_solve(int n, int nrhs, const double *l, int lda, double *b, int ldb);

#ifdef __cplusplus
}
#endif
#endif
