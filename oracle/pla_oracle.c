/*
 * pla_oracle.c -- CPU restatement of PLAPACK R3.0's LU / Cholesky path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP library is compared with.
 * Nothing in the product library links or calls this file.
 *
 * Every routine follows the reference algorithm step for step (citations
 * inline).  The BLAS calls the reference makes (idamax, dswap, dscal, dger,
 * dtrsm, dgemm, dgemv) are written out in the loop order of the netlib
 * reference BLAS, which is what the reference's Dockerfile:9 installs.
 */
#include "pla_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define A_(i, j) a[(size_t)(j) * (size_t)lda + (size_t)(i)]

/* ------------------------------------------------------------------ */
/* drand48 (POSIX closed form), as seeded in TEST/create_problem.c:31 */

static const uint64_t LCG_A = 0x5DEECE66DULL;
static const uint64_t LCG_C = 0xBULL;
static const uint64_t LCG_MASK = (1ULL << 48) - 1;

uint64_t ora_srand48(long seed) {
  return ((((uint64_t)(uint32_t)seed) << 16) | 0x330EULL) & LCG_MASK;
}

double ora_drand48(uint64_t *state) {
  *state = (LCG_A * *state + LCG_C) & LCG_MASK;
  /* glibc builds 1.0 + X*2^-48 in IEEE bits and subtracts 1.0: exactly X*2^-48 */
  return ldexp((double)*state, -48);
}

/* State after k further steps: affine map (a, c) composed by squaring. */
uint64_t ora_lcg_skip(uint64_t state, uint64_t k) {
  uint64_t mul = LCG_A, add = LCG_C;   /* current power-of-two step map */
  uint64_t acc_mul = 1, acc_add = 0;   /* accumulated map */
  while (k) {
    if (k & 1) {
      acc_mul = (acc_mul * mul) & LCG_MASK;
      acc_add = (acc_add * mul + add) & LCG_MASK;
    }
    add = ((mul + 1) * add) & LCG_MASK;
    mul = (mul * mul) & LCG_MASK;
    k >>= 1;
  }
  return (acc_mul * state + acc_add) & LCG_MASK;
}

void ora_fill_uniform(int m, int n, double *a, int lda, long seed, int64_t skip) {
  uint64_t st = ora_lcg_skip(ora_srand48(seed), (uint64_t)skip);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++) A_(i, j) = ora_drand48(&st) * 2.0 - 1.0;
}

/* ------------------------------------------------------------------ */
/* BLAS-1/2 pieces (netlib loop order)                                */

int ora_idamax(int n, const double *x, int incx) {
  /* first index of max |x_i| (strict '>' keeps the first), 0-based */
  if (n <= 0) return -1;
  int best = 0;
  double dmax = fabs(x[0]);
  for (int i = 1; i < n; i++) {
    double v = fabs(x[(size_t)i * incx]);
    if (v > dmax) { dmax = v; best = i; }
  }
  return best;
}

static void swap_rows(int ncols, double *a, int lda, int r1, int r2) {
  if (r1 == r2) return;
  for (int j = 0; j < ncols; j++) {
    double t = A_(r1, j);
    A_(r1, j) = A_(r2, j);
    A_(r2, j) = t;
  }
}

/* ------------------------------------------------------------------ */
/* LU                                                                 */

/* PLA_dlu, SOLVERS/LU/PLA_seq_lu.c:21-46: per column idamax, full-width
   dswap, scale by the reciprocal of the pivot, dger on the trailing block. */
void ora_lu_unblocked(int m, int n, double *a, int lda, int *piv) {
  int mn = m < n ? m : n;
  for (int j = 0; j < mn; j++) {
    int idx = ora_idamax(m - j, &A_(j, j), 1);
    piv[j] = idx;
    swap_rows(n, a, lda, j, j + idx);
    double inv = 1.0 / A_(j, j);
    for (int i = j + 1; i < m; i++) A_(i, j) *= inv;
    /* dger: for each column c, temp = alpha*y(c); A(:,c) += x*temp */
    for (int c = j + 1; c < n; c++) {
      double temp = -A_(j, c);
      if (temp != 0.0)
        for (int i = j + 1; i < m; i++) A_(i, c) += A_(i, j) * temp;
    }
  }
}

void ora_laswp(int ncols, double *a, int lda, int r0, int npiv, const int *piv) {
  for (int i = 0; i < npiv; i++)
    if (piv[i] != 0) swap_rows(ncols, a, lda, r0 + i, r0 + i + piv[i]);
}

/* netlib dgemm, notrans/trans for A and B */
void ora_gemm(char ta, char tb, int m, int n, int k, double alpha, const double *a, int lda,
              const double *b, int ldb, double beta, double *c, int ldc) {
  int tA = (ta == 'T' || ta == 't'), tB = (tb == 'T' || tb == 't');
  for (int j = 0; j < n; j++) {
    double *cj = c + (size_t)j * ldc;
    if (beta == 0.0)
      for (int i = 0; i < m; i++) cj[i] = 0.0;
    else if (beta != 1.0)
      for (int i = 0; i < m; i++) cj[i] *= beta;
    if (!tA) {
      for (int l = 0; l < k; l++) {
        double bl = tB ? b[(size_t)l * ldb + j] : b[(size_t)j * ldb + l];
        double temp = alpha * bl;
        if (temp != 0.0) {
          const double *al = a + (size_t)l * lda;
          for (int i = 0; i < m; i++) cj[i] += temp * al[i];
        }
      }
    } else {
      for (int i = 0; i < m; i++) {
        double s = 0.0;
        const double *ai = a + (size_t)i * lda;
        for (int l = 0; l < k; l++) {
          double bl = tB ? b[(size_t)l * ldb + j] : b[(size_t)j * ldb + l];
          s += ai[l] * bl;
        }
        cj[i] += alpha * s;
      }
    }
  }
}

/* PLA_Syrk (BLAS3/PLA_Syrk.c:13, PLA_Syrk_panpan.c:13-95): every entry of C
   is first multiplied by beta (PLA_Local_scal, PLA_Syrk_panpan.c:22; a
   multiply, so 0 * NaN stays NaN), then the uplo triangle receives
   alpha op(A) op(A)^T in dsyrk's loop order.  op(A) = A (n x k) for
   trans 'N', A^T (A k x n) for 'T'. */
void ora_syrk(char uplo, char trans, int n, int k, double alpha, const double *a, int lda, double beta, double *c,
              int ldc) {
  int lower = (uplo == 'L' || uplo == 'l');
  int tr = (trans == 'T' || trans == 't');
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) c[(size_t)j * ldc + i] *= beta;
  for (int j = 0; j < n; j++) {
    double *cj = c + (size_t)j * ldc;
    int i0 = lower ? j : 0, i1 = lower ? n : j + 1;
    if (!tr) {
      for (int l = 0; l < k; l++) {
        double temp = alpha * A_(j, l);
        if (temp != 0.0)
          for (int i = i0; i < i1; i++) cj[i] += temp * A_(i, l);
      }
    } else {
      for (int i = i0; i < i1; i++) {
        double s = 0.0;
        for (int l = 0; l < k; l++) s += A_(l, i) * A_(l, j);
        cj[i] += alpha * s;
      }
    }
  }
}

/* PLA_LU_right, SOLVERS/LU/PLA_LU.c:37-212, on a 1x1 template. */
/* left = 1: PLA_LU_right; left = 0: PLA_LU_linpack_right
   (SOLVERS/LINPACK/PLA_LU_linpack.c:27-171), whose interchanges reach only
   A_BR, the columns from the panel on (:121-123). */
static void lu_blocked(int m, int n, double *a, int lda, int nb, int *piv, int left) {
  int mn = m < n ? m : n;
  double *panel = (double *)malloc(sizeof(double) * (size_t)m * (size_t)(nb > 0 ? nb : 1));
  for (int k = 0; k < mn; k += nb) {
    int rows = m - k;
    int b = mn - k < nb ? mn - k : nb;                     /* PLA_LU.c:106-107 */
    int ldp = rows;
    /* PLA_Copy(A_1, A_1_mv): panel copy, PLA_LU.c:115-118 */
    for (int j = 0; j < b; j++)
      memcpy(panel + (size_t)j * ldp, &A_(k, k + j), sizeof(double) * rows);
    /* PLA_LU_right_mv, SOLVERS/LU/PLA_LU_mv.c:52-82 */
    for (int j = 0; j < b; j++) {
      double *pc = panel + (size_t)j * ldp;
      int idx = ora_idamax(rows - j, pc + j, 1);            /* PLA_Iamax, :66 */
      piv[k + j] = idx;
      double alpha11 = pc[j + idx];                         /* signed xmax */
      if (idx)                                              /* Pivot_rows_mv, :71 */
        for (int c = 0; c < b; c++) {
          double *pcc = panel + (size_t)c * ldp;
          double t = pcc[j]; pcc[j] = pcc[j + idx]; pcc[j + idx] = t;
        }
      double inv = 1.0 / alpha11;                           /* PLA_Local_inv_scal, :76 */
      for (int i = j + 1; i < rows; i++) pc[i] *= inv;
      for (int c = j + 1; c < b; c++) {                     /* PLA_Local_ger, :78 */
        double *pcc = panel + (size_t)c * ldp;
        double temp = -pcc[j];
        if (temp != 0.0)
          for (int i = j + 1; i < rows; i++) pcc[i] += pc[i] * temp;
      }
    }
    /* PLA_Apply_pivots_to_rows(A_B, piv): every column of rows k..m, :152
       (LINPACK variant: columns k..n only) */
    if (left)
      ora_laswp(n, &A_(k, 0), lda, 0, b, piv + k);
    else
      ora_laswp(n - k, &A_(k, k), lda, 0, b, piv + k);
    /* panel write-back, :160-165 */
    for (int j = 0; j < b; j++)
      memcpy(&A_(k, k + j), panel + (size_t)j * ldp, sizeof(double) * rows);
    int nr = n - k - b;
    if (nr > 0) {
      /* A_12 <- inv(L_11) A_12 (unit lower), :174-182 */
      ora_trsm_left('L', 'N', 'U', b, nr, 1.0, &A_(k, k), lda, &A_(k, k + b), lda);
      /* A_22 <- A_22 - A_21 A_12, :185-186 */
      if (rows - b > 0)
        ora_gemm('N', 'N', rows - b, nr, b, -1.0, &A_(k + b, k), lda, &A_(k, k + b), lda, 1.0,
                 &A_(k + b, k + b), lda);
    }
  }
  free(panel);
}

void ora_lu_blocked(int m, int n, double *a, int lda, int nb, int *piv) { lu_blocked(m, n, a, lda, nb, piv, 1); }

void ora_lu_linpack(int m, int n, double *a, int lda, int nb, int *piv) { lu_blocked(m, n, a, lda, nb, piv, 0); }

/* ------------------------------------------------------------------ */
/* Triangular solves                                                  */

void ora_trsm_left(char uplo, char trans, char diag, int n, int nrhs, double alpha,
                   const double *a, int lda, double *b, int ldb) {
  int lower = (uplo == 'L' || uplo == 'l');
  int tr = (trans == 'T' || trans == 't');
  int unit = (diag == 'U' || diag == 'u');
  for (int c = 0; c < nrhs; c++) {
    double *x = b + (size_t)c * ldb;
    if (alpha != 1.0)
      for (int i = 0; i < n; i++) x[i] *= alpha;
    if (!tr && lower) {                     /* forward, column oriented */
      for (int k = 0; k < n; k++) {
        if (x[k] == 0.0) continue;
        if (!unit) x[k] /= A_(k, k);
        double t = x[k];
        for (int i = k + 1; i < n; i++) x[i] -= t * A_(i, k);
      }
    } else if (!tr && !lower) {             /* backward, column oriented */
      for (int k = n - 1; k >= 0; k--) {
        if (x[k] == 0.0) continue;
        if (!unit) x[k] /= A_(k, k);
        double t = x[k];
        for (int i = 0; i < k; i++) x[i] -= t * A_(i, k);
      }
    } else if (tr && lower) {               /* L^T x = b: backward, dot products */
      for (int i = n - 1; i >= 0; i--) {
        double t = x[i];
        for (int k = i + 1; k < n; k++) t -= A_(k, i) * x[k];
        if (!unit) t /= A_(i, i);
        x[i] = t;
      }
    } else {                                /* U^T x = b: forward, dot products */
      for (int i = 0; i < n; i++) {
        double t = x[i];
        for (int k = 0; k < i; k++) t -= A_(k, i) * x[k];
        if (!unit) t /= A_(i, i);
        x[i] = t;
      }
    }
  }
}

/* B <- alpha B op(A)^{-1}, A n x n, B m x n: dtrsm's right-side column loops
   (the side PLA_Trsm_right_lower / _upper take, BLAS3/PLA_Trsm_right_lower.c:13,
   BLAS3/PLA_Trsm_right_upper.c:13).  op(A) upper-type (upper untransposed,
   lower transposed) runs over the columns left to right, lower-type right to
   left; a non-unit diagonal is applied as a multiply by its reciprocal. */
void ora_trsm_right(char uplo, char trans, char diag, int m, int n, double alpha, const double *a, int lda,
                    double *b, int ldb) {
  int lower = (uplo == 'L' || uplo == 'l');
  int tr = (trans == 'T' || trans == 't');
  int unit = (diag == 'U' || diag == 'u');
#define B_(i, j) b[(size_t)(j) * ldb + (i)]
  if (alpha != 1.0)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < m; i++) B_(i, j) *= alpha;
  /* M(p, q) = op(A)(p, q) */
#define M_(p, q) (tr ? A_(q, p) : A_(p, q))
  int fwd = (lower == tr);
  for (int t = 0; t < n; t++) {
    int j = fwd ? t : n - 1 - t;
    if (!unit) {
      double r = 1.0 / M_(j, j);
      for (int i = 0; i < m; i++) B_(i, j) *= r;
    }
    int q0 = fwd ? j + 1 : 0, q1 = fwd ? n : j;
    for (int q = q0; q < q1; q++) {
      double mq = M_(j, q);
      if (mq != 0.0)
        for (int i = 0; i < m; i++) B_(i, q) -= mq * B_(i, j);
    }
  }
#undef M_
#undef B_
}

void ora_trsv(char uplo, char trans, char diag, int n, const double *a, int lda, double *x) {
  ora_trsm_left(uplo, trans, diag, n, 1, 1.0, a, lda, x, n > 0 ? n : 1);
}

/* ------------------------------------------------------------------ */
/* Cholesky                                                           */

/* PLA_dchol, SOLVERS/Cholesky/PLA_Local_chol.c:103-143 (left-looking). */
int ora_chol_unblocked(int n, double *a, int lda) {
  for (int k = 0; k < n; k++) {
    if (k > 0) {
      /* dgemv("N", n-k, k, -1, A(k,0), lda, A(k,0) row, 1, A(k,k)) */
      for (int j = 0; j < k; j++) {
        double temp = -A_(k, j);
        if (temp != 0.0)
          for (int i = k; i < n; i++) A_(i, k) += temp * A_(i, j);
      }
    }
    if (A_(k, k) <= 0.0) return k + 1;
    A_(k, k) = sqrt(A_(k, k));
    double invert = 1.0 / A_(k, k);
    for (int i = k + 1; i < n; i++) A_(i, k) *= invert;
  }
  return 0;
}

/* dtrsm("R","L","T","N"): B <- B inv(L^T), netlib loop order (reciprocal of
   the diagonal, then column axpys). */
static void trsm_rltn(int m, int nb, const double *l, int ldl, double *b, int ldb) {
  for (int k = 0; k < nb; k++) {
    double temp = 1.0 / l[(size_t)k * ldl + k];
    double *bk = b + (size_t)k * ldb;
    for (int i = 0; i < m; i++) bk[i] *= temp;
    for (int j = k + 1; j < nb; j++) {
      double ljk = l[(size_t)k * ldl + j];
      if (ljk != 0.0) {
        double *bj = b + (size_t)j * ldb;
        for (int i = 0; i < m; i++) bj[i] -= ljk * bk[i];
      }
    }
  }
}

/* PLA_Chol_lower_large, SOLVERS/Cholesky/PLA_Chol.c:159-224. */
int ora_chol_blocked(int n, double *a, int lda, int nb) {
  for (int k = 0; k < n; k += nb) {
    int b = n - k < nb ? n - k : nb;
    int info = ora_chol_unblocked(b, &A_(k, k), lda);          /* :178 */
    if (info) return k + info;                                 /* :181-185 */
    int r = n - k - b;
    if (r > 0) {
      trsm_rltn(r, b, &A_(k, k), lda, &A_(k + b, k), lda);     /* :193-195 */
      /* PLA_Syrk_perform_local_part: lower(A22) -= A21 A21^T, :207-210 */
      double *x = &A_(k + b, k);
      for (int j = 0; j < r; j++) {
        double *cj = &A_(k + b, k + b + j);
        for (int l = 0; l < b; l++) {
          double temp = -x[(size_t)l * lda + j];
          if (temp != 0.0) {
            const double *xl = x + (size_t)l * lda;
            for (int i = j; i < r; i++) cj[i] += temp * xl[i];
          }
        }
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Solves                                                             */

void ora_lu_solve(int n, int nrhs, const double *lu, int lda, const int *piv, double *b, int ldb) {
  ora_laswp(nrhs, b, ldb, 0, n, piv);                      /* PLA_General.c:91 */
  ora_trsm_left('L', 'N', 'U', n, nrhs, 1.0, lu, lda, b, ldb);
  ora_trsm_left('U', 'N', 'N', n, nrhs, 1.0, lu, lda, b, ldb);
}

void ora_chol_solve(int n, int nrhs, const double *l, int lda, double *b, int ldb) {
  ora_trsm_left('L', 'N', 'N', n, nrhs, 1.0, l, lda, b, ldb); /* PLA_Pos_def.c:75-90 */
  ora_trsm_left('L', 'T', 'N', n, nrhs, 1.0, l, lda, b, ldb);
}

/* ------------------------------------------------------------------ */
/* Application interface on one process: every block is local, so the */
/* reference's puts reduce to PLA_Obj_add_to_local_contents /          */
/* PLA_Obj_add_from_local_contents (API/PLA_API_util.c:1519).  The     */
/* reference adds the buffer unscaled; alpha here is the documented    */
/* axpy scalar (equal for alpha = 1).  The product is rounded before   */
/* the add (no fused multiply-add).                                    */

/* PLA_API_axpy_matrix_to_global, API/PLA_API.c:390-588:
   A(dr + i, dc + j) += alpha * buf(i, j). */
void ora_api_to_global(int m, int n, double alpha, const double *buf, int ldb, double *a, int lda, int dr, int dc) {
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++) {
      volatile double t = alpha * buf[(size_t)j * ldb + i];
      A_(dr + i, dc + j) += t;
    }
}

/* PLA_API_axpy_global_to_matrix, API/PLA_API.c:593-792:
   buf(i, j) += alpha * A(dr + i, dc + j). */
void ora_api_from_global(int m, int n, double alpha, const double *a, int lda, int dr, int dc, double *buf,
                         int ldb) {
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++) {
      volatile double t = alpha * A_(dr + i, dc + j);
      buf[(size_t)j * ldb + i] += t;
    }
}

/* PLA_API_multi_axpy_matrix_to_global / _global_to_matrix,
   API/PLA_API.c:797-911: sub-block (i, j) starts at local row
   sum(rsz[0..i)) and column sum(csz[0..j)) and goes to (rd[i], cd[j]);
   blocks in row-major order of (i, j). */
void ora_api_multi(int to_global, int nr, int nc, const int *rsz, const int *csz, double alpha, double *buf,
                   int ldb, double *a, int lda, const int *rd, const int *cd) {
  int li = 0;
  for (int i = 0; i < nr; i++) {
    int lj = 0;
    for (int j = 0; j < nc; j++) {
      double *sub = buf + (size_t)lj * ldb + li;
      if (to_global)
        ora_api_to_global(rsz[i], csz[j], alpha, sub, ldb, a, lda, rd[i], cd[j]);
      else
        ora_api_from_global(rsz[i], csz[j], alpha, a, lda, rd[i], cd[j], sub, ldb);
      lj += csz[j];
    }
    li += rsz[i];
  }
}

/* ------------------------------------------------------------------ */
/* MPI_DOUBLE_COMPLEX LU: PLA_LU_right / PLA_LU_right_mv on complex    */
/* data.  The pivot is izamax's (BLAS1/PLA_Iamax.c:176-230: the first   */
/* maximum of |re| + |im|), a_21 is scaled by the reciprocal of the     */
/* pivot (Local_BLAS/PLA_Local_BLAS1.c:240-250, computed here as        */
/* conj(z) / (re^2 + im^2)), the rank-1 update is zgeru                 */
/* (INCLUDE/PLA_BLAS.h:407), and the trailing steps ztrsm / zgemm.     */

typedef double complex zc;
#define Z_(i, j) z[(size_t)(j) * (size_t)lda + (size_t)(i)]

static int izamax(int n, const zc *x) {
  if (n <= 0) return -1;
  int best = 0;
  double dmax = fabs(creal(x[0])) + fabs(cimag(x[0]));
  for (int i = 1; i < n; i++) {
    double v = fabs(creal(x[i])) + fabs(cimag(x[i]));
    if (v > dmax) { dmax = v; best = i; }
  }
  return best;
}

static zc zrecip(zc v) {
  double d = creal(v) * creal(v) + cimag(v) * cimag(v);
  return creal(v) / d - I * (cimag(v) / d);
}

void ora_zlu_blocked(int m, int n, void *av, int lda, int nb, int *piv) {
  zc *z = (zc *)av;
  int mn = m < n ? m : n;
  for (int k = 0; k < mn; k += nb) {
    int rows = m - k, b = mn - k < nb ? mn - k : nb;
    /* PLA_LU_right_mv on the panel (its own columns only) */
    for (int j = 0; j < b; j++) {
      int idx = izamax(rows - j, &Z_(k + j, k + j));
      piv[k + j] = idx;
      if (idx)
        for (int c = 0; c < b; c++) {
          zc t = Z_(k + j, k + c);
          Z_(k + j, k + c) = Z_(k + j + idx, k + c);
          Z_(k + j + idx, k + c) = t;
        }
      zc inv = zrecip(Z_(k + j, k + j));
      for (int i = k + j + 1; i < m; i++) Z_(i, k + j) *= inv;
      for (int c = j + 1; c < b; c++) {
        zc temp = -Z_(k + j, k + c);
        if (temp != 0.0)
          for (int i = k + j + 1; i < m; i++) Z_(i, k + c) += Z_(i, k + j) * temp;
      }
    }
    /* interchanges on every column outside the panel */
    for (int j = 0; j < b; j++)
      if (piv[k + j]) {
        for (int c = 0; c < n; c++) {
          if (c >= k && c < k + b) continue;
          zc t = Z_(k + j, c);
          Z_(k + j, c) = Z_(k + j + piv[k + j], c);
          Z_(k + j + piv[k + j], c) = t;
        }
      }
    int nr = n - k - b;
    if (nr > 0) {
      /* A_12 <- inv(L_11) A_12 (ztrsm, unit lower, column axpys) */
      for (int c = 0; c < nr; c++)
        for (int l = 0; l < b; l++) {
          zc x = Z_(k + l, k + b + c);
          if (x != 0.0)
            for (int i = l + 1; i < b; i++) Z_(k + i, k + b + c) -= x * Z_(k + i, k + l);
        }
      /* A_22 <- A_22 - A_21 A_12 (zgemm, netlib order) */
      for (int c = 0; c < nr; c++)
        for (int l = 0; l < b; l++) {
          zc temp = -Z_(k + l, k + b + c);
          if (temp != 0.0)
            for (int i = k + b; i < m; i++) Z_(i, k + b + c) += temp * Z_(i, k + l);
        }
    }
  }
}

/* PLA_Chol_lower_large (SOLVERS/Cholesky/PLA_Chol.c:83-234) on complex data:
   PLA_zchol on the diagonal block (PLA_Local_chol.c:210-232: sqrt of the real
   diagonal, a zero imaginary part, the column scaled by the reciprocal, zher
   on the trailing part), A21 <- A21 inv(L11)^H (ztrsm "R","L","C","N"),
   A22 -= A21 A21^H (zherk; the updated diagonal keeps a zero imaginary part).
   Returns 0 or k + j + 1 for the first non-positive real diagonal (a test
   the reference's zchol leaves out and its dchol makes). */
int ora_zchol_blocked(int n, void *av, int lda, int nb) {
  zc *z = (zc *)av;
  for (int k = 0; k < n; k += nb) {
    int b = n - k < nb ? n - k : nb;
    for (int j = 0; j < b; j++) {
      double d = creal(Z_(k + j, k + j));
      if (!(d > 0.0)) return k + j + 1;
      double r = sqrt(d), inv = 1.0 / r;
      Z_(k + j, k + j) = r;
      for (int i = j + 1; i < b; i++) Z_(k + i, k + j) *= inv;
      for (int c = j + 1; c < b; c++) {
        zc y = conj(Z_(k + c, k + j));
        for (int i = c; i < b; i++) Z_(k + i, k + c) -= Z_(k + i, k + j) * y;
        Z_(k + c, k + c) = creal(Z_(k + c, k + c));
      }
    }
    int r = n - k - b;
    if (r <= 0) continue;
    for (int i = k + b; i < n; i++)   /* each row of A21: x <- x inv(L11)^H */
      for (int c = 0; c < b; c++) {
        Z_(i, k + c) /= conj(Z_(k + c, k + c));
        for (int j = c + 1; j < b; j++) Z_(i, k + j) -= conj(Z_(k + j, k + c)) * Z_(i, k + c);
      }
    for (int c = k + b; c < n; c++) { /* lower A22 -= A21 A21^H, column by column */
      for (int l = k; l < k + b; l++) {
        zc y = conj(Z_(c, l));
        if (y != 0.0)
          for (int i = c; i < n; i++) Z_(i, c) -= Z_(i, l) * y;
      }
      Z_(c, c) = creal(Z_(c, c));
    }
  }
  return 0;
}

/* PLA_General_solve on complex factors: interchanges on B, then the unit
   lower and the upper triangular solves (ztrsm column order, division by
   the diagonal). */
void ora_zlu_solve(int n, int nrhs, const void *luv, int lda, const int *piv, void *bv, int ldb) {
  const zc *z = (const zc *)luv;
  zc *bb = (zc *)bv;
  for (int c = 0; c < nrhs; c++) {
    zc *x = bb + (size_t)c * ldb;
    for (int i = 0; i < n; i++)
      if (piv[i]) { zc t = x[i]; x[i] = x[i + piv[i]]; x[i + piv[i]] = t; }
    for (int k = 0; k < n; k++)
      if (x[k] != 0.0)
        for (int i = k + 1; i < n; i++) x[i] -= x[k] * Z_(i, k);
    for (int k = n - 1; k >= 0; k--)
      if (x[k] != 0.0) {
        x[k] /= Z_(k, k);
        for (int i = 0; i < k; i++) x[i] -= x[k] * Z_(i, k);
      }
  }
}
#undef Z_
