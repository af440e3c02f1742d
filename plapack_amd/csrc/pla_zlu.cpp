// pla_zlu.cpp -- PLA_LU and PLA_General_solve for MPI_DOUBLE_COMPLEX
// objects on a 1x1 template.
//
// Reference: SOLVERS/LU/PLA_LU.c:37-212 and SOLVERS/LU/PLA_LU_mv.c:13-101 on
// complex data (the reference's example drivers default to
// MPI_DOUBLE_COMPLEX, EXAMPLES/LU/driver.c:110), SOLVERS/LU/PLA_General.c:67.
//
// Same blocked right-looking algorithm as the real path, on one stream: a
// recursive panel over one-workgroup leaves (pla_complex.hip), interchanges,
// A12 <- inv(L11) A12, A22 -= A21 A12.  Every GEMM -- the trailing update,
// the panel recursion's and the triangular solves' -- is the FP64 MFMA
// kernel on the real embedding of its left operand (a complex m x k block A
// becomes the real 2m x 2k [[re, -im], [im, re]] per element; the right
// operand and the result are their own doubles, 2 ld apart), so the O(n^3)
// work runs at the real GEMM's rate.
#include "pla_internal.hpp"

namespace pla {

namespace {

struct zc {
  double re, im;
};

int split_z(int n, int q) {
  int h = ((n / 2 + q - 1) / q) * q;
  if (h >= n) h = n - q > 0 ? n - q : n / 2;
  if (h <= 0) h = 1;
  return h;
}

// C -= A B, A m x k, B k x n, C m x n (complex, column-major)
int zgemm_sub(hipStream_t s, int m, int n, int k, const zc *A, long lda, const zc *B, long ldb, zc *C, long ldc) {
  if (m <= 0 || n <= 0 || k <= 0) return PLA_SUCCESS;
  double *H = (double *)workspace(sizeof(double) * 4 * (size_t)m * k, WS_ZEMBED);
  if (!H) return PLA_ERR_NO_MEMORY;
  int rc;
  if ((rc = k_zembed(s, m, k, A, lda, H, 2L * m))) return rc;
  return k_gemm(s, 0, false, false, 2 * m, n, 2 * k, -1.0, H, 2L * m, (const double *)B, 2 * ldb, 1.0,
                (double *)C, 2 * ldc, nullptr);
}

// B <- inv(T) B, T n x n lower unit (lower) or upper non-unit, B n x nrhs
int ztrsm_left(hipStream_t s, bool lower, int n, int nrhs, const zc *T, long ldt, zc *B, long ldb) {
  if (n <= 0 || nrhs <= 0) return PLA_SUCCESS;
  const int leaf = ztrsm_leaf_size();
  if (n <= leaf) return k_ztrsm_leaf(s, lower, n, nrhs, T, ldt, B, ldb);
  const int n0 = split_z(n, leaf), n1 = n - n0;
  const zc *T11 = T + (long)n0 * ldt + n0;
  int rc;
  if (lower) {
    if ((rc = ztrsm_left(s, true, n0, nrhs, T, ldt, B, ldb))) return rc;
    if ((rc = zgemm_sub(s, n1, nrhs, n0, T + n0, ldt, B, ldb, B + n0, ldb))) return rc;
    return ztrsm_left(s, true, n1, nrhs, T11, ldt, B + n0, ldb);
  }
  if ((rc = ztrsm_left(s, false, n1, nrhs, T11, ldt, B + n0, ldb))) return rc;
  if ((rc = zgemm_sub(s, n0, nrhs, n1, T + (long)n0 * ldt, ldt, B + n0, ldb, B, ldb))) return rc;
  return ztrsm_left(s, false, n0, nrhs, T, ldt, B, ldb);
}

// Recursive panel LU of the sub-panel whose first row and column are c0 in an
// mp x pw panel; the leaves interchange rows across all pw columns.
int zpanel_rec(hipStream_t s, zc *panel, int mp, int pw, long lda, int *piv, int c0, int w) {
  zc *P = panel + (long)c0 * lda + c0;
  const int m = mp - c0;
  const int leaf = zleaf_width();
  if (w <= leaf) return k_zleaf(s, P, lda, m, w, pw, c0, piv + c0);
  const int w0 = split_z(w, leaf), w1 = w - w0;
  int rc;
  if ((rc = zpanel_rec(s, panel, mp, pw, lda, piv, c0, w0))) return rc;
  zc *Pr = P + (long)w0 * lda;
  if ((rc = ztrsm_left(s, true, w0, w1, P, lda, Pr, lda))) return rc;
  if ((rc = zgemm_sub(s, m - w0, w1, w0, P + w0, lda, Pr, lda, Pr + w0, lda))) return rc;
  return zpanel_rec(s, panel, mp, pw, lda, piv, c0 + w0, w1);
}

}  // namespace

int zlu_factor(hipStream_t s, int m, int n, void *av, long lda, int *piv, int nb) {
  zc *a = (zc *)av;
  const int mn = m < n ? m : n;
  if (nb < 1) nb = 1;
  for (int k = 0, b; k < mn; k += b) {
    b = mn - k < nb ? mn - k : nb;
    zc *Akk = a + (long)k * lda + k;
    int rc;
    if ((rc = zpanel_rec(s, Akk, m - k, b, lda, piv + k, 0, b))) return rc;
    const int nr = n - k - b, mr = m - k - b;
    if (k > 0 && (rc = k_zlaswp(s, a + k, lda, m - k, k, b, piv + k))) return rc;
    if (nr <= 0) continue;
    zc *A12 = Akk + (long)b * lda;
    if ((rc = k_zlaswp(s, A12, lda, m - k, nr, b, piv + k))) return rc;
    if ((rc = ztrsm_left(s, true, b, nr, Akk, lda, A12, lda))) return rc;
    if (mr > 0 && (rc = zgemm_sub(s, mr, nr, b, Akk + b, lda, A12, lda, A12 + b, lda))) return rc;
  }
  return PLA_SUCCESS;
}

int zlu_solve(hipStream_t s, int n, int nrhs, const void *lu, long lda, const int *piv, void *b, long ldb) {
  int rc;
  if ((rc = k_zlaswp(s, b, ldb, n, nrhs, n, piv))) return rc;
  if ((rc = ztrsm_left(s, true, n, nrhs, (const zc *)lu, lda, (zc *)b, ldb))) return rc;
  return ztrsm_left(s, false, n, nrhs, (const zc *)lu, lda, (zc *)b, ldb);
}

}  // namespace pla
