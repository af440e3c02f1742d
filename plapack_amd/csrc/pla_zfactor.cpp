// pla_zfactor.cpp -- PLA_LU and PLA_General_solve for MPI_DOUBLE_COMPLEX
// objects on a 1x1 template.
//
// Reference: SOLVERS/LU/PLA_LU.c:37-212 and SOLVERS/LU/PLA_LU_mv.c:13-101 on
// complex data (the reference's example drivers default to
// MPI_DOUBLE_COMPLEX, EXAMPLES/LU/driver.c:110), SOLVERS/LU/PLA_General.c:67;
// SOLVERS/Cholesky/PLA_Chol.c:83-234 on complex data (conjugate-transposed
// solve, Hermitian rank-k update; EXAMPLES/Cholesky/driver.c:17) and
// SOLVERS/Cholesky/PLA_Pos_def.c:44.
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

// C -= A B, A m x k, B k x n, C m x n (complex, column-major); skipped once
// *abort_flag is set
// (ws: the workspace slot of the embedded operand -- one per stream)
int zgemm_sub(hipStream_t s, int m, int n, int k, const zc *A, long lda, const zc *B, long ldb, zc *C, long ldc,
              const int *abort_flag = nullptr, int ws = WS_ZEMBED) {
  if (m <= 0 || n <= 0 || k <= 0) return PLA_SUCCESS;
  double *H = (double *)workspace(sizeof(double) * 4 * (size_t)m * k, ws);
  if (!H) return PLA_ERR_NO_MEMORY;
  int rc;
  if ((rc = k_zembed(s, m, k, A, lda, H, 2L * m))) return rc;
  return k_gemm(s, 0, false, false, 2 * m, n, 2 * k, -1.0, H, 2L * m, (const double *)B, 2 * ldb, 1.0,
                (double *)C, 2 * ldc, abort_flag);
}

// B <- inv(T) B, T n x n lower or upper, unit or not, B n x nrhs
int ztrsm_left(hipStream_t s, bool lower, bool unit, int n, int nrhs, const zc *T, long ldt, zc *B, long ldb,
               int ws = WS_ZEMBED) {
  if (n <= 0 || nrhs <= 0) return PLA_SUCCESS;
  const int leaf = ztrsm_leaf_size();
  if (n <= leaf) return k_ztrsm_leaf(s, lower, unit, n, nrhs, T, ldt, B, ldb);
  const int n0 = split_z(n, leaf), n1 = n - n0;
  const zc *T11 = T + (long)n0 * ldt + n0;
  int rc;
  if (lower) {
    if ((rc = ztrsm_left(s, true, unit, n0, nrhs, T, ldt, B, ldb, ws))) return rc;
    if ((rc = zgemm_sub(s, n1, nrhs, n0, T + n0, ldt, B, ldb, B + n0, ldb, nullptr, ws))) return rc;
    return ztrsm_left(s, true, unit, n1, nrhs, T11, ldt, B + n0, ldb, ws);
  }
  if ((rc = ztrsm_left(s, false, unit, n1, nrhs, T11, ldt, B + n0, ldb, ws))) return rc;
  if ((rc = zgemm_sub(s, n0, nrhs, n1, T + (long)n0 * ldt, ldt, B + n0, ldb, B, ldb, nullptr, ws))) return rc;
  return ztrsm_left(s, false, unit, n0, nrhs, T, ldt, B, ldb, ws);
}

// X <- X inv(L)^H, X nrows x b, L b x b lower non-unit
int ztrsm_rlc(hipStream_t s, int nrows, int b, const zc *L, long ldl, zc *X, long ldx, const int *info) {
  if (nrows <= 0 || b <= 0) return PLA_SUCCESS;
  const int leaf = ztrsm_leaf_size();
  if (b <= leaf) return k_ztrsm_rlc_leaf(s, nrows, b, L, ldl, X, ldx, info);
  const int b0 = split_z(b, leaf), b1 = b - b0;
  int rc;
  if ((rc = ztrsm_rlc(s, nrows, b0, L, ldl, X, ldx, info))) return rc;
  // X1 -= X0 L10^H, with L10^H (b0 x b1) formed first
  zc *Lh = (zc *)workspace(sizeof(zc) * (size_t)b0 * b1, WS_ZCT);
  if (!Lh) return PLA_ERR_NO_MEMORY;
  if ((rc = k_zconjtrans(s, b1, b0, L + b0, ldl, Lh, b0))) return rc;
  if ((rc = zgemm_sub(s, nrows, b1, b0, X, ldx, Lh, b0, X + (long)b0 * ldx, ldx, info))) return rc;
  return ztrsm_rlc(s, nrows, b1, L + (long)b0 * ldl + b0, ldl, X + (long)b0 * ldx, ldx, info);
}

// lower triangle of C (n x n) -= X X^H, X n x k: diagonal blocks by the
// element kernel, the blocks below them as GEMMs against a formed X0^H
int zherk_lower(hipStream_t s, int n, int k, const zc *X, long ldx, zc *C, long ldc, const int *info) {
  if (n <= 0 || k <= 0) return PLA_SUCCESS;
  if (n <= zherk_diag_max()) return k_zherk_diag(s, n, k, X, ldx, C, ldc, info);
  const int n0 = split_z(n, 128), n1 = n - n0;
  int rc;
  zc *Xh = (zc *)workspace(sizeof(zc) * (size_t)k * n0, WS_ZCT);
  if (!Xh) return PLA_ERR_NO_MEMORY;
  if ((rc = k_zconjtrans(s, n0, k, X, ldx, Xh, k))) return rc;
  if ((rc = zgemm_sub(s, n1, n0, k, X + n0, ldx, Xh, k, C + n0, ldc, info))) return rc;
  if ((rc = zherk_lower(s, n0, k, X, ldx, C, ldc, info))) return rc;
  return zherk_lower(s, n1, k, X + n0, ldx, C + (long)n0 * ldc + n0, ldc, info);
}

// recursive lower Cholesky of a b x b diagonal block; failures go to *info
int zpotrf_block(hipStream_t s, int b, zc *A, long lda, int offset, int *info) {
  if (b <= 0) return PLA_SUCCESS;
  const int leaf = zpotrf_leaf_size();
  if (b <= leaf) return k_zpotrf_leaf(s, b, A, lda, offset, info);
  const int b0 = split_z(b, leaf), b1 = b - b0;
  int rc;
  if ((rc = zpotrf_block(s, b0, A, lda, offset, info))) return rc;
  if ((rc = ztrsm_rlc(s, b1, b0, A, lda, A + b0, lda, info))) return rc;
  if ((rc = zherk_lower(s, b1, b0, A + b0, lda, A + (long)b0 * lda + b0, lda, info))) return rc;
  return zpotrf_block(s, b1, A + (long)b0 * lda + b0, lda, offset + b0, info);
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
  if ((rc = ztrsm_left(s, true, true, w0, w1, P, lda, Pr, lda, WS_ZEMBED_P))) return rc;
  if ((rc = zgemm_sub(s, m - w0, w1, w0, P + w0, lda, Pr, lda, Pr + w0, lda, nullptr, WS_ZEMBED_P))) return rc;
  return zpanel_rec(s, panel, mp, pw, lda, piv, c0 + w0, w1);
}

}  // namespace

int zgemm_minus(hipStream_t s, int m, int n, int k, const void *A, long lda, const void *B, long ldb, void *C,
                long ldc, int ws) {
  return zgemm_sub(s, m, n, k, (const zc *)A, lda, (const zc *)B, ldb, (zc *)C, ldc, nullptr, ws);
}

int ztrsm_left_n(hipStream_t s, bool lower, bool unit, int n, int nrhs, const void *T, long ldt, void *B, long ldb,
                 int ws) {
  return ztrsm_left(s, lower, unit, n, nrhs, (const zc *)T, ldt, (zc *)B, ldb, ws);
}

// Blocked right-looking LU with one step of look-ahead, as lu_factor: the
// panel of step k+1 is factored on the look-ahead stream (its GEMMs on their
// own operand workspace) while the trailing stream finishes step k's update;
// the next panel's columns are interchanged, solved and updated first.  The
// interchanges left of a panel go last in their step, after the update that
// reads the previous panel's L21.
int zlu_factor(hipStream_t sMain, int m, int n, void *av, long lda, int *piv, int nb) {
  zc *a = (zc *)av;
  const int mn = m < n ? m : n;
  if (nb < 1) nb = 1;
  hipEvent_t ev_ready = lib_event(EV_ZLU_READY), ev_panel = lib_event(EV_ZLU_PANEL);
  if (!ev_ready || !ev_panel) return fail(PLA_ERR_NOT_INITIALIZED, "PLAPACK not yet initialized");
  hipStream_t sP, sM;
  int gcap, rc;
  if ((rc = lu_streams_begin(sMain, m, &sP, &sM, &gcap))) return rc;
  auto width_at = [&](int k) { return mn - k < nb ? mn - k : nb; };
  auto panel = [&](int k, int b) -> int {
    return zpanel_rec(sP, a + (long)k * lda + k, m - k, b, lda, piv + k, 0, b);
  };
  auto steps = [&]() -> int {
    int r;
    if ((r = panel(0, width_at(0)))) return r;
    PLA_HIP(hipEventRecord(ev_panel, sP));
    for (int k = 0, b; k < mn; k += b) {
      b = width_at(k);
      PLA_HIP(hipStreamWaitEvent(sM, ev_panel, 0));
      zc *Akk = a + (long)k * lda + k;
      const int kn = k + b, nr = n - kn, mr = m - kn;
      const int bn = (kn < mn && mr > 0) ? width_at(kn) : 0;
      const int n1 = bn > 0 ? bn : nr;
      auto right_part = [&](int c0, int nc) -> int {
        if (nc <= 0) return PLA_SUCCESS;
        zc *A12 = Akk + (long)(b + c0) * lda;
        int q;
        if ((q = k_zlaswp(sM, A12, lda, m - k, nc, b, piv + k))) return q;
        if ((q = ztrsm_left(sM, true, true, b, nc, Akk, lda, A12, lda))) return q;
        return mr > 0 ? zgemm_sub(sM, mr, nc, b, Akk + b, lda, A12, lda, A12 + b, lda) : PLA_SUCCESS;
      };
      if ((r = right_part(0, n1))) return r;
      if (bn > 0) {
        PLA_HIP(hipEventRecord(ev_ready, sM));
        PLA_HIP(hipStreamWaitEvent(sP, ev_ready, 0));
        if ((r = panel(kn, bn))) return r;
        PLA_HIP(hipEventRecord(ev_panel, sP));
      }
      if ((r = right_part(n1, nr - n1))) return r;
      if (k > 0 && (r = k_zlaswp(sM, a + k, lda, m - k, k, b, piv + k))) return r;
    }
    return PLA_SUCCESS;
  };
  rc = steps();
  const int rc_end = lu_streams_end(sMain, sP, sM);
  return rc ? rc : rc_end;
}

int zlu_solve(hipStream_t s, int n, int nrhs, const void *lu, long lda, const int *piv, void *b, long ldb) {
  int rc;
  if ((rc = k_zlaswp(s, b, ldb, n, nrhs, n, piv))) return rc;
  if ((rc = ztrsm_left(s, true, true, n, nrhs, (const zc *)lu, lda, (zc *)b, ldb))) return rc;
  return ztrsm_left(s, false, false, n, nrhs, (const zc *)lu, lda, (zc *)b, ldb);
}

// Blocked lower Cholesky (PLA_Chol_lower_large on complex data): per block,
// the diagonal block (recursive, PLA_zchol leaves), A21 <- A21 inv(L11)^H,
// A22 -= A21 A21^H.  *info_out: 0, or the 1-based index of the first
// non-positive diagonal, with that diagonal block restored to its value
// before its factorization (the real path's convention).
int zchol_factor(hipStream_t s, int n, void *av, long lda, int nb, int *info_out) {
  zc *a = (zc *)av;
  if (nb < 1) nb = 1;
  if (nb > 1024) nb = 1024;
  char *w = (char *)workspace(256 + sizeof(zc) * (size_t)nb * nb, WS_ZINFO);
  if (!w) return PLA_ERR_NO_MEMORY;
  int *info = (int *)w;
  zc *backup = (zc *)(w + 256);
  PLA_HIP(hipMemsetAsync(info, 0, sizeof(int), s));
  int rc;
  for (int k = 0, b; k < n; k += b) {
    b = n - k < nb ? n - k : nb;
    zc *Akk = a + (long)k * lda + k;
    // backup of the diagonal block (as the 2b x b real block of its doubles)
    if ((rc = k_copy2d(s, 2 * b, b, (const double *)Akk, 2 * lda, (double *)backup, 2L * b, info))) return rc;
    if ((rc = zpotrf_block(s, b, Akk, lda, k, info))) return rc;
    const int r = n - k - b;
    if (r <= 0) continue;
    if ((rc = ztrsm_rlc(s, r, b, Akk, lda, Akk + b, lda, info))) return rc;
    if ((rc = zherk_lower(s, r, b, Akk + b, lda, Akk + (long)b * lda + b, lda, info))) return rc;
  }
  int h_info = 0;
  PLA_HIP(hipMemcpyAsync(&h_info, info, sizeof(int), hipMemcpyDeviceToHost, s));
  PLA_HIP(hipStreamSynchronize(s));
  if (h_info > 0) {
    const int k = ((h_info - 1) / nb) * nb, b = n - k < nb ? n - k : nb;
    if ((rc = k_copy2d(s, 2 * b, b, (const double *)backup, 2L * b, (double *)(a + (long)k * lda + k), 2 * lda,
                       nullptr)))
      return rc;
  }
  *info_out = h_info;
  return PLA_SUCCESS;
}

// L L^H X = B: the lower solve, then the upper solve with L^H formed
int zchol_solve(hipStream_t s, int n, int nrhs, const void *l, long lda, void *b, long ldb) {
  int rc;
  if ((rc = ztrsm_left(s, true, false, n, nrhs, (const zc *)l, lda, (zc *)b, ldb))) return rc;
  zc *U = (zc *)workspace(sizeof(zc) * (size_t)n * n, WS_ZCT);
  if (!U) return PLA_ERR_NO_MEMORY;
  if ((rc = k_zconjtrans(s, n, n, l, lda, U, n))) return rc;
  return ztrsm_left(s, false, false, n, nrhs, U, n, (zc *)b, ldb);
}

}  // namespace pla
