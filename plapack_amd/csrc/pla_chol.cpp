// pla_chol.cpp -- PLA_Chol and PLA_Pos_def_solve on a 1x1 template, device resident.
//
// Reference: SOLVERS/Cholesky/PLA_Chol.c:13-234 (blocked right-looking lower
// Cholesky), SOLVERS/Cholesky/PLA_Pos_def.c:44-99 (factor + solve).
//
// Per block step k of width b (PLA_Chol.c:159-224):
//   1. back up A11, factor it (PLA_Local_chol, :178)
//   2. A21 <- A21 inv(L11)^T   (PLA_Local_trsm R/L/T/N, :193-195)
//   3. lower(A22) -= A21 A21^T (PLA_Syrk_perform_local_part, :207-210)
// The upper triangle is never read or written.  A non-positive pivot sets a
// device-side info word; every later kernel sees it and returns at once, so
// the host never waits inside the loop.  On failure the diagonal block that
// failed is restored from its backup: the reference factors a duplicate of
// A11 and breaks before copying it back (:172-185), leaving A11 as it was.
#include "pla_internal.hpp"

using namespace pla;

namespace pla {

// Returns PLA_SUCCESS and sets *info to 0 or to the 1-based index of the
// first non-positive pivot.
int chol_factor(hipStream_t s, int n, double *a, long lda, int nb, int *info_out) {
  int *info = (int *)workspace(sizeof(int), 0);
  double *backup = (double *)workspace(sizeof(double) * (size_t)nb * nb, 1);
  if (!info || !backup) return PLA_ERR_NO_MEMORY;
  PLA_HIP(hipMemsetAsync(info, 0, sizeof(int), s));
  int rc;
  for (int k = 0; k < n; k += nb) {
    const int b = n - k < nb ? n - k : nb;
    double *Akk = a + (long)k * lda + k;
    if ((rc = k_copy2d(s, b, b, Akk, lda, backup, b, info))) return rc;
    if ((rc = potrf_block(s, b, Akk, lda, k, info))) return rc;
    const int r = n - k - b;
    if (r > 0) {
      if ((rc = trsm_rltn(s, r, b, Akk, lda, Akk + b, lda, info))) return rc;
      timing_begin(s, 2);
      rc = k_syrk_lower(s, 1, r, b, -1.0, Akk + b, lda, 1.0, Akk + (long)b * lda + b, lda, info);
      timing_end(s, 2, (double)r * (double)(r + 1) * b);
      if (rc) return rc;
    }
  }
  int h_info = 0;
  PLA_HIP(hipMemcpyAsync(&h_info, info, sizeof(int), hipMemcpyDeviceToHost, s));
  PLA_HIP(hipStreamSynchronize(s));
  if (h_info > 0) {
    const int k = ((h_info - 1) / nb) * nb;
    const int b = n - k < nb ? n - k : nb;
    if ((rc = k_copy2d(s, b, b, backup, b, a + (long)k * lda + k, lda, nullptr))) return rc;
  }
  *info_out = h_info;
  return PLA_SUCCESS;
}

}  // namespace pla

extern "C" {

int PLA_Chol(int uplo, PLA_Obj A) {
  // SOLVERS/Cholesky/PLA_Chol_enter_exit.c:43-47: only the lower triangle
  if (uplo != PLA_LOWER_TRIANGULAR)
    return fail(PLA_ERR_UNSUPPORTED, "PLA_Chol: only uplo = PLA_LOWER_TRIANGULAR is implemented");
  if (!valid(A)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Chol: invalid A");
  if (A->base->datatype != MPI_DOUBLE)
    return fail(PLA_ERR_UNSUPPORTED, "PLA_Chol: only MPI_DOUBLE matrices are supported");
  const int n = A->length;
  if (A->width != n) return fail(PLA_ERR_DIMENSIONS, "PLA_Chol: A is %d x %d", A->length, A->width);
  if (n == 0) return PLA_SUCCESS;
  double *a;
  int rc = dev_ptr(A, (void **)&a);
  if (rc) return rc;
  int info = 0;
  if ((rc = chol_factor(ctx().stream, n, a, ld(A), choose_nb(n, PLA_OP_PAN_PAN), &info))) return rc;
  if ((rc = written(A))) return rc;
  return info;
}

// Factor A = L L^T, then solve L L^T X = B in place (PLA_Pos_def.c:69-90).
// Returns the PLA_Chol failure index (and leaves B untouched) if A is not
// positive definite.
int PLA_Pos_def_solve(PLA_Obj A, PLA_Obj B) {
  int rc = PLA_Chol(PLA_LOWER_TRIANGULAR, A);
  if (rc) return rc;
  if (!valid(B) || B->base->datatype != MPI_DOUBLE)
    return fail(PLA_ERR_INVALID_OBJECT, "PLA_Pos_def_solve: invalid B");
  const int n = A->length;
  if (B->length != n)
    return fail(PLA_ERR_DIMENSIONS, "PLA_Pos_def_solve: B has %d rows, A is %d", B->length, n);
  double *a, *b;
  if ((rc = dev_ptr(A, (void **)&a)) || (rc = dev_ptr(B, (void **)&b))) return rc;
  hipStream_t s = ctx().stream;
  if (B->width <= trsv_max_rhs()) {
    if ((rc = trsv_sf(s, true, false, false, n, B->width, a, ld(A), b, ld(B), nullptr))) return rc;
    if ((rc = trsv_sf(s, true, true, false, n, B->width, a, ld(A), b, ld(B), nullptr))) return rc;
  } else {
    if ((rc = trsm_left(s, true, false, false, n, B->width, a, ld(A), b, ld(B), nullptr))) return rc;
    if ((rc = trsm_left(s, true, true, false, n, B->width, a, ld(A), b, ld(B), nullptr))) return rc;
  }
  return written(B);
}

}  // extern "C"
