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
//
// One step of look-ahead, as in the LU: the column block of step k+1 is
// factored (potrf + trsm) on the look-ahead stream as soon as step k's
// update of those columns is done, while the main stream applies the rest of
// step k's SYRK.
//
// main stream, step k:  wait(panel k) -> lower(A22[:bn, :bn]) -= A21 A21^T,
//                       A22[bn:, :bn] -= A21[bn:] A21[:bn]^T -> signal
//                       -> lower(A22[bn:, bn:]) -= A21[bn:] A21[bn:]^T
// look-ahead stream:    wait(signal) -> back up + factor A11 -> A21 inv(L11)^T
int chol_factor(hipStream_t sMain, int n, double *a, long lda, int nb, int *info_out, const chol_scratch *scr) {
  if (nb > CHOL_MAX_NB) nb = CHOL_MAX_NB;
  if (nb < 1) nb = 1;
  // info: first failure index, written by the panels on sP.  seen: the
  // main stream's snapshot of info taken when panel k is done and panel k+1
  // not yet started, so step k's update runs exactly when the reference's
  // would (every panel <= k succeeded), whatever panel k+1 does meanwhile.
  int *info;
  double *backup;
  if (scr) {  // caller-provided scratch: no allocation, no kernel timing
    if ((size_t)nb * nb > scr->backup_elems)
      return fail(PLA_ERR_DIMENSIONS, "Cholesky scratch holds a %zu-element block backup, nb = %d needs %zu",
                  scr->backup_elems, nb, (size_t)nb * nb);
    info = scr->info;
    backup = scr->backup;
  } else {
    info = (int *)workspace(2 * sizeof(int), 0);
    backup = (double *)workspace(sizeof(double) * (size_t)nb * nb, 1);
    if (!info || !backup) return PLA_ERR_NO_MEMORY;
  }
  int *seen = info + 1;
  hipEvent_t ev_ready = lib_event(EV_CHOL_READY), ev_panel = lib_event(EV_CHOL_PANEL);
  if (!ev_ready || !ev_panel) return fail(PLA_ERR_NOT_INITIALIZED, "PLAPACK not yet initialized");
  const bool timed = !scr;
  PLA_HIP(hipMemsetAsync(info, 0, 2 * sizeof(int), sMain));
  int rc;
  // sP: the panels; sM: the trailing updates (CU-masked apart when
  // PLA_PARAM_CHOL_PANEL_CUS reserves CUs for the panels)
  hipStream_t sP, sM;
  if ((rc = chol_streams_begin(sMain, &sP, &sM))) return rc;
  // factor the column block that starts at diagonal position k (on sP)
  auto panel = [&](int k) -> int {
    const int b = n - k < nb ? n - k : nb;
    double *Akk = a + (long)k * lda + k;
    int rc2;
    if ((rc2 = k_copy2d(sP, b, b, Akk, lda, backup, b, info))) return rc2;
    if ((rc2 = potrf_block(sP, b, Akk, lda, k, info))) return rc2;
    const int r = n - k - b;
    if (r > 0 && (rc2 = trsm_rltn(sP, r, b, Akk, lda, Akk + b, lda, info))) return rc2;
    return PLA_SUCCESS;
  };
  auto steps = [&]() -> int {
  int rc;
  PLA_HIP(hipEventRecord(ev_ready, sM));
  PLA_HIP(hipStreamWaitEvent(sP, ev_ready, 0));
  if ((rc = panel(0))) return rc;
  PLA_HIP(hipEventRecord(ev_panel, sP));
  for (int k = 0; k < n; k += nb) {
    const int b = n - k < nb ? n - k : nb;
    PLA_HIP(hipStreamWaitEvent(sM, ev_panel, 0));
    const int r = n - k - b;
    if (r <= 0) continue;
    PLA_HIP(hipMemcpyAsync(seen, info, sizeof(int), hipMemcpyDeviceToDevice, sM));
    double *A21 = a + (long)k * lda + k + b;
    double *A22 = a + (long)(k + b) * lda + k + b;
    const int bn = r < nb ? r : nb;
    if (timed) timing_begin(sM, 2);
    rc = k_syrk_lower(sM, 1, bn, b, -1.0, A21, lda, 1.0, A22, lda, seen);
    if (!rc && r > bn)
      rc = k_gemm(sM, 1, false, true, r - bn, bn, b, -1.0, A21 + bn, lda, A21, lda, 1.0, A22 + bn, lda,
                  seen);
    if (timed) timing_end(sM, 2, (double)bn * (bn + 1) * b + 2.0 * (r - bn) * (double)bn * b);
    if (rc) return rc;
    PLA_HIP(hipEventRecord(ev_ready, sM));
    PLA_HIP(hipStreamWaitEvent(sP, ev_ready, 0));
    if ((rc = panel(k + b))) return rc;
    PLA_HIP(hipEventRecord(ev_panel, sP));
    if (r > bn) {
      const int rr = r - bn;
      if (timed) timing_begin(sM, 2);
      rc = k_syrk_lower(sM, 1, rr, b, -1.0, A21 + bn, lda, 1.0, A22 + (long)bn * lda + bn, lda, seen);
      if (timed) timing_end(sM, 2, (double)rr * (rr + 1) * b);
      if (rc) return rc;
    }
  }
  return PLA_SUCCESS;
  };
  rc = steps();
  // the caller's stream waits for both streams, also after an error
  const int rc_end = lu_streams_end(sMain, sP, sM);
  if (rc) return rc;
  if (rc_end) return rc_end;
  sM = sMain;
  int h_info = 0;
  PLA_HIP(hipMemcpyAsync(&h_info, info, sizeof(int), hipMemcpyDeviceToHost, sM));
  PLA_HIP(hipStreamSynchronize(sM));
  if (h_info > 0) {
    // the failed block's backup is the last one taken: later panels abort
    // on the info word before copying
    const int k = ((h_info - 1) / nb) * nb;
    const int b = n - k < nb ? n - k : nb;
    if ((rc = k_copy2d(sM, b, b, backup, b, a + (long)k * lda + k, lda, nullptr))) return rc;
  }
  *info_out = h_info;
  return PLA_SUCCESS;
}

}  // namespace pla

extern "C" {

int PLA_Chol(int uplo, PLA_Obj A) {
  if (is_single(A)) return on_promoted({&A}, {true}, [&] { return PLA_Chol(uplo, A); });
  // SOLVERS/Cholesky/PLA_Chol_enter_exit.c:43-47: only the lower triangle
  if (uplo != PLA_LOWER_TRIANGULAR)
    return fail(PLA_ERR_UNSUPPORTED, "PLA_Chol: only uplo = PLA_LOWER_TRIANGULAR is implemented");
  if (!valid(A)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Chol: invalid A");
  if (A->base->datatype != MPI_DOUBLE && A->base->datatype != MPI_DOUBLE_COMPLEX)
    return fail(PLA_ERR_UNSUPPORTED, "PLA_Chol: only MPI_DOUBLE and MPI_DOUBLE_COMPLEX matrices are supported");
  const int n = A->length;
  if (A->width != n) return fail(PLA_ERR_DIMENSIONS, "PLA_Chol: A is %d x %d", A->length, A->width);
  if (n == 0) return PLA_SUCCESS;
  double *a;
  int rc = dev_ptr(A, (void **)&a);
  if (rc) return rc;
  int info = 0;
  if (A->base->datatype == MPI_DOUBLE_COMPLEX) {  // A = L L^H (pla_zfactor.cpp)
    const int nbz = ctx().nb_device > 0 ? ctx().nb_device : (ctx().nb_user > 0 ? ctx().nb_user : 256);
    if ((rc = zchol_factor(ctx().stream, n, a, ld(A), nbz, &info))) return rc;
    if ((rc = written(A))) return rc;
    return info;
  }
  if ((rc = chol_factor(ctx().stream, n, a, ld(A), choose_nb(n, NB_CHOL), &info))) return rc;
  if ((rc = written(A))) return rc;
  return info;
}

// SOLVERS/Cholesky/PLA_Chol_simple.c:13: lower Cholesky with the caller's
// algorithmic block size (the reference composes PLA_Local_chol, PLA_Trsm and
// PLA_Syrk on views; the same three steps run here on the device path).
// nb_alg <= 0 leaves A unchanged, as the reference's loop then ends at once.
// A non-positive pivot returns its 1-based index, as PLA_Chol does (the
// reference's simple variant returns 0 and leaves the block it failed on
// factored as far as it got).
int PLA_Chol_simple(int nb_alg, PLA_Obj A) {
  if (is_single(A)) return on_promoted({&A}, {true}, [&] { return PLA_Chol_simple(nb_alg, A); });
  if (!valid(A)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Chol_simple: invalid A");
  if (A->base->datatype != MPI_DOUBLE && A->base->datatype != MPI_DOUBLE_COMPLEX)
    return fail(PLA_ERR_UNSUPPORTED, "PLA_Chol_simple: only MPI_DOUBLE and MPI_DOUBLE_COMPLEX matrices are supported");
  const int n = A->length;
  if (A->width != n) return fail(PLA_ERR_DIMENSIONS, "PLA_Chol_simple: A is %d x %d", A->length, A->width);
  if (n == 0 || nb_alg <= 0) return PLA_SUCCESS;
  double *a;
  int rc = dev_ptr(A, (void **)&a);
  if (rc) return rc;
  int info = 0;
  if ((rc = A->base->datatype == MPI_DOUBLE_COMPLEX ? zchol_factor(ctx().stream, n, a, ld(A), nb_alg, &info)
                                                    : chol_factor(ctx().stream, n, a, ld(A), nb_alg, &info)))
    return rc;
  if ((rc = written(A))) return rc;
  return info;
}

// Factor A = L L^T, then solve L L^T X = B in place (PLA_Pos_def.c:69-90).
// Returns the PLA_Chol failure index (and leaves B untouched) if A is not
// positive definite.
int PLA_Pos_def_solve(PLA_Obj A, PLA_Obj B) {
  if (is_single(A) || is_single(B))
    return on_promoted({&A, &B}, {true, true}, [&] { return PLA_Pos_def_solve(A, B); });
  int rc = PLA_Chol(PLA_LOWER_TRIANGULAR, A);
  if (rc) return rc;
  if (!valid(B) || B->base->datatype != A->base->datatype)
    return fail(PLA_ERR_INVALID_OBJECT, "PLA_Pos_def_solve: invalid B (or not of A's datatype)");
  const int n = A->length;
  if (B->length != n)
    return fail(PLA_ERR_DIMENSIONS, "PLA_Pos_def_solve: B has %d rows, A is %d", B->length, n);
  double *a, *b;
  if ((rc = dev_ptr(A, (void **)&a)) || (rc = dev_ptr(B, (void **)&b))) return rc;
  hipStream_t s = ctx().stream;
  if (A->base->datatype == MPI_DOUBLE_COMPLEX) {
    if ((rc = zchol_solve(s, n, B->width, a, ld(A), b, ld(B)))) return rc;
    return written(B);
  }
  if (B->width <= trsv_max_rhs()) {
    if ((rc = trsv_sf(s, true, false, false, n, B->width, a, ld(A), b, ld(B), nullptr))) return rc;
    if ((rc = trsv_sf(s, true, true, false, n, B->width, a, ld(A), b, ld(B), nullptr))) return rc;
    if ((rc = leaf_check_error(s, "PLA_Pos_def_solve: triangular solve"))) return rc;
  } else {
    if ((rc = trsm_left(s, true, false, false, n, B->width, a, ld(A), b, ld(B), nullptr))) return rc;
    if ((rc = trsm_left(s, true, true, false, n, B->width, a, ld(A), b, ld(B), nullptr))) return rc;
  }
  return written(B);
}

}  // extern "C"
