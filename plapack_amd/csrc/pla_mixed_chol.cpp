// pla_mixed_chol.cpp -- PLA_Chol_mixed: Cholesky with an FP8 (E4M3, per-
// 256x256-tile power-of-two scales) MFMA trailing update, FP32 storage,
// FP64 panels, then FP64 refinement (classic, falling back to GMRES-IR).
//
// The blocked algorithm is PLA_Chol_lower_large's (SOLVERS/Cholesky/
// PLA_Chol.c:83-234): per block step, factor the diagonal block, solve the
// panel below it, update the trailing lower triangle.  Here the step-k
// panel is factored in FP64 (potrf_block + trsm_rltn on an FP64 copy of the
// FP32 column block), stored back in FP32 for the solves, and quantized to
// E4M3 for the SYRK.  If the FP8-perturbed matrix is not numerically positive
// definite, the factorization is repeated for A + sigma I with a growing
// shift; refinement then runs against the unshifted A.
#include "pla_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

using namespace pla;

namespace {

refine_stats g_last_stats;
double g_last_shift = 0.0;

// FP8-updated blocked Cholesky of W (FP32, lower).  *info: 0 or the 1-based
// index of the first non-positive pivot.
int chol_factor_fp8(hipStream_t s, int n, float *W, long ldw, int *info_out) {
  const int nb = fp8_tile_rows();
  double *P64 = (double *)workspace(sizeof(double) * (size_t)n * nb, WS_CH_P64);
  unsigned char *Lq = (unsigned char *)workspace((size_t)n * nb + 256, WS_CH_LQ);
  int *exps = (int *)workspace(sizeof(int) * ((size_t)n / nb + 2), WS_CH_EXP);
  int *info = (int *)workspace(sizeof(int) * 4, WS_CH_INFO);
  if (!P64 || !Lq || !exps || !info) return PLA_ERR_NO_MEMORY;
  PLA_HIP(hipMemsetAsync(info, 0, sizeof(int), s));
  int rc;
  for (int k = 0; k < n; k += nb) {
    const int b = std::min(nb, n - k), R = n - k - b;
    const long ldp = n - k;
    float *Wk = W + (long)k * ldw + k;
    if ((rc = k_cvt_f32_f64(s, n - k, b, Wk, ldw, P64, ldp))) return rc;
    if ((rc = potrf_block(s, b, P64, ldp, k, info))) return rc;
    if (R > 0 && (rc = trsm_rltn(s, R, b, P64, ldp, P64 + b, ldp, info))) return rc;
    if ((rc = k_cvt_f64_f32(s, n - k, b, P64, ldp, Wk, ldw))) return rc;
    if (R > 0) {
      if ((rc = k_fp8_quantize(s, R, b, P64 + b, ldp, Lq, exps))) return rc;
      timing_begin(s, 4);
      rc = k_fp8_syrk_lower(s, R, Lq, exps, W + (long)(k + b) * ldw + k + b, ldw, info);
      timing_end(s, 4, (double)R * (R + 1) * b);
      if (rc) return rc;
    }
  }
  PLA_HIP(hipMemcpyAsync(info_out, info, sizeof(int), hipMemcpyDeviceToHost, s));
  PLA_HIP(hipStreamSynchronize(s));
  return PLA_SUCCESS;
}

}  // namespace

namespace pla {
void set_last_refine(const refine_stats &st, double shift) {
  g_last_stats = st;
  g_last_shift = shift;
}
}  // namespace pla

extern "C" {

int PLA_Chol_mixed(PLA_Obj A, PLA_Obj B, PLA_Obj X, int precision, double tol, int max_iter, int *iterations,
                   double *backward_error) {
  if (!valid(A) || !valid(B) || !valid(X)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Chol_mixed: invalid object");
  if (A->base->datatype != MPI_DOUBLE || B->base->datatype != MPI_DOUBLE || X->base->datatype != MPI_DOUBLE)
    return fail(PLA_ERR_UNSUPPORTED, "PLA_Chol_mixed: A, B and X must be MPI_DOUBLE");
  if (precision != PLA_MIXED_FP8) return fail(PLA_ERR_INVALID_PARAM, "PLA_Chol_mixed: precision %d", precision);
  const int n = A->length, nrhs = B->width;
  if (A->width != n || B->length != n || X->length != n || X->width != nrhs)
    return fail(PLA_ERR_DIMENSIONS, "PLA_Chol_mixed: A %dx%d, B %dx%d, X %dx%d", A->length, A->width, B->length,
                B->width, X->length, X->width);
  if (nrhs < 1 || nrhs > trsv_max_rhs())
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Chol_mixed: %d right-hand sides (1..%d)", nrhs, trsv_max_rhs());
  if (max_iter < 0) return fail(PLA_ERR_INVALID_PARAM, "PLA_Chol_mixed: max_iter < 0");
  if (n == 0) return PLA_SUCCESS;
  hipStream_t s = ctx().stream;
  double *a, *b, *x;
  int rc;
  if ((rc = dev_ptr(A, (void **)&a)) || (rc = dev_ptr(B, (void **)&b)) || (rc = dev_ptr(X, (void **)&x)))
    return rc;
  const long lda = ld(A), ldb = ld(B), ldx = ld(X);
  const long ldw = n;
  float *W = (float *)workspace(sizeof(float) * (size_t)n * n, WS_CH_W);
  if (!W) return PLA_ERR_NO_MEMORY;
  double anorm = 0.0;
  if ((rc = sym_inf_norm(s, n, a, lda, &anorm))) return rc;
  // factor A + shift I, growing the shift until the FP8-updated factorization
  // completes (E4M3 unit roundoff 2^-4)
  double shift = 0.0;
  int info = 0;
  for (;;) {
    if ((rc = k_cvt_lower_shift(s, n, a, lda, shift, W, ldw))) return rc;
    if ((rc = chol_factor_fp8(s, n, W, ldw, &info))) return rc;
    if (info == 0) break;
    // a shift beyond ||A|| / 4 is far outside the FP8 perturbation of an SPD
    // matrix (unit roundoff 2^-4): A itself is not positive definite
    const double next = shift == 0.0 ? std::ldexp(anorm, -8) : 4.0 * shift;
    if (!(anorm > 0.0) || next > 0.25 * anorm)
      return fail(PLA_ERR_INVALID_PARAM, "PLA_Chol_mixed: A is not positive definite (pivot %d, shift %g)", info,
                  shift);
    shift = next;
  }
  const int nblk = (n + trsv_block_size() - 1) / trsv_block_size();
  const size_t dsz = sizeof(double) * (size_t)nblk * trsv_block_size() * trsv_block_size();
  double *DinvL = (double *)workspace(dsz, WS_CH_DINVL);
  double *DinvLT = (double *)workspace(dsz, WS_CH_DINVLT);
  if (!DinvL || !DinvLT) return PLA_ERR_NO_MEMORY;
  if ((rc = k_inv_diag_f32_op(s, true, false, false, n, W, ldw, DinvL))) return rc;
  if ((rc = k_inv_diag_f32_op(s, true, true, false, n, W, ldw, DinvLT))) return rc;
  refine_ops op;
  op.n = n;
  op.nrhs = nrhs;
  op.residual = [&](const double *Xv, long ldxv, const double *Bv, long ldbv, double *R, long ldr, double *rmax) {
    return sym_residual(s, n, nrhs, a, lda, Xv, ldxv, Bv, ldbv, R, ldr, rmax);
  };
  op.precond = [&](const double *Y, long ldy, double *Z, long ldz) -> int {
    PLA_HIP(hipMemcpy2DAsync(Z, sizeof(double) * ldz, Y, sizeof(double) * ldy, sizeof(double) * n, nrhs,
                             hipMemcpyDeviceToDevice, s));
    int r2;
    if ((r2 = trsv_sf_f32(s, true, false, false, n, nrhs, W, ldw, Z, ldz, DinvL))) return r2;
    return trsv_sf_f32(s, true, true, false, n, nrhs, W, ldw, Z, ldz, DinvLT);
  };
  refine_stats st;
  if ((rc = refine(s, op, anorm, b, ldb, x, ldx, tol, max_iter, true, &st))) return rc;
  if ((rc = leaf_check_error(s))) return rc;
  set_last_refine(st, shift);
  if (iterations) *iterations = st.outer;
  if (backward_error) *backward_error = st.eta;
  if ((rc = written(X))) return rc;
  return refine_outcome("PLA_Chol_mixed", st, tol);
}

int PLA_Mixed_stats(int *outer, int *gmres_steps, int *gmres_corrections, double *shift) {
  if (outer) *outer = g_last_stats.outer;
  if (gmres_steps) *gmres_steps = g_last_stats.inner;
  if (gmres_corrections) *gmres_corrections = g_last_stats.gmres_outer;
  if (shift) *shift = g_last_shift;
  return PLA_SUCCESS;
}

}  // extern "C"
