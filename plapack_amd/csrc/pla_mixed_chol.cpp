// pla_mixed_chol.cpp -- PLA_Chol_mixed: Cholesky with an FP8 (E4M3, per-
// 256x256-tile power-of-two scales) MFMA trailing update, FP32 storage,
// FP64 panels, then FP64 refinement (classic, falling back to GMRES-IR).
//
// The blocked algorithm is PLA_Chol_lower_large's (SOLVERS/Cholesky/
// PLA_Chol.c:83-234): per block step, factor the diagonal block, solve the
// panel below it, update the trailing lower triangle.  Here the step-k
// panel is factored in FP64 (potrf_block + trsm_rltn on an FP64 copy of the
// FP32 column block), stored back in FP32 for the solves, and quantized to
// E4M3 for the update.  If the FP8-perturbed matrix is not numerically
// positive definite, the factorization is repeated for A + sigma I with a
// growing shift; refinement then runs against the unshifted A.
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
//
// One step of look-ahead, as in chol_factor: the column block of step k+1
// is factored on the panel stream as soon as step k's update of those
// columns is done, while the update stream applies the rest of step k's
// update.  The quantized panel and its exponents are double buffered by step
// parity (step k's update may still read them while panel k+1 is quantized).
//
// update stream, step k: wait(panel k) -> lower(C[:bn, :bn]), C[bn:, :bn]
//                        -> signal -> lower(C[bn:, bn:])
// panel stream:          wait(signal) -> FP64 copy of the diagonal block ->
//                        potrf -> inverse -> [wait A21] GEMM -> quantize ->
//                        signal(panel k+1)
// side stream:           A21 to FP64 (beside potrf), X21 to FP32 (beside the
//                        quantization)
int chol_factor_fp8(hipStream_t sMain, int n, float *W, long ldw, int *info_out) {
  const int nb = (ctx().fp8_nb == 256 || ctx().fp8_nb == 1024) ? ctx().fp8_nb : 512;
  const int kq = fp8_row_bytes(nb), QT = fp8_tile_rows();
  const size_t lq_stride = (size_t)n * kq + 256;
  const size_t ex_stride = (size_t)(n / QT + 2) * (kq / QT);
  double *P64 = (double *)workspace(sizeof(double) * (size_t)n * nb, WS_CH_P64);
  double *Linv = (double *)workspace(sizeof(double) * (size_t)nb * nb, WS_CH_LINV);
  double *X21 = (double *)workspace(sizeof(double) * (size_t)n * nb, WS_CH_OUT);
  unsigned char *Lq = (unsigned char *)workspace(2 * lq_stride, WS_CH_LQ);
  int *exps = (int *)workspace(sizeof(int) * 2 * ex_stride, WS_CH_EXP);
  int *info = (int *)workspace(sizeof(int) * 4, WS_CH_INFO);
  if (!P64 || !Linv || !X21 || !Lq || !exps || !info) return PLA_ERR_NO_MEMORY;
  hipEvent_t ev_ready = lib_event(EV_CHOL_READY), ev_panel = lib_event(EV_CHOL_PANEL);
  if (!ev_ready || !ev_panel) return fail(PLA_ERR_NOT_INITIALIZED, "PLAPACK not yet initialized");
  PLA_HIP(hipMemsetAsync(info, 0, sizeof(int), sMain));
  hipStream_t sP, sM;
  int rc;
  if ((rc = chol_streams_begin(sMain, &sP, &sM, ctx().fp8_masked != 0))) return rc;
  // The conversions of the rows below the diagonal block (A21 to FP64 before
  // the panel GEMM, X21 back to FP32 after it) run on a side stream, beside the
  // diagonal block's factorization and the quantization: neither is needed by
  // the next step's update.  Ordering: the side stream starts a step after the
  // panel stream has waited for the update of the step's columns; the panel
  // GEMM waits for the A21 conversion, which (side-stream order) also follows
  // the previous step's X21 conversion, so X21 is not rewritten while read; the
  // main stream waits for the side stream at the end.
  const bool side = ctx().fp8_side_cvt && sP != sM;
  hipStream_t sX = side ? ctx().stream_tr : sP;
  hipEvent_t ev_side = lib_event(EV_FP8_SIDE), ev_a21 = lib_event(EV_FP8_A21), ev_x21 = lib_event(EV_FP8_X21),
             ev_out = lib_event(EV_FP8_OUT);
  if (side && (!sX || !ev_side || !ev_a21 || !ev_x21 || !ev_out))
    return fail(PLA_ERR_NOT_INITIALIZED, "PLAPACK not yet initialized");
  // the column block at diagonal position k, quantized into parity par.
  // L21 = A21 inv(L11)^T is one GEMM against the explicitly inverted L11
  // (inv(L11) = L11 \ I): the triangular solve's chain of 64-column leaves
  // would sit on the critical path of every step, and the factor is only a
  // preconditioner, whose rounding the refinement removes.
  auto panel = [&](int k, int par) -> int {
    const int b = std::min(nb, n - k), R = n - k - b;
    const long ldp = n - k;
    float *Wk = W + (long)k * ldw + k;
    int r2;
    if (side && R > 0) {
      PLA_HIP(hipEventRecord(ev_side, sP));
      PLA_HIP(hipStreamWaitEvent(sX, ev_side, 0));
      if ((r2 = k_cvt_f32_f64(sX, R, b, Wk + b, ldw, P64 + b, ldp))) return r2;
      PLA_HIP(hipEventRecord(ev_a21, sX));
      if ((r2 = k_cvt_f32_f64(sP, b, b, Wk, ldw, P64, ldp))) return r2;
    } else if ((r2 = k_cvt_f32_f64(sP, n - k, b, Wk, ldw, P64, ldp))) {
      return r2;
    }
    if ((r2 = potrf_block(sP, b, P64, ldp, k, info))) return r2;
    if ((r2 = k_cvt_f64_f32(sP, b, b, P64, ldp, Wk, ldw))) return r2;
    if (R <= 0) return PLA_SUCCESS;
    if ((r2 = k_eye(sP, b, Linv, b))) return r2;
    if ((r2 = trsm_left(sP, true, false, false, b, b, P64, ldp, Linv, b, info))) return r2;
    if (side) PLA_HIP(hipStreamWaitEvent(sP, ev_a21, 0));
    if ((r2 = k_gemm(sP, 0, false, true, R, b, b, 1.0, P64 + b, ldp, Linv, b, 0.0, X21, R, info))) return r2;
    if (side) {
      PLA_HIP(hipEventRecord(ev_x21, sP));
      PLA_HIP(hipStreamWaitEvent(sX, ev_x21, 0));
    }
    if ((r2 = k_cvt_f64_f32(sX, R, b, X21, R, Wk + b, ldw))) return r2;
    return k_fp8_quantize(sP, R, b, X21, R, Lq + par * lq_stride, exps + par * ex_stride);
  };
  // update of the trailing block's columns c0 .. c0 + nc (lower part) by panel k
  auto update = [&](int k, int par, int c0, int nc, bool to_bottom) -> int {
    const int b = std::min(nb, n - k), R = n - k - b;
    const unsigned char *L = Lq + par * lq_stride;
    const int *E = exps + par * ex_stride;
    float *C = W + (long)(k + b + c0) * ldw + k + b + c0;
    int r2;
    timing_begin(sM, 4);
    r2 = k_fp8_update(sM, true, nc, nc, b, L, E, c0, c0, C, ldw, info);
    if (!r2 && to_bottom && R > c0 + nc)
      r2 = k_fp8_update(sM, false, R - c0 - nc, nc, b, L, E, c0 + nc, c0, C + nc, ldw, info);
    const double rows = to_bottom ? (double)(R - c0) : (double)nc;
    timing_end(sM, 4, (double)nc * (nc + 1) * b + 2.0 * (rows - nc) * nc * b);
    return r2;
  };
  auto steps = [&]() -> int {
    int r2;
    PLA_HIP(hipEventRecord(ev_ready, sM));
    PLA_HIP(hipStreamWaitEvent(sP, ev_ready, 0));
    if ((r2 = panel(0, 0))) return r2;
    PLA_HIP(hipEventRecord(ev_panel, sP));
    for (int k = 0, step = 0; k < n; k += nb, step++) {
      const int b = std::min(nb, n - k), R = n - k - b, par = step & 1;
      PLA_HIP(hipStreamWaitEvent(sM, ev_panel, 0));
      if (R <= 0) continue;
      const int bn = std::min(nb, R);  // the next panel's columns first
      if ((r2 = update(k, par, 0, bn, true))) return r2;
      PLA_HIP(hipEventRecord(ev_ready, sM));
      PLA_HIP(hipStreamWaitEvent(sP, ev_ready, 0));
      if ((r2 = panel(k + b, par ^ 1))) return r2;
      PLA_HIP(hipEventRecord(ev_panel, sP));
      if (R > bn && (r2 = update(k, par, bn, R - bn, false))) return r2;
    }
    return PLA_SUCCESS;
  };
  rc = steps();
  if (side) {
    PLA_HIP(hipEventRecord(ev_out, sX));
    PLA_HIP(hipStreamWaitEvent(sMain, ev_out, 0));
  }
  const int rc_end = lu_streams_end(sMain, sP, sM);
  if (rc) return rc;
  if (rc_end) return rc_end;
  PLA_HIP(hipMemcpyAsync(info_out, info, sizeof(int), hipMemcpyDeviceToHost, sMain));
  PLA_HIP(hipStreamSynchronize(sMain));
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
  double *dmax_d = (double *)workspace(sizeof(double), WS_CH_D);
  if (!dmax_d) return PLA_ERR_NO_MEMORY;
  double dmax = 0.0;
  if ((rc = k_max_diag(s, n, a, lda, dmax_d))) return rc;
  PLA_HIP(hipMemcpyAsync(&dmax, dmax_d, sizeof(double), hipMemcpyDeviceToHost, s));
  PLA_HIP(hipStreamSynchronize(s));
  // Factor A + shift I, doubling the shift until the FP8-updated
  // factorization completes.  The quantization error of L21 (relative
  // 2^-4 per element) perturbs the Schur complements by a small multiple of
  // 2^-4 max_i A(i, i) in norm at most; after an unshifted attempt fails the
  // shifts start at 2^-6 max_i A(i, i) (measured on B B^T / n + delta I,
  // n = 49152, delta = 4e-4: enough, and at 2^-4 GMRES-IR needed 85 instead
  // of 54 steps); a smaller shift gives a better preconditioner.
  double shift = 0.0;
  int info = 0;
  for (;;) {
    if ((rc = k_cvt_lower_shift(s, n, a, lda, shift, W, ldw))) return rc;
    if ((rc = chol_factor_fp8(s, n, W, ldw, &info))) return rc;
    if (info == 0) break;
    // a shift beyond max A(i, i) is far outside the FP8 perturbation of an
    // SPD matrix: A itself is not positive definite
    const double next = shift == 0.0 ? std::ldexp(dmax, ctx().fp8_shift_exp) : 2.0 * shift;
    if (!(dmax > 0.0) || next > dmax)
      return fail(PLA_ERR_INVALID_PARAM, "PLA_Chol_mixed: A is not positive definite (pivot %d, shift %g)", info,
                  shift);
    shift = next;
  }
  // L^T into the strict upper triangle of W: both solves run untransposed
  if ((rc = k_mirror_lower_f32(s, n, W, ldw))) return rc;
  const int nblk = (n + trsv_block_size() - 1) / trsv_block_size();
  const size_t dsz = sizeof(double) * (size_t)nblk * trsv_block_size() * trsv_block_size();
  double *DinvL = (double *)workspace(dsz, WS_CH_DINVL);
  double *DinvU = (double *)workspace(dsz, WS_CH_DINVLT);
  if (!DinvL || !DinvU) return PLA_ERR_NO_MEMORY;
  if ((rc = k_inv_diag_f32_op(s, true, false, false, n, W, ldw, DinvL))) return rc;
  if ((rc = k_inv_diag_f32_op(s, false, false, false, n, W, ldw, DinvU))) return rc;
  // The refinement's residuals read A as a full symmetric FP64 matrix, built
  // once from its lower triangle: one pass over n^2 contiguous elements per
  // residual instead of the lower triangle read twice, by rows and by columns.
  double *S = (double *)workspace(sizeof(double) * (size_t)n * n, WS_CH_FULL);
  if (!S) return PLA_ERR_NO_MEMORY;
  if ((rc = k_sym_full(s, n, a, lda, S, n))) return rc;
  refine_ops op;
  op.n = n;
  op.nrhs = nrhs;
  // a shifted FP8 factor is a weaker preconditioner: a longer GMRES cycle
  // (its orthogonalization costs little beside the two solves per step)
  op.restart = 100;
  op.gmres_first = ctx().fp8_gmres_first != 0;
  op.residual = [&](const double *Xv, long ldxv, const double *Bv, long ldbv, double *R, long ldr, double *rmax) {
    return residual(s, n, nrhs, S, n, Xv, ldxv, Bv, ldbv, R, ldr, rmax);
  };
  op.precond = [&](const double *Y, long ldy, double *Z, long ldz) -> int {
    PLA_HIP(hipMemcpy2DAsync(Z, sizeof(double) * ldz, Y, sizeof(double) * ldy, sizeof(double) * n, nrhs,
                             hipMemcpyDeviceToDevice, s));
    int r2;
    if ((r2 = trsv_sf_f32(s, true, false, false, n, nrhs, W, ldw, Z, ldz, DinvL))) return r2;
    return trsv_sf_f32(s, false, false, false, n, nrhs, W, ldw, Z, ldz, DinvU);
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
