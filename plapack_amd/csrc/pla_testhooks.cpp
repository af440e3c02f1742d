// pla_testhooks.cpp -- libpla_testhooks.so: C entry points onto internal
// kernels of libpla_mi355x.so, for the tests only (never linked by users).
// All pointers are device pointers; every hook runs on the library stream
// and waits for it.
#include "pla_internal.hpp"

using namespace pla;

extern "C" {

int pla_th_fp8_quantize(int R, int b, const double *P, long ldp, unsigned char *Lq, int *exps) {
  int rc = k_fp8_quantize(ctx().stream, R, b, P, ldp, Lq, exps);
  if (rc) return rc;
  PLA_HIP(hipStreamSynchronize(ctx().stream));
  return PLA_SUCCESS;
}

int pla_th_fp8_syrk(int R, const unsigned char *Lq, const int *exps, float *C, long ldc) {
  int rc = k_fp8_syrk_lower(ctx().stream, R, Lq, exps, C, ldc, nullptr);
  if (rc) return rc;
  PLA_HIP(hipStreamSynchronize(ctx().stream));
  return PLA_SUCCESS;
}

// C -= Lq[ro..] Lq[co..]^T over the first b bytes of each quantized row
// (tri: lower triangle of an M x M block with ro == co).
int pla_th_fp8_update(int tri, int M, int N, int b, const unsigned char *Lq, const int *exps, int ro, int co,
                      float *C, long ldc) {
  int rc = k_fp8_update(ctx().stream, tri != 0, M, N, b, Lq, exps, ro, co, C, ldc, nullptr);
  if (rc) return rc;
  PLA_HIP(hipStreamSynchronize(ctx().stream));
  return PLA_SUCCESS;
}

// W's strict upper triangle <- the transpose of its strict lower one.
int pla_th_mirror_lower(int n, float *W, long ldw) {
  int rc = k_mirror_lower_f32(ctx().stream, n, W, ldw);
  if (rc) return rc;
  PLA_HIP(hipStreamSynchronize(ctx().stream));
  return PLA_SUCCESS;
}

// FP8 Cholesky options: block size (256, 512 or 1024; 0 keeps the current one),
// panel on the CU-masked stream pair (0/1; -1 keeps), the first shift after a
// failed unshifted attempt as 2^shift_exp max A(i, i) (a value >= 0 keeps).
int pla_th_set_fp8_options(int nb, int masked, int shift_exp) {
  if (nb == 256 || nb == 512 || nb == 1024) ctx().fp8_nb = nb;
  if (masked >= 0) ctx().fp8_masked = masked;
  if (shift_exp < 0) ctx().fp8_shift_exp = shift_exp;
  return PLA_SUCCESS;
}

// FP8 Cholesky refinement: GMRES-IR from the first correction (1) or only
// after classic refinement stalls (0).
int pla_th_set_fp8_gmres_first(int on) {
  ctx().fp8_gmres_first = on;
  return PLA_SUCCESS;
}

// FP8 Cholesky panel: the FP32 <-> FP64 conversions of the rows below the
// diagonal block on a side stream (1) or on the panel stream (0); the factor
// is the same either way.
int pla_th_set_fp8_side_cvt(int on) {
  ctx().fp8_side_cvt = on != 0;
  return PLA_SUCCESS;
}

int pla_th_sym_residual(int n, int nrhs, const double *A, long lda, const double *X, long ldx, const double *B,
                        long ldb, double *R, long ldr, double *rmax) {
  return sym_residual(ctx().stream, n, nrhs, A, lda, X, ldx, B, ldb, R, ldr, rmax);
}

int pla_th_sym_inf_norm(int n, const double *A, long lda, double *out) {
  return sym_inf_norm(ctx().stream, n, A, lda, out);
}

// Poll limit of the panel-leaf exchange waits (0 restores the default): a
// tiny limit makes the waits give up, to test the timed-out-exchange path.
int pla_th_set_leaf_spin_limit(unsigned limit) {
  ctx().leaf_spin_limit = limit ? limit : 400000u;
  return PLA_SUCCESS;
}

// Number of panel leaves factored by one workgroup after their exchange gave
// up, since the last call (the count is reset).
int pla_th_leaf_fallbacks(unsigned long long *out) {
  leaf_ws &w = leaf_workspace(0);
  *out = 0;
  if (!w.base) return PLA_SUCCESS;
  PLA_HIP(hipStreamSynchronize(ctx().stream));
  PLA_HIP(hipMemcpy(out, w.state + 1, sizeof(unsigned long long), hipMemcpyDeviceToHost));
  PLA_HIP(hipMemset(w.state + 1, 0, sizeof(unsigned long long)));
  return PLA_SUCCESS;
}

// on: every triangular-solve block that has to wait gives up at its first
// wait (poll limit 0), whatever the timing; off: the default limit.
int pla_th_trsv_give_up_at_first_wait(int on) {
  ctx().trsv_spin_limit = on ? 0u : 2000000u;
  return PLA_SUCCESS;
}

// Columns per independent chain of the untransposed triangular solves
// (1, 2, 4, 8 or 16; 0: every column in one chain).
int pla_th_set_trsv_group_cols(int cols) {
  if (cols != 0 && cols != 1 && cols != 2 && cols != 4 && cols != 8 && cols != 16) return PLA_ERR_INVALID_PARAM;
  ctx().trsv_group_cols = cols;
  return PLA_SUCCESS;
}

}  // extern "C"
