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

// Poll limit of the triangular-solve block waits (0 restores the default).
int pla_th_set_trsv_spin_limit(unsigned limit) {
  ctx().trsv_spin_limit = limit ? limit : 2000000u;
  return PLA_SUCCESS;
}

}  // extern "C"
