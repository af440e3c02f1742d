// pla_zblas.cpp -- PLA_Gemm, PLA_Trsm and PLA_Syrk on MPI_DOUBLE_COMPLEX
// objects (BLAS3/PLA_Gemm.c:13, BLAS3/PLA_Trsm.c:13, BLAS3/PLA_Syrk.c:13 with
// the zgemm / ztrsm / zsyrk local kernels, INCLUDE/PLA_BLAS.h).
//
// Every product runs on the FP64 MFMA GEMM through the real embedding of its
// left operand (pla_complex.hip), so the complex BLAS3 share the real path's
// kernel.  Transposed and conjugated operands, and alpha, are formed
// explicitly into workspaces first (k_zop); beta scales C in place.
#include "pla_internal.hpp"

using namespace pla;

namespace {

// op code of k_zop for a PLA transpose constant
int op_code(int trans) {
  switch (trans) {
    case PLA_NO_TRANSPOSE: return 0;
    case PLA_TRANSPOSE: return 1;
    case PLA_CONJUGATE_TRANSPOSE: return 2;
    case PLA_CONJUGATE: return 3;
    default: return -1;
  }
}

int check_z(PLA_Obj o, const char *who) {
  if (!valid(o)) return fail(PLA_ERR_INVALID_OBJECT, "%s: invalid object", who);
  if (o->base->datatype != MPI_DOUBLE_COMPLEX)
    return fail(PLA_ERR_INVALID_PARAM, "%s: operands of different datatypes", who);
  return PLA_SUCCESS;
}

int scalar_or_one(PLA_Obj o, double *re, double *im) {
  *re = 1.0;
  *im = 0.0;
  return o ? scalar_value_z(o, re, im) : PLA_SUCCESS;
}

}  // namespace

namespace pla {

int zblas_gemm(int transa, int transb, PLA_Obj alpha, PLA_Obj A, PLA_Obj B, PLA_Obj beta, PLA_Obj C) {
  int rc;
  if ((rc = check_z(A, "PLA_Gemm")) || (rc = check_z(B, "PLA_Gemm")) || (rc = check_z(C, "PLA_Gemm"))) return rc;
  const int oa = op_code(transa), ob = op_code(transb);
  if (oa < 0 || ob < 0) return fail(PLA_ERR_INVALID_PARAM, "PLA_Gemm: invalid transa / transb");
  const bool ta = (oa == 1 || oa == 2), tb = (ob == 1 || ob == 2);
  const int m = C->length, n = C->width;
  const int am = ta ? A->width : A->length, k = ta ? A->length : A->width;
  const int bk = tb ? B->width : B->length, bn = tb ? B->length : B->width;
  if (am != m || bn != n || k != bk)
    return fail(PLA_ERR_DIMENSIONS, "PLA_Gemm: nonconformal (%d x %d) * (%d x %d) -> %d x %d", am, k, bk, bn, m, n);
  double ar, ai, br, bi;
  if ((rc = scalar_or_one(alpha, &ar, &ai)) || (rc = scalar_or_one(beta, &br, &bi))) return rc;
  void *a, *b, *c;
  if ((rc = dev_ptr(A, &a)) || (rc = dev_ptr(B, &b)) || (rc = dev_ptr(C, &c))) return rc;
  hipStream_t s = ctx().stream;
  if (m == 0 || n == 0) return PLA_SUCCESS;
  if (!(br == 1.0 && bi == 0.0) && (rc = k_zscale2d(s, m, n, br, bi, c, ld(C)))) return rc;
  if (k > 0 && !(ar == 0.0 && ai == 0.0)) {
    // C -= (-alpha op(A)) op(B)
    void *wa = workspace(16 * (size_t)m * k, WS_ZOPA);
    if (!wa) return PLA_ERR_NO_MEMORY;
    if ((rc = k_zop(s, oa, m, k, -ar, -ai, a, ld(A), wa, m))) return rc;
    const void *bp = b;
    long ldbp = ld(B);
    if (ob != 0) {
      void *wb = workspace(16 * (size_t)k * n, WS_ZOPB);
      if (!wb) return PLA_ERR_NO_MEMORY;
      if ((rc = k_zop(s, ob, k, n, 1.0, 0.0, b, ld(B), wb, k))) return rc;
      bp = wb;
      ldbp = k;
    }
    if ((rc = zgemm_minus(s, m, n, k, wa, m, bp, ldbp, c, ld(C), WS_ZEMBED))) return rc;
  }
  return written(C);
}

int zblas_trsm(int side, int uplo, int trans, int diag, PLA_Obj alpha, PLA_Obj A, PLA_Obj B) {
  int rc;
  if ((rc = check_z(A, "PLA_Trsm")) || (rc = check_z(B, "PLA_Trsm"))) return rc;
  const int op = op_code(trans);
  if (op < 0 || (side != PLA_SIDE_LEFT && side != PLA_SIDE_RIGHT) ||
      (uplo != PLA_LOWER_TRIANGULAR && uplo != PLA_UPPER_TRIANGULAR) ||
      (diag != PLA_UNIT_DIAG && diag != PLA_NONUNIT_DIAG))
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Trsm: invalid side / uplo / trans / diag");
  const bool left = (side == PLA_SIDE_LEFT), lower = (uplo == PLA_LOWER_TRIANGULAR), unit = (diag == PLA_UNIT_DIAG);
  const int nt = A->length;
  if (A->width != nt || (left ? B->length : B->width) != nt)
    return fail(PLA_ERR_DIMENSIONS, "PLA_Trsm: A %d x %d, B %d x %d", A->length, A->width, B->length, B->width);
  double ar, ai;
  if ((rc = scalar_or_one(alpha, &ar, &ai))) return rc;
  void *t, *b;
  if ((rc = dev_ptr(A, &t)) || (rc = dev_ptr(B, &b))) return rc;
  hipStream_t s = ctx().stream;
  const int m = B->length, n = B->width;
  if (m == 0 || n == 0) return PLA_SUCCESS;
  if (left) {
    // op(T) X = alpha B
    if (!(ar == 1.0 && ai == 0.0) && (rc = k_zscale2d(s, m, n, ar, ai, b, ld(B)))) return rc;
    const void *tm = t;
    long ldm = ld(A);
    bool lo = lower;
    if (op != 0) {
      void *wt = workspace(16 * (size_t)nt * nt, WS_ZOPA);
      if (!wt) return PLA_ERR_NO_MEMORY;
      if ((rc = k_zop(s, op, nt, nt, 1.0, 0.0, t, ld(A), wt, nt))) return rc;
      tm = wt;
      ldm = nt;
      if (op != 3) lo = !lower;  // a transpose swaps the triangle; conj keeps it
    }
    if ((rc = ztrsm_left_n(s, lo, unit, nt, n, tm, ldm, b, ld(B), WS_ZEMBED))) return rc;
    return written(B);
  }
  // X op(T) = alpha B  <=>  op(T)^T X^T = alpha B^T
  void *wy = workspace(16 * (size_t)nt * m, WS_ZOPB);
  if (!wy) return PLA_ERR_NO_MEMORY;
  if ((rc = k_zop(s, 1, nt, m, ar, ai, b, ld(B), wy, nt))) return rc;
  // op(T)^T: T^T for N, T for T, conj(T) for C (= H), T^H for conj
  static const int mop[4] = {1, 0, 3, 2};
  const void *tm = t;
  long ldm = ld(A);
  bool lo = lower;
  if (mop[op] != 0) {
    void *wt = workspace(16 * (size_t)nt * nt, WS_ZOPA);
    if (!wt) return PLA_ERR_NO_MEMORY;
    if ((rc = k_zop(s, mop[op], nt, nt, 1.0, 0.0, t, ld(A), wt, nt))) return rc;
    tm = wt;
    ldm = nt;
    if (mop[op] != 3) lo = !lower;
  }
  if ((rc = ztrsm_left_n(s, lo, unit, nt, m, tm, ldm, wy, nt, WS_ZEMBED))) return rc;
  if ((rc = k_zop(s, 1, m, nt, 1.0, 0.0, wy, nt, b, ld(B)))) return rc;
  return written(B);
}

int zblas_syrk(int uplo, int trans, PLA_Obj alpha, PLA_Obj A, PLA_Obj beta, PLA_Obj C) {
  int rc;
  if ((rc = check_z(A, "PLA_Syrk")) || (rc = check_z(C, "PLA_Syrk"))) return rc;
  if (uplo != PLA_LOWER_TRIANGULAR && uplo != PLA_UPPER_TRIANGULAR)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Syrk: invalid uplo");
  // zsyrk: op(A) = A or A^T (no conjugation)
  if (trans != PLA_NO_TRANSPOSE && trans != PLA_TRANSPOSE)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Syrk: MPI_DOUBLE_COMPLEX takes PLA_NO_TRANSPOSE or PLA_TRANSPOSE");
  const bool tr = (trans == PLA_TRANSPOSE);
  const int n = C->length;
  if (C->width != n) return fail(PLA_ERR_DIMENSIONS, "PLA_Syrk: C is %d x %d, not square", C->length, C->width);
  const int an = tr ? A->width : A->length, k = tr ? A->length : A->width;
  if (an != n) return fail(PLA_ERR_DIMENSIONS, "PLA_Syrk: op(A) has %d rows, C is %d x %d", an, n, n);
  double ar, ai, br, bi;
  if ((rc = scalar_or_one(alpha, &ar, &ai)) || (rc = scalar_or_one(beta, &br, &bi))) return rc;
  void *a, *c;
  if ((rc = dev_ptr(A, &a)) || (rc = dev_ptr(C, &c))) return rc;
  hipStream_t s = ctx().stream;
  if (n == 0) return PLA_SUCCESS;
  // C <- beta C on all of C (PLA_Syrk_panpan.c:22)
  if (!(br == 1.0 && bi == 0.0) && (rc = k_zscale2d(s, n, n, br, bi, c, ld(C)))) return rc;
  if (k > 0 && !(ar == 0.0 && ai == 0.0)) {
    // P = alpha op(A) op(A)^T, then the uplo triangle of C += P
    void *wa = workspace(16 * (size_t)n * k, WS_ZOPA);
    void *wb = workspace(16 * (size_t)k * n, WS_ZOPB);
    void *p = workspace(16 * (size_t)n * n, WS_ZTRI);
    if (!wa || !wb || !p) return PLA_ERR_NO_MEMORY;
    if ((rc = k_zop(s, tr ? 1 : 0, n, k, -ar, -ai, a, ld(A), wa, n))) return rc;
    if ((rc = k_zop(s, tr ? 0 : 1, k, n, 1.0, 0.0, a, ld(A), wb, k))) return rc;
    if ((rc = k_zscale2d(s, n, n, 0.0, 0.0, p, n))) return rc;
    if ((rc = zgemm_minus(s, n, n, k, wa, n, wb, k, p, n, WS_ZEMBED))) return rc;
    if ((rc = k_ztri_add(s, n, uplo == PLA_LOWER_TRIANGULAR, p, n, c, ld(C)))) return rc;
  }
  return written(C);
}

}  // namespace pla
