// pla_blas.cpp -- recursive building blocks over the HIP kernels, kernel
// timing, and the global BLAS entry points of the factor-and-solve path.
//
// Reference counterparts:
//   trsm_left   PLA_Local_trsm / PLA_Trsm_left_lower|upper
//               (Local_BLAS/PLA_Local_trsm.c:13-97, BLAS3/PLA_Trsm_left_lower.c:13,
//               BLAS3/PLA_Trsm_left_upper.c:13)
//   trsm_rltn   dtrsm("R","L","T","N") in PLA_Chol_lower_large (SOLVERS/Cholesky/PLA_Chol.c:193-195)
//   lu_panel    PLA_LU_right_mv (SOLVERS/LU/PLA_LU_mv.c:13-101)
//   potrf_block PLA_Local_chol / PLA_dchol (SOLVERS/Cholesky/PLA_Local_chol.c:15-143)
//   PLA_Apply_pivots_to_rows  SOLVERS/LU/PLA_Apply_pivots_new.c:13
//   PLA_Trsv    BLAS2/PLA_Trsv.c:15     PLA_Trsm   BLAS3/PLA_Trsm.c:13
//   PLA_Gemm    BLAS3/PLA_Gemm.c:13     PLA_Gemv   BLAS2/PLA_Gemv.c:13
//   PLA_Syrk    BLAS3/PLA_Syrk.c:13     trsm_right PLA_Trsm_right_lower|upper
//               (BLAS3/PLA_Trsm_right_lower.c:13, BLAS3/PLA_Trsm_right_upper.c:13)
//
// Recursion replaces the reference's fixed-size inner blocking: every split
// puts the bulk of the flops into dgemm_kernel and leaves only <= 64-wide
// (triangular solves, Cholesky) or <= 32-wide (LU panel) leaves to the
// latency-bound kernels.
#include "pla_internal.hpp"

#include <vector>

namespace pla {

// Split point of a recursion over n: about half, rounded up to a multiple of q.
static int split(int n, int q) {
  int h = ((n / 2 + q - 1) / q) * q;
  if (h >= n) h = n - q > 0 ? n - q : n / 2;
  if (h <= 0) h = 1;
  return h;
}

int trsm_left(hipStream_t s, bool lower, bool trans, bool unit, int n, int nrhs, const double *T,
              long ldt, double *B, long ldb, const int *abort_flag) {
  if (n <= 0 || nrhs <= 0) return PLA_SUCCESS;
  const int leaf = trsm_leaf_size();
  if (n <= leaf) return k_trsm_leaf(s, lower, trans, unit, n, nrhs, T, ldt, B, ldb, abort_flag);
  const int n0 = split(n, leaf), n1 = n - n0;
  const double *T00 = T, *T11 = T + (long)n0 * ldt + n0;
  const bool forward = (lower != trans);
  int rc;
  if (forward) {
    // op(T) = [M00 0; M10 M11]; M10 = T10 (lower) or T01^T (upper, transposed)
    const double *M10 = trans ? T + (long)n0 * ldt : T + n0;
    if ((rc = trsm_left(s, lower, trans, unit, n0, nrhs, T00, ldt, B, ldb, abort_flag))) return rc;
    if ((rc = k_gemm(s, 0, trans, false, n1, nrhs, n0, -1.0, M10, ldt, B, ldb, 1.0, B + n0, ldb,
                     abort_flag)))
      return rc;
    return trsm_left(s, lower, trans, unit, n1, nrhs, T11, ldt, B + n0, ldb, abort_flag);
  }
  // op(T) = [M00 M01; 0 M11]; M01 = T01 (upper) or T10^T (lower, transposed)
  const double *M01 = trans ? T + n0 : T + (long)n0 * ldt;
  if ((rc = trsm_left(s, lower, trans, unit, n1, nrhs, T11, ldt, B + n0, ldb, abort_flag))) return rc;
  if ((rc = k_gemm(s, 0, trans, false, n0, nrhs, n1, -1.0, M01, ldt, B + n0, ldb, 1.0, B, ldb,
                   abort_flag)))
    return rc;
  return trsm_left(s, lower, trans, unit, n0, nrhs, T00, ldt, B, ldb, abort_flag);
}

// X op(T) = X0 over column blocks: op(T) = [M00 M01; M10 M11].  Upper-type
// op(T) (M10 = 0) solves X0 first, then X1 -= X0 M01; lower-type (M01 = 0)
// solves X1 first, then X0 -= X1 M10.
int trsm_right(hipStream_t s, bool lower, bool trans, bool unit, int nrows, int n, const double *T, long ldt,
               double *X, long ldx, const int *abort_flag) {
  if (nrows <= 0 || n <= 0) return PLA_SUCCESS;
  const int leaf = trsm_leaf_size();
  if (n <= leaf) return k_trsm_right_leaf(s, lower, trans, unit, nrows, n, T, ldt, X, ldx, abort_flag);
  const int n0 = split(n, leaf), n1 = n - n0;
  const double *T11 = T + (long)n0 * ldt + n0;
  double *X1 = X + (long)n0 * ldx;
  int rc;
  if (lower == trans) {
    // M01 = T01 (upper) or T10^T (lower, transposed)
    const double *M01 = trans ? T + n0 : T + (long)n0 * ldt;
    if ((rc = trsm_right(s, lower, trans, unit, nrows, n0, T, ldt, X, ldx, abort_flag))) return rc;
    if ((rc = k_gemm(s, 0, false, trans, nrows, n1, n0, -1.0, X, ldx, M01, ldt, 1.0, X1, ldx, abort_flag)))
      return rc;
    return trsm_right(s, lower, trans, unit, nrows, n1, T11, ldt, X1, ldx, abort_flag);
  }
  // M10 = T10 (lower) or T01^T (upper, transposed)
  const double *M10 = trans ? T + (long)n0 * ldt : T + n0;
  if ((rc = trsm_right(s, lower, trans, unit, nrows, n1, T11, ldt, X1, ldx, abort_flag))) return rc;
  if ((rc = k_gemm(s, 0, false, trans, nrows, n0, n1, -1.0, X1, ldx, M10, ldt, 1.0, X, ldx, abort_flag))) return rc;
  return trsm_right(s, lower, trans, unit, nrows, n0, T, ldt, X, ldx, abort_flag);
}

int trsm_rltn(hipStream_t s, int nrows, int b, const double *L, long ldl, double *X, long ldx,
              const int *abort_flag) {
  return trsm_right(s, true, true, false, nrows, b, L, ldl, X, ldx, abort_flag);
}

struct panel_rec {
  double *panel;  // panel top-left
  long lda;
  int mp, pw;     // panel rows and width
  int *piv;
  int *pairs;
  int nleaves;
  int leaf;       // leaf width
  int gcap;       // workgroup cap of a leaf launch
};

// Sub-panel whose first row and column are c0 inside the panel.
static int lu_panel_rec(hipStream_t s, panel_rec &pr, int c0, int w) {
  const int leaf = pr.leaf;
  const long lda = pr.lda;
  double *P = pr.panel + c0 * lda + c0;
  const int mp = pr.mp - c0;
  if (w <= leaf) {
    if (pr.nleaves >= LU_MAX_LEAVES)
      return fail(PLA_ERR_INVALID_PARAM, "lu panel: more than %d leaves", LU_MAX_LEAVES);
    lu_leaf_call c;
    c.P = P;
    c.Ppanel = pr.panel + c0;
    c.lda = lda;
    c.mp = mp;
    c.w = w;
    c.pw = pr.pw;
    c.pc0 = c0;
    c.piv = pr.piv + c0;
    c.pairs = pr.pairs + (size_t)pr.nleaves * leaf_pairs_stride();
    c.gcap = pr.gcap;
    c.W = pr.leaf;
    pr.nleaves++;
    return k_lu_leaf(s, c);
  }
  const int w0 = split(w, leaf), w1 = w - w0;
  int rc;
  if ((rc = lu_panel_rec(s, pr, c0, w0))) return rc;
  double *Pr = P + (long)w0 * lda;  // right columns (already interchanged by the leaves)
  if ((rc = trsm_left(s, true, false, true, w0, w1, P, lda, Pr, lda, nullptr))) return rc;
  if ((rc = k_gemm(s, 0, false, false, mp - w0, w1, w0, -1.0, P + w0, lda, Pr, lda, 1.0, Pr + w0, lda,
                   nullptr)))
    return rc;
  return lu_panel_rec(s, pr, c0 + w0, w1);
}

int lu_panel(hipStream_t s, double *P, int mp, int w, long lda, int *piv, int *pairs, int *nleaves, int gcap) {
  // gcap counts CUs; a register-leaf wave takes one SIMD, four per CU
  const bool regs = ctx().leaf_kind == 0;
  int leaf = ctx().leaf_width > 0 ? ctx().leaf_width
                                  : (regs ? rleaf_width_for(mp, gcap * 4) : leaf_width_for(mp, gcap));
  if (leaf <= 0) return fail(PLA_ERR_DIMENSIONS, "lu panel: %d rows are too many for the panel leaf", mp);
  if (regs) gcap *= 4;
  // the leaves of one panel are recorded for the composite interchange
  while ((w + leaf - 1) / leaf > LU_MAX_LEAVES && leaf < leaf_max_width()) leaf *= 2;
  panel_rec pr{P, lda, mp, w, piv, pairs, 0, leaf, gcap};
  const int rc = w > 0 ? lu_panel_rec(s, pr, 0, w) : PLA_SUCCESS;
  *nleaves = pr.nleaves;
  return rc;
}

int potrf_block(hipStream_t s, int b, double *A, long lda, int offset, int *info) {
  if (b <= 0) return PLA_SUCCESS;
  const int leaf = trsm_leaf_size();
  if (b <= leaf) return k_potrf_leaf(s, b, A, lda, offset, info);
  const int b0 = split(b, leaf), b1 = b - b0;
  int rc;
  if ((rc = potrf_block(s, b0, A, lda, offset, info))) return rc;
  if ((rc = trsm_rltn(s, b1, b0, A, lda, A + b0, lda, info))) return rc;
  if ((rc = k_syrk_lower(s, 0, b1, b0, -1.0, A + b0, lda, 1.0, A + (long)b0 * lda + b0, lda, info)))
    return rc;
  return potrf_block(s, b1, A + (long)b0 * lda + b0, lda, offset + b0, info);
}

// ---------------------------------------------------------------------------
// Kernel timing: hipEvent pairs recorded on the library stream around each
// launch of a role; summed on query.
struct role_timer {
  std::vector<hipEvent_t> ev;  // begin/end pairs
  std::vector<double> flops;
  double done_ms = 0.0, done_flops = 0.0;
  long done_launches = 0;
};
static bool g_timing = false;
static role_timer g_timers[8];

bool timing_on() { return g_timing; }

void timing_begin(hipStream_t s, int role) {
  if (!g_timing || role < 1 || role > 7) return;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  (void)hipEventRecord(e, s);
  g_timers[role].ev.push_back(e);
}

void timing_end(hipStream_t s, int role, double flops) {
  if (!g_timing || role < 1 || role > 7) return;
  role_timer &t = g_timers[role];
  if (t.ev.size() % 2 != 1) return;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  (void)hipEventRecord(e, s);
  t.ev.push_back(e);
  t.flops.push_back(flops);
}

static void timing_collect(role_timer &t) {
  const size_t pairs = t.ev.size() / 2;
  for (size_t i = 0; i < pairs; i++) {
    float ms = 0.f;
    (void)hipEventSynchronize(t.ev[2 * i + 1]);
    if (hipEventElapsedTime(&ms, t.ev[2 * i], t.ev[2 * i + 1]) == hipSuccess) t.done_ms += ms;
    t.done_flops += t.flops[i];
    t.done_launches++;
  }
  for (size_t i = 0; i < 2 * pairs; i++) (void)hipEventDestroy(t.ev[i]);
  t.ev.erase(t.ev.begin(), t.ev.begin() + 2 * pairs);
  t.flops.erase(t.flops.begin(), t.flops.begin() + pairs);
}

static void timing_reset() {
  for (auto &t : g_timers) {
    for (auto e : t.ev) (void)hipEventDestroy(e);
    t = role_timer();
  }
}

void timing_release() {
  timing_reset();
  g_timing = false;
}

}  // namespace pla

using namespace pla;

static int check_double(PLA_Obj o, const char *who) {
  if (!valid(o)) return fail(PLA_ERR_INVALID_OBJECT, "%s: invalid object", who);
  if (o->base->datatype != MPI_DOUBLE)
    return fail(PLA_ERR_UNSUPPORTED, "%s: only MPI_DOUBLE objects are supported", who);
  return PLA_SUCCESS;
}

extern "C" {

int PLA_Kernel_timing(int on) {
  if (!ctx().initialized) return fail(PLA_ERR_NOT_INITIALIZED, "PLAPACK not yet initialized");
  timing_reset();
  g_timing = (on != 0);
  return PLA_SUCCESS;
}

int PLA_Kernel_timing_query(int role, long *launches, double *total_ms, double *total_flops) {
  if (role < 1 || role > 7) return fail(PLA_ERR_INVALID_PARAM, "PLA_Kernel_timing_query: role %d", role);
  role_timer &t = g_timers[role];
  timing_collect(t);
  if (launches) *launches = t.done_launches;
  if (total_ms) *total_ms = t.done_ms;
  if (total_flops) *total_flops = t.done_flops;
  return PLA_SUCCESS;
}

int PLA_Apply_pivots_to_rows(PLA_Obj A, PLA_Obj pivots) {
  if (is_single(A)) return on_promoted({&A}, {true}, [&] { return PLA_Apply_pivots_to_rows(A, pivots); });
  const bool cplx = valid(A) && A->base->datatype == MPI_DOUBLE_COMPLEX;
  int rc = cplx ? PLA_SUCCESS : check_double(A, "PLA_Apply_pivots_to_rows");
  if (rc) return rc;
  if (!valid(pivots) || pivots->base->datatype != MPI_INT)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Apply_pivots_to_rows: pivots must be an MPI_INT object");
  if (pivots->width != 1 && pivots->length != 1)
    return fail(PLA_ERR_DIMENSIONS, "PLA_Apply_pivots_to_rows: pivots must be a vector");
  const int np = pivots->width == 1 ? pivots->length : pivots->width;
  if (np > A->length)
    return fail(PLA_ERR_DIMENSIONS, "PLA_Apply_pivots_to_rows: %d pivots for %d rows", np, A->length);
  if (pivots->width != 1 && pivots->base->ldim != 1)
    return fail(PLA_ERR_UNSUPPORTED, "PLA_Apply_pivots_to_rows: strided pivot row vector");
  double *a;
  int *p;
  if ((rc = dev_ptr(A, (void **)&a)) || (rc = dev_ptr(pivots, (void **)&p))) return rc;
  // validate the interchanges on the host: row i + piv[i] must lie in A
  std::vector<int> hp(np);
  PLA_HIP(hipStreamSynchronize(ctx().stream));
  PLA_HIP(hipMemcpy(hp.data(), p, sizeof(int) * np, hipMemcpyDeviceToHost));
  for (int i = 0; i < np; i++)
    if (hp[i] < 0 || i + hp[i] >= A->length)
      return fail(PLA_ERR_INVALID_PARAM, "PLA_Apply_pivots_to_rows: pivot %d (= %d) leaves the matrix", i,
                  hp[i]);
  if ((rc = cplx ? k_zlaswp(ctx().stream, a, ld(A), A->length, A->width, np, p)
                 : k_laswp(ctx().stream, a, ld(A), A->length, A->width, np, p)))
    return rc;
  return written(A);
}

int PLA_Trsm(int side, int uplo, int trans, int diag, PLA_Obj alpha, PLA_Obj A, PLA_Obj B) {
  if (is_single(A) || is_single(B))
    return on_promoted({&alpha, &A, &B}, {false, false, true},
                       [&] { return PLA_Trsm(side, uplo, trans, diag, alpha, A, B); });
  if (valid(A) && A->base->datatype == MPI_DOUBLE_COMPLEX) return zblas_trsm(side, uplo, trans, diag, alpha, A, B);
  int rc;
  if ((rc = check_double(A, "PLA_Trsm")) || (rc = check_double(B, "PLA_Trsm"))) return rc;
  if (uplo != PLA_LOWER_TRIANGULAR && uplo != PLA_UPPER_TRIANGULAR)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Trsm: invalid uplo");
  if (trans != PLA_NO_TRANSPOSE && trans != PLA_TRANSPOSE && trans != PLA_CONJUGATE_TRANSPOSE)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Trsm: invalid trans");
  if (diag != PLA_UNIT_DIAG && diag != PLA_NONUNIT_DIAG)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Trsm: invalid diag");
  const int n = A->length;
  if (A->width != n) return fail(PLA_ERR_DIMENSIONS, "PLA_Trsm: A is %d x %d", A->length, A->width);
  double al = 1.0;
  if (alpha && (rc = scalar_value(alpha, &al))) return rc;
  const bool lower = (uplo == PLA_LOWER_TRIANGULAR);
  const bool tr = (trans != PLA_NO_TRANSPOSE);  // real data: conjugate transpose == transpose
  const bool unit = (diag == PLA_UNIT_DIAG);
  double *a, *b;
  if ((rc = dev_ptr(A, (void **)&a)) || (rc = dev_ptr(B, (void **)&b))) return rc;
  hipStream_t s = ctx().stream;
  if (side == PLA_SIDE_LEFT) {
    if (B->length != n) return fail(PLA_ERR_DIMENSIONS, "PLA_Trsm: B has %d rows, A is %d", B->length, n);
    if ((rc = k_scale2d(s, B->length, B->width, al, b, ld(B)))) return rc;
    if ((rc = trsm_left(s, lower, tr, unit, n, B->width, a, ld(A), b, ld(B), nullptr))) return rc;
  } else if (side == PLA_SIDE_RIGHT) {
    if (B->width != n) return fail(PLA_ERR_DIMENSIONS, "PLA_Trsm: B has %d columns, A is %d", B->width, n);
    if ((rc = k_scale2d(s, B->length, B->width, al, b, ld(B)))) return rc;
    if ((rc = trsm_right(s, lower, tr, unit, B->length, n, a, ld(A), b, ld(B), nullptr))) return rc;
  } else {
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Trsm: invalid side");
  }
  return written(B);
}

int PLA_Trsv(int uplo, int trans, int diag, PLA_Obj A, PLA_Obj x) {
  if (is_single(A) || is_single(x))
    return on_promoted({&A, &x}, {false, true}, [&] { return PLA_Trsv(uplo, trans, diag, A, x); });
  // complex: the one-column case of the complex triangular solve
  if (valid(A) && A->base->datatype == MPI_DOUBLE_COMPLEX) {
    if (!valid(x) || x->width != 1)
      return fail(PLA_ERR_DIMENSIONS, "PLA_Trsv: x must be one column");
    return zblas_trsm(PLA_SIDE_LEFT, uplo, trans, diag, nullptr, A, x);
  }
  int rc;
  if ((rc = check_double(A, "PLA_Trsv")) || (rc = check_double(x, "PLA_Trsv"))) return rc;
  const int n = A->length;
  if (A->width != n) return fail(PLA_ERR_DIMENSIONS, "PLA_Trsv: A is %d x %d", A->length, A->width);
  if (!((x->length == n && x->width == 1) || (x->width == n && x->length == 1)))
    return fail(PLA_ERR_DIMENSIONS, "PLA_Trsv: x is %d x %d, A is %d x %d", x->length, x->width, n, n);
  if (x->width == n && n > 1 && x->length == 1)
    return fail(PLA_ERR_UNSUPPORTED, "PLA_Trsv: row vector x is not supported");
  if (uplo != PLA_LOWER_TRIANGULAR && uplo != PLA_UPPER_TRIANGULAR)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Trsv: invalid uplo");
  if (diag != PLA_UNIT_DIAG && diag != PLA_NONUNIT_DIAG)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Trsv: invalid diag");
  double *a, *b;
  if ((rc = dev_ptr(A, (void **)&a)) || (rc = dev_ptr(x, (void **)&b))) return rc;
  // one sync-free blocked solve (HBM-bound: each element of the triangle read once)
  if ((rc = trsv_sf(ctx().stream, uplo == PLA_LOWER_TRIANGULAR, trans != PLA_NO_TRANSPOSE, diag == PLA_UNIT_DIAG,
                    n, 1, a, ld(A), b, ld(x) > n ? ld(x) : n, nullptr)))
    return rc;
  if ((rc = leaf_check_error(ctx().stream, "PLA_Trsv"))) return rc;
  return written(x);
}

int PLA_Gemm(int transa, int transb, PLA_Obj alpha, PLA_Obj A, PLA_Obj B, PLA_Obj beta, PLA_Obj C) {
  if (is_single(A) || is_single(B) || is_single(C))
    return on_promoted({&alpha, &A, &B, &beta, &C}, {false, false, false, false, true},
                       [&] { return PLA_Gemm(transa, transb, alpha, A, B, beta, C); });
  if (valid(C) && C->base->datatype == MPI_DOUBLE_COMPLEX) return zblas_gemm(transa, transb, alpha, A, B, beta, C);
  int rc;
  if ((rc = check_double(A, "PLA_Gemm")) || (rc = check_double(B, "PLA_Gemm")) ||
      (rc = check_double(C, "PLA_Gemm")))
    return rc;
  const bool ta = (transa != PLA_NO_TRANSPOSE), tb = (transb != PLA_NO_TRANSPOSE);
  const int m = C->length, n = C->width;
  const int am = ta ? A->width : A->length, ak = ta ? A->length : A->width;
  const int bk = tb ? B->width : B->length, bn = tb ? B->length : B->width;
  if (am != m || bn != n || ak != bk)
    return fail(PLA_ERR_DIMENSIONS, "PLA_Gemm: nonconformal (%d x %d) * (%d x %d) -> %d x %d", am, ak, bk,
                bn, m, n);
  double al = 1.0, be = 1.0;
  if ((alpha && (rc = scalar_value(alpha, &al))) || (beta && (rc = scalar_value(beta, &be)))) return rc;
  double *a, *b, *c;
  if ((rc = dev_ptr(A, (void **)&a)) || (rc = dev_ptr(B, (void **)&b)) || (rc = dev_ptr(C, (void **)&c)))
    return rc;
  if ((rc = k_gemm(ctx().stream, 0, ta, tb, m, n, ak, al, a, ld(A), b, ld(B), be, c, ld(C), nullptr)))
    return rc;
  return written(C);
}

// BLAS3/PLA_Gemm_x.c:13: the versions select how the reference distributes
// the product over a mesh (BLAS3/PLA_Gemm_x.c dispatches on them); on one GPU
// they all compute PLA_Gemm's product.
int PLA_Gemm_x(int version, int transa, int transb, PLA_Obj alpha, PLA_Obj A, PLA_Obj B, PLA_Obj beta, PLA_Obj C) {
  if (version != PLA_VERSION_A && version != PLA_VERSION_B && version != PLA_VERSION_C)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Gemm_x: unknown version %d", version);
  return PLA_Gemm(transa, transb, alpha, A, B, beta, C);
}

// C <- beta C, then the uplo triangle of C += alpha op(A) op(A)^T
// (BLAS3/PLA_Syrk_panpan.c:13-95: PLA_Local_scal(beta, C) scales all of C,
// both triangles, before the rank-nb updates of the uplo triangle).
// PLA_NO_TRANS / PLA_CONJ: op(A) = A, n x k; PLA_TRANS / PLA_CONJ_TRANS:
// op(A) = A^T, A k x n (real data: conjugation is the identity).
int PLA_Syrk(int uplo, int trans, PLA_Obj alpha, PLA_Obj A, PLA_Obj beta, PLA_Obj C) {
  if (is_single(A) || is_single(C))
    return on_promoted({&alpha, &A, &beta, &C}, {false, false, false, true},
                       [&] { return PLA_Syrk(uplo, trans, alpha, A, beta, C); });
  if (valid(C) && C->base->datatype == MPI_DOUBLE_COMPLEX) return zblas_syrk(uplo, trans, alpha, A, beta, C);
  int rc;
  if ((rc = check_double(A, "PLA_Syrk")) || (rc = check_double(C, "PLA_Syrk"))) return rc;
  if (uplo != PLA_LOWER_TRIANGULAR && uplo != PLA_UPPER_TRIANGULAR)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Syrk: invalid uplo");
  if (trans != PLA_NO_TRANSPOSE && trans != PLA_TRANSPOSE && trans != PLA_CONJUGATE_TRANSPOSE &&
      trans != PLA_CONJUGATE)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Syrk: invalid trans");
  if (A->objtype != PLA_MATRIX || C->objtype != PLA_MATRIX)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Syrk: A and C must be PLA_MATRIX objects");
  const int n = C->length;
  if (C->width != n) return fail(PLA_ERR_DIMENSIONS, "PLA_Syrk: C is %d x %d, not square", C->length, C->width);
  const bool tr = (trans == PLA_TRANSPOSE || trans == PLA_CONJUGATE_TRANSPOSE);
  if ((tr ? A->width : A->length) != n)
    return fail(PLA_ERR_DIMENSIONS, "PLA_Syrk: A is %d x %d for a %d x %d C", A->length, A->width, n, n);
  const int k = tr ? A->length : A->width;
  double al = 1.0, be = 1.0;
  if ((alpha && (rc = scalar_value(alpha, &al))) || (beta && (rc = scalar_value(beta, &be)))) return rc;
  double *a, *c;
  if ((rc = dev_ptr(A, (void **)&a)) || (rc = dev_ptr(C, (void **)&c))) return rc;
  hipStream_t s = ctx().stream;
  if ((rc = k_scale2d(s, n, n, be, c, ld(C)))) return rc;
  if (k > 0 && al != 0.0 &&
      (rc = k_syrk(s, uplo == PLA_LOWER_TRIANGULAR, tr, n, k, al, a, ld(A), 1.0, c, ld(C))))
    return rc;
  return written(C);
}

static int matrix_norm(PLA_Obj A, PLA_Obj alpha, bool inf_norm, const char *who) {
  if (is_single(A) || is_single(alpha))
    return on_promoted({&A, &alpha}, {false, true}, [&] { return matrix_norm(A, alpha, inf_norm, who); });
  int rc;
  if ((rc = check_double(A, who)) || (rc = check_double(alpha, who))) return rc;
  double *a, v = 0.0;
  if ((rc = dev_ptr(A, (void **)&a))) return rc;
  if ((rc = k_norm(ctx().stream, inf_norm, A->length, A->width, a, ld(A), &v))) return rc;
  void *p;
  if ((rc = dev_ptr(alpha, &p))) return rc;
  PLA_HIP(hipMemcpy(p, &v, sizeof v, hipMemcpyHostToDevice));
  return written(alpha);
}

int PLA_Matrix_infinity_norm(PLA_Obj A, PLA_Obj alpha) {
  return matrix_norm(A, alpha, true, "PLA_Matrix_infinity_norm");
}

int PLA_Matrix_one_norm(PLA_Obj A, PLA_Obj alpha) {
  return matrix_norm(A, alpha, false, "PLA_Matrix_one_norm");
}

int PLA_Gemv(int transa, PLA_Obj alpha, PLA_Obj A, PLA_Obj x, PLA_Obj beta, PLA_Obj y) {
  if (is_single(A) || is_single(x) || is_single(y))
    return on_promoted({&alpha, &A, &x, &beta, &y}, {false, false, false, false, true},
                       [&] { return PLA_Gemv(transa, alpha, A, x, beta, y); });
  if (valid(A) && A->base->datatype == MPI_DOUBLE_COMPLEX) {
    if (!valid(x) || !valid(y) || x->width != 1 || y->width != 1)
      return fail(PLA_ERR_DIMENSIONS, "PLA_Gemv: x and y must be one column");
    return zblas_gemm(transa, PLA_NO_TRANSPOSE, alpha, A, x, beta, y);
  }
  int rc;
  if ((rc = check_double(A, "PLA_Gemv")) || (rc = check_double(x, "PLA_Gemv")) ||
      (rc = check_double(y, "PLA_Gemv")))
    return rc;
  const bool ta = (transa != PLA_NO_TRANSPOSE);
  const int m = ta ? A->width : A->length, k = ta ? A->length : A->width;
  if (x->width != 1 || y->width != 1 || x->length != k || y->length != m)
    return fail(PLA_ERR_DIMENSIONS, "PLA_Gemv: x %d x %d, y %d x %d for op(A) %d x %d", x->length,
                x->width, y->length, y->width, m, k);
  double al = 1.0, be = 1.0;
  if ((alpha && (rc = scalar_value(alpha, &al))) || (beta && (rc = scalar_value(beta, &be)))) return rc;
  double *a, *xv, *yv;
  if ((rc = dev_ptr(A, (void **)&a)) || (rc = dev_ptr(x, (void **)&xv)) || (rc = dev_ptr(y, (void **)&yv)))
    return rc;
  if ((rc = k_gemm(ctx().stream, 0, ta, false, m, 1, k, al, a, ld(A), xv, ld(x), be, yv, ld(y), nullptr)))
    return rc;
  return written(y);
}

}  // extern "C"
