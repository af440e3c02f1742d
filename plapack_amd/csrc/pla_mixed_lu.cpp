// pla_mixed_lu.cpp -- PLA_LU_mixed: LU in FP32 storage with a low-precision
// MFMA trailing update, then FP64 iterative refinement (extension; no
// reference counterpart).  The kernels and the scheme are described in
// pla_mixed.hip.
#include "pla_internal.hpp"

#include <cmath>
#include <vector>

using namespace pla;

namespace pla {

static bool prec_split(int prec) { return prec == PLA_MIXED_FP16X3 || prec == PLA_MIXED_BF16X3; }

// Workspace slots used here (the others belong to pla_kernels / pla_mixed).
enum { WS_W = 10, WS_P64, WS_U64, WS_LH, WS_LL, WS_UH, WS_UL, WS_PERM, WS_DINVL, WS_DINVU, WS_D, WS_R, WS_PAIRS,
       WS_COMP };

// Blocked right-looking LU of the n x n FP32 matrix W (ldw) with one step
// of look-ahead, as lu_factor (pla_lu.cpp) but with FP64 panels and
// triangular solves and a low-precision trailing update.
int lu_factor_mixed(hipStream_t sMain, int prec, int n, float *W, long ldw, int *piv, int nb) {
  nb = (nb / 32) * 32;
  if (nb < 32) nb = 32;
  if (nb > 1024) nb = 1024;
  const int nbp = nb;  // depth of the low-precision copies (a multiple of 32)
  const int esz = 2;   // bytes of one FP16 / BF16 element
  const bool split = prec_split(prec);
  const int stride = leaf_pairs_stride();
  // FP64 panel copies, double buffered by panel parity: step k's solves read
  // L11 from panel k's copy while panel k+1 is factored into the other
  double *P64 = (double *)workspace(sizeof(double) * 2 * (size_t)n * nb, WS_P64);
  double *U64 = (double *)workspace(sizeof(double) * (size_t)n * nb, WS_U64);
  // the low-precision copies are packed k-slices of 32 (k_hgemm): n indices
  // (rows of L, columns of U) per slice, nbp / 32 slices
  const long slice = ((long)n * 32 + 255) / 256 * 256;
  const size_t lbuf = (size_t)slice * (nbp / 32) * esz;  // one copy (one parity of the L copies)
  char *Lh = (char *)workspace(2 * lbuf, WS_LH);
  char *Ll = split ? (char *)workspace(2 * lbuf, WS_LL) : nullptr;
  char *Uh = (char *)workspace(lbuf, WS_UH);
  char *Ul = split ? (char *)workspace(lbuf, WS_UL) : nullptr;
  const size_t pstride = (size_t)LU_MAX_LEAVES * stride;  // one parity of the leaves' pair lists
  int *pairs = (int *)workspace(sizeof(int) * 2 * pstride, WS_PAIRS, true);
  const int cstride = panel_comp_stride();
  // one composite interchange list per step (the left-of-panel interchanges
  // of a step may still be reading theirs while later panels are composed)
  int *comp = (int *)workspace(sizeof(int) * (size_t)lu_step_count(n, n, nb) * cstride, WS_COMP);
  if (!P64 || !U64 || !Lh || !Uh || !pairs || !comp || (split && (!Ll || !Ul))) return PLA_ERR_NO_MEMORY;
  hipEvent_t ev_ready = lib_event(EV_MLU_READY), ev_panel = lib_event(EV_MLU_PANEL);
  if (!ev_ready || !ev_panel) return fail(PLA_ERR_NOT_INITIALIZED, "PLAPACK not yet initialized");
  int nleaves[2] = {0, 0};
  int rc;
  // sP: panels; sM: interchanges right of the panel, U12 solves and trailing
  // updates; sMain: interchanges left of the panel (nothing reads L until the
  // factorization ends)
  hipStream_t sP, sM;
  int gcap;
  if ((rc = lu_streams_begin(sMain, n, &sP, &sM, &gcap))) return rc;
  const bool split_left = (sMain != sM);
  auto width_at = [&](int k) { return lu_step_width(n, n, nb, k); };
  auto steps = [&]() -> int {
    // panel at column k, on the look-ahead stream
    auto P64_of = [&](int par) { return P64 + (size_t)par * n * nb; };
    // (have_p64: the update of the panel's columns already wrote its FP64 copy)
    auto panel = [&](int k, int b, int par, int step, bool have_p64) -> int {
      const int mp = n - k;
      float *Wk = W + (long)k * ldw + k;
      double *P64 = P64_of(par);
      int r;
      if (!have_p64 && (r = k_cvt_f32_f64(sP, mp, b, Wk, ldw, P64, mp))) return r;
      if ((r = lu_panel(sP, P64, mp, b, mp, piv + k, pairs + par * pstride, &nleaves[par], gcap))) return r;
      if ((r = k_panel_compose(sP, pairs + par * pstride, nleaves[par], comp + (size_t)step * cstride)))
        return r;
      if ((r = k_cvt_f64_f32(sP, mp, b, P64, mp, Wk, ldw))) return r;
      return k_panel_out(sP, prec, mp - b, b, (b + 31) / 32 * 32, P64 + b, mp, Lh + par * lbuf,
                         split ? Ll + par * lbuf : nullptr, slice);
    };
    {
      const int b = width_at(0);
      if ((rc = panel(0, b, 0, 0, false))) return rc;
      PLA_HIP(hipEventRecord(ev_panel, sP));
    }
    int step = 0;
    for (int k = 0, b = 0; k < n; k += b, step++) {
      b = width_at(k);
      const int par = step & 1;
      const int kpad = (b + 31) / 32 * 32;  // the update's inner dimension (copies are zero beyond b)
      PLA_HIP(hipStreamWaitEvent(sM, ev_panel, 0));
      const int *cp = comp + (size_t)step * cstride;
      const int nr = n - k - b, mr = nr;
      const int kn = k + b;
      const int bn = kn < n ? width_at(kn) : 0;  // next panel's width
      const int n1 = bn > 0 ? bn : nr;
      float *W12 = W + (long)kn * ldw + k;
      const void *L_h = Lh + par * lbuf;
      const void *L_l = split ? Ll + par * lbuf : nullptr;
      // columns kn + c0 .. kn + c0 + nc: interchanges, U12 = inv(L11) W12 in
      // FP64 (L11: top of panel k's FP64 copy), the low-precision U12 copy,
      // then the update.  The next panel's columns go first: they gate panel
      // k+1; the rest follow once it has been released.
      auto right_part = [&](int c0, int nc) -> int {
        if (nc <= 0) return PLA_SUCCESS;
        int r;
        if ((r = k_panel_swap_f32(sM, W + k, ldw, 0, kn + c0, kn + c0 + nc, cp))) return r;
        float *W12c = W12 + (long)c0 * ldw;
        double *U64c = U64 + (size_t)c0 * b;
        if ((r = k_cvt_f32_f64(sM, b, nc, W12c, ldw, U64c, b))) return r;
        if ((r = trsm_left(sM, true, false, true, b, nc, P64_of(par), n - k, U64c, b, nullptr))) return r;
        if ((r = k_upanel_out(sM, prec, b, kpad, nc, U64c, b, W12c, ldw, Uh, split ? Ul : nullptr, slice, c0)))
          return r;
        timing_begin(sM, 3);
        // the next panel's columns (c0 = 0, bn > 0): the update also writes
        // the next panel's FP64 copy (rows kn.., columns kn..kn+bn, ld mr),
        // so the panel stream starts factoring without a conversion pass
        double *P_next = (c0 == 0 && bn > 0) ? P64_of(par ^ 1) : nullptr;
        r = k_hgemm(sM, prec, mr, nc, kpad, L_h, L_l, slice, Uh, split ? Ul : nullptr, slice, c0, W12c + b, ldw,
                    P_next, mr);
        timing_end(sM, 3, 2.0 * mr * (double)nc * b);
        return r;
      };
      if ((rc = right_part(0, n1))) return rc;
      PLA_HIP(hipEventRecord(ev_ready, sM));
      if (split_left) {  // after the next-panel columns, which gate panel k+1
        PLA_HIP(hipStreamWaitEvent(sMain, ev_ready, 0));
        if ((rc = k_panel_swap_f32(sMain, W + k, ldw, k, 0, 0, cp))) return rc;
      }
      if (bn > 0) {
        PLA_HIP(hipStreamWaitEvent(sP, ev_ready, 0));
        if ((rc = panel(kn, bn, par ^ 1, step + 1, true))) return rc;
        PLA_HIP(hipEventRecord(ev_panel, sP));
      }
      if ((rc = right_part(n1, nr - n1))) return rc;
      // left of the panel last: nothing reads L before the factorization ends
      if (!split_left && (rc = k_panel_swap_f32(sM, W + k, ldw, k, 0, 0, cp))) return rc;
    }
    return PLA_SUCCESS;
  };
  rc = steps();
  // the main stream waits for both streams, also after an error
  const int rc_end = lu_streams_end(sMain, sP, sM);
  if (rc) return rc;
  if (rc_end) return rc_end;
  return leaf_check_error(sMain);
}

}  // namespace pla

extern "C" {

int PLA_LU_mixed(PLA_Obj A, PLA_Obj pivots, PLA_Obj B, PLA_Obj X, int precision, double tol, int max_iter,
                 int *iterations, double *backward_error) {
  if (!valid(A) || !valid(B) || !valid(X) || !valid(pivots))
    return fail(PLA_ERR_INVALID_OBJECT, "PLA_LU_mixed: invalid object");
  if (A->base->datatype != MPI_DOUBLE || B->base->datatype != MPI_DOUBLE || X->base->datatype != MPI_DOUBLE)
    return fail(PLA_ERR_UNSUPPORTED, "PLA_LU_mixed: A, B and X must be MPI_DOUBLE");
  if (pivots->base->datatype != MPI_INT)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_LU_mixed: pivots must be of type MPI_INT");
  const int n = A->length, nrhs = B->width;
  if (A->width != n || B->length != n || X->length != n || X->width != nrhs || pivots->length != n ||
      pivots->width != 1)
    return fail(PLA_ERR_DIMENSIONS, "PLA_LU_mixed: A %dx%d, B %dx%d, X %dx%d, pivots %dx%d", A->length, A->width,
                B->length, B->width, X->length, X->width, pivots->length, pivots->width);
  if (nrhs < 1 || nrhs > trsv_max_rhs())
    return fail(PLA_ERR_UNSUPPORTED, "PLA_LU_mixed: 1 to %d right-hand sides", trsv_max_rhs());
  if (precision < PLA_MIXED_FP16 || precision > PLA_MIXED_BF16X3)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_LU_mixed: unknown precision %d", precision);
  if (max_iter < 0) return fail(PLA_ERR_INVALID_PARAM, "PLA_LU_mixed: max_iter < 0");
  if (n == 0) return PLA_SUCCESS;
  hipStream_t s = ctx().stream;
  double *a, *b, *x;
  int *piv;
  int rc;
  if ((rc = dev_ptr(A, (void **)&a)) || (rc = dev_ptr(B, (void **)&b)) || (rc = dev_ptr(X, (void **)&x)) ||
      (rc = dev_ptr(pivots, (void **)&piv)))
    return rc;
  const long lda = ld(A), ldb = ld(B), ldx = ld(X);

  // factor a single-precision copy
  const long ldw = n;
  float *W = (float *)workspace(sizeof(float) * (size_t)n * n, WS_W);
  if (!W) return PLA_ERR_NO_MEMORY;
  if ((rc = k_cvt_f64_f32(s, n, n, a, lda, W, ldw))) return rc;
  if ((rc = lu_factor_mixed(s, precision, n, W, ldw, piv, choose_nb(n, NB_LU_MIXED)))) return rc;

  // row permutation of the interchanges (position i holds original row perm[i])
  std::vector<int> hp(n), perm(n);
  PLA_HIP(hipMemcpyAsync(hp.data(), piv, sizeof(int) * n, hipMemcpyDeviceToHost, s));
  PLA_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < n; i++) perm[i] = i;
  for (int j = 0; j < n; j++) {
    const int p = j + hp[j];
    if (p < 0 || p >= n) return fail(PLA_ERR_DEVICE, "PLA_LU_mixed: pivot %d out of range", j);
    const int t = perm[j];
    perm[j] = perm[p];
    perm[p] = t;
  }
  const int nblk = (n + trsv_block_size() - 1) / trsv_block_size();
  const size_t dsz = sizeof(double) * (size_t)nblk * trsv_block_size() * trsv_block_size();
  int *dperm = (int *)workspace(sizeof(int) * n, WS_PERM);
  double *DinvL = (double *)workspace(dsz, WS_DINVL);
  double *DinvU = (double *)workspace(dsz, WS_DINVU);
  if (!dperm || !DinvL || !DinvU) return PLA_ERR_NO_MEMORY;
  PLA_HIP(hipMemcpyAsync(dperm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
  if ((rc = k_inv_diag_f32(s, true, true, n, W, ldw, DinvL))) return rc;
  if ((rc = k_inv_diag_f32(s, false, false, n, W, ldw, DinvU))) return rc;

  auto solve = [&](const double *rhs, long ldr, double *out, long ldo) -> int {
    int r;
    if ((r = k_permute_rows(s, n, nrhs, dperm, rhs, ldr, out, ldo))) return r;
    if ((r = trsv_sf_f32(s, true, false, true, n, nrhs, W, ldw, out, ldo, DinvL))) return r;
    return trsv_sf_f32(s, false, false, false, n, nrhs, W, ldw, out, ldo, DinvU);
  };

  double anorm = 0.0;
  if ((rc = k_norm(s, true, n, n, a, lda, &anorm))) return rc;
  refine_ops op;
  op.n = n;
  op.nrhs = nrhs;
  op.residual = [&](const double *Xv, long ldxv, const double *Bv, long ldbv, double *Rv, long ldr, double *rmax) {
    return residual(s, n, nrhs, a, lda, Xv, ldxv, Bv, ldbv, Rv, ldr, rmax);
  };
  op.precond = solve;
  refine_stats st;
  if ((rc = refine(s, op, anorm, b, ldb, x, ldx, tol, max_iter, true, &st))) return rc;
  PLA_HIP(hipStreamSynchronize(s));
  if ((rc = leaf_check_error(s))) return rc;
  set_last_refine(st, 0.0);
  if (iterations) *iterations = st.outer;
  if (backward_error) *backward_error = st.eta;
  if ((rc = written(X)) || (rc = written(pivots))) return rc;
  return refine_outcome("PLA_LU_mixed", st, tol);
}

}  // extern "C"
