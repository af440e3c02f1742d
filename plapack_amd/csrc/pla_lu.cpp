// pla_lu.cpp -- PLA_LU and PLA_General_solve on a 1x1 template, device resident
// (MPI_DOUBLE here; MPI_DOUBLE_COMPLEX objects go to pla_zfactor.cpp).
//
// Reference: SOLVERS/LU/PLA_LU.c:13-212 (blocked right-looking LU with
// partial pivoting), SOLVERS/LU/PLA_General.c:67-119 (factor + solve).
//
// Per block step k of width b (PLA_LU.c:103-192):
//   1. lu_panel on A[k:m, k:k+b]          (PLA_LU_right_mv, :144)
//   2. row interchanges on the columns left and right of the panel
//                                           (PLA_Apply_pivots_to_rows, :152)
//   3. A12 <- inv(L11) A12                  (PLA_Local_trsm, :174-182)
//   4. A22 <- A22 - A21 A12                 (PLA_Local_gemm, :185-186)
// The panel copies PLA_Copy makes between distributions (:115-118, :160-165)
// vanish: every operand is a view into the one HBM buffer.
#include "pla_internal.hpp"

using namespace pla;

static int check_lu_args(PLA_Obj A, PLA_Obj pivots) {
  // SOLVERS/LU/PLA_LU_enter_exit.c:17-105
  if (!valid(A)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_LU: invalid A");
  if (A->base->datatype != MPI_DOUBLE && A->base->datatype != MPI_DOUBLE_COMPLEX)
    return fail(PLA_ERR_UNSUPPORTED, "PLA_LU: only MPI_DOUBLE and MPI_DOUBLE_COMPLEX matrices are supported");
  if (A->objtype != PLA_MATRIX) return fail(PLA_ERR_INVALID_PARAM, "PLA_LU: A must be a PLA_MATRIX");
  if (!valid(pivots)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_LU: invalid pivots");
  if (pivots->base->datatype != MPI_INT)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_LU: pivots must be of type MPI_INT");
  const int mn = A->length < A->width ? A->length : A->width;
  if (pivots->length != mn || pivots->width != 1)
    return fail(PLA_ERR_DIMENSIONS, "PLA_LU: pivots must be %d x 1, got %d x %d", mn, pivots->length,
                pivots->width);
  return PLA_SUCCESS;
}

namespace pla {

struct lu_bufs {
  int *pairs;      // leaves' pair lists, two parities of pstride
  size_t pstride;
  int *comp;       // composite interchange list of each step, cstride apart
  int cstride;
};

static int lu_factor_steps(hipStream_t sP, hipStream_t sM, hipStream_t sL, int gcap, int m, int n, double *a,
                           long lda, int *piv, int nb, bool swap_left, const lu_bufs &bf, hipEvent_t ev_ready,
                           hipEvent_t ev_panel);

// Blocked right-looking LU of the m x n column-major device matrix a, with
// one step of look-ahead: panel k+1 is factored on the look-ahead stream as
// soon as its own columns have received step k's update, while the main
// stream applies the rest of step k's trailing update.
//
// main stream, step k:  wait(panel k) -> full-width interchanges of panel k
//                       -> A12 <- inv(L11) A12 -> A22[:, next panel] -= A21 A12
//                       -> signal -> A22[:, rest] -= A21 A12
// look-ahead stream:    wait(signal) -> factor panel k+1 -> signal(panel k+1)
//
// swap_left = false is the LINPACK variant (SOLVERS/LINPACK/PLA_LU_linpack.c:
// 121-123 interchanges only A_BR, the columns from the panel on): the columns
// of L left of a panel keep the row order of their own step.
int lu_factor(hipStream_t sMain, int m, int n, double *a, long lda, int *piv, int nb, bool swap_left) {
  if (nb > 1024) nb = 1024;
  if (nb < 1) nb = 1;
  const int stride = leaf_pairs_stride();
  const int cstride = panel_comp_stride();
  const size_t pstride = (size_t)LU_MAX_LEAVES * stride;  // one parity of the leaves' pair lists
  int *pairs = (int *)workspace(sizeof(int) * 2 * pstride, 3, true);
  // one composite interchange list per step: the left-of-panel interchanges
  // of a step may still be reading theirs while later panels are composed
  int *comp = (int *)workspace(sizeof(int) * (size_t)lu_step_count(m, m < n ? m : n, nb) * cstride, 4);
  if (!pairs || !comp) return PLA_ERR_NO_MEMORY;
  const lu_bufs bf{pairs, pstride, comp, cstride};
  hipEvent_t ev_ready = lib_event(EV_LU_READY), ev_panel = lib_event(EV_LU_PANEL);
  if (!ev_ready || !ev_panel) return fail(PLA_ERR_NOT_INITIALIZED, "PLAPACK not yet initialized");
  int rc;
  // sP: panels; sM: interchanges right of the panel, A12 solves and trailing
  // updates; the main stream: interchanges left of the panel
  hipStream_t sP, sM;
  int gcap;
  if ((rc = lu_streams_begin(sMain, m, &sP, &sM, &gcap))) return rc;
  rc = lu_factor_steps(sP, sM, sMain, gcap, m, n, a, lda, piv, nb, swap_left, bf, ev_ready, ev_panel);
  // the main stream waits for both streams, also after an error
  const int rc_end = lu_streams_end(sMain, sP, sM);
  if (rc) return rc;
  if (rc_end) return rc_end;
  return leaf_check_error(sMain);
}

// The interchanges of panel k on the columns left of it go to sL when it
// differs from sM, off the trailing-update stream, else to the end of step k
// on sM.  Those columns include panel k-1's L21, which step k-1's trailing
// update reads: sL waits for step k's next-panel columns on sM (ev_ready),
// which follow the whole of step k-1 there, before permuting them.  Step k's
// composite list is comp + k * cstride.
static int lu_factor_steps(hipStream_t sP, hipStream_t sM, hipStream_t sL, int gcap, int m, int n, double *a,
                           long lda, int *piv, int nb, bool swap_left, const lu_bufs &bf, hipEvent_t ev_ready,
                           hipEvent_t ev_panel) {
  int *const pairs = bf.pairs, *const comp = bf.comp;
  const size_t pstride = bf.pstride;
  const int cstride = bf.cstride;
  // panel at column k, b columns, leaves' lists of parity par: factor, then
  // compose its interchanges into list `step`
  auto panel = [&](int k, int b, int par, int step, int *nl) -> int {
    int r;
    if ((r = lu_panel(sP, a + (long)k * lda + k, m - k, b, lda, piv + k, pairs + par * pstride, nl, gcap)))
      return r;
    return k_panel_compose(sP, pairs + par * pstride, *nl, comp + (size_t)step * cstride);
  };
  const bool split_left = (sL != sM);
  const int mn = m < n ? m : n;
  int rc;
  int nleaves[2] = {0, 0};
  // block size of the step at column k: nb, then the tail block size once at
  // most nb_switch rows remain (a narrower panel shortens the critical path
  // when the trailing update no longer hides it)
  auto width_at = [&](int k) { return lu_step_width(m, mn, nb, k); };
  // panel 0, after everything already queued on the main stream
  {
    const int b = width_at(0);
    if ((rc = panel(0, b, 0, 0, &nleaves[0]))) return rc;
    PLA_HIP(hipEventRecord(ev_panel, sP));
  }
  int step = 0;
  for (int k = 0, b = 0; k < mn; k += b, step++) {
    b = width_at(k);
    const int par = step & 1;
    double *Akk = a + (long)k * lda + k;
    PLA_HIP(hipStreamWaitEvent(sM, ev_panel, 0));
    const int nr = n - k - b, mr = m - k - b;
    const int *cp = comp + (size_t)step * cstride;
    // Right of the panel, the next panel's columns go first (interchanges,
    // A12 solve, update): they gate panel k+1.  The other columns follow
    // once panel k+1 has been released.
    const int kn = k + b;
    const int bn = (kn < mn && mr > 0) ? width_at(kn) : 0;  // next panel's width
    const int n1 = bn > 0 ? bn : nr;
    double *A12 = a + (long)kn * lda + k;
    // columns kn + c0 .. kn + c0 + nc: interchanges, then A12 solve and update
    auto swap_part = [&](int c0, int nc) -> int {
      return nc > 0 ? k_panel_swap(sM, a + k, lda, 0, kn + c0, kn + c0 + nc, cp) : PLA_SUCCESS;
    };
    auto update_part = [&](int c0, int nc) -> int {
      if (nc <= 0) return PLA_SUCCESS;
      int r;
      double *A12c = A12 + (long)c0 * lda;
      if ((r = trsm_left(sM, true, false, true, b, nc, Akk, lda, A12c, lda, nullptr))) return r;
      if (mr <= 0) return PLA_SUCCESS;
      timing_begin(sM, 1);
      r = k_gemm(sM, 1, false, false, mr, nc, b, -1.0, Akk + b, lda, A12c, lda, 1.0, A12c + b, lda, nullptr);
      timing_end(sM, 1, 2.0 * mr * (double)nc * b);
      return r;
    };
    if ((rc = swap_part(0, n1)) || (rc = update_part(0, n1))) return rc;
    PLA_HIP(hipEventRecord(ev_ready, sM));
    if (split_left && swap_left) {
      // After step k's next-panel columns: by then step k-1's update no
      // longer reads panel k-1's L21, and the swap does not compete for CUs
      // with the short kernels that gate panel k+1.
      PLA_HIP(hipStreamWaitEvent(sL, ev_ready, 0));
      if ((rc = k_panel_swap(sL, a + k, lda, k, 0, 0, cp))) return rc;
    }
    if (bn > 0) {
      PLA_HIP(hipStreamWaitEvent(sP, ev_ready, 0));
      if ((rc = panel(kn, bn, par ^ 1, step + 1, &nleaves[par ^ 1]))) return rc;
      PLA_HIP(hipEventRecord(ev_panel, sP));
    }
    if ((rc = swap_part(n1, nr - n1)) || (rc = update_part(n1, nr - n1))) return rc;
    // left of the panel last: nothing reads L before the factorization ends
    if (swap_left && !split_left && (rc = k_panel_swap(sM, a + k, lda, k, 0, 0, cp))) return rc;
  }
  return PLA_SUCCESS;
}

}  // namespace pla

extern "C" {

static int lu_entry(PLA_Obj A, PLA_Obj pivots, bool swap_left) {
  int rc = check_lu_args(A, pivots);
  if (rc) return rc;
  const int m = A->length, n = A->width;
  const int mn = m < n ? m : n;
  if (mn == 0) return PLA_SUCCESS;
  double *a;
  int *piv;
  if ((rc = dev_ptr(A, (void **)&a)) || (rc = dev_ptr(pivots, (void **)&piv))) return rc;
  if (A->base->datatype == MPI_DOUBLE_COMPLEX) {
    if (!swap_left) return fail(PLA_ERR_UNSUPPORTED, "PLA_LU_linpack: MPI_DOUBLE_COMPLEX is not supported");
    const int nbz = ctx().nb_device > 0 ? ctx().nb_device : (ctx().nb_user > 0 ? ctx().nb_user : 512);
    if ((rc = zlu_factor(ctx().stream, m, n, a, ld(A), piv, nbz))) return rc;
    if ((rc = written(A))) return rc;
    return written(pivots);
  }
  const int nb = choose_nb(mn, NB_LU);
  if ((rc = lu_factor(ctx().stream, m, n, a, ld(A), piv, nb, swap_left))) return rc;
  if ((rc = written(A))) return rc;
  return written(pivots);
}

int PLA_LU(PLA_Obj A, PLA_Obj pivots) {
  if (is_single(A)) return on_promoted({&A}, {true}, [&] { return PLA_LU(A, pivots); });
  return lu_entry(A, pivots, true);
}

// SOLVERS/LINPACK/PLA_LU_linpack.c:13: the same U, pivots and trailing
// columns as PLA_LU (on the n x (n+1) appended system [A b] the last column
// ends as inv(L) P b); the multipliers of each panel stay in the row order of
// their own step.
int PLA_LU_linpack(PLA_Obj A, PLA_Obj pivots) {
  if (is_single(A)) return on_promoted({&A}, {true}, [&] { return PLA_LU_linpack(A, pivots); });
  return lu_entry(A, pivots, false);
}

// Factor A (PLA_LU), then solve A X = B in place: interchanges on B, then the
// unit-lower and upper triangular solves (PLA_General.c:80-111).
int PLA_General_solve(PLA_Obj A, PLA_Obj ipiv, PLA_Obj B) {
  if (is_single(A) || is_single(B))
    return on_promoted({&A, &B}, {true, true}, [&] { return PLA_General_solve(A, ipiv, B); });
  int rc = PLA_LU(A, ipiv);
  if (rc) return rc;
  if (!valid(B) || B->base->datatype != A->base->datatype)
    return fail(PLA_ERR_INVALID_OBJECT, "PLA_General_solve: invalid B (or not of A's datatype)");
  const int n = A->length;
  if (A->width != n || B->length != n)
    return fail(PLA_ERR_DIMENSIONS, "PLA_General_solve: A %d x %d, B %d x %d", A->length, A->width,
                B->length, B->width);
  double *a, *b;
  int *piv;
  if ((rc = dev_ptr(A, (void **)&a)) || (rc = dev_ptr(B, (void **)&b)) ||
      (rc = dev_ptr(ipiv, (void **)&piv)))
    return rc;
  hipStream_t s = ctx().stream;
  if (A->base->datatype == MPI_DOUBLE_COMPLEX) {
    if ((rc = zlu_solve(s, n, B->width, a, ld(A), piv, b, ld(B)))) return rc;
    return written(B);
  }
  if ((rc = k_laswp(s, b, ld(B), n, B->width, n, piv))) return rc;
  if (B->width <= trsv_max_rhs()) {
    if ((rc = trsv_sf(s, true, false, true, n, B->width, a, ld(A), b, ld(B), nullptr))) return rc;
    if ((rc = trsv_sf(s, false, false, false, n, B->width, a, ld(A), b, ld(B), nullptr))) return rc;
    if ((rc = leaf_check_error(s, "PLA_General_solve: triangular solve"))) return rc;
  } else {
    if ((rc = trsm_left(s, true, false, true, n, B->width, a, ld(A), b, ld(B), nullptr))) return rc;
    if ((rc = trsm_left(s, false, false, false, n, B->width, a, ld(A), b, ld(B), nullptr))) return rc;
  }
  return written(B);
}

// The block-size knob of TEMPLATE/PLA_Environ.c:88-127 (nb_alg), tuned on the
// device: the factorization runs on a scratch n x n matrix (TEST/create_problem.c's
// drand48 input; for Cholesky plus n I, TEST/test_chol.c:121-127) once per
// candidate, and the fastest block size becomes this process's choice for
// order n (after PLA_Set_device_nb and pla_Environ_set_nb_alg, which still
// win).  ms (optional) receives the times of the candidates 128, 256, 512, 1024
// (0 where not tried; 1024 for PLA_LU from n = 16384).  Nothing of the caller's
// is read or written.
int PLA_Autotune_nb(int factorization, int n, int *best_nb, double *ms) {
  if (!best_nb) return fail(PLA_ERR_INVALID_PARAM, "PLA_Autotune_nb: best_nb is NULL");
  if (factorization != PLA_NB_LU && factorization != PLA_NB_CHOL)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Autotune_nb: factorization %d (PLA_NB_LU or PLA_NB_CHOL)",
                factorization);
  if (n < 1) return fail(PLA_ERR_DIMENSIONS, "PLA_Autotune_nb: n = %d", n);
  int rc = ensure_device();
  if (rc) return rc;
  const int cand[4] = {128, 256, 512, 1024};
  const int ncand = (factorization == PLA_NB_LU && n >= 16384) ? 4 : 3;
  double *a = nullptr;
  int *piv = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipStream_t s = ctx().stream;
  auto cleanup = [&]() {
    if (a) (void)hipFree(a);
    if (piv) (void)hipFree(piv);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
  if (hipMalloc(&a, sizeof(double) * (size_t)n * n) != hipSuccess ||
      hipMalloc(&piv, sizeof(int) * (size_t)n) != hipSuccess) {
    cleanup();
    return fail(PLA_ERR_NO_MEMORY, "PLA_Autotune_nb: scratch %d x %d", n, n);
  }
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    cleanup();
    return fail(PLA_ERR_DEVICE, "PLA_Autotune_nb: events");
  }
  int best = 0;
  float best_ms = 0.0f;
  for (int c = 0; c < 4; c++) {
    if (ms) ms[c] = 0.0;
    if (c >= ncand) continue;
    const int nb = cand[c];
    // srand48(0): state (0 << 16) | 0x330E
    if ((rc = k_fill_drand48(s, n, n, a, n, 0x330EULL, 0))) break;
    if (factorization == PLA_NB_CHOL && (rc = k_shift_diag(s, n, a, n, (double)n))) break;
    if (hipEventRecord(e0, s) != hipSuccess) { rc = fail(PLA_ERR_DEVICE, "PLA_Autotune_nb: event"); break; }
    int info = 0;
    if (factorization == PLA_NB_LU) rc = lu_factor(s, n, n, a, n, piv, nb, true);
    else rc = chol_factor(s, n, a, n, nb, &info);
    if (rc) break;
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
      rc = fail(PLA_ERR_DEVICE, "PLA_Autotune_nb: event");
      break;
    }
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, e0, e1);
    if (ms) ms[c] = t;
    if (!best || t < best_ms) {
      best = nb;
      best_ms = t;
    }
  }
  cleanup();
  if (rc) return rc;
  set_tuned_nb(factorization, n, best);
  *best_nb = best;
  return PLA_SUCCESS;
}

}  // extern "C"
