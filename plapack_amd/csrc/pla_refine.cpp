// pla_refine.cpp -- iterative refinement in FP64 with low-precision factors:
// classic refinement, falling back to GMRES-IR (GMRES on the left-
// preconditioned system M^{-1} A d = M^{-1} r, M = the low-precision
// factorization) when classic refinement stops converging.
//
// No counterpart in the reference.  Every vector operation runs on the GPU
// in FP64 (residuals with compensated chunk sums); only GMRES's small
// Hessenberg least-squares problem (at most `restart` columns per
// right-hand side) is kept on the host.
#include "pla_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace pla {

namespace {

double eta_of(int nrhs, double anorm, const std::vector<double> &rmax, const std::vector<double> &xmax,
              const std::vector<double> &bmax) {
  double eta = 0.0;
  for (int q = 0; q < nrhs; q++) {
    const double den = anorm * xmax[q] + bmax[q];
    const double e = den > 0.0 ? rmax[q] / den : (rmax[q] > 0.0 ? INFINITY : 0.0);
    if (!(e <= eta)) eta = e;  // NaN propagates
  }
  return eta;
}

int copy_block(hipStream_t s, int n, int nrhs, const double *X, long ldx, double *Y, long ldy) {
  PLA_HIP(hipMemcpy2DAsync(Y, sizeof(double) * ldy, X, sizeof(double) * ldx, sizeof(double) * n, nrhs,
                           hipMemcpyDeviceToDevice, s));
  return PLA_SUCCESS;
}

// D = GMRES(M^{-1} A, M^{-1} R) for nrhs systems in lockstep, one cycle of
// at most `m` steps (CGS2 orthogonalization), stopping when every preconditioned
// residual estimate is below tau times its initial value.  D: n x nrhs (ld n).
int gmres(hipStream_t s, const refine_ops &op, const double *R, long ldr, double *D, int m, double tau,
          int *steps) {
  const int n = op.n, nrhs = op.nrhs;
  const long vs = (long)n * nrhs;  // one block vector
  double *V = (double *)workspace(sizeof(double) * (size_t)vs * (m + 1), WS_GM_V);
  double *W = (double *)workspace(sizeof(double) * (size_t)vs, WS_GM_W);
  double *T = (double *)workspace(sizeof(double) * (size_t)vs, WS_GM_T);
  double *hd = (double *)workspace(sizeof(double) * (size_t)(m + 2) * nrhs * 2, WS_GM_H);
  if (!V || !W || !T || !hd) return PLA_ERR_NO_MEMORY;
  double *fd = hd + (size_t)(m + 2) * nrhs;  // per-column factors
  int rc;
  std::vector<double> hh((size_t)(m + 1) * nrhs), beta(nrhs), f(nrhs);
  // v0 = M^{-1} R / beta
  if ((rc = op.precond(R, ldr, V, n))) return rc;
  if ((rc = k_mdot(s, n, nrhs, 1, V, vs, n, V, n, hd))) return rc;
  PLA_HIP(hipMemcpyAsync(hh.data(), hd, sizeof(double) * nrhs, hipMemcpyDeviceToHost, s));
  PLA_HIP(hipStreamSynchronize(s));
  for (int q = 0; q < nrhs; q++) beta[q] = std::sqrt(hh[q]);
  PLA_HIP(hipMemcpyAsync(fd, beta.data(), sizeof(double) * nrhs, hipMemcpyHostToDevice, s));
  if ((rc = k_scale_cols(s, n, nrhs, V, n, fd, true))) return rc;
  // Hessenberg columns (upper triangle after rotations), rotations, rhs g
  std::vector<std::vector<double>> H(nrhs, std::vector<double>((size_t)(m + 1) * m, 0.0));
  std::vector<std::vector<double>> cs(nrhs, std::vector<double>(m)), sn(nrhs, std::vector<double>(m));
  std::vector<std::vector<double>> g(nrhs, std::vector<double>(m + 1, 0.0));
  std::vector<char> done(nrhs, 0);
  for (int q = 0; q < nrhs; q++) {
    g[q][0] = beta[q];
    if (!(beta[q] > 0.0)) done[q] = 1;
  }
  int k = 0;
  for (int j = 0; j < m; j++) {
    bool all = true;
    for (int q = 0; q < nrhs; q++) all = all && done[q];
    if (all) break;
    double *Vj = V + (size_t)j * vs;
    // W = M^{-1} A v_j: T = -A v_j, W = M^{-1} T, W = -W
    if ((rc = op.residual(Vj, n, nullptr, 0, T, n, nullptr))) return rc;
    if ((rc = op.precond(T, n, W, n))) return rc;
    for (int q = 0; q < nrhs; q++) f[q] = -1.0;
    PLA_HIP(hipMemcpyAsync(fd, f.data(), sizeof(double) * nrhs, hipMemcpyHostToDevice, s));
    if ((rc = k_scale_cols(s, n, nrhs, W, n, fd, false))) return rc;
    // classical Gram-Schmidt, twice
    std::vector<double> h((size_t)(j + 1) * nrhs, 0.0);
    for (int pass = 0; pass < 2; pass++) {
      if ((rc = k_mdot(s, n, nrhs, j + 1, V, vs, n, W, n, hd))) return rc;
      if ((rc = k_maxpy(s, n, nrhs, j + 1, V, vs, n, hd, -1.0, W, n))) return rc;
      PLA_HIP(hipMemcpyAsync(hh.data(), hd, sizeof(double) * (j + 1) * nrhs, hipMemcpyDeviceToHost, s));
      PLA_HIP(hipStreamSynchronize(s));
      for (size_t e = 0; e < h.size(); e++) h[e] += hh[e];
    }
    if ((rc = k_mdot(s, n, nrhs, 1, W, vs, n, W, n, hd))) return rc;
    PLA_HIP(hipMemcpyAsync(hh.data(), hd, sizeof(double) * nrhs, hipMemcpyDeviceToHost, s));
    PLA_HIP(hipStreamSynchronize(s));
    std::vector<double> hn(nrhs);
    for (int q = 0; q < nrhs; q++) hn[q] = std::sqrt(hh[q]);
    // v_{j+1} = W / ||W||
    PLA_HIP(hipMemcpyAsync(fd, hn.data(), sizeof(double) * nrhs, hipMemcpyHostToDevice, s));
    if ((rc = copy_block(s, n, nrhs, W, n, V + (size_t)(j + 1) * vs, n))) return rc;
    if ((rc = k_scale_cols(s, n, nrhs, V + (size_t)(j + 1) * vs, n, fd, true))) return rc;
    for (int q = 0; q < nrhs; q++) {
      std::vector<double> &Hq = H[q];
      double *col = &Hq[(size_t)j * (m + 1)];
      if (done[q]) continue;  // its solution is fixed; later columns are not used
      for (int i = 0; i <= j; i++) col[i] = h[(size_t)i * nrhs + q];
      col[j + 1] = hn[q];
      for (int i = 0; i < j; i++) {  // earlier rotations
        const double a = col[i], b = col[i + 1];
        col[i] = cs[q][i] * a + sn[q][i] * b;
        col[i + 1] = -sn[q][i] * a + cs[q][i] * b;
      }
      const double a = col[j], b = col[j + 1];
      const double r = std::hypot(a, b);
      cs[q][j] = r > 0.0 ? a / r : 1.0;
      sn[q][j] = r > 0.0 ? b / r : 0.0;
      col[j] = r;
      col[j + 1] = 0.0;
      g[q][j + 1] = -sn[q][j] * g[q][j];
      g[q][j] = cs[q][j] * g[q][j];
    }
    k = j + 1;
    for (int q = 0; q < nrhs; q++) {
      if (done[q]) continue;
      if (std::fabs(g[q][j + 1]) <= tau * beta[q] || !(hn[q] > 0.0)) done[q] = 2 + j;  // converged at step j+1
    }
  }
  // y = H^{-1} g per column (k_q = steps that column took), D = V y
  std::vector<double> y((size_t)k * nrhs, 0.0);
  for (int q = 0; q < nrhs; q++) {
    int kq = k;
    if (done[q] >= 2) kq = done[q] - 1;
    if (done[q] == 1) kq = 0;
    const std::vector<double> &Hq = H[q];
    for (int i = kq - 1; i >= 0; i--) {
      double t = g[q][i];
      for (int c = i + 1; c < kq; c++) t -= Hq[(size_t)c * (m + 1) + i] * y[(size_t)c * nrhs + q];
      const double d = Hq[(size_t)i * (m + 1) + i];
      y[(size_t)i * nrhs + q] = d != 0.0 ? t / d : 0.0;
    }
  }
  if ((rc = k_fill(s, vs, 0.0, D))) return rc;
  if (k > 0) {
    PLA_HIP(hipMemcpyAsync(hd, y.data(), sizeof(double) * k * nrhs, hipMemcpyHostToDevice, s));
    if ((rc = k_maxpy(s, n, nrhs, k, V, vs, n, hd, 1.0, D, n))) return rc;
  }
  *steps = k;
  return PLA_SUCCESS;
}

}  // namespace

int refine(hipStream_t s, const refine_ops &op, double anorm, const double *B, long ldb, double *X, long ldx,
           double tol, int max_iter, bool gmres_fallback, refine_stats *st) {
  const int n = op.n, nrhs = op.nrhs;
  double *Rb = (double *)workspace(sizeof(double) * (size_t)n * nrhs, WS_CH_R);
  double *Db = (double *)workspace(sizeof(double) * (size_t)n * nrhs, WS_CH_D);
  if (!Rb || !Db) return PLA_ERR_NO_MEMORY;
  int rc;
  std::vector<double> bmax(nrhs), xmax(nrhs), rmax(nrhs);
  if ((rc = col_amax(s, n, nrhs, B, ldb, bmax.data()))) return rc;
  if ((rc = op.precond(B, ldb, X, ldx))) return rc;
  refine_stats out;
  bool use_gmres = gmres_fallback && op.gmres_first;
  double prev = INFINITY;
  for (;;) {
    if ((rc = op.residual(X, ldx, B, ldb, Rb, n, rmax.data()))) return rc;
    if ((rc = col_amax(s, n, nrhs, X, ldx, xmax.data()))) return rc;
    double eta = eta_of(nrhs, anorm, rmax, xmax, bmax);
    if (!use_gmres && gmres_fallback && out.outer > 0 && !(eta <= 0.5 * prev)) {
      // classic refinement stalled or diverged: undo a step that made things
      // worse, then continue with GMRES-IR
      use_gmres = true;
      if (!(eta <= prev)) {
        std::vector<double> m1(nrhs, -1.0);
        double *fd = (double *)workspace(sizeof(double) * nrhs, WS_GM_Z);
        if (!fd) return PLA_ERR_NO_MEMORY;
        PLA_HIP(hipMemcpyAsync(fd, m1.data(), sizeof(double) * nrhs, hipMemcpyHostToDevice, s));
        if ((rc = k_scale_cols(s, n, nrhs, Db, n, fd, false))) return rc;
        if ((rc = k_axpy_cols(s, n, nrhs, Db, n, X, ldx))) return rc;
        if ((rc = op.residual(X, ldx, B, ldb, Rb, n, rmax.data()))) return rc;
        if ((rc = col_amax(s, n, nrhs, X, ldx, xmax.data()))) return rc;
        eta = eta_of(nrhs, anorm, rmax, xmax, bmax);
      }
    }
    out.eta = eta;
    if (eta <= tol || out.outer >= max_iter) break;
    if (use_gmres) {
      // the cycle only needs to take eta below tol: its relative tolerance
      // follows from the current eta, with a factor 100 of margin for the
      // gap between the preconditioned residual GMRES sees and eta
      const double tau = std::min(1e-2, std::max(op.gmres_tol, 1e-2 * tol / eta));
      int steps = 0;
      if ((rc = gmres(s, op, Rb, n, Db, op.restart, tau, &steps))) return rc;
      out.inner += steps;
      out.gmres_outer++;
    } else {
      if ((rc = op.precond(Rb, n, Db, n))) return rc;
    }
    if ((rc = k_axpy_cols(s, n, nrhs, Db, n, X, ldx))) return rc;
    prev = eta;
    out.outer++;
  }
  if (st) *st = out;
  return PLA_SUCCESS;
}

int refine_outcome(const char *who, const refine_stats &st, double tol) {
  if (st.eta <= tol) return PLA_SUCCESS;
  return fail(PLA_ERR_NOT_CONVERGED,
              "%s: refinement stopped at backward error %.3e above the tolerance %.3e after %d corrections (%d GMRES "
              "steps); X holds the last iterate",
              who, st.eta, tol, st.outer, st.inner);
}

}  // namespace pla
