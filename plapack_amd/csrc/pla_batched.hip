// pla_batched.hip -- thousands of independent SPD factor-and-solves of uneven
// sizes in one persistent kernel, replayed from a captured hipGraph.
//
// No counterpart in the reference, which factors one distributed matrix at a
// time (SOLVERS/Cholesky/PLA_Chol.c:13-234, PLA_Pos_def.c:44-99).  Each
// system is solved with the same mathematics as PLA_Pos_def_solve: lower
// Cholesky A = L L^T, then L y = b and L^T x = y.
//
// Schedule.  Systems are sorted by size, largest first, and a persistent grid
// of workgroups pulls them from a device counter: the greedy
// longest-processing-time order that keeps every CU busy to the end.
//
// Per system (one workgroup of 256 threads, 4 waves):
//   factor, left-looking by block columns of NB:
//     1. C = A[k0:n, k0:k0+w] - L[k0:n, 0:k0] L[k0:k0+w, 0:k0]^T on MFMA
//        (16x16x4 tiles, FP64 or FP32), written into LDS;
//     2. the tall block column is factored in LDS, one barrier per column
//        (Schur complement form; scaled by 1/sqrt(pivot) when written out);
//     3. the block column of L is written once.
//   Each element of L is written once, and the MFMA update reads earlier
//   columns of L from L2 -- left-looking moves n^3/(6 NB) words per system
//   instead of the right-looking n^3/NB.
//   solve: blocked forward / backward substitution with the right-hand sides
//   in LDS (16-row diagonal blocks solved by wave shuffles).
// Modes:
//   PLA_BATCHED_FP64:    FP64 factor in place (A's lower triangle becomes L)
//                        and FP64 solves; B becomes X.
//   PLA_BATCHED_FP32_IR: FP32 factor into a workspace (A unchanged), then
//                        FP64 refinement: r = b - A x (FP64, lower triangle
//                        of A), d = (L L^T)^{-1} r, x += d, until the normwise
//                        backward error is <= tol.
// Upper triangles of A are never read or written.
#include "pla_internal.hpp"

#include <algorithm>
#include <map>
#include <numeric>
#include <tuple>
#include <vector>

using namespace pla;

namespace {

struct bsys {
  long a_off, b_off;  // element offsets of A_i (n x n, ld n) and B_i (n x nrhs, ld n)
  int n, id;          // order, index in the caller's order
};

constexpr int BT = 256;      // threads per workgroup
constexpr int MAX_RHS = 8;

template <typename LT> struct mfma16;
template <> struct mfma16<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 op(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // lane l, register r of the result holds row (l/16 + 4r), column l%16
  static __device__ inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct mfma16<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 op(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // lane l, register r of the result holds row (4 (l/16) + r), column l%16
  static __device__ inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

struct batched_args {
  const bsys *sys;
  int count;
  double *A, *B;
  float *L32;        // FP32 factors (mode FP32_IR), same offsets as A
  int nrhs, max_iter;
  double tol;
  int *info, *iters;
  double *berr;
  unsigned *counter;
  int rs;            // LDS stride (elements) of a panel column / right-hand side
};

// A is read once per element; reading it around L1 keeps the cache free of
// lines that the same workgroup later overwrites with L (FP64 mode).
__device__ inline double ld_l2(const double *p) {
  const unsigned long long v = __hip_atomic_load((unsigned long long *)p, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double((long long)v);
}

__device__ inline void wait_stores_and_sync() {
  __builtin_amdgcn_s_waitcnt(0);  // every outstanding load and store has completed
  __syncthreads();
}

// Left-looking blocked Cholesky of system d into L (ld n).  Returns 0, or the
// 1-based index of the first non-positive pivot (as PLA_Chol).
// BQ: row tiles of 16 per wave per pass over the block column (4 waves cover
// 64 BQ rows per pass); more tiles reuse each L fragment more, fewer leave
// registers for a second workgroup per CU.
template <typename LT, int NB, int BQ>
__device__ int factor_system(const bsys &d, const batched_args &a, LT *L, LT *P, LT *rsq) {
  using M = mfma16<LT>;
  typedef typename M::v4 v4;
  constexpr int CT = NB / 16;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, l16 = lane & 15;
  const int n = d.n;
  const long ld = n;
  const double *Ad = a.A + d.a_off;
  const int RS = a.rs;
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int w = min(NB, n - k0), R = n - k0;
    const int nrt = (R + 15) >> 4;
    // 1. MFMA update of the block column into P (R x w, column stride RS)
    for (int rt0 = 0; rt0 < nrt; rt0 += 4 * BQ) {
      const int rtb = rt0 + wave;
      const int Q = rtb < nrt ? min(BQ, (nrt - rtb + 3) >> 2) : 0;  // uniform per wave
      v4 acc[BQ][CT];
#pragma unroll
      for (int q = 0; q < BQ; q++)
#pragma unroll
        for (int c = 0; c < CT; c++) acc[q][c] = v4{0, 0, 0, 0};
      if (Q > 0 && k0 > 0) {
        // k0 is a multiple of NB (>= 16): groups of U = 4 MFMA k-steps (16
        // columns of L), the next group's fragments in flight while the
        // current group is multiplied
        constexpr int U = 4;
        LT xa0[U][BQ], yb0[U][CT], xa1[U][BQ], yb1[U][CT];
        auto load = [&](int p0, LT (&xa)[U][BQ], LT (&yb)[U][CT]) {
          // every load from a clamped position with no condition, the
          // selects after them: a guarded or selected load is waited for
          // where it issues
#pragma unroll
          for (int u = 0; u < U; u++) {
            const LT *Lc = L + (long)(p0 + 4 * u + (lane >> 4)) * ld;
#pragma unroll
            for (int c = 0; c < CT; c++) yb[u][c] = Lc[min(k0 + c * 16 + l16, k0 + w - 1)];
#pragma unroll
            for (int q = 0; q < BQ; q++) xa[u][q] = Lc[min(k0 + (rtb + 4 * q) * 16 + l16, n - 1)];
          }
#pragma unroll
          for (int u = 0; u < U; u++) {
#pragma unroll
            for (int c = 0; c < CT; c++) yb[u][c] = c * 16 + l16 < w ? yb[u][c] : (LT)0;
#pragma unroll
            for (int q = 0; q < BQ; q++) {
              const int i = k0 + (rtb + 4 * q) * 16 + l16;
              xa[u][q] = (q < Q && i < n) ? xa[u][q] : (LT)0;
            }
          }
        };
        auto mul = [&](LT (&xa)[U][BQ], LT (&yb)[U][CT]) {
#pragma unroll
          for (int u = 0; u < U; u++)
#pragma unroll
            for (int q = 0; q < BQ; q++)
              if (q < Q) {
#pragma unroll
                for (int c = 0; c < CT; c++) acc[q][c] = M::op(yb[u][c], xa[u][q], acc[q][c]);
              }
        };
        const int ng = k0 / (4 * U);
        load(0, xa0, yb0);
        for (int g = 0; g < ng; g += 2) {
          if (g + 1 < ng) load((g + 1) * 4 * U, xa1, yb1);
          mul(xa0, yb0);
          if (g + 1 < ng) {
            if (g + 2 < ng) load((g + 2) * 4 * U, xa0, yb0);
            mul(xa1, yb1);
          }
        }
      }
      // lane holds C(i = (rtb+4q)*16 + l16, j = c*16 + row(lane, r)), local indices
#pragma unroll
      for (int q = 0; q < BQ; q++)
        if (q < Q) {
          // the tile's A entries read first, at clamped positions inside the
          // block column (no condition, so all are in flight together), then
          // the guarded stores
          double av[CT][4];
#pragma unroll
          for (int c = 0; c < CT; c++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
              const int il = (rtb + 4 * q) * 16 + l16, jl = c * 16 + M::row(lane, r);
              av[c][r] = ld_l2(Ad + (long)(k0 + min(jl, w - 1)) * ld + k0 + min(il, R - 1));
            }
#pragma unroll
          for (int c = 0; c < CT; c++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
              const int il = (rtb + 4 * q) * 16 + l16, jl = c * 16 + M::row(lane, r);
              if (il < R && jl < w && il >= jl) P[jl * RS + il] = (LT)av[c][r] - acc[q][c][r];
            }
        }
    }
    __syncthreads();
    // 2. factor the tall block column in LDS (Schur complement form)
    for (int j = 0; j < w; j++) {
      const LT piv = P[j * RS + j];
      if (!(piv > (LT)0)) return k0 + j + 1;  // same LDS value in every thread
      const LT inv = (LT)1 / piv;
      for (int i = j + 1 + tid; i < R; i += BT) {
        const LT f = P[j * RS + i] * inv;
        const int je = min(w, i + 1);
        for (int jj = j + 1; jj < je; jj++) P[jj * RS + i] -= f * P[j * RS + jj];
      }
      __syncthreads();
    }
    if (tid < w) rsq[tid] = (LT)1 / sqrt(P[tid * RS + tid]);
    __syncthreads();
    // 3. L[k0+i][k0+jj] = S(i, jj) / sqrt(S(jj, jj)), lower part only
    for (int e = tid; e < w * R; e += BT) {
      const int jj = e / R, i = e - jj * R;
      if (i >= jj) L[(long)(k0 + jj) * ld + k0 + i] = P[jj * RS + i] * rsq[jj];
    }
    wait_stores_and_sync();
  }
  return 0;
}

// y <- (L L^T)^{-1} y for nrhs right-hand sides held in LDS (y[r*RS + i]);
// arithmetic in FP64 whatever L's storage type.
template <typename LT>
__device__ __forceinline__ void solve_system(const LT *L, int n, int nrhs, double *y, int RS) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long ld = n;
  const int nwv = (nrhs + 3) >> 2;  // waves busy with the 16-row diagonal solves
  // forward: L y = b
  for (int j0 = 0; j0 < n; j0 += 16) {
    const int jb = min(16, n - j0);
    if (wave < nwv) {
      const int i = lane & 15, r = wave * 4 + (lane >> 4);
      const bool act = r < nrhs && i < jb;
      double v = act ? y[r * RS + j0 + i] : 0.0;
      double lrow[16];  // row i of the diagonal block, loaded before the dependent chain
      const int ic = min(i, jb - 1);
#pragma unroll
      for (int jj = 0; jj < 16; jj++) lrow[jj] = (double)L[(long)(j0 + min(jj, ic)) * ld + j0 + ic];
#pragma unroll
      for (int jj = 0; jj < 16; jj++) lrow[jj] = (act && jj <= i) ? lrow[jj] : 1.0;
#pragma unroll
      for (int jj = 0; jj < 16; jj++) {
        if (jj < jb) {
          if (act && i == jj) v = v / lrow[jj];
          const double yj = __shfl(v, (lane & 48) | jj);
          if (act && i > jj) v -= lrow[jj] * yj;
        }
      }
      if (act) y[r * RS + j0 + i] = v;
    }
    __syncthreads();
    for (int i = j0 + jb + tid; i < n; i += BT) {
      double s[MAX_RHS];
#pragma unroll
      for (int r = 0; r < MAX_RHS; r++) s[r] = 0.0;
      // the block's 16 entries of row i loaded together (clamped), then
      // used in column order
      LT lv[16];
#pragma unroll
      for (int jj = 0; jj < 16; jj++) lv[jj] = L[(long)(j0 + min(jj, jb - 1)) * ld + i];
#pragma unroll
      for (int jj = 0; jj < 16; jj++) {
        if (jj >= jb) break;
        const double l = (double)lv[jj];
#pragma unroll
        for (int r = 0; r < MAX_RHS; r++)
          if (r < nrhs) s[r] += l * y[r * RS + j0 + jj];
      }
#pragma unroll
      for (int r = 0; r < MAX_RHS; r++)
        if (r < nrhs) y[r * RS + i] -= s[r];
    }
    __syncthreads();
  }
  // backward: L^T x = y
  for (int j0 = ((n - 1) >> 4) << 4; j0 >= 0; j0 -= 16) {
    const int jb = min(16, n - j0);
    if (wave < nwv) {
      const int i = lane & 15, r = wave * 4 + (lane >> 4);
      const bool act = r < nrhs && i < jb;
      double v = act ? y[r * RS + j0 + i] : 0.0;
      double lcol[16];  // column i of the diagonal block (row i of its transpose)
      const int ic = min(i, jb - 1);
#pragma unroll
      for (int jj = 0; jj < 16; jj++) lcol[jj] = (double)L[(long)(j0 + ic) * ld + j0 + max(ic, min(jj, jb - 1))];
#pragma unroll
      for (int jj = 0; jj < 16; jj++) lcol[jj] = (act && jj >= i && jj < jb) ? lcol[jj] : 1.0;
#pragma unroll
      for (int jj = 15; jj >= 0; jj--) {
        if (jj < jb) {
          if (act && i == jj) v = v / lcol[jj];
          const double xj = __shfl(v, (lane & 48) | jj);
          if (act && i < jj) v -= lcol[jj] * xj;
        }
      }
      if (act) y[r * RS + j0 + i] = v;
    }
    __syncthreads();
    for (int i = tid; i < j0; i += BT) {
      double s[MAX_RHS];
#pragma unroll
      for (int r = 0; r < MAX_RHS; r++) s[r] = 0.0;
      const LT *Lr = L + (long)i * ld + j0;  // L[j0 + jj][i], jj < jb
      LT lv[16];
#pragma unroll
      for (int jj = 0; jj < 16; jj++) lv[jj] = Lr[min(jj, jb - 1)];
#pragma unroll
      for (int jj = 0; jj < 16; jj++) {
        if (jj >= jb) break;
        const double l = (double)lv[jj];
#pragma unroll
        for (int r = 0; r < MAX_RHS; r++)
          if (r < nrhs) s[r] += l * y[r * RS + j0 + jj];
      }
#pragma unroll
      for (int r = 0; r < MAX_RHS; r++)
        if (r < nrhs) y[r * RS + i] -= s[r];
    }
    __syncthreads();
  }
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// rsd[r][i] = b[i][r] - (A x)[i] with A given by its lower triangle (FP64);
// on the first call also rowsum[i] = sum_j |A[i][j]|.  Returns the normwise
// backward error max_r ||r||_inf / (||A||_inf ||x_r||_inf + ||b_r||_inf) in
// every thread.
__device__ __forceinline__ double residual_system(const double *Ad, const double *Bd, int n, int nrhs, const double *x,
                                  double *rsd, double *rowsum, bool first, double *anorm, int RS,
                                  double *red) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long ld = n;
  for (int e = tid; e < n * nrhs; e += BT) {
    const int r = e / n, i = e - r * n;
    rsd[r * RS + i] = Bd[(long)r * ld + i];
    if (first && r == 0) rowsum[i] = 0.0;
  }
  __syncthreads();
  for (int j = wave; j < n; j += 4) {
    const double *Acol = Ad + (long)j * ld;
    double xj[MAX_RHS], dot[MAX_RHS];
#pragma unroll
    for (int r = 0; r < MAX_RHS; r++) {
      xj[r] = r < nrhs ? x[r * RS + j] : 0.0;
      dot[r] = 0.0;
    }
    double asum = 0.0;
    // four rows' loads in flight per lane (clamped to the column), each
    // row's products applied in row order as before
    for (int i0 = j + lane; i0 < n; i0 += 256) {
      double a4[4];
#pragma unroll
      for (int u = 0; u < 4; u++) a4[u] = Acol[min(i0 + 64 * u, n - 1)];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int i = i0 + 64 * u;
        if (i >= n) break;
        const double aij = a4[u];
        if (i > j) {
#pragma unroll
          for (int r = 0; r < MAX_RHS; r++)
            if (r < nrhs) atomicAdd(&rsd[r * RS + i], -aij * xj[r]);
          if (first) atomicAdd(&rowsum[i], fabs(aij));
        }
#pragma unroll
        for (int r = 0; r < MAX_RHS; r++)
          if (r < nrhs) dot[r] += aij * x[r * RS + i];
        asum += fabs(aij);
      }
    }
#pragma unroll
    for (int r = 0; r < MAX_RHS; r++)
      if (r < nrhs) dot[r] = wave_sum(dot[r]);
    asum = wave_sum(asum);
    if (lane == 0) {
#pragma unroll
      for (int r = 0; r < MAX_RHS; r++)
        if (r < nrhs) atomicAdd(&rsd[r * RS + j], -dot[r]);
      if (first) atomicAdd(&rowsum[j], asum);
    }
  }
  __syncthreads();
  // norms: per rhs max |r|, max |x|, max |b|; ||A||_inf on the first call
  double mr[MAX_RHS], mx[MAX_RHS], mb[MAX_RHS];
#pragma unroll
  for (int r = 0; r < MAX_RHS; r++) mr[r] = mx[r] = mb[r] = 0.0;
  double ma = 0.0;
  for (int i = tid; i < n; i += BT) {
#pragma unroll
    for (int r = 0; r < MAX_RHS; r++)
      if (r < nrhs) {
        mr[r] = fmax(mr[r], fabs(rsd[r * RS + i]));
        mx[r] = fmax(mx[r], fabs(x[r * RS + i]));
        mb[r] = fmax(mb[r], fabs(Bd[(long)r * ld + i]));
      }
    if (first) ma = fmax(ma, rowsum[i]);
  }
#pragma unroll
  for (int r = 0; r < MAX_RHS; r++) {
    mr[r] = wave_max(mr[r]);
    mx[r] = wave_max(mx[r]);
    mb[r] = wave_max(mb[r]);
  }
  ma = wave_max(ma);
  if (lane == 0) {
#pragma unroll
    for (int r = 0; r < MAX_RHS; r++) {
      red[wave * (3 * MAX_RHS + 1) + 3 * r] = mr[r];
      red[wave * (3 * MAX_RHS + 1) + 3 * r + 1] = mx[r];
      red[wave * (3 * MAX_RHS + 1) + 3 * r + 2] = mb[r];
    }
    red[wave * (3 * MAX_RHS + 1) + 3 * MAX_RHS] = ma;
  }
  __syncthreads();
  if (first) {
    double m = 0.0;
    for (int wv = 0; wv < 4; wv++) m = fmax(m, red[wv * (3 * MAX_RHS + 1) + 3 * MAX_RHS]);
    *anorm = m;
  }
  double eta = 0.0;
  for (int r = 0; r < nrhs; r++) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int wv = 0; wv < 4; wv++) {
      a0 = fmax(a0, red[wv * (3 * MAX_RHS + 1) + 3 * r]);
      a1 = fmax(a1, red[wv * (3 * MAX_RHS + 1) + 3 * r + 1]);
      a2 = fmax(a2, red[wv * (3 * MAX_RHS + 1) + 3 * r + 2]);
    }
    const double den = *anorm * a1 + a2;
    const double e = den > 0.0 ? a0 / den : (a0 > 0.0 ? INFINITY : 0.0);
    eta = fmax(eta, e);
  }
  __syncthreads();  // red is reused by the next call
  return eta;
}

template <typename LT, int NB, int BQ, bool IR>
__global__ __launch_bounds__(BT) void batched_chol_kernel(batched_args a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_idx;
  __shared__ LT s_rsq[NB];
  __shared__ double s_red[4 * (3 * MAX_RHS + 1)];
  __shared__ double s_anorm;
  const int tid = threadIdx.x;
  const int RS = a.rs;
  for (;;) {
    if (tid == 0) s_idx = (int)atomicAdd(a.counter, 1u);
    __syncthreads();
    const int idx = s_idx;
    __syncthreads();
    if (idx >= a.count) break;
    const bsys d = a.sys[idx];
    if (d.n <= 0) {
      if (tid == 0) {
        a.info[d.id] = 0;
        if (a.iters) a.iters[d.id] = 0;
        if (a.berr) a.berr[d.id] = 0.0;
      }
      continue;
    }
    LT *L = IR ? (LT *)(a.L32 + d.a_off) : (LT *)(a.A + d.a_off);
    const int fail = factor_system<LT, NB, BQ>(d, a, L, (LT *)smem, s_rsq);
    if (tid == 0) a.info[d.id] = fail;
    if (fail) {
      __syncthreads();
      continue;
    }
    double *Bd = a.B + d.b_off;
    const int n = d.n, nrhs = a.nrhs;
    double *x = (double *)smem;
    if (!IR) {
      for (int e = tid; e < n * nrhs; e += BT) {
        const int r = e / n, i = e - r * n;
        x[r * RS + i] = Bd[(long)r * n + i];
      }
      __syncthreads();
      solve_system<LT>(L, n, nrhs, x, RS);
      for (int e = tid; e < n * nrhs; e += BT) {
        const int r = e / n, i = e - r * n;
        Bd[(long)r * n + i] = x[r * RS + i];
      }
      __syncthreads();
    } else {
      double *rsd = x + (size_t)MAX_RHS * RS;
      double *rowsum = rsd + (size_t)MAX_RHS * RS;
      const double *Ad = a.A + d.a_off;
      for (int e = tid; e < n * nrhs; e += BT) {
        const int r = e / n, i = e - r * n;
        x[r * RS + i] = Bd[(long)r * n + i];
      }
      __syncthreads();
      solve_system<LT>(L, n, nrhs, x, RS);
      int it = 0;
      double eta = 0.0;
      for (;;) {
        eta = residual_system(Ad, Bd, n, nrhs, x, rsd, rowsum, it == 0, &s_anorm, RS, s_red);
        if (eta <= a.tol || it >= a.max_iter) break;
        solve_system<LT>(L, n, nrhs, rsd, RS);
        for (int e = tid; e < n * nrhs; e += BT) {
          const int r = e / n, i = e - r * n;
          x[r * RS + i] += rsd[r * RS + i];
        }
        __syncthreads();
        it++;
      }
      for (int e = tid; e < n * nrhs; e += BT) {
        const int r = e / n, i = e - r * n;
        Bd[(long)r * n + i] = x[r * RS + i];
      }
      if (tid == 0) {
        if (a.iters) a.iters[d.id] = it;
        if (a.berr) a.berr[d.id] = eta;
      }
      __syncthreads();
    }
  }
}

__global__ void batched_reset_kernel(unsigned *counter) {
  if (threadIdx.x == 0) counter[0] = 0u;
}

// ---- kernel selection ----
typedef void (*kfn)(batched_args);
struct kvariant {
  int mode, nb, bq;
  kfn fn;
  int lt_size;
};
const kvariant KVARIANTS[] = {
    {PLA_BATCHED_FP64, 16, 4, batched_chol_kernel<double, 16, 4, false>, 8},
    {PLA_BATCHED_FP64, 16, 8, batched_chol_kernel<double, 16, 8, false>, 8},
    {PLA_BATCHED_FP64, 32, 4, batched_chol_kernel<double, 32, 4, false>, 8},
    {PLA_BATCHED_FP32_IR, 16, 8, batched_chol_kernel<float, 16, 8, true>, 4},
    {PLA_BATCHED_FP32_IR, 32, 4, batched_chol_kernel<float, 32, 4, true>, 4},
    {PLA_BATCHED_FP32_IR, 32, 8, batched_chol_kernel<float, 32, 8, true>, 4},
    {PLA_BATCHED_FP32_IR, 64, 4, batched_chol_kernel<float, 64, 4, true>, 4},
};
constexpr size_t LDS_MAX = 160 * 1024 - 4096;  // dynamic part; the rest is static

const kvariant *find_variant(int mode, int nb, int bq) {
  for (const auto &v : KVARIANTS)
    if (v.mode == mode && v.nb == nb && v.bq == bq) return &v;
  return nullptr;
}

size_t lds_bytes(const kvariant &v, int rs, int nrhs) {
  const size_t panel = (size_t)rs * v.nb * v.lt_size;
  const size_t vec = (size_t)(v.mode == PLA_BATCHED_FP32_IR ? 2 * MAX_RHS + 1 : nrhs) * rs * 8;
  return std::max(panel, vec);
}

struct bconfig {
  int nb, bq, wpc;  // block size, row tiles per wave, workgroups per CU
};
// tuned configuration by problem signature, for the process lifetime
std::map<std::tuple<int, int, int, long, long, int>, bconfig> g_tuned;

}  // namespace

struct PLA_Batched_plan_struct {
  int count = 0, nrhs = 0, mode = 0, max_iter = 0;
  double tol = 0.0;
  double *A = nullptr, *B = nullptr;
  int *info = nullptr, *iters = nullptr;
  double *berr = nullptr;
  int nmax = 0, rs = 0;
  long sum_n = 0, sum_n2 = 0;
  size_t a_elems = 0, b_elems = 0;
  bsys *d_sys = nullptr;
  unsigned *d_counter = nullptr;
  float *d_L32 = nullptr;
  int nb = 0, bq = 0, wpc = 0;  // chosen block size, row tiles per wave, workgroups per CU
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

int launch_into(PLA_Batched_plan p, const kvariant &v, int wpc, double *A, double *B, int *info, int *iters,
                double *berr, hipStream_t s) {
  batched_args a;
  a.sys = p->d_sys;
  a.count = p->count;
  a.A = A;
  a.B = B;
  a.L32 = p->d_L32;
  a.nrhs = p->nrhs;
  a.max_iter = p->max_iter;
  a.tol = p->tol;
  a.info = info;
  a.iters = iters;
  a.berr = berr;
  a.counter = p->d_counter;
  a.rs = p->rs;
  const size_t lds = lds_bytes(v, p->rs, p->nrhs);  // allowed by max_wpc() beforehand
  int grid = ctx().cu_count * wpc;
  if (grid > p->count) grid = p->count > 0 ? p->count : 1;
  hipLaunchKernelGGL(batched_reset_kernel, dim3(1), dim3(64), 0, s, p->d_counter);
  hipLaunchKernelGGL(v.fn, dim3(grid), dim3(BT), lds, s, a);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int max_wpc(const kvariant &v, int rs, int nrhs) {
  const size_t lds = lds_bytes(v, rs, nrhs);
  if (lds > LDS_MAX) return 0;
  int blocks = 0;
  if (hipFuncSetAttribute((const void *)v.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
      hipSuccess)
    return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, v.fn, BT, lds) != hipSuccess) return 0;
  return blocks;
}

int build_graph(PLA_Batched_plan p) {
  if (p->exec) {
    PLA_HIP(hipGraphExecDestroy(p->exec));
    p->exec = nullptr;
  }
  if (p->graph) {
    PLA_HIP(hipGraphDestroy(p->graph));
    p->graph = nullptr;
  }
  const kvariant *v = find_variant(p->mode, p->nb, p->bq);
  if (!v) return fail(PLA_ERR_INVALID_PARAM, "batched: no kernel for nb = %d", p->nb);
  if (max_wpc(*v, p->rs, p->nrhs) < p->wpc)
    return fail(PLA_ERR_INVALID_PARAM, "batched: %d workgroups per CU do not fit", p->wpc);
  hipStream_t s = ctx().stream;
  PLA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
  int rc = launch_into(p, *v, p->wpc, p->A, p->B, p->info, p->iters, p->berr, s);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  p->graph = g;
  PLA_HIP(hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
  return PLA_SUCCESS;
}

// Candidate configurations that fit this plan's sizes.
std::vector<bconfig> candidates(PLA_Batched_plan p) {
  std::vector<bconfig> c;
  for (const auto &v : KVARIANTS) {
    if (v.mode != p->mode) continue;
    const int m = max_wpc(v, p->rs, p->nrhs);
    for (int w = 1; w <= std::min(m, 2); w++) c.push_back({v.nb, v.bq, w});
  }
  return c;
}

}  // namespace

extern "C" {

int PLA_Batched_plan_create(int count, const int *n, const long *a_off, const long *b_off, int nrhs, int mode,
                            double tol, int max_iter, double *A, double *B, int *info, int *iters, double *berr,
                            PLA_Batched_plan *plan) {
  int rc = ensure_device();
  if (rc) return rc;
  if (!plan) return fail(PLA_ERR_INVALID_PARAM, "PLA_Batched_plan_create: plan is NULL");
  *plan = nullptr;
  const int base_mode = mode & 0xff;
  if (base_mode != PLA_BATCHED_FP64 && base_mode != PLA_BATCHED_FP32_IR)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Batched_plan_create: unknown mode %d", mode);
  if (count < 0 || (count > 0 && (!n || !A || !B || !info)))
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Batched_plan_create: count %d with NULL arrays", count);
  if (nrhs < 1 || nrhs > MAX_RHS)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Batched_plan_create: nrhs = %d outside [1, %d]", nrhs, MAX_RHS);
  if (max_iter < 0) return fail(PLA_ERR_INVALID_PARAM, "PLA_Batched_plan_create: max_iter < 0");
  auto p = new PLA_Batched_plan_struct();
  p->count = count;
  p->nrhs = nrhs;
  p->mode = base_mode;
  p->tol = tol;
  p->max_iter = max_iter;
  p->A = A;
  p->B = B;
  p->info = info;
  p->iters = iters;
  p->berr = berr;
  std::vector<bsys> h(count);
  long ao = 0, bo = 0;
  for (int i = 0; i < count; i++) {
    if (n[i] < 0) {
      delete p;
      return fail(PLA_ERR_DIMENSIONS, "PLA_Batched_plan_create: n[%d] = %d", i, n[i]);
    }
    h[i].n = n[i];
    h[i].id = i;
    h[i].a_off = a_off ? a_off[i] : ao;
    h[i].b_off = b_off ? b_off[i] : bo;
    ao += (long)n[i] * n[i];
    bo += (long)n[i] * nrhs;
    p->nmax = std::max(p->nmax, n[i]);
    p->sum_n += n[i];
    p->sum_n2 += (long)n[i] * n[i];
    p->a_elems = std::max(p->a_elems, (size_t)(h[i].a_off + (long)n[i] * n[i]));
    p->b_elems = std::max(p->b_elems, (size_t)(h[i].b_off + (long)n[i] * nrhs));
  }
  // largest first (longest-processing-time order)
  std::stable_sort(h.begin(), h.end(), [](const bsys &x, const bsys &y) { return x.n > y.n; });
  p->rs = ((std::max(p->nmax, 1) + 15) / 16) * 16;
  hipError_t e = hipMalloc(&p->d_sys, sizeof(bsys) * std::max(count, 1));
  if (e == hipSuccess) e = hipMalloc(&p->d_counter, 256);
  if (e == hipSuccess && base_mode == PLA_BATCHED_FP32_IR)
    e = hipMalloc(&p->d_L32, sizeof(float) * std::max(p->a_elems, (size_t)1));
  if (e == hipSuccess && count > 0)
    e = hipMemcpy(p->d_sys, h.data(), sizeof(bsys) * count, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    PLA_Batched_plan_free(&p);
    return hip_fail(e, "PLA_Batched_plan_create: device allocation");
  }
  auto cand = candidates(p);
  if (cand.empty()) {
    const int nm = p->nmax;
    PLA_Batched_plan_free(&p);
    return fail(PLA_ERR_DIMENSIONS, "PLA_Batched_plan_create: n = %d does not fit the batched kernel's LDS",
                nm);
  }
  const auto key = std::make_tuple(p->mode, count, nrhs, p->sum_n, p->sum_n2, p->nmax);
  auto it = g_tuned.find(key);
  bconfig cfg = cand[0];
  if (it != g_tuned.end()) {
    cfg = it->second;
  } else {
    // untuned default: the largest block size that fits at one workgroup per
    // CU, with the most row tiles per wave
    for (auto &c : cand)
      if (c.wpc == 1 && (c.nb > cfg.nb || (c.nb == cfg.nb && c.bq > cfg.bq))) cfg = c;
    cfg.wpc = 1;
  }
  p->nb = cfg.nb;
  p->bq = cfg.bq;
  p->wpc = cfg.wpc;
  if ((rc = build_graph(p))) {
    PLA_Batched_plan_free(&p);
    return rc;
  }
  *plan = p;
  if (mode & PLA_BATCHED_AUTOTUNE) {
    if ((rc = PLA_Batched_plan_autotune(p, nullptr))) {
      PLA_Batched_plan_free(plan);
      return rc;
    }
  }
  return PLA_SUCCESS;
}

int PLA_Batched_plan_execute(PLA_Batched_plan p) {
  if (!p || !p->exec) return fail(PLA_ERR_INVALID_PARAM, "PLA_Batched_plan_execute: invalid plan");
  PLA_HIP(hipGraphLaunch(p->exec, ctx().stream));
  return PLA_SUCCESS;
}

// Times every candidate (block size, workgroups per CU) on scratch copies of
// the plan's A and B (the plan's own buffers are left as they are), keeps the
// fastest and records it for later plans of the same signature.
int PLA_Batched_plan_autotune(PLA_Batched_plan p, double *ms_best) {
  if (!p) return fail(PLA_ERR_INVALID_PARAM, "PLA_Batched_plan_autotune: invalid plan");
  auto cand = candidates(p);
  hipStream_t s = ctx().stream;
  double *As = nullptr, *Bs = nullptr, *berr_s = nullptr;
  int *info_s = nullptr, *iters_s = nullptr;
  const size_t asz = sizeof(double) * std::max(p->a_elems, (size_t)1);
  const size_t bsz = sizeof(double) * std::max(p->b_elems, (size_t)1);
  hipError_t e = hipMalloc(&Bs, bsz);
  if (e == hipSuccess) e = hipMalloc(&info_s, sizeof(int) * std::max(p->count, 1));
  if (e == hipSuccess) e = hipMalloc(&iters_s, sizeof(int) * std::max(p->count, 1));
  if (e == hipSuccess) e = hipMalloc(&berr_s, sizeof(double) * std::max(p->count, 1));
  // FP64 mode factors A in place, so it needs its own copy; FP32_IR only reads A
  if (e == hipSuccess && p->mode == PLA_BATCHED_FP64) e = hipMalloc(&As, asz);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  int rc = PLA_SUCCESS;
  double best = 1e300;
  bconfig best_c{p->nb, p->bq, p->wpc};
  if (e != hipSuccess) rc = hip_fail(e, "PLA_Batched_plan_autotune: scratch allocation");
  for (size_t ci = 0; ci < cand.size() && rc == PLA_SUCCESS; ci++) {
    const kvariant *v = find_variant(p->mode, cand[ci].nb, cand[ci].bq);
    (void)max_wpc(*v, p->rs, p->nrhs);  // sets the kernel's dynamic LDS limit
    double t_min = 1e300;
    for (int rep = 0; rep < 2 && rc == PLA_SUCCESS; rep++) {
      if (As) e = hipMemcpyAsync(As, p->A, asz, hipMemcpyDeviceToDevice, s);
      if (e == hipSuccess) e = hipMemcpyAsync(Bs, p->B, bsz, hipMemcpyDeviceToDevice, s);
      if (e == hipSuccess) e = hipEventRecord(e0, s);
      if (e != hipSuccess) {
        rc = hip_fail(e, "PLA_Batched_plan_autotune");
        break;
      }
      rc = launch_into(p, *v, cand[ci].wpc, As ? As : p->A, Bs, info_s, iters_s, berr_s, s);
      if (rc) break;
      e = hipEventRecord(e1, s);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) {
        rc = hip_fail(e, "PLA_Batched_plan_autotune");
        break;
      }
      t_min = std::min(t_min, (double)ms);
    }
    if (rc == PLA_SUCCESS && t_min < best) {
      best = t_min;
      best_c = cand[ci];
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(As);
  (void)hipFree(Bs);
  (void)hipFree(info_s);
  (void)hipFree(iters_s);
  (void)hipFree(berr_s);
  if (rc) return rc;
  p->nb = best_c.nb;
  p->bq = best_c.bq;
  p->wpc = best_c.wpc;
  g_tuned[std::make_tuple(p->mode, p->count, p->nrhs, p->sum_n, p->sum_n2, p->nmax)] = best_c;
  if (ms_best) *ms_best = best;
  return build_graph(p);
}

int PLA_Batched_plan_config(PLA_Batched_plan p, int *nb, int *workgroups_per_cu) {
  if (!p) return fail(PLA_ERR_INVALID_PARAM, "PLA_Batched_plan_config: invalid plan");
  if (nb) *nb = p->nb;
  if (workgroups_per_cu) *workgroups_per_cu = p->wpc;
  return PLA_SUCCESS;
}

int PLA_Batched_plan_free(PLA_Batched_plan *plan) {
  if (!plan || !*plan) return PLA_SUCCESS;
  PLA_Batched_plan p = *plan;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  (void)hipFree(p->d_sys);
  (void)hipFree(p->d_counter);
  (void)hipFree(p->d_L32);
  delete p;
  *plan = nullptr;
  return PLA_SUCCESS;
}

int PLA_Chol_batched(int count, const int *n, const long *a_off, const long *b_off, int nrhs, int mode,
                     double tol, int max_iter, double *A, double *B, int *info, int *iters, double *berr) {
  PLA_Batched_plan p = nullptr;
  int rc = PLA_Batched_plan_create(count, n, a_off, b_off, nrhs, mode, tol, max_iter, A, B, info, iters, berr,
                                   &p);
  if (rc) return rc;
  rc = PLA_Batched_plan_execute(p);
  if (rc == PLA_SUCCESS) {
    const hipError_t e = hipStreamSynchronize(ctx().stream);
    if (e != hipSuccess) rc = hip_fail(e, "PLA_Chol_batched");
  }
  PLA_Batched_plan_free(&p);
  return rc;
}

}  // extern "C"
