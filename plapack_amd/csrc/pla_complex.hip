// pla_complex.hip -- MPI_DOUBLE_COMPLEX kernels of the LU path.
//
// Reference counterparts: the complex branches of PLA_Iamax
// (BLAS1/PLA_Iamax.c:176-230, izamax: first maximum of |re| + |im|),
// PLA_Local_inv_scal (Local_BLAS/PLA_Local_BLAS1.c:240-250), PLA_Local_ger
// (zgeru, INCLUDE/PLA_BLAS.h:407), PLA_Apply_pivots_to_rows and ztrsm.
//
// The O(n^3) parts of a complex factorization run on the FP64 MFMA GEMM: a
// complex m x k operand A becomes the real 2m x 2k matrix
//   Ahat(2i, 2l) = re a_il,  Ahat(2i, 2l+1) = -im a_il,
//   Ahat(2i+1, 2l) = im a_il, Ahat(2i+1, 2l+1) = re a_il,
// and a complex k x n operand or result with interleaved (re, im) storage is
// already the real 2k x n matrix of its doubles with leading dimension 2 ld;
// then Ahat B is the complex product, 8mnk real flops -- the complex flop
// count -- on the MFMA kernel.  These kernels are the rest: the embedding,
// a one-workgroup panel leaf, the sequential interchanges, the
// triangular-solve leaves, the conjugate transpose, and the Cholesky
// diagonal-block kernels.
#include "pla_internal.hpp"

namespace {

struct zd {
  double re, im;
};

__device__ __forceinline__ zd zmul(zd a, zd b) { return zd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ zd zsub(zd a, zd b) { return zd{a.re - b.re, a.im - b.im}; }
// 1 / v as conj(v) / |v|^2
__device__ __forceinline__ zd zrecip(zd v) {
  const double d = v.re * v.re + v.im * v.im;
  return zd{v.re / d, -v.im / d};
}
// a / b as conj(b) a / |b|^2
__device__ __forceinline__ zd zdiv(zd a, zd b) {
  const double d = b.re * b.re + b.im * b.im;
  return zd{(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
__device__ __forceinline__ double zabs1(zd v) { return fabs(v.re) + fabs(v.im); }

// Ahat (2m x 2n real, ldh) from A (m x n complex, lda): one thread per element.
__global__ __launch_bounds__(256) void zembed_kernel(int m, int n, const zd *__restrict__ A, long lda,
                                                     double *__restrict__ H, long ldh) {
  const long total = (long)m * n;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int j = (int)(e / m), i = (int)(e - (long)j * m);
    const zd a = A[(long)j * lda + i];
    double *h0 = H + (long)(2 * j) * ldh + 2 * i, *h1 = h0 + ldh;
    h0[0] = a.re;
    h0[1] = a.im;
    h1[0] = -a.im;
    h1[1] = a.re;
  }
}

// Sequential interchanges (dlaswp order) of rows i and i + piv[i], i < np,
// on an nrows x ncols complex block: one thread per column.
__global__ __launch_bounds__(256) void zlaswp_kernel(zd *__restrict__ A, long lda, int nrows, int ncols, int np,
                                                     const int *__restrict__ piv) {
  __shared__ int sp[1024];
  const int c = blockIdx.x * 256 + threadIdx.x;
  zd *col = A + (long)c * lda;
  for (int base = 0; base < np; base += 1024) {
    const int cnt = min(1024, np - base);
    __syncthreads();
    for (int t = threadIdx.x; t < cnt; t += 256) sp[t] = piv[base + t];
    __syncthreads();
    if (c < ncols) {
      for (int t = 0; t < cnt; t++) {
        const int p = sp[t], r = base + t;
        if (p > 0 && r + p < nrows) {
          const zd a = col[r], b = col[r + p];
          col[r] = b;
          col[r + p] = a;
        }
      }
    }
  }
}

// Unblocked panel LU of the w leftmost columns of an mp x pw complex panel
// (w <= ZLEAF_W, one workgroup): per column j, izamax over rows j..mp-1 (the
// first maximum of |re| + |im|), the interchange across all pw panel columns,
// a_21 scaled by the pivot's reciprocal, the rank-1 update (zgeru) of the
// leaf's columns j+1..w-1.  piv[j] is the pivot row's offset from row j.
constexpr int ZLEAF_W = 4;
constexpr int ZLEAF_THREADS = 1024;

__global__ __launch_bounds__(ZLEAF_THREADS) void zleaf_kernel(zd *__restrict__ P, long lda, int mp, int w, int pw,
                                                              int pc0, int *__restrict__ piv) {
  // P: the leaf's top-left element; the panel's column 0 is P - pc0 * lda
  __shared__ double sv[ZLEAF_THREADS];
  __shared__ int si[ZLEAF_THREADS];
  __shared__ zd srow[ZLEAF_W];
  const int t = threadIdx.x;
  zd *panel0 = P - (long)pc0 * lda;
  for (int j = 0; j < w && j < mp; j++) {
    zd *cj = P + (long)j * lda;
    // izamax over rows j..mp-1: each thread its strided rows, then a tree
    double best = -1.0;
    int bi = mp;
    for (int i = j + t; i < mp; i += ZLEAF_THREADS) {
      const double v = zabs1(cj[i]);
      if (v > best) { best = v; bi = i; }
    }
    sv[t] = best;
    si[t] = bi;
    __syncthreads();
    for (int h = ZLEAF_THREADS / 2; h > 0; h >>= 1) {
      if (t < h) {
        const double o = sv[t + h];
        const int oi = si[t + h];
        if (o > sv[t] || (o == sv[t] && oi < si[t])) { sv[t] = o; si[t] = oi; }
      }
      __syncthreads();
    }
    const int p = si[0] < mp ? si[0] : j;  // an all-NaN column keeps row j
    if (t == 0) piv[j] = p - j;
    // the interchange across the whole panel (columns -pc0 .. pw-pc0-1 of P)
    if (p != j)
      for (int c = t; c < pw; c += ZLEAF_THREADS) {
        zd *col = panel0 + (long)c * lda;
        const zd a = col[j], b = col[p];
        col[j] = b;
        col[p] = a;
      }
    __threadfence_block();
    __syncthreads();
    // row j of the leaf (after the interchange), then scale and update below it
    if (t < w) srow[t] = P[(long)t * lda + j];
    __syncthreads();
    const zd inv = zrecip(srow[j]);
    for (int i = j + 1 + t; i < mp; i += ZLEAF_THREADS) {
      const zd l = zmul(cj[i], inv);
      cj[i] = l;
      for (int c = j + 1; c < w; c++) {
        zd *x = P + (long)c * lda + i;
        *x = zsub(*x, zmul(l, srow[c]));
      }
    }
    __threadfence_block();
    __syncthreads();
  }
}

// Left triangular-solve leaf for complex data, B <- M^-1 B with M lower or
// upper, unit or not (dividing by the diagonal), b <= ZTRSM_LEAF: a wave per
// column group, lane i owning row i (the real leaf's scheme; complex rows take
// two registers, so the leaf is 32 rows and half the lanes idle).
constexpr int ZTRSM_LEAF = 32;
constexpr int ZTRSM_NC = 4;

__device__ __forceinline__ double lane_bcast_d(double v, int l) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)u, l);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <bool LOWER, bool UNIT>
__global__ __launch_bounds__(256) void ztrsm_leaf_kernel(int b, int ncols, const zd *__restrict__ T, long ldt,
                                                         zd *__restrict__ B, long ldb) {
  __shared__ zd sT[ZTRSM_LEAF][ZTRSM_LEAF + 1];
  for (int e = threadIdx.x; e < b * b; e += 256) {
    const int kcol = e / b, i = e - kcol * b;
    sT[i][kcol] = T[(long)kcol * ldt + i];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  zd mrow[ZTRSM_LEAF];
#pragma unroll
  for (int k = 0; k < ZTRSM_LEAF; k++) {
    const bool in = lane < b && k < b && (LOWER ? k < lane : k > lane);
    mrow[k] = in ? sT[lane][k] : zd{0.0, 0.0};
  }
  const zd diag = (!UNIT && lane < b) ? sT[lane][lane] : zd{1.0, 0.0};
  const int nwaves = gridDim.x * 4;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  for (int c0 = wave * ZTRSM_NC; c0 < ncols; c0 += nwaves * ZTRSM_NC) {
    zd x[ZTRSM_NC];
#pragma unroll
    for (int q = 0; q < ZTRSM_NC; q++)
      x[q] = (c0 + q < ncols && lane < b) ? B[(long)(c0 + q) * ldb + lane] : zd{0.0, 0.0};
#pragma unroll
    for (int s = 0; s < ZTRSM_LEAF; s++) {
      const int k = LOWER ? s : ZTRSM_LEAF - 1 - s;
      if (k < b) {
        if (!UNIT && lane == k) {
#pragma unroll
          for (int q = 0; q < ZTRSM_NC; q++) x[q] = zdiv(x[q], diag);
        }
        zd xk[ZTRSM_NC];
#pragma unroll
        for (int q = 0; q < ZTRSM_NC; q++) xk[q] = zd{lane_bcast_d(x[q].re, k), lane_bcast_d(x[q].im, k)};
        if (LOWER ? lane > k : lane < k) {
#pragma unroll
          for (int q = 0; q < ZTRSM_NC; q++) x[q] = zsub(x[q], zmul(xk[q], mrow[k]));
        }
      }
    }
#pragma unroll
    for (int q = 0; q < ZTRSM_NC; q++)
      if (c0 + q < ncols && lane < b) B[(long)(c0 + q) * ldb + lane] = x[q];
  }
}

// B (n x m) = A^H for the complex m x n block A: 32 x 32 tiles through LDS.
__global__ __launch_bounds__(256) void zconjtrans_kernel(int m, int n, const zd *__restrict__ A, long lda,
                                                         zd *__restrict__ B, long ldb) {
  __shared__ zd tile[32][33];
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  for (int r = threadIdx.y; r < 32; r += 8) {
    const int i = i0 + threadIdx.x, j = j0 + r;  // A(i, j), consecutive i per row of threads
    if (i < m && j < n) tile[r][threadIdx.x] = A[(long)j * lda + i];
  }
  __syncthreads();
  for (int r = threadIdx.y; r < 32; r += 8) {
    const int j = j0 + threadIdx.x, i = i0 + r;  // B(j, i) = conj A(i, j)
    if (i < m && j < n) {
      const zd v = tile[threadIdx.x][r];
      B[(long)i * ldb + j] = zd{v.re, -v.im};
    }
  }
}

// Lower Cholesky of a b x b Hermitian diagonal block (b <= ZPOTRF_LEAF) in
// one workgroup, PLA_zchol's right-looking order
// (SOLVERS/Cholesky/PLA_Local_chol.c:210-232): a(j,j) <- sqrt(re a(j,j)) with
// a zero imaginary part, the column below scaled by its reciprocal, then the
// Hermitian rank-1 update of the lower trailing block (zher: the updated
// diagonal keeps a zero imaginary part).  A non-positive (or NaN) real
// diagonal stops the block and stores offset + j + 1 in *info, as the real
// path's dchol test does (the reference's zchol has no test).
constexpr int ZPOTRF_LEAF = 32;

__global__ __launch_bounds__(256) void zpotrf_leaf_kernel(int b, zd *__restrict__ A, long lda, int offset,
                                                          int *__restrict__ info) {
  if (*info) return;
  __shared__ zd S[ZPOTRF_LEAF][ZPOTRF_LEAF + 1];
  __shared__ int failed;
  const int t = threadIdx.x;
  for (int e = t; e < b * b; e += 256) {
    const int j = e / b, i = e - j * b;
    if (i >= j) S[i][j] = A[(long)j * lda + i];
  }
  if (t == 0) failed = 0;
  __syncthreads();
  for (int j = 0; j < b; j++) {
    const double d = S[j][j].re;
    if (!(d > 0.0)) {
      if (t == 0) failed = j + 1;
      break;
    }
    const double r = sqrt(d), inv = 1.0 / r;
    __syncthreads();
    if (t == 0) S[j][j] = zd{r, 0.0};
    for (int i = j + 1 + t; i < b; i += 256) S[i][j] = zd{S[i][j].re * inv, S[i][j].im * inv};
    __syncthreads();
    // S(i, c) -= S(i, j) conj(S(c, j)) for j < c <= i
    for (int e = t; e < b * b; e += 256) {
      const int c = e / b, i = e - c * b;
      if (c > j && i >= c) {
        const zd x = S[i][j], y = S[c][j];
        const zd u = zd{x.re * y.re + x.im * y.im, x.im * y.re - x.re * y.im};  // x conj(y)
        S[i][c] = zd{S[i][c].re - u.re, i == c ? 0.0 : S[i][c].im - u.im};
      }
    }
    __syncthreads();
  }
  __syncthreads();
  if (failed) {
    if (t == 0) *info = offset + failed;
    return;  // the caller restores the block from its backup
  }
  for (int e = t; e < b * b; e += 256) {
    const int j = e / b, i = e - j * b;
    if (i >= j) A[(long)j * lda + i] = S[i][j];
  }
}

// X <- X inv(L)^H, X nrows x b (b <= ZTRSM_LEAF), L lower non-unit: one thread
// per row (the rows are independent), ztrsm("R","L","C","N")'s order: column k
// divided by conj(L(k,k)), then conj(L(j,k)) x_k subtracted from columns j > k.
__global__ __launch_bounds__(256) void ztrsm_rlc_leaf_kernel(int nrows, int b, const zd *__restrict__ L, long ldl,
                                                             zd *__restrict__ X, long ldx,
                                                             const int *__restrict__ info) {
  if (info && *info) return;
  __shared__ zd sL[ZTRSM_LEAF][ZTRSM_LEAF + 1];
  for (int e = threadIdx.x; e < b * b; e += 256) {
    const int c = e / b, i = e - c * b;
    sL[i][c] = L[(long)c * ldl + i];
  }
  __syncthreads();
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= nrows) return;
  zd v[ZTRSM_LEAF];
#pragma unroll
  for (int c = 0; c < ZTRSM_LEAF; c++) v[c] = c < b ? X[(long)c * ldx + r] : zd{0.0, 0.0};
#pragma unroll
  for (int k = 0; k < ZTRSM_LEAF; k++) {
    if (k < b) {
      const zd d = sL[k][k];
      v[k] = zdiv(v[k], zd{d.re, -d.im});
#pragma unroll
      for (int j = k + 1; j < ZTRSM_LEAF; j++)
        if (j < b) {
          const zd l = sL[j][k];
          v[j] = zsub(v[j], zmul(zd{l.re, -l.im}, v[k]));
        }
    }
  }
#pragma unroll
  for (int c = 0; c < ZTRSM_LEAF; c++)
    if (c < b) X[(long)c * ldx + r] = v[c];
}

// Lower triangle of a diagonal block: C -= X X^H, C n x n (n <= ZHERK_DIAG),
// X n x k; one thread per element (i >= j), the diagonal's imaginary part set
// to zero (zher / zherk semantics).
constexpr int ZHERK_DIAG = 256;

__global__ __launch_bounds__(256) void zherk_diag_kernel(int n, int k, const zd *__restrict__ X, long ldx,
                                                         zd *__restrict__ C, long ldc, const int *__restrict__ info) {
  if (info && *info) return;
  __shared__ zd xi[16][17], xj[16][17];
  const int i = blockIdx.x * 16 + threadIdx.x, j = blockIdx.y * 16 + threadIdx.y;
  if (blockIdx.x < blockIdx.y) return;  // tile entirely above the diagonal
  double sr = 0.0, si = 0.0;
  for (int l0 = 0; l0 < k; l0 += 16) {
    // xi[a][l] = X(16 bx + a, l0 + l), xj[a][l] = X(16 by + a, l0 + l)
    const int l = l0 + threadIdx.y;
    const int ri = blockIdx.x * 16 + threadIdx.x, rj = blockIdx.y * 16 + threadIdx.x;
    xi[threadIdx.x][threadIdx.y] = (ri < n && l < k) ? X[(long)l * ldx + ri] : zd{0.0, 0.0};
    xj[threadIdx.x][threadIdx.y] = (rj < n && l < k) ? X[(long)l * ldx + rj] : zd{0.0, 0.0};
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const zd a = xi[threadIdx.x][q], c = xj[threadIdx.y][q];
      sr += a.re * c.re + a.im * c.im;  // a conj(c)
      si += a.im * c.re - a.re * c.im;
    }
    __syncthreads();
  }
  if (i < n && j < n && i >= j) {
    zd *p = C + (long)j * ldc + i;
    *p = zd{p->re - sr, i == j ? 0.0 : p->im - si};
  }
}

inline unsigned zgrid(long total) {
  long g = (total + 255) / 256;
  if (g > 65535L * 8) g = 65535L * 8;
  return (unsigned)(g > 0 ? g : 1);
}

}  // namespace

namespace pla {

int k_zembed(hipStream_t s, int m, int n, const void *A, long lda, double *H, long ldh) {
  if (m <= 0 || n <= 0) return PLA_SUCCESS;
  hipLaunchKernelGGL(zembed_kernel, dim3(zgrid((long)m * n)), dim3(256), 0, s, m, n, (const zd *)A, lda, H, ldh);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_zlaswp(hipStream_t s, void *A, long lda, int nrows, int ncols, int np, const int *piv) {
  if (ncols <= 0 || np <= 0) return PLA_SUCCESS;
  hipLaunchKernelGGL(zlaswp_kernel, dim3((ncols + 255) / 256), dim3(256), 0, s, (zd *)A, lda, nrows, ncols, np, piv);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int zleaf_width() { return ZLEAF_W; }

int k_zleaf(hipStream_t s, void *P, long lda, int mp, int w, int pw, int pc0, int *piv) {
  if (w <= 0 || mp <= 0) return PLA_SUCCESS;
  if (w > ZLEAF_W) return fail(PLA_ERR_INVALID_PARAM, "complex panel leaf: width %d > %d", w, ZLEAF_W);
  hipLaunchKernelGGL(zleaf_kernel, dim3(1), dim3(ZLEAF_THREADS), 0, s, (zd *)P, lda, mp, w, pw, pc0, piv);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int ztrsm_leaf_size() { return ZTRSM_LEAF; }

int k_ztrsm_leaf(hipStream_t s, bool lower, bool unit, int b, int ncols, const void *T, long ldt, void *B,
                 long ldb) {
  if (b <= 0 || ncols <= 0) return PLA_SUCCESS;
  if (b > ZTRSM_LEAF) return fail(PLA_ERR_INVALID_PARAM, "complex trsm leaf: b = %d", b);
  const long groups = ((long)ncols + ZTRSM_NC - 1) / ZTRSM_NC, wgs = (groups + 3) / 4;
  const dim3 grid((unsigned)(wgs < 512 ? wgs : 512)), blk(256);
  const zd *t = (const zd *)T;
  zd *x = (zd *)B;
  if (lower && unit)
    hipLaunchKernelGGL((ztrsm_leaf_kernel<true, true>), grid, blk, 0, s, b, ncols, t, ldt, x, ldb);
  else if (lower)
    hipLaunchKernelGGL((ztrsm_leaf_kernel<true, false>), grid, blk, 0, s, b, ncols, t, ldt, x, ldb);
  else if (unit)
    hipLaunchKernelGGL((ztrsm_leaf_kernel<false, true>), grid, blk, 0, s, b, ncols, t, ldt, x, ldb);
  else
    hipLaunchKernelGGL((ztrsm_leaf_kernel<false, false>), grid, blk, 0, s, b, ncols, t, ldt, x, ldb);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_zconjtrans(hipStream_t s, int m, int n, const void *A, long lda, void *B, long ldb) {
  if (m <= 0 || n <= 0) return PLA_SUCCESS;
  const dim3 grid((m + 31) / 32, (n + 31) / 32), blk(32, 8);
  if (grid.y > 65535) return fail(PLA_ERR_DIMENSIONS, "conjugate transpose: %d columns", n);
  hipLaunchKernelGGL(zconjtrans_kernel, grid, blk, 0, s, m, n, (const zd *)A, lda, (zd *)B, ldb);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int zpotrf_leaf_size() { return ZPOTRF_LEAF; }

int k_zpotrf_leaf(hipStream_t s, int b, void *A, long lda, int offset, int *info) {
  if (b <= 0) return PLA_SUCCESS;
  if (b > ZPOTRF_LEAF) return fail(PLA_ERR_INVALID_PARAM, "complex Cholesky leaf: b = %d", b);
  hipLaunchKernelGGL(zpotrf_leaf_kernel, dim3(1), dim3(256), 0, s, b, (zd *)A, lda, offset, info);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_ztrsm_rlc_leaf(hipStream_t s, int nrows, int b, const void *L, long ldl, void *X, long ldx, const int *info) {
  if (nrows <= 0 || b <= 0) return PLA_SUCCESS;
  if (b > ZTRSM_LEAF) return fail(PLA_ERR_INVALID_PARAM, "complex right trsm leaf: b = %d", b);
  hipLaunchKernelGGL(ztrsm_rlc_leaf_kernel, dim3((nrows + 255) / 256), dim3(256), 0, s, nrows, b, (const zd *)L,
                     ldl, (zd *)X, ldx, info);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int zherk_diag_max() { return ZHERK_DIAG; }

int k_zherk_diag(hipStream_t s, int n, int k, const void *X, long ldx, void *C, long ldc, const int *info) {
  if (n <= 0) return PLA_SUCCESS;
  if (n > ZHERK_DIAG) return fail(PLA_ERR_INVALID_PARAM, "complex herk diagonal block: n = %d", n);
  const int t = (n + 15) / 16;
  hipLaunchKernelGGL(zherk_diag_kernel, dim3(t, t), dim3(16, 16), 0, s, n, k, (const zd *)X, ldx, (zd *)C, ldc,
                     info);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

}  // namespace pla
