// pla_krylov.hip -- FP64 kernels of the refinement loops: the residual of a
// symmetric matrix given by its lower triangle, and the block-vector
// operations of GMRES (inner products, combinations, column scaling).
//
// No counterpart in the reference (it has no refinement); the symmetric
// residual reads only the lower triangle, as PLA_Chol does
// (SOLVERS/Cholesky/PLA_Chol.c:83-234 never touches the upper one).
#include "pla_internal.hpp"

namespace {

constexpr int SCH = 512;     // columns (N part) / rows (T part) per chunk
constexpr int SMAXRHS = 16;

// N part, chunk c: part[c][q][i] = sum_{j in chunk, j <= i} a(i, j) x(j, q),
// a = A or |A|.  Rows above the chunk are not written (never read).
template <int NR, bool ABS>
__global__ __launch_bounds__(256) void sresid_n_kernel(int n, int nrhs, const double *__restrict__ A, long lda,
                                                       const double *__restrict__ X, long ldx,
                                                       double *__restrict__ part) {
  __shared__ double xs[SCH][NR];
  const int ch = blockIdx.y, j0 = ch * SCH, j1 = min(n, j0 + SCH);
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (blockIdx.x * 256 + 255 < j0) return;  // whole block above the chunk (uniform)
  for (int e = threadIdx.x; e < SCH * NR; e += 256) {
    const int j = e / NR, q = e - j * NR;
    xs[j][q] = (j0 + j < j1 && q < nrhs) ? X[(long)q * ldx + j0 + j] : 0.0;
  }
  __syncthreads();
  if (i >= n || i < j0) return;
  const int je = min(j1, i + 1);
  double s[NR];
#pragma unroll
  for (int q = 0; q < NR; q++) s[q] = 0.0;
  const double *ap = A + (long)j0 * lda + i;
  int j = j0;
  for (; j + 4 <= je; j += 4, ap += 4 * lda) {
    double a0 = ap[0], a1 = ap[lda], a2 = ap[2 * lda], a3 = ap[3 * lda];
    if (ABS) {
      a0 = fabs(a0);
      a1 = fabs(a1);
      a2 = fabs(a2);
      a3 = fabs(a3);
    }
#pragma unroll
    for (int q = 0; q < NR; q++) {
      s[q] = fma(a0, xs[j - j0][q], s[q]);
      s[q] = fma(a1, xs[j - j0 + 1][q], s[q]);
      s[q] = fma(a2, xs[j - j0 + 2][q], s[q]);
      s[q] = fma(a3, xs[j - j0 + 3][q], s[q]);
    }
  }
  for (; j < je; j++, ap += lda) {
    const double a0 = ABS ? fabs(ap[0]) : ap[0];
#pragma unroll
    for (int q = 0; q < NR; q++) s[q] = fma(a0, xs[j - j0][q], s[q]);
  }
  for (int q = 0; q < nrhs; q++) part[((size_t)ch * nrhs + q) * n + i] = s[q];
}

// T part, chunk c of rows j: part[c][q][i] = sum_{j in chunk, j > i} a(j, i) x(j, q)
// (column i of the strict lower triangle).  16 columns per workgroup, 4 per
// wave; lanes read 64 consecutive rows of a column.
template <int NR, bool ABS>
__global__ __launch_bounds__(256) void sresid_t_kernel(int n, int nrhs, const double *__restrict__ A, long lda,
                                                       const double *__restrict__ X, long ldx,
                                                       double *__restrict__ part) {
  __shared__ double xs[SCH][NR];
  const int ch = blockIdx.y, j0 = ch * SCH, j1 = min(n, j0 + SCH);
  const int ifirst = blockIdx.x * 16;
  if (ch < ifirst / SCH) return;  // chunks before the one holding ifirst are never read
  for (int e = threadIdx.x; e < SCH * NR; e += 256) {
    const int j = e / NR, q = e - j * NR;
    xs[j][q] = (j0 + j < j1 && q < nrhs) ? X[(long)q * ldx + j0 + j] : 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int u = 0; u < 4; u++) {
    const int i = ifirst + wave * 4 + u;
    if (i >= n) break;  // uniform per wave
    double s[NR];
#pragma unroll
    for (int q = 0; q < NR; q++) s[q] = 0.0;
    const double *col = A + (long)i * lda;
    // every load unguarded (clamped into the chunk), then selected: a
    // guarded load would be waited for before the next one issued
    double av[SCH / 64];
#pragma unroll
    for (int t = 0; t < SCH / 64; t++) av[t] = col[min(j0 + t * 64 + lane, j1 - 1)];
#pragma unroll
    for (int t = 0; t < SCH / 64; t++) {
      const int j = j0 + t * 64 + lane;
      if (j < j1 && j > i) {
        const double a = ABS ? fabs(av[t]) : av[t];
#pragma unroll
        for (int q = 0; q < NR; q++) s[q] = fma(a, xs[j - j0][q], s[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < NR; q++)
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) s[q] += __shfl_xor(s[q], o);
    if (lane == 0)
      for (int q = 0; q < nrhs; q++) part[((size_t)ch * nrhs + q) * n + i] = s[q];
  }
}

// R = B - (N + T) with the chunks of row i combined in order by compensated
// (Neumaier) addition; B == nullptr means B = 0.  rmax[q] = max |R(:, q)|.
__global__ __launch_bounds__(256) void sresid_final_kernel(int n, int nrhs, int nch, const double *__restrict__ pn,
                                                           const double *__restrict__ pt,
                                                           const double *__restrict__ B, long ldb,
                                                           double *__restrict__ R, long ldr,
                                                           unsigned long long *__restrict__ rmax) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int q = blockIdx.y;
  if (i >= n) return;
  const int ci = i / SCH;
  double s = 0.0, comp = 0.0;
  auto add = [&](double v) {
    const double t = s + v;
    comp += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
    s = t;
  };
  for (int c = 0; c <= ci; c++) add(pn[((size_t)c * nrhs + q) * n + i]);
  for (int c = ci; c < nch; c++) add(pt[((size_t)c * nrhs + q) * n + i]);
  const double bq = B ? B[(long)q * ldb + i] : 0.0;
  const double r = (bq - s) - comp;
  R[(long)q * ldr + i] = r;
  atomicMax(rmax + q, (unsigned long long)__double_as_longlong(fabs(r)));
}

// ---- GMRES block-vector kernels (nrhs independent systems in lockstep) ----
// V_i = V + i * vstride, column q at V_i + q * ldv.

constexpr int DCH = 4096;  // rows per partial inner product

template <int NR>
__global__ __launch_bounds__(256) void mdot_partial_kernel(int n, int nrhs, const double *__restrict__ V,
                                                           long vstride, long ldv, const double *__restrict__ W,
                                                           long ldw, double *__restrict__ part, int nch) {
  __shared__ double red[4][NR];
  const int c = blockIdx.x, i = blockIdx.y;
  const double *Vi = V + (size_t)i * vstride;
  double s[NR];
#pragma unroll
  for (int q = 0; q < NR; q++) s[q] = 0.0;
  const int r1 = min(n, (c + 1) * DCH);
  for (int r = c * DCH + threadIdx.x; r < r1; r += 256) {
    // all loads first, then the products: a load whose use sits in the same
    // guarded block is waited for before the next one issues
    double v[NR], w[NR];
#pragma unroll
    for (int q = 0; q < NR; q++)
      if (q < nrhs) {
        v[q] = Vi[(long)q * ldv + r];
        w[q] = W[(long)q * ldw + r];
      }
#pragma unroll
    for (int q = 0; q < NR; q++)
      if (q < nrhs) s[q] = fma(v[q], w[q], s[q]);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NR; q++) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s[q] += __shfl_xor(s[q], o);
    if (lane == 0) red[wave][q] = s[q];
  }
  __syncthreads();
  if (threadIdx.x < nrhs) {
    const int q = threadIdx.x;
    part[((size_t)i * nch + c) * nrhs + q] = (red[0][q] + red[1][q]) + (red[2][q] + red[3][q]);
  }
}

__global__ void mdot_final_kernel(int nrhs, int k, int nch, const double *__restrict__ part,
                                  double *__restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= k * nrhs) return;
  const int i = e / nrhs, q = e - i * nrhs;
  double s = 0.0;
  for (int c = 0; c < nch; c++) s += part[((size_t)i * nch + c) * nrhs + q];
  out[e] = s;
}

// W(:, q) += alpha sum_{i<k} h[i][q] V_i(:, q)
__global__ __launch_bounds__(256) void maxpy_kernel(int n, int nrhs, int k, const double *__restrict__ V,
                                                    long vstride, long ldv, const double *__restrict__ h,
                                                    double alpha, double *__restrict__ W, long ldw) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  const int q = blockIdx.y;
  if (r >= n) return;
  double s = 0.0;
#pragma unroll 8
  for (int i = 0; i < k; i++) s = fma(h[i * nrhs + q], V[(size_t)i * vstride + (long)q * ldv + r], s);
  W[(long)q * ldw + r] += alpha * s;
}

// W(:, q) *= f(q), f = s[q] or 1 / s[q]; s[q] == 0 leaves the column zero
__global__ __launch_bounds__(256) void scale_cols_kernel(int n, double *__restrict__ W, long ldw,
                                                         const double *__restrict__ s, int inv) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  const int q = blockIdx.y;
  if (r >= n) return;
  const double sq = s[q];
  const double f = inv ? (sq != 0.0 ? 1.0 / sq : 0.0) : sq;
  W[(long)q * ldw + r] *= f;
}

__global__ __launch_bounds__(256) void fill_kernel(long count, double v, double *__restrict__ X) {
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < count; e += (long)gridDim.x * 256) X[e] = v;
}

}  // namespace

namespace pla {

// R = B - A X with A symmetric, lower triangle referenced (B may be nullptr:
// R = -A X); rmax[q] = max |R(:, q)|, synchronous (rmax == nullptr:
// asynchronous, no norms).  abs_mode: |A| instead of A.
int sym_residual(hipStream_t s, int n, int nrhs, const double *A, long lda, const double *X, long ldx,
                 const double *B, long ldb, double *R, long ldr, double *rmax, bool abs_mode) {
  if (n <= 0 || nrhs <= 0) return PLA_SUCCESS;
  if (nrhs > SMAXRHS) return fail(PLA_ERR_INVALID_PARAM, "sym_residual: %d right-hand sides", nrhs);
  const int nch = (n + SCH - 1) / SCH;
  const size_t psz = sizeof(double) * (size_t)nch * nrhs * n;
  double *pn = (double *)workspace(2 * psz + 256, WS_SRES_PART);
  unsigned long long *mx = (unsigned long long *)workspace(sizeof(unsigned long long) * SMAXRHS, WS_SRES_MAX);
  if (!pn || !mx) return PLA_ERR_NO_MEMORY;
  double *pt = (double *)((char *)pn + psz);
  if (abs_mode && nrhs != 1) return fail(PLA_ERR_INVALID_PARAM, "sym_residual: |A| mode takes one column");
  PLA_HIP(hipMemsetAsync(mx, 0, sizeof(unsigned long long) * SMAXRHS, s));
  const dim3 gn((n + 255) / 256, nch), gt((n + 15) / 16, nch);
#define PLA_SR(NR, AB)                                                                              \
  do {                                                                                              \
    hipLaunchKernelGGL((sresid_n_kernel<NR, AB>), gn, dim3(256), 0, s, n, nrhs, A, lda, X, ldx, pn); \
    hipLaunchKernelGGL((sresid_t_kernel<NR, AB>), gt, dim3(256), 0, s, n, nrhs, A, lda, X, ldx, pt); \
  } while (0)
  timing_begin(s, 5);
  if (abs_mode) PLA_SR(1, true);
  else if (nrhs <= 1) PLA_SR(1, false);
  else if (nrhs <= 4) PLA_SR(4, false);
  else if (nrhs <= 8) PLA_SR(8, false);
  else PLA_SR(16, false);
#undef PLA_SR
  // algorithmic bytes: the lower triangle once (both kernels read it), X, the partials
  timing_end(s, 5, 4.0 * n * (n + 1.0) + 8.0 * nrhs * (double)n * (1 + 2 * nch));
  hipLaunchKernelGGL(sresid_final_kernel, dim3((n + 255) / 256, nrhs), dim3(256), 0, s, n, nrhs, nch, pn, pt, B,
                     ldb, R, ldr, mx);
  PLA_HIP(hipGetLastError());
  if (!rmax) return PLA_SUCCESS;
  unsigned long long h[SMAXRHS];
  PLA_HIP(hipMemcpyAsync(h, mx, sizeof(unsigned long long) * nrhs, hipMemcpyDeviceToHost, s));
  PLA_HIP(hipStreamSynchronize(s));
  for (int q = 0; q < nrhs; q++) rmax[q] = __builtin_bit_cast(double, h[q]);
  return PLA_SUCCESS;
}

// ||A||_inf of a symmetric matrix given by its lower triangle.
int sym_inf_norm(hipStream_t s, int n, const double *A, long lda, double *out) {
  *out = 0.0;
  if (n <= 0) return PLA_SUCCESS;
  double *ones = (double *)workspace(sizeof(double) * (size_t)n, WS_ONES);
  double *r = (double *)workspace(sizeof(double) * (size_t)n, WS_NORM_R);
  if (!ones || !r) return PLA_ERR_NO_MEMORY;
  int rc;
  if ((rc = k_fill(s, n, 1.0, ones))) return rc;
  return sym_residual(s, n, 1, A, lda, ones, n, nullptr, 0, r, n, out, true);
}

int k_fill(hipStream_t s, long count, double v, double *X) {
  if (count <= 0) return PLA_SUCCESS;
  long g = (count + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)g), dim3(256), 0, s, count, v, X);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

// out[i * nrhs + q] = <V_i(:, q), W(:, q)>, i < k (device array)
int k_mdot(hipStream_t s, int n, int nrhs, int k, const double *V, long vstride, long ldv, const double *W,
           long ldw, double *out) {
  if (k <= 0) return PLA_SUCCESS;
  if (nrhs > SMAXRHS) return fail(PLA_ERR_INVALID_PARAM, "mdot: %d columns", nrhs);
  const int nch = (n + DCH - 1) / DCH;
  double *part = (double *)workspace(sizeof(double) * (size_t)k * nch * nrhs + 256, WS_MDOT);
  if (!part) return PLA_ERR_NO_MEMORY;
  const dim3 g(nch, k);
  if (nrhs <= 1)
    hipLaunchKernelGGL(mdot_partial_kernel<1>, g, dim3(256), 0, s, n, nrhs, V, vstride, ldv, W, ldw, part, nch);
  else if (nrhs <= 4)
    hipLaunchKernelGGL(mdot_partial_kernel<4>, g, dim3(256), 0, s, n, nrhs, V, vstride, ldv, W, ldw, part, nch);
  else if (nrhs <= 8)
    hipLaunchKernelGGL(mdot_partial_kernel<8>, g, dim3(256), 0, s, n, nrhs, V, vstride, ldv, W, ldw, part, nch);
  else
    hipLaunchKernelGGL(mdot_partial_kernel<16>, g, dim3(256), 0, s, n, nrhs, V, vstride, ldv, W, ldw, part, nch);
  hipLaunchKernelGGL(mdot_final_kernel, dim3((k * nrhs + 255) / 256), dim3(256), 0, s, nrhs, k, nch, part, out);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_maxpy(hipStream_t s, int n, int nrhs, int k, const double *V, long vstride, long ldv, const double *h,
            double alpha, double *W, long ldw) {
  if (k <= 0 || n <= 0) return PLA_SUCCESS;
  hipLaunchKernelGGL(maxpy_kernel, dim3((n + 255) / 256, nrhs), dim3(256), 0, s, n, nrhs, k, V, vstride, ldv, h,
                     alpha, W, ldw);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_scale_cols(hipStream_t s, int n, int nrhs, double *W, long ldw, const double *f, bool inv) {
  if (n <= 0) return PLA_SUCCESS;
  hipLaunchKernelGGL(scale_cols_kernel, dim3((n + 255) / 256, nrhs), dim3(256), 0, s, n, W, ldw, f, inv ? 1 : 0);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

}  // namespace pla
