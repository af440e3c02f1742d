// pla_fp8.hip -- FP8 (OCP E4M3) panel quantization and the FP8 MFMA trailing
// update of the mixed-precision Cholesky.
//
// No counterpart in the reference: PLA_Chol_lower_large's FP64 SYRK
// (SOLVERS/Cholesky/PLA_Chol.c:207-210, BLAS3/PLA_Syrk_perform_local_part.c)
// becomes lower(C) -= L21 L21^T with L21 stored as E4M3 bytes and one
// power-of-two scale per 256 x 256 tile.  The scales enter gfx950's scaled
// 16x16x128 FP8 MFMA as its E8M0 operands, so products of tiles with
// different scales are exact in FP32 and accumulate in FP32 in one pass.
#include "pla_internal.hpp"

namespace {

typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int QT = 256;        // rows and columns of one scale tile
constexpr int KT_MAX = 4;      // scale tiles along k: panels of up to 1024 columns
constexpr int E4M3_MAX = 448;  // largest finite E4M3 value
constexpr int EXP_LIM = 127;   // an E8M0 scale spans 2^-127 .. 2^127

// One workgroup per 256 x 256 tile of the R x b FP64 panel P (column-major;
// blockIdx.x: row tile t, blockIdx.y: k tile kt, KT = gridDim.y):
// exps[t * KT + kt] = the smallest e with max|tile| <= 448 * 2^e, clamped to
// [-127, 127]; Lq row r (kq bytes, row-major) holds E4M3(P[r][k] * 2^-e) in
// bytes k of the tile, zero for k >= b.  The value is rounded to FP32 first,
// then to E4M3 (round to nearest even).
__global__ __launch_bounds__(256) void fp8_quantize_kernel(int R, int b, const double *__restrict__ P, long ldp,
                                                           unsigned char *__restrict__ Lq, int kq,
                                                           int *__restrict__ exps) {
  __shared__ double red[4];
  const int t = blockIdx.x, kt = blockIdx.y, KT = gridDim.y;
  const int r0 = t * QT, rows = min(QT, R - r0);
  const int c0 = kt * QT, cols = max(0, min(QT, b - c0));
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // Thread tid reads row min(tid, rows - 1) of the tile (rows >= 1, cols >= 1:
  // the grid covers ceil(b / QT) column tiles); clamped rows and columns
  // repeat an element already counted, so every load is unguarded and eight
  // are in flight at a time (a guarded load is waited for where it issues).
  const double *pt = P + (long)c0 * ldp + r0 + min(tid, rows - 1);
  double m = 0.0;
  for (int j = 0; j < cols; j += 8) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; u++) x[u] = pt[(long)min(j + u, cols - 1) * ldp];
#pragma unroll
    for (int u = 0; u < 8; u++) m = fmax(m, fabs(x[u]));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if (lane == 0) red[wave] = m;
  __syncthreads();
  const double amax = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  int ex = 0;
  if (amax > 0.0) {
    int e2;
    const double f = frexp(amax / E4M3_MAX, &e2);  // amax / 448 = f 2^e2, f in [0.5, 1)
    ex = (f == 0.5) ? e2 - 1 : e2;
    ex = min(EXP_LIM, max(-EXP_LIM, ex));
  }
  const double sc = ldexp(1.0, -ex);
  if (tid == 0) exps[t * KT + kt] = ex;
  if (tid < rows) {
    const int r = r0 + tid;
    int *dst = (int *)(Lq + (size_t)r * kq + c0);
#pragma unroll 4
    for (int k = 0; k < QT; k += 4) {
      double x[4];
#pragma unroll
      for (int u = 0; u < 4; u++) x[u] = pt[(long)min(k + u, cols - 1) * ldp];
      float v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) v[u] = (k + u < cols) ? (float)(x[u] * sc) : 0.0f;
      int w = 0;
      w = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], w, false);
      w = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], w, true);
      dst[k >> 2] = w;
    }
  }
}

// C(i, j) -= sum_k 2^(er(i, k) + ec(j, k)) Lr(i, k) Lc(j, k) over the M x N
// FP32 block C (column-major): TRI, the lower triangle of an M x M block
// whose row and column operands are the same rows; else every element.
// Lr, Lc: rows of kq E4M3 bytes (row-major), the first row of each at a
// 256-row scale-tile boundary; er, ec: their exponents, KT per row tile.
//
// 128 x 128 output tiles, 4 waves of 64 x 64 (4 x 4 MFMA blocks).  k runs in
// chunks of 128 bytes (one MFMA k block) through double-buffered LDS: the
// next chunk's global loads are in flight while the current one is
// multiplied, then stored to the other buffer, one barrier per chunk.  A
// chunk lies inside one 256-column scale tile and a 128-row output tile
// inside one 256-row scale tile, so each MFMA takes two wave-uniform E8M0
// scales.  MFMA operands: lane l holds row l%16, bytes [32 (l/16),
// 32 (l/16) + 32) of the chunk; result lane l, register r holds row
// 4(l/16) + r, column l%16 of the first operand's rows x the second's.  The
// column side goes first, so lane l ends up with C(i0 + l%16,
// j0 + 4(l/16) + r): 16 consecutive rows of one column per 16 lanes.
constexpr int ST = 128;         // output tile
constexpr int KC = 128;         // k bytes per chunk
constexpr int LROW = KC + 16;   // LDS row stride (bytes)

struct fp8_upd_args {
  int M, N, nks, kq, KT, tiles_m;
  const unsigned char *Lr, *Lc;
  const int *er, *ec;
  float *C;
  long ldc;
  const int *abort_flag;
};

template <bool TRI>
__global__ __launch_bounds__(256, 2) void fp8_update_kernel(fp8_upd_args a) {
  if (a.abort_flag && *a.abort_flag) return;
  int ti, tj;
  if (TRI) {  // lower tile (ti, tj), ti >= tj, from the linear block index
    const long bid = blockIdx.x;
    ti = (int)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
    while ((long)(ti + 1) * (ti + 2) / 2 <= bid) ti++;
    while ((long)ti * (ti + 1) / 2 > bid) ti--;
    tj = (int)(bid - (long)ti * (ti + 1) / 2);
  } else {
    ti = blockIdx.x % a.tiles_m;
    tj = blockIdx.x / a.tiles_m;
  }
  const int i0 = ti * ST, j0 = tj * ST;
  __shared__ __align__(16) unsigned char As[2][ST * LROW];  // row side
  __shared__ __align__(16) unsigned char Bs[2][ST * LROW];  // column side
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int *erow = a.er + (i0 / QT) * a.KT, *ecol = a.ec + (j0 / QT) * a.KT;

  // 128 rows x 128 bytes per operand and chunk: 4 int4 per thread, 8 threads a row
  int4 ra[4], rb[4];
  auto gload = [&](int ks) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int e = tid + 256 * u, r = e >> 3, c = e & 7;
      // rows past M (N) read the last row: they feed only products that are
      // never stored, and an unconditional load is a plain global load (a
      // select between the row and a zero made it a flat load, waited for
      // where it issued)
      ra[u] = *(const int4 *)(a.Lr + (size_t)min(i0 + r, a.M - 1) * a.kq + ks * KC + c * 16);
      rb[u] = *(const int4 *)(a.Lc + (size_t)min(j0 + r, a.N - 1) * a.kq + ks * KC + c * 16);
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int e = tid + 256 * u, r = e >> 3, c = e & 7;
      *(int4 *)(As[buf] + r * LROW + c * 16) = ra[u];
      *(int4 *)(Bs[buf] + r * LROW + c * 16) = rb[u];
    }
  };

  const int wm = wave & 1, wn = wave >> 1, l16 = lane & 15, lq = lane >> 4;
  v4f acc[4][4];
#pragma unroll
  for (int bi = 0; bi < 4; bi++)
#pragma unroll
    for (int bj = 0; bj < 4; bj++) acc[bi][bj] = v4f{0.f, 0.f, 0.f, 0.f};

  gload(0);
  sstore(0);
  __syncthreads();
  for (int ks = 0; ks < a.nks; ks++) {
    const bool more = ks + 1 < a.nks;
    if (more) gload(ks + 1);
    const int kt = (ks * KC) / QT;
    const int scr = 127 + erow[kt], scc = 127 + ecol[kt];
    const unsigned char *as = As[ks & 1], *bs = Bs[ks & 1];
    v8i af[4], bf[4];
#pragma unroll
    for (int bi = 0; bi < 4; bi++) {
      const unsigned char *p = as + (wm * 64 + bi * 16 + l16) * LROW + lq * 32;
      const int4 x0 = *(const int4 *)p, x1 = *(const int4 *)(p + 16);
      af[bi] = v8i{x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
    }
#pragma unroll
    for (int bj = 0; bj < 4; bj++) {
      const unsigned char *p = bs + (wn * 64 + bj * 16 + l16) * LROW + lq * 32;
      const int4 x0 = *(const int4 *)p, x1 = *(const int4 *)(p + 16);
      bf[bj] = v8i{x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
    }
#pragma unroll
    for (int bi = 0; bi < 4; bi++)
#pragma unroll
      for (int bj = 0; bj < 4; bj++)
        // formats 0/0: E4M3 x E4M3; E8M0 scales 127 + e (2^e)
        acc[bi][bj] =
            __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(bf[bj], af[bi], acc[bi][bj], 0, 0, 0, scc, 0, scr);
    if (more) sstore((ks + 1) & 1);
    __syncthreads();
  }
  // C in two halves of 32 elements per lane: each half's loads are all issued
  // before its stores, so they are in flight together
  float *__restrict__ C = a.C;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    float cv[2][4][4];
#pragma unroll
    for (int b2 = 0; b2 < 2; b2++)
#pragma unroll
      for (int bj = 0; bj < 4; bj++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = i0 + wm * 64 + (2 * h + b2) * 16 + l16, j = j0 + wn * 64 + bj * 16 + lq * 4 + r;
          cv[b2][bj][r] = C[(long)min(j, a.N - 1) * a.ldc + min(i, a.M - 1)];  // clamped, unguarded
        }
#pragma unroll
    for (int b2 = 0; b2 < 2; b2++)
#pragma unroll
      for (int bj = 0; bj < 4; bj++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = i0 + wm * 64 + (2 * h + b2) * 16 + l16, j = j0 + wn * 64 + bj * 16 + lq * 4 + r;
          if (i < a.M && j < a.N && (!TRI || j <= i))
            C[(long)j * a.ldc + i] = cv[b2][bj][r] - acc[2 * h + b2][bj][r];
        }
  }
}

// W(i, j) = (float)A(i, j) for i >= j, plus shift on the diagonal; the upper
// triangle of A is not read, W's is left as it is.
__global__ __launch_bounds__(256) void cvt_lower_shift_kernel(int n, const double *__restrict__ A, long lda,
                                                              double shift, float *__restrict__ W, long ldw) {
  const int j = blockIdx.y;
  const long stride = (long)gridDim.x * 256;
  // four rows per thread and step, loads first (clamped to the column)
  for (long i0 = j + blockIdx.x * 256 + threadIdx.x; i0 < n; i0 += 4 * stride) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) v[u] = A[(long)j * lda + min(i0 + u * stride, (long)n - 1)];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const long i = i0 + u * stride;
      if (i < n) W[(long)j * ldw + i] = (float)(i == j ? v[u] + shift : v[u]);
    }
  }
}

// D(j, i) = S(i, j) for i > j, and D(i, j) = S(i, j) for i >= j when D is
// not S: the full symmetric matrix whose lower triangle S holds (D == S: the
// strict upper triangle of S receives the transpose of the strict lower one).
// One workgroup per 64 x 64 tile (ti, tj), ti >= tj, through LDS (padded
// rows: both passes are conflict-free).
template <typename T>
__global__ __launch_bounds__(256) void mirror_lower_kernel(int n, const T *S, long lds, T *D, long ldd) {
  __shared__ T t[64][65];
  const long bid = blockIdx.x;
  int ti = (int)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
  while ((long)(ti + 1) * (ti + 2) / 2 <= bid) ti++;
  while ((long)ti * (ti + 1) / 2 > bid) ti--;
  const int tj = (int)(bid - (long)ti * (ti + 1) / 2);
  const int i0 = ti * 64, j0 = tj * 64;
  const bool copy = (S != D);
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int c = e >> 6, r = e & 63, i = i0 + r, j = j0 + c;
    if (i < n && i >= j) {
      const T v = S[(long)j * lds + i];
      t[c][r] = v;
      if (copy) D[(long)j * ldd + i] = v;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int r = e >> 6, c = e & 63, i = i0 + r, j = j0 + c;
    if (i < n && i > j) D[(long)i * ldd + j] = t[c][r];
  }
}

// *out = max_i A(i, i), one workgroup.
__global__ __launch_bounds__(256) void max_diag_kernel(int n, const double *__restrict__ A, long lda,
                                                       double *__restrict__ out) {
  __shared__ double red[4];
  double m = -HUGE_VAL;
  for (int i = threadIdx.x; i < n; i += 256) m = fmax(m, A[(long)i * lda + i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) *out = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

__global__ __launch_bounds__(256) void eye_kernel(int n, double *__restrict__ T, long ldt) {
  const int j = blockIdx.y;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) T[(long)j * ldt + i] = i == j ? 1.0 : 0.0;
}

long tri_blocks(long t) { return t * (t + 1) / 2; }

}  // namespace

namespace pla {

int k_cvt_lower_shift(hipStream_t s, int n, const double *A, long lda, double shift, float *W, long ldw) {
  if (n <= 0) return PLA_SUCCESS;
  int gx = (n + 255) / 256;
  if (gx > 64) gx = 64;
  hipLaunchKernelGGL(cvt_lower_shift_kernel, dim3(gx, n), dim3(256), 0, s, n, A, lda, shift, W, ldw);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int fp8_tile_rows() { return QT; }
int fp8_row_bytes(int b) { return QT * ((b + QT - 1) / QT); }

int k_fp8_quantize(hipStream_t s, int R, int b, const double *P, long ldp, unsigned char *Lq, int *exps) {
  if (R <= 0) return PLA_SUCCESS;
  if (b < 1 || b > KT_MAX * QT) return fail(PLA_ERR_INVALID_PARAM, "fp8 quantize: panel width %d", b);
  const int kq = fp8_row_bytes(b);
  hipLaunchKernelGGL(fp8_quantize_kernel, dim3((R + QT - 1) / QT, kq / QT), dim3(256), 0, s, R, b, P, ldp, Lq, kq,
                     exps);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_fp8_update(hipStream_t s, bool tri, int M, int N, int b, const unsigned char *Lq, const int *exps, int ro,
                 int co, float *C, long ldc, const int *abort_flag) {
  if (M <= 0 || N <= 0) return PLA_SUCCESS;
  if (b < 1 || b > KT_MAX * QT || ro % QT || co % QT || (tri && (M != N || ro != co)))
    return fail(PLA_ERR_INVALID_PARAM, "fp8 update: b %d, rows from %d, columns from %d", b, ro, co);
  const int kq = fp8_row_bytes(b), KT = kq / QT;
  fp8_upd_args a;
  a.M = M;
  a.N = N;
  a.nks = (b + KC - 1) / KC;
  a.kq = kq;
  a.KT = KT;
  a.tiles_m = (M + ST - 1) / ST;
  a.Lr = Lq + (size_t)ro * kq;
  a.Lc = Lq + (size_t)co * kq;
  a.er = exps + (ro / QT) * KT;
  a.ec = exps + (co / QT) * KT;
  a.C = C;
  a.ldc = ldc;
  a.abort_flag = abort_flag;
  const long tn = (N + ST - 1) / ST;
  const long nblk = tri ? tri_blocks(a.tiles_m) : a.tiles_m * tn;
  if (nblk > (1L << 31) - 1) return fail(PLA_ERR_DIMENSIONS, "fp8 update: %d x %d", M, N);
  if (tri) hipLaunchKernelGGL(fp8_update_kernel<true>, dim3((unsigned)nblk), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(fp8_update_kernel<false>, dim3((unsigned)nblk), dim3(256), 0, s, a);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_fp8_syrk_lower(hipStream_t s, int R, const unsigned char *Lq, const int *exps, float *C, long ldc,
                     const int *abort_flag) {
  return k_fp8_update(s, true, R, R, QT, Lq, exps, 0, 0, C, ldc, abort_flag);
}

int k_mirror_lower_f32(hipStream_t s, int n, float *W, long ldw) {
  if (n <= 1) return PLA_SUCCESS;
  const long nblk = tri_blocks((n + 63) / 64);
  if (nblk > (1L << 31) - 1) return fail(PLA_ERR_DIMENSIONS, "mirror: n = %d", n);
  hipLaunchKernelGGL(mirror_lower_kernel<float>, dim3((unsigned)nblk), dim3(256), 0, s, n, W, ldw, W, ldw);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_sym_full(hipStream_t s, int n, const double *A, long lda, double *S, long lds) {
  if (n <= 0) return PLA_SUCCESS;
  const long nblk = tri_blocks((n + 63) / 64);
  if (nblk > (1L << 31) - 1) return fail(PLA_ERR_DIMENSIONS, "sym_full: n = %d", n);
  hipLaunchKernelGGL(mirror_lower_kernel<double>, dim3((unsigned)nblk), dim3(256), 0, s, n, A, lda, S, lds);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_eye(hipStream_t s, int n, double *T, long ldt) {
  if (n <= 0) return PLA_SUCCESS;
  hipLaunchKernelGGL(eye_kernel, dim3((n + 255) / 256, n), dim3(256), 0, s, n, T, ldt);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_max_diag(hipStream_t s, int n, const double *A, long lda, double *out) {
  if (n <= 0) return fail(PLA_ERR_DIMENSIONS, "max_diag: n = %d", n);
  hipLaunchKernelGGL(max_diag_kernel, dim3(1), dim3(256), 0, s, n, A, lda, out);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

}  // namespace pla
