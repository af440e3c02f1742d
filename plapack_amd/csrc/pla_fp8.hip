// pla_fp8.hip -- FP8 (OCP E4M3) panel quantization and the FP8 MFMA SYRK
// trailing update of the mixed-precision Cholesky.
//
// No counterpart in the reference: PLA_Chol_lower_large's FP64 SYRK
// (SOLVERS/Cholesky/PLA_Chol.c:207-210, BLAS3/PLA_Syrk_perform_local_part.c)
// becomes lower(C) -= L21 L21^T with L21 stored as E4M3 bytes and one
// power-of-two scale per 256 x 256 tile; products are exact in FP32 and
// accumulate in FP32 on gfx950's 16x16x128 FP8 MFMA.
#include "pla_internal.hpp"

namespace {

typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int QT = 256;       // rows per scale tile (and bytes per quantized row)
constexpr int E4M3_MAX = 448;  // largest finite E4M3 value

// One workgroup per 256-row tile of the panel P (R x b FP64, column-major):
// exps[t] = the smallest e with max|tile| <= 448 * 2^e, and row r of Lq
// (QT bytes, row-major) = E4M3(P[r][k] * 2^-e) for k < b, zero for k >= b.
// The value is rounded to FP32 first, then to E4M3 (round to nearest even).
__global__ __launch_bounds__(256) void fp8_quantize_kernel(int R, int b, const double *__restrict__ P, long ldp,
                                                           unsigned char *__restrict__ Lq, int *__restrict__ exps) {
  __shared__ double red[4];
  const int t = blockIdx.x, r0 = t * QT, rows = min(QT, R - r0);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double m = 0.0;
  for (int e = tid; e < rows * b; e += 256) {
    const int j = e / rows, i = e - j * rows;
    m = fmax(m, fabs(P[(long)j * ldp + r0 + i]));
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
  }
  const double sc = ldexp(1.0, -ex);
  if (tid == 0) exps[t] = ex;
  if (tid < rows) {
    const int r = r0 + tid;
    int *dst = (int *)(Lq + (size_t)r * QT);
    for (int k = 0; k < QT; k += 4) {
      float v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) v[u] = (k + u < b) ? (float)(P[(long)(k + u) * ldp + r] * sc) : 0.0f;
      int w = 0;
      w = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], w, false);
      w = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], w, true);
      dst[k >> 2] = w;
    }
  }
}

// lower(C) -= 2^(e_i + e_j) Lq_i Lq_j^T for the R x R lower triangle of C
// (FP32, column-major), Lq R x 256 E4M3 bytes (row-major), exps per 256 rows.
// 128 x 128 tiles (4 waves of 64 x 64), the full k extent staged in LDS.
// MFMA operands: lane l holds row l%16, k bytes [32 (l/16), 32 (l/16) + 32)
// of the 128-wide k block; result lane l, register r holds row 4(l/16) + r,
// column l%16.  The column-side rows are fed first, so lane l ends up with
// C(i0 + l%16, j0 + 4(l/16) + r): 16 consecutive rows of one column.
constexpr int ST = 128;             // output tile
constexpr int QLDS = QT + 16;       // LDS row stride (bytes)

__global__ __launch_bounds__(256, 2) void fp8_syrk_kernel(int R, const unsigned char *__restrict__ Lq,
                                                         const int *__restrict__ exps, float *__restrict__ C,
                                                         long ldc, const int *__restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  // lower tile (ti, tj), ti >= tj, from the linear block index
  const long bid = blockIdx.x;
  int ti = (int)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
  while ((long)(ti + 1) * (ti + 2) / 2 <= bid) ti++;
  while ((long)ti * (ti + 1) / 2 > bid) ti--;
  const int tj = (int)(bid - (long)ti * (ti + 1) / 2);
  const int i0 = ti * ST, j0 = tj * ST;
  __shared__ __align__(16) unsigned char As[ST * QLDS];  // row side
  __shared__ __align__(16) unsigned char Bs[ST * QLDS];  // column side
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int e = tid; e < ST * (QT / 16); e += 256) {
    const int r = e / (QT / 16), c16 = e - r * (QT / 16);
    int4 va = {0, 0, 0, 0}, vb = {0, 0, 0, 0};
    if (i0 + r < R) va = *(const int4 *)(Lq + (size_t)(i0 + r) * QT + c16 * 16);
    if (j0 + r < R) vb = *(const int4 *)(Lq + (size_t)(j0 + r) * QT + c16 * 16);
    *(int4 *)(As + r * QLDS + c16 * 16) = va;
    *(int4 *)(Bs + r * QLDS + c16 * 16) = vb;
  }
  __syncthreads();
  const int wm = wave & 1, wn = wave >> 1, l16 = lane & 15, lq = lane >> 4;
  v4f acc[4][4];
#pragma unroll
  for (int bi = 0; bi < 4; bi++)
#pragma unroll
    for (int bj = 0; bj < 4; bj++) acc[bi][bj] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < QT / 128; ks++) {
    v8i af[4], bf[4];
#pragma unroll
    for (int bi = 0; bi < 4; bi++) {
      const unsigned char *p = As + (wm * 64 + bi * 16 + l16) * QLDS + ks * 128 + lq * 32;
      const int4 x0 = *(const int4 *)p, x1 = *(const int4 *)(p + 16);
      af[bi] = v8i{x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
    }
#pragma unroll
    for (int bj = 0; bj < 4; bj++) {
      const unsigned char *p = Bs + (wn * 64 + bj * 16 + l16) * QLDS + ks * 128 + lq * 32;
      const int4 x0 = *(const int4 *)p, x1 = *(const int4 *)(p + 16);
      bf[bj] = v8i{x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
    }
#pragma unroll
    for (int bi = 0; bi < 4; bi++)
#pragma unroll
      for (int bj = 0; bj < 4; bj++)
        // formats 0/0: E4M3 x E4M3; E8M0 scales 127 = 2^0 (tile scales are applied below)
        acc[bi][bj] =
            __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(bf[bj], af[bi], acc[bi][bj], 0, 0, 0, 127, 0, 127);
  }
#pragma unroll
  for (int bi = 0; bi < 4; bi++) {
    const int i = i0 + wm * 64 + bi * 16 + l16;
    if (i >= R) continue;
    const float si = ldexpf(1.0f, exps[i / QT]);
#pragma unroll
    for (int bj = 0; bj < 4; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int j = j0 + wn * 64 + bj * 16 + lq * 4 + r;
        if (j <= i) {
          const float sc = si * ldexpf(1.0f, exps[j / QT]);
          float *cp = C + (long)j * ldc + i;
          *cp = fmaf(-sc, acc[bi][bj][r], *cp);
        }
      }
  }
}

// W(i, j) = (float)A(i, j) for i >= j, plus shift on the diagonal; the upper
// triangle of A is not read, W's is left as it is.
__global__ __launch_bounds__(256) void cvt_lower_shift_kernel(int n, const double *__restrict__ A, long lda,
                                                              double shift, float *__restrict__ W, long ldw) {
  const int j = blockIdx.y;
  for (int i = j + blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    double v = A[(long)j * lda + i];
    if (i == j) v += shift;
    W[(long)j * ldw + i] = (float)v;
  }
}

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

int k_fp8_quantize(hipStream_t s, int R, int b, const double *P, long ldp, unsigned char *Lq, int *exps) {
  if (R <= 0) return PLA_SUCCESS;
  if (b > QT) return fail(PLA_ERR_INVALID_PARAM, "fp8 quantize: panel width %d > %d", b, QT);
  hipLaunchKernelGGL(fp8_quantize_kernel, dim3((R + QT - 1) / QT), dim3(256), 0, s, R, b, P, ldp, Lq, exps);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_fp8_syrk_lower(hipStream_t s, int R, const unsigned char *Lq, const int *exps, float *C, long ldc,
                     const int *abort_flag) {
  if (R <= 0) return PLA_SUCCESS;
  const long t = (R + ST - 1) / ST;
  const long nblk = t * (t + 1) / 2;
  if (nblk > (1L << 31) - 1) return fail(PLA_ERR_DIMENSIONS, "fp8 syrk: %d rows", R);
  hipLaunchKernelGGL(fp8_syrk_kernel, dim3((unsigned)nblk), dim3(256), 0, s, R, Lq, exps, C, ldc, abort_flag);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

}  // namespace pla
