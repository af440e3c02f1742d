// pla_convert.hip -- element conversions between the reference's datatypes
// and the complex operand kernels of the complex BLAS3 entry points.
//
// The reference computes MPI_FLOAT / MPI_COMPLEX objects with the single-
// precision BLAS (INCLUDE/PLA_BLAS.h: sgemm, ctrsm, ...).  Here those objects
// are widened to MPI_DOUBLE / MPI_DOUBLE_COMPLEX workspaces, computed on the
// FP64 kernels and rounded back (round to nearest even) -- the results are
// those of the double-precision computation, rounded once.
//
// Complex BLAS3 (BLAS3/PLA_Gemm.c, PLA_Trsm.c, PLA_Syrk.c on
// MPI_DOUBLE_COMPLEX) runs on the real embedding of pla_complex.hip; these
// kernels form its operands: alpha op(A) for op in {N, T, C, conj}, beta C,
// and the triangle-masked accumulation of a complex SYRK.
#include "pla_internal.hpp"

namespace {

struct zd {
  double re, im;
};
struct zf {
  float re, im;
};

unsigned grid_for(long total) {
  long g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  return (unsigned)(g < 1 ? 1 : g);
}

template <typename S, typename D>
__global__ __launch_bounds__(256) void convert2d_kernel(int m, int n, const S *__restrict__ src, long lds_,
                                                        D *__restrict__ dst, long ldd) {
  const long total = (long)m * n;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const long i = t % m, j = t / m;
    dst[i + j * ldd] = (D)src[i + j * lds_];
  }
}

// B (mo x no) = alpha op(A): op 0 = A, 1 = A^T, 2 = A^H, 3 = conj(A)
__global__ __launch_bounds__(256) void zop_kernel(int mo, int no, int op, double are, double aim,
                                                  const zd *__restrict__ A, long lda, zd *__restrict__ B, long ldb) {
  const long total = (long)mo * no;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const long i = t % mo, j = t / mo;
    zd v = (op == 1 || op == 2) ? A[j + i * lda] : A[i + j * lda];
    if (op >= 2) v.im = -v.im;
    B[i + j * ldb] = zd{are * v.re - aim * v.im, are * v.im + aim * v.re};
  }
}

// C = beta C (beta == 0: C = 0, whatever C held)
__global__ __launch_bounds__(256) void zscale2d_kernel(int m, int n, double bre, double bim, zd *__restrict__ C,
                                                       long ldc) {
  const long total = (long)m * n;
  const bool zero = (bre == 0.0 && bim == 0.0);
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const long i = t % m, j = t / m;
    const zd c = C[i + j * ldc];
    C[i + j * ldc] = zero ? zd{0.0, 0.0} : zd{bre * c.re - bim * c.im, bre * c.im + bim * c.re};
  }
}

// the uplo triangle (diagonal included) of C += P, both n x n
__global__ __launch_bounds__(256) void ztri_add_kernel(int n, bool lower, const zd *__restrict__ P, long ldp,
                                                       zd *__restrict__ C, long ldc) {
  const long total = (long)n * n;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const long i = t % n, j = t / n;
    if (lower ? i >= j : i <= j) {
      const zd p = P[i + j * ldp];
      zd c = C[i + j * ldc];
      c.re += p.re;
      c.im += p.im;
      C[i + j * ldc] = c;
    }
  }
}

}  // namespace

namespace pla {

int k_convert2d(hipStream_t s, int m, int n, const void *src, long lds_, MPI_Datatype sdt, void *dst, long ldd,
                MPI_Datatype ddt) {
  if (m <= 0 || n <= 0) return PLA_SUCCESS;
  const unsigned g = grid_for((long)m * n * 2);
  // complex data is the real 2m x n matrix of its parts (leading dimension 2 ld)
  if (sdt == MPI_FLOAT && ddt == MPI_DOUBLE)
    hipLaunchKernelGGL((convert2d_kernel<float, double>), dim3(g), dim3(256), 0, s, m, n, (const float *)src, lds_,
                       (double *)dst, ldd);
  else if (sdt == MPI_DOUBLE && ddt == MPI_FLOAT)
    hipLaunchKernelGGL((convert2d_kernel<double, float>), dim3(g), dim3(256), 0, s, m, n, (const double *)src, lds_,
                       (float *)dst, ldd);
  else if (sdt == MPI_COMPLEX && ddt == MPI_DOUBLE_COMPLEX)
    hipLaunchKernelGGL((convert2d_kernel<float, double>), dim3(g), dim3(256), 0, s, 2 * m, n, (const float *)src,
                       2 * lds_, (double *)dst, 2 * ldd);
  else if (sdt == MPI_DOUBLE_COMPLEX && ddt == MPI_COMPLEX)
    hipLaunchKernelGGL((convert2d_kernel<double, float>), dim3(g), dim3(256), 0, s, 2 * m, n, (const double *)src,
                       2 * lds_, (float *)dst, 2 * ldd);
  else
    return fail(PLA_ERR_UNSUPPORTED, "conversion between these datatypes is not supported");
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_zop(hipStream_t s, int op, int mo, int no, double are, double aim, const void *A, long lda, void *B, long ldb) {
  if (mo <= 0 || no <= 0) return PLA_SUCCESS;
  hipLaunchKernelGGL(zop_kernel, dim3(grid_for((long)mo * no)), dim3(256), 0, s, mo, no, op, are, aim,
                     (const zd *)A, lda, (zd *)B, ldb);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_zscale2d(hipStream_t s, int m, int n, double bre, double bim, void *C, long ldc) {
  if (m <= 0 || n <= 0) return PLA_SUCCESS;
  hipLaunchKernelGGL(zscale2d_kernel, dim3(grid_for((long)m * n)), dim3(256), 0, s, m, n, bre, bim, (zd *)C, ldc);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_ztri_add(hipStream_t s, int n, bool lower, const void *P, long ldp, void *C, long ldc) {
  if (n <= 0) return PLA_SUCCESS;
  hipLaunchKernelGGL(ztri_add_kernel, dim3(grid_for((long)n * n)), dim3(256), 0, s, n, lower, (const zd *)P, ldp,
                     (zd *)C, ldc);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

}  // namespace pla
