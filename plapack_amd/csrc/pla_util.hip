// pla_util.hip -- the level-1/2 helpers and utilities that PLAPACK programs
// call around the factor-and-solve path (the residual checks and problem
// generators of TEST/ and EXAMPLES/), on a 1x1 template, device resident:
//
//   PLA_Axpy   BLAS1/PLA_Axpy.c:13     y <- alpha x + y
//   PLA_Nrm2   BLAS1/PLA_Nrm2.c:13     alpha <- ||x||_2 (dnrm2 / snrm2 / dznrm2)
//   PLA_Symv   BLAS2/PLA_Symv.c:13     y <- alpha A x + beta y, A symmetric
//   PLA_Hemv   BLAS2/PLA_Hemv.c:13     y <- alpha A x + beta y, A Hermitian
//   PLA_Shift  UTIL/PLA_Shift.c:15     diag(A) += value
//   PLA_Abort, PLA_malloc / PLA_calloc / PLA_free (OBJ/PLA_Diagnostics.c:108,
//   OBJ/PLA_Memory.c:18-80)
//
// None of these is on the factorization's critical path: the kernels are
// simple, coalesced along columns, and accumulate in FP64.
#include "pla_internal.hpp"

#include <dlfcn.h>

namespace pla {
namespace {

struct zd {
  double re, im;
};
__host__ __device__ inline zd operator+(zd a, zd b) { return {a.re + b.re, a.im + b.im}; }
__host__ __device__ inline zd operator*(zd a, zd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__host__ __device__ inline zd conj(zd a) { return {a.re, -a.im}; }

// arithmetic type of each element type: FP64 for the real types, zd for complex
template <typename T> struct acc;
template <> struct acc<float> {
  using type = double;
  __device__ static double load(float v) { return v; }
  __device__ static float store(double v) { return (float)v; }
  __device__ static double cj(double v) { return v; }
  __device__ static double real_only(double v) { return v; }
  __device__ static double zero() { return 0.0; }
};
template <> struct acc<double> {
  using type = double;
  __device__ static double load(double v) { return v; }
  __device__ static double store(double v) { return v; }
  __device__ static double cj(double v) { return v; }
  __device__ static double real_only(double v) { return v; }
  __device__ static double zero() { return 0.0; }
};
template <> struct acc<zd> {
  using type = zd;
  __device__ static zd load(zd v) { return v; }
  __device__ static zd store(zd v) { return v; }
  __device__ static zd cj(zd v) { return conj(v); }
  __device__ static zd real_only(zd v) { return {v.re, 0.0}; }
  __device__ static zd zero() { return {0.0, 0.0}; }
};

template <typename T, typename S>
__global__ __launch_bounds__(256) void axpy2d_kernel(int m, int n, S alpha, const T *__restrict__ X, long ldx,
                                                     T *__restrict__ Y, long ldy) {
  using A = acc<T>;
  const long total = (long)m * n;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const long i = t % m, j = t / m;
    Y[i + j * ldy] = A::store(alpha * A::load(X[i + j * ldx]) + A::load(Y[i + j * ldy]));
  }
}

__device__ inline double abs_max_part(double v) { return fabs(v); }
__device__ inline double abs_max_part(float v) { return fabs((double)v); }
__device__ inline double abs_max_part(zd v) { return fmax(fabs(v.re), fabs(v.im)); }
__device__ inline double sq_scaled(double v, double s) { return (v * s) * (v * s); }
__device__ inline double sq_scaled(float v, double s) { return ((double)v * s) * ((double)v * s); }
__device__ inline double sq_scaled(zd v, double s) { return (v.re * s) * (v.re * s) + (v.im * s) * (v.im * s); }

// block reduction of a double by `op` (256 threads)
template <typename Op>
__device__ double block_reduce(double v, Op op) {
  __shared__ double red[4];
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o));
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) red[w] = v;
  __syncthreads();
  v = red[0];
  for (int i = 1; i < (int)(blockDim.x >> 6); i++) v = op(v, red[i]);
  __syncthreads();
  return v;
}

// Norm in two passes over x (n elements, stride inc), one workgroup each:
// the largest |component| first, then the sum of squares scaled by its
// reciprocal, as dnrm2 scales to avoid overflow.  out[0] = ||x||_2.
template <typename T>
__global__ __launch_bounds__(256) void nrm2_kernel(int n, const T *__restrict__ X, long inc, double *out) {
  double mx = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) mx = fmax(mx, abs_max_part(X[(long)i * inc]));
  mx = block_reduce(mx, [](double a, double b) { return fmax(a, b); });
  if (!(mx > 0.0) || isinf(mx)) {  // zero vector, inf or nan: nothing to scale
    if (threadIdx.x == 0) out[0] = mx;
    return;
  }
  const double s = 1.0 / mx;
  double ss = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) ss += sq_scaled(X[(long)i * inc], s);
  ss = block_reduce(ss, [](double a, double b) { return a + b; });
  if (threadIdx.x == 0) out[0] = mx * sqrt(ss);
}

// t_i = sum over j of the stored triangle's row i, diagonal included:
// lower: j <= i, upper: j >= i.  Thread per row; for a fixed j consecutive
// threads read consecutive rows of column j (coalesced).  The Hermitian
// diagonal contributes its real part only.
template <typename T, bool LOWER, bool HERM>
__global__ __launch_bounds__(256) void tri_rows_kernel(int n, const T *__restrict__ Am, long lda,
                                                       const T *__restrict__ x, T *__restrict__ t) {
  using A = acc<T>;
  using V = typename A::type;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V s = A::zero();
  const int j0 = LOWER ? 0 : i, j1 = LOWER ? i : n - 1;
  for (int j = j0; j <= j1; j++) {
    V a = A::load(Am[i + (long)j * lda]);
    if (HERM && j == i) a = A::real_only(a);
    s = s + a * A::load(x[j]);
  }
  t[i] = A::store(s);
}

// y_j = alpha (t_j + u_j) + beta y_j with u_j the strict triangle's column j
// applied transposed (conjugated when Hermitian): lower: sum over i > j of
// A(i,j) x_i, upper: sum over i < j.  One wave per column.
template <typename T, bool LOWER, bool HERM>
__global__ __launch_bounds__(256) void tri_cols_kernel(int n, const T *__restrict__ Am, long lda,
                                                       const T *__restrict__ x, const T *__restrict__ t,
                                                       typename acc<T>::type alpha, typename acc<T>::type beta,
                                                       bool beta_zero, T *__restrict__ y) {
  using A = acc<T>;
  using V = typename A::type;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
  if (j >= n) return;
  const int i0 = LOWER ? j + 1 : 0, i1 = LOWER ? n : j;
  V s = A::zero();
  for (int i = i0 + l; i < i1; i += 64) {
    V a = A::load(Am[i + (long)j * lda]);
    if (HERM) a = A::cj(a);
    s = s + a * A::load(x[i]);
  }
  // wave sum (component-wise for complex)
  if constexpr (sizeof(V) == sizeof(double)) {
    for (int o = 32; o > 0; o >>= 1) s = s + __shfl_xor(s, o);
  } else {
    for (int o = 32; o > 0; o >>= 1) {
      V r;
      r.re = __shfl_xor(s.re, o);
      r.im = __shfl_xor(s.im, o);
      s = s + r;
    }
  }
  if (l == 0) {
    const V v = alpha * (A::load(t[j]) + s);
    y[j] = A::store(beta_zero ? v : v + beta * A::load(y[j]));
  }
}

template <typename T>
__global__ __launch_bounds__(256) void shift_diag_kernel(int nd, T *__restrict__ Am, long lda, double v) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d < nd) Am[d + (long)d * lda] = (T)((double)Am[d + (long)d * lda] + v);
}

unsigned grid_for(long total) {
  long g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  return (unsigned)(g < 1 ? 1 : g);
}

// alpha of an object's datatype, read to the host
int scalar_of(PLA_Obj o, MPI_Datatype dt, zd *v) {
  if (!valid(o)) return fail(PLA_ERR_INVALID_OBJECT, "invalid scalar object");
  if (o->base->datatype != dt) return fail(PLA_ERR_INVALID_PARAM, "scalar datatype differs from the operands'");
  if (dt != MPI_DOUBLE_COMPLEX) {
    double r;
    int rc = scalar_value(o, &r);
    if (rc) return rc;
    *v = {r, 0.0};
    return PLA_SUCCESS;
  }
  void *p;
  int rc = dev_ptr(o, &p);
  if (rc) return rc;
  PLA_HIP(hipStreamSynchronize(ctx().stream));
  PLA_HIP(hipMemcpy(v, p, sizeof(zd), hipMemcpyDeviceToHost));
  return PLA_SUCCESS;
}

bool real_or_complex(MPI_Datatype dt) { return dt == MPI_DOUBLE || dt == MPI_FLOAT || dt == MPI_DOUBLE_COMPLEX; }

template <typename T>
int symv_launch(hipStream_t s, bool lower, bool herm, int n, const void *A, long lda, const void *x, void *t,
                typename acc<T>::type al, typename acc<T>::type be, bool beta_zero, void *y) {
  const T *a = (const T *)A, *xv = (const T *)x;
  T *tv = (T *)t, *yv = (T *)y;
  const dim3 gr((n + 255) / 256), gc((n + 3) / 4);
#define PLA_SYMV_CASE(LO, HE)                                                                            \
  do {                                                                                                   \
    hipLaunchKernelGGL((tri_rows_kernel<T, LO, HE>), gr, dim3(256), 0, s, n, a, lda, xv, tv);            \
    hipLaunchKernelGGL((tri_cols_kernel<T, LO, HE>), gc, dim3(256), 0, s, n, a, lda, xv, (const T *)tv, \
                       al, be, beta_zero, yv);                                                           \
  } while (0)
  if (lower && herm) PLA_SYMV_CASE(true, true);
  else if (lower) PLA_SYMV_CASE(true, false);
  else if (herm) PLA_SYMV_CASE(false, true);
  else PLA_SYMV_CASE(false, false);
#undef PLA_SYMV_CASE
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int symv_common(const char *who, bool herm, int uplo, PLA_Obj alpha, PLA_Obj A, PLA_Obj x, PLA_Obj beta, PLA_Obj y) {
  if (!valid(A) || !valid(x) || !valid(y)) return fail(PLA_ERR_INVALID_OBJECT, "%s: invalid object", who);
  const MPI_Datatype dt = A->base->datatype;
  if (herm ? dt != MPI_DOUBLE_COMPLEX : (dt != MPI_DOUBLE && dt != MPI_FLOAT))
    return fail(PLA_ERR_UNSUPPORTED, "%s: A must be %s", who, herm ? "MPI_DOUBLE_COMPLEX" : "MPI_DOUBLE or MPI_FLOAT");
  if (x->base->datatype != dt || y->base->datatype != dt)
    return fail(PLA_ERR_INVALID_PARAM, "%s: x and y must have A's datatype", who);
  if (uplo != PLA_LOWER_TRIANGULAR && uplo != PLA_UPPER_TRIANGULAR)
    return fail(PLA_ERR_INVALID_PARAM, "%s: uplo %d", who, uplo);
  const int n = A->length;
  if (A->width != n || x->length != n || y->length != n || x->width != 1 || y->width != 1)
    return fail(PLA_ERR_DIMENSIONS, "%s: A %d x %d, x %d x %d, y %d x %d", who, A->length, A->width, x->length,
                x->width, y->length, y->width);
  if (n == 0) return PLA_SUCCESS;
  zd al{1.0, 0.0}, be{1.0, 0.0};
  int rc;
  if ((alpha && (rc = scalar_of(alpha, dt, &al))) || (beta && (rc = scalar_of(beta, dt, &be)))) return rc;
  void *a, *xv, *yv;
  if ((rc = dev_ptr(A, &a)) || (rc = dev_ptr(x, &xv)) || (rc = dev_ptr(y, &yv))) return rc;
  const size_t es = (size_t)A->base->typesize;
  void *t = workspace(es * (size_t)n, WS_UTIL);
  if (!t) return PLA_ERR_NO_MEMORY;
  const bool lower = (uplo == PLA_LOWER_TRIANGULAR);
  const bool beta_zero = (be.re == 0.0 && be.im == 0.0);
  hipStream_t s = ctx().stream;
  if (dt == MPI_DOUBLE_COMPLEX) rc = symv_launch<zd>(s, lower, true, n, a, ld(A), xv, t, al, be, beta_zero, yv);
  else if (dt == MPI_DOUBLE) rc = symv_launch<double>(s, lower, false, n, a, ld(A), xv, t, al.re, be.re, beta_zero, yv);
  else rc = symv_launch<float>(s, lower, false, n, a, ld(A), xv, t, al.re, be.re, beta_zero, yv);
  if (rc) return rc;
  return written(y);
}

}  // namespace

int k_shift_diag(hipStream_t s, int nd, double *a, long lda, double v) {
  if (nd <= 0) return PLA_SUCCESS;
  hipLaunchKernelGGL(shift_diag_kernel<double>, dim3((nd + 255) / 256), dim3(256), 0, s, nd, a, lda, v);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

}  // namespace pla

using namespace pla;

extern "C" {

int PLA_Axpy(PLA_Obj alpha, PLA_Obj x, PLA_Obj y) {
  if (valid(y) && y->base->datatype == MPI_COMPLEX)
    return on_promoted({&alpha, &x, &y}, {false, false, true}, [&] { return PLA_Axpy(alpha, x, y); });
  if (!valid(x) || !valid(y)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Axpy: invalid object");
  const MPI_Datatype dt = y->base->datatype;
  if (!real_or_complex(dt))
    return fail(PLA_ERR_UNSUPPORTED, "PLA_Axpy: MPI_DOUBLE, MPI_FLOAT or MPI_DOUBLE_COMPLEX objects only");
  if (x->base->datatype != dt) return fail(PLA_ERR_INVALID_PARAM, "PLA_Axpy: x and y differ in datatype");
  if (x->length != y->length || x->width != y->width)
    return fail(PLA_ERR_DIMENSIONS, "PLA_Axpy: x %d x %d, y %d x %d", x->length, x->width, y->length, y->width);
  const int m = y->length, n = y->width;
  if (m == 0 || n == 0) return PLA_SUCCESS;
  zd al;
  int rc = scalar_of(alpha, dt, &al);
  if (rc) return rc;
  void *xv, *yv;
  if ((rc = dev_ptr(x, &xv)) || (rc = dev_ptr(y, &yv))) return rc;
  hipStream_t s = ctx().stream;
  const unsigned g = grid_for((long)m * n);
  if (dt == MPI_DOUBLE_COMPLEX)
    hipLaunchKernelGGL((axpy2d_kernel<zd, zd>), dim3(g), dim3(256), 0, s, m, n, al, (const zd *)xv, (long)ld(x),
                       (zd *)yv, (long)ld(y));
  else if (dt == MPI_DOUBLE)
    hipLaunchKernelGGL((axpy2d_kernel<double, double>), dim3(g), dim3(256), 0, s, m, n, al.re,
                       (const double *)xv, (long)ld(x), (double *)yv, (long)ld(y));
  else
    hipLaunchKernelGGL((axpy2d_kernel<float, double>), dim3(g), dim3(256), 0, s, m, n, al.re, (const float *)xv,
                       (long)ld(x), (float *)yv, (long)ld(y));
  PLA_HIP(hipGetLastError());
  return written(y);
}

int PLA_Nrm2(PLA_Obj x, PLA_Obj alpha) {
  if (valid(x) && x->base->datatype == MPI_COMPLEX)
    return on_promoted({&x, &alpha}, {false, true}, [&] { return PLA_Nrm2(x, alpha); });
  if (!valid(x) || !valid(alpha)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Nrm2: invalid object");
  const MPI_Datatype dt = x->base->datatype;
  if (!real_or_complex(dt))
    return fail(PLA_ERR_UNSUPPORTED, "PLA_Nrm2: MPI_DOUBLE, MPI_FLOAT or MPI_DOUBLE_COMPLEX vectors only");
  const MPI_Datatype rt = dt == MPI_FLOAT ? MPI_FLOAT : MPI_DOUBLE;
  if (alpha->base->datatype != rt)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Nrm2: alpha must be %s", rt == MPI_FLOAT ? "MPI_FLOAT" : "MPI_DOUBLE");
  if (x->width != 1 && x->length != 1)
    return fail(PLA_ERR_DIMENSIONS, "PLA_Nrm2: x is %d x %d, not a vector", x->length, x->width);
  const int n = x->width == 1 ? x->length : x->width;
  const long inc = x->width == 1 ? 1 : ld(x);
  void *xv, *av;
  int rc;
  if ((rc = dev_ptr(x, &xv)) || (rc = dev_ptr(alpha, &av))) return rc;
  double *out = (double *)workspace(sizeof(double), WS_UTIL_NORM);
  if (!out) return PLA_ERR_NO_MEMORY;
  hipStream_t s = ctx().stream;
  if (n == 0) {
    PLA_HIP(hipMemsetAsync(out, 0, sizeof(double), s));
  } else if (dt == MPI_DOUBLE_COMPLEX) {
    hipLaunchKernelGGL(nrm2_kernel<zd>, dim3(1), dim3(256), 0, s, n, (const zd *)xv, inc, out);
  } else if (dt == MPI_DOUBLE) {
    hipLaunchKernelGGL(nrm2_kernel<double>, dim3(1), dim3(256), 0, s, n, (const double *)xv, inc, out);
  } else {
    hipLaunchKernelGGL(nrm2_kernel<float>, dim3(1), dim3(256), 0, s, n, (const float *)xv, inc, out);
  }
  PLA_HIP(hipGetLastError());
  double h = 0.0;
  PLA_HIP(hipMemcpyAsync(&h, out, sizeof h, hipMemcpyDeviceToHost, s));
  PLA_HIP(hipStreamSynchronize(s));
  if (rt == MPI_FLOAT) {
    const float f = (float)h;
    PLA_HIP(hipMemcpy(av, &f, sizeof f, hipMemcpyHostToDevice));
  } else {
    PLA_HIP(hipMemcpy(av, &h, sizeof h, hipMemcpyHostToDevice));
  }
  return written(alpha);
}

int PLA_Symv(int uplo, PLA_Obj alpha, PLA_Obj A, PLA_Obj x, PLA_Obj beta, PLA_Obj y) {
  return symv_common("PLA_Symv", false, uplo, alpha, A, x, beta, y);
}

int PLA_Hemv(int uplo, PLA_Obj alpha, PLA_Obj A, PLA_Obj x, PLA_Obj beta, PLA_Obj y) {
  if (valid(A) && A->base->datatype == MPI_COMPLEX)
    return on_promoted({&alpha, &A, &x, &beta, &y}, {false, false, false, false, true},
                       [&] { return PLA_Hemv(uplo, alpha, A, x, beta, y); });
  return symv_common("PLA_Hemv", true, uplo, alpha, A, x, beta, y);
}

int PLA_Shift(PLA_Obj obj, MPI_Datatype datatype, void *value) {
  if (!valid(obj)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Shift: invalid object");
  if (!value) return fail(PLA_ERR_INVALID_PARAM, "PLA_Shift: value is NULL");
  double v;
  if (datatype == MPI_DOUBLE) v = *(const double *)value;
  else if (datatype == MPI_FLOAT) v = *(const float *)value;
  else return fail(PLA_ERR_UNSUPPORTED, "PLA_Shift: value must be MPI_DOUBLE or MPI_FLOAT");
  const MPI_Datatype dt = obj->base->datatype;
  if (dt != MPI_DOUBLE && dt != MPI_FLOAT)
    return fail(PLA_ERR_UNSUPPORTED, "PLA_Shift: obj must be MPI_DOUBLE or MPI_FLOAT");
  const int nd = obj->length < obj->width ? obj->length : obj->width;
  if (nd == 0) return PLA_SUCCESS;
  void *a;
  int rc = dev_ptr(obj, &a);
  if (rc) return rc;
  hipStream_t s = ctx().stream;
  if (dt == MPI_DOUBLE)
    hipLaunchKernelGGL(shift_diag_kernel<double>, dim3((nd + 255) / 256), dim3(256), 0, s, nd, (double *)a,
                       (long)ld(obj), v);
  else
    hipLaunchKernelGGL(shift_diag_kernel<float>, dim3((nd + 255) / 256), dim3(256), 0, s, nd, (float *)a,
                       (long)ld(obj), v);
  PLA_HIP(hipGetLastError());
  return written(obj);
}

int PLA_Abort(char *message, int line, char *file) {
  fprintf(stderr, "PLAPACK abort: %s (line %d of %s)\n", message ? message : "", line, file ? file : "?");
  fflush(stderr);
  typedef int (*initialized_fn)(int *);
  typedef int (*abort_fn)(MPI_Comm, int);
  initialized_fn initd = (initialized_fn)dlsym(RTLD_DEFAULT, "MPI_Initialized");
  abort_fn mpi_abort = (abort_fn)dlsym(RTLD_DEFAULT, "MPI_Abort");
  int flag = 0;
  if (initd && mpi_abort && initd(&flag) == 0 && flag) mpi_abort(MPI_COMM_WORLD, 1);
  exit(1);
}

void *PLA_malloc(size_t size) { return malloc(size); }

void *PLA_calloc(size_t size, size_t typesize) { return calloc(size, typesize); }

void *PLA_free(void *buffer) {
  free(buffer);
  return NULL;
}

}  // extern "C"
