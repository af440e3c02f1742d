// pla_mixed.hip -- mixed-precision LU with FP64 iterative refinement, and the
// FP64-accurate solve kernels it needs (gfx950).
//
// No reference counterpart: this is the precision mode added beside PLA_LU
// (SOLVERS/LU/PLA_LU.c:13) and PLA_General_solve (SOLVERS/LU/PLA_General.c:67).
//
// Factorization.  A (FP64) is copied to an FP32 working matrix W.  The
// blocked right-looking LU of PLA_LU.c:37-212 runs on W with one step of
// look-ahead; per step:
//   panel   W[k:, k:k+b] -> FP64, lu_panel (the FP64 leaf kernels, partial
//           pivoting identical to PLA_LU), back to W, and a low-precision
//           row-major copy of L21;
//   rows    the panel's interchanges on the rest of W (panel_swap, FP32);
//   U12     W12 -> FP64, inv(L11) W12 by the FP64 recursive trsm, back to W,
//           and a low-precision column-major copy of U12;
//   update  W22 -= L21 U12 by hgemm_kernel: FP16 or BF16 MFMA
//           (v_mfma_f32_16x16x32_{f16,bf16}) with FP32 accumulation;
//           FP16x3 splits both operands into hi + lo halves and adds
//           hi*hi + hi*lo + lo*hi (about 22 significant bits).
// Refinement.  x0 = U\L\(P b) with the FP32 factors in FP64 arithmetic; then
// r = b - A x in FP64 (chunked sums combined with compensated addition),
// eta = ||r||_inf / (||A||_inf ||x||_inf + ||b||_inf), d = U\L\(P r), x += d,
// until eta <= tol or max_iter.
//
// Triangular solves (trsv_sf_kernel, also used by PLA_Trsv / the solves of
// the plain path for up to 16 right-hand sides).  One workgroup per 64-row
// block in solve order (untransposed solves: per 64-row block and group of
// at most 4 columns, each group its own chain).  Each workgroup accumulates
// the contributions of the earlier blocks as their solutions appear, then
// multiplies by the inverted diagonal block.  Workgroups publish their
// solution blocks as tagged 64-bit words (tag | 32-bit half), so readers
// need no flags or fences.  A workgroup waits only on workgroups with
// smaller ids, which were dispatched before it.
#include "pla_internal.hpp"

#include <climits>
#include <utility>
#include <cstdint>
#include <vector>

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bhalf8 __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));

template <typename H>
struct hvec;
template <>
struct hvec<_Float16> {
  typedef half8 type;
};
template <>
struct hvec<__bf16> {
  typedef bhalf8 type;
};

__device__ inline f4v mfma16(half8 a, half8 b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}
__device__ inline f4v mfma16(bhalf8 a, bhalf8 b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// C (m x n FP32, column-major) -= A (m x k) * B (k x n), A and B in FP16/BF16
// in the packed k-slice layout below.  128 x 128 block tile, 4 waves of
// 64 x 64 as 4 x 4 MFMA 16x16x32 tiles, K staged 32 at a time through
// double-buffered LDS.  MFMA 16x16x32 layout (lane l): A row l%16,
// k = 8*(l/16)..+7; B column l%16, same k; result lane l holds
// D[4*(l/16) + r][l%16].
//
// Packed k-slice layout of an operand with `index` rows of A (columns of B):
// k-slice t (k = 32t .. 32t+31) is a block of slice_stride elements; inside
// it, index r owns 32 consecutive elements (64 bytes, four 16-byte chunks),
// chunk c (k = 32t + 8c ..) stored at chunk position c ^ ((r >> 1) & 3).  A
// 128-index x 32-k tile is then one contiguous 8 KB block: the staging loads
// are fully used 128-byte lines, and its image in LDS, chunk for chunk, is
// the swizzled tile whose fragment reads (16 lanes: 16 indices, one chunk)
// and staging writes are free of bank conflicts (measured with
// SQ_LDS_BANK_CONFLICT; rows padded to 40 elements cost 4 conflict cycles per
// read).
constexpr int HBM = 128, HBN = 128, HKS = 32;

// Offset of element (index r, depth kk) in the packed k-slice layout.
__device__ __forceinline__ long kslice_off(long r, int kk, long slice_stride) {
  const int c = (kk >> 3) & 3;
  return (long)(kk >> 5) * slice_stride + r * 32 + ((c ^ (int)((r >> 1) & 3)) << 3) + (kk & 7);
}

// A: rows 0..m-1 of a packed operand (slice stride sa); B: columns
// boff..boff+n-1 of a packed operand (slice stride sb, boff a multiple of 8).
template <typename H, bool SPLIT>
__global__ __launch_bounds__(256, 2) void hgemm_kernel(int m, int n, int k, const H *__restrict__ Ah,
                                                       const H *__restrict__ Al, long sa,
                                                       const H *__restrict__ Bh, const H *__restrict__ Bl,
                                                       long sb, long boff, float *__restrict__ C, long ldc,
                                                       double *__restrict__ P, long ldp, int tiles_m,
                                                       int tiles_n, int vec_ok) {
  typedef typename hvec<H>::type V;
  constexpr int NS = SPLIT ? 2 : 1;
  __shared__ V As[2][NS][HBM * 4];
  __shared__ V Bs[2][NS][HBN * 4];

  const int ntiles = tiles_m * tiles_n;
  const int nblk = gridDim.x;  // a multiple of 8
  const int bid = (blockIdx.x & 7) * (nblk >> 3) + (blockIdx.x >> 3);
  if (bid >= ntiles) return;
  constexpr int GROUP = 8;
  const int group = bid / (GROUP * tiles_n);
  const int first_m = group * GROUP;
  const int gsz = min(tiles_m - first_m, GROUP);
  const int within = bid - group * GROUP * tiles_n;
  const int i0 = (first_m + within % gsz) * HBM, j0 = (within / gsz) * HBN;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int l16 = lane & 15, lq = lane >> 4;

  f4v acc[4][4];
#pragma unroll
  for (int bi = 0; bi < 4; bi++)
#pragma unroll
    for (int bj = 0; bj < 4; bj++) acc[bi][bj] = f4v{0.f, 0.f, 0.f, 0.f};

  // Global loads run one K step ahead of the MFMAs (staged through registers).
  // Thread tid moves 16-byte chunks tid and tid + 256 of each 8 KB tile; the
  // chunk's LDS slot is its position in the tile.
  V ra[NS][2], rb[NS][2];
  auto load_tile = [&](int kt) {
#pragma unroll
    for (int e = 0; e < 2; e++) {
      const int c = tid + 256 * e, row = c >> 2;
      // Rows (columns) past the edge read the last one instead: they feed
      // only rows (columns) of the product that are never stored, and the
      // loads carry no condition, so they stay in flight during the MFMAs
      // (a guarded or selected load is waited for where it is issued).
      const int gi = min(i0 + row, m - 1), gj = min(j0 + row, n - 1);
      const long ao = (long)kt * sa + (long)gi * 32 + (c & 3) * 8;
      const long bo = (long)kt * sb + (boff + gj) * 32 + (c & 3) * 8;
#pragma unroll
      for (int s = 0; s < NS; s++) {
        const H *Ap = s ? Al : Ah;
        const H *Bp = s ? Bl : Bh;
        ra[s][e] = *(const V *)(Ap + ao);
        rb[s][e] = *(const V *)(Bp + bo);
      }
    }
  };
  // fragment rows 16 bi + l16 (+ 64 wm / wn): chunk lq of the row sits at
  // lq ^ ((l16 >> 1) & 3) (tile rows start at multiples of 8)
  const int rchunk = l16 * 4 + (lq ^ ((l16 >> 1) & 3));
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int e = 0; e < 2; e++) {
#pragma unroll
      for (int s = 0; s < NS; s++) {
        As[buf][s][tid + 256 * e] = ra[s][e];
        Bs[buf][s][tid + 256 * e] = rb[s][e];
      }
    }
  };

  const int nk = k / HKS;
  if (nk > 0) {
    load_tile(0);
    store_tile(0);
  }
  __syncthreads();
  for (int kt = 0; kt < nk; kt++) {
    const int cur = kt & 1;
    if (kt + 1 < nk) load_tile(kt + 1);
    V af[NS][4], bf[NS][4];
#pragma unroll
    for (int s = 0; s < NS; s++) {
#pragma unroll
      for (int bi = 0; bi < 4; bi++) af[s][bi] = As[cur][s][(wm * 64 + bi * 16) * 4 + rchunk];
#pragma unroll
      for (int bj = 0; bj < 4; bj++) bf[s][bj] = Bs[cur][s][(wn * 64 + bj * 16) * 4 + rchunk];
    }
#pragma unroll
    for (int bi = 0; bi < 4; bi++)
#pragma unroll
      for (int bj = 0; bj < 4; bj++) {
        acc[bi][bj] = mfma16(af[0][bi], bf[0][bj], acc[bi][bj]);
        if (SPLIT) {
          acc[bi][bj] = mfma16(af[0][bi], bf[NS - 1][bj], acc[bi][bj]);
          acc[bi][bj] = mfma16(af[NS - 1][bi], bf[0][bj], acc[bi][bj]);
        }
      }
    if (kt + 1 < nk) store_tile(cur ^ 1);
    __syncthreads();
  }

  if (vec_ok && i0 + HBM <= m && j0 + HBN <= n) {
    // a whole tile: all 16 loads of C in flight together
    f4v cv[4][4];
#pragma unroll
    for (int bi = 0; bi < 4; bi++)
#pragma unroll
      for (int bj = 0; bj < 4; bj++)
        cv[bi][bj] = *(const f4v *)(C + (long)(j0 + wn * 64 + bj * 16 + l16) * ldc + i0 + wm * 64 + bi * 16 + lq * 4);
#pragma unroll
    for (int bi = 0; bi < 4; bi++)
#pragma unroll
      for (int bj = 0; bj < 4; bj++) {
        cv[bi][bj] -= acc[bi][bj];
        *(f4v *)(C + (long)(j0 + wn * 64 + bj * 16 + l16) * ldc + i0 + wm * 64 + bi * 16 + lq * 4) = cv[bi][bj];
      }
    if (P) {  // the FP64 copy of the updated block (exact)
#pragma unroll
      for (int bi = 0; bi < 4; bi++)
#pragma unroll
        for (int bj = 0; bj < 4; bj++) {
          double *pp = P + (long)(j0 + wn * 64 + bj * 16 + l16) * ldp + i0 + wm * 64 + bi * 16 + lq * 4;
#pragma unroll
          for (int r = 0; r < 4; r++) pp[r] = (double)cv[bi][bj][r];
        }
    }
    return;
  }
#pragma unroll
  for (int bi = 0; bi < 4; bi++)
#pragma unroll
    for (int bj = 0; bj < 4; bj++) {
      const int row0 = i0 + wm * 64 + bi * 16 + lq * 4;
      const int col = j0 + wn * 64 + bj * 16 + l16;
      if (col >= n) continue;
      float *cp = C + (long)col * ldc + row0;
      if (vec_ok && row0 + 3 < m) {
        f4v c4 = *(f4v *)cp;
        c4 -= acc[bi][bj];
        *(f4v *)cp = c4;
        if (P) {
#pragma unroll
          for (int r = 0; r < 4; r++) P[(long)col * ldp + row0 + r] = (double)c4[r];
        }
      } else {
#pragma unroll
        for (int r = 0; r < 4; r++)
          if (row0 + r < m) {
            const float v = cp[r] - acc[bi][bj][r];
            cp[r] = v;
            if (P) P[(long)col * ldp + row0 + r] = (double)v;
          }
      }
    }
}

// ---------------------------------------------------------------------------
// Precision conversions of m x n column-major blocks: grid y walks the
// columns, grid x the rows, four consecutive rows per thread (coalesced, no
// index division).
template <typename S, typename D>
__global__ __launch_bounds__(256) void cvt2d_kernel(int m, int n, const S *__restrict__ A, long lda,
                                                    D *__restrict__ B, long ldb) {
  for (int j = blockIdx.y; j < n; j += gridDim.y) {
    const S *a = A + (long)j * lda;
    D *b = B + (long)j * ldb;
    for (int i0 = (blockIdx.x * 256 + threadIdx.x) * 4; i0 < m; i0 += gridDim.x * 1024) {
      // unguarded loads (clamped to the column's last row): all four in flight
      S x[4];
#pragma unroll
      for (int r = 0; r < 4; r++) x[r] = a[min(i0 + r, m - 1)];
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (i0 + r < m) b[i0 + r] = (D)x[r];
    }
  }
}

template <typename S, typename D>
static void launch_cvt2d(hipStream_t s, int m, int n, const S *A, long lda, D *B, long ldb) {
  const int gx = (m + 1023) / 1024 < 64 ? (m + 1023) / 1024 : 64;
  const int gy = n < 65535 ? n : 65535;
  hipLaunchKernelGGL((cvt2d_kernel<S, D>), dim3(gx, gy), dim3(256), 0, s, m, n, A, lda, B, ldb);
}

template <typename H, bool SPLIT>
__device__ inline void to_low(float v, H *hi, H *lo, long idx) {
  const H h = (H)v;
  hi[idx] = h;
  if (SPLIT) lo[idx] = (H)(v - (float)h);
}

// Rows of a factored panel below its diagonal block (P: FP64, rows x b, ldp),
// rounded through FP32 as stored in W, as a packed k-slice copy (index = row,
// depth = column; slice stride sl; depths b .. kpad-1 zero).  Consecutive
// threads write consecutive packed elements.
template <typename H, bool SPLIT>
__global__ __launch_bounds__(256) void panel_out_kernel(int rows, int b, int kpad, const double *__restrict__ P,
                                                        long ldp, H *__restrict__ Lh, H *__restrict__ Ll, long sl) {
  const long per_slice = (long)rows * 32, total = per_slice * (kpad / 32);
#pragma unroll 4
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int t = (int)(e / per_slice);
    const long w = e - (long)t * per_slice;
    const int i = (int)(w >> 5), pos = (int)(w & 31);
    const int c = t * 32 + ((((pos >> 3) ^ ((i >> 1) & 3))) << 3) + (pos & 7);
    const double x = P[(long)min(c, b - 1) * ldp + i];  // unguarded (b >= 1)
    const float v = c < b ? (float)x : 0.f;
    to_low<H, SPLIT>(v, Lh, Ll, (long)t * sl + w);
  }
}

// U12 (FP64, b x nr, ldu) back into W (FP32) and as a packed k-slice copy of
// columns col0 .. col0 + nr - 1 (index = column, depth = row; slice stride
// su; depths b .. kpad-1 zero).
template <typename H, bool SPLIT>
__global__ __launch_bounds__(256) void upanel_out_kernel(int b, int kpad, int nr, const double *__restrict__ U,
                                                         long ldu, float *__restrict__ W, long ldw,
                                                         H *__restrict__ Uh, H *__restrict__ Ul, long su, long col0) {
  const long total = (long)nr * kpad;
#pragma unroll 4
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int j = (int)(e / kpad), r = (int)(e - (long)j * kpad);
    const double x = U[(long)j * ldu + min(r, b - 1)];  // unguarded (b >= 1)
    const float v = r < b ? (float)x : 0.f;
    if (r < b) W[(long)j * ldw + r] = v;
    to_low<H, SPLIT>(v, Uh, Ul, kslice_off(col0 + j, r, su));
  }
}

// ---------------------------------------------------------------------------
// Triangular solves.  Effective matrix M = op(A) (TRANS: A^T); it is lower
// triangular when LOWER != TRANS.  TB-row blocks.
constexpr int TB = 64;
constexpr int MAXRHS = 16;

template <typename T, bool LOWER, bool TRANS>
__device__ inline double op_elem(const T *A, long lda, int r, int c) {
  return TRANS ? (double)A[(long)r * lda + c] : (double)A[(long)c * lda + r];
}

// Inverse of every diagonal block of M, as FP64 column-major TB x TB blocks.
template <typename T, bool LOWER, bool TRANS, bool UNIT>
__global__ __launch_bounds__(TB) void inv_diag_kernel(int n, const T *__restrict__ A, long lda,
                                                      double *__restrict__ Dinv) {
  constexpr bool LOWEFF = (LOWER != TRANS);
  __shared__ double M[TB][TB + 1];
  __shared__ double X[TB][TB + 1];  // X[i][c]: column c of the inverse
  const int blk = blockIdx.x, r0 = blk * TB, bs = min(TB, n - r0);
  const int c = threadIdx.x;
  for (int i = 0; i < TB; i++) {
    double v = 0.0;
    if (i < bs && c < bs) {
      const bool in_tri = LOWEFF ? (c <= i) : (c >= i);
      if (in_tri) v = (UNIT && c == i) ? 1.0 : op_elem<T, LOWER, TRANS>(A, lda, r0 + i, r0 + c);
    }
    M[i][c] = (i < bs && c < bs) ? v : (i == c ? 1.0 : 0.0);
  }
  __syncthreads();
  // column c of the inverse: M x = e_c
  if (LOWEFF) {
    for (int i = 0; i < TB; i++) {
      double t = (i == c) ? 1.0 : 0.0;
      if (i > c)
        for (int q = c; q < i; q++) t = fma(-M[i][q], X[q][c], t);
      X[i][c] = (i < c) ? 0.0 : t / M[i][i];
    }
  } else {
    for (int i = TB - 1; i >= 0; i--) {
      double t = (i == c) ? 1.0 : 0.0;
      if (i < c)
        for (int q = i + 1; q <= c; q++) t = fma(-M[i][q], X[q][c], t);
      X[i][c] = (i > c) ? 0.0 : t / M[i][i];
    }
  }
  __syncthreads();
  double *D = Dinv + (size_t)blk * TB * TB;
  for (int i = 0; i < TB; i++) D[(size_t)c * TB + i] = X[i][c];
}

__device__ inline void st_dev64(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned long long ld_dev64(const unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait until block `w` of this launch has published (its ready word equals
// want): one thread polls the single word, so waiting blocks do not stream
// the solution words through the memory system while the block is being
// computed.  Returns false, uniformly, when the bounded wait gives up.
__device__ inline bool wait_ready(const unsigned long long *__restrict__ w, unsigned long long want,
                                  unsigned spin_limit, int *__restrict__ err, int *s_flag) {
  if (spin_limit == 0) return false;  // a zero poll limit gives up at the first wait (uniform)
  if (threadIdx.x == 0) {
    unsigned it = 0;
    int f = 0;
    while (ld_dev64(w) != want) {
      ++it;
      if (it >= spin_limit || ((it & 63) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
        f = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    *s_flag = f;
  }
  __syncthreads();
  return *s_flag == 0;
}

// After the solution words of block blk are issued by every thread: the
// ready word.  It may become visible before some of the words (no fence:
// a device-scope fence here would write back and invalidate the XCD's whole
// L2 at every block, and waiting for the stores would put their round trip
// on the chain); readers check every word's tag and re-read, so it only
// tells them when to start reading.
__device__ inline void publish_ready(unsigned long long *__restrict__ w, unsigned long long want) {
  __syncthreads();
  if (threadIdx.x == 0) st_dev64(w, want);
}

// One poll of a published solution block: rows 0..cs-1, columns q0..q0+nq-1
// of the words at src into x[c][q] (zero outside).  Every thread issues all
// of its word loads before it looks at any (a loop that tested each pair
// before loading the next waited for one round trip per pair).  Returns
// whether every tag this thread read matched.
template <int NR>
__device__ __forceinline__ int poll_words(const unsigned long long *__restrict__ src, int cs, int q0, int nq,
                                          unsigned long long tag, double (*x)[NR]) {
  constexpr int PW = (TB * NR + 255) / 256;
  const int tid = threadIdx.x;
  if (TB * NR < 256 && tid >= TB * NR) return 1;  // whole waves (TB * NR is a multiple of 64)
  unsigned long long lo[PW], hi[PW];
#pragma unroll
  for (int j = 0; j < PW; j++) {
    const int e = tid + 256 * j, c = e / NR, q = e % NR;
    // outside the block: the first word, always in bounds, read and ignored
    const size_t w = (c < cs && q < nq) ? ((size_t)c * MAXRHS + q0 + q) * 2 : 0;
    lo[j] = ld_dev64(src + w);
    hi[j] = ld_dev64(src + w + 1);
  }
  int ok = 1;
#pragma unroll
  for (int j = 0; j < PW; j++) {
    const int e = tid + 256 * j, c = e / NR, q = e % NR;
    const bool in = c < cs && q < nq;
    if (in && ((lo[j] & 0xFFFFFFFF00000000ULL) != tag || (hi[j] & 0xFFFFFFFF00000000ULL) != tag)) ok = 0;
    const double v = __longlong_as_double((long long)(((hi[j] & 0xFFFFFFFFULL) << 32) | (lo[j] & 0xFFFFFFFFULL)));
    x[c][q] = in ? v : 0.0;
  }
  return ok;
}

// Solve M X = X in place (n x nrhs, nrhs <= NR <= 16).  xs: exchange words
// [nblk][TB][MAXRHS][2], then one ready word per block; tag unique per launch.
// A block's solution words carry the tag; readers re-read them until every
// tag matches.  From NR = 4 on a ready word (tag | 1) also follows the words,
// and a block waiting for any but its immediate predecessor waits for that
// word before reading the words, so the blocks that wait do not stream the
// words through the memory system (at n = 49152: 8 columns 8.0 -> 5.5 ms, 16
// columns 15.3 -> 10.1 ms; with 1 or 2 columns the extra round trip cost
// more than it saved).
//
// Thread t owns row t%64 of its block and the columns c = t/64 (mod 4) of
// every tile; its partial sums stay in registers across tiles and the four
// partials of a row are combined once at the end.  Per step the next tile is
// loaded into registers while the current one is used, and the solution
// block it needs is read together with the tile (one round trip when the
// block is already published).
template <typename T, bool LOWER, bool TRANS, int NR>
__global__ __launch_bounds__(256) void trsv_sf_kernel(int n, int nrhs, const T *__restrict__ A, long lda,
                                                      const double *__restrict__ Dinv, double *__restrict__ X,
                                                      long ldx, unsigned long long *__restrict__ xs, unsigned tag32,
                                                      unsigned spin_limit, int *__restrict__ err) {
  constexpr bool LOWEFF = (LOWER != TRANS);
  __shared__ double tile[TB][TB + 1];
  __shared__ double xb[TB][NR];
  __shared__ double red[4][TB][NR];
  __shared__ int s_flag;
  constexpr bool READY = NR >= 4;  // ready words: measured faster from 4 columns on
  const int nblk = (n + TB - 1) / TB;
  unsigned long long *const rdy = xs + (size_t)nblk * TB * MAXRHS * 2;
  const int b = blockIdx.x;
  const int blk = LOWEFF ? b : nblk - 1 - b;
  const int r0 = blk * TB, bs = min(TB, n - r0);
  const int tid = threadIdx.x;
  const unsigned long long tag = (unsigned long long)tag32 << 32;
  const int row = tid & 63, cq = tid >> 6;

  double part[NR];
#pragma unroll
  for (int q = 0; q < NR; q++) part[q] = 0.0;

  // tile of step s in registers: 16 elements per thread
  T treg[16];
  auto load_tile = [&](int s) {
    const int iblk = LOWEFF ? s : nblk - 1 - s;
    const int c0 = iblk * TB, cs = min(TB, n - c0);
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const int e = tid + 256 * k;
      int i, c;
      if (TRANS) { c = e & 63; i = e >> 6; }
      else { i = e & 63; c = e >> 6; }
      treg[k] = (i < bs && c < cs) ? (TRANS ? A[(long)(r0 + i) * lda + c0 + c] : A[(long)(c0 + c) * lda + r0 + i])
                                   : (T)0;
    }
  };
  bool failed = false;
  if (b > 0) load_tile(0);
  for (int s = 0; s < b; s++) {
    const int iblk = LOWEFF ? s : nblk - 1 - s;
    const int cs = min(TB, n - iblk * TB);
    const unsigned long long *src = xs + (size_t)iblk * TB * MAXRHS * 2;
    __syncthreads();  // tile and xb of the previous step consumed
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const int e = tid + 256 * k;
      int i, c;
      if (TRANS) { c = e & 63; i = e >> 6; }
      else { i = e & 63; c = e >> 6; }
      tile[i][c] = (double)treg[k];
    }
    if (s + 1 < b) load_tile(s + 1);  // next tile in flight during this step
    // the ready word only while the awaited block is not the chain's head:
    // the immediate predecessor's words are polled directly (one block at a
    // time does that), which saves the ready word's round trip on the chain
    if (READY && s + 1 < b && !wait_ready(rdy + iblk, tag | 1, spin_limit, err, &s_flag)) {
      failed = true;
      break;
    }
    unsigned it = 0;
    for (;;) {
      const int ok = poll_words<NR>(src, cs, 0, nrhs, tag, xb);
      // (a zero poll limit gives up at the first wait, data or not)
      if (__syncthreads_and(ok) && spin_limit != 0) break;
      ++it;
      const int bail = (it >= spin_limit) || (tid == 0 && (it & 63) == 0 &&
                                               __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      if (__syncthreads_or(bail)) {
        failed = true;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    if (failed) break;
    for (int c = cq; c < cs; c += 4) {
      const double a = tile[row][c];
#pragma unroll
      for (int q = 0; q < NR; q++) part[q] = fma(-a, xb[c][q], part[q]);
    }
  }
  if (failed && tid == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NR; q++) red[cq][row][q] = part[q];
  __syncthreads();
  // acc = rhs + partials; x_blk = inv(M_bb) acc (xb reused for acc)
  for (int e = tid; e < TB * NR; e += 256) {
    const int i = e / NR, q = e - i * NR;
    double v = 0.0;
    if (i < bs && q < nrhs) v = X[(long)q * ldx + r0 + i] + ((red[0][i][q] + red[1][i][q]) + (red[2][i][q] + red[3][i][q]));
    xb[i][q] = v;
  }
  __syncthreads();
  const double *D = Dinv + (size_t)blk * TB * TB;
  unsigned long long *dst = xs + (size_t)blk * TB * MAXRHS * 2;
  for (int e = tid; e < bs * nrhs; e += 256) {
    const int i = e % bs, q = e / bs;
    double t = 0.0;
    for (int c = 0; c < bs; c++) t = fma(D[(size_t)c * TB + i], xb[c][q], t);
    X[(long)q * ldx + r0 + i] = t;
    const unsigned long long bits = (unsigned long long)__double_as_longlong(t);
    st_dev64(dst + ((size_t)i * MAXRHS + q) * 2, tag | (bits & 0xFFFFFFFFULL));
    st_dev64(dst + ((size_t)i * MAXRHS + q) * 2 + 1, tag | (bits >> 32));
  }
  if (READY) publish_ready(rdy + blk, tag | 1);
}

// Untransposed solves (M = A, column-major): the chain of dependent blocks is
// the cost, so everything a block needs apart from its predecessors' solution
// blocks is fetched when it starts: the inverted diagonal block (thread t
// keeps row t % 64, columns 16 (t / 64) .. +15) and the right-hand side, both
// pinned in registers.  Thread t owns row t % 64 and columns 16 (t / 64) ..
// +15 of every 64 x 64 tile, read straight from A (coalesced: 64 rows of one
// column per wave), one tile ahead of its use.  Per tile only the solution
// block comes through LDS.
template <typename T, bool LOWER, int NR>
__global__ __launch_bounds__(256, (NR == 1 && sizeof(T) == 4) ? 4 : 1) void trsv_nt_kernel(int n, int nrhs, const T *__restrict__ A, long lda,
                                                      const double *__restrict__ Dinv, double *__restrict__ X,
                                                      long ldx, unsigned long long *__restrict__ xs, unsigned tag32,
                                                      unsigned spin_limit, int *__restrict__ err) {
  constexpr int RPT = (TB * NR + 255) / 256;  // right-hand-side entries per thread
  __shared__ double xb[2][TB][NR];
  __shared__ double red[4][TB][NR];
  __shared__ int s_flag;
  constexpr bool READY = NR >= 4;  // ready words: measured faster from 4 columns on
  const int nblk = (n + TB - 1) / TB;
  // Column groups: columns q0 .. q0 + NR - 1 form an independent chain.  In
  // every run of 8 G blocks (G groups) the first 8 are row blocks 8p .. 8p+7
  // of group 0, the next 8 the same row blocks of group 1, and so on: each
  // chain is dispatched in its own order, and the G blocks of one row block
  // land on the same XCD (blocks are spread over the 8 XCDs round-robin), so
  // the factor tiles they share are read through that XCD's L2.
  const int G = (nrhs + NR - 1) / NR;
  int b = blockIdx.x, grp = 0;
  if (G > 1) {
    const int r = (int)(blockIdx.x % (8u * G));
    b = (int)(blockIdx.x / (8u * G)) * 8 + (r & 7);
    grp = r >> 3;
    if (b >= nblk) return;  // padding of the last run (waited on by no block)
  }
  const int q0 = grp * NR, nq = min(NR, nrhs - q0);
  unsigned long long *const rdy = xs + (size_t)nblk * TB * MAXRHS * 2 + (size_t)grp * nblk;
  const int blk = LOWER ? b : nblk - 1 - b;
  const int r0 = blk * TB, bs = min(TB, n - r0);
  const int tid = threadIdx.x, row = tid & 63, cg = tid >> 6;
  const unsigned long long tag = (unsigned long long)tag32 << 32;
  double dv[16];
  {
    const double *D = Dinv + (size_t)blk * TB * TB;
#pragma unroll
    for (int c = 0; c < 16; c++) dv[c] = D[(size_t)(cg * 16 + c) * TB + row];
  }
  double rhs[RPT];
#pragma unroll
  for (int j = 0; j < RPT; j++) {
    const int e = tid + 256 * j, i = e / NR, q = e % NR;
    rhs[j] = (e < TB * NR && i < bs && q < nq) ? X[(long)(q0 + q) * ldx + r0 + i] : 0.0;
  }
  // wait for them here, while the block has nothing else to do
#pragma unroll
  for (int c = 0; c < 16; c++) asm volatile("" : "+v"(dv[c]));
#pragma unroll
  for (int j = 0; j < RPT; j++) asm volatile("" : "+v"(rhs[j]));
  double part[NR];
#pragma unroll
  for (int q = 0; q < NR; q++) part[q] = 0.0;
  T tr[16];
  auto load_tile = [&](int s) {
    const int iblk = LOWER ? s : nblk - 1 - s;
    const int c0 = iblk * TB, cs = min(TB, n - c0);
#pragma unroll
    for (int c = 0; c < 16; c++) {
      const int cc = cg * 16 + c;
      tr[c] = (row < bs && cc < cs) ? A[(long)(c0 + cc) * lda + r0 + row] : (T)0;
    }
  };
  bool failed = false;
  if (b > 0) load_tile(0);
  for (int s = 0; s < b; s++) {
    const int iblk = LOWER ? s : nblk - 1 - s;
    const int cs = min(TB, n - iblk * TB);
    const unsigned long long *src = xs + (size_t)iblk * TB * MAXRHS * 2;
    double (*x)[NR] = xb[s & 1];
    // the ready word only while the awaited block is not the chain's head:
    // the immediate predecessor's words are polled directly (one block at a
    // time does that), which saves the ready word's round trip on the chain
    if (READY && s + 1 < b && !wait_ready(rdy + iblk, tag | 1, spin_limit, err, &s_flag)) {
      failed = true;
      break;
    }
    unsigned it = 0;
    for (;;) {
      const int ok = poll_words<NR>(src, cs, q0, nq, tag, x);
      // (a zero poll limit gives up at the first wait, data or not)
      if (__syncthreads_and(ok) && spin_limit != 0) break;
      ++it;
      const int bail = (it >= spin_limit) || (tid == 0 && (it & 63) == 0 &&
                                               __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      if (__syncthreads_or(bail)) {
        failed = true;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    if (failed) break;
    double a[16];
#pragma unroll
    for (int c = 0; c < 16; c++) a[c] = (double)tr[c];
    if (s + 1 < b) load_tile(s + 1);
#pragma unroll
    for (int c = 0; c < 16; c++)
#pragma unroll
      for (int q = 0; q < NR; q++) part[q] = fma(-a[c], x[cg * 16 + c][q], part[q]);
  }
  if (failed && tid == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
  for (int q = 0; q < NR; q++) red[cg][row][q] = part[q];
  __syncthreads();
  // acc = rhs + partials (into xb[b & 1], not read since the last barrier),
  // then x = inv(M_bb) acc
  double (*acc)[NR] = xb[b & 1];
#pragma unroll
  for (int j = 0; j < RPT; j++) {
    const int e = tid + 256 * j, i = e / NR, q = e % NR;
    if (e < TB * NR) acc[i][q] = rhs[j] + ((red[0][i][q] + red[1][i][q]) + (red[2][i][q] + red[3][i][q]));
  }
  __syncthreads();
  double y[NR];
#pragma unroll
  for (int q = 0; q < NR; q++) y[q] = 0.0;
#pragma unroll
  for (int c = 0; c < 16; c++)
#pragma unroll
    for (int q = 0; q < NR; q++) y[q] = fma(dv[c], acc[cg * 16 + c][q], y[q]);
  __syncthreads();  // red consumed
#pragma unroll
  for (int q = 0; q < NR; q++) red[cg][row][q] = y[q];
  __syncthreads();
  unsigned long long *dst = xs + (size_t)blk * TB * MAXRHS * 2;
  for (int e = tid; e < bs * nq; e += 256) {
    const int i = e % bs, q = e / bs;
    const double t = (red[0][i][q] + red[1][i][q]) + (red[2][i][q] + red[3][i][q]);
    X[(long)(q0 + q) * ldx + r0 + i] = t;
    const unsigned long long bits = (unsigned long long)__double_as_longlong(t);
    st_dev64(dst + ((size_t)i * MAXRHS + q0 + q) * 2, tag | (bits & 0xFFFFFFFFULL));
    st_dev64(dst + ((size_t)i * MAXRHS + q0 + q) * 2 + 1, tag | (bits >> 32));
  }
  if (READY) publish_ready(rdy + blk, tag | 1);
}

// ---------------------------------------------------------------------------
// Residual R = B - A X (FP64, nrhs <= NR <= 16): per 512-column chunk a plain
// FMA sum, then the chunks combined in order with compensated (Neumaier)
// addition; max |R| per right-hand side into rmax (as ordered bits).
constexpr int RCH = 512;

template <int NR>
__global__ __launch_bounds__(256) void resid_partial_kernel(int n, int nrhs, const double *__restrict__ A,
                                                            long lda, const double *__restrict__ X, long ldx,
                                                            double *__restrict__ part) {
  __shared__ double xs[RCH][NR];
  const int ch = blockIdx.y, j0 = ch * RCH, j1 = min(n, j0 + RCH);
  for (int e = threadIdx.x; e < RCH * NR; e += 256) {
    const int j = e / NR, q = e - j * NR;
    xs[j][q] = (j0 + j < j1 && q < nrhs) ? X[(long)q * ldx + j0 + j] : 0.0;
  }
  __syncthreads();
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double s[NR];
#pragma unroll
  for (int q = 0; q < NR; q++) s[q] = 0.0;
  const double *ap = A + (long)j0 * lda + i;
  int j = j0;
  for (; j + 4 <= j1; j += 4, ap += 4 * lda) {
    const double a0 = ap[0], a1 = ap[lda], a2 = ap[2 * lda], a3 = ap[3 * lda];
#pragma unroll
    for (int q = 0; q < NR; q++) {
      s[q] = fma(a0, xs[j - j0][q], s[q]);
      s[q] = fma(a1, xs[j - j0 + 1][q], s[q]);
      s[q] = fma(a2, xs[j - j0 + 2][q], s[q]);
      s[q] = fma(a3, xs[j - j0 + 3][q], s[q]);
    }
  }
  for (; j < j1; j++, ap += lda) {
    const double a0 = ap[0];
#pragma unroll
    for (int q = 0; q < NR; q++) s[q] = fma(a0, xs[j - j0][q], s[q]);
  }
  for (int q = 0; q < nrhs; q++) part[((size_t)ch * nrhs + q) * n + i] = s[q];
}

__global__ __launch_bounds__(256) void resid_final_kernel(int n, int nrhs, int nch, const double *__restrict__ part,
                                                          const double *__restrict__ B, long ldb,
                                                          double *__restrict__ R, long ldr,
                                                          unsigned long long *__restrict__ rmax) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int q = blockIdx.y;
  if (i >= n) return;
  double s = 0.0, comp = 0.0;
  for (int c = 0; c < nch; c++) {
    const double v = part[((size_t)c * nrhs + q) * n + i];
    const double t = s + v;
    comp += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
    s = t;
  }
  const double bq = B ? B[(long)q * ldb + i] : 0.0;
  const double r = (bq - s) - comp;
  R[(long)q * ldr + i] = r;
  atomicMax(rmax + q, (unsigned long long)__double_as_longlong(fabs(r)));
}

// max |X(:, q)| per column into out (as ordered bits)
__global__ __launch_bounds__(256) void col_amax_kernel(int n, const double *__restrict__ X, long ldx,
                                                       unsigned long long *__restrict__ out) {
  __shared__ double red[256];
  const int q = blockIdx.y;
  double m = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) m = fmax(m, fabs(X[(long)q * ldx + i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out + q, (unsigned long long)__double_as_longlong(red[0]));
}

// Y(i, q) = X(perm[i], q)
__global__ __launch_bounds__(256) void permute_rows_kernel(int n, int nrhs, const int *__restrict__ perm,
                                                           const double *__restrict__ X, long ldx,
                                                           double *__restrict__ Y, long ldy) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int src = perm[i];
  for (int q = 0; q < nrhs; q++) Y[(long)q * ldy + i] = X[(long)q * ldx + src];
}

// X += D
__global__ __launch_bounds__(256) void axpy_cols_kernel(int n, int nrhs, const double *__restrict__ D, long ldd,
                                                        double *__restrict__ X, long ldx) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  for (int q = 0; q < nrhs; q++) X[(long)q * ldx + i] += D[(long)q * ldd + i];
}

inline unsigned grid1(long total) {
  long g = (total + 1023) / 1024;
  if (g < 1) g = 1;
  if (g > 1L << 20) g = 1L << 20;
  return (unsigned)g;
}

}  // namespace

namespace pla {

// ---- launchers ---------------------------------------------------------------

int k_hgemm(hipStream_t s, int prec, int m, int n, int k, const void *Ah, const void *Al, long sa, const void *Bh,
            const void *Bl, long sb, long boff, float *C, long ldc, double *P, long ldp) {
  if (m <= 0 || n <= 0 || k <= 0) return PLA_SUCCESS;
  if (k % HKS || sa % 256 || sb % 256 || boff % 8)
    return fail(PLA_ERR_INVALID_PARAM, "hgemm: k %% 32, slice strides %% 256, column offset %% 8");
  if (sa < (long)m * 32 || sb < (boff + n) * 32) return fail(PLA_ERR_INVALID_PARAM, "hgemm: slice too small");
  if (P && ldp < m) return fail(PLA_ERR_INVALID_PARAM, "hgemm: FP64 copy leading dimension %ld < %d", ldp, m);
  const int tm = (m + HBM - 1) / HBM, tn = (n + HBN - 1) / HBN;
  const int nblk = ((tm * tn + 7) / 8) * 8;
  const int vec_ok = (ldc % 4 == 0) && (((uintptr_t)C & 15) == 0);
  dim3 grid(nblk), blk(256);
#define PLA_HG(H, SP)                                                                                             \
  hipLaunchKernelGGL((hgemm_kernel<H, SP>), grid, blk, 0, s, m, n, k, (const H *)Ah, (const H *)Al, sa,           \
                     (const H *)Bh, (const H *)Bl, sb, boff, C, ldc, P, ldp, tm, tn, vec_ok)
  switch (prec) {
    case PLA_MIXED_FP16: PLA_HG(_Float16, false); break;
    case PLA_MIXED_FP16X3: PLA_HG(_Float16, true); break;
    case PLA_MIXED_BF16: PLA_HG(__bf16, false); break;
    case PLA_MIXED_BF16X3: PLA_HG(__bf16, true); break;
    default: return fail(PLA_ERR_INVALID_PARAM, "hgemm: precision %d", prec);
  }
#undef PLA_HG
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_cvt_f64_f32(hipStream_t s, int m, int n, const double *A, long lda, float *W, long ldw) {
  if (m <= 0 || n <= 0) return PLA_SUCCESS;
  launch_cvt2d(s, m, n, A, lda, W, ldw);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_cvt_f32_f64(hipStream_t s, int m, int n, const float *W, long ldw, double *P, long ldp) {
  if (m <= 0 || n <= 0) return PLA_SUCCESS;
  launch_cvt2d(s, m, n, W, ldw, P, ldp);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

static bool is_bf16(int prec) { return prec == PLA_MIXED_BF16 || prec == PLA_MIXED_BF16X3; }
static bool is_split(int prec) { return prec == PLA_MIXED_FP16X3 || prec == PLA_MIXED_BF16X3; }

int k_panel_out(hipStream_t s, int prec, int rows, int b, int kpad, const double *P, long ldp, void *Lh, void *Ll,
                long sl) {
  if (rows <= 0) return PLA_SUCCESS;
  if (b < 1 || kpad % 32 || kpad < b || sl < (long)rows * 32) return fail(PLA_ERR_INVALID_PARAM, "panel_out: layout");
  dim3 grid(grid1((long)rows * kpad)), blk(256);
#define PLA_PO(H, SP) \
  hipLaunchKernelGGL((panel_out_kernel<H, SP>), grid, blk, 0, s, rows, b, kpad, P, ldp, (H *)Lh, (H *)Ll, sl)
  if (is_bf16(prec)) {
    if (is_split(prec)) PLA_PO(__bf16, true); else PLA_PO(__bf16, false);
  } else {
    if (is_split(prec)) PLA_PO(_Float16, true); else PLA_PO(_Float16, false);
  }
#undef PLA_PO
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_upanel_out(hipStream_t s, int prec, int b, int kpad, int nr, const double *U, long ldu, float *W, long ldw,
                 void *Uh, void *Ul, long su, long col0) {
  if (nr <= 0) return PLA_SUCCESS;
  if (b < 1 || kpad % 32 || kpad < b || su < (col0 + nr) * 32) return fail(PLA_ERR_INVALID_PARAM, "upanel_out: layout");
  dim3 grid(grid1((long)nr * kpad)), blk(256);
#define PLA_UO(H, SP)                                                                                           \
  hipLaunchKernelGGL((upanel_out_kernel<H, SP>), grid, blk, 0, s, b, kpad, nr, U, ldu, W, ldw, (H *)Uh, (H *)Ul, \
                     su, col0)
  if (is_bf16(prec)) {
    if (is_split(prec)) PLA_UO(__bf16, true); else PLA_UO(__bf16, false);
  } else {
    if (is_split(prec)) PLA_UO(_Float16, true); else PLA_UO(_Float16, false);
  }
#undef PLA_UO
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

// ---- sync-free triangular solve --------------------------------------------

static unsigned g_trsv_tag = 1;

template <typename T, int NR>
static int trsv_launch(hipStream_t s, bool lower, bool trans, int n, int nrhs, const T *A, long lda, double *X,
                       long ldx, const double *Dinv, unsigned long long *xs, unsigned tag, int *err) {
  const int nblk = (n + TB - 1) / TB;
  // untransposed: ceil(nrhs / NR) column groups, each its own chain, in runs
  // of 8 blocks per group (the kernel's block order)
  const int G = (nrhs + NR - 1) / NR;
  if (trans && G > 1) return fail(PLA_ERR_INVALID_PARAM, "trsv: column groups in a transposed solve");
  const unsigned grid_nt = G > 1 ? 8u * (unsigned)G * (unsigned)((nblk + 7) / 8) : (unsigned)nblk;
#define PLA_TSF(LO, TR)                                                                                      \
  hipLaunchKernelGGL((trsv_sf_kernel<T, LO, TR, NR>), dim3(nblk), dim3(256), 0, s, n, nrhs, A, lda, Dinv, X, \
                     ldx, xs, tag, ctx().trsv_spin_limit, err)
#define PLA_TNT(LO)                                                                                          \
  hipLaunchKernelGGL((trsv_nt_kernel<T, LO, NR>), dim3(grid_nt), dim3(256), 0, s, n, nrhs, A, lda, Dinv, X, \
                     ldx, xs, tag, ctx().trsv_spin_limit, err)
  if (lower && !trans) PLA_TNT(true);
  else if (lower && trans) PLA_TSF(true, true);
  else if (!lower && !trans) PLA_TNT(false);
  else PLA_TSF(false, true);
#undef PLA_TSF
#undef PLA_TNT
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

template <typename T>
static int trsv_sf_impl(hipStream_t s, bool lower, bool trans, bool unit, int n, int nrhs, const T *A, long lda,
                        double *X, long ldx, double *Dinv) {
  if (n <= 0 || nrhs <= 0) return PLA_SUCCESS;
  if (nrhs > MAXRHS) return fail(PLA_ERR_INVALID_PARAM, "trsv: %d right-hand sides (max %d)", nrhs, MAXRHS);
  const int nblk = (n + TB - 1) / TB;
  // solution words, then one ready word per block and column group
  const size_t xs_words = (size_t)nblk * TB * MAXRHS * 2 + (size_t)nblk * MAXRHS;
  unsigned long long *xs = (unsigned long long *)workspace(sizeof(unsigned long long) * xs_words, 5, true);
  if (!xs) return PLA_ERR_NO_MEMORY;
  int *err = leaf_error_word();
  if (!Dinv) {
    Dinv = (double *)workspace(sizeof(double) * (size_t)nblk * TB * TB, 6);
    if (!Dinv) return PLA_ERR_NO_MEMORY;
#define PLA_INV(LO, TR, UN) \
  hipLaunchKernelGGL((inv_diag_kernel<T, LO, TR, UN>), dim3(nblk), dim3(TB), 0, s, n, A, lda, Dinv)
    if (lower && !trans && unit) PLA_INV(true, false, true);
    else if (lower && !trans && !unit) PLA_INV(true, false, false);
    else if (lower && trans && unit) PLA_INV(true, true, true);
    else if (lower && trans && !unit) PLA_INV(true, true, false);
    else if (!lower && !trans && unit) PLA_INV(false, false, true);
    else if (!lower && !trans && !unit) PLA_INV(false, false, false);
    else if (!lower && trans && unit) PLA_INV(false, true, true);
    else PLA_INV(false, true, false);
#undef PLA_INV
    PLA_HIP(hipGetLastError());
  }
  if (g_trsv_tag > 0xF0000000u) {
    PLA_HIP(hipMemsetAsync(xs, 0, sizeof(unsigned long long) * xs_words, s));
    g_trsv_tag = 1;
  }
  const unsigned tag = g_trsv_tag++;
  int rc;
  timing_begin(s, 6);
  // Untransposed solves split the columns into groups of at most
  // trsv_group_cols, each group an independent chain of blocks: a chain's
  // link costs more the more columns it carries, the factor is read once per
  // group.
  const int gc = ctx().trsv_group_cols;
  const int width = (!trans && gc > 0 && nrhs > gc) ? gc : nrhs;
  if (width <= 1) rc = trsv_launch<T, 1>(s, lower, trans, n, nrhs, A, lda, X, ldx, Dinv, xs, tag, err);
  else if (width <= 2) rc = trsv_launch<T, 2>(s, lower, trans, n, nrhs, A, lda, X, ldx, Dinv, xs, tag, err);
  else if (width <= 4) rc = trsv_launch<T, 4>(s, lower, trans, n, nrhs, A, lda, X, ldx, Dinv, xs, tag, err);
  else if (width <= 8) rc = trsv_launch<T, 8>(s, lower, trans, n, nrhs, A, lda, X, ldx, Dinv, xs, tag, err);
  else rc = trsv_launch<T, MAXRHS>(s, lower, trans, n, nrhs, A, lda, X, ldx, Dinv, xs, tag, err);
  // algorithmic bytes: the triangle once, the inverted diagonal blocks once, X read and written
  timing_end(s, 6, sizeof(T) * 0.5 * n * (n + 1.0) + 8.0 * TB * (double)n + 16.0 * nrhs * (double)n);
  return rc;
}

int trsv_sf(hipStream_t s, bool lower, bool trans, bool unit, int n, int nrhs, const double *A, long lda, double *X,
            long ldx, double *Dinv) {
  return trsv_sf_impl<double>(s, lower, trans, unit, n, nrhs, A, lda, X, ldx, Dinv);
}
int trsv_sf_f32(hipStream_t s, bool lower, bool trans, bool unit, int n, int nrhs, const float *A, long lda,
                double *X, long ldx, double *Dinv) {
  return trsv_sf_impl<float>(s, lower, trans, unit, n, nrhs, A, lda, X, ldx, Dinv);
}

int k_inv_diag_f32(hipStream_t s, bool lower, bool unit, int n, const float *A, long lda, double *Dinv) {
  const int nblk = (n + TB - 1) / TB;
  if (n <= 0) return PLA_SUCCESS;
  if (lower && unit)
    hipLaunchKernelGGL((inv_diag_kernel<float, true, false, true>), dim3(nblk), dim3(TB), 0, s, n, A, lda, Dinv);
  else if (lower)
    hipLaunchKernelGGL((inv_diag_kernel<float, true, false, false>), dim3(nblk), dim3(TB), 0, s, n, A, lda, Dinv);
  else if (unit)
    hipLaunchKernelGGL((inv_diag_kernel<float, false, false, true>), dim3(nblk), dim3(TB), 0, s, n, A, lda, Dinv);
  else
    hipLaunchKernelGGL((inv_diag_kernel<float, false, false, false>), dim3(nblk), dim3(TB), 0, s, n, A, lda, Dinv);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

// inverse diagonal blocks of op(A) for trsv_sf_f32 with the same (lower, trans, unit)
int k_inv_diag_f32_op(hipStream_t s, bool lower, bool trans, bool unit, int n, const float *A, long lda,
                      double *Dinv) {
  const int nblk = (n + TB - 1) / TB;
  if (n <= 0) return PLA_SUCCESS;
#define PLA_INV(LO, TR, UN) \
  hipLaunchKernelGGL((inv_diag_kernel<float, LO, TR, UN>), dim3(nblk), dim3(TB), 0, s, n, A, lda, Dinv)
  if (lower && !trans && unit) PLA_INV(true, false, true);
  else if (lower && !trans && !unit) PLA_INV(true, false, false);
  else if (lower && trans && unit) PLA_INV(true, true, true);
  else if (lower && trans && !unit) PLA_INV(true, true, false);
  else if (!lower && !trans && unit) PLA_INV(false, false, true);
  else if (!lower && !trans && !unit) PLA_INV(false, false, false);
  else if (!lower && trans && unit) PLA_INV(false, true, true);
  else PLA_INV(false, true, false);
#undef PLA_INV
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int trsv_block_size() { return TB; }
int trsv_max_rhs() { return MAXRHS; }

// ---- residual and vector norms -----------------------------------------------

// R = B - A X (B == nullptr: R = -A X); rmax[q] = max |R(:, q)| (double),
// synchronous; rmax == nullptr: asynchronous, no norms.
int residual(hipStream_t s, int n, int nrhs, const double *A, long lda, const double *X, long ldx,
             const double *B, long ldb, double *R, long ldr, double *rmax) {
  if (nrhs > MAXRHS) return fail(PLA_ERR_INVALID_PARAM, "residual: %d right-hand sides", nrhs);
  const int nch = (n + RCH - 1) / RCH;
  double *part = (double *)workspace(sizeof(double) * (size_t)nch * nrhs * n + 256, 7);
  unsigned long long *mx = (unsigned long long *)workspace(sizeof(unsigned long long) * MAXRHS, 8);
  if (!part || !mx) return PLA_ERR_NO_MEMORY;
  PLA_HIP(hipMemsetAsync(mx, 0, sizeof(unsigned long long) * MAXRHS, s));
  const dim3 rg((n + 255) / 256, nch);
  timing_begin(s, 5);
  if (nrhs <= 1)
    hipLaunchKernelGGL(resid_partial_kernel<1>, rg, dim3(256), 0, s, n, nrhs, A, lda, X, ldx, part);
  else if (nrhs <= 4)
    hipLaunchKernelGGL(resid_partial_kernel<4>, rg, dim3(256), 0, s, n, nrhs, A, lda, X, ldx, part);
  else if (nrhs <= 8)
    hipLaunchKernelGGL(resid_partial_kernel<8>, rg, dim3(256), 0, s, n, nrhs, A, lda, X, ldx, part);
  else
    hipLaunchKernelGGL(resid_partial_kernel<MAXRHS>, rg, dim3(256), 0, s, n, nrhs, A, lda, X, ldx, part);
  // algorithmic bytes: A once, X once per column chunk (cached), the partials written once
  timing_end(s, 5, 8.0 * n * (double)n + 8.0 * nrhs * (double)n * (1 + nch));
  hipLaunchKernelGGL(resid_final_kernel, dim3((n + 255) / 256, nrhs), dim3(256), 0, s, n, nrhs, nch, part, B, ldb,
                     R, ldr, mx);
  PLA_HIP(hipGetLastError());
  if (!rmax) return PLA_SUCCESS;
  unsigned long long h[MAXRHS];
  PLA_HIP(hipMemcpyAsync(h, mx, sizeof(unsigned long long) * nrhs, hipMemcpyDeviceToHost, s));
  PLA_HIP(hipStreamSynchronize(s));
  for (int q = 0; q < nrhs; q++) rmax[q] = __builtin_bit_cast(double, h[q]);
  return PLA_SUCCESS;
}

int col_amax(hipStream_t s, int n, int nrhs, const double *X, long ldx, double *out) {
  if (nrhs > MAXRHS) return fail(PLA_ERR_INVALID_PARAM, "col_amax: %d columns", nrhs);
  unsigned long long *mx = (unsigned long long *)workspace(sizeof(unsigned long long) * MAXRHS, 9);
  if (!mx) return PLA_ERR_NO_MEMORY;
  PLA_HIP(hipMemsetAsync(mx, 0, sizeof(unsigned long long) * MAXRHS, s));
  int gx = (n + 255) / 256;
  if (gx > 1024) gx = 1024;
  hipLaunchKernelGGL(col_amax_kernel, dim3(gx, nrhs), dim3(256), 0, s, n, X, ldx, mx);
  PLA_HIP(hipGetLastError());
  unsigned long long h[MAXRHS];
  PLA_HIP(hipMemcpyAsync(h, mx, sizeof(unsigned long long) * nrhs, hipMemcpyDeviceToHost, s));
  PLA_HIP(hipStreamSynchronize(s));
  for (int q = 0; q < nrhs; q++) out[q] = __builtin_bit_cast(double, h[q]);
  return PLA_SUCCESS;
}

int k_permute_rows(hipStream_t s, int n, int nrhs, const int *perm, const double *X, long ldx, double *Y,
                   long ldy) {
  hipLaunchKernelGGL(permute_rows_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, nrhs, perm, X, ldx, Y, ldy);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_axpy_cols(hipStream_t s, int n, int nrhs, const double *D, long ldd, double *X, long ldx) {
  hipLaunchKernelGGL(axpy_cols_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, nrhs, D, ldd, X, ldx);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

}  // namespace pla
