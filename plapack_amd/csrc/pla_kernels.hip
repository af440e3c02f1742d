// pla_kernels.hip -- HIP kernels for gfx950 (AMD Instinct MI355X) behind the
// PLAPACK dense factor-and-solve path.
//
// Kernels and the reference routine each one replaces:
//   dgemm_kernel      PLA_Local_gemm (Local_BLAS/PLA_Local_gemm.c:102) and the
//                     dgemm/dsyrk calls of PLA_Syrk_perform_local_part
//                     (BLAS3/PLA_Syrk_perform_local_part.c:54,133,136):
//                     FP64 MFMA (v_mfma_f64_16x16x4_f64), 128x128 block
//                     tile, LDS-staged 16-deep K panels, double buffered,
//                     next-step loads interleaved with the MFMAs.
//   lu_leaf_kernel    PLA_LU_right_mv (SOLVERS/LU/PLA_LU_mv.c:52-82):
//                     unblocked panel LU of up to 32 columns; PLA_Iamax
//                     (BLAS1/PLA_Iamax.c:37-124) becomes a workgroup
//                     reduction plus a cross-workgroup exchange of
//                     (value, index, row) candidates, Pivot_rows_mv
//                     (PLA_Apply_pivots_new.c:208-305) a row exchange in
//                     LDS, PLA_Local_inv_scal (PLA_Local_BLAS1.c:258) a
//                     reciprocal-multiply, PLA_Local_ger a rank-1 update.
//   laswp_kernel      PLA_Apply_pivots_to_rows (PLA_Apply_pivots_new.c:13-179)
//   trsm_leaf_kernel  PLA_Local_trsm / dtrsm (Local_BLAS/PLA_Local_trsm.c:67),
//                     left side, a wave per group of columns, a lane per row.
//   trsm_right_leaf_kernel  dtrsm("R", ...) of PLA_Local_trsm and of
//                     PLA_Chol.c:193-195 ("R","L","T","N"), one thread per row.
//   potrf_leaf_kernel PLA_dchol (SOLVERS/Cholesky/PLA_Local_chol.c:103-143)
//   fill_drand48      TEST/create_problem.c:31-52 (drand48()*2-1, column-major)
//
// All matrices are column-major double with a 64-bit leading dimension.
#include "pla_internal.hpp"
#include "pla_leaf_fallback.hpp"

#include <climits>
#include <cstdint>

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// DGEMM: C = beta*C + alpha*op(A)*op(B), m x n result, inner dimension k.
// Block tile 128 x 128, 4 waves in a 2 x 2 arrangement, 64 x 64 per wave as a
// 4 x 4 grid of 16 x 16 MFMA tiles.  K is staged through LDS 16 at a time,
// double buffered, one barrier per 16-deep step.
// MFMA f64 16x16x4 operand layout (lane l): A[l%16][l/16], B[l/16][l%16];
// result lane l, register r holds D[l/16 + 4r][l%16].  The kernel feeds the
// B fragment as the A operand and vice versa, so D is the transposed C tile
// and lane l holds C(l%16, l/16 + 4r): 16 consecutive lanes cover 16
// consecutive rows of one column (128-byte coalesced loads and stores of C).
// TRI: 0 the whole of C; 1 (2) only the lower (upper) triangle, row >= col
// (row <= col), of C is read or written and tiles entirely outside it exit at
// once (SYRK).
// ROLE only separates kernel symbols (the trailing update is ROLE 1) so that
// profiles can attribute time to the LU/Cholesky trailing update.
//
// Two tile paths share the LDS layout and the MFMA step:
//  * interior tiles (all 128 x 128 in range) with alpha = +-1 load operand
//    pairs with 16-byte loads from a uniform (SGPR) base plus one 32-bit lane
//    offset, and interleave the next step's loads and LDS stores with the
//    current step's 64 MFMAs (sched_group_barrier), so no wave idles in a
//    separate load/store phase.  alpha = -1 is applied exactly through the
//    sign of the accumulator: acc = -(beta C), acc += op(A) op(B), C = -acc
//    (round-to-nearest is symmetric, so this equals beta C - op(A) op(B)
//    accumulated in the same order);
//  * edge tiles and general alpha use element loads with bounds checks and
//    scale A by alpha when it is staged.
constexpr int GBM = 128, GBN = 128, GBK = 16;
constexpr int APAD = 4;   // As row stride GBM + APAD doubles
constexpr int BPAD = 1;   // Bs row stride GBK + BPAD doubles
typedef double dbl2u __attribute__((ext_vector_type(2), aligned(8)));

struct gemm_smem {
  double As[2][GBK][GBM + APAD];  // [depth][row]
  double Bs[2][GBN][GBK + BPAD];  // [column][depth]
};

// One 16-deep MFMA step of a wave's 64 x 64 block from LDS buffer cur.
__device__ __forceinline__ void gemm_mfma_step(const gemm_smem &sm, int cur, int wm, int wn, int l16, int lq,
                                               dbl4 (&acc)[4][4]) {
#pragma unroll
  for (int kq = 0; kq < 4; kq++) {
    double af[4], bf[4];
#pragma unroll
    for (int bi = 0; bi < 4; bi++) af[bi] = sm.As[cur][kq * 4 + lq][wm * 64 + bi * 16 + l16];
#pragma unroll
    for (int bj = 0; bj < 4; bj++) bf[bj] = sm.Bs[cur][wn * 64 + bj * 16 + l16][kq * 4 + lq];
#pragma unroll
    for (int bi = 0; bi < 4; bi++)
#pragma unroll
      for (int bj = 0; bj < 4; bj++)
        acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[bj], af[bi], acc[bi][bj], 0, 0, 0);
  }
}

template <int TRI>
__device__ __forceinline__ bool tri_in(int gi, int gj) {
  return TRI == 0 || (TRI == 1 ? gi >= gj : gi <= gj);
}

// acc = sgn * beta * C over the tile (0 outside C or outside the TRI triangle)
template <int TRI>
__device__ __forceinline__ void gemm_load_c(int m, int n, double beta, double sgn, const double *__restrict__ C,
                                            long ldc, int i0, int j0, dbl4 (&acc)[4][4]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1, l16 = lane & 15, lq = lane >> 4;
#pragma unroll
  for (int bi = 0; bi < 4; bi++)
#pragma unroll
    for (int bj = 0; bj < 4; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gi = i0 + wm * 64 + bi * 16 + l16;
        const int gj = j0 + wn * 64 + bj * 16 + lq + 4 * r;
        // read at the clamped position (always inside C) and select: an
        // unguarded load does not wait for its data before the next is issued
        const double c = C[(long)min(gj, n - 1) * ldc + min(gi, m - 1)];
        const bool in = beta != 0.0 && gi < m && gj < n && tri_in<TRI>(gi, gj);
        acc[bi][bj][r] = in ? sgn * beta * c : 0.0;
      }
}

template <int TRI>
__device__ __forceinline__ void gemm_store_c(int m, int n, double sgn, double *__restrict__ C, long ldc, int i0,
                                             int j0, const dbl4 (&acc)[4][4]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1, l16 = lane & 15, lq = lane >> 4;
#pragma unroll
  for (int bi = 0; bi < 4; bi++)
#pragma unroll
    for (int bj = 0; bj < 4; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gi = i0 + wm * 64 + bi * 16 + l16;
        const int gj = j0 + wn * 64 + bj * 16 + lq + 4 * r;
        if (gi < m && gj < n && tri_in<TRI>(gi, gj)) C[(long)gj * ldc + gi] = sgn * acc[bi][bj][r];
      }
}

// The same for a tile wholly inside C.  The loads carry no per-element
// condition, so all 64 are in flight together (a guarded load is followed
// by its own wait); elements outside the TRI triangle are read but never
// stored back.
template <int TRI>
__device__ __forceinline__ void gemm_load_c_full(double beta, double sgn, const double *__restrict__ C, long ldc,
                                                 int i0, int j0, dbl4 (&acc)[4][4]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1, l16 = lane & 15, lq = lane >> 4;
  if (beta == 0.0) {
#pragma unroll
    for (int bi = 0; bi < 4; bi++)
#pragma unroll
      for (int bj = 0; bj < 4; bj++) acc[bi][bj] = dbl4{0.0, 0.0, 0.0, 0.0};
    return;
  }
  const double f = sgn * beta;
  const double *cw = C + (long)(j0 + wn * 64 + lq) * ldc + i0 + wm * 64 + l16;
#pragma unroll
  for (int bi = 0; bi < 4; bi++)
#pragma unroll
    for (int bj = 0; bj < 4; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) acc[bi][bj][r] = f * cw[(long)(bj * 16 + 4 * r) * ldc + bi * 16];
}

template <int TRI>
__device__ __forceinline__ void gemm_store_c_full(double sgn, double *__restrict__ C, long ldc, int i0, int j0,
                                                  const dbl4 (&acc)[4][4]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1, l16 = lane & 15, lq = lane >> 4;
  double *cw = C + (long)(j0 + wn * 64 + lq) * ldc + i0 + wm * 64 + l16;
  // a tile of the triangle's interior needs no per-element test
  const bool whole = TRI == 0 || (TRI == 1 ? i0 >= j0 + GBN - 1 : j0 >= i0 + GBM - 1);
  if (whole) {
#pragma unroll
    for (int bi = 0; bi < 4; bi++)
#pragma unroll
      for (int bj = 0; bj < 4; bj++)
#pragma unroll
        for (int r = 0; r < 4; r++) cw[(long)(bj * 16 + 4 * r) * ldc + bi * 16] = sgn * acc[bi][bj][r];
    return;
  }
#pragma unroll
  for (int bi = 0; bi < 4; bi++)
#pragma unroll
    for (int bj = 0; bj < 4; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (tri_in<TRI>(i0 + wm * 64 + bi * 16 + l16, j0 + wn * 64 + bj * 16 + lq + 4 * r))
          cw[(long)(bj * 16 + 4 * r) * ldc + bi * 16] = sgn * acc[bi][bj][r];
}

// Interior tile, alpha = sgn = +-1.  Operand tiles are 128 (index) x 16
// (depth).  An index-contiguous operand (A, or B transposed) is read as index
// pairs (2p, 2p+1), p = tid % 64, at depths tid / 64 + 4e; a depth-contiguous
// one (A transposed, or B) as depth pairs (2q, 2q+1), q = tid % 8, of indices
// tid / 8 + 32e; e = 0..3.
template <bool TA, bool TB, int TRI>
__device__ __forceinline__ void gemm_tile_fast(gemm_smem &sm, int m, int n, int k, double sgn,
                                               const double *__restrict__ A, long lda, const double *__restrict__ B,
                                               long ldb, double beta, double *__restrict__ C, long ldc, int i0,
                                               int j0) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1, l16 = lane & 15, lq = lane >> 4;
  const int ax = TA ? (tid >> 3) : 2 * (tid & 63), ak = TA ? 2 * (tid & 7) : (tid >> 6);
  const int bx = TB ? 2 * (tid & 63) : (tid >> 3), bk = TB ? (tid >> 6) : 2 * (tid & 7);
  const unsigned voa = (unsigned)((TA ? (long)ax * lda + ak : (long)ak * lda + ax) * 8);
  const unsigned vob = (unsigned)((TB ? (long)bk * ldb + bx : (long)bx * ldb + bk) * 8);
  auto pa = [&](int kb, int e) {
    const double *base = TA ? A + (long)(i0 + 32 * e) * lda + kb : A + (long)(kb + 4 * e) * lda + i0;
    return (const double *)((const char *)base + voa);
  };
  auto pb = [&](int kb, int e) {
    const double *base = TB ? B + (long)(kb + 4 * e) * ldb + j0 : B + (long)(j0 + 32 * e) * ldb + kb;
    return (const double *)((const char *)base + vob);
  };
  dbl2u ra[4], rb[4];
  auto load = [&](int kb) {
#pragma unroll
    for (int e = 0; e < 4; e++) ra[e] = *(const dbl2u *)pa(kb, e);
#pragma unroll
    for (int e = 0; e < 4; e++) rb[e] = *(const dbl2u *)pb(kb, e);
  };
  // last, partial 16-deep step: depths >= k read as zero.  The loads read
  // clamped depths (no condition, so they are all in flight together) and
  // the selects come after them.
  auto load_tail = [&](int kb) {
#pragma unroll
    for (int e = 0; e < 4; e++) {
      if (TA) {
        const double *row = A + (long)(i0 + 32 * e + ax) * lda;
        ra[e].x = row[min(kb + ak, k - 1)];
        ra[e].y = row[min(kb + ak + 1, k - 1)];
      } else {
        ra[e] = *(const dbl2u *)(A + (long)min(kb + ak + 4 * e, k - 1) * lda + i0 + ax);
      }
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
      if (TB) {
        rb[e] = *(const dbl2u *)(B + (long)min(kb + bk + 4 * e, k - 1) * ldb + j0 + bx);
      } else {
        const double *col = B + (long)(j0 + 32 * e + bx) * ldb;
        rb[e].x = col[min(kb + bk, k - 1)];
        rb[e].y = col[min(kb + bk + 1, k - 1)];
      }
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int d0 = TA ? kb + ak : kb + ak + 4 * e, d1 = TA ? d0 + 1 : d0;
      ra[e].x = d0 < k ? ra[e].x : 0.0;
      ra[e].y = d1 < k ? ra[e].y : 0.0;
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int d0 = TB ? kb + bk + 4 * e : kb + bk, d1 = TB ? d0 : d0 + 1;
      rb[e].x = d0 < k ? rb[e].x : 0.0;
      rb[e].y = d1 < k ? rb[e].y : 0.0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int e = 0; e < 4; e++) {
      if (!TA) {
        sm.As[buf][ak + 4 * e][ax] = ra[e].x;
        sm.As[buf][ak + 4 * e][ax + 1] = ra[e].y;
      } else {
        sm.As[buf][ak][ax + 32 * e] = ra[e].x;
        sm.As[buf][ak + 1][ax + 32 * e] = ra[e].y;
      }
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
      if (!TB) {
        sm.Bs[buf][bx + 32 * e][bk] = rb[e].x;
        sm.Bs[buf][bx + 32 * e][bk + 1] = rb[e].y;
      } else {
        sm.Bs[buf][bx][bk + 4 * e] = rb[e].x;
        sm.Bs[buf][bx + 1][bk + 4 * e] = rb[e].y;
      }
    }
  };
  const int nkf = k / GBK, nk = (k + GBK - 1) / GBK;
  dbl4 acc[4][4];
  if (nk > 0) {
    if (nkf > 0) load(0);
    else load_tail(0);
  }
  gemm_load_c_full<TRI>(beta, sgn, C, ldc, i0, j0, acc);
  if (nk > 0) store(0);
  __syncthreads();
  int kt = 0;
  for (; kt + 1 < nkf; kt++) {
    const int cur = kt & 1;
    load((kt + 1) * GBK);
    gemm_mfma_step(sm, cur, wm, wn, l16, lq, acc);
    store(cur ^ 1);
    // 64 MFMAs in four groups of 16 (one per 4-deep slice).  The 8 loads of
    // the next step are spread over the first group; each slice's fragment
    // reads are issued two MFMAs before the end of the previous group; the 8
    // LDS stores of the next step (whose loads have landed by then) are spread
    // over the last group.
    __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
#pragma unroll
    for (int i = 0; i < 8; i++) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, 6, 0);
#pragma unroll
    for (int q = 0; q < 3; q++) {
      __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
      if (q < 2) __builtin_amdgcn_sched_group_barrier(0x008, 14, 0);
    }
#pragma unroll
    for (int i = 0; i < 7; i++) {
      __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
    __syncthreads();
  }
  if (nkf > 0 && nk > nkf) {
    load_tail(nkf * GBK);
    gemm_mfma_step(sm, kt & 1, wm, wn, l16, lq, acc);
    store((kt & 1) ^ 1);
    __syncthreads();
    kt++;
  }
  if (nk > 0) gemm_mfma_step(sm, kt & 1, wm, wn, l16, lq, acc);
  gemm_store_c_full<TRI>(sgn, C, ldc, i0, j0, acc);
}

// Any tile, any alpha: element loads with bounds checks, A scaled by alpha.
template <bool TA, bool TB, int TRI>
__device__ __forceinline__ void gemm_tile_checked(gemm_smem &sm, int m, int n, int k, double alpha,
                                                  const double *__restrict__ A, long lda,
                                                  const double *__restrict__ B, long ldb, double beta,
                                                  double *__restrict__ C, long ldc, int i0, int j0) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1, l16 = lane & 15, lq = lane >> 4;
  dbl4 acc[4][4];
  gemm_load_c<TRI>(m, n, beta, 1.0, C, ldc, i0, j0, acc);
  double ra[8], rb[8];
  auto load_tile = [&](int kt) {
    const int kb = kt * GBK;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int idx = tid + 256 * e;
      int i, kk;
      if (!TA) { i = idx & 127; kk = idx >> 7; }
      else { kk = idx & 15; i = idx >> 4; }
      const int gi = i0 + i, gk = kb + kk;
      const int ci = min(gi, m - 1), ck = min(gk, k - 1);
      const double v = TA ? A[(long)ci * lda + ck] : A[(long)ck * lda + ci];
      ra[e] = (gi < m && gk < k) ? v : 0.0;
    }
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int idx = tid + 256 * e;
      int j, kk;
      if (!TB) { kk = idx & 15; j = idx >> 4; }
      else { j = idx & 127; kk = idx >> 7; }
      const int gj = j0 + j, gk = kb + kk;
      const int cj = min(gj, n - 1), ck = min(gk, k - 1);
      const double v = TB ? B[(long)ck * ldb + cj] : B[(long)cj * ldb + ck];
      rb[e] = (gj < n && gk < k) ? v : 0.0;
    }
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int idx = tid + 256 * e;
      int i, kk;
      if (!TA) { i = idx & 127; kk = idx >> 7; }
      else { kk = idx & 15; i = idx >> 4; }
      sm.As[buf][kk][i] = alpha * ra[e];
    }
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int idx = tid + 256 * e;
      int j, kk;
      if (!TB) { kk = idx & 15; j = idx >> 4; }
      else { j = idx & 127; kk = idx >> 7; }
      sm.Bs[buf][j][kk] = rb[e];
    }
  };
  const int nk = (k + GBK - 1) / GBK;
  if (nk > 0) {
    load_tile(0);
    store_tile(0);
  }
  __syncthreads();
  for (int kt = 0; kt < nk; kt++) {
    const int cur = kt & 1;
    if (kt + 1 < nk) load_tile(kt + 1);
    gemm_mfma_step(sm, cur, wm, wn, l16, lq, acc);
    if (kt + 1 < nk) store_tile(cur ^ 1);
    __syncthreads();
  }
  gemm_store_c<TRI>(m, n, 1.0, C, ldc, i0, j0, acc);
}

// Tile b of the lower triangle (diagonal included) of a tiles x tiles grid,
// in bands of `group` tile rows: inside a band the columns left of it (all
// the band's rows), then the band's own triangle column by column.  False
// once b is past the last tile.
__device__ inline bool tri_tile(int b, int tiles, int group, int *tm, int *tn) {
  for (int r0 = 0; r0 < tiles; r0 += group) {
    const int h = min(group, tiles - r0);
    const int full = r0 * h;
    const int cnt = full + h * (h + 1) / 2;
    if (b < cnt) {
      if (b < full) {
        *tn = b / h;
        *tm = r0 + b % h;
        return true;
      }
      b -= full;
      for (int c = 0; c < h; c++) {
        if (b < h - c) {
          *tn = r0 + c;
          *tm = r0 + c + b;
          return true;
        }
        b -= h - c;
      }
    }
    b -= cnt;
  }
  return false;
}

// fast: alpha is +-1 and every operand lane offset fits 32 bits (host check)
template <bool TA, bool TB, int TRI, int ROLE>
__global__ __launch_bounds__(256, 2) void dgemm_kernel(
    int m, int n, int k, double alpha, const double *__restrict__ A, long lda,
    const double *__restrict__ B, long ldb, double beta, double *__restrict__ C, long ldc,
    int tiles_m, int tiles_n, const int *__restrict__ abort_flag, int fast) {
  if (abort_flag && *abort_flag) return;
  // Consecutive workgroup ids are dispatched round-robin over the 8 XCDs;
  // give each XCD a contiguous range of tiles so its L2 sees neighbours.
  const int nblk = gridDim.x;  // a multiple of 8
  int bid = (blockIdx.x & 7) * (nblk >> 3) + (blockIdx.x >> 3);
  constexpr int GROUP = 8;  // tile rows per group: A rows are reused inside a group
  int tm, tn;
  if (TRI == 0) {
    if (bid >= tiles_m * tiles_n) return;
    const int group = bid / (GROUP * tiles_n);
    const int first_m = group * GROUP;
    const int gsz = min(tiles_m - first_m, GROUP);
    const int within = bid - group * GROUP * tiles_n;
    tm = first_m + within % gsz;
    tn = within / gsz;
  } else {
    // only the triangle's tiles are launched (a square tile grid), so every
    // XCD's contiguous range holds the same number of them
    if (!tri_tile(bid, tiles_m, GROUP, &tm, &tn)) return;
    if (TRI == 2) {
      const int t = tm;
      tm = tn;
      tn = t;
    }
  }
  const int i0 = tm * GBM, j0 = tn * GBN;
  __shared__ gemm_smem sm;
  if (fast && i0 + GBM <= m && j0 + GBN <= n)
    gemm_tile_fast<TA, TB, TRI>(sm, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, i0, j0);
  else
    gemm_tile_checked<TA, TB, TRI>(sm, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, i0, j0);
}

// ---------------------------------------------------------------------------
// Tall-skinny DGEMM, n <= 64 and k <= 64: C = beta C + alpha A op(B) with A
// untransposed (the panel recursion's updates, m up to the whole column).
// One workgroup per 64 rows, one wave per 16 of them: alpha op(B) (k x n) is
// staged once in LDS, each lane loads its A entries (16 consecutive rows of
// one column per quarter-wave) for all of K up front, and the wave runs
// ceil(n/16) x ceil(k/4) MFMAs.  The 128 x 128 tiles of dgemm_kernel would
// leave most of their MFMA work and every operand load on the bounds-checked
// edge path for these shapes.
constexpr int TS_N = 64, TS_K = 64;

template <bool TB>
__global__ __launch_bounds__(256) void dgemm_tall_kernel(int m, int n, int k, double alpha,
                                                         const double *__restrict__ A, long lda,
                                                         const double *__restrict__ B, long ldb, double beta,
                                                         double *__restrict__ C, long ldc,
                                                         const int *__restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  __shared__ double Bs[TS_K][TS_N + 1];  // Bs[kk][j] = alpha op(B)(kk, j), zero outside k x n
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, lq = lane >> 4;
  // Every load below reads a clamped position (inside the operand: k, n >= 1)
  // and selects zero outside it.  Unguarded loads are all in flight at once;
  // a guarded load would wait for its data before the next one issued.
#pragma unroll
  for (int e0 = 0; e0 < TS_K * TS_N; e0 += 256) {
    const int e = e0 + tid;
    const int kk = e & 63, j = e >> 6;  // consecutive threads: consecutive depth (untransposed B is depth-contiguous)
    const int ck = min(kk, k - 1), cj = min(j, n - 1);
    const double v = TB ? B[(long)ck * ldb + cj] : B[(long)cj * ldb + ck];
    Bs[kk][j] = (kk < k && j < n) ? alpha * v : 0.0;
  }
  const int row = blockIdx.x * 64 + wave * 16 + l16;
  const bool rok = row < m;
  const int crow = min(row, m - 1);
  const int nkq = (k + 3) >> 2, njt = (n + 15) >> 4;
  // all loads first, with no condition at all (the selects come after, so
  // no load waits for its data before the next one issues)
  double a[TS_K / 4];
#pragma unroll
  for (int q = 0; q < TS_K / 4; q++) a[q] = A[(long)min(4 * q + lq, k - 1) * lda + crow];
  dbl4 acc[TS_N / 16];
#pragma unroll
  for (int t = 0; t < TS_N / 16; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) acc[t][r] = C[(long)min(16 * t + lq + 4 * r, n - 1) * ldc + crow];
#pragma unroll
  for (int q = 0; q < TS_K / 4; q++) a[q] = (rok && 4 * q + lq < k) ? a[q] : 0.0;
#pragma unroll
  for (int t = 0; t < TS_N / 16; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int col = 16 * t + lq + 4 * r;
      acc[t][r] = (beta != 0.0 && rok && col < n) ? beta * acc[t][r] : 0.0;
    }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < TS_K / 4; q++) {
    if (q < nkq) {
#pragma unroll
      for (int t = 0; t < TS_N / 16; t++)
        if (t < njt) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(Bs[4 * q + lq][16 * t + l16], a[q], acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < TS_N / 16; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int col = 16 * t + lq + 4 * r;
      if (rok && col < n) C[(long)col * ldc + row] = acc[t][r];
    }
}

// ---------------------------------------------------------------------------
// Unblocked panel LU of a tall leaf (mp rows, w <= 32 columns) in one
// persistent grid of G workgroups; workgroup g keeps rows [g*R, g*R+R) in LDS.
//
// Rows are never moved while the leaf is factored.  Each row carries its
// logical position pos[r] (initially its row number); interchanging rows j
// and p only relabels them (pivot row -> j, the row at j -> p).  Rows at
// positions < j are the pivot rows already chosen.  Per column j:
//   1. local candidate: largest |L(r, j)| over rows at positions >= j, ties to
//      the smallest position (the first maximum, as idamax and the tie rule
//      of PLA_Iamax.c:108-115);
//   2. wave 0 publishes its candidate row, then (value, position);
//   3. every workgroup reads all G (value, position) words, picks the same
//      winner, reads the winner's row, and applies the relabelling, the
//      reciprocal scaling (PLA_Local_BLAS1.c:258) and the rank-1 update.
// At the end rows are written to their final positions, the interchanges
// are replayed into at most 2w (destination, source) row pairs, and the
// pairs are applied to the panel's other columns (so the recursive panel
// needs no separate row-interchange launches).  Workgroup 0 also stores the
// pairs for the full-width interchange of the panel.
//
// Exchange protocol: every 64-bit slot word carries the 32-bit tag of the
// column it belongs to next to 32 bits of payload (a double is split into two
// words).  Writers use plain device-scope stores, readers device-scope loads
// and re-read until every word they need carries the current tag, so no
// fences are needed.  Two round trips per column: the G small (value,
// position) records, then the one winning row -- reading every candidate
// row instead costs several times more (measured on MI355X: 2.5 us vs
// 8.7 us per exchange at G = 64).  Slots are double buffered by column
// parity: no workgroup can publish column j+2 before every workgroup has
// finished reading column j.  Tags only grow across launches.  A bounded
// spin sets *err and lets every wave exit.
constexpr int LEAF_W = 32;
constexpr int LEAF_THREADS = 256;
constexpr int LEAF_MAXG = 128;
constexpr int CAND_WORDS = 4;            // value lo/hi, position, pad
constexpr int ROW_WORDS = 2 * LEAF_W;    // candidate row, lo/hi halves
constexpr int SLOT_WORDS = CAND_WORDS + ROW_WORDS;
constexpr int PAIRS_MAX = 2 * LEAF_W;
constexpr int PAIRS_STRIDE = 1 + 2 * PAIRS_MAX;  // count, then (dst, src) pairs

struct leaf_args {
  double *P;                // leaf top-left (its row 0 is the leaf's first diagonal row)
  double *Ppanel;           // panel column 0, at the leaf's first row
  long lda;
  int mp, w;                // rows from the leaf's first row to the panel bottom, leaf width
  int pw, pc0;              // panel width, leaf's first column inside the panel
  int R;                    // rows per workgroup
  int *piv;                 // w entries: pivot offset relative to row j
  int *pairs;               // PAIRS_STRIDE ints, positions relative to the panel's first row
  unsigned long long *slots;  // [2][G][CAND_WORDS] then [2][G][ROW_WORDS]
  int *err;
  unsigned long long *state;  // the launch's committed / abandoned word
  char *fscr;                 // workgroup 0's scratch if the launch is abandoned
  unsigned tag0, spin_limit;
};

__device__ inline void argmax3(double &bv, int &bi, int &br, double v, int i, int r) {
  if (v > bv || (v == bv && i < bi)) { bv = v; bi = i; br = r; }
}

__device__ inline void st_dev(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned long long ld_dev(const unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double ld_dev(const double *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int ld_dev(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned dbl_half(double v, int hi) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return hi ? (unsigned)(b >> 32) : (unsigned)b;
}
__device__ inline double dbl_from(unsigned lo, unsigned hi) {
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Bounded wait: every lane counts the same iterations, and the state word is
// read by lane 0, so the decision is wave-uniform.  Returns true when the
// launch is abandoned: this wait gave up first, or another participant's did.
// A wait that gives up after another participant committed the launch keeps
// waiting: the records it waits for exist.
__device__ inline bool wave_spin_fail(unsigned &it, const leaf_args &a) {
  ++it;
  const bool timed_out = it > a.spin_limit;
  if (timed_out || (it & 63) == 1) {
    int st = 0;
    if ((threadIdx.x & 63) == 0)
      st = timed_out ? leaffb::decide(a.state, a.tag0, 2) : leaffb::state(a.state, a.tag0);
    st = __shfl(st, 0, 64);
    if (st == 2) return true;
    if (timed_out) it = 0;
  }
  __builtin_amdgcn_s_sleep(1);
  return false;
}

// Wave-wide (|v|, position, row) argmax; result valid in every lane.
__device__ inline void wave_argmax3(double &bv, int &bi, int &br) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double ov = __shfl_xor(bv, off, 64);
    const int oi = __shfl_xor(bi, off, 64);
    const int orr = __shfl_xor(br, off, 64);
    argmax3(bv, bi, br, ov, oi, orr);
  }
}

// Apply (dst, src) row pairs to one column with one wave: every lane loads
// its source before any lane stores (one load instruction, then one store).
__device__ inline void wave_gather(double *col, int cnt, const int *dst, const int *src) {
  const int lane = threadIdx.x & 63;
  if (lane < cnt) {
    const double v = ld_dev(col + src[lane]);
    col[dst[lane]] = v;
  }
}

__global__ __launch_bounds__(LEAF_THREADS) void lu_leaf_kernel(leaf_args a) {
  extern __shared__ double lds[];
  const int w = a.w, R = a.R;
  double *L = lds;                                // L[c*R + r]
  int *pos = (int *)(lds + (size_t)w * R);        // logical position of local row r
  __shared__ double red_v[2][LEAF_THREADS / 64];
  __shared__ int red_i[2][LEAF_THREADS / 64], red_r[2][LEAF_THREADS / 64];
  __shared__ double prow_s[LEAF_THREADS / 64][LEAF_W];  // each wave's copy of the winning row
  __shared__ int s_p[LEAF_W];
  __shared__ int s_dst[PAIRS_MAX], s_src[PAIRS_MAX];
  __shared__ int s_cnt, s_outcome;
  bool dead = false;  // a wait of this wave found the launch abandoned

  const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int r0 = g * R;
  const int rows = max(0, min(a.mp - r0, R));
  // a bounded wait that gave up earlier on this stream (the error word): leave the panel untouched
  // and record no interchanges
  if (ld_dev(a.err)) {
    if (g == 0 && tid == 0) a.pairs[0] = 0;
    return;
  }

  for (int e = tid; e < rows * w; e += LEAF_THREADS) {
    const int c = e / rows, r = e - c * rows;
    L[c * R + r] = a.P[(long)c * a.lda + r0 + r];
  }
  for (int r = tid; r < rows; r += LEAF_THREADS) pos[r] = r0 + r;
  __syncthreads();

  unsigned long long *cand_slots = a.slots;
  unsigned long long *row_slots = a.slots + (size_t)2 * LEAF_MAXG * CAND_WORDS;
  // this thread's candidate for column 0 (every thread owns rows tid, tid+256, ...)
  double bv = -1.0;
  int bi = INT_MAX, br = -1;
  for (int r = tid; r < rows; r += LEAF_THREADS) argmax3(bv, bi, br, fabs(L[r]), pos[r], r);

  for (int j = 0; j < w; j++) {
    const int par = j & 1;
    const unsigned long long tag = (unsigned long long)(a.tag0 + (unsigned)j) << 32;

    // workgroup candidate: wave reduction, then the 4 wave results (one barrier;
    // the arrays alternate by column so a wave running ahead cannot clobber them)
    wave_argmax3(bv, bi, br);
    if (lane == 0) {
      red_v[par][wave] = bv;
      red_i[par][wave] = bi;
      red_r[par][wave] = br;
    }
    __syncthreads();
    bv = red_v[par][0];
    bi = red_i[par][0];
    br = red_r[par][0];
    for (int v = 1; v < LEAF_THREADS / 64; v++) argmax3(bv, bi, br, red_v[par][v], red_i[par][v], red_r[par][v]);

    // publish (wave 0): the candidate row first, then (value, position)
    if (wave == 0) {
      const bool have = (br >= 0);
      const int c = lane >> 1;
      st_dev(row_slots + ((size_t)par * G + g) * ROW_WORDS + lane,
             tag | dbl_half((have && c < w) ? L[c * R + br] : 0.0, lane & 1));
      if (lane < 3) {
        const unsigned pl = lane < 2 ? dbl_half(have ? L[j * R + br] : 0.0, lane) : (unsigned)bi;
        st_dev(cand_slots + ((size_t)par * G + g) * CAND_WORDS + lane, tag | pl);
      }
    }

    // round trip 1, per wave: every workgroup's (value, position); lane l
    // holds the records of workgroups l and l + 64, then a wave reduction
    double gv = -1.0;
    int gi = INT_MAX, gw = -1;
    {
      const unsigned long long *cs = cand_slots + (size_t)par * G * CAND_WORDS;
      unsigned it = 0;
      for (;;) {
        int ok = 1;
        gv = -1.0;
        gi = INT_MAX;
        gw = -1;
        for (int t = lane; t < G; t += 64) {
          const unsigned long long w0 = ld_dev(cs + (size_t)t * CAND_WORDS);
          const unsigned long long w1 = ld_dev(cs + (size_t)t * CAND_WORDS + 1);
          const unsigned long long w2 = ld_dev(cs + (size_t)t * CAND_WORDS + 2);
          if (((w0 & w1 & w2) >> 32) != (tag >> 32) || ((w0 | w1 | w2) >> 32) != (tag >> 32)) ok = 0;
          const int ci = (int)(unsigned)w2;
          if (ci >= j && ci < a.mp) argmax3(gv, gi, gw, fabs(dbl_from((unsigned)w0, (unsigned)w1)), ci, t);
        }
        if (__all(ok)) break;
        if (wave_spin_fail(it, a)) {
          dead = true;
          break;
        }
      }
      wave_argmax3(gv, gi, gw);
    }

    // round trip 2, per wave: the winner's row (all-NaN column: no winner)
    {
      const unsigned long long *rs = row_slots + ((size_t)par * G + (gw >= 0 ? gw : 0)) * ROW_WORDS;
      unsigned it = 0;
      for (;;) {
        int ok = 1;
        double v = __longlong_as_double(0x7FF8000000000000LL);
        if (lane < w && gw >= 0) {
          const unsigned long long lo = ld_dev(rs + 2 * lane), hi = ld_dev(rs + 2 * lane + 1);
          if ((lo >> 32) != (tag >> 32) || (hi >> 32) != (tag >> 32)) ok = 0;
          v = dbl_from((unsigned)lo, (unsigned)hi);
        }
        if (lane < w) prow_s[wave][lane] = v;
        if (__all(ok)) break;
        if (wave_spin_fail(it, a)) {
          dead = true;
          break;
        }
      }
      __builtin_amdgcn_wave_barrier();
    }

    const double *prow = prow_s[wave];
    const int p = (gw >= 0) ? gi : j;
    const double inv = 1.0 / prow[j];
    // relabel, scale, rank-1 update; the next column's candidate on the way
    bv = -1.0;
    bi = INT_MAX;
    br = -1;
    for (int r = tid; r < rows; r += LEAF_THREADS) {
      int q = pos[r];
      if (q == p) q = j;
      else if (q == j) q = p;
      pos[r] = q;
      if (q > j) {
        const double l = L[j * R + r] * inv;
        L[j * R + r] = l;
        for (int c = j + 1; c < w; c++) L[c * R + r] = fma(l, -prow[c], L[c * R + r]);
        if (j + 1 < w) argmax3(bv, bi, br, fabs(L[(j + 1) * R + r]), q, r);
      }
    }
    if (tid == 0) {
      s_p[j] = p;
      if (g == 0) a.piv[j] = p - j;
    }
  }
  // the workgroup decides the launch: committed (every wave saw every record)
  // or abandoned (some wait gave up first); the outcome is the same for all.
  // An abandoned launch writes nothing, and workgroup 0 factors the leaf
  // alone from memory.
  {
    const int any_dead = __syncthreads_or(dead);
    if (tid == 0) s_outcome = leaffb::decide(a.state, a.tag0, any_dead ? 2 : 1);
    __syncthreads();
    if (s_outcome == 2) {
      if (g == 0) {
        if (tid == 0) atomicAdd(a.state + 1, 1ULL);  // fallbacks run, for the tests
        leaffb::factor(a.P, a.Ppanel, a.lda, a.mp, a.w, a.pw, a.pc0, a.piv, a.pairs, a.fscr);
      }
      return;
    }
  }

  // rows to their final positions (a permutation: no two rows collide)
  for (int e = tid; e < rows * w; e += LEAF_THREADS) {
    const int c = e / rows, r = e - c * rows;
    a.P[(long)c * a.lda + pos[r]] = L[c * R + r];
  }

  // replay the interchanges on (position -> source row) entries held by wave 0
  if (tid < 64) {
    const int lane = tid;
    int key = -1, val = -1, cnt = 0;
    for (int j = 0; j < w; j++) {
      const int p = s_p[j];
      if (p == j) continue;
      unsigned long long mj = __ballot(key == j);
      if (!mj) {
        if (lane == cnt) { key = j; val = j; }
        mj = 1ULL << cnt;
        cnt++;
      }
      unsigned long long mp_ = __ballot(key == p);
      if (!mp_) {
        if (lane == cnt) { key = p; val = p; }
        mp_ = 1ULL << cnt;
        cnt++;
      }
      const int lj = __ffsll((long long)mj) - 1, lp = __ffsll((long long)mp_) - 1;
      const int vj = __shfl(val, lj, 64), vp = __shfl(val, lp, 64);
      if (lane == lj) val = vp;
      if (lane == lp) val = vj;
    }
    const bool moved = (key >= 0 && key != val);
    const unsigned long long mm = __ballot(moved);
    const int slot = __popcll(mm & ((1ULL << lane) - 1ULL));
    if (moved) {
      s_dst[slot] = key;
      s_src[slot] = val;
      if (g == 0) {
        a.pairs[1 + 2 * slot] = key + a.pc0;
        a.pairs[2 + 2 * slot] = val + a.pc0;
      }
    }
    if (lane == 0) {
      s_cnt = __popcll(mm);
      if (g == 0) a.pairs[0] = __popcll(mm);
    }
  }
  __syncthreads();

  // the same interchanges on the panel's other columns, one column per wave
  const int cnt = s_cnt;
  const int others = a.pw - w;
  if (cnt > 0) {
    for (int q = g * (LEAF_THREADS / 64) + (tid >> 6); q < others; q += G * (LEAF_THREADS / 64)) {
      const int cc = q < a.pc0 ? q : q + w;
      wave_gather(a.Ppanel + (long)cc * a.lda, cnt, s_dst, s_src);
    }
  }
}

// Composite interchange of one panel: the leaves' (dst, src) row lists,
// applied one after the other, are replayed on a position -> source-row map
// kept in an LDS hash table (one workgroup), and the positions whose source
// changed become one list of at most 2*nb (dst, src) pairs: count, then pairs.
constexpr int COMP_MAX = 2048;  // 2 * the largest panel width (1024)
constexpr int HASH_SZ = 4096;

__device__ inline int hslot(int key) { return (int)(((unsigned)key * 2654435761u) >> 20); }  // 12 bits

__global__ __launch_bounds__(256) void panel_compose_kernel(const int *__restrict__ pairs, int nleaves,
                                                            int *__restrict__ out, const int *__restrict__ err) {
  // after a bounded wait gave up (the error word) the leaves' lists are not trustworthy: no interchanges
  if (ld_dev(err)) {
    if (threadIdx.x == 0) out[0] = 0;
    return;
  }
  __shared__ int hk[HASH_SZ], hv[HASH_SZ];
  __shared__ int val[PAIRS_MAX];
  __shared__ int ncomp;
  for (int e = threadIdx.x; e < HASH_SZ; e += 256) hk[e] = -1;
  if (threadIdx.x == 0) ncomp = 0;
  __syncthreads();
  const int t = threadIdx.x;
  for (int q = 0; q < nleaves; q++) {
    const int *pl = pairs + (size_t)q * PAIRS_STRIDE;
    const int cnt = pl[0];
    if (t < cnt) {  // new[dst] = old[src]: read every source first
      const int src = pl[2 + 2 * t];
      int h = hslot(src), v = src;
      for (int probe = 0; probe < HASH_SZ; probe++, h = (h + 1) & (HASH_SZ - 1)) {
        const int key = hk[h];
        if (key == src) { v = hv[h]; break; }
        if (key == -1) break;
      }
      val[t] = v;
    }
    __syncthreads();
    if (t < cnt) {  // then insert or update the destinations (distinct within a leaf)
      const int dst = pl[1 + 2 * t];
      int h = hslot(dst);
      for (int probe = 0; probe < HASH_SZ; probe++, h = (h + 1) & (HASH_SZ - 1)) {
        const int prev = atomicCAS(&hk[h], -1, dst);
        if (prev == -1 || prev == dst) { hv[h] = val[t]; break; }
      }
    }
    __syncthreads();
  }
  for (int e = t; e < HASH_SZ; e += 256) {
    const int key = hk[e];
    if (key >= 0 && hv[e] != key) {
      const int slot = atomicAdd(&ncomp, 1);
      if (slot < COMP_MAX) {
        out[1 + 2 * slot] = key;
        out[2 + 2 * slot] = hv[e];
      }
    }
  }
  __syncthreads();
  if (t == 0) out[0] = min(ncomp, COMP_MAX);
}

// Full-width interchanges of one panel: for every column outside the panel
// (ranges [0, nleft) and [c1, c2)) apply the composite (dst, src) list, one
// column per wave: every lane loads all its sources before any lane stores.
// Each lane reads its (dst, src) entries from the list itself (16 KB at most,
// cached), so the kernel uses no LDS and can run beside trailing-update and
// panel workgroups that hold most of a CU's LDS.
template <typename T>
__global__ __launch_bounds__(256) void panel_swap_kernel(T *__restrict__ A, long lda, int nleft, int c1, int c2,
                                                         const int *__restrict__ comp, const int *__restrict__ err) {
  if (ld_dev(err)) return;  // a bounded wait gave up: the list is not trustworthy
  const int cnt = min(comp[0], COMP_MAX);
  if (cnt <= 0) return;
  const int total = nleft + (c2 - c1);
  const int lane = threadIdx.x & 63;
  constexpr int PER = COMP_MAX / 64;
  const int chunks = (cnt + 63) >> 6;  // uniform: chunks of 64 pairs in use
  // The list is the same for every column: each lane keeps its (dst, src)
  // entries in registers.  Lanes past the list's end repeat its last entry,
  // so every load is unguarded (a guarded load waits for its data before
  // the next one issues) and only the stores test the position.
  int dst[PER], src[PER];
#pragma unroll
  for (int i = 0; i < PER; i++) {
    const int p = min(lane + 64 * i, cnt - 1);
    dst[i] = comp[1 + 2 * p];
    src[i] = comp[2 + 2 * p];
  }
  for (int q = blockIdx.x * 4 + (threadIdx.x >> 6); q < total; q += gridDim.x * 4) {
    const int col = q < nleft ? q : c1 + (q - nleft);
    T *cp = A + (long)col * lda;
    T v[PER];
    // every source is read before any destination is written
#pragma unroll
    for (int i = 0; i < PER; i++)
      if (i < chunks) v[i] = cp[src[i]];
#pragma unroll
    for (int i = 0; i < PER; i++)
      if (i < chunks && lane + 64 * i < cnt) cp[dst[i]] = v[i];
  }
}

// ---------------------------------------------------------------------------
// Forward sequential row interchanges (LAPACK dlaswp semantics): for
// i = 0..np-1, swap rows i and i + piv[i] of columns [0, ncols).  One thread
// per column; the pivots are staged in LDS 1024 at a time.
__global__ __launch_bounds__(256) void laswp_kernel(double *__restrict__ A, long lda, int nrows,
                                                    int ncols, int np, const int *__restrict__ piv) {
  __shared__ int sp[1024];
  const int c = blockIdx.x * 256 + threadIdx.x;
  double *col = A + (long)c * lda;
  for (int base = 0; base < np; base += 1024) {
    const int cnt = min(1024, np - base);
    __syncthreads();
    for (int t = threadIdx.x; t < cnt; t += 256) sp[t] = piv[base + t];
    __syncthreads();
    if (c < ncols) {
      for (int t = 0; t < cnt; t++) {
        const int p = sp[t];
        const int r = base + t;
        if (p > 0 && r + p < nrows) {  // out-of-range entries are ignored
          const double a = col[r], b = col[r + p];
          col[r] = b;
          col[r + p] = a;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Triangular-solve leaves: blocks of at most 64 rows / columns.
constexpr int TRSM_LEAF = 64;

// Left-side triangular solve of a b x b leaf (b <= 64): B <- op(T)^-1 B for
// ncols columns.  A wave takes NC columns at a time with lane i owning row i
// of each, so every column is read and written as one coalesced 512-byte
// row of lanes (a thread per column would touch one cache line -- and for a
// wide matrix one page -- per lane and element).  Lane i keeps row i of
// M = op(T) in registers (staged through LDS); the solve is dtrsm's column
// order: at step k lane k finalizes x_k (divided by M(k,k) unless unit), it
// is broadcast with v_readlane, and the lanes of the rows still open
// subtract M(i,k) x_k.  Forward (lower-type M) takes k upwards, backward
// downwards; each row receives its updates in the same order, with the same
// fused multiply-adds, as a dot product over k would.
constexpr int TRSM_NC = 4;

__device__ __forceinline__ double lane_bcast(double v, int l) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)u, l);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <bool LOWER, bool TRANS, bool UNIT>
__global__ __launch_bounds__(256) void trsm_leaf_kernel(int b, int ncols, const double *__restrict__ T,
                                                        long ldt, double *__restrict__ B, long ldb,
                                                        const int *__restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  constexpr bool FORWARD = (LOWER != TRANS);  // op(T) lower-type
  const int lane = threadIdx.x & 63;
  // row `lane` of M = op(T), strictly on the solved side of the diagonal,
  // read straight from T (64 x 64 at most, cached): no LDS, so the leaf can
  // run beside trailing-update workgroups that hold all of a CU's LDS
  // (clamped indices: every load is in bounds and issued before any is
  // used; a guarded load was waited for where it issued)
  const int lc = min(lane, b - 1);
  double mrow[TRSM_LEAF];
#pragma unroll
  for (int k = 0; k < TRSM_LEAF; k++) {
    const int kc = min(k, b - 1);
    mrow[k] = TRANS ? T[(long)lc * ldt + kc] : T[(long)kc * ldt + lc];
  }
  double diag = UNIT ? 1.0 : T[(long)lc * ldt + lc];
#pragma unroll
  for (int k = 0; k < TRSM_LEAF; k++) {
    const bool in = lane < b && k < b && (FORWARD ? k < lane : k > lane);
    mrow[k] = in ? mrow[k] : 0.0;
  }
  diag = lane < b ? diag : 1.0;
  const int nwaves = gridDim.x * 4;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  for (int c0 = wave * TRSM_NC; c0 < ncols; c0 += nwaves * TRSM_NC) {
    double x[TRSM_NC];
#pragma unroll
    for (int q = 0; q < TRSM_NC; q++) x[q] = B[(long)min(c0 + q, ncols - 1) * ldb + lc];
#pragma unroll
    for (int q = 0; q < TRSM_NC; q++) x[q] = (c0 + q < ncols && lane < b) ? x[q] : 0.0;
#pragma unroll
    for (int s = 0; s < TRSM_LEAF; s++) {
      const int k = FORWARD ? s : TRSM_LEAF - 1 - s;
      if (k < b) {
        if (!UNIT && lane == k) {
#pragma unroll
          for (int q = 0; q < TRSM_NC; q++) x[q] = x[q] / diag;
        }
        double xk[TRSM_NC];
#pragma unroll
        for (int q = 0; q < TRSM_NC; q++) xk[q] = lane_bcast(x[q], k);
        if (FORWARD ? lane > k : lane < k) {
#pragma unroll
          for (int q = 0; q < TRSM_NC; q++) x[q] = fma(-mrow[k], xk[q], x[q]);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < TRSM_NC; q++)
      if (c0 + q < ncols && lane < b) B[(long)(c0 + q) * ldb + lane] = x[q];
  }
}

// Right-side triangular solve of a b x b leaf (b <= 64): X op(T) = X0 for
// nrows rows, one thread per row (the rows are independent).  M = op(T) is
// staged in LDS as sM[i][j] = M(i, j); upper-type M (FWD) is solved over the
// columns left to right, x_j = (x0_j - sum_{i<j} x_i M(i,j)) / M(j,j), lower-
// type M right to left.  A non-unit diagonal is applied as a multiply by its
// reciprocal, the order of dtrsm's right-side loops.  Every lane reads the
// same sM entry at a time (LDS broadcast).
template <bool FWD, bool UNIT>
__global__ __launch_bounds__(256) void trsm_right_leaf_kernel(int nrows, int b, const double *__restrict__ T,
                                                              long ldt, int trans, double *__restrict__ X, long ldx,
                                                              const int *__restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  __shared__ double sM[TRSM_LEAF][TRSM_LEAF + 1];
  {
    // the b x b triangle block: all of a thread's loads in flight at once
    constexpr int PER = TRSM_LEAF * TRSM_LEAF / 256;
    double tv[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
      // past b * b: the last element again (in bounds, not stored)
      const int e = min((int)threadIdx.x + 256 * q, b * b - 1);
      const int col = e / b, row = e - col * b;  // T(row, col), read down the columns
      tv[q] = T[(long)col * ldt + row];
    }
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int e = threadIdx.x + 256 * q;
      const int col = e / b, row = e - col * b;
      if (e < b * b) {
        if (trans)
          sM[col][row] = tv[q];
        else
          sM[row][col] = tv[q];
      }
    }
  }
  __syncthreads();
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= nrows) return;
  double v[TRSM_LEAF];
#pragma unroll
  for (int c = 0; c < TRSM_LEAF; c++) v[c] = X[(long)min(c, b - 1) * ldx + r];
#pragma unroll
  for (int c = 0; c < TRSM_LEAF; c++) v[c] = c < b ? v[c] : 0.0;
#pragma unroll
  for (int q = 0; q < TRSM_LEAF; q++) {
    const int kk = FWD ? q : TRSM_LEAF - 1 - q;
    if (kk < b) {
      if (!UNIT) v[kk] *= 1.0 / sM[kk][kk];
      if (FWD) {
#pragma unroll
        for (int j = kk + 1; j < TRSM_LEAF; j++)
          if (j < b) v[j] = fma(-sM[kk][j], v[kk], v[j]);
      } else {
#pragma unroll
        for (int j = 0; j < kk; j++) v[j] = fma(-sM[kk][j], v[kk], v[j]);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < TRSM_LEAF; c++)
    if (c < b) X[(long)c * ldx + r] = v[c];
}

// ---------------------------------------------------------------------------
// v of lane `src` (a compile-time lane after unrolling), as a scalar
__device__ __forceinline__ double lane_value(double v, int src) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)bits, src), hi = __builtin_amdgcn_readlane((int)(bits >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Unblocked lower Cholesky of a b x b diagonal block (b <= 64) in one wave:
// PLA_dchol's left-looking order (dgemv update of column k, test
// a(k,k) <= 0, sqrt, scale by the reciprocal).  Lane i holds row i in
// registers; L(k, j) reaches every lane by readlane, so a column costs no
// barrier and no LDS round trip (the LDS version took about 60 us per block
// beside other work).  The sums and roundings are the same as with the
// block in LDS.  On failure *info receives offset + k + 1 (1-based global
// index) and the block is not written; the caller restores it from its
// backup.
__global__ __launch_bounds__(64) void potrf_leaf_kernel(int b, double *__restrict__ Ablk, long lda,
                                                        int offset, int *__restrict__ info) {
  if (*info) return;
  const int i = threadIdx.x;
  double a[TRSM_LEAF];  // a[j] = A(i, j), j <= i; L(i, j) once column j is done
  {
    // every load of the row in flight at once (clamped: always in bounds)
#pragma unroll
    for (int j = 0; j < TRSM_LEAF; j++) a[j] = Ablk[(long)min(j, b - 1) * lda + min(i, b - 1)];
#pragma unroll
    for (int j = 0; j < TRSM_LEAF; j++) a[j] = (j < b && i < b && i >= j) ? a[j] : 0.0;
  }
  // no early exit inside the unrolled loop (a[] stays in registers): a
  // failed pivot stops the updates through `fail`
  int fail = 0;
#pragma unroll
  for (int kk = 0; kk < TRSM_LEAF; kk++) {
    if (kk < b && !fail) {  // uniform
      double y = a[kk];
#pragma unroll
      for (int j = 0; j < kk; j++) y = fma(-lane_value(a[j], kk), a[j], y);
      if (i >= kk) a[kk] = y;
      const double d = lane_value(a[kk], kk);
      if (d <= 0.0) {
        fail = kk + 1;
      } else {
        const double sq = sqrt(d);
        const double invert = 1.0 / sq;
        if (i == kk) a[kk] = sq;
        else if (i > kk) a[kk] *= invert;
      }
    }
  }
  if (fail) {
    if (i == 0) atomicCAS(info, 0, offset + fail);
    return;
  }
#pragma unroll
  for (int j = 0; j < TRSM_LEAF; j++)
    if (j < b && i < b && i >= j) Ablk[(long)j * lda + i] = a[j];
}

// ---------------------------------------------------------------------------
// 2-D copy (m x n, column-major), optionally skipped after a failure.
__global__ __launch_bounds__(256) void copy2d_kernel(int m, int n, const double *__restrict__ S, long lds_,
                                                     double *__restrict__ D, long ldd,
                                                     const int *__restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  const long total = (long)m * n;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int j = (int)(e / m), i = (int)(e - (long)j * m);
    D[(long)j * ldd + i] = S[(long)j * lds_ + i];
  }
}

// Application-interface transfers (API/PLA_API.c:390-792).  A packed M x N
// block (column-major, leading dimension M) travels between host and HBM in
// one DMA copy; these kernels move it between the packed form and its place in
// the global object.  Local row i / column j of the block lands on global row
// rmap[i] / column cmap[j] (or r0 + i / c0 + j without a map), so many
// sub-blocks of a multi-axpy go in one launch.  Rows run over x (coalesced
// within a column), columns over y.
// MPI_DOUBLE_COMPLEX elements of the application interface: alpha x + d with
// the complex product rounded before the add
struct api_z {
  double re, im;
};
__device__ __forceinline__ api_z operator*(api_z a, api_z b) {
#pragma clang fp contract(off)
  return api_z{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ api_z operator+(api_z a, api_z b) { return api_z{a.re + b.re, a.im + b.im}; }

template <typename T>
__global__ __launch_bounds__(256) void api_scatter_add_kernel(int M, int N, T alpha, const T *__restrict__ src,
                                                              T *__restrict__ dst, long ldd,
                                                              const int *__restrict__ rmap,
                                                              const int *__restrict__ cmap, int r0, int c0) {
#pragma clang fp contract(off)  // alpha * x rounded, then added: no fused multiply-add
  for (int j = blockIdx.y; j < N; j += gridDim.y) {
    const long gj = cmap ? cmap[j] : c0 + j;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < M; i += gridDim.x * 256) {
      const long gi = rmap ? rmap[i] : r0 + i;
      T *d = dst + gj * ldd + gi;
      *d = *d + alpha * src[(long)j * M + i];
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void api_gather_kernel(int M, int N, const T *__restrict__ src, long lds_,
                                                         T *__restrict__ dst, const int *__restrict__ rmap,
                                                         const int *__restrict__ cmap, int r0, int c0) {
  for (int j = blockIdx.y; j < N; j += gridDim.y) {
    const long gj = cmap ? cmap[j] : c0 + j;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < M; i += gridDim.x * 256) {
      const long gi = rmap ? rmap[i] : r0 + i;
      dst[(long)j * M + i] = src[gj * lds_ + gi];
    }
  }
}

// Scale an m x n block by alpha.
__global__ __launch_bounds__(256) void scale2d_kernel(int m, int n, double alpha, double *__restrict__ D,
                                                      long ldd) {
  const long total = (long)m * n;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int j = (int)(e / m), i = (int)(e - (long)j * m);
    D[(long)j * ldd + i] *= alpha;
  }
}

// ---------------------------------------------------------------------------
// Matrix norms (BLAS2/PLA_Matrix_infinity_norm.c:13, PLA_Matrix_one_norm.c:13):
// absolute row (column) sums, then their maximum.  Row sums are formed in two
// passes -- per 256-column chunk, then over chunks in order -- so the result is
// deterministic; the maximum of non-negative doubles is an integer maximum of
// their bit patterns.
constexpr int NORM_CHUNK = 256;

__global__ __launch_bounds__(256) void row_asum_partial_kernel(int m, int n, const double *__restrict__ A,
                                                               long lda, double *__restrict__ part) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  const int j0 = blockIdx.y * NORM_CHUNK, j1 = min(n, j0 + NORM_CHUNK);
  double s = 0.0;
  for (int j = j0; j < j1; j++) s += fabs(A[(long)j * lda + i]);
  part[(long)blockIdx.y * m + i] = s;
}

__global__ __launch_bounds__(256) void row_asum_max_kernel(int m, int nch, const double *__restrict__ part,
                                                           unsigned long long *__restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  double s = 0.0;
  for (int c = 0; c < nch; c++) s += part[(long)c * m + i];
  atomicMax(out, (unsigned long long)__double_as_longlong(s));
}

__global__ __launch_bounds__(256) void col_asum_max_kernel(int m, const double *__restrict__ A, long lda,
                                                           unsigned long long *__restrict__ out) {
  __shared__ double red[256];
  const double *col = A + (long)blockIdx.x * lda;
  double s = 0.0;
  for (int i = threadIdx.x; i < m; i += 256) s += fabs(col[i]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(red[0]));
}

// ---------------------------------------------------------------------------
// drand48 fill: element (i, j) takes draw number skip + j*m + i after
// srand48(seed); value = X * 2^-48 * 2 - 1 (exact in double).  Each thread
// jumps ahead once (affine LCG map squared log2 times) and then steps
// through a run of 64 consecutive draws.
constexpr uint64_t LCG_A = 0x5DEECE66DULL, LCG_C = 0xBULL, LCG_MASK = (1ULL << 48) - 1;

__device__ inline uint64_t lcg_skip(uint64_t x, uint64_t k) {
  uint64_t mul = LCG_A, add = LCG_C, am = 1, aa = 0;
  while (k) {
    if (k & 1) {
      am = (am * mul) & LCG_MASK;
      aa = (aa * mul + add) & LCG_MASK;
    }
    add = ((mul + 1) * add) & LCG_MASK;
    mul = (mul * mul) & LCG_MASK;
    k >>= 1;
  }
  return (am * x + aa) & LCG_MASK;
}

__global__ __launch_bounds__(256) void fill_drand48_kernel(int m, int n, double *__restrict__ D, long ldd,
                                                           uint64_t state0, uint64_t skip) {
  const uint64_t total = (uint64_t)m * (uint64_t)n;
  const uint64_t run = 64;
  const uint64_t start = ((uint64_t)blockIdx.x * 256 + threadIdx.x) * run;
  if (start >= total) return;
  uint64_t x = lcg_skip(state0, skip + start);
  const uint64_t end = start + run < total ? start + run : total;
  for (uint64_t e = start; e < end; e++) {
    x = (LCG_A * x + LCG_C) & LCG_MASK;
    const double v = ldexp((double)x, -48) * 2.0 - 1.0;
    const uint64_t j = e / (uint64_t)m, i = e - j * (uint64_t)m;
    D[(long)j * ldd + (long)i] = v;
  }
}

inline unsigned grid1d(long total, int per_block) {
  long g = (total + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > 65535L * 64) g = 65535L * 64;
  return (unsigned)g;
}

}  // namespace

namespace pla {

template <bool TA, bool TB, int TRI, int ROLE>
static void launch_gemm(hipStream_t s, int m, int n, int k, double alpha, const double *A, long lda,
                        const double *B, long ldb, double beta, double *C, long ldc,
                        const int *abort_flag) {
  const int tm = (m + GBM - 1) / GBM, tn = (n + GBN - 1) / GBN;
  // triangular updates (square C: m == n, GBM == GBN) launch only the
  // triangle's tm (tm + 1) / 2 tiles
  const long nt = TRI ? (long)tm * (tm + 1) / 2 : (long)tm * tn;
  const int nblk = (int)(((nt + 7) / 8) * 8);
  // the interior path addresses a lane's operand pairs as a 32-bit byte offset
  // of at most (31 ld + 128) doubles from a uniform base
  const long lim = (1L << 32) / 8 - 128;
  const int fast = (alpha == 1.0 || alpha == -1.0) && 31 * lda < lim && 31 * ldb < lim;
  hipLaunchKernelGGL((dgemm_kernel<TA, TB, TRI, ROLE>), dim3(nblk), dim3(256), 0, s, m, n, k, alpha,
                     A, lda, B, ldb, beta, C, ldc, tm, tn, abort_flag, fast);
}

int k_gemm(hipStream_t s, int role, bool ta, bool tb, int m, int n, int k, double alpha,
           const double *A, long lda, const double *B, long ldb, double beta, double *C, long ldc,
           const int *abort_flag) {
  if (m <= 0 || n <= 0) return PLA_SUCCESS;
  if (!ta && n <= TS_N && k <= TS_K && k > 0 && m >= 256 && role == 0) {
    const unsigned grid = (unsigned)((m + 63) / 64);
    if (tb)
      hipLaunchKernelGGL((dgemm_tall_kernel<true>), dim3(grid), dim3(256), 0, s, m, n, k, alpha, A, lda, B, ldb, beta,
                         C, ldc, abort_flag);
    else
      hipLaunchKernelGGL((dgemm_tall_kernel<false>), dim3(grid), dim3(256), 0, s, m, n, k, alpha, A, lda, B, ldb,
                         beta, C, ldc, abort_flag);
    PLA_HIP(hipGetLastError());
    return PLA_SUCCESS;
  }
  if ((long)((m + GBM - 1) / GBM) * ((n + GBN - 1) / GBN) > (1L << 31) - 8)
    return fail(PLA_ERR_DIMENSIONS, "gemm: too many tiles");
  if (role == 1 && !ta && !tb)
    launch_gemm<false, false, 0, 1>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, abort_flag);
  else if (!ta && !tb)
    launch_gemm<false, false, 0, 0>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, abort_flag);
  else if (!ta && tb)
    launch_gemm<false, true, 0, 0>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, abort_flag);
  else if (ta && !tb)
    launch_gemm<true, false, 0, 0>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, abort_flag);
  else
    launch_gemm<true, true, 0, 0>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, abort_flag);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_syrk_lower(hipStream_t s, int role, int n, int k, double alpha, const double *X, long ldx,
                 double beta, double *C, long ldc, const int *abort_flag) {
  if (n <= 0) return PLA_SUCCESS;
  if (role == 1)
    launch_gemm<false, true, 1, 1>(s, n, n, k, alpha, X, ldx, X, ldx, beta, C, ldc, abort_flag);
  else
    launch_gemm<false, true, 1, 0>(s, n, n, k, alpha, X, ldx, X, ldx, beta, C, ldc, abort_flag);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_syrk(hipStream_t s, bool lower, bool trans, int n, int k, double alpha, const double *X, long ldx, double beta,
           double *C, long ldc) {
  if (n <= 0) return PLA_SUCCESS;
  if ((long)((n + GBM - 1) / GBM) * ((n + GBN - 1) / GBN) > (1L << 31) - 8)
    return fail(PLA_ERR_DIMENSIONS, "syrk: too many tiles");
  // C (n x n) += alpha op(X) op(X)^T with op(X) = X (n x k) or X^T (X k x n)
  if (lower && !trans)
    launch_gemm<false, true, 1, 0>(s, n, n, k, alpha, X, ldx, X, ldx, beta, C, ldc, nullptr);
  else if (lower)
    launch_gemm<true, false, 1, 0>(s, n, n, k, alpha, X, ldx, X, ldx, beta, C, ldc, nullptr);
  else if (!trans)
    launch_gemm<false, true, 2, 0>(s, n, n, k, alpha, X, ldx, X, ldx, beta, C, ldc, nullptr);
  else
    launch_gemm<true, false, 2, 0>(s, n, n, k, alpha, X, ldx, X, ldx, beta, C, ldc, nullptr);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_laswp(hipStream_t s, double *A, long lda, int nrows, int ncols, int np, const int *piv) {
  if (ncols <= 0 || np <= 0) return PLA_SUCCESS;
  hipLaunchKernelGGL(laswp_kernel, dim3((ncols + 255) / 256), dim3(256), 0, s, A, lda, nrows, ncols, np,
                     piv);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_trsm_leaf(hipStream_t s, bool lower, bool trans, bool unit, int b, int ncols, const double *T,
                long ldt, double *B, long ldb, const int *abort_flag) {
  if (b <= 0 || ncols <= 0) return PLA_SUCCESS;
  if (b > TRSM_LEAF) return fail(PLA_ERR_INVALID_PARAM, "trsm leaf: b = %d > %d", b, TRSM_LEAF);
  // four waves per workgroup, TRSM_NC columns per wave and pass; at most 512
  // workgroups (grid-stride beyond)
  const long groups = ((long)ncols + TRSM_NC - 1) / TRSM_NC;
  const long wgs = (groups + 3) / 4;
  dim3 grid((unsigned)(wgs < 512 ? wgs : 512)), blk(256);
#define PLA_TRSM_CASE(LO, TR, UN)                                                                     \
  if (lower == LO && trans == TR && unit == UN) {                                                     \
    hipLaunchKernelGGL((trsm_leaf_kernel<LO, TR, UN>), grid, blk, 0, s, b, ncols, T, ldt, B, ldb,     \
                       abort_flag);                                                                   \
  }
  PLA_TRSM_CASE(true, false, true)
  PLA_TRSM_CASE(true, false, false)
  PLA_TRSM_CASE(true, true, true)
  PLA_TRSM_CASE(true, true, false)
  PLA_TRSM_CASE(false, false, true)
  PLA_TRSM_CASE(false, false, false)
  PLA_TRSM_CASE(false, true, true)
  PLA_TRSM_CASE(false, true, false)
#undef PLA_TRSM_CASE
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_trsm_right_leaf(hipStream_t s, bool lower, bool trans, bool unit, int nrows, int b, const double *T, long ldt,
                      double *X, long ldx, const int *abort_flag) {
  if (nrows <= 0 || b <= 0) return PLA_SUCCESS;
  if (b > TRSM_LEAF) return fail(PLA_ERR_INVALID_PARAM, "trsm right leaf: b = %d", b);
  const dim3 grid((nrows + 255) / 256), blk(256);
  const bool fwd = (lower == trans);  // op(T) upper-type
  if (fwd && unit)
    hipLaunchKernelGGL((trsm_right_leaf_kernel<true, true>), grid, blk, 0, s, nrows, b, T, ldt, (int)trans, X, ldx,
                       abort_flag);
  else if (fwd)
    hipLaunchKernelGGL((trsm_right_leaf_kernel<true, false>), grid, blk, 0, s, nrows, b, T, ldt, (int)trans, X, ldx,
                       abort_flag);
  else if (unit)
    hipLaunchKernelGGL((trsm_right_leaf_kernel<false, true>), grid, blk, 0, s, nrows, b, T, ldt, (int)trans, X, ldx,
                       abort_flag);
  else
    hipLaunchKernelGGL((trsm_right_leaf_kernel<false, false>), grid, blk, 0, s, nrows, b, T, ldt, (int)trans, X,
                       ldx, abort_flag);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_potrf_leaf(hipStream_t s, int b, double *A, long lda, int offset, int *info) {
  if (b <= 0) return PLA_SUCCESS;
  if (b > TRSM_LEAF) return fail(PLA_ERR_INVALID_PARAM, "potrf leaf: b = %d", b);
  hipLaunchKernelGGL(potrf_leaf_kernel, dim3(1), dim3(64), 0, s, b, A, lda, offset, info);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

// grid of the API transfer kernels: enough row blocks to cover M (at most
// 64), columns over y (at most 65535, grid-stride beyond)
static dim3 api_grid(int M, int N) {
  const int gx = (M + 255) / 256 < 64 ? (M + 255) / 256 : 64;
  const int gy = N < 65535 ? N : 65535;
  return dim3(gx > 0 ? gx : 1, gy > 0 ? gy : 1);
}

template <typename T>
static int api_scatter_t(hipStream_t s, int M, int N, const void *alpha, const void *src, void *dst, long ldd,
                         const int *rmap, const int *cmap, int r0, int c0) {
  const T a = *(const T *)alpha;
  hipLaunchKernelGGL((api_scatter_add_kernel<T>), api_grid(M, N), dim3(256), 0, s, M, N, a, (const T *)src,
                     (T *)dst, ldd, rmap, cmap, r0, c0);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

template <typename T>
static int api_gather_t(hipStream_t s, int M, int N, const void *src, long lds_, void *dst, const int *rmap,
                        const int *cmap, int r0, int c0) {
  hipLaunchKernelGGL((api_gather_kernel<T>), api_grid(M, N), dim3(256), 0, s, M, N, (const T *)src, lds_, (T *)dst,
                     rmap, cmap, r0, c0);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_api_scatter_add(hipStream_t s, int elem, int M, int N, const void *alpha, const void *src, void *dst, long ldd,
                      const int *rmap, const int *cmap, int r0, int c0) {
  if (M <= 0 || N <= 0) return PLA_SUCCESS;
  switch (elem) {
    case API_F64: return api_scatter_t<double>(s, M, N, alpha, src, dst, ldd, rmap, cmap, r0, c0);
    case API_F32: return api_scatter_t<float>(s, M, N, alpha, src, dst, ldd, rmap, cmap, r0, c0);
    case API_I32: return api_scatter_t<int>(s, M, N, alpha, src, dst, ldd, rmap, cmap, r0, c0);
    case API_Z64: return api_scatter_t<api_z>(s, M, N, alpha, src, dst, ldd, rmap, cmap, r0, c0);
  }
  return fail(PLA_ERR_UNSUPPORTED, "api scatter: element kind %d", elem);
}

int k_api_gather(hipStream_t s, int elem, int M, int N, const void *src, long lds_, void *dst, const int *rmap,
                 const int *cmap, int r0, int c0) {
  if (M <= 0 || N <= 0) return PLA_SUCCESS;
  switch (elem) {
    case API_F64: return api_gather_t<double>(s, M, N, src, lds_, dst, rmap, cmap, r0, c0);
    case API_F32: return api_gather_t<float>(s, M, N, src, lds_, dst, rmap, cmap, r0, c0);
    case API_I32: return api_gather_t<int>(s, M, N, src, lds_, dst, rmap, cmap, r0, c0);
    case API_Z64: return api_gather_t<api_z>(s, M, N, src, lds_, dst, rmap, cmap, r0, c0);
  }
  return fail(PLA_ERR_UNSUPPORTED, "api gather: element kind %d", elem);
}

int k_copy2d(hipStream_t s, int m, int n, const double *S, long lds_, double *D, long ldd,
             const int *abort_flag) {
  if (m <= 0 || n <= 0) return PLA_SUCCESS;
  hipLaunchKernelGGL(copy2d_kernel, dim3(grid1d((long)m * n, 256 * 4)), dim3(256), 0, s, m, n, S, lds_,
                     D, ldd, abort_flag);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_scale2d(hipStream_t s, int m, int n, double alpha, double *D, long ldd) {
  if (m <= 0 || n <= 0 || alpha == 1.0) return PLA_SUCCESS;
  hipLaunchKernelGGL(scale2d_kernel, dim3(grid1d((long)m * n, 256 * 4)), dim3(256), 0, s, m, n, alpha,
                     D, ldd);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_norm(hipStream_t s, bool inf_norm, int m, int n, const double *A, long lda, double *result) {
  *result = 0.0;
  if (m <= 0 || n <= 0) return PLA_SUCCESS;
  const int nch = (n + NORM_CHUNK - 1) / NORM_CHUNK;
  const size_t part_bytes = inf_norm ? sizeof(double) * (size_t)nch * m : 0;
  char *ws = (char *)workspace(part_bytes + 256, 2);
  if (!ws) return PLA_ERR_NO_MEMORY;
  unsigned long long *out = (unsigned long long *)(ws + part_bytes);
  PLA_HIP(hipMemsetAsync(out, 0, sizeof(unsigned long long), s));
  if (inf_norm) {
    hipLaunchKernelGGL(row_asum_partial_kernel, dim3((m + 255) / 256, nch), dim3(256), 0, s, m, n, A, lda,
                       (double *)ws);
    hipLaunchKernelGGL(row_asum_max_kernel, dim3((m + 255) / 256), dim3(256), 0, s, m, nch,
                       (const double *)ws, out);
  } else {
    hipLaunchKernelGGL(col_asum_max_kernel, dim3(n), dim3(256), 0, s, m, A, lda, out);
  }
  PLA_HIP(hipGetLastError());
  unsigned long long bits = 0;
  PLA_HIP(hipMemcpyAsync(&bits, out, sizeof bits, hipMemcpyDeviceToHost, s));
  PLA_HIP(hipStreamSynchronize(s));
  *result = __builtin_bit_cast(double, bits);
  return PLA_SUCCESS;
}

int k_fill_drand48(hipStream_t s, int m, int n, double *D, long ldd, uint64_t state0, uint64_t skip) {
  if (m <= 0 || n <= 0) return PLA_SUCCESS;
  const uint64_t total = (uint64_t)m * n;
  const uint64_t threads = (total + 63) / 64;
  const uint64_t blocks = (threads + 255) / 256;
  if (blocks > 0x7fffffffULL) return fail(PLA_ERR_DIMENSIONS, "fill: too large");
  hipLaunchKernelGGL(fill_drand48_kernel, dim3((unsigned)blocks), dim3(256), 0, s, m, n, D, ldd, state0,
                     skip);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

// Panel leaf: chooses the grid (about 256 rows per workgroup, at most
// LEAF_MAXG workgroups, all co-resident) and draws fresh exchange tags.

// Workgroups of a leaf over mp rows: about leaf_rows rows each, at most
// LEAF_MAXG, the CU count and gcap.  Every workgroup must be resident at once
// (they exchange pivot candidates every column): with more than 80 KiB of LDS
// each, one fits per CU, so G must not exceed the CUs the stream may use.
static int leaf_grid(int mp, int gcap) {
  int maxG = ctx().cu_count > 0 ? ctx().cu_count : 256;
  if (maxG > LEAF_MAXG) maxG = LEAF_MAXG;
  if (gcap > 0 && maxG > gcap) maxG = gcap;
  const int target = ctx().leaf_rows > 0 ? ctx().leaf_rows : 256;
  int G = (mp + target - 1) / target;
  if (G > maxG) G = maxG;
  if (G < 1) G = 1;
  return G;
}
constexpr size_t LEAF_LDS_MAX = 150 * 1024;
static size_t leaf_lds_bytes(int R, int w) { return (size_t)w * R * sizeof(double) + (size_t)R * sizeof(int); }

int leaf_width_for(int mp, int gcap) {
  const int G = leaf_grid(mp, gcap);
  const int R = (mp + G - 1) / G;
  for (int w = LEAF_W; w >= 8; w /= 2)
    if (leaf_lds_bytes(R, w) <= LEAF_LDS_MAX) return w;
  return 0;
}

int k_lu_leaf(hipStream_t s, const lu_leaf_call &c) {
  if (ctx().leaf_kind == 0) return k_lu_rleaf(s, c, c.W);
  if (c.mp <= 0 || c.w <= 0) return PLA_SUCCESS;
  if (c.w > LEAF_W || c.mp < c.w) return fail(PLA_ERR_INVALID_PARAM, "lu leaf: %d x %d", c.mp, c.w);
  const int G = leaf_grid(c.mp, c.gcap);
  const int R = (c.mp + G - 1) / G;
  const size_t lds_bytes = leaf_lds_bytes(R, c.w);
  if (lds_bytes > LEAF_LDS_MAX)
    return fail(PLA_ERR_DIMENSIONS, "lu leaf: panel of %d rows exceeds the LDS budget", c.mp);
  leaf_ws &ws = leaf_workspace((size_t)2 * LEAF_MAXG * SLOT_WORDS * sizeof(unsigned long long));
  if (!ws.base) return PLA_ERR_NO_MEMORY;
  unsigned tag0;
  int rc = leaf_tags(s, (unsigned)c.w, &tag0);
  if (rc) return rc;
  leaf_args a;
  a.P = c.P;
  a.Ppanel = c.Ppanel;
  a.lda = c.lda;
  a.mp = c.mp;
  a.w = c.w;
  a.pw = c.pw;
  a.pc0 = c.pc0;
  a.R = R;
  a.piv = c.piv;
  a.pairs = c.pairs;
  a.slots = ws.slots;
  a.err = ws.err;
  a.state = ws.state;
  a.fscr = (char *)workspace(leaffb::scratch_bytes(c.mp), WS_LEAF_FB);
  if (!a.fscr) return PLA_ERR_NO_MEMORY;
  a.tag0 = tag0;
  a.spin_limit = ctx().leaf_spin_limit;
  static bool lds_attr = false;  // allow more than 64 KiB of dynamic LDS (gfx950 has 160 KiB)
  if (!lds_attr) {
    PLA_HIP(hipFuncSetAttribute((const void *)lu_leaf_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                150 * 1024));
    lds_attr = true;
  }
  hipLaunchKernelGGL(lu_leaf_kernel, dim3(G), dim3(LEAF_THREADS), lds_bytes, s, a);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_panel_compose(hipStream_t s, const int *pairs, int nleaves, int *comp) {
  if (nleaves <= 0) return PLA_SUCCESS;
  if (nleaves > LU_MAX_LEAVES) return fail(PLA_ERR_INVALID_PARAM, "panel compose: %d leaves", nleaves);
  int *err = leaf_error_word();
  if (!err) return PLA_ERR_NO_MEMORY;
  hipLaunchKernelGGL(panel_compose_kernel, dim3(1), dim3(256), 0, s, pairs, nleaves, comp, (const int *)err);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

template <typename T>
static int panel_swap_launch(hipStream_t s, T *A, long lda, int nleft, int c1, int c2, const int *comp) {
  const int total = nleft + (c2 > c1 ? c2 - c1 : 0);
  if (total <= 0) return PLA_SUCCESS;
  int blocks = (total + 3) / 4;
  const int cap = 16 * (ctx().cu_count > 0 ? ctx().cu_count : 256);
  if (blocks > cap) blocks = cap;
  int *err = leaf_error_word();
  if (!err) return PLA_ERR_NO_MEMORY;
  timing_begin(s, 7);
  hipLaunchKernelGGL(panel_swap_kernel<T>, dim3(blocks), dim3(256), 0, s, A, lda, nleft, c1, c2 > c1 ? c2 : c1,
                     comp, (const int *)err);
  // the rows moved are known only on the device: the columns swapped are the unit
  timing_end(s, 7, (double)total);
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

int k_panel_swap(hipStream_t s, double *A, long lda, int nleft, int c1, int c2, const int *comp) {
  return panel_swap_launch<double>(s, A, lda, nleft, c1, c2, comp);
}

int k_panel_swap_f32(hipStream_t s, float *A, long lda, int nleft, int c1, int c2, const int *comp) {
  return panel_swap_launch<float>(s, A, lda, nleft, c1, c2, comp);
}

int panel_comp_stride() { return 1 + 2 * COMP_MAX; }

int leaf_pairs_stride() { return PAIRS_STRIDE > rleaf_pairs_stride() ? PAIRS_STRIDE : rleaf_pairs_stride(); }
int leaf_max_width() { return LEAF_W; }
int trsm_leaf_size() { return TRSM_LEAF; }

}  // namespace pla
