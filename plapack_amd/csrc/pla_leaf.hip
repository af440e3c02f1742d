// pla_leaf.hip -- the unblocked panel-LU leaf with its rows in registers.
//
// Replaces PLA_LU_right_mv (SOLVERS/LU/PLA_LU_mv.c:52-82) for a leaf of up to
// W columns of the panel: per column PLA_Iamax (BLAS1/PLA_Iamax.c:37-124)
// becomes a wave reduction plus an exchange of (value, position, row)
// records between waves, Pivot_rows_mv (PLA_Apply_pivots_new.c:208-305)
// a relabelling of logical row positions, PLA_Local_inv_scal
// (Local_BLAS/PLA_Local_BLAS1.c:258) one reciprocal and a multiply, and
// PLA_Local_ger (Local_BLAS/PLA_Local_BLAS2.c:578-656) a rank-1 update.
//
// Layout.  Each wave of 64 lanes is one participant and uses no LDS; lane l
// keeps rows r0 + l + 64 i (i < RPL) of the wave's share of the leaf, all W
// columns, in registers (RPL * W = 64 doubles per lane, about 160 VGPRs).
// Waves are launched four to a workgroup (normally one per SIMD of a CU), so a leaf of
// G waves occupies G / 4 CUs.  Beside the FP16x3 trailing update (two
// workgroups of 4 waves of 156 VGPRs and 80 KB of LDS per CU) a leaf
// workgroup fits without displacing it; beside the FP64 update (256 VGPRs per
// wave, two waves per SIMD) it takes the place of one update workgroup on
// each of its G / 4 CUs, not on G CUs as single-wave workgroups spread over
// the chip would.  All waves of one leaf must be resident at once (they
// exchange pivot candidates every column); when a bounded wait gives up
// first, the launch is abandoned and workgroup 0 factors the leaf alone
// (pla_leaf_fallback.hpp).
//
// Per column j (rows are never moved; each carries its logical position,
// initially its row number; positions < j hold the pivot rows chosen so far):
//   1. candidate: the largest |L(r, j)| over rows at positions >= j, ties to
//      the smallest position (idamax's first maximum, and the tie rule of
//      PLA_Iamax.c:108-115), reduced over the wave;
//   2. publish (position, value) in three dense header arrays (one word per
//      workgroup each, so a poll of all workgroups reads 3 G contiguous
//      words) and the candidate row, columns j.. only, stored straight from
//      the lane that holds it;
//   3. read every workgroup's (position, value), pick the same winner, then
//      read the winner's row;
//   4. relabel (pivot row -> j, the row at j -> p), scale by the reciprocal
//      of the pivot, rank-1 update; then the candidate of column j+1 is
//      published (its row must carry this step's update in every column).
// Exchange words carry the column's 32-bit tag beside 32 bits of payload;
// readers re-read until the tags match (no fences, no flags).  Records are
// double buffered by column parity.  A wave that finishes every column
// commits the launch; a bounded wait that gives up before that abandons it,
// and every wave then stops without writing the matrix.
//
// At the end rows are written to their final positions, the interchanges are
// replayed into at most 2w (destination, source) row pairs, and the pairs are
// applied to the panel's other columns, so the recursive panel needs no
// separate row-interchange launches.
#include "pla_internal.hpp"
#include "pla_leaf_fallback.hpp"

#include <climits>
#include <cstdint>
#include <utility>

namespace {

constexpr int RL_MAXG = 512;     // waves of one leaf
constexpr int RL_WG_WAVES = 4;   // waves per workgroup (one per SIMD of a CU)
constexpr int RL_MAXW = 32;
constexpr int RL_ROWW = 2 * RL_MAXW;                    // words of one candidate row
constexpr int RL_PAR = 3 * RL_MAXG + RL_MAXG * RL_ROWW;  // words of one column parity
constexpr int RL_PAIRS = 2 * RL_MAXW;
constexpr int RL_PAIRS_STRIDE = 1 + 2 * RL_PAIRS;  // count, then (dst, src) pairs

struct rleaf_args {
  double *P;                // leaf top-left (its row 0 is the leaf's first diagonal row)
  double *Ppanel;           // panel column 0, at the leaf's first row
  long lda;
  int mp, w;                // rows from the leaf's first row to the panel bottom, leaf width (<= W)
  int pw, pc0;              // panel width, leaf's first column inside the panel
  int *piv;                 // w entries: pivot offset relative to row j
  int *pairs;               // RL_PAIRS_STRIDE ints, positions relative to the leaf's first row + pc0
  unsigned long long *slots;  // per column parity: headers [3][RL_MAXG], rows [RL_MAXG][RL_ROWW]
  int *err;
  unsigned long long *state;  // the launch's committed / abandoned word
  char *fscr;                 // workgroup 0's scratch if the launch is abandoned
  unsigned tag0, spin_limit;
};

__device__ inline void st_dev(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned long long ld_dev(const unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int ld_dev(const int *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline double ld_dev(const double *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned lo32(double v) { return (unsigned)(unsigned long long)__double_as_longlong(v); }
__device__ inline unsigned hi32(double v) { return (unsigned)((unsigned long long)__double_as_longlong(v) >> 32); }
__device__ inline double dbl_from(unsigned lo, unsigned hi) {
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// f(integral_constant<int, 0>) ... f(integral_constant<int, N-1>): loop bodies
// with compile-time indices, so register arrays are never indexed at run time.
template <typename F, int... Js>
__device__ inline void static_for_impl(F &&f, std::integer_sequence<int, Js...>) {
  (f(std::integral_constant<int, Js>{}), ...);
}
template <int N, typename F>
__device__ inline void static_for(F &&f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// (|v|, position) order of the pivot search: larger value, then smaller position.
__device__ inline bool better(double v, int p, double bv, int bp) { return v > bv || (v == bv && p < bp); }

// Wave-wide argmax of (value, position) carrying a third int; valid in every lane.
__device__ inline void wave_argmax(double &bv, int &bp, int &bx) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double ov = __shfl_xor(bv, off, 64);
    const int op = __shfl_xor(bp, off, 64);
    const int ox = __shfl_xor(bx, off, 64);
    if (better(ov, op, bv, bp)) {
      bv = ov;
      bp = op;
      bx = ox;
    }
  }
}

// Bounded wait (the counter, and so every decision, is wave-uniform).
// Returns true when the launch is abandoned: this wait gave up first, or
// another wave's did.  A wait that gives up after another wave committed the
// launch keeps waiting: the records it waits for exist.
__device__ inline bool spin_fail(unsigned &it, const rleaf_args &a) {
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

template <int W, int RPL>
__device__ inline void leaf_finish(const rleaf_args &a, double (&v)[RPL][W], const int (&pos)[RPL], int ph, int g,
                                   int G, int lane);

template <int W>
__global__ __launch_bounds__(256, 2) void lu_rleaf_kernel(rleaf_args a) {
  constexpr int RPL = 64 / W;  // rows per lane
  constexpr int R = 64 * RPL;  // rows per wave
  // G waves in workgroups of RL_WG_WAVES; wave g is one participant
  const int G = gridDim.x * RL_WG_WAVES, g = blockIdx.x * RL_WG_WAVES + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int w = a.w, mp = a.mp;
  const long lda = a.lda;
  if (ld_dev(a.err)) {  // a bounded wait gave up earlier on this stream (the error word): no interchanges
    if (g == 0 && lane == 0) a.pairs[0] = 0;
    return;
  }

  double v[RPL][W];
  int pos[RPL];
#pragma unroll
  for (int i = 0; i < RPL; i++) {
    const int r = g * R + lane + 64 * i;
    pos[i] = r < mp ? r : -1;
  }
#pragma unroll
  for (int c = 0; c < W; c++)
#pragma unroll
    for (int i = 0; i < RPL; i++)
      v[i][c] = (c < w && pos[i] >= 0) ? a.P[(long)c * lda + pos[i]] : 0.0;

  unsigned long long *slots = a.slots;

  // candidate of column j over this lane's rows, reduced over the wave,
  // then published with the candidate row (j is a compile-time constant)
  auto publish = [&](auto J) {
    constexpr int j = decltype(J)::value;
    double bv = -1.0;
    int bp = INT_MAX, bx = -1;
#pragma unroll
    for (int i = 0; i < RPL; i++) {
      const double x = fabs(v[i][j]);
      if (pos[i] >= j && better(x, pos[i], bv, bp)) {
        bv = x;
        bp = pos[i];
        bx = i;
      }
    }
    int who = bx >= 0 ? lane * RPL + bx : -1;  // lane and row slot of the candidate
    wave_argmax(bv, bp, who);
    const unsigned long long tag = (unsigned long long)(a.tag0 + (unsigned)j) << 32;
    unsigned long long *par = slots + (size_t)(j & 1) * RL_PAR;
    // the candidate row, columns j.. (words 2c, 2c + 1: lo, hi of column c),
    // stored by the lane that holds it
    // (the slot is found by its position, one test per slot, so that every
    // register index stays a constant)
    if (who >= 0 && lane == who / RPL) {
      unsigned long long *row = par + 3 * RL_MAXG + (size_t)g * RL_ROWW;
#pragma unroll
      for (int i = 0; i < RPL; i++) {
        if (pos[i] == bp) {
#pragma unroll
          for (int c = j; c < W; c++) {
            st_dev(row + 2 * c, tag | lo32(v[i][c]));
            st_dev(row + 2 * c + 1, tag | hi32(v[i][c]));
          }
        }
      }
    }
    if (lane < 3) st_dev(par + lane * RL_MAXG + g, tag | (lane == 0 ? (unsigned)bp : lane == 1 ? lo32(bv) : hi32(bv)));
  };

  publish(std::integral_constant<int, 0>{});
  int ph = 0;         // lane j: the pivot position chosen for column j
  bool dead = false;  // the launch was abandoned: stop, write nothing
  static_for<W>([&](auto J) {
    constexpr int j = decltype(J)::value;
    if (j < w && !dead) {
      const unsigned tagv = a.tag0 + (unsigned)j;
      const unsigned long long *par = slots + (size_t)(j & 1) * RL_PAR;
      // every workgroup's (position, value)
      double gv = -1.0;
      int gp = INT_MAX, gw = -1;
      {
        unsigned it = 0;
        for (;;) {
          int ok = 1;
          gv = -1.0;
          gp = INT_MAX;
          gw = -1;
          for (int t = lane; t < G; t += 64) {
            const unsigned long long w0 = ld_dev(par + t), w1 = ld_dev(par + RL_MAXG + t),
                                     w2 = ld_dev(par + 2 * RL_MAXG + t);
            if ((unsigned)(w0 >> 32) != tagv || (unsigned)(w1 >> 32) != tagv || (unsigned)(w2 >> 32) != tagv) ok = 0;
            const int ci = (int)(unsigned)w0;
            const double cv = dbl_from((unsigned)w1, (unsigned)w2);
            if (ci >= j && ci < mp && better(cv, ci, gv, gp)) {
              gv = cv;
              gp = ci;
              gw = t;
            }
          }
          if (__all(ok)) break;
          if (spin_fail(it, a)) {
            dead = true;
            break;
          }
        }
        wave_argmax(gv, gp, gw);
      }
      if (dead) return;
      // the winner's row (all-NaN column: no winner, no interchange); lane l in
      // [2j, 2W) holds row word l, so column c is in lanes 2c (low) and 2c + 1 (high)
      unsigned rw = 0u;
      {
        const unsigned long long *rt = par + 3 * RL_MAXG + (size_t)(gw >= 0 ? gw : 0) * RL_ROWW;
        unsigned it = 0;
        for (;;) {
          int ok = 1;
          if (gw >= 0 && lane >= 2 * j && lane < 2 * W) {
            const unsigned long long x = ld_dev(rt + lane);
            if ((unsigned)(x >> 32) != tagv) ok = 0;
            rw = (unsigned)x;
          }
          if (__all(ok)) break;
          if (spin_fail(it, a)) {
            dead = true;
            break;
          }
        }
      }
      if (dead) return;
      // wave-uniform element c of the pivot row
      auto prow = [&](int c) -> double {
        if (gw < 0) return __longlong_as_double(0x7FF8000000000000LL);
        return dbl_from((unsigned)__builtin_amdgcn_readlane((int)rw, 2 * c),
                        (unsigned)__builtin_amdgcn_readlane((int)rw, 2 * c + 1));
      };
      const int p = gw >= 0 ? gp : j;
      if (lane == j) ph = p;
      if (g == 0 && lane == 0) a.piv[j] = p - j;
      const double inv = 1.0 / prow(j);
#pragma unroll
      for (int i = 0; i < RPL; i++) {
        int q = pos[i];
        if (q == p) q = j;
        else if (q == j) q = p;
        pos[i] = q;
        if (q > j) v[i][j] *= inv;
      }
      // rank-1 update, then the next column's candidate (its row must carry
      // this step's update in every column)
#pragma unroll
      for (int c = 0; c < W; c++) {
        if (c > j && c < w) {
          const double pc = prow(c);
#pragma unroll
          for (int i = 0; i < RPL; i++)
            if (pos[i] > j) v[i][c] = fma(v[i][j], -pc, v[i][c]);
        }
      }
      if constexpr (j + 1 < W) {
        if (j + 1 < w) publish(std::integral_constant<int, j + 1>{});
      }
    }
  });

  // every wave decides the launch: committed (this wave saw every record) or
  // abandoned (some wait gave up first); the outcome is the same for all
  int outcome = 0;
  if (lane == 0) outcome = leaffb::decide(a.state, a.tag0, dead ? 2 : 1);
  outcome = __shfl(outcome, 0, 64);
  if (outcome == 1) leaf_finish(a, v, pos, ph, g, G, lane);
  if (blockIdx.x == 0) {  // workgroup 0 is always resident with itself
    __syncthreads();
    if (outcome == 2) {
      if (threadIdx.x == 0) atomicAdd(a.state + 1, 1ULL);  // fallbacks run, for the tests
      leaffb::factor(a.P, a.Ppanel, a.lda, a.mp, a.w, a.pw, a.pc0, a.piv, a.pairs, a.fscr);
    }
  }
}

// The committed launch's end: rows to their final positions (a permutation:
// no two rows collide), the interchanges replayed into (dst, src) pairs, and
// the pairs applied to the panel's other columns.
template <int W, int RPL>
__device__ inline void leaf_finish(const rleaf_args &a, double (&v)[RPL][W], const int (&pos)[RPL], int ph, int g,
                                   int G, int lane) {
  const int w = a.w;
  const long lda = a.lda;
#pragma unroll
  for (int c = 0; c < W; c++)
#pragma unroll
    for (int i = 0; i < RPL; i++)
      if (c < w && pos[i] >= 0) a.P[(long)c * lda + pos[i]] = v[i][c];

  // replay the interchanges on (position -> source row) entries held in lanes
  int key = -1, val = -1, cnt = 0;
  for (int j = 0; j < w; j++) {
    {
      const int p = __shfl(ph, j, 64);
      if (p != j) {
        unsigned long long mj = __ballot(key == j);
        if (!mj) {
          if (lane == cnt) { key = j; val = j; }
          mj = 1ULL << cnt;
          cnt++;
        }
        unsigned long long mq = __ballot(key == p);
        if (!mq) {
          if (lane == cnt) { key = p; val = p; }
          mq = 1ULL << cnt;
          cnt++;
        }
        const int lj = __ffsll((long long)mj) - 1, lq = __ffsll((long long)mq) - 1;
        const int vj = __shfl(val, lj, 64), vq = __shfl(val, lq, 64);
        if (lane == lj) val = vq;
        if (lane == lq) val = vj;
      }
    }
  }
  const bool moved = (key >= 0 && key != val);
  const unsigned long long mm = __ballot(moved);
  const int npairs = __popcll(mm);
  // compact the moved entries: lane k ends up with the k-th (dst, src) pair
  int dst = -1, src = -1;
  {
    unsigned long long rest = mm;
    for (int k = 0; k < npairs; k++) {
      const int from = __ffsll((long long)rest) - 1;
      rest &= rest - 1;
      const int kd = __shfl(key, from, 64), ks = __shfl(val, from, 64);
      if (lane == k) {
        dst = kd;
        src = ks;
      }
    }
  }
  if (g == 0) {
    if (lane < npairs) {
      a.pairs[1 + 2 * lane] = dst + a.pc0;
      a.pairs[2 + 2 * lane] = src + a.pc0;
    }
    if (lane == 0) a.pairs[0] = npairs;
  }

  // the same interchanges on the panel's other columns, one column per workgroup
  const int others = a.pw - w;
  if (npairs > 0) {
    for (int q = g; q < others; q += G) {
      const int cc = q < a.pc0 ? q : q + w;
      double *col = a.Ppanel + (long)cc * lda;
      if (lane < npairs) {
        const double x = ld_dev(col + src);
        col[dst] = x;
      }
    }
  }
}

}  // namespace

namespace pla {

int rleaf_pairs_stride() { return RL_PAIRS_STRIDE; }

// Rows per workgroup of the register leaf of width W (64 doubles per lane).
static int rleaf_rows(int W) { return 64 * (64 / W); }

int rleaf_width_for(int mp, int gcap) {
  int maxg = RL_MAXG;
  if (gcap > 0 && gcap < maxg) maxg = gcap;
  // widest leaf within 256 waves, else the widest within the cap
  for (int lim : {256, maxg})
    for (int W = RL_MAXW; W >= 8; W /= 2)
      if ((mp + rleaf_rows(W) - 1) / rleaf_rows(W) <= (lim < maxg ? lim : maxg)) return W;
  return 0;
}

int k_lu_rleaf(hipStream_t s, const lu_leaf_call &c, int W) {
  if (c.mp <= 0 || c.w <= 0) return PLA_SUCCESS;
  if ((W != 8 && W != 16 && W != 32) || c.w > W || c.mp < c.w)
    return fail(PLA_ERR_INVALID_PARAM, "lu leaf: %d x %d in a width-%d leaf", c.mp, c.w, W);
  const int R = rleaf_rows(W);
  // whole workgroups of waves; the waves past the last row hold no rows
  const int G = ((c.mp + R - 1) / R + RL_WG_WAVES - 1) / RL_WG_WAVES * RL_WG_WAVES;
  if (G > RL_MAXG || (c.gcap > 0 && G > c.gcap))
    return fail(PLA_ERR_DIMENSIONS, "lu leaf: %d rows need %d waves of width %d", c.mp, G, W);
  const dim3 grid(G / RL_WG_WAVES), block(64 * RL_WG_WAVES);
  const size_t slot_bytes = (size_t)2 * RL_PAR * sizeof(unsigned long long);
  leaf_ws &ws = leaf_workspace(slot_bytes);
  if (!ws.base) return PLA_ERR_NO_MEMORY;
  unsigned tag0;
  int rc = leaf_tags(s, (unsigned)c.w, &tag0);
  if (rc) return rc;
  rleaf_args a;
  a.P = c.P;
  a.Ppanel = c.Ppanel;
  a.lda = c.lda;
  a.mp = c.mp;
  a.w = c.w;
  a.pw = c.pw;
  a.pc0 = c.pc0;
  a.piv = c.piv;
  a.pairs = c.pairs;
  a.slots = ws.slots;
  a.err = ws.err;
  a.state = ws.state;
  a.fscr = (char *)workspace(leaffb::scratch_bytes(c.mp), WS_LEAF_FB);
  if (!a.fscr) return PLA_ERR_NO_MEMORY;
  a.tag0 = tag0;
  a.spin_limit = ctx().leaf_spin_limit;
  switch (W) {
    case 8: hipLaunchKernelGGL(lu_rleaf_kernel<8>, grid, block, 0, s, a); break;
    case 16: hipLaunchKernelGGL(lu_rleaf_kernel<16>, grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL(lu_rleaf_kernel<32>, grid, block, 0, s, a); break;
  }
  PLA_HIP(hipGetLastError());
  return PLA_SUCCESS;
}

}  // namespace pla
