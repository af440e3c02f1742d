// pla_leaf_fallback.hpp -- what the panel-LU leaves (pla_leaf.hip, pla_kernels.hip)
// do when their cross-workgroup exchange gives up.
//
// A leaf's participants (waves or workgroups) must all be resident at once.
// When one of them cannot become resident in time (a bounded wait gives up),
// the launch is abandoned instead of failing: every participant of a launch
// decides, through one state word, whether the launch is committed (a
// participant finished all columns: the records exist and every waiter will
// see them) or abandoned (a wait gave up first).  No participant writes the
// matrix before the last column, so an abandoned panel is untouched; then
// workgroup 0 alone -- always resident with itself -- factors the leaf from
// memory with the same arithmetic in the same order (PLA_LU_mv.c:52-82:
// iamax with ties to the first position, interchange, reciprocal scale,
// rank-1 update), and produces the same pivots, factors, row pairs and
// other-column interchanges as the exchange would have.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>

namespace leaffb {

// The launch's state word: (tag0 << 32) | 1 committed, | 2 abandoned; a word
// carrying another tag is undecided.
__device__ inline int state(const unsigned long long *st, unsigned tag0) {
  const unsigned long long cur = __hip_atomic_load(st, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return (unsigned)(cur >> 32) == tag0 ? (int)(cur & 3ULL) : 0;
}

// Decide `want` (1 or 2) unless already decided; returns the decided state.
__device__ inline int decide(unsigned long long *st, unsigned tag0, int want) {
  unsigned long long cur = __hip_atomic_load(st, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (;;) {
    if ((unsigned)(cur >> 32) == tag0) return (int)(cur & 3ULL);
    const unsigned long long nv = ((unsigned long long)tag0 << 32) | (unsigned long long)want;
    const unsigned long long prev = atomicCAS(st, cur, nv);
    if (prev == cur) return want;
    cur = prev;
  }
}

// Bytes of the fallback's scratch for a leaf of mp rows.
constexpr int MAXW = 32;
__host__ __device__ inline size_t scratch_bytes(int mp) {
  return sizeof(double) * (4 + MAXW + 2 * MAXW * MAXW) + sizeof(int) * (256 + (size_t)mp);
}

// (|v|, position) order of the pivot search: larger value, then smaller position.
__device__ inline bool better(double v, int p, double bv, int bp) { return v > bv || (v == bv && p < bp); }

// Replay of the interchanges p[0..w) (p[j]: the position swapped with j) on a
// position -> source-row map held one entry per lane, entries created in
// order of first use; returns, in lane k < *npairs, the k-th moved entry as
// (destination, source) positions.  One wave.
__device__ inline void replay_pairs(const int *p, int w, int *dst, int *src, int *npairs) {
  const int lane = threadIdx.x & 63;
  int key = -1, val = -1, cnt = 0;
  for (int j = 0; j < w; j++) {
    const int pj = p[j];
    if (pj == j) continue;
    unsigned long long mj = __ballot(key == j);
    if (!mj) {
      if (lane == cnt) {
        key = j;
        val = j;
      }
      mj = 1ULL << cnt;
      cnt++;
    }
    unsigned long long mq = __ballot(key == pj);
    if (!mq) {
      if (lane == cnt) {
        key = pj;
        val = pj;
      }
      mq = 1ULL << cnt;
      cnt++;
    }
    const int lj = __ffsll((long long)mj) - 1, lq = __ffsll((long long)mq) - 1;
    const int vj = __shfl(val, lj, 64), vq = __shfl(val, lq, 64);
    if (lane == lj) val = vq;
    if (lane == lq) val = vj;
  }
  const bool moved = (key >= 0 && key != val);
  const unsigned long long mm = __ballot(moved);
  *npairs = __popcll(mm);
  *dst = -1;
  *src = -1;
  unsigned long long rest = mm;
  for (int k = 0; k < *npairs; k++) {
    const int from = __ffsll((long long)rest) - 1;
    rest &= rest - 1;
    const int kd = __shfl(key, from, 64), ks = __shfl(val, from, 64);
    if (lane == k) {
      *dst = kd;
      *src = ks;
    }
  }
}

// The leaf of w <= 32 columns and mp rows at P (column-major, lda) factored by
// one workgroup of 256 threads straight from memory.  Ppanel: panel column 0
// at the leaf's first row (pw columns, the leaf at pc0); piv: w pivot offsets;
// pairs: count, then (dst + pc0, src + pc0); scr: scratch_bytes(mp) of device
// memory.  Uses no LDS (the register leaf has none).
__device__ inline void factor(double *P, double *Ppanel, long lda, int mp, int w, int pw, int pc0, int *piv,
                              int *pairs, char *scr) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double *red_v = (double *)scr;           // [4]
  double *prow = red_v + 4;                // [MAXW]
  double *rowbuf = prow + MAXW;            // [2 MAXW][MAXW]
  int *ibase = (int *)(rowbuf + 2 * MAXW * MAXW);
  int *red_p = ibase, *red_r = ibase + 4;  // [4] each
  int *ph = ibase + 8;                     // [MAXW]
  int *mrow = ibase + 8 + MAXW;            // [2 MAXW] rows whose position changed
  int *mcnt = ibase + 8 + 3 * MAXW;        // [1]
  int *pd = mcnt + 1, *ps = pd + 2 * MAXW; // [2 MAXW] each: the (dst, src) pairs
  int *pos = ibase + 256;                  // [mp]
  for (int i = tid; i < mp; i += 256) pos[i] = i;
  if (tid == 0) *mcnt = 0;
  __syncthreads();
  for (int j = 0; j < w; j++) {
    double bv = -1.0;
    int bp = INT_MAX, br = -1;
    for (int i = tid; i < mp; i += 256) {
      const int q = pos[i];
      if (q >= j) {
        const double x = fabs(P[(long)j * lda + i]);
        if (better(x, q, bv, bp)) {
          bv = x;
          bp = q;
          br = i;
        }
      }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const double ov = __shfl_xor(bv, off, 64);
      const int op = __shfl_xor(bp, off, 64), orr = __shfl_xor(br, off, 64);
      if (better(ov, op, bv, bp)) {
        bv = ov;
        bp = op;
        br = orr;
      }
    }
    if (lane == 0) {
      red_v[wave] = bv;
      red_p[wave] = bp;
      red_r[wave] = br;
    }
    __syncthreads();
    double gv = red_v[0];
    int gp = red_p[0], gr = red_r[0];
    for (int v = 1; v < 4; v++)
      if (better(red_v[v], red_p[v], gv, gp)) {
        gv = red_v[v];
        gp = red_p[v];
        gr = red_r[v];
      }
    // the winner's row (all-NaN column: no winner, no interchange)
    const int p = gr >= 0 ? gp : j;
    if (tid < w && tid >= j)
      prow[tid] = gr >= 0 ? P[(long)tid * lda + gr] : __longlong_as_double(0x7FF8000000000000LL);
    if (tid == 0) {
      piv[j] = p - j;
      ph[j] = p;
    }
    __syncthreads();
    const double inv = 1.0 / prow[j];
    for (int i = tid; i < mp; i += 256) {
      int q = pos[i];
      if (q == p) q = j;
      else if (q == j) q = p;
      pos[i] = q;
      if (q > j) {
        const double l = P[(long)j * lda + i] * inv;
        P[(long)j * lda + i] = l;
        for (int c = j + 1; c < w; c++) P[(long)c * lda + i] = fma(l, -prow[c], P[(long)c * lda + i]);
      }
    }
    __syncthreads();
  }
  // rows to their final positions: the rows whose position changed (at most
  // 2w) through the scratch row buffer
  for (int i = tid; i < mp; i += 256)
    if (pos[i] != i) mrow[atomicAdd(mcnt, 1)] = i;
  __syncthreads();
  const int cnt = *mcnt;
  for (int e = tid; e < cnt * w; e += 256) {
    const int k = e % cnt, c = e / cnt;
    rowbuf[k * MAXW + c] = P[(long)c * lda + mrow[k]];
  }
  __syncthreads();
  for (int e = tid; e < cnt * w; e += 256) {
    const int k = e % cnt, c = e / cnt;
    P[(long)c * lda + pos[mrow[k]]] = rowbuf[k * MAXW + c];
  }
  __syncthreads();
  // the pair list, by the replay every leaf uses
  int dst = -1, src = -1, npairs = 0;
  if (wave == 0) {
    replay_pairs(ph, w, &dst, &src, &npairs);
    if (lane < npairs) {
      pairs[1 + 2 * lane] = dst + pc0;
      pairs[2 + 2 * lane] = src + pc0;
    }
    if (lane == 0) {
      pairs[0] = npairs;
      *mcnt = npairs;
    }
    // the pairs themselves, for the other waves
    if (lane < npairs) {
      pd[lane] = dst;
      ps[lane] = src;
    }
  }
  __syncthreads();
  npairs = *mcnt;
  // the same interchanges on the panel's other columns, one column per wave:
  // every lane loads its source before any lane stores
  const int others = pw - w;
  if (npairs > 0) {
    const int d = lane < npairs ? pd[lane] : 0, sr = lane < npairs ? ps[lane] : 0;
    for (int q = wave; q < others; q += 4) {
      const int cc = q < pc0 ? q : q + w;
      double *col = Ppanel + (long)cc * lda;
      if (lane < npairs) {
        const double x = __hip_atomic_load(col + sr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        col[d] = x;
      }
    }
  }
}

}  // namespace leaffb
