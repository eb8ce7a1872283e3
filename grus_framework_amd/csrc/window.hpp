// window.hpp -- row reduction over an in-edge CSR cut into fixed windows.
//
// The pull side of PageRank (sum of contributions) and of SSSP's large
// rounds (minimum of dist[u] + w) are the same computation: for every row v,
// reduce a per-edge value over the row's edges.  Rows have any length (RMAT
// hubs span hundreds of windows), so the edge array is cut into windows of
// WIN edges, one per workgroup, WIN_EPT consecutive edges per thread
// (merge-style CSR SpMV):
//   1. each thread computes the values of its WIN_EPT edges into registers
//      (two 16-byte column loads when the columns are in one piece);
//   2. the window's segments are staged in LDS: segment 0 is the head (the
//      part of a row that began in an earlier window, possibly empty), then
//      every non-empty row owned by the window in order; empty owned rows
//      are finished on the spot;
//   3. each thread finds the segment of its first edge by binary search and
//      walks its edges: segments that begin and end inside its range are
//      finished directly, the one it enters keeps its partial until the
//      carry is known, the one it leaves open is its scan value (and is
//      finished by the next thread when it ends exactly at that boundary);
//   4. a segmented inclusive scan over threads (keys = open segment, 64-lane
//      shuffles then the waves through LDS) gives every thread the reduction
//      of the open partials of the threads before it in the same segment.
// The head partial goes to op.head(k), the partial of a row that continues
// past the window to op.tail(k); window_fixup completes those rows.  The
// order of every reduction is fixed, so results are bit-identical run to
// run even for floating point.
//
// An Op provides: type T; identity(); combine(a, b); value(edge, column);
// finish(row, total); empty(row); head(k, t) / tail(k, t) and head_at(k) /
// tail_at(k).
#pragma once

#include "advance.hpp"

namespace mgg {

constexpr int WIN_EPT = 8;              // edges per thread
constexpr int WIN = WIN_EPT * BLOCK;    // edges per window

// win_row[k] = first row r with off[r] >= k*WIN (V if none), k in [0, nwin]
void build_window_rows(Context &ctx, const uint64_t *off, uint64_t V, uint64_t nwin,
                       uint32_t *win_row);

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// LDS the window reduction needs (kept out of the template so that callers
// can add their own shared arrays)
struct WindowShared {
  uint32_t start[WIN + 2];  // segment starts relative to the window's first edge
  uint32_t row[WIN + 2];    // row of segment j >= 1
  uint32_t cnt[BLOCK / WAVE];
  uint32_t wkey[BLOCK / WAVE];
  uint64_t wsum[BLOCK / WAVE];  // holds an Op::T of at most 8 bytes
};

template <class Op>
__device__ __forceinline__ void window_reduce(const CsrView &in, const uint32_t *win_row,
                                              uint64_t k, Op &op, WindowShared &sh) {
  using T = typename Op::T;
  static_assert(sizeof(T) <= sizeof(uint64_t), "partials are staged in 8-byte slots");
  constexpr int NW = BLOCK / WAVE;
  constexpr uint32_t NONE = 0xFFFFFFFFu;
  const uint64_t e0 = k * WIN;
  const uint64_t e1 = e0 + WIN < in.E ? e0 + WIN : in.E;
  const uint32_t len = (uint32_t)(e1 - e0);
  const uint32_t tid = threadIdx.x, lane = lane_id(), wid = tid / WAVE;
  const uint32_t p0 = tid * WIN_EPT;

  // 1. per-edge values
  T v[WIN_EPT];
  {
    const uint64_t g0 = e0 + p0;
    // a thread's WIN_EPT columns lie wholly on one side of the HBM / host
    // split (a multiple of 64) and are 32-byte aligned on either side
    if (p0 + WIN_EPT <= len) {
      const u32x4 *cp = reinterpret_cast<const u32x4 *>(
          g0 < in.split ? in.col + g0 : in.col_far + (g0 - in.split));
      const u32x4 c0 = __builtin_nontemporal_load(cp);
      const u32x4 c1 = __builtin_nontemporal_load(cp + 1);
      v[0] = op.value(g0 + 0, c0.x);
      v[1] = op.value(g0 + 1, c0.y);
      v[2] = op.value(g0 + 2, c0.z);
      v[3] = op.value(g0 + 3, c0.w);
      v[4] = op.value(g0 + 4, c1.x);
      v[5] = op.value(g0 + 5, c1.y);
      v[6] = op.value(g0 + 6, c1.z);
      v[7] = op.value(g0 + 7, c1.w);
    } else {
#pragma unroll
      for (int q = 0; q < WIN_EPT; q++) {
        v[q] = op.identity();
        if (p0 + q < len) v[q] = op.value(g0 + q, edge_col(in, g0 + q));
      }
    }
  }

  // 2. stage the segments (order-preserving compaction of non-empty rows)
  const uint32_t r_begin = win_row[k], r_end = win_row[k + 1];
  uint32_t n = 0;  // non-empty owned rows so far
  for (uint32_t rb = r_begin; rb < r_end; rb += BLOCK) {
    const uint32_t r = rb + tid;
    uint64_t a = 0, b = 0;
    if (r < r_end) {
      a = in.off[r];
      b = in.off[r + 1];
      if (b == a) op.empty(r);
    }
    const bool ne = r < r_end && b > a;
    const uint64_t m = __ballot(ne);
    if (lane == 0) sh.cnt[wid] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
      before += (w < (int)wid) ? sh.cnt[w] : 0u;
      total += sh.cnt[w];
    }
    if (ne) {
      const uint32_t j = 1 + n + before + lane_rank(m);
      sh.start[j] = (uint32_t)(a - e0);
      sh.row[j] = r;
    }
    n += total;
    __syncthreads();  // cnt is reused by the next chunk
  }
  if (tid == 0) {
    sh.start[0] = 0;
    sh.start[n + 1] = len;
  }
  __syncthreads();

  // 3. walk
  const bool active = p0 < len;
  uint32_t key = NONE;      // open segment at the end of this thread's range
  T x = op.identity();      // its partial
  uint32_t first_j = NONE;  // segment entered at p0, if this thread closes it
  T first_sum = op.identity();
  uint32_t j_start = NONE;  // segment of edge p0
  if (active) {
    uint32_t lo = 0, hi = n + 1;  // start[lo] <= p0 < start[hi]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (sh.start[mid] <= p0) lo = mid;
      else hi = mid;
    }
    uint32_t j = lo;
    j_start = lo;
    uint32_t end = sh.start[j + 1];
    bool first = true;
    T sum = op.identity();
#pragma unroll
    for (int q = 0; q < WIN_EPT; q++) {
      const uint32_t p = p0 + q;
      if (p < len) {
        while (p >= end) {  // segment j closes before edge p
          if (first) {
            first_j = j;
            first_sum = sum;
            first = false;
          } else if (j == 0) {
            op.head(k, sum);
          } else {
            op.finish(sh.row[j], sum);
          }
          sum = op.identity();
          j++;
          end = sh.start[j + 1];
        }
        sum = op.combine(sum, v[q]);
      }
    }
    key = j;
    x = sum;
  }

  // 4. segmented inclusive scan of the open partials
  T S = x;
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    const T y = __shfl_up(S, off);
    const uint32_t ky = __shfl_up(key, off);
    if (lane >= (uint32_t)off && ky == key) S = op.combine(S, y);
  }
  if (lane == WAVE - 1) {
    *reinterpret_cast<T *>(&sh.wsum[wid]) = S;
    sh.wkey[wid] = key;
  }
  __syncthreads();
  // carry into this wave: reduction of segment wkey[wid-1] through wave wid-1
  T cw = op.identity();
  uint32_t kw = NONE;
  if (wid > 0) {
    kw = sh.wkey[wid - 1];
    cw = *reinterpret_cast<const T *>(&sh.wsum[wid - 1]);
    for (int w = (int)wid - 2; w >= 0 && sh.wkey[w] == kw; w--)
      cw = op.combine(cw, *reinterpret_cast<const T *>(&sh.wsum[w]));
  }
  if (key == kw && kw != NONE) S = op.combine(S, cw);
  T prevS = __shfl_up(S, 1);
  uint32_t prevK = __shfl_up(key, 1);
  if (lane == 0) {
    prevS = cw;
    prevK = kw;
  }
  if (!active) return;
  // the previous thread's open segment ended exactly at p0: finish it here
  if (prevK != NONE && prevK != j_start) {
    if (prevK == 0) op.head(k, prevS);
    else op.finish(sh.row[prevK], prevS);
  }
  if (first_j != NONE) {
    const T total = op.combine(prevK == first_j ? prevS : op.identity(), first_sum);
    if (first_j == 0) op.head(k, total);
    else op.finish(sh.row[first_j], total);
  }
  // the window's last segment is closed (or left open) by the thread of the
  // last edge
  if (len - p0 <= WIN_EPT) {
    if (key == 0) {
      op.head(k, S);
    } else {
      const uint32_t r = sh.row[key];
      if (in.off[r + 1] > e1) op.tail(k, S);
      else op.finish(r, S);
    }
  }
}

// rows that span windows: the owner window k holds the tail partial; add
// the head partials of the following windows the row reaches, in order
template <class Op>
__device__ __forceinline__ void window_fixup(const CsrView &in, const uint32_t *win_row,
                                             uint64_t k, Op &op) {
  const uint32_t r_begin = win_row[k], r_end = win_row[k + 1];
  if (r_end == r_begin) return;  // owns no row
  const uint32_t r = r_end - 1;  // last owned row
  const uint64_t b = in.off[r + 1];
  const uint64_t e1 = (k + 1) * WIN < in.E ? (k + 1) * WIN : in.E;
  if (b <= e1) return;  // completed inside its window
  typename Op::T s = op.tail_at(k);
  const uint64_t klast = (b - 1) / WIN;
  for (uint64_t j = k + 1; j <= klast; j++) s = op.combine(s, op.head_at(j));
  op.finish(r, s);
}

}  // namespace mgg
