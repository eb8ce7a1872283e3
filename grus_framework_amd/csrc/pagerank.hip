// pagerank.hip -- PageRank, pull direction, FP32, fixed iteration count.
//
// Reference: samples/pagerank/pr.cu (pr_init :17-33, pr_dd0 :34-43,
// updater :78-90, pull driver :134-168).  The reference runs an
// epsilon-thresholded, unnormalised delta-push with float atomics (order
// dependent) and its pull path never refreshes delta2 (kernel_pull skips
// job.prepare, include/kernel.cuh:128-132).  Here PageRank is the Jacobi
// iteration that the reference's delta recurrence computes, normalised by V
// so that an L1 tolerance is meaningful:
//     r0 = (1-a)/V,   r_{k+1}(v) = (1-a)/V + a * sum_{(u,v) in E} r_k(u)/outdeg(u)
// with the out-degree of the source (pr.cu:73-75, 83; the reference divides
// by the CSC degree in its pull path, :146, which only agrees on symmetric
// graphs).  Dangling mass is dropped, as in the reference.
//
// Kernel (SpMV over the in-edge CSR): the edge array is cut into windows of
// PR_WIN edges, one per workgroup, PR_EPT consecutive edges per thread.
// Each thread gathers contrib[col[e]] for its edges into registers and the
// window is reduced by row segments (merge-style): rows that start and end
// inside the window are finished there, the segment of the row that entered
// from the left (head) goes to head_sum[k] and the partial of the row that
// leaves to the right to tail_sum[k].  A fixup pass completes every row
// that spans windows by adding its partials in window order.  Every sum has
// a fixed order, so ranks are bit-identical run to run.  Each iteration
// writes only the next contrib (r/outdeg); the last one writes the ranks.
//
// Blocked layout (propagation blocking, pr_layout = 3): the destinations
// are cut into bins of 2^bits vertices and the out-edges of each source into
// groups, maximal runs whose destinations share a bin (all edges of a group
// carry the same value contrib[u]).  Group values live in val[], bin-major
// and by source inside a bin.  Per iteration
//   phase 1  k_pb_scatter: val[gpos[j]] = contrib[u] for each group j of u,
//   phase 2  the window kernel over the in-edges with col replaced by
//            gidx[e], the group of edge e, so a window gathers from one
//            bin's slice of val[] (windows in XCD order, so that slice stays
//            in one L2).
// The terms of every row and their order are those of the plain pull (in-
// edges by ascending source), so the ranks are bit-identical to it.  On RMAT
// s24 the plain pull is faster: the gathers it saves hit L2 but still move a
// line each, and the scatter costs more than they did; so it is opt-in.
#include <algorithm>
#include <cstring>

#include "advance.hpp"
#include "apps.hpp"
#include "prims.hpp"

namespace mgg {

constexpr int PR_EPT = 8;                // edges per thread
constexpr int PR_WIN = PR_EPT * BLOCK;    // edges per window
constexpr int PB_BITS = 14;   // blocked layout: 16K destinations per bin

struct PrWork {
  Buf contrib[2], win_row, head_sum, tail_sum;
  uint64_t nwin = 0;
  const void *built_for = nullptr;  // offsets the window table was built on
  uint64_t cap_v = 0;
  // degree-sorted relabelled copy (symmetric, HBM-resident graphs)
  Buf r_off, r_col, perm;  // perm[new id] = original id
  bool relabelled = false;
  // blocked layout: goff[u] = first group of source u (groups numbered in
  // CSR order, goff[V] = G); gpos[j] = slot of group j in val; gidx[e] =
  // slot read by in-edge e; val = group values of the current iteration
  Buf pb_goff, pb_gpos, pb_gidx, pb_val;
  uint64_t pb_G = 0;
  int pb_bits = 0;
  const void *pb_for = nullptr;  // offsets the layout was built on
};

namespace pr {

// win_row[k] = first row r with off[r] >= k*PR_WIN (V if none), k in [0, nwin]
__global__ void k_window_rows(const uint64_t *off, uint64_t V, uint64_t nwin, uint32_t *win_row) {
  uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; r <= V; r += stride) {
    // rows are owned by the window of their start offset; row V is a sentinel
    int64_t kprev = (r == 0) ? -1 : (int64_t)(off[r - 1] / PR_WIN);
    int64_t kcur = (r == V) ? (int64_t)nwin : (int64_t)(off[r] / PR_WIN);
    for (int64_t k = kprev + 1; k <= kcur; k++) win_row[k] = (uint32_t)r;
  }
}

// key = (degree descending, id ascending) for the relabelling sort
__global__ void k_degree_keys(const uint64_t *off, uint64_t V, uint64_t *keys) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < V; i += stride) keys[i] = ((0xFFFFFFFFull - (off[i + 1] - off[i])) << 32) | i;
}

__global__ void k_perm_from_keys(const uint64_t *keys, uint64_t V, uint32_t *perm,
                                 uint32_t *inv) {
  uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; r < V; r += stride) {
    uint32_t v = (uint32_t)keys[r];
    perm[r] = v;
    inv[v] = (uint32_t)r;
  }
}

// relabelled edge keys (inv[u] << 32 | inv[v]), one wave per row
__global__ void k_relabel_edges(CsrView g, const uint32_t *inv, uint64_t *keys) {
  uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE;
  uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / WAVE;
  for (uint64_t u = wave; u < g.V; u += nw) {
    uint64_t nu = (uint64_t)inv[u] << 32;
    for (uint64_t e = g.off[u] + lane_id(); e < g.off[u + 1]; e += WAVE)
      keys[e] = nu | inv[edge_col(g, e)];
  }
}

__global__ void k_scatter_back(const float *r_new, const uint32_t *perm, uint64_t V, float *out) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < V; i += stride) out[perm[i]] = r_new[i];
}

// ---- blocked layout construction (once per graph and bin width) ----

// flag[e] = 1 when CSR edge e starts a group: the first edge of its row or
// a destination bin different from the previous edge's (wave per row)
__global__ void k_pb_flags(const uint64_t *off, const uint32_t *col, uint64_t V, int bits,
                           uint32_t *flag) {
  uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE;
  uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / WAVE;
  for (uint64_t u = wave; u < V; u += nw) {
    const uint64_t a = off[u], b = off[u + 1];
    for (uint64_t e = a + lane_id(); e < b; e += WAVE)
      flag[e] = (e == a || (col[e] >> bits) != (col[e - 1] >> bits)) ? 1u : 0u;
  }
}

// group sort keys (bin << 32 | source), indexed by group id = excl[e] for
// the group's first edge e
__global__ void k_pb_gkeys(const uint64_t *off, const uint32_t *col, uint64_t V, int bits,
                           const uint32_t *flag, const uint32_t *excl, uint64_t *gkey) {
  uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE;
  uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / WAVE;
  for (uint64_t u = wave; u < V; u += nw)
    for (uint64_t e = off[u] + lane_id(); e < off[u + 1]; e += WAVE)
      if (flag[e]) gkey[excl[e]] = ((uint64_t)(col[e] >> bits) << 32) | u;
}

// transpose keys (dst << 32 | src) of every edge; excl[] becomes the group
// id of each edge in place (excl[e] + flag[e] - 1)
__global__ void k_pb_tkeys(const uint64_t *off, const uint32_t *col, uint64_t V,
                           const uint32_t *flag, uint32_t *excl_to_gid, uint64_t *tkey) {
  uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE;
  uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / WAVE;
  for (uint64_t u = wave; u < V; u += nw)
    for (uint64_t e = off[u] + lane_id(); e < off[u + 1]; e += WAVE) {
      tkey[e] = ((uint64_t)col[e] << 32) | u;
      excl_to_gid[e] = excl_to_gid[e] + flag[e] - 1u;
    }
}

// goff[u] = number of groups of the rows before u, for u in [0, V]
__global__ void k_pb_goff(const uint64_t *off, uint64_t V, uint64_t E, const uint32_t *excl,
                          uint64_t G, uint64_t *goff) {
  uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; u <= V; u += stride) goff[u] = off[u] < E ? excl[off[u]] : G;
}

__global__ void k_pb_iota(uint32_t *a, uint64_t n) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) a[i] = (uint32_t)i;
}

// slot p of the bin-major order holds group sorted_gid[p]
__global__ void k_pb_gpos(const uint32_t *sorted_gid, uint64_t G, uint32_t *gpos) {
  uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; p < G; p += stride) gpos[sorted_gid[p]] = (uint32_t)p;
}

// in-edge p (transpose order) reads the slot of its group
__global__ void k_pb_gidx(const uint32_t *sorted_gid, uint64_t E, const uint32_t *gpos,
                          uint32_t *gidx) {
  uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; p < E; p += stride) gidx[p] = gpos[sorted_gid[p]];
}

// ---- blocked layout, phase 1 ----
constexpr uint32_t PB_CHUNK = 16;  // sources per workgroup (measured best of 4..256)

// val[gpos[j]] = contrib[u] for every group j of every source u.  A
// workgroup takes PB_CHUNK consecutive sources (XCD-contiguous chunks),
// stages their group offsets and values in LDS and walks their groups with
// coalesced gpos reads; the owner of group j is found by binary search.
__global__ __launch_bounds__(BLOCK) void k_pb_scatter(const uint64_t *goff, const uint32_t *gpos,
                                                      const float *contrib, float *val,
                                                      uint64_t V) {
  constexpr uint32_t chunk = PB_CHUNK;
  __shared__ uint64_t s_off[PB_CHUNK + 1];
  __shared__ float s_c[PB_CHUNK];
  const uint64_t u0 = xcd_block() * chunk;
  if (u0 >= V) return;
  const uint32_t n = (uint32_t)(V - u0 < chunk ? V - u0 : chunk);
  if (threadIdx.x < n) {
    s_off[threadIdx.x] = goff[u0 + threadIdx.x];
    s_c[threadIdx.x] = contrib[u0 + threadIdx.x];
  }
  if (threadIdx.x == 0) s_off[n] = goff[u0 + n];
  __syncthreads();
  const uint64_t j1 = s_off[n];
  for (uint64_t j = s_off[0] + threadIdx.x; j < j1; j += BLOCK) {
    uint32_t lo = 0, hi = n;  // s_off[lo] <= j < s_off[hi]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_off[mid] <= j) lo = mid;
      else hi = mid;
    }
    val[__builtin_nontemporal_load(&gpos[j])] = s_c[lo];
  }
}

__global__ void k_init_contrib(const uint64_t *out_off, uint64_t V, float base, float *contrib) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < V; i += stride) {
    uint64_t d = out_off[i + 1] - out_off[i];
    contrib[i] = d ? base / (float)d : 0.0f;  // pr_init pr.cu:26,28 (r0 / deg)
  }
}

__device__ __forceinline__ float finish_row(float sum, float base, float alpha) {
  return base + alpha * sum;
}

// a completed row writes its rank (last iteration) or next contrib
template <bool kLast>
__device__ __forceinline__ void finish(uint32_t r, float s, const uint64_t *out_off, float *next,
                                       float base, float alpha) {
  const float rank = finish_row(s, base, alpha);
  if (kLast) {
    next[r] = rank;
  } else {
    const uint64_t d = out_off[r + 1] - out_off[r];
    next[r] = d ? rank / (float)d : 0.0f;
  }
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// One window of PR_WIN in-edges per workgroup, PR_EPT consecutive edges per
// thread, reduced by row segments (merge-style CSR SpMV):
//   1. each thread gathers the contributions of its PR_EPT edges into
//      registers (two 16-byte column loads, non-temporal);
//   2. the window's segments are staged in LDS: segment 0 is the head (the
//      part of a row that began in an earlier window, possibly empty), then
//      every non-empty row owned by the window in order; empty owned rows
//      are finished on the spot with a zero sum;
//   3. each thread finds the segment of its first edge by binary search and
//      walks its edges: segments that begin and end inside its range are
//      finished directly, the one it enters keeps its partial until the
//      carry is known, the one it leaves open is its scan value (and is
//      finished by the next thread when it ends exactly at that boundary);
//   4. a segmented inclusive scan over threads (keys = open segment, 64-lane
//      shuffles then the 4 waves through LDS) gives every thread the sum of
//      the open partials of the threads before it in the same segment.
// Sums are sequential within a thread and a fixed tree across threads, so
// results are bit-identical run to run.  The head partial goes to
// head_sum[k], the partial of a row that continues past the window to
// tail_sum[k]; k_fixup completes those rows.
template <bool kLast>
__global__ __launch_bounds__(BLOCK) void k_pull_window(CsrView in, const uint64_t *out_off,
                                                       const uint32_t *win_row,
                                                       const float *contrib, float *next,
                                                       float *head_sum, float *tail_sum,
                                                       float base, float alpha, uint64_t nwin,
                                                       bool xcd_order) {
  constexpr int NW = BLOCK / WAVE;
  constexpr uint32_t NONE = 0xFFFFFFFFu;
  __shared__ uint32_t s_start[PR_WIN + 2];  // segment starts relative to e0
  __shared__ uint32_t s_row[PR_WIN + 2];    // row of segment j >= 1
  __shared__ uint32_t s_cnt[NW];
  __shared__ float s_wsum[NW];
  __shared__ uint32_t s_wkey[NW];
  // blocked layout: consecutive windows gather from the same bin, so they
  // are kept on one XCD (its L2); the plain pull gathers from anywhere and
  // runs slightly faster in dispatch order
  const uint64_t k = xcd_order ? xcd_block() : blockIdx.x;
  if (k >= nwin) return;
  const uint64_t e0 = k * PR_WIN;
  const uint64_t e1 = e0 + PR_WIN < in.E ? e0 + PR_WIN : in.E;
  const uint32_t len = (uint32_t)(e1 - e0);
  const uint32_t tid = threadIdx.x, lane = lane_id(), wid = tid / WAVE;
  const uint32_t p0 = tid * PR_EPT;

  // 1. gather
  float v[PR_EPT];
  {
    const uint64_t g0 = e0 + p0;
    // a thread's PR_EPT columns lie wholly on one side of the HBM / host
    // split (a multiple of 64) and are 32-byte aligned on either side
    if (p0 + PR_EPT <= len) {
      const u32x4 *cp = reinterpret_cast<const u32x4 *>(
          g0 < in.split ? in.col + g0 : in.col_far + (g0 - in.split));
      const u32x4 c0 = __builtin_nontemporal_load(cp);
      const u32x4 c1 = __builtin_nontemporal_load(cp + 1);
      v[0] = contrib[c0.x];
      v[1] = contrib[c0.y];
      v[2] = contrib[c0.z];
      v[3] = contrib[c0.w];
      v[4] = contrib[c1.x];
      v[5] = contrib[c1.y];
      v[6] = contrib[c1.z];
      v[7] = contrib[c1.w];
    } else {
#pragma unroll
      for (int q = 0; q < PR_EPT; q++) {
        v[q] = 0.0f;
        if (p0 + q < len) v[q] = contrib[edge_col(in, g0 + q)];
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
      if (b == a) finish<kLast>(r, 0.0f, out_off, next, base, alpha);
    }
    const bool ne = r < r_end && b > a;
    const uint64_t m = __ballot(ne);
    if (lane == 0) s_cnt[wid] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
      before += (w < (int)wid) ? s_cnt[w] : 0u;
      total += s_cnt[w];
    }
    if (ne) {
      const uint32_t j = 1 + n + before + lane_rank(m);
      s_start[j] = (uint32_t)(a - e0);
      s_row[j] = r;
    }
    n += total;
    __syncthreads();  // s_cnt is reused by the next chunk
  }
  if (tid == 0) {
    s_start[0] = 0;
    s_start[n + 1] = len;
  }
  __syncthreads();

  // 3. walk
  const bool active = p0 < len;
  uint32_t key = NONE;        // open segment at the end of this thread's range
  float x = 0.0f;             // its partial
  uint32_t first_j = NONE;    // segment entered at p0, if this thread closes it
  float first_sum = 0.0f;
  uint32_t j_start = NONE;    // segment of edge p0
  if (active) {
    uint32_t lo = 0, hi = n + 1;  // s_start[lo] <= p0 < s_start[hi]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_start[mid] <= p0) lo = mid;
      else hi = mid;
    }
    uint32_t j = lo;
    j_start = lo;
    uint32_t end = s_start[j + 1];
    bool first = true;
    float sum = 0.0f;
#pragma unroll
    for (int q = 0; q < PR_EPT; q++) {
      const uint32_t p = p0 + q;
      if (p < len) {
        while (p >= end) {  // segment j closes before edge p
          if (first) {
            first_j = j;
            first_sum = sum;
            first = false;
          } else if (j == 0) {
            head_sum[k] = sum;
          } else {
            finish<kLast>(s_row[j], sum, out_off, next, base, alpha);
          }
          sum = 0.0f;
          j++;
          end = s_start[j + 1];
        }
        sum += v[q];
      }
    }
    key = j;
    x = sum;
    if (first) first_j = NONE;  // never closed a segment: all of it is open
  }

  // 4. segmented inclusive scan of the open partials
  float S = x;
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    const float y = __shfl_up(S, off);
    const uint32_t ky = __shfl_up(key, off);
    if (lane >= (uint32_t)off && ky == key) S += y;
  }
  if (lane == WAVE - 1) {
    s_wsum[wid] = S;
    s_wkey[wid] = key;
  }
  __syncthreads();
  // carry into this wave: total of segment s_wkey[wid-1] through wave wid-1
  float cw = 0.0f;
  uint32_t kw = NONE;
  if (wid > 0) {
    kw = s_wkey[wid - 1];
    cw = s_wsum[wid - 1];
    for (int w = (int)wid - 2; w >= 0 && s_wkey[w] == kw; w--) cw += s_wsum[w];
  }
  if (key == kw && kw != NONE) S += cw;
  float prevS = __shfl_up(S, 1);
  uint32_t prevK = __shfl_up(key, 1);
  if (lane == 0) {
    prevS = cw;
    prevK = kw;
  }
  if (!active) return;
  // the previous thread's open segment ended exactly at p0: finish it here
  if (prevK != NONE && prevK != j_start) {
    if (prevK == 0) head_sum[k] = prevS;
    else finish<kLast>(s_row[prevK], prevS, out_off, next, base, alpha);
  }
  if (first_j != NONE) {
    const float total = (prevK == first_j ? prevS : 0.0f) + first_sum;
    if (first_j == 0) head_sum[k] = total;
    else finish<kLast>(s_row[first_j], total, out_off, next, base, alpha);
  }
  // the window's last segment is closed (or left open) by the thread of the
  // last edge
  if (p0 < len && len - p0 <= PR_EPT) {
    if (key == 0) {
      head_sum[k] = S;
    } else {
      const uint32_t r = s_row[key];
      if (in.off[r + 1] > e1) tail_sum[k] = S;
      else finish<kLast>(r, S, out_off, next, base, alpha);
    }
  }
}

// rows that span windows: owner window k has a tail partial; add the head
// partials of the following windows the row reaches, in order
template <bool kLast>
__global__ void k_fixup(CsrView in, const uint64_t *out_off, const uint32_t *win_row,
                        const float *head_sum, const float *tail_sum, float *next, float base,
                        float alpha, uint64_t nwin) {
  uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (k >= nwin) return;
  uint32_t r_begin = win_row[k], r_end = win_row[k + 1];
  if (r_end == r_begin) return;  // owns no row
  uint32_t r = r_end - 1;        // last owned row
  uint64_t b = in.off[r + 1];
  uint64_t e1 = (k + 1) * PR_WIN < in.E ? (k + 1) * PR_WIN : in.E;
  if (b <= e1) return;  // completed inside its window
  float s = tail_sum[k];
  uint64_t klast = (b - 1) / PR_WIN;
  for (uint64_t j = k + 1; j <= klast; j++) s += head_sum[j];
  float rank = finish_row(s, base, alpha);
  if (kLast) {
    next[r] = rank;
  } else {
    uint64_t d = out_off[r + 1] - out_off[r];
    next[r] = d ? rank / (float)d : 0.0f;
  }
}

}  // namespace pr

namespace {
// Degree-sorted relabelling (one-time, cached): new id = rank of the vertex
// by degree (descending, ties by id), so the contributions of high-degree
// vertices -- the targets of most gathers on a power-law graph -- share cache
// lines instead of being spread over the whole contrib array.
void build_relabelled(Graph &g, PrWork &w) {
  Context &ctx = *g.ctx;
  const uint64_t V = g.V, E = g.E;
  Buf ka, kb, inv;
  ka.alloc(std::max(V, E) * 8);
  kb.alloc(std::max(V, E) * 8);
  inv.alloc(V * 4);
  w.perm.alloc(V * 4);
  pr::k_degree_keys<<<grid_cap(ctx, V, BLOCK), BLOCK, 0, ctx.stream>>>(g.off.as<uint64_t>(), V,
                                                                      ka.as<uint64_t>());
  GRUS_LAUNCH_CHECK();
  uint64_t *sk = sort_keys_u64(ctx, ka.as<uint64_t>(), kb.as<uint64_t>(), V);
  pr::k_perm_from_keys<<<grid_cap(ctx, V, BLOCK), BLOCK, 0, ctx.stream>>>(
      sk, V, w.perm.as<uint32_t>(), inv.as<uint32_t>());
  GRUS_LAUNCH_CHECK();
  pr::k_relabel_edges<<<grid_cap(ctx, V * WAVE, BLOCK), BLOCK, 0, ctx.stream>>>(
      g.csr(), inv.as<uint32_t>(), ka.as<uint64_t>());
  GRUS_LAUNCH_CHECK();
  uint64_t *se = sort_keys_u64(ctx, ka.as<uint64_t>(), kb.as<uint64_t>(), E);
  w.r_off.alloc((V + 1) * 8);
  w.r_col.alloc(E * 4);
  csr_from_sorted_keys(ctx, se, E, V, w.r_off.as<uint64_t>(), w.r_col.as<uint32_t>());
  ctx.sync();
  w.relabelled = true;
}

int bits_for(uint64_t x) {  // bits needed to represent x
  int b = 0;
  while (x) {
    b++;
    x >>= 1;
  }
  return b;
}

// Blocked layout over the out-edge CSR (HBM).  Returns false, leaving the
// plain pull in charge, when the device lacks the memory for the build
// (peak ~ 28E + 8G + 8V bytes with G <= E groups, freed as it goes).
bool build_blocked(Graph &g, PrWork &w, int bits) {
  Context &ctx = *g.ctx;
  const CsrView out = g.csr();
  const uint64_t V = g.V, E = g.E;
  size_t free_b = 0, total_b = 0;
  GRUS_HIP(hipMemGetInfo(&free_b, &total_b));
  const uint64_t need = 36 * E + 16 * V + (1ull << 30);
  if (free_b < need) return false;
  w.pb_goff.release();
  w.pb_gpos.release();
  w.pb_gidx.release();
  w.pb_val.release();
  w.pb_for = nullptr;
  const unsigned rows_grid = grid_cap(ctx, V * WAVE, BLOCK);
  Buf flag, excl;
  flag.alloc(E * 4);
  excl.alloc(E * 4);
  pr::k_pb_flags<<<rows_grid, BLOCK, 0, ctx.stream>>>(out.off, out.col, V, bits,
                                                      flag.as<uint32_t>());
  GRUS_LAUNCH_CHECK();
  scan_excl_u32(ctx, flag.as<uint32_t>(), excl.as<uint32_t>(), E);
  uint32_t last[2];
  GRUS_HIP(hipMemcpyAsync(&last[0], excl.as<uint32_t>() + E - 1, 4, hipMemcpyDeviceToHost,
                          ctx.stream));
  GRUS_HIP(hipMemcpyAsync(&last[1], flag.as<uint32_t>() + E - 1, 4, hipMemcpyDeviceToHost,
                          ctx.stream));
  ctx.sync();
  const uint64_t G = (uint64_t)last[0] + last[1];
  w.pb_goff.alloc((V + 1) * 8);
  pr::k_pb_goff<<<grid_cap(ctx, V + 1, BLOCK), BLOCK, 0, ctx.stream>>>(
      out.off, V, E, excl.as<uint32_t>(), G, w.pb_goff.as<uint64_t>());
  GRUS_LAUNCH_CHECK();
  // bin-major slots of the groups
  w.pb_gpos.alloc(G * 4);
  {
    Buf ka, kb, va, vb;
    ka.alloc(G * 8);
    kb.alloc(G * 8);
    va.alloc(G * 4);
    vb.alloc(G * 4);
    pr::k_pb_gkeys<<<rows_grid, BLOCK, 0, ctx.stream>>>(out.off, out.col, V, bits,
                                                        flag.as<uint32_t>(), excl.as<uint32_t>(),
                                                        ka.as<uint64_t>());
    GRUS_LAUNCH_CHECK();
    pr::k_pb_iota<<<grid_cap(ctx, G, BLOCK), BLOCK, 0, ctx.stream>>>(va.as<uint32_t>(), G);
    GRUS_LAUNCH_CHECK();
    uint32_t *sorted_gid = nullptr;
    sort_pairs_u64_u32(ctx, ka.as<uint64_t>(), kb.as<uint64_t>(), va.as<uint32_t>(),
                       vb.as<uint32_t>(), G, &sorted_gid, 32 + bits_for((V - 1) >> bits));
    pr::k_pb_gpos<<<grid_cap(ctx, G, BLOCK), BLOCK, 0, ctx.stream>>>(sorted_gid, G,
                                                                     w.pb_gpos.as<uint32_t>());
    GRUS_LAUNCH_CHECK();
    ctx.sync();
  }
  // the group slot of every in-edge, in transpose (dst, src) order
  {
    Buf ka, kb, vb;
    ka.alloc(E * 8);
    kb.alloc(E * 8);
    vb.alloc(E * 4);
    pr::k_pb_tkeys<<<rows_grid, BLOCK, 0, ctx.stream>>>(out.off, out.col, V, flag.as<uint32_t>(),
                                                        excl.as<uint32_t>(), ka.as<uint64_t>());
    GRUS_LAUNCH_CHECK();
    ctx.sync();
    flag.release();
    uint32_t *sorted_gid = nullptr;
    sort_pairs_u64_u32(ctx, ka.as<uint64_t>(), kb.as<uint64_t>(), excl.as<uint32_t>(),
                       vb.as<uint32_t>(), E, &sorted_gid, 32 + bits_for(V - 1));
    w.pb_gidx.alloc(E * 4);
    pr::k_pb_gidx<<<grid_cap(ctx, E, BLOCK), BLOCK, 0, ctx.stream>>>(
        sorted_gid, E, w.pb_gpos.as<uint32_t>(), w.pb_gidx.as<uint32_t>());
    GRUS_LAUNCH_CHECK();
    ctx.sync();
  }
  w.pb_val.alloc(G * 4);
  w.pb_G = G;
  w.pb_bits = bits;
  w.pb_for = out.off;
  return true;
}
}  // namespace

void run_pagerank(Graph &g, float alpha, int iters, grus_stats *st) {
  Context &ctx = *g.ctx;
  ctx.activate();
  if (iters < 0) raise(GRUS_E_INVALID, "iteration count");
  if (!(alpha >= 0.0f && alpha <= 1.0f)) raise(GRUS_E_INVALID, "damping factor");
  if (g.V == 0) raise(GRUS_E_INVALID, "empty graph");
  if (!g.pr) g.pr = std::make_shared<PrWork>();
  PrWork &w = *g.pr;
  const uint64_t V = g.V;
  const uint32_t layout = ctx.tuning.pr_layout;
  const int bits = ctx.tuning.pr_bin_bits ? (int)ctx.tuning.pr_bin_bits : PB_BITS;
  if (layout > 3 || bits < 6 || bits > 24) raise(GRUS_E_INVALID, "PageRank layout tuning");
  // relabelling only on request: on RMAT inputs the natural numbering
  // already spreads the hubs over many cache lines, while degree order packs
  // the hottest contributions into a few lines and measured ~5x slower
  const bool relabel = g.symmetric && g.placement == GRUS_PLACE_HBM && layout == 1;
  if (relabel && !w.relabelled) build_relabelled(g, w);
  CsrView out = g.csr();
  CsrView in = g.csc();
  // blocked layout on request only (HBM graphs whose group and edge ids fit
  // 32 bits): on RMAT s24 its scatter costs more than the gathers it saves
  bool blocked = !relabel && layout == 3 && g.placement == GRUS_PLACE_HBM &&
                 g.E > 0 && g.E <= (1ull << 31) && V < (1ull << 32);
  if (blocked && (w.pb_for != out.off || w.pb_bits != bits)) blocked = build_blocked(g, w, bits);
  if (blocked) {
    in.col = w.pb_gidx.as<const uint32_t>();
    in.col_far = nullptr;
    in.split = in.E;
  }
  if (relabel) {
    out.off = w.r_off.as<const uint64_t>();
    out.col = w.r_col.as<const uint32_t>();
    out.w = nullptr;
    out.split = out.E;
    in = out;
  }
  if (w.built_for != in.off || w.cap_v < V) {
    w.nwin = in.E / PR_WIN + 1;
    w.win_row.alloc((w.nwin + 1) * 4);
    w.head_sum.alloc(w.nwin * 4);
    w.tail_sum.alloc(w.nwin * 4);
    w.contrib[0].alloc(V * 4);
    w.contrib[1].alloc(V * 4);
    pr::k_window_rows<<<grid_cap(ctx, V + 1, BLOCK), BLOCK, 0, ctx.stream>>>(in.off, V, w.nwin,
                                                                           w.win_row.as<uint32_t>());
    GRUS_LAUNCH_CHECK();
    w.built_for = in.off;
    w.cap_v = V;
  }
  g.result.ensure(V * 4);
  float *rank = g.result.as<float>();
  const float base = (float)((1.0 - (double)alpha) / (double)V);
  const unsigned nwin = (unsigned)(div_up(w.nwin, NXCD) * NXCD);  // a multiple of NXCD
  const unsigned scat_grid = (unsigned)(div_up(div_up(V, pr::PB_CHUNK), NXCD) * NXCD);
  const unsigned fix_grid = (unsigned)div_up(w.nwin, BLOCK);
  GRUS_HIP(hipEventRecord(ctx.ev0, ctx.stream));
  pr::k_init_contrib<<<grid_cap(ctx, V, BLOCK), BLOCK, 0, ctx.stream>>>(out.off, V, base,
                                                                      w.contrib[0].as<float>());
  GRUS_LAUNCH_CHECK();
  uint64_t launches = 1;
  if (iters == 0) {  // r0 = (1-a)/V everywhere
    uint32_t bits;
    memcpy(&bits, &base, 4);
    fill_u32(ctx, reinterpret_cast<uint32_t *>(rank), V, bits);
    launches++;
  }
  int cur = 0;
  for (int it = 0; it < iters; it++) {
    const bool last = (it == iters - 1);
    float *next = (last && !relabel) ? rank : w.contrib[cur ^ 1].as<float>();
    const float *c = w.contrib[cur].as<float>();
    if (blocked) {  // phase 1: group values; phase 2 gathers them through gidx
      pr::k_pb_scatter<<<scat_grid, BLOCK, 0, ctx.stream>>>(
          w.pb_goff.as<uint64_t>(), w.pb_gpos.as<uint32_t>(), c, w.pb_val.as<float>(), V);
      GRUS_LAUNCH_CHECK();
      c = w.pb_val.as<float>();
      launches++;
    }
    if (last) {
      pr::k_pull_window<true><<<nwin, BLOCK, 0, ctx.stream>>>(
          in, out.off, w.win_row.as<uint32_t>(), c, next, w.head_sum.as<float>(),
          w.tail_sum.as<float>(), base, alpha, w.nwin, blocked);
      GRUS_LAUNCH_CHECK();
      pr::k_fixup<true><<<fix_grid, BLOCK, 0, ctx.stream>>>(
          in, out.off, w.win_row.as<uint32_t>(), w.head_sum.as<float>(), w.tail_sum.as<float>(),
          next, base, alpha, w.nwin);
    } else {
      pr::k_pull_window<false><<<nwin, BLOCK, 0, ctx.stream>>>(
          in, out.off, w.win_row.as<uint32_t>(), c, next, w.head_sum.as<float>(),
          w.tail_sum.as<float>(), base, alpha, w.nwin, blocked);
      GRUS_LAUNCH_CHECK();
      pr::k_fixup<false><<<fix_grid, BLOCK, 0, ctx.stream>>>(
          in, out.off, w.win_row.as<uint32_t>(), w.head_sum.as<float>(), w.tail_sum.as<float>(),
          next, base, alpha, w.nwin);
    }
    GRUS_LAUNCH_CHECK();
    launches += 2;
    cur ^= 1;
  }
  if (relabel && iters > 0) {  // ranks back to the original vertex ids
    pr::k_scatter_back<<<grid_cap(ctx, V, BLOCK), BLOCK, 0, ctx.stream>>>(
        w.contrib[cur].as<float>(), w.perm.as<uint32_t>(), V, rank);
    GRUS_LAUNCH_CHECK();
    launches++;
  }
  GRUS_HIP(hipEventRecord(ctx.ev1, ctx.stream));
  GRUS_HIP(hipEventSynchronize(ctx.ev1));
  float ms = 0;
  GRUS_HIP(hipEventElapsedTime(&ms, ctx.ev0, ctx.ev1));
  if (st) {
    *st = grus_stats{};
    st->device_ms = ms;
    st->iterations = (uint32_t)iters;
    st->vertices_reached = V;
    st->edges_reached = in.E;
    st->kernel_launches = launches;
  }
}

}  // namespace mgg
