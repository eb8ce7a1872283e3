// pagerank.hip -- PageRank, FP32 ranks, fixed iteration count.
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
// Layouts (grus_tuning::pr_layout): the binned push (4, automatic on HBM
// graphs that fit it), the plain pull (2, everything else), propagation
// blocking over the pull (3) and a degree-sorted relabelled pull (1).
//
// Binned push (pr_layout = 4; build_binned, k_bp_scatter, k_bp_accum):
// destinations are cut into bins of 2^14 vertices and the out-edges of each
// source into groups whose destinations share a bin.  Phase 1 writes one
// value per group into per-bin runs; phase 2 streams each bin's 16-bit edge
// entries and group values and adds them into LDS accumulators as 64-bit
// fixed point (floor(x * 2^62)).  Integer sums are exact and order-free, so
// the ranks are bit-identical run to run and whatever the bin width or task
// split, and closer to the FP64 oracle than an FP32 sum.  Every global
// access is a sequential stream; the random accesses hit LDS.
//
// Plain pull (SpMV over the in-edge CSR): the edge array is cut into windows of
// WIN edges, one per workgroup, WIN_EPT consecutive edges per thread.
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
#include "window.hpp"

namespace mgg {

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
  // binned push layout (pr_layout = 4); see the comment above
  // build_binned.  p1 = phase-1 stream (u16); ch_u0 / ch_i0 / ch_r0 =
  // first source / stream position / run of every phase-1 chunk; rtab =
  // slot offset of every run; rbase = runs started before every slab;
  // val = group values (bin-major);
  // d2 = phase-2 edge stream (u16); gs512 = first group of every slab;
  // task / multi = phase-2 tasks and the bins split over several tasks;
  // gacc = fixed-point sums of split bins (kept zero between iterations)
  Buf bp_p1, bp_ch_u0, bp_ch_i0, bp_ch_r0, bp_rtab, bp_rbase, bp_val, bp_d2, bp_gs512, bp_task,
      bp_multi, bp_gacc;
  uint64_t bp_G = 0, bp_nch = 0, bp_ntask = 0, bp_nmulti = 0;
  uint32_t bp_nbins = 0;
  int bp_bits = 0;
  const void *bp_for = nullptr;
  std::vector<uint32_t> bp_bstart;  // first phase-2 edge of every bin (host)
  uint64_t bp_task_edges = 0;       // task size the task table was cut with
  // SSSP pull rounds over the same layout: w2 = edge weights in phase-2
  // stream order (slab-transposed like d2); gmin = minima of the bins split
  // over several tasks (all 0xFFFFFFFF between rounds)
  Buf bp_w2, bp_gmin;
  int bp_w2_bytes = 4;  // 1, 2 or 4: the narrowest that holds the largest weight
  const void *bp_w2_for = nullptr;
  const void *bp_unusable_for = nullptr;  // offsets a failed build was tried on
};

__global__ void k_window_rows(const uint64_t *off, uint64_t V, uint64_t nwin, uint32_t *win_row) {
  uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; r <= V; r += stride) {
    // rows are owned by the window of their start offset; row V is a sentinel
    int64_t kprev = (r == 0) ? -1 : (int64_t)(off[r - 1] / WIN);
    int64_t kcur = (r == V) ? (int64_t)nwin : (int64_t)(off[r] / WIN);
    for (int64_t k = kprev + 1; k <= kcur; k++) win_row[k] = (uint32_t)r;
  }
}

void build_window_rows(Context &ctx, const uint64_t *off, uint64_t V, uint64_t nwin,
                       uint32_t *win_row) {
  k_window_rows<<<grid_cap(ctx, V + 1, BLOCK), BLOCK, 0, ctx.stream>>>(off, V, nwin, win_row);
  GRUS_LAUNCH_CHECK();
}

namespace pr {

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
__device__ __forceinline__ void finish(uint64_t r, float s, const uint64_t *out_off, float *next,
                                       float base, float alpha) {
  const float rank = finish_row(s, base, alpha);
  if (kLast) {
    next[r] = rank;
  } else {
    const uint64_t d = out_off[r + 1] - out_off[r];
    next[r] = d ? rank / (float)d : 0.0f;
  }
}

// ---- binned push layout (pr_layout = 4), per iteration ----
constexpr int BP_MAX_BITS = 14;       // phase-2 accumulators: 2^14 x u64 = 128 KB of LDS
constexpr uint32_t BP_MAX_BINS = 8192;  // phase-1 directory row in LDS: <= 32 KB
constexpr int BP_BLOCK = 1024;        // phase-2 workgroup (one per CU: its LDS)
constexpr int BP_EPT = 8;             // phase-2 entries per lane and step (16-byte load)
constexpr int BP_SLAB = WAVE * BP_EPT;  // entries per wave step
constexpr int BP_FIX = 62;            // sums are kept as floor(x * 2^62) in u64
constexpr uint32_t BP_SPAN = 8192;    // sources per phase-1 chunk (staged in LDS)
constexpr int BP_UNROLL = 8;          // loads in flight per thread in the streaming loops

struct BpTask {
  uint32_t e0, e1;  // edge range of the phase-2 stream
  uint32_t bin;
  uint32_t single;  // 1: the bin's only task, which finishes its rows
};

// floor(x * 2^62) for a finite x in [0, 4): contributions are ranks
// divided by degrees, and ranks sum to at most 1.  With y = x * 2^30
// (exact), the high word is trunc(y) and the low word trunc(fract(y) *
// 2^32); fract(y) is exact and has at most 24 significant bits, so both
// conversions are exact.
__device__ __forceinline__ unsigned long long to_fixed(float x) {
  static_assert(BP_FIX == 62, "the split below assumes 2^62");
  const float y = x * 0x1p30f;
  const uint32_t hi = (uint32_t)y;
  const uint32_t lo = (uint32_t)(__builtin_amdgcn_fractf(y) * 0x1p32f);
  return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ float from_fixed(unsigned long long s) {
  return (float)((double)s * 0x1p-62);
}

// Phase 1: val[slot(j)] = contrib[src(j)] for every group j.  Workgroup c
// takes the groups of one source chunk in (bin, source) order, so the
// stores of each bin are one contiguous run.  The stream holds one u16 per
// group: bit 15 starts a run, bits 0..12 the source within the chunk.
// Slabs of BP_SLAB entries are stored transposed (lane l of a wave gets
// the entries l, l + 64, ... with one 16-byte load), so each row of 64
// stores is consecutive within a run.  The run of an entry is rbase[slab]
// plus a wave prefix count of the run flags; rtab[run] turns a stream
// position into its slot (u32 arithmetic, slots < 2^32).  Chunks run in
// dispatch order, so the chunks in flight write neighbouring runs of every
// bin (keeping consecutive chunks on one XCD instead measured 6% slower on
// RMAT s24).  The chunk's contributions (at most BP_SPAN sources) and run
// offsets are staged in LDS first: read straight from memory, the stream's
// source order made each wave's loads touch dozens of lines.
//
// The value of a source is any 4-byte pattern Src gives: PageRank's
// contribution, or (SSSP pull rounds) the distance of a frontier vertex.
struct ContribValue {
  const float *contrib;
  __device__ __forceinline__ uint32_t operator()(uint32_t u) const {
    return __float_as_uint(contrib[u]);
  }
};
struct FrontDistValue {  // 0xFFFFFFFF: not in the frontier bitmap
  const uint64_t *front;
  const uint32_t *dist;
  __device__ __forceinline__ uint32_t operator()(uint32_t u) const {
    const uint32_t d = dist[u];  // both loads issued before either is used
    return bm_test(front, u) ? d : 0xFFFFFFFFu;
  }
};
struct StampDistValue {  // 0xFFFFFFFF: not stamped as the current queue
  const uint32_t *stamp;
  uint32_t id;
  const uint32_t *dist;
  __device__ __forceinline__ uint32_t operator()(uint32_t u) const {
    const uint32_t d = dist[u];
    return stamp[u] == id ? d : 0xFFFFFFFFu;
  }
};

//
// The values are stored as VT: u32, or u16 for SSSP rounds whose frontier
// distances are all below 0xFFFF (0xFFFF then marks "not in the frontier").
template <class Src, typename VT = uint32_t>
__global__ __launch_bounds__(BLOCK) void k_bp_scatter(const uint32_t *ch_u0, const uint32_t *ch_i0,
                                                      const uint32_t *ch_r0, const uint16_t *p1,
                                                      const uint32_t *rbase, const uint32_t *rtab,
                                                      uint32_t nbins, Src src, VT *val) {
  extern __shared__ uint32_t sr[];  // the chunk's run offsets (<= nbins), then BP_SPAN values
  uint32_t *sc = sr + nbins;
  const uint64_t c = blockIdx.x;
  const uint32_t u0 = ch_u0[c], span = ch_u0[c + 1] - u0;  // span >= 1
  const uint32_t r0 = ch_r0[c], nr = ch_r0[c + 1] - r0;    // nr <= nbins
  // every staging loop issues BP_UNROLL loads (indices clamped in range)
  // before it uses any: one memory latency per round instead of one per
  // element
  for (uint32_t jb = 0; jb < nr; jb += BLOCK * BP_UNROLL) {
    uint32_t x[BP_UNROLL];
#pragma unroll
    for (int q = 0; q < BP_UNROLL; q++) {
      const uint32_t j = jb + threadIdx.x + BLOCK * q;
      x[q] = rtab[r0 + (j < nr ? j : nr - 1u)];
    }
#pragma unroll
    for (int q = 0; q < BP_UNROLL; q++) {
      const uint32_t j = jb + threadIdx.x + BLOCK * q;
      if (j < nr) sr[j] = x[q];
    }
  }
  for (uint32_t jb = 0; jb < span; jb += BLOCK * BP_UNROLL) {
    uint32_t x[BP_UNROLL];
#pragma unroll
    for (int q = 0; q < BP_UNROLL; q++) {
      const uint32_t j = jb + threadIdx.x + BLOCK * q;
      x[q] = src(u0 + (j < span ? j : span - 1u));
    }
#pragma unroll
    for (int q = 0; q < BP_UNROLL; q++) {
      const uint32_t j = jb + threadIdx.x + BLOCK * q;
      if (j < span) sc[j] = x[q];
    }
  }
  __syncthreads();
  const uint32_t i0 = ch_i0[c], i1 = ch_i0[c + 1];
  if (i0 == i1) return;  // a chunk of sources without out-edges
  constexpr uint32_t NW = BLOCK / WAVE;
  const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
  const uint32_t s1 = (i1 + BP_SLAB - 1u) / BP_SLAB;
  uint32_t s = i0 / BP_SLAB + wid;
  u32x4 nx = {0, 0, 0, 0};
  uint32_t nrb = 0;
  if (s < s1) {
    nx = __builtin_nontemporal_load(
        reinterpret_cast<const u32x4 *>(p1 + (uint64_t)s * BP_SLAB + lane * BP_EPT));
    nrb = rbase[s];
  }
  for (; s < s1; s += NW) {
    const u32x4 cur = nx;
    const uint32_t rb = nrb;  // runs started before the slab
    if (s + NW < s1) {  // the next slab loads while this one is stored
      nx = __builtin_nontemporal_load(
          reinterpret_cast<const u32x4 *>(p1 + (uint64_t)(s + NW) * BP_SLAB + lane * BP_EPT));
      nrb = rbase[s + NW];
    }
    const uint32_t wd[4] = {cur.x, cur.y, cur.z, cur.w};
    uint32_t ent[BP_EPT], rq[BP_EPT];
    uint32_t run = rb - 1u;
#pragma unroll
    for (int q = 0; q < BP_EPT; q++) {
      ent[q] = (q & 1) ? (wd[q >> 1] >> 16) : (wd[q >> 1] & 0xFFFFu);
      const uint64_t bq = __ballot(ent[q] >> 15);
      rq[q] = run + lane_rank(bq) + (ent[q] >> 15);
      run += (uint32_t)__popcll(bq);
    }
    const uint32_t sb = s * BP_SLAB;
    if (sb >= i0 && sb + BP_SLAB <= i1) {  // wave-uniform: the whole slab is the chunk's
#pragma unroll
      for (int q = 0; q < BP_EPT; q++) {
        const uint32_t i = sb + WAVE * q + lane;
        val[sr[rq[q] - r0] + i] = (VT)sc[ent[q] & 0x1FFFu];
      }
    } else {
#pragma unroll
      for (int q = 0; q < BP_EPT; q++) {
        const uint32_t i = sb + WAVE * q + lane;
        if (i >= i0 && i < i1) val[sr[rq[q] - r0] + i] = (VT)sc[ent[q] & 0x1FFFu];
      }
    }
  }
}

// Phase 2: one task = a piece of one destination bin's edge stream.  Each
// stream entry is a u16: bit 15 starts a group, bits 0..14 the destination
// within the bin.  A wave step takes one slab of BP_SLAB entries (8 per
// lane, one 16-byte load; slabs are stored transposed, so lane l gets the
// entries l, l + 64, ...).  The group index of every entry follows from
// the slab's first group (gs512) and a wave prefix count of the flags (one
// ballot per row of 64 entries).  The slab's groups, at most BP_SLAB of them,
// are loaded into the wave's LDS window with coalesced loads; gathered
// lane by lane from memory they touched a dozen lines per load (RMAT s24:
// 0.74 -> 0.48 ms without those loads).  The values are added into the
// LDS accumulators as fixed-point integers, so the sums do not depend on
// the order of the atomics.  A bin's only task finishes its rows; pieces of
// split bins add their sums into gacc, which k_bp_finish completes.
template <bool kLast>
__global__ __launch_bounds__(BP_BLOCK) void k_bp_accum(
    const BpTask *tasks, const uint16_t *d2, const uint32_t *gs512, const float *val, uint64_t V,
    int bits, const uint64_t *out_off, float *next, unsigned long long *gacc, float base,
    float alpha) {
  __shared__ unsigned long long acc[1 << BP_MAX_BITS];
  __shared__ float win[BP_BLOCK / WAVE][BP_SLAB];
  const BpTask t = tasks[blockIdx.x];
  const uint64_t v0 = (uint64_t)t.bin << bits;
  const uint32_t nb = (uint32_t)((V - v0) < (1ull << bits) ? (V - v0) : (1ull << bits));
  for (uint32_t i = threadIdx.x; i < nb; i += BP_BLOCK) acc[i] = 0ull;
  __syncthreads();
  constexpr uint32_t NW = BP_BLOCK / WAVE;
  const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
  float *wv = win[wid];
  const uint32_t s1 = (uint32_t)((t.e1 + (uint64_t)BP_SLAB - 1) / BP_SLAB);
  // software pipeline: the next slab's entries load while this one runs
  uint32_t s = t.e0 / BP_SLAB + wid;
  u32x4 nx = {0, 0, 0, 0};
  uint32_t ngs = 0;
  if (s < s1) {
    nx = __builtin_nontemporal_load(
        reinterpret_cast<const u32x4 *>(d2 + (uint64_t)s * BP_SLAB + lane * BP_EPT));
    ngs = gs512[s];
  }
  for (; s < s1; s += NW) {
    const u32x4 cur = nx;
    const uint32_t gbase = ngs;  // groups that start before the slab
    const uint32_t wd[4] = {cur.x, cur.y, cur.z, cur.w};
    // lane l holds the slab's entries 64 q + l, q = 0..7 (transposed layout)
    uint32_t ent[BP_EPT];
#pragma unroll
    for (int q = 0; q < BP_EPT; q++) ent[q] = (q & 1) ? (wd[q >> 1] >> 16) : (wd[q >> 1] & 0xFFFFu);
    // group of entry 64 q + l: gbase - 1 + (flags of the rows before q) +
    // (flags of row q up to lane l); row totals are wave-uniform
    uint32_t gq[BP_EPT];
    uint32_t run = gbase - 1u;
    uint64_t b0 = 0;
#pragma unroll
    for (int q = 0; q < BP_EPT; q++) {
      const uint64_t bq = __ballot(ent[q] >> 15);
      if (q == 0) b0 = bq;
      gq[q] = run + lane_rank(bq) + (ent[q] >> 15);
      run += (uint32_t)__popcll(bq);
    }
    // window = the slab's groups: the one its first entry continues (when
    // that entry starts none) and every group started inside it
    const uint32_t w0 = (b0 & 1ull) ? gbase : gbase - 1u;
    // at most BP_SLAB groups from w0 on; val is padded by BP_SLAB, so the
    // window loads need no bound
    float wl[BP_SLAB / WAVE];
#pragma unroll
    for (int j = 0; j < BP_SLAB / WAVE; j++) wl[j] = val[w0 + lane + WAVE * j];
    // the next slab's entries are requested after the window, so waiting
    // for the window does not wait for them (unconditional: past the
    // task's last slab it reloads that slab, which keeps the wait counts
    // exact)
    {
      const uint32_t sn = s + NW < s1 ? s + NW : s1 - 1u;
      nx = __builtin_nontemporal_load(
          reinterpret_cast<const u32x4 *>(d2 + (uint64_t)sn * BP_SLAB + lane * BP_EPT));
      ngs = gs512[sn];
    }
#pragma unroll
    for (int j = 0; j < BP_SLAB / WAVE; j++) wv[lane + WAVE * j] = wl[j];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float x[BP_EPT];
#pragma unroll
    for (int q = 0; q < BP_EPT; q++) x[q] = wv[gq[q] - w0];  // neighbouring lanes: near addresses
    const uint32_t sb = s * BP_SLAB;
    if (sb >= t.e0 && sb + BP_SLAB <= t.e1) {  // wave-uniform: the whole slab is the task's
#pragma unroll
      for (int q = 0; q < BP_EPT; q++)
        atomicAdd(&acc[ent[q] & ((1u << BP_MAX_BITS) - 1u)], to_fixed(x[q]));
    } else {
      // a slab shared with a neighbouring bin (or padding): entries outside
      // the task add zero; their destination bits stay inside the array
#pragma unroll
      for (int q = 0; q < BP_EPT; q++) {
        const uint32_t pe = sb + WAVE * q + lane;
        const unsigned long long keep = 0ull - (unsigned long long)(pe >= t.e0 && pe < t.e1);
        atomicAdd(&acc[ent[q] & ((1u << BP_MAX_BITS) - 1u)], to_fixed(x[q]) & keep);
      }
    }
    // the window is rewritten by the next slab only after these reads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  __syncthreads();
  constexpr int R = (1 << BP_MAX_BITS) / BP_BLOCK;  // rows per thread at most
  if (!t.single) {
#pragma unroll
    for (int r = 0; r < R; r++) {
      const uint32_t i = threadIdx.x + BP_BLOCK * r;
      if (i < nb) {
        const unsigned long long sum = acc[i];
        if (sum) atomicAdd(&gacc[v0 + i], sum);
      }
    }
  } else if (kLast) {
#pragma unroll
    for (int r = 0; r < R; r++) {
      const uint32_t i = threadIdx.x + BP_BLOCK * r;
      if (i < nb) next[v0 + i] = finish_row(from_fixed(acc[i]), base, alpha);
    }
  } else {
    // degrees of all the thread's rows are requested before any is used
#pragma unroll
    for (int rb = 0; rb < R; rb += BP_UNROLL) {
      uint64_t a[BP_UNROLL], b[BP_UNROLL];
#pragma unroll
      for (int q = 0; q < BP_UNROLL; q++) {
        const uint32_t i = threadIdx.x + BP_BLOCK * (rb + q);
        const uint64_t v = v0 + (i < nb ? i : nb - 1u);
        a[q] = out_off[v];
        b[q] = out_off[v + 1];
      }
#pragma unroll
      for (int q = 0; q < BP_UNROLL; q++) {
        const uint32_t i = threadIdx.x + BP_BLOCK * (rb + q);
        if (i < nb) {
          const float rank = finish_row(from_fixed(acc[i]), base, alpha);
          const uint64_t d = b[q] - a[q];
          next[v0 + i] = d ? rank / (float)d : 0.0f;
        }
      }
    }
  }
}

// rows of the bins split over several tasks; clears gacc for the next use
template <bool kLast>
__global__ __launch_bounds__(BLOCK) void k_bp_finish(const uint32_t *mbins, int bits, uint64_t V,
                                                     unsigned long long *gacc,
                                                     const uint64_t *out_off, float *next,
                                                     float base, float alpha) {
  const uint32_t per = (1u << bits) / BLOCK > 0 ? (1u << bits) / BLOCK : 1u;
  const uint64_t bin = mbins[blockIdx.x / per];
  const uint64_t i = (uint64_t)(blockIdx.x % per) * BLOCK + threadIdx.x;
  const uint64_t v = (bin << bits) + i;
  if (i >= (1ull << bits) || v >= V) return;
  const unsigned long long s = gacc[v];
  gacc[v] = 0ull;
  finish<kLast>(v, from_fixed(s), out_off, next, base, alpha);
}

// ---- SSSP pull rounds over the binned layout ----
// The finish of one bin's rows (v0 + i, i < nb), BP_BLOCK threads with the
// rows i = tid + BP_BLOCK r: v takes best(i) when it is below dist[v], then
// joins the near queue (below the threshold; its out-degree counts toward
// the next round's edges) or the far list (once per far epoch).  Only the
// finisher of v writes dist[v], so distances need no atomics.  The workgroup
// appends both lists with one global atomic each (a block scan of the
// per-thread counts): an atomic per wave on the two counters serialised
// ~10^5 waves per round.
constexpr int BPR = (1 << BP_MAX_BITS) / BP_BLOCK;  // rows per thread at most
static_assert(BPR <= 32, "row bits per thread");

template <class Best>
__device__ __forceinline__ void finish_rows(uint64_t v0, uint32_t nb, Best best_of,
                                            const BinnedMinArgs &a) {
  constexpr uint32_t NW = BP_BLOCK / WAVE;
  __shared__ uint32_t s_w[NW];
  __shared__ uint32_t s_base[2];
  __shared__ unsigned long long s_e[NW];
  const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
  uint32_t nearm = 0, farm = 0;
  uint64_t edges = 0;
#pragma unroll
  for (int r = 0; r < BPR; r++) {
    const uint32_t i = threadIdx.x + BP_BLOCK * r;
    if (i >= nb) continue;
    const uint32_t best = best_of(i);
    const uint64_t v = v0 + i;
    if (best < INFINIT && best < a.dist[v]) {
      a.dist[v] = best;
      if (best < a.T_thr) {
        a.near_stamp[v] = a.near_id;
        nearm |= 1u << r;
        edges += a.out_off[v + 1] - a.out_off[v];
      } else if (a.far_stamp[v] != a.far_id) {
        a.far_stamp[v] = a.far_id;
        farm |= 1u << r;
      }
    }
  }
  // block exclusive scan of (near count << 16 | far count): both stay below
  // 2^15 per workgroup, so the halves never carry into each other
  const uint32_t mine = ((uint32_t)__popc(nearm) << 16) | (uint32_t)__popc(farm);
  uint32_t inc = mine;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const uint32_t y = __shfl_up(inc, o);
    if (lane >= (uint32_t)o) inc += y;
  }
  const unsigned long long ew = wave_sum((unsigned long long)edges);
  if (lane == WAVE - 1) s_w[wid] = inc;
  if (lane == 0) s_e[wid] = ew;
  __syncthreads();
  uint32_t before = 0, total = 0;
  unsigned long long etot = 0;
#pragma unroll
  for (uint32_t w = 0; w < NW; w++) {
    before += w < wid ? s_w[w] : 0u;
    total += s_w[w];
    etot += s_e[w];
  }
  if (threadIdx.x == 0) {
    s_base[0] = (total >> 16) ? atomicAdd(&a.cnt->next_count, total >> 16) : 0u;
    s_base[1] = (total & 0xFFFFu) ? atomicAdd(&a.cnt->far_count, total & 0xFFFFu) : 0u;
    if (etot) atomicAdd((unsigned long long *)&a.cnt->next_edges, etot);
  }
  __syncthreads();
  const uint32_t excl = before + inc - mine;
  uint32_t pn = s_base[0] + (excl >> 16), pf = s_base[1] + (excl & 0xFFFFu);
  for (uint32_t m = nearm; m; m &= m - 1)
    a.qn[pn++] = (uint32_t)(v0 + threadIdx.x + BP_BLOCK * (uint32_t)__ffs(m) - BP_BLOCK);
  for (uint32_t m = farm; m; m &= m - 1)
    a.qfar[pf++] = (uint32_t)(v0 + threadIdx.x + BP_BLOCK * (uint32_t)__ffs(m) - BP_BLOCK);
}

// Phase 2 of an SSSP pull round: the phase-1 values are the distances of
// frontier sources (0xFFFFFFFF for the others); entry e of a bin's stream
// offers val[group(e)] + w2[e] to its destination, and the minimum is kept
// in LDS (exact and order-independent).  The stream walk is k_bp_accum's;
// w2 is stored like d2, so lane l's 8, 16 or 32 bytes of a slab (weights of
// 1, 2 or 4 bytes) are the weights of its entries l, l + 64, ...  A bin's
// only task finishes its rows; pieces of split bins take the minimum into
// gmin, which k_bp_min_finish completes.

// lane l's 8 weights of a slab (entries l, l + 64, ...)
template <typename WT>
struct W2Slab {
  uint32_t v[BP_EPT];
  __device__ __forceinline__ void load(const WT *w2, uint64_t o) {
    if constexpr (sizeof(WT) == 1) {
      const uint64_t x = __builtin_nontemporal_load(reinterpret_cast<const uint64_t *>(w2 + o));
#pragma unroll
      for (int q = 0; q < BP_EPT; q++) v[q] = (uint32_t)(x >> (8 * q)) & 0xFFu;
    } else if constexpr (sizeof(WT) == 2) {
      const u32x4 x = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(w2 + o));
      const uint32_t wd[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
      for (int q = 0; q < BP_EPT; q++) v[q] = (q & 1) ? (wd[q >> 1] >> 16) : (wd[q >> 1] & 0xFFFFu);
    } else {
      const u32x4 x0 = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(w2 + o));
      const u32x4 x1 = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(w2 + o) + 1);
      v[0] = x0.x; v[1] = x0.y; v[2] = x0.z; v[3] = x0.w;
      v[4] = x1.x; v[5] = x1.y; v[6] = x1.z; v[7] = x1.w;
    }
  }
};

template <typename WT, typename VT>
__global__ __launch_bounds__(BP_BLOCK) void k_bp_min(const BpTask *tasks, const uint16_t *d2,
                                                     const WT *w2, const uint32_t *gs512,
                                                     const VT *val, uint64_t V, int bits,
                                                     BinnedMinArgs a, uint32_t *gmin) {
  __shared__ uint32_t acc[1 << BP_MAX_BITS];
  __shared__ uint32_t win[BP_BLOCK / WAVE][BP_SLAB];
  const BpTask t = tasks[blockIdx.x];
  const uint64_t v0 = (uint64_t)t.bin << bits;
  const uint32_t nb = (uint32_t)((V - v0) < (1ull << bits) ? (V - v0) : (1ull << bits));
  for (uint32_t i = threadIdx.x; i < nb; i += BP_BLOCK) acc[i] = 0xFFFFFFFFu;
  __syncthreads();
  constexpr uint32_t NW = BP_BLOCK / WAVE;
  const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
  uint32_t *wv = win[wid];
  const uint32_t s1 = (uint32_t)((t.e1 + (uint64_t)BP_SLAB - 1) / BP_SLAB);
  uint32_t s = t.e0 / BP_SLAB + wid;
  u32x4 nx = {0, 0, 0, 0};
  W2Slab<WT> nw{};
  uint32_t ngs = 0;
  if (s < s1) {
    const uint64_t o = (uint64_t)s * BP_SLAB + lane * BP_EPT;
    nx = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(d2 + o));
    nw.load(w2, o);
    ngs = gs512[s];
  }
  for (; s < s1; s += NW) {
    const u32x4 cur = nx;
    const W2Slab<WT> cw = nw;
    const uint32_t gbase = ngs;
    const uint32_t wd[4] = {cur.x, cur.y, cur.z, cur.w};
    const uint32_t *wt = cw.v;
    uint32_t ent[BP_EPT];
#pragma unroll
    for (int q = 0; q < BP_EPT; q++) ent[q] = (q & 1) ? (wd[q >> 1] >> 16) : (wd[q >> 1] & 0xFFFFu);
    uint32_t gq[BP_EPT];
    uint32_t run = gbase - 1u;
    uint64_t b0 = 0;
#pragma unroll
    for (int q = 0; q < BP_EPT; q++) {
      const uint64_t bq = __ballot(ent[q] >> 15);
      if (q == 0) b0 = bq;
      gq[q] = run + lane_rank(bq) + (ent[q] >> 15);
      run += (uint32_t)__popcll(bq);
    }
    const uint32_t w0 = (b0 & 1ull) ? gbase : gbase - 1u;
    uint32_t wl[BP_SLAB / WAVE];
#pragma unroll
    for (int j = 0; j < BP_SLAB / WAVE; j++) {
      const uint32_t x = val[w0 + lane + WAVE * j];
      wl[j] = (sizeof(VT) == 2 && x == 0xFFFFu) ? 0xFFFFFFFFu : x;  // u16: 0xFFFF = none
    }
    {
      const uint32_t sn = s + NW < s1 ? s + NW : s1 - 1u;
      const uint64_t o = (uint64_t)sn * BP_SLAB + lane * BP_EPT;
      nx = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(d2 + o));
      nw.load(w2, o);
      ngs = gs512[sn];
    }
#pragma unroll
    for (int j = 0; j < BP_SLAB / WAVE; j++) wv[lane + WAVE * j] = wl[j];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    uint32_t x[BP_EPT];
#pragma unroll
    for (int q = 0; q < BP_EPT; q++) x[q] = wv[gq[q] - w0];
    const uint32_t sb = s * BP_SLAB;
    const bool whole = sb >= t.e0 && sb + BP_SLAB <= t.e1;  // wave-uniform
#pragma unroll
    for (int q = 0; q < BP_EPT; q++) {
      const uint64_t c = (uint64_t)x[q] + wt[q];
      const uint32_t pe = sb + WAVE * q + lane;
      const bool mine = whole || (pe >= t.e0 && pe < t.e1);
      if (mine && x[q] < INFINIT && c < INFINIT)
        atomicMin(&acc[ent[q] & ((1u << BP_MAX_BITS) - 1u)], (uint32_t)c);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  __syncthreads();
  if (!t.single) {  // block-uniform
#pragma unroll
    for (int r = 0; r < BPR; r++) {
      const uint32_t i = threadIdx.x + BP_BLOCK * r;
      if (i < nb && acc[i] != 0xFFFFFFFFu) atomicMin(&gmin[v0 + i], acc[i]);
    }
    return;
  }
  finish_rows(v0, nb, [&](uint32_t i) { return acc[i]; }, a);
}

// rows of the bins split over several tasks, one workgroup per bin; resets
// gmin for the next round
__global__ __launch_bounds__(BP_BLOCK) void k_bp_min_finish(const uint32_t *mbins, int bits,
                                                            uint64_t V, uint32_t *gmin,
                                                            BinnedMinArgs a) {
  const uint64_t v0 = (uint64_t)mbins[blockIdx.x] << bits;
  const uint32_t nb = (uint32_t)((V - v0) < (1ull << bits) ? (V - v0) : (1ull << bits));
  finish_rows(v0, nb, [&](uint32_t i) {
    const uint32_t best = gmin[v0 + i];
    if (best != 0xFFFFFFFFu) gmin[v0 + i] = 0xFFFFFFFFu;
    return best;
  }, a);
}

// weight stream keys: CSR edge e -> (bin of its destination, weight); the
// same keys as the d2 sort, so the same stable sort puts the weights in d2's
// order
__global__ void k_bp_w2_keys(const uint64_t *off, const uint32_t *col, const uint32_t *w,
                             uint64_t V, int bits, uint64_t *key, uint32_t *wt) {
  uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE;
  uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / WAVE;
  for (uint64_t u = wave; u < V; u += nw)
    for (uint64_t e = off[u] + lane_id(); e < off[u + 1]; e += WAVE) {
      key[e] = col[e] >> bits;
      wt[e] = w[e];
    }
}

// w2: the sorted weights (narrowed to WT, which holds the largest), padded
// with zeros to a multiple of BP_SLAB, each slab stored transposed as d2
template <typename WT>
__global__ void k_bp_w2(const uint32_t *wt, uint64_t E, uint64_t Epad, WT *w2) {
  uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; p < Epad; p += stride) {
    const uint64_t r = p % BP_SLAB;
    const uint64_t phys = p - r + (r % WAVE) * BP_EPT + r / WAVE;
    w2[phys] = p < E ? (WT)wt[p] : (WT)0;
  }
}

__global__ void k_max_u32(const uint32_t *a, uint64_t n, uint32_t *out) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint32_t m = 0;
  for (; i < n; i += stride) m = a[i] > m ? a[i] : m;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t y = __shfl_xor(m, o);
    m = y > m ? y : m;
  }
  if (lane_id() == 0 && m) atomicMax(out, m);
}

// ---- binned push layout construction ----

// keys of the phase-1 order: (chunk << 16 | bin) of every group, with
// chunk = the chunk whose group range holds j
__global__ void k_bp_p1_keys(const uint64_t *gkey, uint64_t G, const uint32_t *ch_i0,
                             uint32_t nch, uint64_t *key, uint32_t *ids) {
  uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; j < G; j += stride) {
    uint32_t lo = 0, hi = nch;  // ch_i0[lo] <= j < ch_i0[hi]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (ch_i0[mid] <= j) lo = mid;
      else hi = mid;
    }
    key[j] = ((uint64_t)lo << 16) | (gkey[j] >> 32);
    ids[j] = (uint32_t)j;
  }
}

// group keys by bin only (stable sort keeps the source order inside a bin)
__global__ void k_bp_bin_keys(const uint64_t *gkey, uint64_t G, uint64_t *key, uint32_t *ids) {
  uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; j < G; j += stride) {
    key[j] = gkey[j] >> 32;
    ids[j] = (uint32_t)j;
  }
}

// slot[group] = its position in (bin, source) order
__global__ void k_bp_slots(const uint32_t *sorted, uint64_t G, uint32_t *slot) {
  uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; s < G; s += stride) slot[sorted[s]] = (uint32_t)s;
}

// phase-1 stream entry and run data: position i of chunk c holds group
// sorted[i]; a run starts at the chunk's first position and wherever the
// bin changes, and records slot - position.  The u16 entries are written
// slab-transposed (as the phase-2 stream).
__global__ void k_bp_p1(const uint32_t *sorted, const uint64_t *gkey, const uint32_t *slot,
                        uint64_t G, uint64_t Gpad, const uint32_t *ch_u0, const uint32_t *ch_i0,
                        uint32_t nch, uint16_t *p1, uint32_t *rflag, uint32_t *rdel) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < Gpad; i += stride) {
    const uint64_t r = i % BP_SLAB;
    const uint64_t phys = i - r + (r % WAVE) * BP_EPT + r / WAVE;
    if (i >= G) {
      p1[phys] = 0;
      continue;
    }
    uint32_t lo = 0, hi = nch;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (ch_i0[mid] <= i) lo = mid;
      else hi = mid;
    }
    const uint32_t j = sorted[i];
    const uint32_t bin = (uint32_t)(gkey[j] >> 32);
    const uint32_t src = (uint32_t)gkey[j];
    const bool run_start = (i == ch_i0[lo]) || (uint32_t)(gkey[sorted[i - 1]] >> 32) != bin;
    p1[phys] = (uint16_t)((run_start ? 0x8000u : 0u) | (src - ch_u0[lo]));
    rflag[i] = run_start ? 1u : 0u;
    rdel[i] = run_start ? slot[j] - (uint32_t)i : 0u;
  }
}

// run table, runs before every slab and first run of every chunk
__global__ void k_bp_runs(const uint32_t *rflag, const uint32_t *rexcl, const uint32_t *rdel,
                          uint64_t G, uint64_t nslab, const uint32_t *ch_i0, uint32_t nch,
                          uint32_t R, uint32_t *rtab, uint32_t *rbase, uint32_t *ch_r0) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < G; i += stride) {
    if (rflag[i]) rtab[rexcl[i]] = rdel[i];
    if (i % BP_SLAB == 0 && i / BP_SLAB < nslab) rbase[i / BP_SLAB] = rexcl[i];
  }
  i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  for (; i <= nch; i += stride) ch_r0[i] = ch_i0[i] < G ? rexcl[ch_i0[i]] : R;
}

// phase-2 stream keys: CSR edge e -> (bin of its destination, entry)
__global__ void k_bp_d2_keys(const uint64_t *off, const uint32_t *col, uint64_t V, int bits,
                             const uint32_t *flag, uint64_t *key, uint32_t *ent) {
  uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE;
  uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / WAVE;
  const uint32_t mask = (1u << bits) - 1u;
  for (uint64_t u = wave; u < V; u += nw)
    for (uint64_t e = off[u] + lane_id(); e < off[u + 1]; e += WAVE) {
      const uint32_t v = col[e];
      key[e] = v >> bits;
      ent[e] = (flag[e] << 15) | (v & mask);
    }
}

// d2 (u16, padded with zeros to a multiple of BP_SLAB, each slab stored
// transposed), its group flags as
// u32 for the scan, and the first edge of every bin
__global__ void k_bp_d2(const uint64_t *key, const uint32_t *ent, uint64_t E, uint64_t Epad,
                        uint32_t nbins, uint16_t *d2, uint32_t *f, uint32_t *bstart) {
  uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; p < Epad; p += stride) {
    // slab-transposed: logical entry r = 64 q + l of a slab is stored in
    // lane l's 16 bytes, slot q, so one 16-byte load per lane gives each
    // lane the entries l, l + 64, ..., l + 448
    const uint64_t r = p % BP_SLAB;
    const uint64_t phys = p - r + (r % WAVE) * BP_EPT + r / WAVE;
    if (p >= E) {
      d2[phys] = 0;
      continue;
    }
    d2[phys] = (uint16_t)ent[p];
    f[p] = ent[p] >> 15;
    const uint64_t kb = key[p];
    const uint64_t ka = p ? key[p - 1] + 1 : 0;  // bins (key[p-1], key[p]] start here
    for (uint64_t b = ka; b <= kb; b++) bstart[b] = (uint32_t)p;
    if (p == E - 1)
      for (uint64_t b = kb + 1; b <= nbins; b++) bstart[b] = (uint32_t)E;
  }
}

// gs512[s] = groups that start before slab s
__global__ void k_bp_gs512(const uint32_t *excl, uint64_t E, uint64_t ns, uint32_t *gs512) {
  uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; k < ns; k += stride) gs512[k] = excl[k * BP_SLAB < E ? k * BP_SLAB : E - 1];
}

// PageRank as a window reduction: sum of contrib[u] over the in-edges
template <bool kLast>
struct PrOp {
  using T = float;
  const float *contrib;
  float *next;
  const uint64_t *out_off;
  float base, alpha;
  float *head_sum, *tail_sum;
  __device__ T identity() const { return 0.0f; }
  __device__ T combine(T a, T b) const { return a + b; }
  __device__ T value(uint64_t, uint32_t c) const { return contrib[c]; }
  __device__ void finish(uint32_t r, T s) const {
    pr::finish<kLast>(r, s, out_off, next, base, alpha);
  }
  __device__ void empty(uint32_t r) const { finish(r, 0.0f); }
  __device__ void head(uint64_t k, T s) const { head_sum[k] = s; }
  __device__ void tail(uint64_t k, T s) const { tail_sum[k] = s; }
  __device__ T head_at(uint64_t k) const { return head_sum[k]; }
  __device__ T tail_at(uint64_t k) const { return tail_sum[k]; }
};

template <bool kLast>
__global__ __launch_bounds__(BLOCK) void k_pull_window(CsrView in, const uint32_t *win_row,
                                                       PrOp<kLast> op, uint64_t nwin,
                                                       bool xcd_order) {
  __shared__ WindowShared sh;
  // blocked layout: consecutive windows gather from the same bin, so they
  // are kept on one XCD (its L2); the plain pull gathers from anywhere and
  // runs slightly faster in dispatch order
  const uint64_t k = xcd_order ? xcd_block() : blockIdx.x;
  if (k >= nwin) return;
  window_reduce(in, win_row, k, op, sh);
}

template <bool kLast>
__global__ void k_fixup(CsrView in, const uint32_t *win_row, PrOp<kLast> op, uint64_t nwin) {
  uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (k >= nwin) return;
  window_fixup(in, win_row, k, op);
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

// Binned push layout (pr_layout = 4), built once per graph and bin width.
// Destinations are cut into bins of 2^bits vertices (bits <= 14) and the
// out-edges of each source into groups: maximal runs whose destinations
// share a bin.  All edges of a group carry the same value contrib[u].
//   phase 1  writes one value per group into val[], bin-major and by
//            source inside a bin.  Sources are cut into chunks of at most
//            32K groups and 8K sources; a chunk's groups are streamed in
//            (bin, source) order (p1: bin << 16 | source - first source),
//            so each bin receives one contiguous run per chunk.
//   phase 2  streams each bin's edges in (source, destination) order as
//            u16 entries (group-start flag, destination in the bin) next to
//            the bin's slice of val[], and adds the values into LDS
//            accumulators.  Bins with more than BP_TASK_EDGES edges are
//            split into several tasks whose partial sums meet in gacc.
// Every access of both phases is a sequential stream except the LDS
// atomics, the reads of contrib inside one chunk's source range and the
// per-bin runs of phase 1.  Returns false (the plain pull takes over) when
// the graph does not fit the layout or the device lacks memory for the
// build (peak ~ 40E + 16V bytes).
constexpr uint32_t BP_CHUNK_GROUPS = 32768;
constexpr uint64_t BP_TASK_EDGES = 1ull << 18;  // automatic task size

bool build_binned(Graph &g, PrWork &w, int bits) {
  Context &ctx = *g.ctx;
  const CsrView out = g.csr();
  const uint64_t V = g.V, E = g.E;
  const uint64_t nbins = div_up(V, 1ull << bits);
  if (E == 0 || E >= (1ull << 32) - 1024 || V >= (1ull << 32) || nbins > pr::BP_MAX_BINS)
    return false;
  size_t free_b = 0, total_b = 0;
  GRUS_HIP(hipMemGetInfo(&free_b, &total_b));
  if (free_b < 48 * E + 16 * V + (1ull << 30)) return false;
  for (Buf *b : {&w.bp_p1, &w.bp_ch_u0, &w.bp_ch_i0, &w.bp_ch_r0, &w.bp_rtab, &w.bp_rbase,
                 &w.bp_val, &w.bp_d2, &w.bp_gs512, &w.bp_task, &w.bp_multi, &w.bp_gacc,
                 &w.bp_w2})
    b->release();
  w.bp_for = nullptr;
  w.bp_w2_for = nullptr;
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
  // groups per source (goff) -> phase-1 chunks, on the host (a source has
  // at most nbins <= 8192 groups, so it always fits a chunk)
  static const uint64_t chunk_groups = [] {
    const char *e = getenv("GRUS_BP_CHUNK_GROUPS");  // experiments only
    const long v = e ? atol(e) : 0;
    return v >= (long)pr::BP_MAX_BINS ? (uint64_t)v : (uint64_t)BP_CHUNK_GROUPS;
  }();
  std::vector<uint32_t> ch_u0, ch_i0;
  {
    Buf goff;
    goff.alloc((V + 1) * 8);
    pr::k_pb_goff<<<grid_cap(ctx, V + 1, BLOCK), BLOCK, 0, ctx.stream>>>(
        out.off, V, E, excl.as<uint32_t>(), G, goff.as<uint64_t>());
    GRUS_LAUNCH_CHECK();
    std::vector<uint64_t> h(V + 1);
    GRUS_HIP(hipMemcpyAsync(h.data(), goff.p, (V + 1) * 8, hipMemcpyDeviceToHost, ctx.stream));
    ctx.sync();
    uint64_t u0 = 0;
    while (u0 < V) {
      uint64_t u1 = u0 + 1;
      while (u1 < V && u1 - u0 < pr::BP_SPAN && h[u1 + 1] - h[u0] <= chunk_groups) u1++;
      ch_u0.push_back((uint32_t)u0);
      ch_i0.push_back((uint32_t)h[u0]);
      u0 = u1;
    }
    ch_i0.push_back((uint32_t)G);
  }
  const uint32_t nch = (uint32_t)ch_u0.size();
  ch_u0.push_back((uint32_t)V);  // sentinel: the last chunk's end
  w.bp_ch_u0.alloc((nch + 1) * 4);
  w.bp_ch_i0.alloc((nch + 1) * 4);
  GRUS_HIP(hipMemcpyAsync(w.bp_ch_u0.p, ch_u0.data(), (nch + 1) * 4, hipMemcpyHostToDevice,
                          ctx.stream));
  GRUS_HIP(hipMemcpyAsync(w.bp_ch_i0.p, ch_i0.data(), (nch + 1) * 4, hipMemcpyHostToDevice,
                          ctx.stream));
  // group keys (bin << 32 | source), by group id
  Buf gkey, slot;
  gkey.alloc(G * 8);
  pr::k_pb_gkeys<<<rows_grid, BLOCK, 0, ctx.stream>>>(out.off, out.col, V, bits,
                                                      flag.as<uint32_t>(), excl.as<uint32_t>(),
                                                      gkey.as<uint64_t>());
  GRUS_LAUNCH_CHECK();
  slot.alloc(G * 4);
  {  // slots: (bin, source) order
    Buf ka, kb, va, vb;
    ka.alloc(G * 8);
    kb.alloc(G * 8);
    va.alloc(G * 4);
    vb.alloc(G * 4);
    pr::k_bp_bin_keys<<<grid_cap(ctx, G, BLOCK), BLOCK, 0, ctx.stream>>>(
        gkey.as<uint64_t>(), G, ka.as<uint64_t>(), va.as<uint32_t>());
    GRUS_LAUNCH_CHECK();
    uint32_t *sorted = nullptr;
    sort_pairs_u64_u32(ctx, ka.as<uint64_t>(), kb.as<uint64_t>(), va.as<uint32_t>(),
                       vb.as<uint32_t>(), G, &sorted, std::max(1, bits_for(nbins - 1)));
    pr::k_bp_slots<<<grid_cap(ctx, G, BLOCK), BLOCK, 0, ctx.stream>>>(sorted, G,
                                                                     slot.as<uint32_t>());
    GRUS_LAUNCH_CHECK();
    ctx.sync();
  }
  const uint64_t Gpad = div_up(G, pr::BP_SLAB) * pr::BP_SLAB;
  const uint64_t nslab1 = Gpad / pr::BP_SLAB;
  w.bp_p1.alloc(Gpad * 2);
  w.bp_rbase.alloc(nslab1 * 4);
  w.bp_ch_r0.alloc((nch + 1) * 4);
  {  // phase-1 stream: (chunk, bin, source) order
    Buf ka, kb, va, vb;
    ka.alloc(G * 8);
    kb.alloc(G * 8);
    va.alloc(G * 4);
    vb.alloc(G * 4);
    pr::k_bp_p1_keys<<<grid_cap(ctx, G, BLOCK), BLOCK, 0, ctx.stream>>>(
        gkey.as<uint64_t>(), G, w.bp_ch_i0.as<uint32_t>(), nch, ka.as<uint64_t>(),
        va.as<uint32_t>());
    GRUS_LAUNCH_CHECK();
    uint32_t *sorted = nullptr;
    sort_pairs_u64_u32(ctx, ka.as<uint64_t>(), kb.as<uint64_t>(), va.as<uint32_t>(),
                       vb.as<uint32_t>(), G, &sorted, 16 + bits_for(nch));
    // run flags, offsets and their scan reuse buffers the sort is done with:
    // the other value buffer, and ka (2G u32)
    uint32_t *rflag = sorted == va.as<uint32_t>() ? vb.as<uint32_t>() : va.as<uint32_t>();
    uint32_t *rdel = ka.as<uint32_t>();
    uint32_t *rexcl = rdel + G;
    pr::k_bp_p1<<<grid_cap(ctx, Gpad, BLOCK), BLOCK, 0, ctx.stream>>>(
        sorted, gkey.as<uint64_t>(), slot.as<uint32_t>(), G, Gpad, w.bp_ch_u0.as<uint32_t>(),
        w.bp_ch_i0.as<uint32_t>(), nch, w.bp_p1.as<uint16_t>(), rflag, rdel);
    GRUS_LAUNCH_CHECK();
    scan_excl_u32(ctx, rflag, rexcl, G);
    uint32_t lastr[2];
    GRUS_HIP(hipMemcpyAsync(&lastr[0], rexcl + G - 1, 4, hipMemcpyDeviceToHost, ctx.stream));
    GRUS_HIP(hipMemcpyAsync(&lastr[1], rflag + G - 1, 4, hipMemcpyDeviceToHost, ctx.stream));
    ctx.sync();
    const uint32_t R = lastr[0] + lastr[1];
    w.bp_rtab.alloc((uint64_t)std::max<uint32_t>(R, 1) * 4);
    pr::k_bp_runs<<<grid_cap(ctx, G, BLOCK), BLOCK, 0, ctx.stream>>>(
        rflag, rexcl, rdel, G, nslab1, w.bp_ch_i0.as<uint32_t>(), nch, R,
        w.bp_rtab.as<uint32_t>(), w.bp_rbase.as<uint32_t>(), w.bp_ch_r0.as<uint32_t>());
    GRUS_LAUNCH_CHECK();
    ctx.sync();
  }
  gkey.release();
  slot.release();
  // phase-2 stream: CSR edges stably sorted by destination bin
  const uint64_t Epad = div_up(E, pr::BP_SLAB) * pr::BP_SLAB;
  const uint64_t nk = Epad / pr::BP_SLAB;
  std::vector<uint32_t> bstart(nbins + 1);
  w.bp_d2.alloc(Epad * 2);
  w.bp_gs512.alloc(nk * 4);
  {
    Buf ka, kb, vb, bs;
    ka.alloc(E * 8);
    kb.alloc(E * 8);
    vb.alloc(E * 4);
    bs.alloc((nbins + 1) * 4);
    pr::k_bp_d2_keys<<<rows_grid, BLOCK, 0, ctx.stream>>>(out.off, out.col, V, bits,
                                                         flag.as<uint32_t>(), ka.as<uint64_t>(),
                                                         excl.as<uint32_t>());
    GRUS_LAUNCH_CHECK();
    uint32_t *ent = nullptr;
    uint64_t *key = sort_pairs_u64_u32(ctx, ka.as<uint64_t>(), kb.as<uint64_t>(),
                                       excl.as<uint32_t>(), vb.as<uint32_t>(), E, &ent,
                                       std::max(1, bits_for(nbins - 1)));
    // flags of the sorted stream go to whichever entry buffer is free
    uint32_t *f = (ent == excl.as<uint32_t>()) ? vb.as<uint32_t>() : excl.as<uint32_t>();
    pr::k_bp_d2<<<grid_cap(ctx, Epad, BLOCK), BLOCK, 0, ctx.stream>>>(
        key, ent, E, Epad, (uint32_t)nbins, w.bp_d2.as<uint16_t>(), f, bs.as<uint32_t>());
    GRUS_LAUNCH_CHECK();
    scan_excl_u32(ctx, f, flag.as<uint32_t>(), E);  // flag now holds the exclusive scan
    pr::k_bp_gs512<<<grid_cap(ctx, nk, BLOCK), BLOCK, 0, ctx.stream>>>(
        flag.as<uint32_t>(), E, nk, w.bp_gs512.as<uint32_t>());
    GRUS_LAUNCH_CHECK();
    GRUS_HIP(hipMemcpyAsync(bstart.data(), bs.p, (nbins + 1) * 4, hipMemcpyDeviceToHost,
                            ctx.stream));
    ctx.sync();
  }
  flag.release();
  excl.release();
  w.bp_val.alloc((G + pr::BP_SLAB) * 4);  // a window may read up to BP_SLAB past the last group
  GRUS_HIP(hipMemsetAsync(w.bp_val.p, 0, (G + pr::BP_SLAB) * 4, ctx.stream));
  ctx.sync();  // host vectors above are copied before they go out of scope
  w.bp_bstart = std::move(bstart);
  w.bp_task_edges = 0;
  w.bp_G = G;
  w.bp_nch = nch;
  w.bp_nbins = (uint32_t)nbins;
  w.bp_bits = bits;
  w.bp_for = out.off;
  return true;
}

// phase-2 tasks: every bin's edge range cut into equal pieces of at most
// task_edges edges (an empty bin still gets one task: its rows are finished)
void build_binned_tasks(Graph &g, PrWork &w, uint64_t task_edges) {
  Context &ctx = *g.ctx;
  std::vector<pr::BpTask> tasks;
  std::vector<uint32_t> multi;
  for (uint64_t b = 0; b < w.bp_nbins; b++) {
    const uint64_t e0 = w.bp_bstart[b], e1 = w.bp_bstart[b + 1];
    const uint64_t n = e1 - e0;
    const uint64_t parts = n > task_edges ? div_up(n, task_edges) : 1;
    if (parts > 1) multi.push_back((uint32_t)b);
    for (uint64_t q = 0; q < parts; q++)
      tasks.push_back(pr::BpTask{(uint32_t)(e0 + n * q / parts),
                                 (uint32_t)(e0 + n * (q + 1) / parts), (uint32_t)b,
                                 parts == 1 ? 1u : 0u});
  }
  w.bp_task.alloc(tasks.size() * sizeof(pr::BpTask));
  GRUS_HIP(hipMemcpyAsync(w.bp_task.p, tasks.data(), tasks.size() * sizeof(pr::BpTask),
                          hipMemcpyHostToDevice, ctx.stream));
  w.bp_multi.release();
  if (!multi.empty()) {
    w.bp_multi.alloc(multi.size() * 4);
    GRUS_HIP(hipMemcpyAsync(w.bp_multi.p, multi.data(), multi.size() * 4, hipMemcpyHostToDevice,
                            ctx.stream));
    if (!w.bp_gacc.p) {
      w.bp_gacc.alloc(g.V * 8);
      GRUS_HIP(hipMemsetAsync(w.bp_gacc.p, 0, g.V * 8, ctx.stream));
    }
  }
  ctx.sync();  // the host vectors are copied before they go out of scope
  w.bp_ntask = tasks.size();
  w.bp_nmulti = multi.size();
  w.bp_task_edges = task_edges;
}

// iterations over the binned push layout (graph already built for it)
void run_binned(Graph &g, PrWork &w, float alpha, int iters, grus_stats *st) {
  Context &ctx = *g.ctx;
  const uint64_t V = g.V;
  const CsrView out = g.csr();
  w.contrib[0].ensure(V * 4);
  g.result.ensure(V * 4);
  float *rank = g.result.as<float>();
  const float base = (float)((1.0 - (double)alpha) / (double)V);
  const int bits = w.bp_bits;
  const unsigned fin_grid =
      (unsigned)(w.bp_nmulti * std::max<uint64_t>(1, (1ull << bits) / BLOCK));
  float *contrib = w.contrib[0].as<float>();
  const pr::BpTask *tasks = w.bp_task.as<const pr::BpTask>();
  unsigned long long *gacc = w.bp_gacc.as<unsigned long long>();
  GRUS_HIP(hipEventRecord(ctx.ev0, ctx.stream));
  pr::k_init_contrib<<<grid_cap(ctx, V, BLOCK), BLOCK, 0, ctx.stream>>>(out.off, V, base, contrib);
  GRUS_LAUNCH_CHECK();
  uint64_t launches = 1;
  if (iters == 0) {  // r0 = (1-a)/V everywhere
    uint32_t b;
    memcpy(&b, &base, 4);
    fill_u32(ctx, reinterpret_cast<uint32_t *>(rank), V, b);
    launches++;
  }
  for (int it = 0; it < iters; it++) {
    const bool last = (it == iters - 1);
    // phase 1 reads contrib; phase 2 overwrites it (or writes the ranks)
    float *next = last ? rank : contrib;
    pr::k_bp_scatter<<<(unsigned)w.bp_nch, BLOCK, (w.bp_nbins + pr::BP_SPAN) * 4, ctx.stream>>>(
        w.bp_ch_u0.as<uint32_t>(), w.bp_ch_i0.as<uint32_t>(), w.bp_ch_r0.as<uint32_t>(),
        w.bp_p1.as<uint16_t>(), w.bp_rbase.as<uint32_t>(), w.bp_rtab.as<uint32_t>(), w.bp_nbins,
        pr::ContribValue{contrib}, w.bp_val.as<uint32_t>());
    GRUS_LAUNCH_CHECK();
    if (last) {
      pr::k_bp_accum<true><<<(unsigned)w.bp_ntask, pr::BP_BLOCK, 0, ctx.stream>>>(
          tasks, w.bp_d2.as<uint16_t>(), w.bp_gs512.as<uint32_t>(), w.bp_val.as<float>(), V, bits,
          out.off, next, gacc, base, alpha);
      GRUS_LAUNCH_CHECK();
      if (w.bp_nmulti)
        pr::k_bp_finish<true><<<fin_grid, BLOCK, 0, ctx.stream>>>(
            w.bp_multi.as<uint32_t>(), bits, V, gacc, out.off, next, base, alpha);
    } else {
      pr::k_bp_accum<false><<<(unsigned)w.bp_ntask, pr::BP_BLOCK, 0, ctx.stream>>>(
          tasks, w.bp_d2.as<uint16_t>(), w.bp_gs512.as<uint32_t>(), w.bp_val.as<float>(), V, bits,
          out.off, next, gacc, base, alpha);
      GRUS_LAUNCH_CHECK();
      if (w.bp_nmulti)
        pr::k_bp_finish<false><<<fin_grid, BLOCK, 0, ctx.stream>>>(
            w.bp_multi.as<uint32_t>(), bits, V, gacc, out.off, next, base, alpha);
    }
    GRUS_LAUNCH_CHECK();
    launches += w.bp_nmulti ? 3 : 2;
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
    st->edges_reached = g.E;
    st->kernel_launches = launches;
  }
}
// the weight stream of the binned layout (for SSSP pull rounds): the CSR
// weights sorted stably by destination bin, as d2 was, so w2 lines up with
// d2 entry by entry
bool build_binned_weights(Graph &g, PrWork &w) {
  Context &ctx = *g.ctx;
  const CsrView out = g.csr();
  const uint64_t V = g.V, E = g.E;
  if (!out.w || out.split < E) return false;
  size_t free_b = 0, total_b = 0;
  GRUS_HIP(hipMemGetInfo(&free_b, &total_b));
  const uint64_t Epad = div_up(E, pr::BP_SLAB) * pr::BP_SLAB;
  if (free_b < 24 * E + 4 * Epad + (1ull << 30)) return false;
  w.bp_w2.release();
  {
    Buf ka, kb, va, vb;
    ka.alloc(E * 8);
    kb.alloc(E * 8);
    va.alloc(E * 4);
    vb.alloc(E * 4);
    pr::k_bp_w2_keys<<<grid_cap(ctx, V * WAVE, BLOCK), BLOCK, 0, ctx.stream>>>(
        out.off, out.col, out.w, V, w.bp_bits, ka.as<uint64_t>(), va.as<uint32_t>());
    GRUS_LAUNCH_CHECK();
    // the stream is as narrow as the largest weight allows (RMAT inputs
    // with weights below 256 read one byte per edge instead of four)
    uint32_t *wmax = kb.as<uint32_t>();
    GRUS_HIP(hipMemsetAsync(wmax, 0, 4, ctx.stream));
    pr::k_max_u32<<<grid_cap(ctx, E, BLOCK, 4), BLOCK, 0, ctx.stream>>>(va.as<uint32_t>(), E,
                                                                       wmax);
    GRUS_LAUNCH_CHECK();
    uint32_t hmax = 0;
    GRUS_HIP(hipMemcpyAsync(&hmax, wmax, 4, hipMemcpyDeviceToHost, ctx.stream));
    ctx.sync();
    w.bp_w2_bytes = hmax < 256u ? 1 : hmax < 65536u ? 2 : 4;
    uint32_t *wt = nullptr;
    sort_pairs_u64_u32(ctx, ka.as<uint64_t>(), kb.as<uint64_t>(), va.as<uint32_t>(),
                       vb.as<uint32_t>(), E, &wt, std::max(1, bits_for(w.bp_nbins - 1)));
    w.bp_w2.alloc(Epad * (uint64_t)w.bp_w2_bytes);
    const unsigned grid = grid_cap(ctx, Epad, BLOCK);
    if (w.bp_w2_bytes == 1)
      pr::k_bp_w2<<<grid, BLOCK, 0, ctx.stream>>>(wt, E, Epad, w.bp_w2.as<uint8_t>());
    else if (w.bp_w2_bytes == 2)
      pr::k_bp_w2<<<grid, BLOCK, 0, ctx.stream>>>(wt, E, Epad, w.bp_w2.as<uint16_t>());
    else
      pr::k_bp_w2<<<grid, BLOCK, 0, ctx.stream>>>(wt, E, Epad, w.bp_w2.as<uint32_t>());
    GRUS_LAUNCH_CHECK();
    ctx.sync();  // the sort buffers are freed on scope exit
  }
  w.bp_w2_for = out.off;
  return true;
}
}  // namespace

bool sssp_binned_ready(Graph &g) {
  if (g.placement != GRUS_PLACE_HBM || g.V == 0) return false;
  Context &ctx = *g.ctx;
  const CsrView out = g.csr();
  if (!out.w) return false;
  if (!g.pr) g.pr = std::make_shared<PrWork>();
  PrWork &w = *g.pr;
  if (w.bp_unusable_for == out.off) return false;
  if (w.bp_for != out.off) {  // a layout PageRank already built is used as it is
    const int bits = std::min(ctx.tuning.pr_bin_bits ? (int)ctx.tuning.pr_bin_bits : PB_BITS,
                              pr::BP_MAX_BITS);
    if (!build_binned(g, w, bits)) {
      w.bp_unusable_for = out.off;
      return false;
    }
  }
  if (w.bp_task_edges == 0)
    build_binned_tasks(g, w, ctx.tuning.pr_task_edges ? ctx.tuning.pr_task_edges : BP_TASK_EDGES);
  if (w.bp_w2_for != out.off && !build_binned_weights(g, w)) {
    w.bp_unusable_for = out.off;
    return false;
  }
  if (w.bp_gmin.bytes < g.V * 4) {
    w.bp_gmin.alloc(g.V * 4);
    fill_u32(ctx, w.bp_gmin.as<uint32_t>(), g.V, 0xFFFFFFFFu);
  }
  return true;
}

namespace {
template <class Src, typename VT>
void binned_min_launch(Graph &g, PrWork &w, const BinnedMinArgs &a, Src src) {
  Context &ctx = *g.ctx;
  const int bits = w.bp_bits;
  VT *val = w.bp_val.as<VT>();
  pr::k_bp_scatter<<<(unsigned)w.bp_nch, BLOCK, (w.bp_nbins + pr::BP_SPAN) * 4, ctx.stream>>>(
      w.bp_ch_u0.as<uint32_t>(), w.bp_ch_i0.as<uint32_t>(), w.bp_ch_r0.as<uint32_t>(),
      w.bp_p1.as<uint16_t>(), w.bp_rbase.as<uint32_t>(), w.bp_rtab.as<uint32_t>(), w.bp_nbins,
      src, val);
  GRUS_LAUNCH_CHECK();
  const pr::BpTask *tasks = w.bp_task.as<const pr::BpTask>();
  const uint16_t *d2 = w.bp_d2.as<uint16_t>();
  const uint32_t *gs = w.bp_gs512.as<uint32_t>();
  uint32_t *gmin = w.bp_gmin.as<uint32_t>();
  const unsigned nt = (unsigned)w.bp_ntask;
  if (w.bp_w2_bytes == 1)
    pr::k_bp_min<<<nt, pr::BP_BLOCK, 0, ctx.stream>>>(tasks, d2, w.bp_w2.as<uint8_t>(), gs,
                                                      (const VT *)val, g.V, bits, a, gmin);
  else if (w.bp_w2_bytes == 2)
    pr::k_bp_min<<<nt, pr::BP_BLOCK, 0, ctx.stream>>>(tasks, d2, w.bp_w2.as<uint16_t>(), gs,
                                                      (const VT *)val, g.V, bits, a, gmin);
  else
    pr::k_bp_min<<<nt, pr::BP_BLOCK, 0, ctx.stream>>>(tasks, d2, w.bp_w2.as<uint32_t>(), gs,
                                                      (const VT *)val, g.V, bits, a, gmin);
  GRUS_LAUNCH_CHECK();
  if (w.bp_nmulti) {
    pr::k_bp_min_finish<<<(unsigned)w.bp_nmulti, pr::BP_BLOCK, 0, ctx.stream>>>(
        w.bp_multi.as<uint32_t>(), bits, g.V, gmin, a);
    GRUS_LAUNCH_CHECK();
  }
}
}  // namespace

void sssp_binned_round(Graph &g, const BinnedMinArgs &a) {
  PrWork &w = *g.pr;
  // the frontier's distances are below the round's near threshold, so a
  // threshold up to 0xFFFF lets the group values travel as u16 (half the
  // value stream written in phase 1 and read in phase 2)
  const bool narrow = a.T_thr <= 0xFFFFu;
  if (a.front_stamp) {
    const pr::StampDistValue src{a.front_stamp, a.front_id, a.dist};
    if (narrow) binned_min_launch<pr::StampDistValue, uint16_t>(g, w, a, src);
    else binned_min_launch<pr::StampDistValue, uint32_t>(g, w, a, src);
  } else {
    const pr::FrontDistValue src{a.front, a.dist};
    if (narrow) binned_min_launch<pr::FrontDistValue, uint16_t>(g, w, a, src);
    else binned_min_launch<pr::FrontDistValue, uint32_t>(g, w, a, src);
  }
}

void run_pagerank(Graph &g, float alpha, int iters, grus_stats *st) {
  Context &ctx = *g.ctx;
  ctx.activate();
  if (iters < 0) raise(GRUS_E_INVALID, "iteration count");
  if (!(alpha >= 0.0f && alpha <= 1.0f)) raise(GRUS_E_INVALID, "damping factor");
  if (g.V == 0) raise(GRUS_E_INVALID, "empty graph");
  g.make_room(Graph::App::PR);
  if (!g.pr) g.pr = std::make_shared<PrWork>();
  PrWork &w = *g.pr;
  const uint64_t V = g.V;
  const uint32_t layout = ctx.tuning.pr_layout;
  const int bits = ctx.tuning.pr_bin_bits ? (int)ctx.tuning.pr_bin_bits : PB_BITS;
  if (layout > 4 || bits < 6 || bits > 24) raise(GRUS_E_INVALID, "PageRank layout tuning");
  CsrView out = g.csr();
  // binned push (automatic choice on HBM graphs that fit it; bins of at
  // most 2^14 destinations, the LDS accumulators)
  if ((layout == 0 || layout == 4) && g.placement == GRUS_PLACE_HBM) {
    const int bb = std::min(bits, pr::BP_MAX_BITS);
    bool binned = true;
    if (w.bp_for != out.off || w.bp_bits != bb) binned = build_binned(g, w, bb);
    const uint64_t te = ctx.tuning.pr_task_edges ? ctx.tuning.pr_task_edges : BP_TASK_EDGES;
    if (binned && w.bp_task_edges != te) build_binned_tasks(g, w, te);
    if (binned) {
      run_binned(g, w, alpha, iters, st);
      return;
    }
  }
  // relabelling only on request: on RMAT inputs the natural numbering
  // already spreads the hubs over many cache lines, while degree order packs
  // the hottest contributions into a few lines and measured ~5x slower
  const bool relabel = g.symmetric && g.placement == GRUS_PLACE_HBM && layout == 1;
  if (relabel && !w.relabelled) build_relabelled(g, w);
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
    w.nwin = in.E / WIN + 1;
    w.win_row.alloc((w.nwin + 1) * 4);
    w.head_sum.alloc(w.nwin * 4);
    w.tail_sum.alloc(w.nwin * 4);
    w.contrib[0].alloc(V * 4);
    w.contrib[1].alloc(V * 4);
    build_window_rows(ctx, in.off, V, w.nwin, w.win_row.as<uint32_t>());
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
      const pr::PrOp<true> op{c, next, out.off, base, alpha, w.head_sum.as<float>(),
                              w.tail_sum.as<float>()};
      pr::k_pull_window<true><<<nwin, BLOCK, 0, ctx.stream>>>(in, w.win_row.as<uint32_t>(), op,
                                                             w.nwin, blocked);
      GRUS_LAUNCH_CHECK();
      pr::k_fixup<true><<<fix_grid, BLOCK, 0, ctx.stream>>>(in, w.win_row.as<uint32_t>(), op,
                                                           w.nwin);
    } else {
      const pr::PrOp<false> op{c, next, out.off, base, alpha, w.head_sum.as<float>(),
                               w.tail_sum.as<float>()};
      pr::k_pull_window<false><<<nwin, BLOCK, 0, ctx.stream>>>(in, w.win_row.as<uint32_t>(), op,
                                                              w.nwin, blocked);
      GRUS_LAUNCH_CHECK();
      pr::k_fixup<false><<<fix_grid, BLOCK, 0, ctx.stream>>>(in, w.win_row.as<uint32_t>(), op,
                                                            w.nwin);
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
