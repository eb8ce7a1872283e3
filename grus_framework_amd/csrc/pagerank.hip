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
