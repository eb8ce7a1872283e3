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
// Kernel (HBM-bound SpMV over the in-edge CSR): the edge array is cut into
// windows of PR_WIN edges; workgroup k gathers contrib[col[e]] for the
// window's edges with fully coalesced column reads into LDS, then
//   - rows that start and end inside the window are summed from LDS by one
//     lane each and written (rank or next contrib),
//   - the segment of the row that entered from the left (head) is reduced
//     by the whole workgroup (fixed tree) into head_sum[k],
//   - the row that leaves to the right keeps its partial in tail_sum[k].
// A fixup pass completes every row that spans windows by adding its
// partials in window order.  Every sum has a fixed order, so ranks are
// bit-identical run to run.  Each iteration writes only the next contrib
// (r/outdeg); the last one writes the ranks.
#include <algorithm>
#include <cstring>

#include "advance.hpp"
#include "apps.hpp"
#include "prims.hpp"

namespace mgg {

constexpr int PR_WIN = 2048;  // edges per window (8 per lane)

struct PrWork {
  Buf contrib[2], win_row, head_sum, tail_sum;
  uint64_t nwin = 0;
  const void *built_for = nullptr;  // offsets the window table was built on
  uint64_t cap_v = 0;
  // degree-sorted relabelled copy (symmetric, HBM-resident graphs)
  Buf r_off, r_col, perm;  // perm[new id] = original id
  bool relabelled = false;
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

constexpr uint32_t PR_LONG = 16;  // owned-row segments longer than this go to a wave

// a completed row writes its rank (last iteration) or next contrib; a row
// that continues past the window leaves its partial for the fixup pass
template <bool kLast>
__device__ __forceinline__ void store_row(uint32_t r, float s, bool partial, uint64_t k,
                                          const CsrView &in, const uint64_t *out_off,
                                          float *next, float *tail_sum, float base,
                                          float alpha) {
  if (partial) {
    tail_sum[k] = s;
    return;
  }
  float rank = finish_row(s, base, alpha);
  if (kLast) {
    next[r] = rank;
  } else {
    uint64_t d = out_off[r + 1] - out_off[r];
    next[r] = d ? rank / (float)d : 0.0f;
  }
}

template <bool kLast>
__global__ __launch_bounds__(BLOCK) void k_pull_window(CsrView in, const uint64_t *out_off,
                                                       const uint32_t *win_row,
                                                       const float *contrib, float *next,
                                                       float *head_sum, float *tail_sum,
                                                       float base, float alpha, uint64_t nwin) {
  __shared__ float vals[PR_WIN];
  __shared__ float red[BLOCK / WAVE];
  const uint64_t k = blockIdx.x;
  if (k >= nwin) return;
  const uint64_t e0 = k * PR_WIN;
  const uint64_t e1 = e0 + PR_WIN < in.E ? e0 + PR_WIN : in.E;
  const uint32_t r_begin = win_row[k], r_end = win_row[k + 1];
  // gather (coalesced column reads)
#pragma unroll
  for (int j = 0; j < PR_WIN / BLOCK; j++) {
    uint64_t e = e0 + (uint64_t)j * BLOCK + threadIdx.x;
    float x = 0.0f;
    if (e < e1) {
      // the column stream is read once per iteration: load it non-temporally
      // so it does not evict the contrib lines that later windows reuse
      uint32_t c = e < in.split ? __builtin_nontemporal_load(&in.col[e]) : in.col_far[e - in.split];
      x = contrib[c];
    }
    vals[j * BLOCK + threadIdx.x] = x;
  }
  __syncthreads();
  // head segment: edges [e0, start of first owned row) of the row entering from the left
  const uint64_t head_end = (r_begin < in.V) ? (in.off[r_begin] < e1 ? in.off[r_begin] : e1) : e1;
  if (k > 0) {
    float s = 0.0f;
    for (uint64_t e = e0 + threadIdx.x; e < head_end; e += BLOCK) s += vals[e - e0];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane_id() == 0) red[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      float t = 0.0f;
#pragma unroll
      for (int i = 0; i < BLOCK / WAVE; i++) t += red[i];
      head_sum[k] = t;
    }
  }
  // owned rows: short segments are summed by one lane, long ones (queued in
  // LDS) by a whole wave with a fixed shuffle tree, so a window dominated by
  // a few long rows does not serialise on one lane
  __shared__ uint32_t s_long[PR_WIN / PR_LONG + 1];
  __shared__ uint32_t s_nlong;
  if (threadIdx.x == 0) s_nlong = 0;
  __syncthreads();
  for (uint32_t r = r_begin + threadIdx.x; r < r_end; r += BLOCK) {
    uint64_t a = in.off[r], b = in.off[r + 1];
    uint64_t bb = b < e1 ? b : e1;
    if (bb - a > PR_LONG) {
      s_long[atomicAdd(&s_nlong, 1u)] = r;
      continue;
    }
    float s = 0.0f;
    for (uint64_t e = a; e < bb; e++) s += vals[e - e0];
    store_row<kLast>(r, s, b > e1, k, in, out_off, next, tail_sum, base, alpha);
  }
  __syncthreads();
  const uint32_t nlong = s_nlong;
  for (uint32_t i = threadIdx.x / WAVE; i < nlong; i += BLOCK / WAVE) {
    const uint32_t r = s_long[i];
    uint64_t a = in.off[r], b = in.off[r + 1];
    uint64_t bb = b < e1 ? b : e1;
    float s = 0.0f;
    for (uint64_t e = a + lane_id(); e < bb; e += WAVE) s += vals[e - e0];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane_id() == 0) store_row<kLast>(r, s, b > e1, k, in, out_off, next, tail_sum, base, alpha);
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
  const uint32_t mode = ctx.tuning.pr_relabel;
  // off by default: on RMAT inputs the natural numbering already spreads the
  // hubs over many cache lines, while degree order packs the hottest
  // contributions into a few lines and measured ~5x slower (L2 hot-spotting)
  const bool relabel = g.symmetric && g.placement == GRUS_PLACE_HBM && mode == 1;
  if (relabel && !w.relabelled) build_relabelled(g, w);
  CsrView out = g.csr();
  CsrView in = g.csc();
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
  const unsigned nwin = (unsigned)w.nwin;
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
    if (last) {
      pr::k_pull_window<true><<<nwin, BLOCK, 0, ctx.stream>>>(
          in, out.off, w.win_row.as<uint32_t>(), c, next, w.head_sum.as<float>(),
          w.tail_sum.as<float>(), base, alpha, w.nwin);
      GRUS_LAUNCH_CHECK();
      pr::k_fixup<true><<<fix_grid, BLOCK, 0, ctx.stream>>>(
          in, out.off, w.win_row.as<uint32_t>(), w.head_sum.as<float>(), w.tail_sum.as<float>(),
          next, base, alpha, w.nwin);
    } else {
      pr::k_pull_window<false><<<nwin, BLOCK, 0, ctx.stream>>>(
          in, out.off, w.win_row.as<uint32_t>(), c, next, w.head_sum.as<float>(),
          w.tail_sum.as<float>(), base, alpha, w.nwin);
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
