// cc.hip -- connected components, label = minimum vertex id of the component.
//
// Reference: samples/cc/cc.cu (CCInit :35-40 label = id, updater :69-79
// atomicMin(label[dst], label[src]), push driver :155-190 from a full
// frontier).  Its fixed point on a symmetric graph is the minimum vertex id
// of each component; both methods below reach exactly that labelling.
//
// method 0 (default) -- hooking + pointer jumping with neighbour sampling
//   (Sutton et al., IPDPS'18 "Afforest"): link every vertex to its first two
//   neighbours, compress, find the most frequent root by sampling, then
//   link the remaining edges only for vertices outside that component.  A
//   link always hooks the larger root under the smaller one with a CAS, so
//   parent[x] <= x holds throughout and every root is the minimum of its
//   tree.  The giant component's edges (most of an RMAT graph) are never
//   read after the first two per vertex.
// method 1 -- the reference's min-label propagation through the push
//   operator (advance.hpp), frontier deduplicated by round stamps.
#include <algorithm>
#include <unordered_map>
#include <vector>

#include "advance.hpp"
#include "apps.hpp"
#include "device.hpp"

namespace mgg {

struct CcWork {
  Buf q[2], stamp, sample;
  PushWork push;
  uint64_t cap_v = 0;
  uint32_t stamp_id = 0;
};

namespace cc {

__global__ void k_init(uint32_t *label, uint64_t V) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < V; i += stride) label[i] = (uint32_t)i;  // CCInit cc.cu:38
}

__device__ __forceinline__ uint32_t ld(const uint32_t *p) {
  return __atomic_load_n(p, __ATOMIC_RELAXED);
}

// hook the larger root under the smaller (CAS); parent[x] <= x is invariant
__device__ __forceinline__ void link(uint32_t u, uint32_t v, uint32_t *parent) {
  uint32_t p1 = ld(&parent[u]), p2 = ld(&parent[v]);
  while (p1 != p2) {
    uint32_t hi = p1 > p2 ? p1 : p2, lo = p1 > p2 ? p2 : p1;
    uint32_t ph = ld(&parent[hi]);
    if (ph == lo) return;
    if (ph == hi && atomicCAS(&parent[hi], hi, lo) == hi) return;
    p1 = ld(&parent[ld(&parent[hi])]);
    p2 = ld(&parent[lo]);
  }
}

// link neighbour number r (0-based) of every vertex that has one
__global__ void k_link_round(CsrView g, uint32_t r, uint32_t *parent) {
  uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; u < g.V; u += stride) {
    uint64_t e = g.off[u] + r;
    if (e < g.off[u + 1]) link((uint32_t)u, edge_col(g, e), parent);
  }
}

__global__ void k_compress(uint32_t *parent, uint64_t V) {
  uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; v < V; v += stride) {
    uint32_t p = ld(&parent[v]);
    uint32_t pp = ld(&parent[p]);
    while (p != pp) {
      p = pp;
      pp = ld(&parent[p]);
    }
    parent[v] = p;
  }
}

__global__ void k_sample(const uint32_t *parent, uint64_t V, uint32_t n, uint64_t seed,
                         uint32_t *out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = parent[mix64(seed + i) % V];
}

// wave per vertex outside the giant component: link edges from number `from` on
__global__ __launch_bounds__(BLOCK) void k_link_rest(CsrView g, uint32_t from, uint32_t giant,
                                                     uint32_t *parent) {
  uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE;
  uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) / WAVE;
  for (uint64_t u = wave; u < g.V; u += nwaves) {
    if (ld(&parent[u]) == giant) continue;
    uint64_t end = g.off[u + 1];
    for (uint64_t e = g.off[u] + from + lane_id(); e < end; e += WAVE)
      link((uint32_t)u, edge_col(g, e), parent);
  }
}

__global__ void k_iota(uint32_t *q, uint64_t n) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) q[i] = (uint32_t)i;
}

// job_t of samples/cc/cc.cu:42-67 for min-label propagation
struct job_t {
  static constexpr bool kFar = false;
  uint32_t *label;
  uint32_t *stamp;
  const uint64_t *out_off;
  uint32_t stamp_id;
  // updater cc.cu:69-79
  __device__ __forceinline__ bool update(vtx_t src, vtx_t dst, eid_t, weight_t, uint32_t *) const {
    uint32_t ls = ld(&label[src]);
    if (ld(&label[dst]) <= ls) return false;
    return atomicMin(&label[dst], ls) > ls;
  }
  // generator cc.cu:86-91 (the byte CAS flag becomes a round stamp)
  __device__ __forceinline__ int generate(bool updated, vtx_t dst, uint32_t) const {
    // stamps are unique per round: a plain read showing this round's is exact
    if (!updated || stamp[dst] == stamp_id) return 0;
    return atomicExch(&stamp[dst], stamp_id) != stamp_id ? 1 : 0;
  }
  __device__ __forceinline__ uint64_t out_degree(vtx_t v) const {
    return out_off[v + 1] - out_off[v];
  }
};

}  // namespace cc

void run_cc(Graph &g, int method, grus_stats *st) {
  Context &ctx = *g.ctx;
  ctx.activate();
  if (method != 0 && method != 1) raise(GRUS_E_INVALID, "cc method");
  if (!g.symmetric)
    raise(GRUS_E_UNSUPPORTED, "connected components need a symmetric graph (weak CC)");
  g.make_room(Graph::App::CC);
  if (!g.cc) g.cc = std::make_shared<CcWork>();
  CcWork &w = *g.cc;
  const uint64_t V = g.V;
  g.result.ensure(V * 4);
  uint32_t *label = g.result.as<uint32_t>();
  const CsrView gr = g.csr();
  uint64_t launches = 0;
  uint32_t iterations = 0;
  GRUS_HIP(hipEventRecord(ctx.ev0, ctx.stream));
  cc::k_init<<<grid_cap(ctx, V, BLOCK), BLOCK, 0, ctx.stream>>>(label, V);
  GRUS_LAUNCH_CHECK();
  launches++;
  if (method == 0) {
    const uint32_t kRounds = 2;
    for (uint32_t r = 0; r < kRounds; r++) {
      cc::k_link_round<<<grid_cap(ctx, V, BLOCK), BLOCK, 0, ctx.stream>>>(gr, r, label);
      GRUS_LAUNCH_CHECK();
      cc::k_compress<<<grid_cap(ctx, V, BLOCK), BLOCK, 0, ctx.stream>>>(label, V);
      GRUS_LAUNCH_CHECK();
      launches += 2;
    }
    // most frequent root among 1024 pseudo-random vertices
    const uint32_t ns = 1024;
    w.sample.ensure(ns * 4);
    cc::k_sample<<<div_up(ns, BLOCK), BLOCK, 0, ctx.stream>>>(label, V, ns, 0x5EEDull,
                                                             w.sample.as<uint32_t>());
    GRUS_LAUNCH_CHECK();
    std::vector<uint32_t> hs(ns);
    GRUS_HIP(hipMemcpyAsync(hs.data(), w.sample.p, ns * 4, hipMemcpyDeviceToHost, ctx.stream));
    ctx.sync();
    std::unordered_map<uint32_t, uint32_t> freq;
    uint32_t giant = hs[0], best = 0;
    for (uint32_t x : hs) {
      uint32_t c = ++freq[x];
      if (c > best || (c == best && x < giant)) best = c, giant = x;
    }
    cc::k_link_rest<<<grid_cap(ctx, V * WAVE, BLOCK, 8), BLOCK, 0, ctx.stream>>>(gr, kRounds,
                                                                               giant, label);
    GRUS_LAUNCH_CHECK();
    cc::k_compress<<<grid_cap(ctx, V, BLOCK), BLOCK, 0, ctx.stream>>>(label, V);
    GRUS_LAUNCH_CHECK();
    launches += 3;
    iterations = kRounds + 1;
  } else {
    if (w.cap_v < V) {
      w.q[0].alloc((V + 1) * 4);
      w.q[1].alloc((V + 1) * 4);
      w.stamp.alloc((V + 1) * 4);
      fill_u32(ctx, w.stamp.as<uint32_t>(), V + 1, 0);
      w.push.ensure(V);
      w.cap_v = V;
      w.stamp_id = 0;
    }
    auto *cnt = reinterpret_cast<StepCounters *>(ctx.d_ctl.p);
    const StepCounters *hc = reinterpret_cast<const StepCounters *>(ctx.h_ctl);
    cc::k_iota<<<grid_cap(ctx, V, BLOCK), BLOCK, 0, ctx.stream>>>(w.q[0].as<uint32_t>(), V);
    GRUS_LAUNCH_CHECK();
    uint64_t nf = V, mf = g.E;
    int cur = 0;
    while (nf) {
      if (iterations > V + 1) raise(GRUS_E_HIP, "label propagation did not converge");
      if (w.stamp_id >= 0xFFFFFF00u) {
        fill_u32(ctx, w.stamp.as<uint32_t>(), V + 1, 0);
        w.stamp_id = 0;
      }
      GRUS_HIP(hipMemsetAsync(cnt, 0, sizeof(StepCounters), ctx.stream));
      push_scan_degrees(ctx, w.push, gr.off, w.q[cur].as<uint32_t>(), (uint32_t)nf);
      cc::job_t job{label, w.stamp.as<uint32_t>(), gr.off, ++w.stamp_id};
      push_advance(ctx, w.push, gr, w.q[cur].as<uint32_t>(), (uint32_t)nf, mf, job,
                   w.q[cur ^ 1].as<uint32_t>(), cnt);
      launches += 4;
      ctx.read_ctl(2);
      nf = hc->next_count;
      mf = hc->next_edges;
      cur ^= 1;
      iterations++;
    }
  }
  GRUS_HIP(hipEventRecord(ctx.ev1, ctx.stream));
  GRUS_HIP(hipEventSynchronize(ctx.ev1));
  float ms = 0;
  GRUS_HIP(hipEventElapsedTime(&ms, ctx.ev0, ctx.ev1));
  if (st) {
    *st = grus_stats{};
    st->device_ms = ms;
    st->iterations = iterations;
    st->vertices_reached = V;
    st->edges_reached = g.E;
    st->kernel_launches = launches;
  }
}

}  // namespace mgg
