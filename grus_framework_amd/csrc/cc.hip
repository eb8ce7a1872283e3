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
#include <cstring>
#include <string>

#include "advance.hpp"
#include "apps.hpp"
#include "device.hpp"
#include "prims.hpp"

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

// The device buffers of one CC run beyond the graph and the labels.
struct CcSpace {
  uint32_t *sample = nullptr;           // method 0: kSample roots
  uint32_t *q[2] = {nullptr, nullptr};  // method 1: frontier queues, V + 1 each
  uint32_t *stamp = nullptr;            // method 1: round stamps, V + 1
  uint32_t *stamp_id = nullptr;         // method 1: last stamp handed out (host)
  PushWork *push = nullptr;             // method 1: the push operator's scans
};

namespace {
constexpr uint32_t kSample = 1024;  // roots sampled to find the giant component
static_assert(kSample * 4 <= 4096, "the sample is read back through the pinned control block");

void cc_core(Graph &g, int method, uint32_t *label, CcSpace &sp, grus_stats *st) {
  Context &ctx = *g.ctx;
  const uint64_t V = g.V;
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
    // most frequent root among kSample pseudo-random vertices (ties: the
    // smaller root), read back through the pinned control block
    cc::k_sample<<<div_up(kSample, BLOCK), BLOCK, 0, ctx.stream>>>(label, V, kSample, 0x5EEDull,
                                                                  sp.sample);
    GRUS_LAUNCH_CHECK();
    GRUS_HIP(hipMemcpyAsync(ctx.h_ctl, sp.sample, kSample * 4, hipMemcpyDeviceToHost, ctx.stream));
    ctx.sync();
    uint32_t hs[kSample];
    memcpy(hs, ctx.h_ctl, sizeof hs);
    std::sort(hs, hs + kSample);
    uint32_t giant = hs[0], best = 0;
    for (uint32_t i = 0; i < kSample;) {
      uint32_t j = i;
      while (j < kSample && hs[j] == hs[i]) j++;
      if (j - i > best) best = j - i, giant = hs[i];  // ascending: ties keep the smaller
      i = j;
    }
    cc::k_link_rest<<<grid_cap(ctx, V * WAVE, BLOCK, 8), BLOCK, 0, ctx.stream>>>(gr, kRounds,
                                                                               giant, label);
    GRUS_LAUNCH_CHECK();
    cc::k_compress<<<grid_cap(ctx, V, BLOCK), BLOCK, 0, ctx.stream>>>(label, V);
    GRUS_LAUNCH_CHECK();
    launches += 3;
    iterations = kRounds + 1;
  } else {
    auto *cnt = reinterpret_cast<StepCounters *>(ctx.d_ctl.p);
    const StepCounters *hc = reinterpret_cast<const StepCounters *>(ctx.h_ctl);
    uint32_t &stamp_id = *sp.stamp_id;
    cc::k_iota<<<grid_cap(ctx, V, BLOCK), BLOCK, 0, ctx.stream>>>(sp.q[0], V);
    GRUS_LAUNCH_CHECK();
    uint64_t nf = V, mf = g.E;
    int cur = 0;
    while (nf) {
      if (iterations > V + 1) raise(GRUS_E_HIP, "label propagation did not converge");
      if (stamp_id >= 0xFFFFFF00u) {
        fill_u32(ctx, sp.stamp, V + 1, 0);
        stamp_id = 0;
      }
      fill_u64(ctx, reinterpret_cast<uint64_t *>(cnt), sizeof(StepCounters) / 8, 0);
      push_scan_degrees(ctx, *sp.push, gr.off, sp.q[cur], (uint32_t)nf);
      cc::job_t job{label, sp.stamp, gr.off, ++stamp_id};
      push_advance(ctx, *sp.push, gr, sp.q[cur], (uint32_t)nf, mf, job, sp.q[cur ^ 1], cnt);
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

void cc_check(Graph &g, int method) {
  g.ctx->activate();
  if (method != 0 && method != 1) raise(GRUS_E_INVALID, "cc method");
  if (!g.symmetric)
    raise(GRUS_E_UNSUPPORTED, "connected components need a symmetric graph (weak CC)");
}
}  // namespace

void run_cc(Graph &g, int method, grus_stats *st) {
  cc_check(g, method);
  g.make_room(Graph::App::CC);
  if (!g.cc) g.cc = std::make_shared<CcWork>();
  CcWork &w = *g.cc;
  Context &ctx = *g.ctx;
  const uint64_t V = g.V;
  g.result.ensure(V * 4);
  CcSpace sp;
  if (method == 0) {
    w.sample.ensure(kSample * 4);
    sp.sample = w.sample.as<uint32_t>();
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
    sp.q[0] = w.q[0].as<uint32_t>();
    sp.q[1] = w.q[1].as<uint32_t>();
    sp.stamp = w.stamp.as<uint32_t>();
    sp.stamp_id = &w.stamp_id;
    sp.push = &w.push;
  }
  cc_core(g, method, g.result.as<uint32_t>(), sp, st);
}

uint64_t cc_scratch_bytes(Context &ctx, uint64_t V, int method) {
  if (method == 0) return kSample * 4;
  // queues and stamps (3 x (V + 1) u32), degree and scan (2 x (V + 2) u64),
  // the scan's temporaries, each start aligned to 256 bytes
  auto up = [](uint64_t x) { return (x + 255) & ~255ull; };
  return 3 * up((V + 1) * 4) + 2 * up((V + 2) * 8) + up(scan_excl_u64_bytes(ctx, V + 1));
}

void run_cc_in(Graph &g, int method, uint32_t *label, void *scratch, uint64_t scratch_bytes,
               grus_stats *st) {
  cc_check(g, method);
  Context &ctx = *g.ctx;
  const uint64_t V = g.V;
  const uint64_t need = cc_scratch_bytes(ctx, V, method);
  if (need > scratch_bytes)
    raise(GRUS_E_BUDGET, "connected components (method " + std::to_string(method) + ") on " +
                             std::to_string(V) + " vertices need " + std::to_string(need) +
                             " bytes of scratch; " + std::to_string(scratch_bytes) +
                             " are available");
  auto *base = static_cast<unsigned char *>(scratch);
  auto up = [](uint64_t x) { return (x + 255) & ~255ull; };
  CcSpace sp;
  PushWork push;  // borrowed buffers only: its own Bufs stay empty
  uint32_t stamp_id = 0;
  if (method == 0) {
    sp.sample = reinterpret_cast<uint32_t *>(base);
  } else {
    uint64_t at = 0;
    auto take = [&](uint64_t bytes) {
      unsigned char *p = base + at;
      at += up(bytes);
      return p;
    };
    sp.q[0] = reinterpret_cast<uint32_t *>(take((V + 1) * 4));
    sp.q[1] = reinterpret_cast<uint32_t *>(take((V + 1) * 4));
    sp.stamp = reinterpret_cast<uint32_t *>(take((V + 1) * 4));
    push.deg_ext = reinterpret_cast<uint64_t *>(take((V + 2) * 8));
    push.scan_ext = reinterpret_cast<uint64_t *>(take((V + 2) * 8));
    push.tmp_bytes = up(scan_excl_u64_bytes(ctx, V + 1));
    push.tmp = take(push.tmp_bytes);
    // the scratch keeps the last call's contents: stamps start from zero
    fill_u32(ctx, sp.stamp, V + 1, 0);
    sp.stamp_id = &stamp_id;
    sp.push = &push;
  }
  cc_core(g, method, label, sp, st);
}

}  // namespace mgg
