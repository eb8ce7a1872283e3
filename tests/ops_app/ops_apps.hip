// ops_apps.hip -- four apps written against the public operator surface
// (include/grus_ops.hpp) the way the reference's samples are written against
// its kernel / frontier templates: a job_t, updater / generator /
// pull_selector functors, and a host loop "while (!F.finish()) { K(G, F,
// job); F.Next(); job.itr++; }".  Built by __graft_entry__.build() into
// libops_apps.so; tests/test_ops_gpu.py checks every app against the
// oracle.  This is user code: it allocates its own job state.
#include <hip/hip_runtime.h>

#include "../../include/grus_ops.hpp"

using namespace grus;

namespace {

__global__ void k_init_label(uint32_t *label, uint64_t n, uint32_t src, uint32_t at_src,
                             uint32_t others, bool iota) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    label[i] = iota ? (uint32_t)i : (i == src ? at_src : others);
}

uint32_t *new_label(uint64_t n, hipStream_t s, uint32_t src, uint32_t at_src, uint32_t others,
                    bool iota) {
  uint32_t *p = nullptr;
  hip_check(hipMalloc(&p, (n ? n : 1) * 4), "hipMalloc");
  k_init_label<<<1024, 256, 0, s>>>(p, n, src, at_src, others, iota);
  hip_check(hipGetLastError(), "init");
  return p;
}

int finish(uint32_t *label, uint64_t n, hipStream_t s, uint32_t *out) {
  hipError_t e = hipMemcpyAsync(out, label, n * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(label);
  return e == hipSuccess ? 0 : GRUS_E_HIP;
}

// ---- BFS (samples/bfs/bfs.cu:25-87) ----
namespace bfs {
struct job_t {
  uint32_t *label;
  uint32_t itr;
  void prepare() {}
};
struct updater {
  __device__ bool operator()(vtx_t, vtx_t dst, eid_t, job_t &job) const {
    if (job.label[dst] > job.itr + 1) {  // the reference's check-then-store:
      job.label[dst] = job.itr + 1;      // every writer stores the same level
      return true;
    }
    return false;
  }
};
struct generator {
  __device__ void operator()(bool updated, const next_t &next, vtx_t dst) const {
    if (updated) next.push(dst);
  }
};
struct pull_selector {
  __device__ bool operator()(vtx_t id, job_t &job) const { return job.label[id] == INFINIT; }
};
}  // namespace bfs

// ---- SSSP (samples/sssp/sssp.cu:25-60, relax with an atomic min) ----
namespace sssp {
struct job_t {
  uint32_t *dist;
  graph_t G;
  void prepare() {}
};
struct updater {
  __device__ bool operator()(vtx_t src, vtx_t dst, eid_t e, job_t &job) const {
    const uint32_t nd = job.dist[src] + job.G.weight(e);
    if (nd >= job.dist[dst]) return false;
    return atomicMin(&job.dist[dst], nd) > nd;
  }
};
using generator = bfs::generator;
}  // namespace sssp

// ---- CC, min-label propagation (samples/cc/cc.cu:42-91) ----
namespace cc {
struct job_t {
  uint32_t *label;
  void prepare() {}
};
struct updater {
  __device__ bool operator()(vtx_t src, vtx_t dst, eid_t, job_t &job) const {
    const uint32_t ls = job.label[src];
    if (job.label[dst] <= ls) return false;
    return atomicMin(&job.label[dst], ls) > ls;
  }
};
using generator = bfs::generator;
}  // namespace cc

template <class F>
int guarded(F &&f) {
  try {
    f();
    return 0;
  } catch (const error &e) {
    return e.code;
  } catch (...) {
    return GRUS_E_HIP;
  }
}

}  // namespace

extern "C" {

// BFS levels from src: push over a work queue (BFS_single_gpu,
// samples/bfs/bfs.cu:135-173) or pull over bitmaps (BFS_pull_single_gpu,
// :104-134); *iters = levels run
int ops_bfs(grus_ctx *ctx, grus_graph *g, uint32_t src, int pull, uint32_t *out,
            uint32_t *iters) {
  return guarded([&] {
    const hipStream_t s = stream_of(ctx);
    const graph_t G = graph_view(g, pull != 0);
    if (src >= G.numNode) throw error(GRUS_E_INVALID, "source vertex out of range");
    bfs::job_t job{new_label(G.numNode, s, src, 0, INFINIT, false), 0};
    if (pull) {
      Frontier<BITMAP> F;
      F.Init((vtx_t)G.numNode, src, s);
      kernel_pull<bfs::updater, bfs::generator, bfs::pull_selector, bfs::job_t> K;
      while (!F.finish()) {
        K(G, F, job);
        F.Next();
        job.itr++;
      }
    } else {
      Frontier<BDF> F;
      F.Init((vtx_t)G.numNode, src, s);
      kernel<graph_t, Frontier<BDF>, bfs::updater, bfs::generator, bfs::job_t> K;
      while (!F.finish()) {
        K(G, F, job);
        F.Next();
        job.itr++;
      }
    }
    *iters = job.itr;
    check(finish(job.label, G.numNode, s, out));
  });
}

// SSSP distances from src (Bellman-Ford over a work queue)
int ops_sssp(grus_ctx *ctx, grus_graph *g, uint32_t src, uint32_t *out, uint32_t *iters) {
  return guarded([&] {
    const hipStream_t s = stream_of(ctx);
    const graph_t G = graph_view(g, false);
    if (src >= G.numNode) throw error(GRUS_E_INVALID, "source vertex out of range");
    sssp::job_t job{new_label(G.numNode, s, src, 0, INFINIT, false), G};
    Frontier<BDF> F;
    F.Init((vtx_t)G.numNode, src, s);
    kernel<graph_t, Frontier<BDF>, sssp::updater, sssp::generator, sssp::job_t> K;
    uint32_t it = 0;
    while (!F.finish()) {
      K(G, F, job);
      F.Next();
      it++;
    }
    *iters = it;
    check(finish(job.dist, G.numNode, s, out));
  });
}

// CC labels (minimum vertex id): push over bitmaps from a full frontier
// (CC_single_gpu, samples/cc/cc.cu:155-190)
int ops_cc(grus_ctx *ctx, grus_graph *g, uint32_t *out, uint32_t *iters) {
  return guarded([&] {
    const hipStream_t s = stream_of(ctx);
    const graph_t G = graph_view(g, false);
    cc::job_t job{new_label(G.numNode, s, 0, 0, 0, true)};
    Frontier<BITMAP> F;
    F.Init((vtx_t)G.numNode, 0, s, true);
    kernel<graph_t, Frontier<BITMAP>, cc::updater, cc::generator, cc::job_t> K;
    uint32_t it = 0;
    while (!F.finish()) {
      K(G, F, job);
      F.Next();
      it++;
    }
    *iters = it;
    check(finish(job.label, G.numNode, s, out));
  });
}

}  // extern "C"
