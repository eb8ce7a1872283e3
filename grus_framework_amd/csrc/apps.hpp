// apps.hpp -- host drivers of the traversal apps (one per reference sample).
#pragma once

#include "graph.hpp"

namespace mgg {

// samples/bfs/bfs.cu:104-173
void run_bfs(Graph &g, uint32_t src, int dir, grus_stats *st);
// samples/bfs_2s/bfs_2s.cu:94-138 (two-stage: frontier rows gathered into a
// device buffer of stage_edges edges, 0 = default, then expanded)
void run_bfs_2s(Graph &g, uint32_t src, uint64_t stage_edges, grus_stats *st);
// samples/sssp/sssp.cu:91-170
void run_sssp(Graph &g, uint32_t src, grus_stats *st);
// samples/pagerank/pr.cu:134-205 (redefined as fixed-iteration pull)
void run_pagerank(Graph &g, float alpha, int iters, grus_stats *st);
// samples/cc/cc.cu:124-190
void run_cc(Graph &g, int method, grus_stats *st);
// the same into the caller's device labels, with every other device buffer
// carved from `scratch` (no allocation); GRUS_E_BUDGET when it is too small
void run_cc_in(Graph &g, int method, uint32_t *labels_dev, void *scratch, uint64_t scratch_bytes,
               grus_stats *st);
uint64_t cc_scratch_bytes(Context &ctx, uint64_t V, int method);
// the allocation-free entry point (noalloc.hip): run_cc_in over the
// library's per-device scratch area, set up by scratch_setup at context init
void scratch_setup(int device);
void run_cc_noalloc(Graph &g, int method, uint32_t *labels_dev, grus_stats *st);
// batched multi-source (no reference counterpart); results at g.ms output
void run_bfs_multi(Graph &g, const uint32_t *srcs, int k, uint32_t *out_dev, grus_stats *st);
void run_sssp_multi(Graph &g, const uint32_t *srcs, int k, uint32_t *out_dev, grus_stats *st);

// one SSSP pull round over the binned push layout (pagerank.hip): every
// out-edge (u, v) of a frontier vertex u offers dist[u] + w to v, and v
// takes the minimum when it is below dist[v]; improved vertices join the
// near queue (below T_thr) or the far list (once per far epoch), as in the
// window pull round of sssp.hip
struct StepCounters;
struct BinnedMinArgs {
  // the frontier: vertices u with front_stamp[u] == front_id, or (when
  // front_stamp is null) the set bits of the bitmap front
  const uint32_t *front_stamp;
  uint32_t front_id;
  const uint64_t *front;
  uint32_t *dist;
  const uint64_t *out_off;
  uint32_t T_thr, near_id, far_id;
  uint32_t *near_stamp, *far_stamp;
  uint32_t *qn, *qfar;  // near queue; far list at its current end
  StepCounters *cnt;    // next_count, next_edges, far_count
};
// builds the binned layout, its tasks and its weight stream when missing;
// false where the layout does not apply (then pull rounds use the window)
bool sssp_binned_ready(Graph &g);
void sssp_binned_round(Graph &g, const BinnedMinArgs &a);

// autotuner (tune.cpp)
void autotune(Graph &g, int kind, const uint32_t *srcs, int n, grus_tuning *out);
uint64_t graph_signature(const Graph &g);

}  // namespace mgg
