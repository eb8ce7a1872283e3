// apps.hpp -- host drivers of the traversal apps (one per reference sample).
#pragma once

#include "graph.hpp"

namespace mgg {

// samples/bfs/bfs.cu:104-173
void run_bfs(Graph &g, uint32_t src, int dir, grus_stats *st);
// samples/sssp/sssp.cu:91-170
void run_sssp(Graph &g, uint32_t src, grus_stats *st);
// samples/pagerank/pr.cu:134-205 (redefined as fixed-iteration pull)
void run_pagerank(Graph &g, float alpha, int iters, grus_stats *st);
// samples/cc/cc.cu:124-190
void run_cc(Graph &g, int method, grus_stats *st);
// batched multi-source (no reference counterpart); results at g.ms output
void run_bfs_multi(Graph &g, const uint32_t *srcs, int k, uint32_t *out_dev, grus_stats *st);
void run_sssp_multi(Graph &g, const uint32_t *srcs, int k, uint32_t *out_dev, grus_stats *st);

// autotuner (tune.cpp)
void autotune(Graph &g, int kind, const uint32_t *srcs, int n, grus_tuning *out);
uint64_t graph_signature(const Graph &g);

}  // namespace mgg
