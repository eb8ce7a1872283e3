// tune.cpp -- the autotuner: picks the traversal thresholds for one graph.
//
// No reference counterpart (the reference hard-codes BLOCK_SIZE and its
// frontier thresholds, include/common.cuh:30, include/frontier.cuh:140).
// For a graph and a sample of sources it measures the device time of the
// app under candidate settings and keeps the fastest:
//   BFS   coordinate descent over alpha (push->pull: m_f > m_u/alpha), beta
//         (pull->push: n_f < V/beta) and the bottom-up lane/wave cut-off
//         (edges a lane scans before handing a vertex to a wave)
//   SSSP  bucket width delta, and the frontier size at which a round pulls
//   PR    edge layout (binned push / blocked / plain pull), the bin width
//         and the binned layout's task size
// Every BFS / SSSP candidate computes the same result (the settings only
// change the schedule), so tuning never changes those answers.  PageRank
// layouts sum in different orders (ranks agree to float rounding); within
// the binned layout the bin width and task size leave the ranks bit-
// identical, since every edge adds the same fixed-point term to the same
// integer sum.  The winner is written into
// the context's tuning and returned; grus_graph_signature() gives a key
// under which callers can cache it.
#include <algorithm>
#include <vector>

#include "apps.hpp"
#include "graph.hpp"

namespace mgg {

namespace {

// device time of the app over the sample: per source the best of kReps runs
// (timing noise only ever adds time), summed over the sources
constexpr int kReps = 3;
// a candidate must beat the incumbent by this fraction to replace it, so
// that noise cannot walk the setting away from a good default
constexpr double kMinGain = 0.02;

double time_runs(Graph &g, int kind, const std::vector<uint32_t> &srcs) {
  double total = 0;
  grus_stats st;
  const size_t n = kind == GRUS_TUNE_PR ? 1 : srcs.size();
  for (size_t i = 0; i < n; i++) {
    double best = 1e30;
    for (int r = 0; r < kReps; r++) {
      if (kind == GRUS_TUNE_BFS) run_bfs(g, srcs[i], GRUS_DIR_AUTO, &st);
      else if (kind == GRUS_TUNE_SSSP) run_sssp(g, srcs[i], &st);
      else run_pagerank(g, 0.85f, 5, &st);
      best = std::min(best, (double)st.device_ms);
    }
    total += best;
  }
  return total;
}

template <class T>
void sweep(Graph &g, int kind, const std::vector<uint32_t> &srcs, T grus_tuning::*field,
           const std::vector<T> &cands, double *best_ms) {
  Context &ctx = *g.ctx;
  T keep = ctx.tuning.*field;
  for (T c : cands) {
    if (c == keep) continue;
    ctx.tuning.*field = c;
    double ms = time_runs(g, kind, srcs);
    if (ms < *best_ms * (1.0 - kMinGain)) {
      *best_ms = ms;
      keep = c;
    }
  }
  ctx.tuning.*field = keep;
}

}  // namespace

void autotune(Graph &g, int kind, const uint32_t *srcs, int n, grus_tuning *out) {
  if (kind != GRUS_TUNE_BFS && kind != GRUS_TUNE_SSSP && kind != GRUS_TUNE_PR)
    raise(GRUS_E_INVALID, "autotune kind");
  if (kind != GRUS_TUNE_PR && (n < 1 || !srcs)) raise(GRUS_E_INVALID, "need sample sources");
  std::vector<uint32_t> s(srcs, srcs + (kind == GRUS_TUNE_PR ? 0 : n));
  for (uint32_t x : s)
    if (x >= g.V) raise(GRUS_E_INVALID, "sample source out of range");
  Context &ctx = *g.ctx;
  time_runs(g, kind, s);  // warm-up: workspaces allocated, caches built
  double best = time_runs(g, kind, s);  // the incumbent: the current tuning
  if (kind == GRUS_TUNE_BFS) {
    for (int pass = 0; pass < 2; pass++) {
      sweep<double>(g, kind, s, &grus_tuning::bfs_alpha, {2, 4, 8, 14, 24, 48, 96, 192}, &best);
      sweep<double>(g, kind, s, &grus_tuning::bfs_beta, {4, 8, 24, 64, 256, 1024}, &best);
      sweep<uint32_t>(g, kind, s, &grus_tuning::bu_thread_edges, {4, 8, 16, 32, 64, 128}, &best);
    }
  } else if (kind == GRUS_TUNE_SSSP) {
    // candidates around the automatic choice (0 = automatic)
    std::vector<uint32_t> cands = {0};
    for (uint32_t d : {32u, 64u, 128u, 256u, 512u, 1024u, 2048u, 4096u, 8192u, 16384u, 65536u})
      cands.push_back(d);
    sweep<uint32_t>(g, kind, s, &grus_tuning::sssp_delta, cands, &best);
    sweep<uint32_t>(g, kind, s, &grus_tuning::sssp_pull_div, {4, 8, 16, 64, 128, 0xFFFFFFFFu},
                    &best);
    if (g.max_deg <= 32)  // the persistent path
      sweep<uint32_t>(g, kind, s, &grus_tuning::sssp_local_hops, {8, 32, 256}, &best);
  } else {
    // the relabelled layout is left out: it only pays on graphs whose
    // numbering scatters the hubs, and it costs a full relabelling build
    sweep<uint32_t>(g, kind, s, &grus_tuning::pr_layout, {2, 3, 4}, &best);
    if (ctx.tuning.pr_layout == 3)
      sweep<uint32_t>(g, kind, s, &grus_tuning::pr_bin_bits, {12, 14, 16, 18}, &best);
    if (ctx.tuning.pr_layout == 4 || ctx.tuning.pr_layout == 0) {
      sweep<uint32_t>(g, kind, s, &grus_tuning::pr_bin_bits, {12, 13, 14}, &best);
      sweep<uint32_t>(g, kind, s, &grus_tuning::pr_task_edges, {1u << 17, 1u << 18, 1u << 19},
                      &best);
    }
  }
  if (out) *out = ctx.tuning;
}

uint64_t graph_signature(const Graph &g) {
  // V, E, weightedness, max degree and its vertex, isolated count: cheap to
  // get and distinct for the synthetic and loaded inputs we meet in practice
  uint64_t h = 0xcbf29ce484222325ull;
  auto mixin = [&](uint64_t x) {
    for (int i = 0; i < 8; i++) {
      h ^= (x >> (8 * i)) & 0xFF;
      h *= 0x100000001b3ull;
    }
  };
  mixin(g.V);
  mixin(g.E);
  mixin(g.weighted);
  mixin(g.symmetric);
  mixin(g.max_deg);
  mixin(g.max_deg_vertex);
  mixin(g.isolated);
  return h;
}

}  // namespace mgg
