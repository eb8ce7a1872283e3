// capi.cpp -- the extern "C" boundary of libgrus_hip.so (include/grus_c.h).
//
// Every entry point runs its body under guard(): engine errors become a
// negative status code plus a thread-local message; nothing exits the
// process (the reference's H_ERR does, include/common.cuh:57-68).
#include <cstring>
#include <memory>
#include <new>
#include <string>

#include "apps.hpp"
#include "graph.hpp"
#include "multi.hpp"
#include "serve.hpp"

struct grus_ctx {
  std::unique_ptr<mgg::Context> c;
};
struct grus_graph {
  std::unique_ptr<mgg::Graph> g;
  int last_result = 0;  // 0 none, 1 u32 per vertex, 2 f32 per vertex
};
struct grus_server {
  std::unique_ptr<mgg::Server> s;
};

namespace {
thread_local std::string g_err;

template <class F>
int guard(F &&f) {
  try {
    f();
    return GRUS_OK;
  } catch (const mgg::Error &e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_err = "host allocation failed";
    return GRUS_E_NOMEM;
  } catch (const std::exception &e) {
    g_err = e.what();
    return GRUS_E_INVALID;
  }
}

void need(const void *p, const char *what) {
  if (!p) mgg::raise(GRUS_E_INVALID, std::string("null argument: ") + what);
}

void copy_out(grus_graph *g, void *out) {
  if (!out) return;
  GRUS_HIP(hipMemcpy(out, g->g->result.p, g->g->V * 4, hipMemcpyDeviceToHost));
}
}  // namespace

extern "C" {

const char *grus_last_error(void) { return g_err.c_str(); }
const char *grus_version(void) { return "grus-mi355x 0.1 (gfx950)"; }

int grus_device_count(int *count) {
  return guard([&] {
    need(count, "count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      n = 0;
    }
    *count = n;
  });
}

int grus_init(int device, grus_ctx **out) {
  return guard([&] {
    need(out, "out");
    *out = nullptr;
    auto c = std::make_unique<grus_ctx>();
    c->c = std::make_unique<mgg::Context>(device);  // leaves `device` current
    mgg::scratch_setup(device);  // the scratch area of grus_cc_noalloc
    *out = c.release();
  });
}

int grus_fini(grus_ctx *ctx) {
  return guard([&] { delete ctx; });
}

int grus_get_tuning(grus_ctx *ctx, grus_tuning *out) {
  return guard([&] {
    need(ctx, "ctx");
    need(out, "out");
    *out = ctx->c->tuning;
  });
}

int grus_set_cu_mask(grus_ctx *ctx, const uint32_t *mask, uint32_t mask_words) {
  return guard([&] {
    need(ctx, "ctx");
    if (mask_words) need(mask, "mask");
    ctx->c->set_cu_mask(mask, mask_words);
  });
}

int grus_cu_count(grus_ctx *ctx, int *active, int *total) {
  return guard([&] {
    need(ctx, "ctx");
    if (active) *active = ctx->c->cu_count;
    if (total) *total = ctx->c->cu_total;
  });
}

int grus_set_device_budget(grus_ctx *ctx, uint64_t bytes) {
  return guard([&] {
    need(ctx, "ctx");
    mgg::devmem_set_budget(ctx->c->device, bytes);
  });
}

int grus_device_memory(grus_ctx *ctx, uint64_t *held, uint64_t *peak, int reset) {
  return guard([&] {
    need(ctx, "ctx");
    const int d = ctx->c->device;
    if (held) *held = mgg::devmem_live(d);
    if (peak) *peak = mgg::devmem_peak(d);
    if (reset) mgg::devmem_reset_peak(d);
  });
}

int grus_set_tuning(grus_ctx *ctx, const grus_tuning *in) {
  return guard([&] {
    need(ctx, "ctx");
    need(in, "in");
    if (in->bfs_alpha < 0 || in->bfs_beta < 0) mgg::raise(GRUS_E_INVALID, "negative threshold");
    if (in->pr_layout > 4) mgg::raise(GRUS_E_INVALID, "pr_layout");
    if (in->pr_bin_bits != 0 && (in->pr_bin_bits < 6 || in->pr_bin_bits > 24))
      mgg::raise(GRUS_E_INVALID, "pr_bin_bits outside [6, 24]");
    if (in->sssp_local_hops > 256) mgg::raise(GRUS_E_INVALID, "sssp_local_hops above 256");
    if (in->pr_task_edges != 0 && in->pr_task_edges < 256)
      mgg::raise(GRUS_E_INVALID, "pr_task_edges below 256");
    if (in->sssp_pull_layout > 2) mgg::raise(GRUS_E_INVALID, "sssp_pull_layout");
    ctx->c->tuning = *in;
  });
}

// ---------------------------------------------------------------- .gr I/O
int grus_gr_header(const char *path, uint64_t *V, uint64_t *E, int *weighted) {
  return guard([&] {
    need(path, "path");
    mgg::GrHeader h = mgg::gr_read_header(path);
    if (V) *V = h.V;
    if (E) *E = h.E;
    if (weighted) *weighted = h.size_edge_ty != 0;
  });
}

int grus_gr_read(const char *path, uint64_t *off, uint32_t *col, uint32_t *w) {
  return guard([&] {
    need(path, "path");
    need(off, "off");
    mgg::GrHeader h = mgg::gr_read_header(path);
    if (h.E) need(col, "col");
    mgg::gr_read(path, off, col, w);
  });
}

int grus_gr_write(const char *path, uint64_t V, uint64_t E, const uint64_t *off,
                  const uint32_t *col, const uint32_t *w) {
  return guard([&] {
    need(path, "path");
    need(off, "off");
    if (E) need(col, "col");
    mgg::gr_write(path, V, E, off, col, w);
  });
}

// ---------------------------------------------------------------- graphs
int grus_graph_from_csr(grus_ctx *ctx, uint64_t V, uint64_t E, const uint64_t *off,
                        const uint32_t *col, const uint32_t *w, int symmetric,
                        const grus_place *place, grus_graph **out) {
  return guard([&] {
    need(ctx, "ctx");
    need(out, "out");
    *out = nullptr;
    auto g = std::make_unique<grus_graph>();
    g->g = mgg::graph_from_host(ctx->c.get(), V, E, off, col, w, symmetric != 0, place);
    *out = g.release();
  });
}

int grus_graph_load_gr(grus_ctx *ctx, const char *path, int need_weight, int symmetric,
                       const grus_place *place, grus_graph **out) {
  return guard([&] {
    need(ctx, "ctx");
    need(path, "path");
    need(out, "out");
    *out = nullptr;
    mgg::GrHeader h = mgg::gr_read_header(path);
    std::unique_ptr<uint64_t[]> off(new uint64_t[h.V + 1]);
    std::unique_ptr<uint32_t[]> col(new uint32_t[h.E ? h.E : 1]);
    std::unique_ptr<uint32_t[]> w;
    if (need_weight) w.reset(new uint32_t[h.E ? h.E : 1]);
    mgg::gr_read(path, off.get(), col.get(), w.get());
    auto g = std::make_unique<grus_graph>();
    g->g = mgg::graph_from_host(ctx->c.get(), h.V, h.E, off.get(), col.get(), w.get(),
                                symmetric != 0, place);
    *out = g.release();
  });
}

int grus_graph_gen_rmat(grus_ctx *ctx, int scale, int edge_factor, double a, double b, double c,
                        uint64_t seed, uint32_t wlo, uint32_t whi, uint64_t wseed,
                        const grus_place *place, grus_graph **out) {
  return guard([&] {
    need(ctx, "ctx");
    need(out, "out");
    *out = nullptr;
    auto g = std::make_unique<grus_graph>();
    g->g = mgg::graph_gen_rmat(ctx->c.get(), scale, edge_factor, a, b, c, seed, wlo, whi, wseed,
                               place);
    *out = g.release();
  });
}

int grus_graph_gen_grid(grus_ctx *ctx, uint32_t rows, uint32_t cols, uint32_t wlo, uint32_t whi,
                        uint64_t wseed, const grus_place *place, grus_graph **out) {
  return guard([&] {
    need(ctx, "ctx");
    need(out, "out");
    *out = nullptr;
    auto g = std::make_unique<grus_graph>();
    g->g = mgg::graph_gen_grid(ctx->c.get(), rows, cols, wlo, whi, wseed, place);
    *out = g.release();
  });
}

int grus_graph_view(grus_graph *g, int in_edges, grus_csr_view *out) {
  return guard([&] {
    need(g, "g");
    need(out, "out");
    if (in_edges != 0 && in_edges != 1) mgg::raise(GRUS_E_INVALID, "in_edges is 0 or 1");
    mgg::Graph &G = *g->g;
    G.ctx->activate();
    const mgg::CsrView v = in_edges ? G.csc() : G.csr();
    *out = grus_csr_view{v.off, v.col, v.w, v.col_far, v.w_far, v.split, v.V, v.E};
  });
}

int grus_ctx_stream(grus_ctx *ctx, void **hip_stream) {
  return guard([&] {
    need(ctx, "ctx");
    need(hip_stream, "hip_stream");
    *hip_stream = static_cast<void *>(ctx->c->stream);
  });
}

int grus_graph_info_get(grus_graph *g, grus_graph_info *out) {
  return guard([&] {
    need(g, "g");
    need(out, "out");
    mgg::Graph &G = *g->g;
    *out = grus_graph_info{};
    out->V = G.V;
    out->E = G.E;
    out->weighted = G.weighted;
    out->symmetric = G.symmetric;
    out->placement = G.placement;
    out->max_degree_vertex = G.max_deg_vertex;
    out->max_degree = G.max_deg;
    out->isolated = G.isolated;
    out->device_bytes = G.device_bytes();
    out->host_bytes = G.host_bytes;
  });
}

int grus_graph_copy_csr(grus_graph *g, uint64_t *off, uint32_t *col, uint32_t *w) {
  return guard([&] {
    need(g, "g");
    mgg::graph_copy_csr(*g->g, off, col, w);
  });
}

int grus_graph_destroy(grus_graph *g) {
  return guard([&] { delete g; });
}

// ---------------------------------------------------------------- apps
int grus_bfs(grus_graph *g, uint32_t src, int dir_mode, uint32_t *levels_out, grus_stats *stats) {
  return guard([&] {
    need(g, "g");
    g->last_result = 0;
    mgg::run_bfs(*g->g, src, dir_mode, stats);
    g->last_result = 1;
    copy_out(g, levels_out);
  });
}

int grus_bfs_2s(grus_graph *g, uint32_t src, uint64_t stage_edges, uint32_t *levels_out,
                grus_stats *stats) {
  return guard([&] {
    need(g, "g");
    g->last_result = 0;
    mgg::run_bfs_2s(*g->g, src, stage_edges, stats);
    g->last_result = 1;
    copy_out(g, levels_out);
  });
}

int grus_sssp(grus_graph *g, uint32_t src, uint32_t *dist_out, grus_stats *stats) {
  return guard([&] {
    need(g, "g");
    g->last_result = 0;
    mgg::run_sssp(*g->g, src, stats);
    g->last_result = 1;
    copy_out(g, dist_out);
  });
}

int grus_pagerank(grus_graph *g, float alpha, int iters, float *rank_out, grus_stats *stats) {
  return guard([&] {
    need(g, "g");
    g->last_result = 0;
    mgg::run_pagerank(*g->g, alpha, iters, stats);
    g->last_result = 2;
    copy_out(g, rank_out);
  });
}

int grus_cc(grus_graph *g, int method, uint32_t *labels_out, grus_stats *stats) {
  return guard([&] {
    need(g, "g");
    g->last_result = 0;
    mgg::run_cc(*g->g, method, stats);
    g->last_result = 1;
    copy_out(g, labels_out);
  });
}

int grus_cc_noalloc(grus_graph *g, int method, uint32_t *labels_dev, grus_stats *stats) {
  return guard([&] {
    need(g, "g");
    need(labels_dev, "labels_dev");
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, labels_dev) != hipSuccess) {
      (void)hipGetLastError();
      mgg::raise(GRUS_E_INVALID, "labels_dev is not device memory");
    }
    if (a.type != hipMemoryTypeDevice && a.type != hipMemoryTypeManaged)
      mgg::raise(GRUS_E_INVALID, "labels_dev is not device memory");
    if (a.type == hipMemoryTypeDevice && a.device != g->g->ctx->device)
      mgg::raise(GRUS_E_INVALID, "labels_dev is on another device than the graph");
    mgg::run_cc_noalloc(*g->g, method, labels_dev, stats);
  });
}

int grus_result_copy(grus_graph *g, void *out) {
  return guard([&] {
    need(g, "g");
    need(out, "out");
    if (!g->last_result) mgg::raise(GRUS_E_INVALID, "no result held for this graph");
    copy_out(g, out);
  });
}

int grus_result_digest(grus_graph *g, uint64_t *reached, uint64_t *checksum) {
  return guard([&] {
    need(g, "g");
    if (g->last_result != 1) mgg::raise(GRUS_E_INVALID, "no u32 result held for this graph");
    uint64_t r = 0, c = 0;
    mgg::row_digest(*g->g->ctx, g->g->result.as<uint32_t>(), g->g->V, &r, &c);
    if (reached) *reached = r;
    if (checksum) *checksum = c;
  });
}

int grus_bfs_multi(grus_graph *g, const uint32_t *srcs, int k, uint32_t *out, grus_stats *stats) {
  return guard([&] {
    need(g, "g");
    need(srcs, "srcs");
    mgg::run_bfs_multi(*g->g, srcs, k, out, stats);
  });
}

int grus_sssp_multi(grus_graph *g, const uint32_t *srcs, int k, uint32_t *out,
                    grus_stats *stats) {
  return guard([&] {
    need(g, "g");
    need(srcs, "srcs");
    mgg::run_sssp_multi(*g->g, srcs, k, out, stats);
  });
}

// ---------------------------------------------------------------- tuning
int grus_autotune(grus_graph *g, int kind, const uint32_t *srcs, int n, grus_tuning *out) {
  return guard([&] {
    need(g, "g");
    mgg::autotune(*g->g, kind, srcs, n, out);
    g->last_result = 0;
  });
}

int grus_graph_signature(grus_graph *g, uint64_t *sig) {
  return guard([&] {
    need(g, "g");
    need(sig, "sig");
    *sig = mgg::graph_signature(*g->g);
  });
}

// ---------------------------------------------------------------- serving
int grus_server_create(grus_graph *g, const grus_server_opts *opts, grus_server **out) {
  return guard([&] {
    need(g, "g");
    need(out, "out");
    *out = nullptr;
    auto s = std::make_unique<grus_server>();
    s->s = mgg::make_server(*g->g, opts);
    *out = s.release();
  });
}

int grus_server_submit(grus_server *s, int kind, uint32_t src, uint32_t *out_host,
                       uint64_t *id_out) {
  return guard([&] {
    need(s, "s");
    uint64_t id = s->s->submit(kind, src, out_host);
    if (id_out) *id_out = id;
  });
}

int grus_server_poll(grus_server *s, grus_query_result *res, int timeout_ms) {
  return guard([&] {
    need(s, "s");
    need(res, "res");
    s->s->poll(res, timeout_ms);
  });
}

int grus_server_destroy(grus_server *s) {
  return guard([&] { delete s; });
}

int grus_server_stats_get(grus_server *s, grus_server_stats *out) {
  return guard([&] {
    need(s, "s");
    need(out, "out");
    s->s->stats(out);
  });
}

}  // extern "C"
