/*
 * grus_c.h -- C-ABI of libgrus_hip.so, the MI355X (gfx950) traversal engine.
 *
 * The reference (wpybtw/Grus_framework) exposes its engine only as CUDA C++
 * header templates; there is no FFI.  Each entry point below replaces the
 * reference interface named in its comment (paths relative to the reference
 * root).  Inside the library the reference's operator surface is kept:
 * graph / frontier / job classes and app functors (updater, generator,
 * pull_selector) compiled into the HIP translation units.
 *
 * Conventions
 *   - Every function returns GRUS_OK (0) or a negative GRUS_E_* code and
 *     never exits the process (the reference's H_ERR calls exit(),
 *     include/common.cuh:57-68).  grus_last_error() gives a thread-local
 *     message for the last failure on the calling thread.
 *   - Handles own all device memory they allocate; the caller owns every
 *     host buffer it passes.  Inputs are copied in at create time; results
 *     are copied out before a call returns (output pointers may be NULL to
 *     leave the result on the device, see grus_result_copy).
 *   - One context = one device + one HIP stream.  Calls on one context or on
 *     graphs of one context are serialized by the caller (not reentrant).
 *     A server (grus_server_*) owns its own worker thread and streams.
 *   - Vertex ids are uint32; edge offsets are uint64 (the reference's u32
 *     xadj, include/graph_loader.cuh:118, overflows past 2^32 edges).
 *   - Unreached vertices hold GRUS_INFINIT (0xFFFFFFFE, the reference's
 *     INFINIT, include/common.cuh:52).
 */
#ifndef GRUS_C_H
#define GRUS_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GRUS_INFINIT 0xFFFFFFFEu

/* status codes */
#define GRUS_OK 0
#define GRUS_E_INVALID -1     /* bad argument */
#define GRUS_E_NOMEM -2       /* device or host allocation failed */
#define GRUS_E_HIP -3         /* HIP runtime error */
#define GRUS_E_IO -4          /* file open/read/write failure */
#define GRUS_E_FORMAT -5      /* malformed .gr file */
#define GRUS_E_BUDGET -6      /* request exceeds the memory budget / cap */
#define GRUS_E_TIMEOUT -7     /* deadline passed */
#define GRUS_E_UNSUPPORTED -8 /* operation not available for this graph */
#define GRUS_E_NODEVICE -9    /* no usable GPU */
#define GRUS_E_CLOSED -10     /* server shut down */

typedef struct grus_ctx grus_ctx;
typedef struct grus_graph grus_graph;
typedef struct grus_server grus_server;

/* Placement of graph arrays.  Replaces graph_base::Set_Mem_Policy
 * (include/graph.cuh:170-251) and its --opt/--pf/--ab/--rm/--pl flags
 * (include/app.cuh:27-31). */
#define GRUS_PLACE_HBM 0      /* hipMalloc: whole CSR resident in HBM */
#define GRUS_PLACE_MANAGED 1  /* hipMallocManaged + advice/prefetch per flags */
#define GRUS_PLACE_CAPPED 2   /* offsets + edge prefix in HBM up to hbm_cap,
                                 remaining edges in host-coherent pinned memory */
typedef struct grus_place {
  int mode;               /* GRUS_PLACE_* */
  uint64_t hbm_cap_bytes; /* CAPPED: device bytes the graph may use (0 = no cap) */
  int prefetch;           /* MANAGED: --pf  prefetch arrays to the device */
  int accessed_by;        /* MANAGED: --ab  hipMemAdviseSetAccessedBy */
  int read_mostly;        /* MANAGED: --rm  hipMemAdviseSetReadMostly */
  int preferred_location; /* MANAGED: --pl  hipMemAdviseSetPreferredLocation */
  uint64_t device_budget_bytes; /* CAPPED: cap of the whole device footprint
                             (0 = none).  The graph's HBM part is sized so that
                             it, every other device buffer the library holds
                             and the largest app workspace fit under it, and
                             the budget is then installed for the device as
                             grus_set_device_budget does (the reference caps
                             the whole device, tools/mem_limit.cu:6-18, and
                             places from cudaMemGetInfo, include/graph.cuh:
                             171-205).  The workspace reserve is
                             48 (V + 2) + 16 (E / 2048 + 2) + 64 MiB bytes.
                             The device build of a generated graph and its
                             temporaries come before the cap, as the
                             reference's loader fills host memory first. */
} grus_place;

/* Direction policy for BFS (samples/bfs/bfs.cu:135-173 push, :104-134 pull) */
#define GRUS_DIR_AUTO 0 /* direction-optimizing: push <-> pull switch */
#define GRUS_DIR_PUSH 1 /* top-down only (the reference's "plain method") */
#define GRUS_DIR_PULL 2 /* bottom-up only */

typedef struct grus_stats {
  double device_ms;         /* device time of the traversal (HIP events) */
  uint32_t iterations;      /* levels / rounds / PR iterations */
  uint32_t pull_iterations; /* how many of them ran bottom-up (pull) */
  uint64_t vertices_reached; /* V_r */
  uint64_t edges_reached;    /* E_r: sum of out-degrees of reached vertices */
  uint64_t kernel_launches;
  uint64_t reserved[4];
} grus_stats;

/* Tunables of the traversal engine (the autotuner writes these). */
typedef struct grus_tuning {
  double bfs_alpha;       /* push->pull when m_f > m_u / alpha   (default 14) */
  double bfs_beta;        /* pull->push when n_f < V / beta      (default 64) */
  uint32_t bu_thread_edges; /* bottom-up: edges a lane scans before handing
                               the vertex to a wave (default 32) */
  uint32_t sssp_delta;    /* SSSP bucket width (0 = automatic: mean weight x
                             max(2, 32 / mean degree); on low-degree graphs,
                             which run one persistent launch, x max(2, 80 /
                             mean degree); unweighted graphs: one bucket) */
  uint32_t pr_layout;     /* PageRank edge layout: 0 automatic (= 4 where it
                             applies, else 2); 1 degree-sorted relabelled copy
                             (symmetric HBM graphs); 2 plain pull over the
                             in-edges; 3 blocked (propagation blocking over
                             bins of 2^pr_bin_bits destinations; HBM graphs
                             with E <= 2^31, else 2); 4 binned push (values
                             streamed per destination bin of at most 2^14
                             vertices and summed in LDS as 64-bit fixed point;
                             HBM graphs with E < 2^32 and at most 8192 bins,
                             else 2).  Layouts 2 and 3 give bit-identical
                             ranks; 1 and 4 sum in another order or precision
                             (same ranks up to float rounding).  Every layout
                             is bit-identical run to run. */
  uint32_t pr_bin_bits;   /* blocked / binned layout bin width, log2 (0 =
                             automatic: 14; layout 4 caps it at 14) */
  uint32_t sssp_pull_div; /* SSSP: a round pulls (each unsettled vertex takes
                             the min over its in-neighbours in the frontier)
                             when the frontier's out-edges exceed
                             E / sssp_pull_div; 0 = automatic (16),
                             0xFFFFFFFF = push only */
  uint32_t sssp_local_hops; /* SSSP on low-degree graphs (one persistent
                             launch): at most this many passes per round in
                             which a workgroup relaxes the vertices it has
                             improved, before the grid barrier (what is
                             still owed after the last pass goes to the next
                             round); 0 = automatic (256), at most 256 */
  uint32_t pr_task_edges; /* binned PageRank layout: edges per accumulation
                             task; a bin with more edges is split over several
                             tasks (0 = automatic, 2^18) */
  uint32_t sssp_pull_layout; /* SSSP pull rounds: 0 automatic (= 2 where it
                             applies, else 1); 1 window reduction over the
                             in-edges; 2 binned push over PageRank's binned
                             layout plus a weight stream in its order
                             (HBM-resident weighted graphs that layout 4 of
                             pr_layout applies to, else 1).  Distances are
                             identical either way. */
  uint32_t reserved[2];
} grus_tuning;

/* ---- autotuning (no reference counterpart: the reference hard-codes its
 * thresholds, include/frontier.cuh:140) ----
 * Times the app on g under candidate settings (BFS: direction-switch alpha,
 * beta and the bottom-up lane/wave cut-off; SSSP: bucket width; PR:
 * edge layout and bin width) for the sample sources, keeps the fastest in the context's
 * tuning and returns it.  Results never depend on the tuning. */
#define GRUS_TUNE_BFS 0
#define GRUS_TUNE_SSSP 1
#define GRUS_TUNE_PR 2
int grus_autotune(grus_graph *g, int kind, const uint32_t *srcs, int n, grus_tuning *out);
/* Key for caching a tuning per graph (V, E, degree extremes, flags). */
int grus_graph_signature(grus_graph *g, uint64_t *sig);

/* ---- errors / version ---- */
const char *grus_last_error(void);
const char *grus_version(void);

/* ---- context ---- */
/* Replaces cudaSetDevice(FLAGS_device) + implicit default stream
 * (samples/bfs/bfs.cu:139, include/app.cuh:16). */
int grus_init(int device, grus_ctx **out);
int grus_fini(grus_ctx *ctx);
int grus_device_count(int *count);
int grus_get_tuning(grus_ctx *ctx, grus_tuning *out);
int grus_set_tuning(grus_ctx *ctx, const grus_tuning *in);
/* Restrict the context's work to a set of compute units: bit i of mask
 * (32 CUs per word) enables CU i, as hipExtStreamCreateWithCUMask numbers
 * them; mask_words = 0 restores every CU.  Replaces the reference's
 * co-location tool, which parks spinning kernels on the SMs to take them
 * away (tools/sm_limit.cu:5-25, used with MPS, tools/mps_start.sh:6-7).
 * Results never depend on the mask.  Servers created afterwards run on their
 * own streams and are not restricted. */
int grus_set_cu_mask(grus_ctx *ctx, const uint32_t *mask, uint32_t mask_words);
/* CUs the context may use and the device's total. */
int grus_cu_count(grus_ctx *ctx, int *active, int *total);
/* Cap the device memory the library may hold on the context's device
 * (0 = no cap): an allocation past it fails with GRUS_E_BUDGET, and under a
 * cap every app releases the workspaces other apps left cached on its graph.
 * Replaces the reference's memory-limit tool, a balloon cudaMalloc of all
 * but X GB (tools/mem_limit.cu:6-18).  Fails with GRUS_E_BUDGET when the
 * library already holds more. */
int grus_set_device_budget(grus_ctx *ctx, uint64_t bytes);
/* Device bytes the library holds on the context's device (every context
 * and graph of the process), and the peak since the last reset; reset != 0
 * restarts the peak from the bytes held now.  No reference counterpart. */
int grus_device_memory(grus_ctx *ctx, uint64_t *held, uint64_t *peak, int reset);

/* ---- host-side .gr I/O (no device needed) ----
 * Galois .gr v1: u64 header {1, sizeEdgeTy, N, M}, u64[N] end offsets,
 * u32[M] destinations, 4-byte pad if M odd and weighted, u32[M] weights.
 * Replaces graph_loader::ReadGraphGR (include/graph_loader.cuh:65-159).
 * grus_gr_read fills caller buffers: off[V+1] (off[0]=0), col[E], and
 * w[E] when w != NULL (all ones if the file has no weights). */
int grus_gr_header(const char *path, uint64_t *V, uint64_t *E, int *weighted);
int grus_gr_read(const char *path, uint64_t *off, uint32_t *col, uint32_t *w);
int grus_gr_write(const char *path, uint64_t V, uint64_t E, const uint64_t *off,
                  const uint32_t *col, const uint32_t *w /* NULL = unweighted */);

/* ---- graphs ----
 * symmetric != 0 declares that the CSR equals its transpose (then the CSC
 * used by pull kernels aliases the CSR); otherwise the CSC is built on the
 * device when a pull traversal first needs it, weights carried (the
 * reference's csr_tocsc drops them, include/graph.cuh:294). */
/* graph_t<CSR> from caller arrays (include/graph.cuh:254-268) */
int grus_graph_from_csr(grus_ctx *ctx, uint64_t V, uint64_t E, const uint64_t *off,
                        const uint32_t *col, const uint32_t *w, int symmetric,
                        const grus_place *place, grus_graph **out);
/* graph_loader::Load (include/graph_loader.cuh:160-166) + Set_Mem_Policy */
int grus_graph_load_gr(grus_ctx *ctx, const char *path, int need_weight, int symmetric,
                       const grus_place *place, grus_graph **out);
/* Synthetic inputs generated on the device (the reference has none):
 * RMAT(scale, edge_factor, a, b, c, seed) symmetrized, deduplicated,
 * self-loops dropped, rows sorted; weights (if wlo>0) uniform in [wlo,whi]
 * per undirected edge from wseed.  Grid: rows x cols 4-neighbour lattice,
 * id = r*cols + c. */
int grus_graph_gen_rmat(grus_ctx *ctx, int scale, int edge_factor, double a, double b,
                        double c, uint64_t seed, uint32_t wlo, uint32_t whi,
                        uint64_t wseed, const grus_place *place, grus_graph **out);
int grus_graph_gen_grid(grus_ctx *ctx, uint32_t rows, uint32_t cols, uint32_t wlo,
                        uint32_t whi, uint64_t wseed, const grus_place *place,
                        grus_graph **out);
typedef struct grus_graph_info {
  uint64_t V, E;
  int weighted, symmetric, placement;
  uint32_t max_degree_vertex; /* lowest id among max out-degree (loader :132-137) */
  uint64_t max_degree;
  uint64_t isolated;          /* vertices with out-degree 0 */
  uint64_t device_bytes;      /* HBM held by the graph */
  uint64_t host_bytes;        /* host-resident edge bytes (CAPPED/MANAGED) */
} grus_graph_info;
int grus_graph_info_get(grus_graph *g, grus_graph_info *out);
/* Device view of a graph's edges for user operators (include/grus_ops.hpp):
 * graph_t<CSR> / graph_t<CSC> as the reference's kernels receive them
 * (include/graph.cuh:113-268, passed by value into kernels).  Edges
 * [split, E) are in host memory (capped placement; split == E otherwise),
 * as graph_chunk::access_edge splits them (include/graph.cuh:70-75).  The
 * pointers stay valid until the graph is destroyed. */
typedef struct grus_csr_view {
  const uint64_t *off;     /* V + 1 row offsets */
  const uint32_t *col;     /* edges [0, split) */
  const uint32_t *w;       /* their weights, NULL when unweighted */
  const uint32_t *col_far; /* edges [split, E) */
  const uint32_t *w_far;
  uint64_t split, V, E;
} grus_csr_view;
/* in_edges = 0: out-edges (CSR); 1: in-edges (CSC, built on the device on
 * first use for a directed graph, as csr_tocsc does, include/graph.cuh:
 * 273-303; aliases the CSR for a symmetric one) */
int grus_graph_view(grus_graph *g, int in_edges, grus_csr_view *out);
/* The context's HIP stream (a hipStream_t), on which every library call of
 * the context runs: user kernels launched on it are ordered with them.
 * grus_set_cu_mask replaces the stream; query it again afterwards. */
int grus_ctx_stream(grus_ctx *ctx, void **hip_stream);
/* Copy the CSR back to host buffers (any pointer may be NULL). */
int grus_graph_copy_csr(grus_graph *g, uint64_t *off, uint32_t *col, uint32_t *w);
int grus_graph_destroy(grus_graph *g);

/* ---- apps (single source) ---- */
/* BFS levels.  Replaces BFS_single_gpu / BFS_pull_single_gpu
 * (samples/bfs/bfs.cu:104-173) with kernel<...Frontier<BDF>...> /
 * kernel_pull (include/kernel.cuh:124-165). */
int grus_bfs(grus_graph *g, uint32_t src, int dir_mode, uint32_t *levels_out,
             grus_stats *stats);
/* Two-stage top-down BFS, the reference's active-subgraph streaming
 * (BFS_2S_single_gpu, samples/bfs_2s/bfs_2s.cu:94-138, with Frontier_2S<BDF>,
 * Subgraph<NORMAL> and kernel_2s, include/frontier_2s.cuh:27-109,
 * include/subgraph.cuh:46-165, include/kernel_2s.cuh:36-92): each level's
 * frontier rows are gathered by a kernel into a device buffer of
 * stage_edges edges (0 = 2^26; within a device budget, what it leaves) and
 * expanded from there; a larger frontier goes in pieces, the gather of the
 * next overlapping the expansion of the current.  Meant for capped graphs,
 * whose host-resident rows then cross the link once per level in bulk.
 * stats->reserved[0] = staging capacity used (edges).  Same levels as
 * grus_bfs. */
int grus_bfs_2s(grus_graph *g, uint32_t src, uint64_t stage_edges, uint32_t *levels_out,
                grus_stats *stats);
/* SSSP distances (samples/sssp/sssp.cu:123-170; the relax is an atomic
 * min here, fixing the lost update at :54-55).  Requires weights (or unit
 * weights when the graph has none). */
int grus_sssp(grus_graph *g, uint32_t src, uint32_t *dist_out, grus_stats *stats);
/* PageRank pull, FP32, fixed iterations (samples/pagerank/pr.cu:134-168,
 * redefined: r0=(1-a)/V, r'=(1-a)/V + a*sum r(u)/outdeg(u)). */
int grus_pagerank(grus_graph *g, float alpha, int iters, float *rank_out,
                  grus_stats *stats);
/* Connected components: label = min vertex id of the component
 * (samples/cc/cc.cu:155-190).  method 0 = hooking + pointer jumping,
 * 1 = min-label propagation through the push operator (the reference's
 * algorithm). */
int grus_cc(grus_graph *g, int method, uint32_t *labels_out, grus_stats *stats);
/* ---- allocation-free variant (no counterpart in the reference) ----
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc.  Computes what
 * grus_cc computes, into the caller's device buffer labels_dev (V u32 on the
 * graph's device), and calls no hipMalloc, hipFree or allocating routine:
 * every other device buffer comes from a scratch area of
 * GRUS_NOALLOC_SCRATCH_BYTES that the library holds on each device until the
 * process ends (set up by grus_init).  Method 0 needs 4 KiB of it at any V;
 * method 1 needs 28 (V + 2) bytes plus scan temporaries (V up to about
 * 1.1 M); a larger graph fails with GRUS_E_BUDGET.  The call runs on the
 * context's stream (grus_ctx_stream) and returns when it is done; no other
 * stream may write labels_dev meanwhile (a fill on the null stream does not
 * order with it).  Calls that overlap on one device race on the scratch
 * area; the result is not held for grus_result_copy / grus_result_digest. */
#define GRUS_NOALLOC_SCRATCH_BYTES (32ull << 20)
int grus_cc_noalloc(grus_graph *g, int method, uint32_t *labels_dev, grus_stats *stats);
/* Copy the last app result held on the device for g (V elements of 4 B). */
int grus_result_copy(grus_graph *g, void *out);
/* Digest of the last u32 result held for g (BFS/SSSP/CC): reached = entries
 * below GRUS_INFINIT, checksum = sum over v of mix64(v << 32 | value)
 * mod 2^64 (mix64 = the splitmix64 output function); order independent, so
 * a served query's digest can be compared with a single-query run. */
int grus_result_digest(grus_graph *g, uint64_t *reached, uint64_t *checksum);

/* ---- batched multi-source (no reference counterpart: the north-star
 * batching pattern; each source gives exactly the single-source result) ----
 * BFS: k <= 64 sources packed as bits of one u64 per vertex.  SSSP: k <= 128
 * sources, two per bit (16-bit distances; 32-bit once a query needs it).
 * out is k*V u32 (row i = result of srcs[i]); NULL keeps results on the
 * device. */
int grus_bfs_multi(grus_graph *g, const uint32_t *srcs, int k, uint32_t *out,
                   grus_stats *stats);
int grus_sssp_multi(grus_graph *g, const uint32_t *srcs, int k, uint32_t *out,
                    grus_stats *stats);

/* ---- serving (no reference counterpart) ----
 * A server keeps the graph and per-query state on the device, admits
 * queries as lane slots and state memory free up, batches up to 64 queries
 * of one kind per traversal, and keeps admitting while batches run.  A
 * query is reported at the end of the step in which its lane finishes. */
#define GRUS_Q_BFS 0
#define GRUS_Q_SSSP 1
typedef struct grus_server_opts {
  int max_inflight_batches;   /* concurrent batches per query kind (engines,
                                 each with its own stream and worker thread),
                                 1..8, default 4: engines overlap each other's
                                 small steps and host round trips (more
                                 throughput at saturation; under moderate load
                                 fewer engines give lower tail latency) */
  int max_batch;              /* lanes per batch, 1..64, default 64; a BFS
                                 lane holds one query, an SSSP lane two */
  uint64_t state_budget_bytes; /* device bytes all engines' per-query state
                                 may hold together (0 = no cap).  Engines take
                                 their state when a batch needs one and it
                                 fits under this cap and under the device
                                 budget; until then waiting queries stay
                                 queued.  The first engine of each kind holds
                                 its state from creation (lanes shrink when
                                 one engine per kind at max_batch lanes would
                                 not fit), the others give theirs back after
                                 idle_release_ms without a query */
  int copy_results;           /* 1: copy each result to the host buffer given
                                 at submit; 0: keep a digest only */
  int admit_min_free;         /* past its admission window, an engine that
                                 still has running queries admits waiting ones
                                 only once at least this many of its lanes are
                                 free (an idle engine always admits).  1 =
                                 admit into any free lane at once (fully
                                 continuous); larger values keep a batch's
                                 queries in step, which shrinks the union of
                                 their frontiers.  0 = default: max_batch; at
                                 most max_batch */
  int admit_window;           /* a batch admits arrivals into its free lanes
                                 during its first admit_window steps after the
                                 engine started from idle, whatever
                                 admit_min_free says (0 = default 2; negative =
                                 no window) */
  int solo_max;               /* while at most solo_max queries of a kind
                                 wait, an idle engine's worker serves them one
                                 at a time with the single-source kernels.
                                 0 = default: while the engines of the kind,
                                 one query each at a time, would empty the
                                 queue within one batch's time (running
                                 means of both; a batch counted as 32
                                 single queries until one has run; at most
                                 8 waiting until a single query has run);
                                 negative = never */
  int idle_release_ms;        /* an engine other than the first of its kind
                                 that has run no query for this long gives
                                 its device state back.  0 = default: 1000
                                 when state_budget_bytes or a device budget
                                 caps memory, else never; negative = never */
} grus_server_opts;
/* Memory-based admission, as counted by the server (no reference
 * counterpart). */
typedef struct grus_server_stats {
  int engines_held[2];      /* engines holding device state, per kind */
  uint64_t state_bytes;     /* the device bytes of that state */
  uint64_t engines_started; /* engines that took state (including the first) */
  uint64_t engines_released;
  uint64_t memory_waits;    /* times a batch found no room for an engine */
} grus_server_stats;
typedef struct grus_query_result {
  uint64_t id;
  int status;            /* GRUS_OK or error */
  uint64_t reached;      /* vertices with a finite value */
  uint64_t checksum;     /* digest of the V-element result, as grus_result_digest */
  double submit_us, start_us, done_us; /* server clock */
} grus_query_result;
int grus_server_create(grus_graph *g, const grus_server_opts *opts, grus_server **out);
/* out_host: V u32 for the result when copy_results=1 (may be NULL). */
int grus_server_submit(grus_server *s, int kind, uint32_t src, uint32_t *out_host,
                       uint64_t *id_out);
/* Wait up to timeout_ms (<0 = forever) for one finished query. */
int grus_server_poll(grus_server *s, grus_query_result *res, int timeout_ms);
int grus_server_destroy(grus_server *s);
int grus_server_stats_get(grus_server *s, grus_server_stats *out);

#ifdef __cplusplus
}
#endif
#endif /* GRUS_C_H */
