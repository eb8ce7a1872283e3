// graph.hpp -- context, device buffers and the CSR/CSC graph container.
//
// Mirrors graph_base / graph_t<CSR> / graph_t<CSC> (reference
// include/graph.cuh:113-321) and graph_loader (include/graph_loader.cuh).
// Differences by design: offsets are u64; the graph owns and frees its
// memory; placement is an explicit mode (HBM / managed / capped) instead of
// Set_Mem_Policy's flags over managed memory (include/graph.cuh:170-251);
// for symmetric graphs the CSC aliases the CSR instead of a host transpose.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "common.hpp"

namespace mgg {

enum class Mem { None, Device, Managed, HostPinned };

// Library-wide accounting of device memory (hipMalloc'd Buf memory, per
// device): bytes held now, the peak since the last reset, and an optional
// budget.  With a budget set, an allocation that would take the bytes held
// above it fails with GRUS_E_BUDGET instead of reaching the device -- the
// whole-device cap of the reference's memory-limit tool (tools/mem_limit.cu:
// 6-18), counted per allocation instead of by a balloon.  Managed and pinned
// host memory are not counted.
uint64_t devmem_live(int dev);
uint64_t devmem_peak(int dev);
void devmem_reset_peak(int dev);
uint64_t devmem_budget(int dev);
void devmem_set_budget(int dev, uint64_t bytes);  // 0 = no budget

// Owning untyped buffer.
struct Buf {
  void *p = nullptr;
  size_t bytes = 0;
  Mem kind = Mem::None;
  int dev = -1;  // device of a Mem::Device allocation (for the accounting)
  Buf() = default;
  Buf(const Buf &) = delete;
  Buf &operator=(const Buf &) = delete;
  Buf(Buf &&o) noexcept { *this = std::move(o); }
  Buf &operator=(Buf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      kind = o.kind;
      dev = o.dev;
      o.p = nullptr;
      o.bytes = 0;
      o.kind = Mem::None;
      o.dev = -1;
    }
    return *this;
  }
  ~Buf() { release(); }
  void alloc(size_t nbytes, Mem k = Mem::Device);
  // grow-only (contents not preserved)
  void ensure(size_t nbytes, Mem k = Mem::Device) {
    if (nbytes > bytes || kind != k) alloc(nbytes, k);
  }
  void release();
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

// scan buffers of the edge-balanced push operator (advance.hpp)
struct PushWork {
  Buf deg, scan;
  // borrowed buffers of a caller that cannot allocate (used when set), and
  // the scan's temporary storage (null: the context's, grown on demand)
  uint64_t *deg_ext = nullptr, *scan_ext = nullptr;
  void *tmp = nullptr;
  size_t tmp_bytes = 0;
  void ensure(uint64_t V) {
    deg.ensure((V + 2) * 8);
    scan.ensure((V + 2) * 8);
  }
  uint64_t *deg_ptr() const { return deg_ext ? deg_ext : deg.as<uint64_t>(); }
  uint64_t *scan_ptr() const { return scan_ext ? scan_ext : scan.as<uint64_t>(); }
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int cu_count = 256;  // CUs the context's stream may use (all, or its CU mask)
  int cu_total = 256;  // CUs of the device
  bool cu_masked = false;
  // other contexts run kernels on the device at the same time (server
  // workers): no cooperative persistent launches, whose grid must be resident
  bool shared_device = false;
  int wall_khz = 100000;  // device wall clock rate (wall_clock64 ticks per ms)
  grus_tuning tuning{};
  uint64_t *h_ctl = nullptr;  // pinned host mirror of the control block
  Buf d_ctl;                  // device control block (counters)
  Buf scan_tmp;               // rocPRIM temporary storage, grown on demand
  explicit Context(int dev);
  ~Context();
  void sync() { GRUS_HIP(hipStreamSynchronize(stream)); }
  void activate() const { GRUS_HIP(hipSetDevice(device)); }
  // copy the first n u64 of the device control block to h_ctl and wait
  void read_ctl(size_t n);
  // one kernel copies the first n u64 of the control block to h_ctl and
  // zeroes them on the device; then wait.  Replaces a copy and the next
  // step's memset (two launches) in per-level read-backs.
  void publish_ctl(size_t n);
  // replace the stream by one restricted to the CUs set in mask (words == 0:
  // all CUs); waits for the old stream first
  void set_cu_mask(const uint32_t *mask, uint32_t words);
  uint64_t wall_ticks_per_ms() const { return (uint64_t)(wall_khz > 0 ? wall_khz : 100000); }
};

// POD view of a CSR (or CSC) handed to kernels.  Edges [0, split) are read
// from col/w; edges [split, E) from col_far/w_far (host-resident remainder
// in capped placement; split == E otherwise).
struct CsrView {
  const uint64_t *off;
  const uint32_t *col;
  const uint32_t *w;
  const uint32_t *col_far;
  const uint32_t *w_far;
  uint64_t split;
  uint64_t V, E;
};

struct BfsWork;
struct SsspWork;
struct PrWork;
struct CcWork;
class MsEngine;

struct Graph {
  Context *ctx = nullptr;
  uint64_t V = 0, E = 0;
  bool weighted = false;
  bool symmetric = false;
  int placement = GRUS_PLACE_HBM;
  Buf off, col, w;           // CSR (device, managed, or HBM prefix in capped mode)
  Buf col_far, w_far;        // capped mode: host-resident edge remainder
  Buf first_col;             // capped mode: first neighbour of every row (GRUS_NO_VERTEX
                             // if empty), in HBM, so a bottom-up first test never
                             // crosses the host link
  uint64_t split = 0;        // first edge held in col_far
  Buf in_off, in_col, in_w;  // CSC of a directed graph (built on first pull use)
  bool csc_ready = false;
  // derived at creation
  uint32_t max_deg_vertex = 0;
  uint64_t max_deg = 0;
  uint64_t isolated = 0;
  Buf unreachable_bm;  // bit v set: in-degree 0 (never discovered by a traversal)
  bool unreachable_ready = false;
  uint64_t unreachable_count = 0;  // vertices with in-degree 0 (set by unreachable())
  uint64_t host_bytes = 0;
  // workspaces and last result
  Buf result;
  std::shared_ptr<BfsWork> bfs;
  std::shared_ptr<SsspWork> sssp;
  std::shared_ptr<PrWork> pr;
  std::shared_ptr<CcWork> cc;
  std::shared_ptr<MsEngine> ms;

  explicit Graph(Context *c);
  ~Graph();
  // Under a device budget, an app first releases the workspaces other apps
  // left cached on this graph, so that the budget has to hold the largest
  // app's workspace, not their sum (workspace_reserve sizes that).
  enum class App { BFS, SSSP, PR, CC, MULTI };
  void make_room(App keep);
  CsrView csr() const;
  CsrView csc();  // in-edges; aliases csr() when symmetric
  uint64_t words() const { return (V + 63) / 64; }
  void compute_derived();         // max degree vertex, isolated count
  const uint64_t *unreachable();  // in-degree-0 bitmap (builds CSC if directed)
  uint64_t device_bytes() const;
};

// graph construction (graph.hip)
std::unique_ptr<Graph> graph_from_host(Context *ctx, uint64_t V, uint64_t E, const uint64_t *off,
                                       const uint32_t *col, const uint32_t *w, bool symmetric,
                                       const grus_place *place);
std::unique_ptr<Graph> graph_gen_rmat(Context *ctx, int scale, int ef, double a, double b,
                                      double c, uint64_t seed, uint32_t wlo, uint32_t whi,
                                      uint64_t wseed, const grus_place *place);
std::unique_ptr<Graph> graph_gen_grid(Context *ctx, uint32_t rows, uint32_t cols, uint32_t wlo,
                                      uint32_t whi, uint64_t wseed, const grus_place *place);
void graph_copy_csr(Graph &g, uint64_t *off, uint32_t *col, uint32_t *w);
// CSR (off[V+1], col[m]) from m sorted (row << 32 | col) keys (asynchronous)
void csr_from_sorted_keys(Context &ctx, const uint64_t *keys, uint64_t m, uint64_t V,
                          uint64_t *off, uint32_t *col);
// (re)place a device-built CSR per the placement request
void graph_apply_placement(Graph &g, const grus_place *place);
// A graph that borrows g's arrays (CSR, CSC, first neighbours, in-degree-0
// bitmap, all built on g first) and runs on ctx with workspaces of its own:
// single-source apps of several threads on one graph, one view each.  g must
// outlive the view.
std::unique_ptr<Graph> graph_view(Graph &g, Context *ctx);
// Upper bound of the device bytes one app (BFS, SSSP, PageRank pull, CC) on
// a capped graph of V vertices and E edges allocates beside the graph: its
// workspace, the result array, control blocks and scan temporaries.
uint64_t workspace_reserve(uint64_t V, uint64_t E);

// .gr I/O (gr_io.cpp; host only)
struct GrHeader {
  uint64_t V, E, size_edge_ty;
};
GrHeader gr_read_header(const std::string &path);
void gr_read(const std::string &path, uint64_t *off, uint32_t *col, uint32_t *w);
void gr_write(const std::string &path, uint64_t V, uint64_t E, const uint64_t *off,
              const uint32_t *col, const uint32_t *w);

}  // namespace mgg
