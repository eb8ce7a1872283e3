// graph.hip -- device buffers, context, graph construction on the device.
//
// Reference counterparts: graph_loader::Load (include/graph_loader.cuh:160)
// and graph_base (include/graph.cuh:113-252).  Generation, deduplication,
// CSR assembly and the CSC transpose all run on the GPU (radix sort +
// scan), so RMAT scale 24 is built in well under a second instead of the
// reference's host-side serial loops (csr_tocsc, include/graph.cuh:273-303).
#include <algorithm>
#include <atomic>
#include <cstring>

#include "device.hpp"
#include "gen.hpp"
#include "graph.hpp"
#include "prims.hpp"

namespace mgg {

// ------------------------------------------------------------------ accounting
namespace {
constexpr int MAX_DEV = 64;
std::atomic<uint64_t> g_live[MAX_DEV], g_peak[MAX_DEV], g_budget[MAX_DEV];
int dev_slot(int dev) {
  if (dev < 0 || dev >= MAX_DEV) raise(GRUS_E_NODEVICE, "device index out of range");
  return dev;
}
}  // namespace

uint64_t devmem_live(int dev) { return g_live[dev_slot(dev)].load(); }
uint64_t devmem_peak(int dev) { return g_peak[dev_slot(dev)].load(); }
void devmem_reset_peak(int dev) { g_peak[dev_slot(dev)].store(g_live[dev].load()); }
uint64_t devmem_budget(int dev) { return g_budget[dev_slot(dev)].load(); }
void devmem_set_budget(int dev, uint64_t bytes) {
  const int d = dev_slot(dev);
  if (bytes && g_live[d].load() > bytes)
    raise(GRUS_E_BUDGET, "device budget below the bytes the library already holds (" +
                             std::to_string(g_live[d].load()) + ")");
  g_budget[d].store(bytes);
}

uint64_t workspace_reserve(uint64_t V, uint64_t E) {
  // per vertex, the largest app: SSSP's two near and two far queues and two
  // stamp arrays (24 B), the push operator's degree scan (16 B), the result
  // (4 B) and bitmaps; BFS needs 42 B, PageRank's pull 12 B, CC 30 B.  Per
  // 2048-edge window, the pull rounds' window tables (12 B).  Plus control
  // blocks, rocPRIM temporaries and the in-degree-0 bitmap.
  return 48 * (V + 2) + 16 * (E / 2048 + 2) + (64ull << 20);
}

// ------------------------------------------------------------------ Buf
void Buf::release() {
  if (!p) return;
  switch (kind) {
    case Mem::Device:
      (void)hipFree(p);
      if (dev >= 0 && dev < MAX_DEV) g_live[dev].fetch_sub(bytes);
      break;
    case Mem::Managed:
      (void)hipFree(p);
      break;
    case Mem::HostPinned:
      (void)hipHostFree(p);
      break;
    default:
      break;
  }
  p = nullptr;
  bytes = 0;
  kind = Mem::None;
}

void Buf::alloc(size_t nbytes, Mem k) {
  release();
  if (nbytes == 0) nbytes = 8;
  switch (k) {
    case Mem::Device: {
      int d = 0;
      GRUS_HIP(hipGetDevice(&d));
      d = dev_slot(d);
      const uint64_t held = g_live[d].fetch_add(nbytes) + nbytes;
      const uint64_t budget = g_budget[d].load();
      if (budget && held > budget) {
        g_live[d].fetch_sub(nbytes);
        raise(GRUS_E_BUDGET, "device budget of " + std::to_string(budget) +
                                 " bytes exceeded: " + std::to_string(held - nbytes) +
                                 " held, " + std::to_string(nbytes) + " requested");
      }
      const hipError_t e = hipMalloc(&p, nbytes);
      if (e != hipSuccess) {
        g_live[d].fetch_sub(nbytes);
        p = nullptr;
        GRUS_HIP(e);
      }
      uint64_t pk = g_peak[d].load();
      while (held > pk && !g_peak[d].compare_exchange_weak(pk, held)) {
      }
      dev = d;
      break;
    }
    case Mem::Managed:
      GRUS_HIP(hipMallocManaged(&p, nbytes, hipMemAttachGlobal));
      break;
    case Mem::HostPinned:
      GRUS_HIP(hipHostMalloc(&p, nbytes, hipHostMallocMapped));
      break;
    default:
      raise(GRUS_E_INVALID, "bad memory kind");
  }
  bytes = nbytes;
  kind = k;
}

// ------------------------------------------------------------------ Context
Context::Context(int dev) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    (void)hipGetLastError();
    raise(GRUS_E_NODEVICE, "no HIP device available");
  }
  if (dev < 0 || dev >= n) raise(GRUS_E_NODEVICE, "device index out of range");
  device = dev;
  activate();
  hipDeviceProp_t prop;
  GRUS_HIP(hipGetDeviceProperties(&prop, dev));
  cu_count = cu_total = prop.multiProcessorCount;
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
    wall_khz = khz;
  GRUS_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  GRUS_HIP(hipEventCreate(&ev0));
  GRUS_HIP(hipEventCreate(&ev1));
  GRUS_HIP(hipHostMalloc((void **)&h_ctl, 4096, hipHostMallocDefault));
  d_ctl.alloc(4096);
  GRUS_HIP(hipMemsetAsync(d_ctl.p, 0, 4096, stream));
  GRUS_HIP(hipStreamSynchronize(stream));
  tuning.bfs_alpha = 14.0;
  tuning.bfs_beta = 64.0;
  tuning.bu_thread_edges = 32;
  tuning.sssp_delta = 0;
}

Context::~Context() {
  (void)hipSetDevice(device);
  if (stream) (void)hipStreamSynchronize(stream);
  d_ctl.release();
  scan_tmp.release();
  if (h_ctl) (void)hipHostFree(h_ctl);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
}

void Context::set_cu_mask(const uint32_t *mask, uint32_t words) {
  activate();
  int n = 0;
  for (uint32_t i = 0; i < words; i++)
    for (int b = 0; b < 32; b++)
      if ((mask[i] >> b) & 1u && (int)(i * 32 + b) < cu_total) n++;
  if (words && n == 0) raise(GRUS_E_INVALID, "CU mask selects no compute unit of the device");
  hipStream_t s = nullptr;
  if (words) {
    GRUS_HIP(hipExtStreamCreateWithCUMask(&s, words, mask));
  } else {
    GRUS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  }
  sync();
  GRUS_HIP(hipStreamDestroy(stream));
  stream = s;
  cu_masked = words != 0;
  cu_count = words ? n : cu_total;
}

namespace {
__global__ void k_publish_ctl(uint64_t *d, uint64_t *h, uint32_t n) {
  const uint32_t i = threadIdx.x;
  if (i < n) {
    h[i] = d[i];
    d[i] = 0;
  }
}
}  // namespace

void Context::publish_ctl(size_t n) {
  if (n > 64) raise(GRUS_E_INVALID, "publish_ctl: at most 64 words");
  k_publish_ctl<<<1, 64, 0, stream>>>(d_ctl.as<uint64_t>(), h_ctl, (uint32_t)n);
  GRUS_LAUNCH_CHECK();
  GRUS_HIP(hipStreamSynchronize(stream));
}

void Context::read_ctl(size_t n) {
  GRUS_HIP(hipMemcpyAsync(h_ctl, d_ctl.p, n * sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
  GRUS_HIP(hipStreamSynchronize(stream));
}

// ------------------------------------------------------------------ Graph
Graph::Graph(Context *c) : ctx(c) {}

void Graph::make_room(App keep) {
  if (!devmem_budget(ctx->device)) return;  // unbudgeted: keep every cache
  GRUS_HIP(hipStreamSynchronize(ctx->stream));
  if (keep != App::BFS) bfs.reset();
  if (keep != App::SSSP) sssp.reset();
  // SSSP's binned pull rounds run on PageRank's layout (HBM graphs): kept
  if (keep != App::PR && !(keep == App::SSSP && placement == GRUS_PLACE_HBM)) pr.reset();
  if (keep != App::CC) cc.reset();
  if (keep != App::MULTI) ms.reset();
}

CsrView Graph::csr() const {
  CsrView v;
  v.off = off.as<const uint64_t>();
  v.col = col.as<const uint32_t>();
  v.w = weighted ? w.as<const uint32_t>() : nullptr;
  v.col_far = col_far.as<const uint32_t>();
  v.w_far = weighted ? w_far.as<const uint32_t>() : nullptr;
  v.split = (placement == GRUS_PLACE_CAPPED) ? split : E;
  v.V = V;
  v.E = E;
  return v;
}

uint64_t Graph::device_bytes() const {
  uint64_t b = 0;
  for (const Buf *x :
       {&off, &col, &w, &in_off, &in_col, &in_w, &unreachable_bm, &result, &first_col})
    if (x->kind == Mem::Device) b += x->bytes;
  return b;
}

namespace {

__global__ void k_first_col(const uint64_t *off, const uint32_t *col, uint64_t V, uint32_t *out) {
  uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; v < V; v += stride) out[v] = off[v] < off[v + 1] ? col[off[v]] : GRUS_NO_VERTEX;
}

// ---- derived quantities ----
__global__ void k_degree_stats(const uint64_t *off, uint64_t V, unsigned long long *best,
                               unsigned long long *isolated) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long mine = 0, iso = 0;
  for (; i < V; i += stride) {
    uint64_t d = off[i + 1] - off[i];
    // highest degree first, then lowest id (graph_loader.cuh:132-137 uses
    // std::max_element, i.e. the first maximum)
    unsigned long long key = (d << 32) | (0xFFFFFFFFull - i);
    mine = key > mine ? key : mine;
    iso += (d == 0);
  }
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long other = __shfl_xor(mine, o);
    mine = other > mine ? other : mine;
  }
  iso = wave_sum(iso);
  if (lane_id() == 0) {
    atomicMax(best, mine);
    atomicAdd(isolated, iso);
  }
}

// bit v of bm set when row v of (CSC) offsets is empty
__global__ void k_empty_rows_bitmap(const uint64_t *off, uint64_t V, uint64_t *bm, uint64_t W) {
  uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE;
  uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / WAVE;
  for (uint64_t wd = wave; wd < W; wd += nw) {
    uint64_t v = wd * 64 + lane_id();
    bool empty = (v >= V) || (off[v + 1] == off[v]);
    uint64_t m = __ballot(empty);
    if (lane_id() == 0) bm[wd] = m;
  }
}

// set bits of bm among the first V
__global__ void k_count_bits(const uint64_t *bm, uint64_t W, uint64_t V, unsigned long long *n) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long c = 0;
  for (; i < W; i += stride) {
    uint64_t m = bm[i];
    if ((i + 1) * 64 > V) m &= (V % 64) ? ((1ull << (V % 64)) - 1ull) : ~0ull;
    c += __popcll(m);
  }
  c = wave_sum(c);
  if (lane_id() == 0 && c) atomicAdd(n, c);
}

// ---- RMAT ----
__global__ void k_rmat_keys(RmatSpec s, uint64_t M, uint64_t *keys) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < M; i += stride) {
    uint32_t u, v;
    rmat_sample(s, i, u, v);
    uint64_t k0 = ~0ull, k1 = ~0ull;  // self-loops become sentinels
    if (u != v) {
      k0 = ((uint64_t)u << 32) | v;
      k1 = ((uint64_t)v << 32) | u;
    }
    keys[2 * i] = k0;
    keys[2 * i + 1] = k1;
  }
}

// compaction of sorted keys; pos = exclusive scan of unique flags
__global__ void k_unique_scatter(const uint64_t *keys, const uint64_t *pos, uint64_t n,
                                 uint64_t *out) {
  uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; j < n; j += stride) {
    uint64_t x = keys[j];
    if (x != ~0ull && (j == 0 || keys[j - 1] != x)) out[pos[j]] = x;
  }
}

// off[r] = first j with row(j) >= r, for sorted (row<<32 | col) keys[0..m)
__global__ void k_offsets_from_sorted(const uint64_t *keys, uint64_t m, uint64_t V,
                                      uint64_t *off) {
  uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; j <= m; j += stride) {
    int64_t prev = (j == 0) ? -1 : (int64_t)(keys[j - 1] >> 32);
    int64_t cur = (j == m) ? (int64_t)V : (int64_t)(keys[j] >> 32);
    for (int64_t r = prev + 1; r <= cur; r++) off[r] = j;
  }
}

__global__ void k_split_keys(const uint64_t *keys, uint64_t m, uint32_t *col, uint32_t *w,
                             uint64_t wkey, uint32_t wlo, uint32_t whi) {
  uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; j < m; j += stride) {
    uint64_t k = keys[j];
    uint32_t u = (uint32_t)(k >> 32), v = (uint32_t)k;
    col[j] = v;
    if (w) w[j] = edge_weight(u, v, wkey, wlo, whi);
  }
}

// ---- grid ----
__global__ void k_grid_degree(uint32_t R, uint32_t C, uint64_t *deg) {
  uint64_t n = (uint64_t)R * C;
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    uint32_t r = (uint32_t)(i / C), c = (uint32_t)(i % C);
    deg[i] = (r > 0) + (c > 0) + (c + 1 < C) + (r + 1 < R);
  }
}

__global__ void k_grid_fill(uint32_t R, uint32_t C, const uint64_t *off, uint32_t *col,
                            uint32_t *w, uint64_t wkey, uint32_t wlo, uint32_t whi) {
  uint64_t n = (uint64_t)R * C;
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    uint32_t r = (uint32_t)(i / C), c = (uint32_t)(i % C);
    uint64_t e = off[i];
    uint64_t nb[4];
    int k = 0;
    if (r > 0) nb[k++] = i - C;
    if (c > 0) nb[k++] = i - 1;
    if (c + 1 < C) nb[k++] = i + 1;
    if (r + 1 < R) nb[k++] = i + C;
    for (int j = 0; j < k; j++) {
      col[e + j] = (uint32_t)nb[j];
      if (w) w[e + j] = edge_weight((uint32_t)i, (uint32_t)nb[j], wkey, wlo, whi);
    }
  }
}

// ---- transpose: key = (dst << 32) | src for every CSR edge ----
__global__ void k_transpose_keys(const uint64_t *off, const uint32_t *col, uint64_t V,
                                 uint64_t *keys) {
  uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE;
  uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / WAVE;
  for (uint64_t u = wave; u < V; u += nw)
    for (uint64_t e = off[u] + lane_id(); e < off[u + 1]; e += WAVE)
      keys[e] = ((uint64_t)col[e] << 32) | u;
}

__global__ void k_iota_u32(uint32_t *a, uint64_t n) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) a[i] = (uint32_t)i;
}

__global__ void k_gather_u32(const uint32_t *src, const uint32_t *idx, uint64_t n, uint32_t *dst) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[idx[i]];
}

__global__ void k_low32(const uint64_t *keys, uint64_t n, uint32_t *out) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = (uint32_t)keys[i];
}

unsigned grid_for(Context *ctx, uint64_t n, unsigned block = BLOCK) {
  uint64_t g = div_up(n ? n : 1, block);
  uint64_t cap = (uint64_t)ctx->cu_count * 32;
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

}  // namespace

void Graph::compute_derived() {
  auto *best = reinterpret_cast<unsigned long long *>(ctx->d_ctl.p);
  GRUS_HIP(hipMemsetAsync(best, 0, 16, ctx->stream));
  if (V) {
    k_degree_stats<<<grid_for(ctx, V), BLOCK, 0, ctx->stream>>>(off.as<uint64_t>(), V, best,
                                                                best + 1);
    GRUS_LAUNCH_CHECK();
  }
  ctx->read_ctl(2);
  uint64_t key = ctx->h_ctl[0];
  max_deg = key >> 32;
  max_deg_vertex = V ? (uint32_t)(0xFFFFFFFFull - (key & 0xFFFFFFFFull)) : 0;
  isolated = ctx->h_ctl[1];
}

CsrView Graph::csc() {
  if (symmetric) return csr();
  if (!csc_ready) {
    if (placement == GRUS_PLACE_CAPPED)
      raise(GRUS_E_UNSUPPORTED, "pull traversal of a directed graph needs HBM placement");
    // keys (dst<<32|src) sorted with the CSR edge index as value, so that
    // weights can be permuted along (the reference drops them,
    // include/graph.cuh:294)
    Buf ka, kb, va, vb;
    ka.alloc(E * 8);
    kb.alloc(E * 8);
    va.alloc(E * 4);
    vb.alloc(E * 4);
    k_transpose_keys<<<grid_for(ctx, V * WAVE), BLOCK, 0, ctx->stream>>>(
        off.as<uint64_t>(), col.as<uint32_t>(), V, ka.as<uint64_t>());
    GRUS_LAUNCH_CHECK();
    k_iota_u32<<<grid_for(ctx, E), BLOCK, 0, ctx->stream>>>(va.as<uint32_t>(), E);
    GRUS_LAUNCH_CHECK();
    uint32_t *perm = nullptr;
    uint64_t *sorted = sort_pairs_u64_u32(*ctx, ka.as<uint64_t>(), kb.as<uint64_t>(),
                                          va.as<uint32_t>(), vb.as<uint32_t>(), E, &perm);
    in_off.alloc((V + 1) * 8);
    in_col.alloc(E * 4);
    k_offsets_from_sorted<<<grid_for(ctx, E + 1), BLOCK, 0, ctx->stream>>>(
        sorted, E, V, in_off.as<uint64_t>());
    GRUS_LAUNCH_CHECK();
    k_low32<<<grid_for(ctx, E), BLOCK, 0, ctx->stream>>>(sorted, E, in_col.as<uint32_t>());
    GRUS_LAUNCH_CHECK();
    if (weighted) {
      in_w.alloc(E * 4);
      k_gather_u32<<<grid_for(ctx, E), BLOCK, 0, ctx->stream>>>(w.as<uint32_t>(), perm, E,
                                                               in_w.as<uint32_t>());
      GRUS_LAUNCH_CHECK();
    }
    ctx->sync();
    csc_ready = true;
  }
  CsrView v;
  v.off = in_off.as<const uint64_t>();
  v.col = in_col.as<const uint32_t>();
  v.w = weighted ? in_w.as<const uint32_t>() : nullptr;
  v.col_far = nullptr;
  v.w_far = nullptr;
  v.split = E;
  v.V = V;
  v.E = E;
  return v;
}

const uint64_t *Graph::unreachable() {
  if (!unreachable_ready) {
    CsrView in = csc();
    uint64_t W = words();
    unreachable_bm.alloc((W + 1) * 8);
    k_empty_rows_bitmap<<<grid_for(ctx, W * WAVE), BLOCK, 0, ctx->stream>>>(
        in.off, V, unreachable_bm.as<uint64_t>(), W);
    GRUS_LAUNCH_CHECK();
    Buf n;
    n.alloc(8);
    GRUS_HIP(hipMemsetAsync(n.p, 0, 8, ctx->stream));
    k_count_bits<<<grid_for(ctx, W), BLOCK, 0, ctx->stream>>>(unreachable_bm.as<uint64_t>(), W,
                                                             V, n.as<unsigned long long>());
    GRUS_LAUNCH_CHECK();
    GRUS_HIP(hipMemcpyAsync(&unreachable_count, n.p, 8, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    unreachable_ready = true;
  }
  return unreachable_bm.as<const uint64_t>();
}

Graph::~Graph() {
  if (ctx) (void)hipSetDevice(ctx->device);
}

std::unique_ptr<Graph> graph_view(Graph &g, Context *ctx) {
  g.csc();          // directed graphs: the in-edges, built once on g
  g.unreachable();  // the in-degree-0 bitmap
  auto v = std::make_unique<Graph>(ctx);
  v->V = g.V;
  v->E = g.E;
  v->weighted = g.weighted;
  v->symmetric = g.symmetric;
  v->placement = g.placement;
  v->split = g.split;
  v->max_deg_vertex = g.max_deg_vertex;
  v->max_deg = g.max_deg;
  v->isolated = g.isolated;
  v->host_bytes = g.host_bytes;
  // borrowed buffers: kind None, so the view never frees them
  auto borrow = [](Buf &dst, const Buf &src) {
    dst.p = src.p;
    dst.bytes = src.bytes;
    dst.kind = Mem::None;
  };
  borrow(v->off, g.off);
  borrow(v->col, g.col);
  borrow(v->w, g.w);
  borrow(v->col_far, g.col_far);
  borrow(v->w_far, g.w_far);
  borrow(v->first_col, g.first_col);
  borrow(v->in_off, g.in_off);
  borrow(v->in_col, g.in_col);
  borrow(v->in_w, g.in_w);
  v->csc_ready = g.csc_ready;
  borrow(v->unreachable_bm, g.unreachable_bm);
  v->unreachable_ready = true;
  v->unreachable_count = g.unreachable_count;
  return v;
}

// ------------------------------------------------------------------ builders
std::unique_ptr<Graph> graph_from_host(Context *ctx, uint64_t V, uint64_t E, const uint64_t *off,
                                       const uint32_t *col, const uint32_t *w, bool symmetric,
                                       const grus_place *place) {
  if (!off || (E && !col)) raise(GRUS_E_INVALID, "null CSR array");
  if (V > 0xFFFFFFFFull) raise(GRUS_E_INVALID, "vertex ids are 32-bit");
  if (off[0] != 0 || off[V] != E) raise(GRUS_E_INVALID, "offsets must run from 0 to E");
  ctx->activate();
  auto g = std::make_unique<Graph>(ctx);
  g->V = V;
  g->E = E;
  g->weighted = (w != nullptr);
  g->symmetric = symmetric;
  g->off.alloc((V + 1) * 8);
  g->col.alloc(E * 4);
  GRUS_HIP(hipMemcpy(g->off.p, off, (V + 1) * 8, hipMemcpyHostToDevice));
  if (E) GRUS_HIP(hipMemcpy(g->col.p, col, E * 4, hipMemcpyHostToDevice));
  if (w) {
    g->w.alloc(E * 4);
    if (E) GRUS_HIP(hipMemcpy(g->w.p, w, E * 4, hipMemcpyHostToDevice));
  }
  g->compute_derived();
  graph_apply_placement(*g, place);
  return g;
}

std::unique_ptr<Graph> graph_gen_rmat(Context *ctx, int scale, int ef, double a, double b,
                                      double c, uint64_t seed, uint32_t wlo, uint32_t whi,
                                      uint64_t wseed, const grus_place *place) {
  if (scale < 1 || scale > 31 || ef < 1 || ef > 1024)
    raise(GRUS_E_INVALID, "rmat scale must be in [1,31] and edge factor in [1,1024]");
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) raise(GRUS_E_INVALID, "bad rmat a/b/c");
  if (wlo > whi) raise(GRUS_E_INVALID, "weight range");
  ctx->activate();
  uint64_t V = 1ull << scale;
  uint64_t M = (uint64_t)ef << scale;
  uint64_t n = 2 * M;
  Buf ka, kb, pos;
  ka.alloc(n * 8);
  kb.alloc(n * 8);
  RmatSpec spec = rmat_spec(scale, a, b, c, seed);
  k_rmat_keys<<<grid_for(ctx, M), BLOCK, 0, ctx->stream>>>(spec, M, ka.as<uint64_t>());
  GRUS_LAUNCH_CHECK();
  uint64_t *sorted = sort_keys_u64(*ctx, ka.as<uint64_t>(), kb.as<uint64_t>(), n);
  uint64_t *other = (sorted == ka.as<uint64_t>()) ? kb.as<uint64_t>() : ka.as<uint64_t>();
  pos.alloc(n * 8);
  scan_unique_flags(*ctx, sorted, pos.as<uint64_t>(), n, ~0ull);
  // count = pos[n-1] + flag(n-1)
  uint64_t last_pos = 0, last2[2] = {0, 0};
  GRUS_HIP(hipMemcpyAsync(&last_pos, pos.as<uint64_t>() + n - 1, 8, hipMemcpyDeviceToHost,
                          ctx->stream));
  GRUS_HIP(hipMemcpyAsync(last2, sorted + n - 2, 16, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  uint64_t m = last_pos + ((last2[1] != ~0ull && last2[1] != last2[0]) ? 1 : 0);
  k_unique_scatter<<<grid_for(ctx, n), BLOCK, 0, ctx->stream>>>(sorted, pos.as<uint64_t>(), n,
                                                               other);
  GRUS_LAUNCH_CHECK();
  pos.release();
  auto g = std::make_unique<Graph>(ctx);
  g->V = V;
  g->E = m;
  g->symmetric = true;
  g->weighted = (wlo > 0);
  g->off.alloc((V + 1) * 8);
  g->col.alloc(m * 4);
  if (g->weighted) g->w.alloc(m * 4);
  k_offsets_from_sorted<<<grid_for(ctx, m + 1), BLOCK, 0, ctx->stream>>>(other, m, V,
                                                                         g->off.as<uint64_t>());
  GRUS_LAUNCH_CHECK();
  k_split_keys<<<grid_for(ctx, m), BLOCK, 0, ctx->stream>>>(
      other, m, g->col.as<uint32_t>(), g->weighted ? g->w.as<uint32_t>() : nullptr,
      weight_key(wseed), wlo, whi);
  GRUS_LAUNCH_CHECK();
  ctx->sync();
  ka.release();
  kb.release();
  g->compute_derived();
  graph_apply_placement(*g, place);
  return g;
}

std::unique_ptr<Graph> graph_gen_grid(Context *ctx, uint32_t R, uint32_t C, uint32_t wlo,
                                      uint32_t whi, uint64_t wseed, const grus_place *place) {
  if (R == 0 || C == 0 || (uint64_t)R * C > 0xFFFFFFFEull) raise(GRUS_E_INVALID, "grid size");
  if (wlo > whi) raise(GRUS_E_INVALID, "weight range");
  ctx->activate();
  uint64_t V = (uint64_t)R * C;
  uint64_t E = 2ull * ((uint64_t)R * (C - 1) + (uint64_t)C * (R - 1));
  auto g = std::make_unique<Graph>(ctx);
  g->V = V;
  g->E = E;
  g->symmetric = true;
  g->weighted = (wlo > 0);
  Buf deg;
  deg.alloc((V + 1) * 8);
  GRUS_HIP(hipMemsetAsync(deg.p, 0, (V + 1) * 8, ctx->stream));
  k_grid_degree<<<grid_for(ctx, V), BLOCK, 0, ctx->stream>>>(R, C, deg.as<uint64_t>());
  GRUS_LAUNCH_CHECK();
  g->off.alloc((V + 1) * 8);
  scan_excl_u64(*ctx, deg.as<uint64_t>(), g->off.as<uint64_t>(), V + 1);
  g->col.alloc(E * 4);
  if (g->weighted) g->w.alloc(E * 4);
  k_grid_fill<<<grid_for(ctx, V), BLOCK, 0, ctx->stream>>>(
      R, C, g->off.as<uint64_t>(), g->col.as<uint32_t>(),
      g->weighted ? g->w.as<uint32_t>() : nullptr, weight_key(wseed), wlo, whi);
  GRUS_LAUNCH_CHECK();
  ctx->sync();
  g->compute_derived();
  graph_apply_placement(*g, place);
  return g;
}

void csr_from_sorted_keys(Context &ctx, const uint64_t *keys, uint64_t m, uint64_t V,
                          uint64_t *off, uint32_t *col) {
  k_offsets_from_sorted<<<grid_for(&ctx, m + 1), BLOCK, 0, ctx.stream>>>(keys, m, V, off);
  GRUS_LAUNCH_CHECK();
  if (m) {
    k_low32<<<grid_for(&ctx, m), BLOCK, 0, ctx.stream>>>(keys, m, col);
    GRUS_LAUNCH_CHECK();
  }
}

void graph_copy_csr(Graph &g, uint64_t *off, uint32_t *col, uint32_t *w) {
  g.ctx->activate();
  g.ctx->sync();
  if (off) GRUS_HIP(hipMemcpy(off, g.off.p, (g.V + 1) * 8, hipMemcpyDefault));
  uint64_t near = (g.placement == GRUS_PLACE_CAPPED) ? g.split : g.E;
  if (col && g.E) {
    if (near) GRUS_HIP(hipMemcpy(col, g.col.p, near * 4, hipMemcpyDefault));
    if (near < g.E) memcpy(col + near, g.col_far.p, (g.E - near) * 4);
  }
  if (w && g.E) {
    if (!g.weighted) {
      for (uint64_t i = 0; i < g.E; i++) w[i] = 1;
    } else {
      if (near) GRUS_HIP(hipMemcpy(w, g.w.p, near * 4, hipMemcpyDefault));
      if (near < g.E) memcpy(w + near, g.w_far.p, (g.E - near) * 4);
    }
  }
}

// ------------------------------------------------------------------ placement
// HBM: arrays stay in hipMalloc memory.
// MANAGED: arrays move to hipMallocManaged memory with the reference's hints
//   (include/graph.cuh:206-250: --pf prefetch, --ab AccessedBy, --rm
//   ReadMostly, --pl PreferredLocation).
// CAPPED: offsets and the first `split` edges stay in HBM, sized so the graph
//   uses at most hbm_cap_bytes; the remaining edges (and weights) move to
//   pinned host memory that kernels read over the host link -- the explicit
//   form of graph_chunk::access_edge's local/global split
//   (include/graph.cuh:70-75) and of Set_Mem_Policy's "prefetch what fits"
//   branch (:178-196).
void graph_apply_placement(Graph &g, const grus_place *place) {
  int mode = place ? place->mode : GRUS_PLACE_HBM;
  g.placement = mode;
  if (mode == GRUS_PLACE_HBM) return;
  Context *ctx = g.ctx;
  if (mode == GRUS_PLACE_MANAGED) {
    auto move = [&](Buf &b) {
      if (!b.p) return;
      Buf m;
      m.alloc(b.bytes, Mem::Managed);
      GRUS_HIP(hipMemcpy(m.p, b.p, b.bytes, hipMemcpyDefault));
      b = std::move(m);
      if (place->read_mostly)
        GRUS_HIP(hipMemAdvise(b.p, b.bytes, hipMemAdviseSetReadMostly, ctx->device));
      if (place->preferred_location)
        GRUS_HIP(hipMemAdvise(b.p, b.bytes, hipMemAdviseSetPreferredLocation, ctx->device));
      if (place->accessed_by)
        GRUS_HIP(hipMemAdvise(b.p, b.bytes, hipMemAdviseSetAccessedBy, ctx->device));
      if (place->prefetch) GRUS_HIP(hipMemPrefetchAsync(b.p, b.bytes, ctx->device, ctx->stream));
    };
    move(g.off);
    move(g.col);
    move(g.w);
    ctx->sync();
    g.host_bytes = g.col.bytes + (g.weighted ? g.w.bytes : 0);
    return;
  }
  if (mode != GRUS_PLACE_CAPPED) raise(GRUS_E_INVALID, "unknown placement mode");
  if (!g.symmetric) raise(GRUS_E_UNSUPPORTED, "capped placement needs a symmetric graph");
  uint64_t cap = place->hbm_cap_bytes;
  uint64_t per_edge = g.weighted ? 8 : 4;
  // always resident: the offsets and the first neighbour of every row
  uint64_t fixed = (g.V + 1) * 8 + g.V * 4;
  const uint64_t budget = place->device_budget_bytes;
  if (budget) {
    // whole-device cap (the reference caps the device, tools/mem_limit.cu:6-18,
    // and places from what cudaMemGetInfo leaves, include/graph.cuh:171-205):
    // the graph gets what the budget leaves after the bytes the library holds
    // for other purposes and the largest app workspace
    ctx->scan_tmp.release();  // the build's sort / scan temporaries
    const uint64_t mine = g.off.bytes + g.col.bytes + g.w.bytes + g.first_col.bytes;
    const uint64_t others = devmem_live(ctx->device) - mine;
    const uint64_t need = others + fixed + workspace_reserve(g.V, g.E);
    if (budget < need)
      raise(GRUS_E_BUDGET, "device budget below the offsets, first neighbours and app "
                           "workspaces (" + std::to_string(need) + " bytes)");
    const uint64_t room = budget - need + fixed;
    cap = cap ? std::min(cap, room) : room;
  }
  if (cap && cap < fixed)
    raise(GRUS_E_BUDGET, "HBM cap below the offsets and first-neighbour arrays (12 bytes per vertex)");
  g.first_col.alloc(std::max<uint64_t>(g.V, 1) * 4);
  k_first_col<<<grid_for(ctx, g.V), BLOCK, 0, ctx->stream>>>(
      g.off.as<uint64_t>(), g.col.as<uint32_t>(), g.V, g.first_col.as<uint32_t>());
  GRUS_LAUNCH_CHECK();
  ctx->sync();
  uint64_t fit = cap ? (cap - fixed) / per_edge : g.E;
  uint64_t split = std::min<uint64_t>(fit, g.E);
  // a multiple of 64 edges, so that vector column loads never straddle the
  // HBM / host boundary and stay aligned on both sides of it
  if (split < g.E) split &= ~63ull;
  g.split = split;
  uint64_t far = g.E - split;
  if (far) {
    g.col_far.alloc(far * 4, Mem::HostPinned);
    GRUS_HIP(hipMemcpy(g.col_far.p, g.col.as<uint32_t>() + split, far * 4, hipMemcpyDeviceToHost));
    if (g.weighted) {
      g.w_far.alloc(far * 4, Mem::HostPinned);
      GRUS_HIP(hipMemcpy(g.w_far.p, g.w.as<uint32_t>() + split, far * 4, hipMemcpyDeviceToHost));
    }
  }
  // shrink the device arrays to the resident prefix
  auto shrink = [&](Buf &b, uint64_t n) {
    if (!b.p) return;
    Buf s;
    s.alloc(n * 4);
    if (n) GRUS_HIP(hipMemcpy(s.p, b.p, n * 4, hipMemcpyDeviceToDevice));
    b = std::move(s);
  };
  shrink(g.col, split);
  if (g.weighted) shrink(g.w, split);
  g.host_bytes = far * per_edge;
  // from here on every device allocation counts against the budget
  if (budget) devmem_set_budget(ctx->device, budget);
}

}  // namespace mgg
