// grus_ops.hpp -- the public frontier / operator surface (HIP C++, gfx950).
//
// The reference's extension point: write a job_t plus updater / generator /
// pull_selector functors and drive them with kernel<...> / kernel_pull<...>
// over a Frontier until it is empty (include/frontier.cuh:27-127,
// include/kernel.cuh:124-203, functor contract samples/bfs/bfs.cu:25-87).
// This header keeps that surface -- the same class and functor names, the
// same Init / Next / finish loop -- for user code compiled with hipcc for
// gfx950 against libgrus_hip.so.  The graph comes from a grus_graph handle
// (grus_graph_view) and kernels run on the context's stream
// (grus_ctx_stream), so user operators and library calls are ordered.
//
// Functor contract (default-constructed, as in the reference):
//   bool updater_t()(vtx_t src, vtx_t dst, eid_t edge_id, job_t &job)
//   void generator_t()(bool updated, const next_t &next, vtx_t dst)
//        -- next.push(dst) puts dst in the next frontier (deduplicated)
//   bool pull_selector_t()(vtx_t id, job_t &job)  -- pull: scan id's in-edges?
//   job_t::prepare()  -- host, called before every push (kernel.cuh:147)
// The edge id is the position in the CSR (push) or CSC (pull), as the
// reference passes it; graph_t::weight(edge_id) reads its weight.
//
// What changes against the reference, by design:
//   - frontier flags are bitmaps of u64 words (one bit per vertex, 8x less
//     traffic than char flags); a BDF frontier's next set is such a bitmap,
//     compacted into the work queue by Next() with one atomic per workgroup;
//   - the push over a work queue is edge-balanced for 64-lane waves: a wave
//     takes contiguous 64-edge chunks of the frontier's edge space (one edge
//     per lane) and finds each chunk's owner with one ballot over 64 queue
//     entries, so hubs spread over the whole GPU and low-degree vertices do
//     not idle lanes (the reference gives a 32-lane warp to each vertex);
//   - the pull is degree-binned: a lane scans up to PULL_LANE_EDGES in-edges
//     of its own vertex, and the wave then scans the longer rows together;
//   - errors throw grus::error instead of exiting.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

#include "grus_c.h"

namespace grus {

using vtx_t = uint32_t;
using weight_t = uint32_t;
using eid_t = uint64_t;
constexpr uint32_t INFINIT = GRUS_INFINIT;

// frontier representations (include/common.cuh's frontierType names)
enum frontierType { BDF, BITMAP };

struct error : std::runtime_error {
  int code;
  error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
inline void check(int rc) {
  if (rc != GRUS_OK) throw error(rc, grus_last_error());
}
inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw error(GRUS_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
  }
}

// graph_t<CSR> / graph_t<CSC> as kernels see them (include/graph.cuh:113-268):
// the edges of v are [xadj[v], xadj[v+1]); in capped placement the edges from
// `split` on live in host memory (graph_chunk::access_edge, graph.cuh:70-75)
struct graph_t {
  const uint64_t *xadj;
  const vtx_t *adjncy;
  const weight_t *adjwgt;  // null: unit weights
  const vtx_t *adjncy_far;
  const weight_t *adjwgt_far;
  uint64_t split;
  uint64_t numNode, numEdge;
  __device__ __forceinline__ eid_t begin(vtx_t v) const { return xadj[v]; }
  __device__ __forceinline__ eid_t end(vtx_t v) const { return xadj[v + 1]; }
  __device__ __forceinline__ vtx_t dst(eid_t e) const {
    return e < split ? adjncy[e] : adjncy_far[e - split];
  }
  __device__ __forceinline__ weight_t weight(eid_t e) const {
    if (!adjwgt) return 1u;
    return e < split ? adjwgt[e] : adjwgt_far[e - split];
  }
};

// out-edges (CSR) or in-edges (CSC; built on first use for directed graphs)
inline graph_t graph_view(grus_graph *g, bool in_edges = false) {
  grus_csr_view v{};
  check(grus_graph_view(g, in_edges ? 1 : 0, &v));
  return graph_t{v.off, v.col, v.w, v.col_far, v.w_far, v.split, v.V, v.E};
}
inline hipStream_t stream_of(grus_ctx *ctx) {
  void *s = nullptr;
  check(grus_ctx_stream(ctx, &s));
  return static_cast<hipStream_t>(s);
}

// The next frontier as a generator sees it.
struct next_t {
  unsigned long long *bits;  // one bit per vertex
  uint32_t *active;          // BITMAP frontiers: set when any bit is pushed
  __device__ __forceinline__ void push(vtx_t v) const {
    unsigned long long *w = &bits[v >> 6];
    const unsigned long long b = 1ull << (v & 63);
    if (!(*w & b)) atomicOr(w, b);  // a plain read first: most repeats stop here
    if (active) *active = 1u;       // every writer stores the same value
  }
};

namespace detail {

constexpr int WAVE = 64;
constexpr int BLOCK = 256;
constexpr uint32_t CHUNK = 64;           // push: edges per wave task, one per lane
constexpr uint32_t SCAN_ITEMS = 4;       // scan: items per thread
constexpr uint32_t SCAN_TILE = BLOCK * SCAN_ITEMS;

__device__ __forceinline__ uint32_t lane() { return __lane_id(); }
__device__ __forceinline__ uint32_t lane_rank(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ uint64_t readlane64(uint64_t x, uint32_t l) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)x, l);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(x >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}
// inclusive scan over the workgroup (BLOCK threads); returns the total in *tot
__device__ __forceinline__ uint32_t block_scan_incl(uint32_t x, uint32_t *lds, uint32_t *tot) {
  const uint32_t wid = threadIdx.x / WAVE;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if ((int)lane() >= o) x += y;
  }
  if (lane() == WAVE - 1) lds[wid] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (uint32_t w = 0; w < BLOCK / WAVE; w++) {
    const uint32_t s = lds[w];
    if (w < wid) before += s;
    all += s;
  }
  __syncthreads();
  *tot = all;
  return x + before;
}

__global__ void k_fill_u64(unsigned long long *a, uint64_t n, unsigned long long v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    a[i] = v;
}
__global__ void k_set_u32(uint32_t *p, uint32_t v) { *p = v; }
__global__ void k_seed(unsigned long long *bits, vtx_t src) {
  bits[src >> 6] |= 1ull << (src & 63);
}
// all vertices [0, n) set; bits past n stay clear
__global__ void k_full(unsigned long long *bits, uint64_t n) {
  const uint64_t W = (n + 63) / 64;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < W;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t rem = n - i * 64;
    bits[i] = rem >= 64 ? ~0ull : ((1ull << rem) - 1);
  }
}

// queue <- the set bits of bits[0..W) (ascending within a workgroup), bits
// cleared; one atomic per workgroup reserves its slots
__global__ __launch_bounds__(BLOCK) void k_bits_to_queue(unsigned long long *bits, uint64_t W,
                                                         vtx_t *queue, uint32_t *count) {
  __shared__ uint32_t lds[BLOCK / WAVE];
  __shared__ uint32_t s_base;
  for (uint64_t b0 = (uint64_t)blockIdx.x * BLOCK; b0 < W; b0 += (uint64_t)gridDim.x * BLOCK) {
    const uint64_t i = b0 + threadIdx.x;
    unsigned long long w = 0;
    if (i < W) {
      w = bits[i];
      if (w) bits[i] = 0;
    }
    const uint32_t c = (uint32_t)__popcll(w);
    uint32_t tot = 0;
    const uint32_t incl = block_scan_incl(c, lds, &tot);
    if (threadIdx.x == 0) s_base = tot ? atomicAdd(count, tot) : 0u;
    __syncthreads();
    uint32_t slot = s_base + incl - c;
    for (; w; w &= w - 1) queue[slot++] = (vtx_t)(i * 64 + (uint64_t)(__ffsll(w) - 1));
    __syncthreads();  // s_base reusable
  }
}

// cnt[i] = 64-edge chunks of queue entry i (i < F); cnt[F] = 0
__global__ void k_chunk_counts(graph_t G, const vtx_t *q, uint32_t F, uint32_t *cnt) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= F;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t c = 0;
    if (i < F) {
      const vtx_t u = q[i];
      c = (uint32_t)((G.xadj[u + 1] - G.xadj[u] + CHUNK - 1) / CHUNK);
    }
    cnt[i] = c;
  }
}

// three-pass exclusive scan of n u32 (tiles of SCAN_TILE): tile scans and
// tile sums, one workgroup scans the tile sums, then each tile adds its base
__global__ __launch_bounds__(BLOCK) void k_scan_tiles(const uint32_t *in, uint32_t *out,
                                                      uint64_t n, uint32_t *tile_sum) {
  __shared__ uint32_t lds[BLOCK / WAVE];
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS], s = 0;
#pragma unroll
  for (uint32_t k = 0; k < SCAN_ITEMS; k++) {
    v[k] = base + k < n ? in[base + k] : 0u;
    s += v[k];
  }
  uint32_t tot = 0;
  uint32_t run = block_scan_incl(s, lds, &tot) - s;
#pragma unroll
  for (uint32_t k = 0; k < SCAN_ITEMS; k++) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}
__global__ __launch_bounds__(BLOCK) void k_scan_sums(uint32_t *sum, uint64_t ntiles) {
  __shared__ uint32_t lds[BLOCK / WAVE];
  uint32_t carry = 0;
  for (uint64_t t0 = 0; t0 < ntiles; t0 += BLOCK) {
    const uint64_t i = t0 + threadIdx.x;
    const uint32_t x = i < ntiles ? sum[i] : 0u;
    uint32_t tot = 0;
    const uint32_t incl = block_scan_incl(x, lds, &tot);
    if (i < ntiles) sum[i] = carry + incl - x;
    carry += tot;
  }
}
__global__ void k_scan_add(uint32_t *out, uint64_t n, const uint32_t *tile_base) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] += tile_base[i / SCAN_TILE];
}

// Edge-balanced push over queue q[0..F): chunk c of the frontier's edge space
// is edges [cs[j] .. ) of the queue entry j that owns it, 64 per chunk, one
// per lane.  A wave takes a contiguous run of chunks; the queue entries its
// chunks belong to are read 64 at a time, one per lane, and the owner of a
// chunk is the highest lane whose scan value is <= c (one ballot).
template <class updater_t, class generator_t, class job_t>
__global__ __launch_bounds__(BLOCK) void k_push_queue(graph_t G, const vtx_t *q, uint32_t F,
                                                      const uint32_t *cs, uint32_t chunks,
                                                      job_t job, next_t next) {
  const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) / WAVE;
  const uint64_t per = (chunks + nwaves - 1) / nwaves;
  const uint64_t c0 = wave * per;
  const uint64_t c1 = c0 + per < chunks ? c0 + per : chunks;
  if (c0 >= c1) return;
  uint32_t lo = 0, hi = F + 1;  // last j with cs[j] <= c0
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (cs[mid] <= c0) lo = mid + 1;
    else hi = mid;
  }
  uint32_t jb = lo - 1;
  uint32_t w_cs = 0, w_u = 0;
  uint64_t w_beg = 0, w_end = 0, wend = 0;
  auto load = [&]() {  // queue entries jb .. jb + 63, one per lane
    const uint64_t j = (uint64_t)jb + lane();
    w_cs = j <= F ? cs[j] : 0xFFFFFFFFu;  // cs[F] = chunks: never an owner
    w_u = j < F ? q[j] : 0u;
    w_beg = j < F ? G.xadj[w_u] : 0ull;
    w_end = j < F ? G.xadj[w_u + 1] : 0ull;
    wend = (uint64_t)jb + WAVE <= F ? cs[jb + WAVE] : chunks;
  };
  load();
  for (uint64_t c = c0; c < c1; c++) {
    while (c >= wend) {
      jb += WAVE;
      load();
    }
    const uint64_t below = __ballot(w_cs <= c);
    const uint32_t l = 63u - (uint32_t)__clzll((long long)below);
    const vtx_t u = __builtin_amdgcn_readlane(w_u, l);
    const eid_t e = readlane64(w_beg, l) + (c - __builtin_amdgcn_readlane(w_cs, l)) * CHUNK +
                    lane();
    if (e < readlane64(w_end, l)) {
      const vtx_t v = G.dst(e);
      generator_t()(updater_t()(u, v, e, job), next, v);
    }
  }
}

// push over a bitmap frontier (include/kernel.cuh:104-122, flag to flag):
// a wave per 64-vertex word, its set vertices in turn, lanes over the edges
template <class updater_t, class generator_t, class job_t>
__global__ __launch_bounds__(BLOCK) void k_push_bitmap(graph_t G, const unsigned long long *cur,
                                                       uint64_t W, job_t job, next_t next) {
  const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) / WAVE;
  for (uint64_t i = wave; i < W; i += nwaves) {
    for (unsigned long long m = cur[i]; m; m &= m - 1) {
      const vtx_t u = (vtx_t)(i * 64 + (uint64_t)(__ffsll(m) - 1));
      const eid_t e1 = G.end(u);
      for (eid_t e = G.begin(u) + lane(); e < e1; e += WAVE) {
        const vtx_t v = G.dst(e);
        generator_t()(updater_t()(u, v, e, job), next, v);
      }
    }
  }
}

constexpr uint32_t PULL_LANE_EDGES = 32;  // pull: in-edges a lane scans alone

// pull over in-edges (include/kernel.cuh:18-40): every vertex the selector
// accepts calls the updater for each in-neighbour in the current frontier
template <class updater_t, class generator_t, class pull_selector_t, class job_t>
__global__ __launch_bounds__(BLOCK) void k_pull_bitmap(graph_t Gin,
                                                       const unsigned long long *cur, job_t job,
                                                       next_t next) {
  const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) / WAVE;
  const uint64_t W = (Gin.numNode + 63) / 64;
  for (uint64_t i = wave; i < W; i += nwaves) {
    const uint64_t v64 = i * 64 + lane();
    const vtx_t v = (vtx_t)v64;
    bool sel = v64 < Gin.numNode && pull_selector_t()(v, job);
    eid_t e = 0, e1 = 0;
    if (sel) {
      e = Gin.begin(v);
      e1 = Gin.end(v);
      const eid_t lim = e1 - e > PULL_LANE_EDGES ? e + PULL_LANE_EDGES : e1;
      for (; e < lim; e++) {
        const vtx_t u = Gin.dst(e);
        if ((cur[u >> 6] >> (u & 63)) & 1ull) generator_t()(updater_t()(u, v, e, job), next, v);
      }
    }
    // the longer rows, together: every lane strides the row of one vertex
    for (uint64_t longm = __ballot(sel && e < e1); longm; longm &= longm - 1) {
      const uint32_t l = (uint32_t)(__ffsll((unsigned long long)longm) - 1);
      const vtx_t lv = __builtin_amdgcn_readlane(v, l);
      const eid_t le1 = readlane64(e1, l);
      for (eid_t le = readlane64(e, l) + lane(); le < le1; le += WAVE) {
        const vtx_t u = Gin.dst(le);
        if ((cur[u >> 6] >> (u & 63)) & 1ull) generator_t()(updater_t()(u, lv, le, job), next, lv);
      }
    }
  }
}

inline unsigned grid_for(uint64_t items, unsigned per_block = BLOCK) {
  uint64_t g = (items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;  // grid-stride loops cover the rest
  return (unsigned)g;
}

template <class T>
T *dalloc(uint64_t n) {
  void *p = nullptr;
  hip_check(hipMalloc(&p, (n ? n : 1) * sizeof(T)), "hipMalloc");
  return static_cast<T *>(p);
}

}  // namespace detail

template <frontierType type>
class Frontier;

// Frontier<BDF> (include/frontier.cuh:27-76): the current frontier is a work
// queue, the next one a bitmap that Next() compacts into the queue.
template <>
class Frontier<BDF> {
 public:
  vtx_t numNode = 0;
  vtx_t *queue = nullptr;        // the work items, wl_sz of them
  uint32_t *count = nullptr;     // device: queue length during Next()
  unsigned long long *bits = nullptr;  // the next frontier
  uint32_t *cnt = nullptr, *cs = nullptr, *tile = nullptr;  // push scans
  vtx_t wl_sz = 0;
  hipStream_t stream = nullptr;

  Frontier() = default;
  Frontier(const Frontier &) = delete;
  Frontier &operator=(const Frontier &) = delete;
  ~Frontier() { release(); }

  void Init(vtx_t n, vtx_t src, hipStream_t s, bool full = false) {
    release();
    numNode = n;
    stream = s;
    const uint64_t W = words();
    queue = detail::dalloc<vtx_t>((uint64_t)n + 1);
    count = detail::dalloc<uint32_t>(1);
    bits = detail::dalloc<unsigned long long>(W + 1);
    cnt = detail::dalloc<uint32_t>((uint64_t)n + 2);
    cs = detail::dalloc<uint32_t>((uint64_t)n + 2);
    tile = detail::dalloc<uint32_t>((n + 2) / detail::SCAN_TILE + 2);
    detail::k_fill_u64<<<detail::grid_for(W + 1), detail::BLOCK, 0, stream>>>(bits, W + 1, 0);
    hip_check(hipGetLastError(), "frontier init");
    // the first work queue: every vertex, or the source
    if (full) {
      detail::k_full<<<detail::grid_for(W), detail::BLOCK, 0, stream>>>(bits, n);
    } else {
      if (src >= n) throw error(GRUS_E_INVALID, "source vertex out of range");
      detail::k_seed<<<1, 1, 0, stream>>>(bits, src);
    }
    hip_check(hipGetLastError(), "frontier init");
    Next();
  }
  // wl_c <- the next set (include/frontier.cuh:64-70); blocking size read
  void Next() {
    detail::k_set_u32<<<1, 1, 0, stream>>>(count, 0u);
    detail::k_bits_to_queue<<<detail::grid_for(words()), detail::BLOCK, 0, stream>>>(
        bits, words(), queue, count);
    hip_check(hipGetLastError(), "Frontier<BDF>::Next");
    hip_check(hipMemcpyAsync(&wl_sz, count, 4, hipMemcpyDeviceToHost, stream), "size read");
    hip_check(hipStreamSynchronize(stream), "size read");
  }
  bool finish() const { return wl_sz == 0; }
  vtx_t get_work_size_h() const { return wl_sz; }
  next_t next() const { return next_t{bits, nullptr}; }
  uint64_t words() const { return ((uint64_t)numNode + 63) / 64; }

 private:
  void release() {
    for (void *p : {(void *)queue, (void *)count, (void *)bits, (void *)cnt, (void *)cs,
                    (void *)tile})
      if (p) (void)hipFree(p);
    queue = nullptr;
    count = nullptr;
    bits = nullptr;
    cnt = cs = tile = nullptr;
  }
};

// Frontier<BITMAP> (include/frontier.cuh:77-127): current and next bitmaps
// and a device flag that any push into the next one sets.
template <>
class Frontier<BITMAP> {
 public:
  vtx_t numNode = 0;
  unsigned long long *flag1 = nullptr, *flag2 = nullptr;  // current, next
  uint32_t *active_d = nullptr;
  uint32_t active_h = 1;
  bool finished = false;
  hipStream_t stream = nullptr;

  Frontier() = default;
  Frontier(const Frontier &) = delete;
  Frontier &operator=(const Frontier &) = delete;
  ~Frontier() { release(); }

  void Init(vtx_t n, vtx_t src, hipStream_t s, bool full = false) {
    release();
    numNode = n;
    stream = s;
    const uint64_t W = words();
    flag1 = detail::dalloc<unsigned long long>(W + 1);
    flag2 = detail::dalloc<unsigned long long>(W + 1);
    active_d = detail::dalloc<uint32_t>(1);
    detail::k_fill_u64<<<detail::grid_for(W + 1), detail::BLOCK, 0, stream>>>(flag1, W + 1, 0);
    detail::k_fill_u64<<<detail::grid_for(W + 1), detail::BLOCK, 0, stream>>>(flag2, W + 1, 0);
    detail::k_set_u32<<<1, 1, 0, stream>>>(active_d, 0u);
    if (full) {
      detail::k_full<<<detail::grid_for(W), detail::BLOCK, 0, stream>>>(flag1, n);
    } else {
      if (src >= n) throw error(GRUS_E_INVALID, "source vertex out of range");
      detail::k_seed<<<1, 1, 0, stream>>>(flag1, src);
    }
    hip_check(hipGetLastError(), "Frontier<BITMAP>::Init");
    finished = (n == 0);
  }
  // swap (frontier.cuh:108-117): the next set becomes current; finished when
  // no generator pushed anything during the step; blocking flag read
  void Next() {
    hip_check(hipMemcpyAsync(&active_h, active_d, 4, hipMemcpyDeviceToHost, stream), "flag read");
    hip_check(hipStreamSynchronize(stream), "flag read");
    finished = active_h == 0;
    std::swap(flag1, flag2);
    detail::k_fill_u64<<<detail::grid_for(words() + 1), detail::BLOCK, 0, stream>>>(
        flag2, words() + 1, 0);
    detail::k_set_u32<<<1, 1, 0, stream>>>(active_d, 0u);
    hip_check(hipGetLastError(), "Frontier<BITMAP>::Next");
  }
  bool finish() const { return finished; }
  next_t next() const { return next_t{flag2, active_d}; }
  uint64_t words() const { return ((uint64_t)numNode + 63) / 64; }

 private:
  void release() {
    for (void *p : {(void *)flag1, (void *)flag2, (void *)active_d})
      if (p) (void)hipFree(p);
    flag1 = flag2 = nullptr;
    active_d = nullptr;
  }
};

// ---- operators ----
// push over a work queue: job.prepare(), then every edge of every queued
// vertex through updater -> generator
template <class updater_t, class generator_t, class job_t>
void push(const graph_t &G, Frontier<BDF> &F, job_t &job) {
  job.prepare();
  const uint32_t n = F.wl_sz;
  if (n == 0) return;
  if ((uint64_t)n + G.numEdge / detail::CHUNK >= 0xFFFFFFFFull)
    throw error(GRUS_E_UNSUPPORTED, "push: more than 2^32 edge chunks");
  const hipStream_t s = F.stream;
  detail::k_chunk_counts<<<detail::grid_for((uint64_t)n + 1), detail::BLOCK, 0, s>>>(
      G, F.queue, n, F.cnt);
  const uint64_t items = (uint64_t)n + 1;
  const uint64_t tiles = (items + detail::SCAN_TILE - 1) / detail::SCAN_TILE;
  detail::k_scan_tiles<<<(unsigned)tiles, detail::BLOCK, 0, s>>>(F.cnt, F.cs, items, F.tile);
  detail::k_scan_sums<<<1, detail::BLOCK, 0, s>>>(F.tile, tiles);
  detail::k_scan_add<<<detail::grid_for(items), detail::BLOCK, 0, s>>>(F.cs, items, F.tile);
  hip_check(hipGetLastError(), "push: chunk scan");
  uint32_t chunks = 0;  // cs[n]: the number of chunks
  hip_check(hipMemcpyAsync(&chunks, F.cs + n, 4, hipMemcpyDeviceToHost, s), "chunk count");
  hip_check(hipStreamSynchronize(s), "chunk count");
  if (chunks == 0) return;
  detail::k_push_queue<updater_t, generator_t, job_t>
      <<<detail::grid_for((uint64_t)chunks * detail::WAVE), detail::BLOCK, 0, s>>>(
          G, F.queue, n, F.cs, chunks, job, F.next());
  hip_check(hipGetLastError(), "push");
}

// push over a bitmap frontier
template <class updater_t, class generator_t, class job_t>
void push(const graph_t &G, Frontier<BITMAP> &F, job_t &job) {
  job.prepare();
  detail::k_push_bitmap<updater_t, generator_t, job_t>
      <<<detail::grid_for(F.words() * detail::WAVE), detail::BLOCK, 0, F.stream>>>(
          G, F.flag1, F.words(), job, F.next());
  hip_check(hipGetLastError(), "push");
}

// pull over the in-edges (Gin = graph_view(g, true)) of a bitmap frontier
template <class updater_t, class generator_t, class pull_selector_t, class job_t>
void pull(const graph_t &Gin, Frontier<BITMAP> &F, job_t &job) {
  detail::k_pull_bitmap<updater_t, generator_t, pull_selector_t, job_t>
      <<<detail::grid_for(F.words() * detail::WAVE), detail::BLOCK, 0, F.stream>>>(
          Gin, F.flag1, job, F.next());
  hip_check(hipGetLastError(), "pull");
}

// the reference's dispatch functors (include/kernel.cuh:124-203)
template <class graph_type, class frontier_t, class updater_t, class generator_t, class job_t>
class kernel {
 public:
  void operator()(const graph_type &G, frontier_t &F, job_t &job) {
    push<updater_t, generator_t, job_t>(G, F, job);
  }
};
template <class updater_t, class generator_t, class pull_selector_t, class job_t>
class kernel_pull {
 public:
  void operator()(const graph_t &Gin, Frontier<BITMAP> &F, job_t &job) {
    pull<updater_t, generator_t, pull_selector_t, job_t>(Gin, F, job);
  }
};

}  // namespace grus
