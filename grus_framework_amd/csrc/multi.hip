// multi.hip -- batched multi-source BFS and SSSP engines (see multi.hpp).
//
// BFS lanes: visit/front/next are u64 masks per vertex.  A step expands the
// frontier either by push (edge-balanced operator over the queue of frontier
// vertices: next[v] |= front[u] & ~visit[v]) or by pull (a lane per vertex
// ORs front[] over its in-edges until every lane still missing v is
// satisfied), then a finalize pass commits the new bits, writes the levels
// (row per lane, coalesced: the wave writes one row at a time over 64
// consecutive vertices) and rebuilds the frontier queue.
// SSSP lanes: distances are stored vertex-major, dist[v*64 + lane], so one
// relaxation of edge (u,v) for all lanes is a single coalesced 64-lane
// atomicMin over 256 bytes.  Rounds are Bellman-Ford over the frontier
// (front bit = the lane's distance at u improved last round); a wave owns a
// chunk of up to 64 edges of one frontier vertex.
#include <algorithm>
#include <cstring>
#include <string>

#include "advance.hpp"
#include "device.hpp"
#include "multi.hpp"
#include "prims.hpp"

namespace mgg {

namespace ms {

constexpr uint32_t K_PULL = 32;
constexpr uint32_t CHUNK = 64;       // SSSP relax: out-edges per wave task (one per lane)
#ifndef SSSP_SPARSE_LANES
#define SSSP_SPARSE_LANES 16
#endif
// at most this many active lanes at a vertex: edge-parallel relax
constexpr int SPARSE_LANES = SSSP_SPARSE_LANES;
#ifndef RELAX_U16_EDGES
#define RELAX_U16_EDGES 4
#endif
constexpr uint32_t RELAX_U16 = RELAX_U16_EDGES;  // 16-bit relax: edges whose CAS are in flight together
static_assert(CHUNK == WAVE, "the edge-parallel relax covers a chunk with one edge per lane");

struct LaneStart {
  uint32_t s[64];
};

__device__ __forceinline__ uint64_t wave_or(uint64_t x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o);
  return x;
}

// order-independent digest of a result row: sum of mix64(v<<32 | value)
__device__ __forceinline__ uint64_t digest_term(uint64_t v, uint32_t val) {
  return mix64((v << 32) | val);
}

__global__ void k_reset_masks(uint64_t *visit, uint64_t *front, uint64_t *next, uint64_t V,
                              uint64_t keep) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < V; i += stride) {
    if (visit) visit[i] &= keep;
    front[i] &= keep;
    next[i] &= keep;
  }
}

// BFS rows (row-major [lane][V]) of lanes in mask <- INFINIT
__global__ void k_reset_rows(uint32_t *level, uint64_t V, uint64_t mask) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < V; i += stride) {
    uint64_t m = mask;
    while (m) {
      int b = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      level[(uint64_t)b * V + i] = INFINIT;
    }
  }
}

// SSSP rows: 128 entries per vertex, 16 bits (0xFFFF unreached) or 32 bits
constexpr uint32_t INF16 = 0xFFFFu;
__device__ __forceinline__ uint32_t wide_dist(uint16_t x) { return x == INF16 ? INFINIT : x; }
__device__ __forceinline__ uint32_t wide_dist(uint32_t x) { return x; }
template <class T>
constexpr uint32_t inf_of() {
  return sizeof(T) == 2 ? INF16 : INFINIT;
}
__device__ __forceinline__ bool slot_in(uint64_t lo, uint64_t hi, uint32_t s) {
  return s < 64 ? ((lo >> s) & 1ull) : ((hi >> (s - 64)) & 1ull);
}

// SSSP entries of the slots in (lo, hi) <- unreached
template <class T>
__global__ void k_reset_slots(T *dist, uint64_t V, uint64_t lo, uint64_t hi) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < V * 128; i += stride)
    if (slot_in(lo, hi, (uint32_t)(i & 127))) dist[i] = (T)inf_of<T>();
}

// 16-bit entries to 32-bit entries (0xFFFF -> INFINIT)
__global__ void k_widen(const uint16_t *in, uint32_t *out, uint64_t n) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = wide_dist(in[i]);
}

// sources into their slots: BFS slot = lane, SSSP slot s in lane s / 2
__global__ void k_inject(const uint32_t *srcs, const uint32_t *slots, int n, bool bfs,
                         bool narrow, uint64_t *visit, uint64_t *front, uint32_t *vals,
                         uint64_t V, uint32_t *queue, MsCounters *cnt) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t s = srcs[k], slot = slots[k];
  const uint32_t lane = bfs ? slot : slot >> 1;
  const unsigned long long bit = 1ull << lane;
  if (bfs) {
    atomicOr((unsigned long long *)&visit[s], bit);
    vals[(uint64_t)lane * V + s] = 0;  // level[lane][s]
  } else if (narrow) {
    reinterpret_cast<uint16_t *>(vals)[(uint64_t)s * 128 + slot] = 0;  // dist[s][slot]
  } else {
    vals[(uint64_t)s * 128 + slot] = 0;
  }
  unsigned long long old = atomicOr((unsigned long long *)&front[s], bit);
  if (old == 0) queue[atomicAdd(&cnt->qcount, 1u)] = s;  // not yet queued
}

// ---- BFS pull ----
__global__ __launch_bounds__(BLOCK) void k_bfs_pull(CsrView in, const uint64_t *visit,
                                                    const uint64_t *front, uint64_t *next,
                                                    uint64_t running, uint32_t *longq,
                                                    MsCounters *cnt) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t first = blockIdx.x * (uint64_t)blockDim.x + (threadIdx.x & ~(uint64_t)(WAVE - 1));
  for (uint64_t base = first; base < in.V; base += stride) {
    uint64_t v = base + lane_id();
    bool is_long = false;
    if (v < in.V) {
      uint64_t need = running & ~visit[v];
      if (need) {
        uint64_t e = in.off[v], e1 = in.off[v + 1];
        uint64_t lim = (e1 - e > K_PULL) ? e + K_PULL : e1;
        uint64_t acc = 0;
        for (; e < lim; e++) {
          acc |= front[edge_col(in, e)];
          if ((acc & need) == need) break;
        }
        acc &= need;
        is_long = (acc != need) && lim < e1;
        if (acc) next[v] = acc;
      }
    }
    uint32_t slot = wave_append(&cnt->long_count, is_long);
    if (is_long) longq[slot] = (uint32_t)v;
  }
}

__global__ __launch_bounds__(BLOCK) void k_bfs_pull_long(CsrView in, const uint64_t *visit,
                                                         const uint64_t *front, uint64_t *next,
                                                         uint64_t running, const uint32_t *longq,
                                                         const MsCounters *cnt) {
  // a workgroup per long row, its waves striding over the row (a hub's row
  // would otherwise keep one wave busy long after the rest have finished);
  // each wave stops once its own bits cover the lanes still looking, and
  // ORs what it found into next[v], which only ever holds bits of need
  const uint32_t wid = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
  const uint32_t n = cnt->long_count;
  for (uint64_t i = blockIdx.x; i < n; i += gridDim.x) {
    const uint32_t v = longq[i];
    const uint64_t need = running & ~visit[v];
    uint64_t acc = next[v];
    const uint64_t e1 = in.off[v + 1];
    for (uint64_t base = in.off[v] + K_PULL + (uint64_t)wid * WAVE;
         base < e1 && (acc & need) != need; base += (uint64_t)nw * WAVE) {
      const uint64_t e = base + lane_id();
      const uint64_t x = e < e1 ? front[edge_col(in, e)] : 0ull;
      acc |= wave_or(x);
    }
    if (lane_id() == 0 && (acc & need))
      atomicOr((unsigned long long *)&next[v], (unsigned long long)(acc & need));
  }
}

// ---- BFS push job for the edge-balanced operator ----
struct push_job_t {
  static constexpr bool kFar = false;
  const uint64_t *front, *visit;
  uint64_t *next;
  __device__ __forceinline__ bool update(vtx_t src, vtx_t dst, eid_t, weight_t, uint32_t *) const {
    uint64_t m = front[src] & ~visit[dst];
    if (m && (next[dst] & m) != m) atomicOr((unsigned long long *)&next[dst], m);
    return false;
  }
  __device__ __forceinline__ int generate(bool, vtx_t, uint32_t) const { return 0; }
  __device__ __forceinline__ uint64_t out_degree(vtx_t) const { return 0; }
};

// ---- BFS finalize: commit new bits, levels, queue, lane counts ----
__global__ __launch_bounds__(BLOCK) void k_bfs_finalize(const uint64_t *off, uint64_t *visit,
                                                        uint64_t *front, uint64_t *next,
                                                        uint32_t *level, uint64_t V,
                                                        uint64_t running, uint32_t t,
                                                        LaneStart st, uint32_t *queue,
                                                        MsCounters *cnt) {
  __shared__ unsigned long long s_alive;
  __shared__ uint32_t s_app[BLOCK / WAVE + 1];
  if (threadIdx.x == 0) s_alive = 0;
  __syncthreads();
  uint64_t edges = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < V; base += stride) {
    uint64_t v = base + threadIdx.x;
    uint64_t nw = 0;
    if (v < V) {
      uint64_t nx = next[v];
      if (nx) {
        next[v] = 0;
        uint64_t vis = visit[v];
        nw = nx & ~vis & running;
        if (nw) visit[v] = vis | nw;
      }
      if (front[v] != nw) front[v] = nw;
    }
    uint64_t wm = wave_or(nw);
    if (lane_id() == 0 && wm) atomicOr(&s_alive, (unsigned long long)wm);
    while (wm) {  // level rows: one coalesced 64-vertex store per lane present
      int i = __ffsll((unsigned long long)wm) - 1;
      wm &= wm - 1;
      if ((nw >> i) & 1ull) level[(uint64_t)i * V + v] = t + 1 - st.s[i];
    }
    uint32_t slot = block_append(&cnt->qcount, nw != 0, s_app);
    if (nw) {
      queue[slot] = (uint32_t)v;
      edges += off[v + 1] - off[v];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && s_alive) atomicOr((unsigned long long *)&cnt->alive, s_alive);
  block_add_u64(edges, (unsigned long long *)&cnt->front_edges);
}

// ---- SSSP ----
// chunk counts: ceil(deg(q[i]) / CHUNK), plus a trailing zero
__global__ void k_chunk_counts(const uint64_t *off, const uint32_t *q, uint32_t F, uint64_t *c) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i <= F; i += stride) {
    uint64_t d = 0;
    if (i < F) {
      uint32_t u = q[i];
      d = (off[u + 1] - off[u] + CHUNK - 1) / CHUNK;
    }
    c[i] = d;
  }
}

__device__ __forceinline__ uint64_t readlane_u64(uint64_t x, uint32_t i) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)x, i);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(x >> 32), i);
  return ((uint64_t)hi << 32) | lo;
}

// Lowers 16-bit half `half` of *word to nd (a previous read of the word is
// `old`); returns once the half is at most nd.  Values only decrease, so
// every failed compare-and-swap means another lane lowered the word and the
// loop ends after finitely many of them.
__device__ __forceinline__ void cas_min16(uint32_t *word, uint32_t half, uint32_t nd,
                                          uint32_t old) {
  for (;;) {
    const uint32_t cur = half ? (old >> 16) : (old & 0xFFFFu);
    if (nd >= cur) return;
    const uint32_t nw = half ? ((old & 0xFFFFu) | (nd << 16)) : ((old & 0xFFFF0000u) | nd);
    const uint32_t prev = atomicCAS(word, old, nw);
    if (prev == old) return;
    old = prev;
  }
}

// the two distances of lane p at vertex v
template <bool NARROW>
__device__ __forceinline__ uint2 ld_pair(const uint32_t *dist, uint64_t v, uint32_t p) {
  if constexpr (NARROW) {
    const uint32_t w = dist[v * 64 + p];
    return make_uint2(w & 0xFFFFu, w >> 16);
  } else {
    return reinterpret_cast<const uint2 *>(dist)[v * 64 + p];
  }
}

// Relaxation of one 64-edge chunk of frontier vertex u per wave task, for
// 128 queries in 64 lanes of two.  Lane l loads edge e+l of the chunk
// (coalesced), and the edges are then broadcast to the wave with readlane.
// Distances only decrease within a round, so a plain read of the pair at v
// that already shows a new distance below it proves that it drops this
// round: the next-frontier bit and the slot's "improved" bit are set from
// that read.
//   32-bit entries: atomicMin per improved slot, issued without waiting for
//   the old values; the row reads of 8 edges are in flight before any.
//   16-bit entries: the lane lowers its 32-bit word by compare-and-swap
//   (the device has no 16-bit atomic minimum), the CAS of 8 edges issued
//   before any result is awaited; the rare failures retry from the value
//   returned.  A slot whose new distance would be 0xFFFF or more where it
//   has none is flagged in cnt->overflow instead of stored.
// The chunk count is read from the scan (CS[F]); the grid, sized by the
// host from an estimate, only sets how many waves share the chunks.
template <bool NARROW>
__global__ __launch_bounds__(BLOCK) void k_sssp_relax(CsrView g, const uint32_t *q, uint32_t F,
                                                      const uint64_t *CS, const uint64_t *front,
                                                      uint64_t *next, uint32_t *dist,
                                                      MsCounters *cnt) {
  constexpr uint32_t INF = NARROW ? INF16 : INFINIT;
  const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) / WAVE;
  const uint32_t lane = lane_id();
  const uint64_t chunks = CS[F];
  // slots that improved a distance / that overflowed: bit p = lane p's slot
  uint64_t imp_a = 0, imp_b = 0, ovf_a = 0, ovf_b = 0;
  bool dimp_a = false, dimp_b = false, dovf_a = false, dovf_b = false;
  // each wave takes a contiguous run of chunks (equal counts, so equal work
  // within one chunk).  The queue entries its chunks belong to are read 64
  // at a time, one per lane (scan value, vertex, row bounds, front word):
  // the owner of chunk c is then the highest lane whose scan value is <= c
  // (a ballot), and the entry's fields are broadcast with readlane
  const uint64_t per = (chunks + nwaves - 1) / nwaves;
  const uint64_t c0 = wave * per;
  const uint64_t c1 = c0 + per < chunks ? c0 + per : chunks;
  if (c0 < c1) {
    uint32_t jb = upper_bound_u64(CS, 0, F + 1, c0) - 1;  // owner of c0
    uint64_t w_cs = 0, w_off = 0, w_end = 0, w_fm = 0, wend = 0;
    uint32_t w_u = 0;
    auto load_window = [&]() {  // queue entries jb .. jb+63
      const uint64_t jl = (uint64_t)jb + lane;
      w_cs = jl <= F ? CS[jl] : ~0ull;  // CS[F] = chunks: never an owner
      const bool in = jl < F;
      w_u = in ? q[jl] : 0u;
      w_off = in ? g.off[w_u] : 0ull;
      w_end = in ? g.off[w_u + 1] : 0ull;
      w_fm = in ? front[w_u] : 0ull;
      wend = (uint64_t)jb + WAVE <= F ? CS[jb + WAVE] : chunks;  // first chunk past the window
    };
    load_window();
    for (uint64_t c = c0; c < c1; c++) {
      while (c >= wend) {  // (queued vertices without out-edges own no chunk)
        jb += WAVE;
        load_window();
      }
      const uint64_t below = __ballot(w_cs <= c);  // never 0: lane 0's scan value is <= c
      const uint32_t l = 63u - (uint32_t)__clzll((long long)below);
      const uint32_t u = __builtin_amdgcn_readlane(w_u, l);
      const uint64_t e = readlane_u64(w_off, l) + (c - readlane_u64(w_cs, l)) * CHUNK;
      uint64_t e1 = readlane_u64(w_end, l);
      if (e1 > e + CHUNK) e1 = e + CHUNK;
      const uint32_t n = (uint32_t)(e1 - e);  // 1..64, wave-uniform
      const bool valid = lane < n;
      const uint32_t my_v = valid ? edge_col(g, e + lane) : u;
      const uint32_t my_w = valid ? edge_w(g, e + lane) : 0u;
      const uint64_t fm = readlane_u64(w_fm, l);
      const uint2 du_l = ld_pair<NARROW>(dist, u, lane);  // lane l: its two queries' distances at u
      if (__popcll(fm) <= SPARSE_LANES) {
        // few lanes active at u: lanes take different edges and loop over
        // the active lane bits (wave-uniform loop)
        uint64_t nxt = 0;
        for (uint64_t m = fm; m; m &= m - 1) {
          const uint32_t qb = (uint32_t)(__ffsll((unsigned long long)m) - 1);
          const uint32_t da = __builtin_amdgcn_readlane(du_l.x, qb);
          const uint32_t db = __builtin_amdgcn_readlane(du_l.y, qb);
          if (!valid) continue;
          const uint64_t na = (uint64_t)da + my_w, nb = (uint64_t)db + my_w;
          const uint2 cur = ld_pair<NARROW>(dist, my_v, qb);
          const bool ia = da != INF && na < cur.x, ib = db != INF && nb < cur.y;
          if (NARROW) {
            if (da != INF && !ia && cur.x == INF) ovf_a |= 1ull << qb;
            if (db != INF && !ib && cur.y == INF) ovf_b |= 1ull << qb;
          }
          if (!ia && !ib) continue;
          if (ia) imp_a |= 1ull << qb;
          if (ib) imp_b |= 1ull << qb;
          nxt |= 1ull << qb;
          if constexpr (NARROW) {
            uint32_t *wp = &dist[(uint64_t)my_v * 64 + qb];
            const uint32_t old = (cur.y << 16) | cur.x;
            if (ia) cas_min16(wp, 0, (uint32_t)na, old);
            if (ib) cas_min16(wp, 1, (uint32_t)nb, *wp);
          } else {
            uint32_t *pp = &dist[((uint64_t)my_v * 64 + qb) * 2];
            if (ia) atomicMin(pp, (uint32_t)na);
            if (ib) atomicMin(pp + 1, (uint32_t)nb);
          }
        }
        if (nxt) atomicOr((unsigned long long *)&next[my_v], nxt);
        continue;
      }
      const bool act = (fm >> lane) & 1ull;
      const bool act_a = act && du_l.x != INF, act_b = act && du_l.y != INF;
      constexpr uint32_t U = NARROW ? RELAX_U16 : 8;
      for (uint32_t k0 = 0; k0 < n; k0 += U) {
        uint32_t vv[U];
        if constexpr (NARROW) {
          // the lane's 32-bit word per edge: read, lowered value, CAS result
          uint32_t wd[U], want[U];
#pragma unroll
          for (uint32_t k = 0; k < U; k++) {  // issue up to U row reads
            const uint32_t jj = k0 + k < n ? k0 + k : n - 1;
            vv[k] = __builtin_amdgcn_readlane(my_v, jj);
            wd[k] = dist[(uint64_t)vv[k] * 64 + lane];
          }
#pragma unroll
          for (uint32_t k = 0; k < U; k++) {
            const bool in = k0 + k < n;
            // weights clamped to 2^16: any larger one gives no 16-bit distance,
            // and the sums stay within 32 bits
            const uint32_t w = min(__builtin_amdgcn_readlane(my_w, in ? k0 + k : 0), 0x10000u);
            const uint32_t ca = wd[k] & 0xFFFFu, cb = wd[k] >> 16;
            const uint32_t na = du_l.x + w, nb = du_l.y + w;
            const bool ia = in && act_a && na < ca, ib = in && act_b && nb < cb;
            dovf_a |= in && act_a && !ia && ca == INF;
            dovf_b |= in && act_b && !ib && cb == INF;
            dimp_a |= ia;
            dimp_b |= ib;
            const uint64_t m = __ballot(ia || ib);
            if (m && lane == 0) atomicOr((unsigned long long *)&next[vv[k]], m);
            want[k] = ((ib ? nb : cb) << 16) | (ia ? na : ca);
          }
#pragma unroll
          for (uint32_t k = 0; k < U; k++)  // every CAS in flight before any result is used
            if (want[k] != wd[k])
              want[k] = atomicCAS(&dist[(uint64_t)vv[k] * 64 + lane], wd[k], want[k]) == wd[k]
                            ? wd[k]
                            : want[k];
#pragma unroll
          for (uint32_t k = 0; k < U; k++) {
            if (want[k] == wd[k]) continue;  // stored, or nothing to store
            // another wave lowered the word first: lower each half from its value
            uint32_t *wp = &dist[(uint64_t)vv[k] * 64 + lane];
            cas_min16(wp, 0, want[k] & 0xFFFFu, *wp);
            cas_min16(wp, 1, want[k] >> 16, *wp);
          }
        } else {
          uint2 dd[U];
#pragma unroll
          for (uint32_t k = 0; k < U; k++) {  // issue up to U row reads
            const uint32_t jj = k0 + k < n ? k0 + k : n - 1;
            vv[k] = __builtin_amdgcn_readlane(my_v, jj);
            dd[k] = ld_pair<false>(dist, vv[k], lane);
          }
#pragma unroll
          for (uint32_t k = 0; k < U; k++) {
            const bool in = k0 + k < n;
            const uint32_t w = __builtin_amdgcn_readlane(my_w, in ? k0 + k : 0);
            const uint64_t na = (uint64_t)du_l.x + w, nb = (uint64_t)du_l.y + w;
            const bool ia = in && act_a && na < dd[k].x, ib = in && act_b && nb < dd[k].y;
            dimp_a |= ia;
            dimp_b |= ib;
            const uint64_t m = __ballot(ia || ib);
            if (m && lane == 0) atomicOr((unsigned long long *)&next[vv[k]], m);
            uint32_t *pp = &dist[((uint64_t)vv[k] * 64 + lane) * 2];
            if (ia) atomicMin(pp, (uint32_t)na);
            if (ib) atomicMin(pp + 1, (uint32_t)nb);
          }
        }
      }
    }
  }
  imp_a = wave_or(imp_a) | __ballot(dimp_a);
  imp_b = wave_or(imp_b) | __ballot(dimp_b);
  if (lane == 0) {
    if (imp_a) atomicOr((unsigned long long *)&cnt->slot_alive[0], imp_a);
    if (imp_b) atomicOr((unsigned long long *)&cnt->slot_alive[1], imp_b);
  }
  if (NARROW) {
    ovf_a = wave_or(ovf_a) | __ballot(dovf_a);
    ovf_b = wave_or(ovf_b) | __ballot(dovf_b);
    if (lane == 0) {
      if (ovf_a) atomicOr((unsigned long long *)&cnt->overflow[0], ovf_a);
      if (ovf_b) atomicOr((unsigned long long *)&cnt->overflow[1], ovf_b);
    }
  }
}

// A workgroup owns a contiguous span of `items` x BLOCK vertices (items <= 32,
// one bit per item in `bits`): it commits the next bits of its span, counts
// its new frontier vertices, reserves their queue slots with one atomic and
// then writes them.  One queue atomic per workgroup instead of one per
// workgroup-wide step of a grid-stride loop: the single counter address is
// the serial point of the pass.
__global__ __launch_bounds__(BLOCK) void k_sssp_finalize(const uint64_t *off, uint64_t *front,
                                                         uint64_t *next, uint64_t V, uint32_t items,
                                                         uint64_t running, uint32_t *queue,
                                                         MsCounters *cnt) {
  __shared__ unsigned long long s_alive;
  __shared__ uint32_t s_wpre[BLOCK / WAVE];
  __shared__ uint32_t s_base;
  if (threadIdx.x == 0) s_alive = 0;
  __syncthreads();
  const uint64_t first = (uint64_t)blockIdx.x * items * BLOCK + threadIdx.x;
  uint32_t bits = 0;
  uint64_t alive = 0, edges = 0;
  for (uint32_t i = 0; i < items; i++) {
    const uint64_t v = first + (uint64_t)i * BLOCK;
    if (v < V) {
      const uint64_t nx = next[v];
      if (nx) next[v] = 0;
      const uint64_t nw = nx & running;
      if (front[v] != nw) front[v] = nw;
      if (nw) {
        bits |= 1u << i;
        alive |= nw;
        edges += off[v + 1] - off[v];
      }
    }
  }
  const uint64_t wm = wave_or(alive);
  if (lane_id() == 0 && wm) atomicOr(&s_alive, (unsigned long long)wm);
  // exclusive scan of the per-thread counts over the workgroup
  const uint32_t c = __popc(bits);
  uint32_t incl = c;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o);
    if (lane_id() >= (uint32_t)o) incl += y;
  }
  const uint32_t wid = threadIdx.x / WAVE;
  if (lane_id() == WAVE - 1) s_wpre[wid] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (uint32_t w = 0; w < BLOCK / WAVE; w++) {
      const uint32_t x = s_wpre[w];
      s_wpre[w] = tot;
      tot += x;
    }
    s_base = tot ? atomicAdd(&cnt->qcount, tot) : 0u;
    if (s_alive) atomicOr((unsigned long long *)&cnt->alive, s_alive);
  }
  __syncthreads();
  uint32_t slot = s_base + s_wpre[wid] + incl - c;
  for (uint32_t m = bits; m; m &= m - 1) {
    const uint32_t i = (uint32_t)(__ffs(m) - 1);
    queue[slot++] = (uint32_t)(first + (uint64_t)i * BLOCK);
  }
  block_add_u64(edges, (unsigned long long *)&cnt->front_edges);
}

// the step's counters to the pinned host mirror, cleared for the next step
__global__ void k_publish(uint64_t *d, uint64_t *h, uint32_t n) {
  const uint32_t i = threadIdx.x;
  if (i < n) {
    h[i] = d[i];
    d[i] = 0;
  }
}

// ---- extraction ----
template <class T>
__global__ void k_gather_col(const T *dist, uint64_t V, uint32_t slot, uint32_t *row) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < V; i += stride) row[i] = wide_dist(dist[i * 128 + slot]);
}

// the block's sums land in out[0], out[1] with one atomic pair per block
// (atomics from every wave on one address serialise at the L2)
__device__ __forceinline__ void block_add2(unsigned long long h, unsigned long long n,
                                           unsigned long long *out) {
  __shared__ unsigned long long s_h[BLOCK / WAVE], s_n[BLOCK / WAVE];
  h = wave_sum(h);
  n = wave_sum(n);
  const uint32_t wid = threadIdx.x / WAVE;
  if (lane_id() == 0) {
    s_h[wid] = h;
    s_n[wid] = n;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t w = 1; w < BLOCK / WAVE; w++) {
      h += s_h[w];
      n += s_n[w];
    }
    atomicAdd(&out[0], h);
    atomicAdd(&out[1], n);
  }
}

__global__ __launch_bounds__(BLOCK) void k_digest(const uint32_t *row, uint64_t V,
                                                  unsigned long long *out) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long h = 0, n = 0;
  for (; i < V; i += stride) {
    uint32_t x = row[i];
    h += digest_term(i, x);
    n += (x != INFINIT);
  }
  block_add2(h, n, out);
}

// digests of the finished BFS lanes (rows of the row-major [lanes][V]
// state): blockIdx.y picks the k-th lane of the mask
__global__ __launch_bounds__(BLOCK) void k_digest_rows(const uint32_t *vals, uint64_t V,
                                                       uint64_t mask, unsigned long long *out) {
  uint64_t m = mask;
  for (uint32_t k = 0; k < blockIdx.y; k++) m &= m - 1;  // drop the lowest set bits
  if (!m) return;  // block-uniform
  const uint32_t lane = (uint32_t)(__ffsll((unsigned long long)m) - 1);
  const uint32_t *row = vals + (uint64_t)lane * V;
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long h = 0, n = 0;
  for (; i < V; i += stride) {
    const uint32_t x = row[i];
    h += digest_term(i, x);
    n += (x != INFINIT);
  }
  block_add2(h, n, out + 2 * lane);
}

// digests of the finished SSSP slots in one sweep of the vertex-major
// [V][128] state: lane p of every wave reads the entries of slots 2p and
// 2p+1 (coalesced rows) and sums the terms of those of its slots in the mask
template <class T>
__global__ __launch_bounds__(BLOCK) void k_digest_cols(const T *vals, uint64_t V, uint64_t lo,
                                                       uint64_t hi, unsigned long long *out) {
  constexpr int NW = BLOCK / WAVE;
  __shared__ unsigned long long s_h[NW][2 * WAVE], s_n[NW][2 * WAVE];
  const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
  const bool mine_a = slot_in(lo, hi, 2 * lane), mine_b = slot_in(lo, hi, 2 * lane + 1);
  const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) / WAVE;
  unsigned long long ha = 0, na = 0, hb = 0, nb = 0;
  if (mine_a || mine_b) {
    for (uint64_t v = wave; v < V; v += nwaves) {
      const uint32_t xa = wide_dist(vals[v * 128 + 2 * lane]);
      const uint32_t xb = wide_dist(vals[v * 128 + 2 * lane + 1]);
      ha += digest_term(v, xa);
      na += (xa != INFINIT);
      hb += digest_term(v, xb);
      nb += (xb != INFINIT);
    }
  }
  s_h[wid][2 * lane] = ha;
  s_n[wid][2 * lane] = na;
  s_h[wid][2 * lane + 1] = hb;
  s_n[wid][2 * lane + 1] = nb;
  __syncthreads();
  if (wid == 0) {
    for (int w = 1; w < NW; w++) {
      ha += s_h[w][2 * lane];
      na += s_n[w][2 * lane];
      hb += s_h[w][2 * lane + 1];
      nb += s_n[w][2 * lane + 1];
    }
    if (mine_a) {
      atomicAdd(&out[4 * lane], ha);
      atomicAdd(&out[4 * lane + 1], na);
    }
    if (mine_b) {
      atomicAdd(&out[4 * lane + 2], hb);
      atomicAdd(&out[4 * lane + 3], nb);
    }
  }
}

}  // namespace ms

void row_digest(Context &ctx, const uint32_t *row, uint64_t V, uint64_t *reached,
                uint64_t *checksum) {
  auto *d = reinterpret_cast<unsigned long long *>(ctx.d_ctl.p) + 48;
  GRUS_HIP(hipMemsetAsync(d, 0, 16, ctx.stream));
  ms::k_digest<<<grid_cap(ctx, V, BLOCK, 4), BLOCK, 0, ctx.stream>>>(row, V, d);
  GRUS_LAUNCH_CHECK();
  ctx.read_ctl(50);
  *checksum = ctx.h_ctl[48];
  *reached = ctx.h_ctl[49];
}

// ------------------------------------------------------------------ engine
MsEngine::MsEngine(Graph &g, MsKind kind, int lanes, bool own_stream)
    : g_(g), kind_(kind), lanes_(std::max(1, std::min(lanes, 64))) {
  slots_ = kind_ == MsKind::BFS ? lanes_ : 2 * lanes_;
  if (own_stream) {
    own_ctx_ = std::make_unique<Context>(g.ctx->device);
    own_ctx_->tuning = g.ctx->tuning;
    own_ctx_->shared_device = true;  // engines of a server run side by side
    ctx_ = own_ctx_.get();
  } else {
    ctx_ = g.ctx;
  }
  init();
}

MsEngine::MsEngine(Graph &g, MsKind kind, int lanes, Context *ctx)
    : g_(g), kind_(kind), lanes_(std::max(1, std::min(lanes, 64))), ctx_(ctx) {
  slots_ = kind_ == MsKind::BFS ? lanes_ : 2 * lanes_;
  init();
}

void MsEngine::init() {
  Context &ctx = *ctx_;
  Graph &g = g_;
  ctx.activate();
  const uint64_t V = g.V;
  if (kind_ == MsKind::BFS) visit_.alloc((V + 1) * 8);
  front_.alloc((V + 1) * 8);
  next_.alloc((V + 1) * 8);
  queue_.alloc((V + 128) * 4);  // duplicates only from restarted slots' sources
  longq_.alloc((V + 1) * 4);
  row_.alloc((V + 1) * 4);
  digest_.alloc(128 * 16);
  srcs_dev_.alloc(1024);
  cnt_.alloc(sizeof(MsCounters));
  push_cnt_.alloc(sizeof(StepCounters));
  hmirror_.alloc(256, Mem::HostPinned);
  if (kind_ == MsKind::BFS) {
    vals_.alloc((uint64_t)lanes_ * V * 4 + 8);  // level[lane][v]
    fill_u64(ctx, visit_.as<uint64_t>(), V + 1, 0);
  } else {
#ifdef SSSP_START_WIDE
    narrow_ = false;
    vals_.alloc(V * 128 * 4 + 8);
#else
    narrow_ = true;
    vals_.alloc(V * 128 * 2 + 8);  // dist[v][slot], 16 bits until a slot needs more
#endif
  }
  fill_u64(ctx, front_.as<uint64_t>(), V + 1, 0);
  fill_u64(ctx, next_.as<uint64_t>(), V + 1, 0);
  GRUS_HIP(hipMemsetAsync(cnt_.p, 0, sizeof(MsCounters), ctx.stream));
  // a queue holds at most V + 128 vertices (restarted slots' sources may
  // repeat), and its degree scan one entry more; the scan's temporaries are
  // the engine's own, so that a step allocates nothing (under a device
  // budget another engine's start may hold the headroom at that moment)
  push_.ensure(V + 128);
  const size_t tb = std::max<size_t>(256, scan_excl_u64_bytes(ctx, V + 129));
  scan_tmp_.alloc(tb);
  push_.tmp = scan_tmp_.p;
  push_.tmp_bytes = tb;
  dirty_ = all_slots();  // state starts uninitialised
  ctx.sync();
}

MsEngine::~MsEngine() {
  if (ctx_) (void)hipStreamSynchronize(ctx_->stream);
}

uint64_t MsEngine::state_bytes() const {
  return visit_.bytes + front_.bytes + next_.bytes + vals_.bytes + queue_.bytes + longq_.bytes +
         row_.bytes + push_.deg.bytes + push_.scan.bytes + scan_tmp_.bytes;
}

uint64_t MsEngine::lanes_of(SlotMask m) const {
  if (kind_ == MsKind::BFS) return (uint64_t)m;
  uint64_t lanes = 0;
  for (; m; m &= m - 1) lanes |= 1ull << (slot_lowest(m) >> 1);
  return lanes;
}

void MsEngine::reset_slots(SlotMask mask) {
  if (!mask) return;
  Context &ctx = *ctx_;
  const uint64_t V = g_.V;
  if (kind_ == MsKind::BFS) {
    const uint64_t lanes = (uint64_t)mask;
    ms::k_reset_masks<<<grid_cap(ctx, V, BLOCK), BLOCK, 0, ctx.stream>>>(
        visit_.as<uint64_t>(), front_.as<uint64_t>(), next_.as<uint64_t>(), V, ~lanes);
    GRUS_LAUNCH_CHECK();
    ms::k_reset_rows<<<grid_cap(ctx, V, BLOCK), BLOCK, 0, ctx.stream>>>(vals_.as<uint32_t>(), V,
                                                                       lanes);
    GRUS_LAUNCH_CHECK();
    return;
  }
  // a lane's frontier bits go only when neither of its slots stays busy
  // (bits of a reset slot's lane are harmless: unreached entries relax
  // nothing)
  const uint64_t quiet = lanes_of(mask) & ~lanes_of(busy_ & ~mask);
  if (quiet) {
    ms::k_reset_masks<<<grid_cap(ctx, V, BLOCK), BLOCK, 0, ctx.stream>>>(
        nullptr, front_.as<uint64_t>(), next_.as<uint64_t>(), V, ~quiet);
    GRUS_LAUNCH_CHECK();
  }
  const uint64_t lo = (uint64_t)mask, hi = (uint64_t)(mask >> 64);
  if (narrow_) {
    ms::k_reset_slots<uint16_t><<<grid_cap(ctx, V * 128, BLOCK), BLOCK, 0, ctx.stream>>>(
        vals_.as<uint16_t>(), V, lo, hi);
  } else {
    ms::k_reset_slots<uint32_t><<<grid_cap(ctx, V * 128, BLOCK), BLOCK, 0, ctx.stream>>>(
        vals_.as<uint32_t>(), V, lo, hi);
  }
  GRUS_LAUNCH_CHECK();
}

// sources into their slots and onto the frontier queue (qn_ entries so far)
void MsEngine::inject(const uint32_t *srcs, const int *slots, int n) {
  Context &ctx = *ctx_;
  auto *cnt = cnt_.as<MsCounters>();
  auto *sd = srcs_dev_.as<uint32_t>();
  for (int k0 = 0; k0 < n; k0 += 64) {  // 64 sources per launch (srcs_dev_ holds 128 words)
    const int m = std::min(64, n - k0);
    uint32_t hs[128];  // sources then slots
    for (int k = 0; k < m; k++) {
      hs[k] = srcs[k0 + k];
      hs[64 + k] = (uint32_t)slots[k0 + k];
      src_[slots[k0 + k]] = srcs[k0 + k];
    }
    MsCounters zero{};
    zero.qcount = (uint32_t)qn_;
    GRUS_HIP(hipMemcpyAsync(cnt, &zero, sizeof zero, hipMemcpyHostToDevice, ctx.stream));
    GRUS_HIP(hipMemcpyAsync(sd, hs, sizeof hs, hipMemcpyHostToDevice, ctx.stream));
    ms::k_inject<<<1, 64, 0, ctx.stream>>>(sd, sd + 64, m, kind_ == MsKind::BFS, narrow_,
                                           visit_.as<uint64_t>(), front_.as<uint64_t>(),
                                           vals_.as<uint32_t>(), g_.V, queue_.as<uint32_t>(), cnt);
    GRUS_LAUNCH_CHECK();
    const MsCounters h = publish();
    qn_ = h.qcount;
  }
}

void MsEngine::widen(SlotMask restart) {
  Context &ctx = *ctx_;
  const uint64_t n = g_.V * 128;
  Buf wide;
  wide.alloc(n * 4 + 8);
  ms::k_widen<<<grid_cap(ctx, n, BLOCK), BLOCK, 0, ctx.stream>>>(vals_.as<uint16_t>(),
                                                                 wide.as<uint32_t>(), n);
  GRUS_LAUNCH_CHECK();
  ctx.sync();
  vals_ = std::move(wide);
  narrow_ = false;
  if (!restart) return;
  // the flagged slots start over from their sources in 32-bit entries
  reset_slots(restart);
  uint32_t srcs[128];
  int slots[128], k = 0;
  for (SlotMask m = restart; m; m &= m - 1) {
    const int sl = slot_lowest(m);
    srcs[k] = src_[sl];
    slots[k++] = sl;
  }
  inject(srcs, slots, k);
}

std::vector<int> MsEngine::admit(const uint32_t *srcs, int n) {
  Context &ctx = *ctx_;
  ctx.activate();
  std::vector<int> used;
  SlotMask freem = all_slots() & ~busy_;
  for (int k = 0; k < n && freem; k++) {
    if (srcs[k] >= g_.V) raise(GRUS_E_INVALID, "source vertex out of range");
    const int sl = slot_lowest(freem);
    freem &= freem - 1;
    used.push_back(sl);
  }
  const int m = (int)used.size();
  if (!m) return used;
  SlotMask mask = 0;
  for (int k = 0; k < m; k++) {
    mask |= slot_bit(used[k]);
    start_[used[k]] = (uint32_t)t_;
  }
  reset_slots(mask & dirty_);
  dirty_ &= ~mask;
  // the frontier queue holds qn_ entries; injected sources are appended
  inject(srcs, used.data(), m);
  busy_ |= mask;
  return used;
}

// read the counters and clear them on the device in one launch (the next
// step starts from zero counters without a memset)
MsCounters MsEngine::publish() {
  Context &ctx = *ctx_;
  static_assert(sizeof(MsCounters) % 8 == 0 && sizeof(MsCounters) <= 128, "counter block");
  ms::k_publish<<<1, 64, 0, ctx.stream>>>(cnt_.as<uint64_t>(), hmirror_.as<uint64_t>(),
                                          (uint32_t)(sizeof(MsCounters) / 8));
  GRUS_LAUNCH_CHECK();
  ctx.sync();
  MsCounters h;
  memcpy(&h, hmirror_.p, sizeof h);
  return h;
}

// one device u64 to the host through the pinned mirror (no pageable staging)
uint64_t MsEngine::read_u64(const uint64_t *d) {
  Context &ctx = *ctx_;
  uint64_t *hm = hmirror_.as<uint64_t>() + 16;
  GRUS_HIP(hipMemcpyAsync(hm, d, 8, hipMemcpyDeviceToHost, ctx.stream));
  ctx.sync();
  return *hm;
}

std::vector<int> MsEngine::step() {
  Context &ctx = *ctx_;
  ctx.activate();
  std::vector<int> done;
  if (!busy_) return done;
  const uint64_t V = g_.V;
  const CsrView out = g_.csr();
  auto *cnt = cnt_.as<MsCounters>();
  // running slots: busy slots whose frontier was not empty (finished slots
  // are reported and leave the running set at once); their lanes
  const SlotMask running = busy_ & ~finished_;
  const uint64_t run_lanes = lanes_of(running);
  // the counters are zero here: cleared by the previous publish()
  if (kind_ == MsKind::BFS) {
    const CsrView in = g_.csc();
    const bool pull = (double)front_edges_ > (double)g_.E / 24.0;
    if (pull) {
      ms::k_bfs_pull<<<grid_cap(ctx, V, BLOCK, 8), BLOCK, 0, ctx.stream>>>(
          in, visit_.as<uint64_t>(), front_.as<uint64_t>(), next_.as<uint64_t>(), run_lanes,
          longq_.as<uint32_t>(), cnt);
      GRUS_LAUNCH_CHECK();
      ms::k_bfs_pull_long<<<grid_cap(ctx, V * WAVE, BLOCK, 8), BLOCK, 0, ctx.stream>>>(
          in, visit_.as<uint64_t>(), front_.as<uint64_t>(), next_.as<uint64_t>(), run_lanes,
          longq_.as<uint32_t>(), cnt);
      GRUS_LAUNCH_CHECK();
    } else if (qn_) {
      push_scan_degrees(ctx, push_, out.off, queue_.as<uint32_t>(), (uint32_t)qn_);
      const uint64_t total = read_u64(push_.scan_ptr() + qn_);
      ms::push_job_t job{front_.as<uint64_t>(), visit_.as<uint64_t>(), next_.as<uint64_t>()};
      push_advance(ctx, push_, out, queue_.as<uint32_t>(), (uint32_t)qn_, total, job,
                   nullptr, push_cnt_.as<StepCounters>());
    }
    ms::LaneStart st;
    for (int i = 0; i < 64; i++) st.s[i] = start_[i];
    // (a pull step left long_count set; finalize uses the other counters,
    // which are still zero)
    ms::k_bfs_finalize<<<grid_cap(ctx, V, BLOCK, 8), BLOCK, 0, ctx.stream>>>(
        out.off, visit_.as<uint64_t>(), front_.as<uint64_t>(), next_.as<uint64_t>(),
        vals_.as<uint32_t>(), V, run_lanes, (uint32_t)t_, st, queue_.as<uint32_t>(), cnt);
    GRUS_LAUNCH_CHECK();
  } else {
    if (qn_) {
      uint64_t *cs = push_.scan_ptr();
      ms::k_chunk_counts<<<grid_cap(ctx, qn_ + 1, BLOCK), BLOCK, 0, ctx.stream>>>(
          out.off, queue_.as<uint32_t>(), (uint32_t)qn_, push_.deg_ptr());
      GRUS_LAUNCH_CHECK();
      scan_excl_u64_with(ctx, push_.deg_ptr(), cs, qn_ + 1, push_.tmp, push_.tmp_bytes);
      // no read-back of the chunk count: at most one chunk per queued vertex
      // plus one per CHUNK of its edges (front_edges_ misses the degrees of
      // sources admitted since the last step, which only sizes the grid)
      const uint64_t est = qn_ + front_edges_ / ms::CHUNK;
      const unsigned grid = grid_cap(ctx, est * WAVE, BLOCK, 8);
      if (narrow_)
        ms::k_sssp_relax<true><<<grid, BLOCK, 0, ctx.stream>>>(
            out, queue_.as<uint32_t>(), (uint32_t)qn_, cs, front_.as<uint64_t>(),
            next_.as<uint64_t>(), vals_.as<uint32_t>(), cnt);
      else
        ms::k_sssp_relax<false><<<grid, BLOCK, 0, ctx.stream>>>(
            out, queue_.as<uint32_t>(), (uint32_t)qn_, cs, front_.as<uint64_t>(),
            next_.as<uint64_t>(), vals_.as<uint32_t>(), cnt);
      GRUS_LAUNCH_CHECK();
    }
    // spans of up to 32 x BLOCK vertices, about 8 workgroups per CU
    const uint64_t cap_v = (uint64_t)grid_cap(ctx, V, BLOCK, 8) * BLOCK;
    const uint32_t items = (uint32_t)std::min<uint64_t>(32, std::max<uint64_t>(1, (V + cap_v - 1) / cap_v));
    const uint64_t span = (uint64_t)items * BLOCK;
    const unsigned fin_grid = (unsigned)std::max<uint64_t>(1, (V + span - 1) / span);
    ms::k_sssp_finalize<<<fin_grid, BLOCK, 0, ctx.stream>>>(
        out.off, front_.as<uint64_t>(), next_.as<uint64_t>(), V, items, run_lanes,
        queue_.as<uint32_t>(), cnt);
    GRUS_LAUNCH_CHECK();
  }
  const MsCounters h = publish();
  t_++;
  qn_ = h.qcount;
  front_edges_ = h.front_edges;
  SlotMask alive, restart = 0;
  if (kind_ == MsKind::BFS) {
    alive = (SlotMask)h.alive;
  } else {
    // slot 2p / 2p+1 improved a distance <=> its frontier is not empty
    alive = 0;
    SlotMask ovf = 0;
    for (int p = 0; p < 64; p++) {
      if ((h.slot_alive[0] >> p) & 1ull) alive |= slot_bit(2 * p);
      if ((h.slot_alive[1] >> p) & 1ull) alive |= slot_bit(2 * p + 1);
      if ((h.overflow[0] >> p) & 1ull) ovf |= slot_bit(2 * p);
      if ((h.overflow[1] >> p) & 1ull) ovf |= slot_bit(2 * p + 1);
    }
    restart = ovf & busy_;
    if (ovf) widen(restart);  // the flagged slots run again, in 32-bit entries
  }
  for (SlotMask m = running & ~alive & ~restart; m; m &= m - 1) {
    const int sl = slot_lowest(m);
    done.push_back(sl);
    finished_ |= slot_bit(sl);
  }
  return done;
}

void MsEngine::extract(int slot, uint32_t *host_out, uint64_t *reached, uint64_t *checksum,
                       uint32_t *device_out) {
  Context &ctx = *ctx_;
  ctx.activate();
  const uint64_t V = g_.V;
  uint32_t *row;
  if (kind_ == MsKind::BFS) {
    row = vals_.as<uint32_t>() + (uint64_t)slot * V;
  } else {
    row = device_out ? device_out : row_.as<uint32_t>();
    if (narrow_)
      ms::k_gather_col<uint16_t><<<grid_cap(ctx, V, BLOCK), BLOCK, 0, ctx.stream>>>(
          vals_.as<uint16_t>(), V, (uint32_t)slot, row);
    else
      ms::k_gather_col<uint32_t><<<grid_cap(ctx, V, BLOCK), BLOCK, 0, ctx.stream>>>(
          vals_.as<uint32_t>(), V, (uint32_t)slot, row);
    GRUS_LAUNCH_CHECK();
  }
  if (device_out && row != device_out)
    GRUS_HIP(hipMemcpyAsync(device_out, row, V * 4, hipMemcpyDeviceToDevice, ctx.stream));
  if (reached || checksum) {
    auto *d = digest_.as<unsigned long long>();
    GRUS_HIP(hipMemsetAsync(d, 0, 16, ctx.stream));
    ms::k_digest<<<grid_cap(ctx, V, BLOCK, 4), BLOCK, 0, ctx.stream>>>(row, V, d);
    GRUS_LAUNCH_CHECK();
    unsigned long long h[2];
    GRUS_HIP(hipMemcpyAsync(h, d, 16, hipMemcpyDeviceToHost, ctx.stream));
    if (host_out)
      GRUS_HIP(hipMemcpyAsync(host_out, row, V * 4, hipMemcpyDeviceToHost, ctx.stream));
    ctx.sync();
    if (checksum) *checksum = h[0];
    if (reached) *reached = h[1];
  } else if (host_out) {
    GRUS_HIP(hipMemcpyAsync(host_out, row, V * 4, hipMemcpyDeviceToHost, ctx.stream));
    ctx.sync();
  }
}

void MsEngine::digest(SlotMask mask, uint64_t *reached, uint64_t *checksum) {
  if (!mask) return;
  Context &ctx = *ctx_;
  ctx.activate();
  const uint64_t V = g_.V;
  auto *d = digest_.as<unsigned long long>();
  GRUS_HIP(hipMemsetAsync(d, 0, 128 * 16, ctx.stream));
  if (kind_ == MsKind::BFS) {
    const uint64_t lanes = (uint64_t)mask;
    dim3 grid(grid_cap(ctx, V, BLOCK, 2), (unsigned)__builtin_popcountll(lanes));
    ms::k_digest_rows<<<grid, BLOCK, 0, ctx.stream>>>(vals_.as<uint32_t>(), V, lanes, d);
  } else if (narrow_) {
    ms::k_digest_cols<uint16_t><<<grid_cap(ctx, V * WAVE, BLOCK), BLOCK, 0, ctx.stream>>>(
        vals_.as<uint16_t>(), V, (uint64_t)mask, (uint64_t)(mask >> 64), d);
  } else {
    ms::k_digest_cols<uint32_t><<<grid_cap(ctx, V * WAVE, BLOCK), BLOCK, 0, ctx.stream>>>(
        vals_.as<uint32_t>(), V, (uint64_t)mask, (uint64_t)(mask >> 64), d);
  }
  GRUS_LAUNCH_CHECK();
  unsigned long long h[256];
  GRUS_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, ctx.stream));
  ctx.sync();
  for (SlotMask m = mask; m; m &= m - 1) {
    const int sl = slot_lowest(m);
    checksum[sl] = h[2 * sl];
    reached[sl] = h[2 * sl + 1];
  }
}

void MsEngine::release(int slot) {
  const SlotMask bit = slot_bit(slot);
  busy_ &= ~bit;
  finished_ &= ~bit;
  dirty_ |= bit;
}

// ------------------------------------------------------------------ batch API
namespace {
void run_multi(Graph &g, MsKind kind, const uint32_t *srcs, int k, uint32_t *out,
               grus_stats *st) {
  const int max_k = kind == MsKind::BFS ? 64 : 128;
  if (k < 1 || k > max_k)
    raise(GRUS_E_INVALID, std::string("between 1 and ") + std::to_string(max_k) +
                              " sources per batch");
  g.make_room(Graph::App::MULTI);
  if (!g.ms || g.ms->kind() != kind) g.ms = std::make_shared<MsEngine>(g, kind, 64, false);
  MsEngine &e = *g.ms;
  Context &ctx = e.ctx();
  for (SlotMask m = e.busy_mask(); m; m &= m - 1) e.release(slot_lowest(m));
  GRUS_HIP(hipEventRecord(ctx.ev0, ctx.stream));
  std::vector<int> slots = e.admit(srcs, k);
  SlotMask pending = 0;
  for (int sl : slots) pending |= slot_bit(sl);
  uint64_t steps = 0;
  while (pending) {
    if (++steps > 2 * g.V + 4) raise(GRUS_E_HIP, "multi-source traversal did not converge");
    for (int sl : e.step()) pending &= ~slot_bit(sl);
  }
  GRUS_HIP(hipEventRecord(ctx.ev1, ctx.stream));
  GRUS_HIP(hipEventSynchronize(ctx.ev1));
  float ms = 0;
  GRUS_HIP(hipEventElapsedTime(&ms, ctx.ev0, ctx.ev1));
  // the reached counts of every slot in one pass over the engine's state;
  // the rows are gathered one by one only when they are copied out
  SlotMask all = 0;
  for (int sl : slots) all |= slot_bit(sl);
  uint64_t reached[128], checksum[128];
  e.digest(all, reached, checksum);
  uint64_t vr = 0;
  for (int sl : slots) vr += reached[sl];
  if (out)
    for (int i = 0; i < k; i++) e.extract(slots[i], out + (uint64_t)i * g.V, nullptr, nullptr);
  for (int sl : slots) e.release(sl);
  if (st) {
    *st = grus_stats{};
    st->device_ms = ms;
    st->iterations = (uint32_t)steps;
    st->vertices_reached = vr;
  }
}
}  // namespace

void run_bfs_multi(Graph &g, const uint32_t *srcs, int k, uint32_t *out, grus_stats *st) {
  run_multi(g, MsKind::BFS, srcs, k, out, st);
}

void run_sssp_multi(Graph &g, const uint32_t *srcs, int k, uint32_t *out, grus_stats *st) {
  run_multi(g, MsKind::SSSP, srcs, k, out, st);
}

}  // namespace mgg
