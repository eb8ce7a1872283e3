// bfs.hip -- breadth-first search levels, direction-optimizing.
//
// Reference: samples/bfs/bfs.cu (BFSInit :18-23, job_t :25-49, updater
// :51-60, generator :61-79, pull_selector :80-87, push driver :135-173,
// pull driver :104-134) on kernel.cuh's push_kernel / pull_kernel.
//
// Semantics kept: level[src] = 0, level[v] = iteration + 1 at first
// discovery, unreached = INFINIT.  Levels are unique, so the result is
// bit-exact whatever the discovery order.
//
// MI355X design:
//   top-down  (push): edge-balanced LB operator (advance.hpp).  The updater
//     claims dst with an atomicOr on a visited bitmap (V/8 bytes, L2
//     resident) instead of the reference's non-atomic label test, so each
//     vertex is written once and enters the next queue exactly once.
//   bottom-up (pull): one wave per 64-vertex bitmap word; a lane owns a
//     vertex that is still unvisited (the reference's pull_selector
//     label==INFINIT, read from the visited bitmap) and scans its in-edges
//     for a parent in the current frontier bitmap, stopping at the first
//     hit (the reference scans every in-edge, kernel.cuh:31-37).  After
//     `bu_thread_edges` edges a vertex is handed to a workgroup-cooperative
//     scan (its waves 64 edges per step each, ballot early exit).  The wave ballots its found
//     lanes and writes the next frontier word and visited word with plain
//     stores -- no atomics -- and its level stores are coalesced.
//   direction switch (Beamer et al., SC'12): push->pull when the
//     frontier's out-edges m_f exceed m_u/alpha (m_u = edges of unvisited
//     vertices), pull->push when the frontier has shrunk below V/beta.
//     Replaces Frontier<BDF_AUTO>'s fraction thresholds
//     (include/frontier.cuh:140,206-215).
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>

#include "advance.hpp"
#include "apps.hpp"

namespace mgg {

struct BfsWork {
  Buf visited, front[2], q[2], longq;
  Buf stamp;            // large top-down steps: byte per vertex, = the step's stamp when found
  uint8_t last_stamp = 0;
  Buf cand;  // sparse bottom-up step: the unvisited vertices
  Buf first_in;                     // first in-neighbour of every vertex (GRUS_NO_VERTEX if none)
  const void *first_for = nullptr;  // in-edge columns first_in was built from
  PushWork push;
  uint64_t cap_v = 0;
  // two-stage steps: two staging buffers of frontier rows, the stream that
  // fills one while the other is expanded, and the events that order them
  Buf stage[2];
  hipStream_t copy_stream = nullptr;
  hipEvent_t ev_ready[2] = {nullptr, nullptr}, ev_free[2] = {nullptr, nullptr}, ev_scan = nullptr;
  ~BfsWork() {
    if (copy_stream) (void)hipStreamSynchronize(copy_stream);
    for (hipEvent_t e : {ev_ready[0], ev_ready[1], ev_free[0], ev_free[1], ev_scan})
      if (e) (void)hipEventDestroy(e);
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
  }
  void ensure_stage(uint64_t edges) {
    if (!copy_stream) {
      GRUS_HIP(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking));
      for (int b = 0; b < 2; b++) {
        GRUS_HIP(hipEventCreateWithFlags(&ev_ready[b], hipEventDisableTiming));
        GRUS_HIP(hipEventCreateWithFlags(&ev_free[b], hipEventDisableTiming));
      }
      GRUS_HIP(hipEventCreateWithFlags(&ev_scan, hipEventDisableTiming));
    }
    for (auto &b : stage) b.ensure(std::max<uint64_t>(edges, 1) * 4);
  }
  void ensure(Context &ctx, uint64_t V) {
    if (V <= cap_v) return;
    uint64_t W = (V + 63) / 64;
    visited.alloc((W + 1) * 8);
    stamp.alloc(W * 64 + 64);
    // cleared on the context's stream, ahead of the kernels that read it
    // (a null-stream memset is not ordered with that non-blocking stream)
    fill_u32(ctx, stamp.as<uint32_t>(), W * 16 + 16, 0);
    last_stamp = 0;
    front[0].alloc((W + 1) * 8);
    front[1].alloc((W + 1) * 8);
    q[0].alloc((V + 1) * 4);
    q[1].alloc((V + 1) * 4);
    longq.alloc((V + 1) * 4);
    cand.alloc((V + 1) * 4);
    push.ensure(V);
    cap_v = V;
  }
};

namespace bfs {

// job_t of samples/bfs/bfs.cu:25-49 as the device-side job of one level
struct job_t {
  static constexpr bool kFar = false;
  uint32_t *label;
  uint64_t *visited;
  const uint64_t *out_off;
  uint32_t next_level;  // job.itr + 1
  // updater (bfs.cu:51-60): first discovery sets the level; the bitmap
  // claim makes the check-then-store race-free
  __device__ __forceinline__ bool update(vtx_t, vtx_t dst, eid_t, weight_t, uint32_t *) const {
    if (bm_test(visited, dst)) return false;
    if (!bm_claim(visited, dst)) return false;
    label[dst] = next_level;
    return true;
  }
  // generator (bfs.cu:67-71): a claimed vertex joins the next frontier once
  __device__ __forceinline__ int generate(bool updated, vtx_t, uint32_t) const {
    return updated ? 1 : 0;
  }
  __device__ __forceinline__ uint64_t out_degree(vtx_t v) const {
    return out_off[v + 1] - out_off[v];
  }
};

// Updater of a large top-down step (many edges): no atomics.  An unvisited
// destination gets its level and the step's stamp with plain stores; every
// writer stores the same values, so concurrent writes agree.  Nothing is
// appended: k_stamp_commit turns the stamps into the visited / frontier
// bits and counts (one pass over V bytes instead of an atomic per claim).
struct stamp_job_t {
  static constexpr bool kFar = false;
  uint32_t *label;
  const uint64_t *visited;
  uint8_t *stamp;
  uint32_t next_level;
  uint8_t step_stamp;
  __device__ __forceinline__ bool update(vtx_t, vtx_t dst, eid_t, weight_t, uint32_t *) const {
    if (bm_test(visited, dst)) return false;
    label[dst] = next_level;
    stamp[dst] = step_stamp;
    return false;
  }
  __device__ __forceinline__ int generate(bool, vtx_t, uint32_t) const { return 0; }
  __device__ __forceinline__ uint64_t out_degree(vtx_t) const { return 0; }
};

// after a stamped step: word w's new vertices = its bytes equal to the
// stamp; they join visited and the next frontier bitmap (every word
// written), and the counters get their number and out-degree sum
__global__ __launch_bounds__(BLOCK) void k_stamp_commit(const uint8_t *stamp, uint8_t step_stamp,
                                                        const uint64_t *out_off, uint64_t *visited,
                                                        uint64_t *front_next, uint64_t W,
                                                        uint64_t V, StepCounters *cnt) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint32_t found = 0;
  uint64_t edges = 0;
  for (; i < W; i += stride) {
    const uint64_t *b = reinterpret_cast<const uint64_t *>(stamp + i * 64);
    const uint64_t rep = 0x0101010101010101ull * step_stamp;
    uint64_t m = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      // 8 bytes at a time: y has a zero byte where the stamp matches; the
      // high bit of each byte of nz is set exactly where y's byte is not
      // zero (no carry crosses a byte); the multiply gathers those 8 bits
      const uint64_t y = b[k] ^ rep;
      const uint64_t nz = (((y & 0x7F7F7F7F7F7F7F7Full) + 0x7F7F7F7F7F7F7F7Full) | y) &
                          0x8080808080808080ull;
      const uint64_t z = (nz ^ 0x8080808080808080ull) >> 7;  // bit 8j set: byte j matches
      m |= ((z * 0x0102040810204080ull) >> 56) << (8 * k);
    }
    front_next[i] = m;
    if (m) {
      visited[i] |= m;
      found += (uint32_t)__popcll(m);
      for (uint64_t r = m; r; r &= r - 1) {
        const uint64_t v = i * 64 + (uint64_t)(__ffsll((unsigned long long)r) - 1);
        if (v < V) edges += out_off[v + 1] - out_off[v];
      }
    }
  }
  block_add_counts(found, edges, &cnt->next_count, (unsigned long long *)&cnt->next_edges);
}

__global__ void k_init(uint32_t *label, uint64_t V, uint64_t *visited, const uint64_t *unreachable,
                       uint64_t W, uint32_t src, uint32_t *q0) {
  const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  // levels (BFSInit bfs.cu:21): INFINIT, 0 at the source, four per 16-byte
  // store (the array is 256-byte aligned), the last V % 4 one by one
  const uint64_t V4 = V / 4;
  uint4 *l4 = reinterpret_cast<uint4 *>(label);
  for (uint64_t i = t; i < V4; i += stride) {
    uint4 x = make_uint4(INFINIT, INFINIT, INFINIT, INFINIT);
    if (i == (uint64_t)(src >> 2)) {
      const uint32_t c = src & 3;
      x.x = c == 0 ? 0u : x.x;
      x.y = c == 1 ? 0u : x.y;
      x.z = c == 2 ? 0u : x.z;
      x.w = c == 3 ? 0u : x.w;
    }
    l4[i] = x;
  }
  for (uint64_t i = V4 * 4 + t; i < V; i += stride) label[i] = i == src ? 0u : INFINIT;
  for (uint64_t i = t; i < W; i += stride) {
    uint64_t m = unreachable ? unreachable[i] : 0ull;
    if (i == (src >> 6)) m |= 1ull << (src & 63);
    visited[i] = m;
  }
  if (t == 0) q0[0] = src;
}

// first[v] = lowest-id in-neighbour of v (rows are sorted), GRUS_NO_VERTEX
// for an empty row
__global__ void k_first_in(CsrView in, uint32_t *first) {
  uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; v < in.V; v += stride)
    first[v] = in.off[v] < in.off[v + 1] ? edge_col(in, in.off[v]) : GRUS_NO_VERTEX;
}

// bottom-up step over every bitmap word (see file comment).  A wave takes
// PULL_UNROLL consecutive words (512 vertices) per step and issues their
// visited words, first in-neighbours and row offsets together -- all three
// coalesced and independent of each other -- then the frontier bits of the
// first in-neighbours, so the first test of a vertex is two memory
// latencies, the second an L2-resident bitmap word (reading the first
// neighbour through the offsets added a dependent, uncoalesced column
// load).  Most RMAT vertices find a parent at their first (lowest-id)
// in-neighbour.
constexpr int PULL_UNROLL = 8;
constexpr int PULL_BATCH = 4;  // in-edges tested per dependent load round
__global__ __launch_bounds__(BLOCK) void k_pull(CsrView in, const uint64_t *out_off,
                                                const uint32_t *first_in, const uint64_t *front,
                                                uint64_t *visited, uint64_t *front_next,
                                                uint32_t *label, uint32_t next_level, uint32_t K,
                                                uint32_t *longq, StepCounters *cnt, uint64_t W) {
  uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE;
  uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) / WAVE;
  uint32_t lane = lane_id();
  uint32_t my_found = 0;
  uint64_t my_edges = 0;
  for (uint64_t w0 = wave * PULL_UNROLL; w0 < W; w0 += nwaves * PULL_UNROLL) {
    uint64_t vis[PULL_UNROLL], beg[PULL_UNROLL], end[PULL_UNROLL];
    bool active[PULL_UNROLL], found[PULL_UNROLL];
    uint32_t first[PULL_UNROLL];
    // visited words, first in-neighbours and offsets: independent loads
#pragma unroll
    for (int j = 0; j < PULL_UNROLL; j++) {
      const uint64_t v = (w0 + j) * 64 + lane;
      const bool in_range = v < in.V;
      vis[j] = (w0 + j < W) ? visited[w0 + j] : ~0ull;
      first[j] = in_range ? first_in[v] : GRUS_NO_VERTEX;
      beg[j] = in_range ? in.off[v] : 0;
      end[j] = in_range ? in.off[v + 1] : 0;
    }
#pragma unroll
    for (int j = 0; j < PULL_UNROLL; j++) {
      // pull_selector (bfs.cu:80-87): label == INFINIT, read as a clear
      // visited bit (bits past V are set at init)
      active[j] = ((vis[j] >> lane) & 1ull) == 0;
      found[j] = active[j] && first[j] != GRUS_NO_VERTEX && bm_test(front, first[j]);
    }
#pragma unroll
    for (int j = 0; j < PULL_UNROLL; j++) {
      if (vis[j] == ~0ull) {
        if (lane == 0 && w0 + j < W) front_next[w0 + j] = 0;
        continue;
      }
      uint32_t v = (uint32_t)((w0 + j) * 64 + lane);
      bool is_long = false;
      if (active[j] && !found[j] && beg[j] < end[j]) {
        uint64_t e = beg[j] + 1;
        uint64_t lim = (end[j] - beg[j] > K) ? beg[j] + K : end[j];
        // PULL_BATCH columns, then their PULL_BATCH frontier words, are
        // loaded independently: one column -> bitmap latency chain per batch
        // instead of per edge (a batch may read past the parent it finds)
        for (; e < lim && !found[j]; e += PULL_BATCH) {
          uint32_t c[PULL_BATCH];
#pragma unroll
          for (int q = 0; q < PULL_BATCH; q++) c[q] = (e + q < lim) ? edge_col(in, e + q) : 0u;
          uint64_t fw[PULL_BATCH];
#pragma unroll
          for (int q = 0; q < PULL_BATCH; q++) fw[q] = (e + q < lim) ? front[c[q] >> 6] : 0ull;
#pragma unroll
          for (int q = 0; q < PULL_BATCH; q++) found[j] = found[j] || ((fw[q] >> (c[q] & 63)) & 1ull);
        }
        is_long = !found[j] && lim < end[j];
      }
      uint64_t m = __ballot(found[j]);
      if (lane == 0) {
        front_next[w0 + j] = m;
        if (m) visited[w0 + j] = vis[j] | m;
      }
      if (found[j]) {
        label[v] = next_level;
        my_found++;
        // out-degree: the in-edge row itself on a symmetric graph
        my_edges += (out_off == in.off) ? end[j] - beg[j] : out_off[v + 1] - out_off[v];
      }
      uint32_t slot = wave_append(&cnt->long_count, is_long);
      if (is_long) longq[slot] = v;
    }
  }
  block_add_counts(my_found, my_edges, &cnt->next_count,
                   (unsigned long long *)&cnt->next_edges);
}

// Bottom-up step over a candidate list instead of the whole bitmap, for
// late steps with few unvisited vertices left: the list is every clear bit
// of the visited bitmap (in-degree-0 vertices and the bits past V are set
// at init), compacted just before, its length *ncp on the device.  A lane
// per candidate, the same first test and batched scan as k_pull; found
// vertices set their visited and next-frontier bits with atomics
// (candidates of one word are spread over lanes), front_next must be zero
// on entry.  (Reading every vertex's first neighbour and offsets, as the
// dense step does, costs ~20 bytes per vertex whatever is left.)
constexpr int SPARSE_UNROLL = 4;
// scan the list when it holds under V / 16 vertices (RMAT s24, pinned
// thresholds: V/4 507, V/8 531, V/16 534, V/32 518 GTEPS)
constexpr uint64_t SPARSE_DIV = 16;
// a top-down step over more than V / STAMP_DIV edges claims with stamps
// (RMAT s24, pinned thresholds: V/16 580, V/32 584, V/64 583, V/128 582
// GTEPS; 551 without stamps)
constexpr uint64_t STAMP_DIV = 64;
// Several small top-down levels in one launch of one workgroup: while the
// frontier has at most SL_F vertices and SL_E out-edges (and the direction
// rule keeps pushing), the workgroup expands a level, claims the new
// vertices in the visited bitmap, writes their levels, appends them to the
// other queue and goes on, with workgroup barriers between levels instead
// of a return to the host.  The first and last levels of most traversals
// (a source of low degree, the tail of the reached component) are such
// levels; each costs a host round trip of ~20 us when launched alone.
constexpr uint32_t SL_T = 1024;                  // threads of the workgroup
constexpr uint32_t SL_PER = 4;                   // edges per thread per level
constexpr uint32_t SL_F = SL_T;                  // frontier vertices staged in LDS
constexpr uint64_t SL_E = (uint64_t)SL_T * SL_PER;
constexpr uint32_t SL_LEVELS = 64;               // levels per launch at most

// what the launch leaves behind (in the control block, after the counters)
struct SmallLevels {
  uint64_t mf, mu, v_add, e_add;  // frontier edges, unexplored edges, reached sums
  uint32_t nf, levels, cur, prev_nf;
};

__global__ __launch_bounds__(SL_T) void k_small_levels(CsrView g, const uint64_t *out_off,
                                                       uint32_t *q0, uint32_t *q1, int cur0,
                                                       uint32_t nf0, uint64_t mf0, uint64_t mu0,
                                                       double alpha, uint32_t *label,
                                                       uint64_t *visited, uint32_t level0,
                                                       SmallLevels *out) {
  __shared__ uint64_t sS[SL_F + 1];
  __shared__ uint64_t sRow[SL_F];
  __shared__ uint64_t s_wsum[SL_T / WAVE];
  __shared__ uint32_t s_nn;
  __shared__ unsigned long long s_mn;
  uint32_t *q[2] = {q0, q1};
  const uint32_t t = threadIdx.x, lane = lane_id(), wid = t / WAVE;
  int cur = cur0;
  uint32_t nf = nf0, prev = 0, lv = 0;
  uint64_t mf = mf0, mu = mu0, vadd = 0, eadd = 0;
  // every thread holds the same nf / mf / mu / lv, so the loop is uniform.
  // The host chose push for the first level (its pull -> push switch does
  // not test alpha); later levels follow the host's push -> pull rule
  while (nf > 0 && lv < SL_LEVELS && nf <= SL_F && mf <= SL_E &&
         !(lv > 0 && alpha > 0 && (double)mf > (double)mu / alpha)) {
    // stage the frontier (thread t owns entry t) and scan its degrees
    uint64_t d = 0;
    if (t < nf) {
      const uint32_t u = q[cur][t];
      const uint64_t a = g.off[u];
      sRow[t] = a;
      d = g.off[u + 1] - a;
    }
    uint64_t x = d;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
      const uint64_t y = __shfl_up(x, o);
      if ((int)lane >= o) x += y;
    }
    if (lane == WAVE - 1) s_wsum[wid] = x;
    __syncthreads();
    uint64_t before = 0, sum = 0;
    for (uint32_t w2 = 0; w2 < SL_T / WAVE; w2++) {
      before += (w2 < wid) ? s_wsum[w2] : 0ull;
      sum += s_wsum[w2];
    }
    if (t < nf) sS[t] = before + x - d;
    if (t == 0) {
      sS[nf] = sum;
      s_nn = 0;
      s_mn = 0;
    }
    __syncthreads();
    // the level's edges, SL_PER per thread: columns first, then the claims
    uint32_t dst[SL_PER];
#pragma unroll
    for (uint32_t r = 0; r < SL_PER; r++) {
      const uint64_t k = t + (uint64_t)r * SL_T;
      dst[r] = GRUS_NO_VERTEX;
      if (k < sum) {
        const uint32_t j = upper_bound_u64(sS, 0, nf + 1, k) - 1;
        dst[r] = edge_col(g, sRow[j] + (k - sS[j]));
      }
    }
    uint64_t my_e = 0;
#pragma unroll
    for (uint32_t r = 0; r < SL_PER; r++) {
      const uint32_t v = dst[r];
      if (v == GRUS_NO_VERTEX || bm_test(visited, v) || !bm_claim(visited, v)) continue;
      label[v] = level0 + lv + 1;
      q[cur ^ 1][atomicAdd(&s_nn, 1u)] = v;
      my_e += out_off[v + 1] - out_off[v];
    }
    if (my_e) atomicAdd(&s_mn, (unsigned long long)my_e);
    __syncthreads();
    // (s_nn / s_mn are reset only after the next level's first barrier)
    prev = nf;
    nf = s_nn;
    mf = s_mn;
    mu = mu >= mf ? mu - mf : 0;
    vadd += nf;
    eadd += mf;
    cur ^= 1;
    lv++;
  }
  if (t == 0) {
    out->mf = mf;
    out->mu = mu;
    out->v_add = vadd;
    out->e_add = eadd;
    out->nf = nf;
    out->levels = lv;
    out->cur = (uint32_t)cur;
    out->prev_nf = prev;
  }
}

__global__ __launch_bounds__(BLOCK) void k_pull_sparse(
    CsrView in, const uint64_t *out_off, const uint32_t *first_in, const uint64_t *front,
    uint64_t *visited, uint64_t *front_next, uint32_t *label, uint32_t next_level, uint32_t K,
    const uint32_t *cand, const uint32_t *ncp, uint32_t *longq, StepCounters *cnt) {
  uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE;
  uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) / WAVE;
  uint32_t lane = lane_id();
  const uint32_t nc = *ncp;
  uint32_t my_found = 0;
  uint64_t my_edges = 0;
  for (uint64_t c0 = wave * SPARSE_UNROLL * WAVE; c0 < nc; c0 += nwaves * SPARSE_UNROLL * WAVE) {
    uint32_t v[SPARSE_UNROLL], first[SPARSE_UNROLL];
    uint64_t beg[SPARSE_UNROLL], end[SPARSE_UNROLL];
    bool valid[SPARSE_UNROLL], found[SPARSE_UNROLL];
#pragma unroll
    for (int j = 0; j < SPARSE_UNROLL; j++) {
      const uint64_t i = c0 + (uint64_t)j * WAVE + lane;
      valid[j] = i < nc;
      v[j] = valid[j] ? cand[i] : 0u;
    }
#pragma unroll
    for (int j = 0; j < SPARSE_UNROLL; j++) {
      first[j] = valid[j] ? first_in[v[j]] : GRUS_NO_VERTEX;
      beg[j] = valid[j] ? in.off[v[j]] : 0;
      end[j] = valid[j] ? in.off[v[j] + 1] : 0;
    }
#pragma unroll
    for (int j = 0; j < SPARSE_UNROLL; j++)
      found[j] = first[j] != GRUS_NO_VERTEX && bm_test(front, first[j]);
#pragma unroll
    for (int j = 0; j < SPARSE_UNROLL; j++) {
      bool is_long = false;
      if (valid[j] && !found[j] && beg[j] < end[j]) {
        uint64_t e = beg[j] + 1;
        uint64_t lim = (end[j] - beg[j] > K) ? beg[j] + K : end[j];
        for (; e < lim && !found[j]; e += PULL_BATCH) {
          uint32_t c[PULL_BATCH];
#pragma unroll
          for (int q = 0; q < PULL_BATCH; q++) c[q] = (e + q < lim) ? edge_col(in, e + q) : 0u;
          uint64_t fw[PULL_BATCH];
#pragma unroll
          for (int q = 0; q < PULL_BATCH; q++) fw[q] = (e + q < lim) ? front[c[q] >> 6] : 0ull;
#pragma unroll
          for (int q = 0; q < PULL_BATCH; q++) found[j] = found[j] || ((fw[q] >> (c[q] & 63)) & 1ull);
        }
        is_long = !found[j] && lim < end[j];
      }
      if (found[j]) {
        const uint32_t u = v[j];
        label[u] = next_level;
        const unsigned long long bit = 1ull << (u & 63);
        atomicOr((unsigned long long *)&front_next[u >> 6], bit);
        atomicOr((unsigned long long *)&visited[u >> 6], bit);
        my_found++;
        my_edges += (out_off == in.off) ? end[j] - beg[j] : out_off[u + 1] - out_off[u];
      }
      const uint32_t slot = wave_append(&cnt->long_count, is_long);
      if (is_long) longq[slot] = v[j];
    }
  }
  block_add_counts(my_found, my_edges, &cnt->next_count,
                   (unsigned long long *)&cnt->next_edges);
}

// workgroup-cooperative scan of the vertices whose first K in-edges held no parent
__global__ __launch_bounds__(BLOCK) void k_pull_long(CsrView in, const uint64_t *out_off,
                                                     const uint64_t *front, uint64_t *visited,
                                                     uint64_t *front_next, uint32_t *label,
                                                     uint32_t next_level, uint32_t K,
                                                     const uint32_t *longq, StepCounters *cnt) {
  // a workgroup per long row, its waves striding over the row and stopping
  // as soon as one of them sees a parent (an LDS flag): a row whose parent
  // lies deep in it no longer holds one wave while the rest have finished
  __shared__ uint32_t s_found;
  const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
  const uint32_t n = cnt->long_count;
  uint32_t my_found = 0;
  uint64_t my_edges = 0;
  for (uint64_t i = blockIdx.x; i < n; i += gridDim.x) {  // block-uniform
    if (threadIdx.x == 0) s_found = 0;
    __syncthreads();
    const uint32_t v = longq[i];
    const uint64_t end = in.off[v + 1];
    for (uint64_t base = in.off[v] + K + (uint64_t)wid * WAVE; base < end;
         base += (uint64_t)nw * WAVE) {
      if (__hip_atomic_load(&s_found, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
      const uint64_t e = base + lane;
      const bool hit = e < end && bm_test(front, edge_col(in, e));
      if (__ballot(hit)) {
        if (lane == 0) __hip_atomic_store(&s_found, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        break;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_found) {
      label[v] = next_level;
      unsigned long long bit = 1ull << (v & 63);
      atomicOr((unsigned long long *)&front_next[v >> 6], bit);
      atomicOr((unsigned long long *)&visited[v >> 6], bit);
      my_found++;
      my_edges += out_off[v + 1] - out_off[v];
    }
    __syncthreads();  // s_found read before the next row resets it
  }
  if (threadIdx.x == 0 && my_found) {
    atomicAdd(&cnt->next_count, my_found);
    atomicAdd((unsigned long long *)&cnt->next_edges, my_edges);
  }
}

}  // namespace bfs

void run_bfs(Graph &g, uint32_t src, int dir, grus_stats *st) {
  Context &ctx = *g.ctx;
  ctx.activate();
  if (src >= g.V) raise(GRUS_E_INVALID, "source vertex out of range");
  if (dir < GRUS_DIR_AUTO || dir > GRUS_DIR_PULL) raise(GRUS_E_INVALID, "direction mode");
  g.make_room(Graph::App::BFS);
  if (!g.bfs) g.bfs = std::make_shared<BfsWork>();
  BfsWork &w = *g.bfs;
  w.ensure(ctx, g.V);
  g.result.ensure(g.V * 4);
  const uint64_t V = g.V, W = g.words();
  const CsrView out = g.csr();
  const CsrView in = (dir == GRUS_DIR_PUSH) ? out : g.csc();
  // push-only runs need no in-degrees (and so no CSC of a directed graph)
  const uint64_t *unreach = (dir == GRUS_DIR_PUSH) ? nullptr : g.unreachable();
  // first in-neighbours for the bottom-up first test: the capped placement
  // keeps its own resident copy; otherwise built once per in-edge array
  const uint32_t *first_in = nullptr;
  if (dir != GRUS_DIR_PUSH) {
    if (g.first_col.p) {
      first_in = g.first_col.as<uint32_t>();
    } else {
      if (w.first_for != in.col || w.first_in.bytes < g.V * 4) {
        w.first_in.alloc(std::max<uint64_t>(g.V, 1) * 4);
        bfs::k_first_in<<<grid_cap(ctx, g.V, BLOCK), BLOCK, 0, ctx.stream>>>(
            in, w.first_in.as<uint32_t>());
        GRUS_LAUNCH_CHECK();
        w.first_for = in.col;
      }
      first_in = w.first_in.as<uint32_t>();
    }
  }
  uint32_t *label = g.result.as<uint32_t>();
  auto *cnt = reinterpret_cast<StepCounters *>(ctx.d_ctl.p);
  const double alpha = ctx.tuning.bfs_alpha > 0 ? ctx.tuning.bfs_alpha : 14.0;
  const double beta = ctx.tuning.bfs_beta > 0 ? ctx.tuning.bfs_beta : 64.0;
  const uint32_t K = ctx.tuning.bu_thread_edges ? ctx.tuning.bu_thread_edges : 32;

  uint64_t src_deg = 0;
  {
    uint64_t o[2];
    GRUS_HIP(hipMemcpy(o, g.off.as<uint64_t>() + src, 16, hipMemcpyDeviceToHost));
    src_deg = o[1] - o[0];
  }
  uint64_t launches = 0;
  GRUS_HIP(hipMemsetAsync(cnt, 0, sizeof(StepCounters), ctx.stream));
  GRUS_HIP(hipEventRecord(ctx.ev0, ctx.stream));
  bfs::k_init<<<grid_cap(ctx, std::max(V, W), BLOCK), BLOCK, 0, ctx.stream>>>(
      label, V, w.visited.as<uint64_t>(), unreach, W, src, w.q[0].as<uint32_t>());
  GRUS_LAUNCH_CHECK();
  launches++;

  int cur = 0;               // index of the current queue / bitmap
  bool have_queue = true;    // current frontier available as a queue
  bool have_bitmap = false;  // ... as a bitmap
  bool pull = (dir == GRUS_DIR_PULL);
  uint64_t nf = 1, mf = src_deg, prev_nf = 0;
  uint64_t mu = g.E >= src_deg ? g.E - src_deg : 0;
  uint64_t v_reached = 1, e_reached = src_deg;
  uint32_t level = 0, pull_levels = 0;
  // a bottom-up step with few unvisited vertices left scans a list of them
  // instead of every bitmap word; the host knows how many are left from
  // the vertices reached so far and the in-degree-0 count
  const uint64_t n_unreach = unreach ? g.unreachable_count : 0;
  // GRUS_TRACE=1 prints one line per level (mode, frontier, host time)
  static const bool trace = getenv("GRUS_TRACE") != nullptr;
  auto t_level = std::chrono::steady_clock::now();
  while (nf) {
    // every level discovers at least one new vertex, so V levels is a bound
    if (level > V) raise(GRUS_E_HIP, "bfs did not converge within V levels");
    if (dir == GRUS_DIR_AUTO) {
      if (!pull && (double)mf > (double)mu / alpha) pull = true;
      else if (pull && (double)nf < (double)V / beta && nf < prev_nf) pull = false;
    }
    bool sparse = false;  // (the counters are zero: cleared before the loop, then by publish_ctl)
    bool stamped = false;
    if (pull) {
      if (!have_bitmap) {
        queue_to_bitmap(ctx, w.q[cur].as<uint32_t>(), (uint32_t)nf, w.front[cur].as<uint64_t>(), W);
        launches += 2;
      }
      // unvisited vertices left (an upper bound: the source may be one of
      // the in-degree-0 vertices)
      const uint64_t left = V + 1 > v_reached + n_unreach ? V + 1 - v_reached - n_unreach : 0;
      sparse = left * bfs::SPARSE_DIV < V;
      if (sparse) {
        fill_u64(ctx, w.front[cur ^ 1].as<uint64_t>(), W, 0ull);
        clear_bits_to_queue(ctx, w.visited.as<uint64_t>(), W, w.cand.as<uint32_t>(),
                            &cnt->rem_count);
        const uint64_t waves = div_up(std::max<uint64_t>(left, 1), (uint64_t)bfs::SPARSE_UNROLL * WAVE);
        bfs::k_pull_sparse<<<grid_cap(ctx, waves * WAVE, BLOCK, 8), BLOCK, 0, ctx.stream>>>(
            in, out.off, first_in, w.front[cur].as<uint64_t>(), w.visited.as<uint64_t>(),
            w.front[cur ^ 1].as<uint64_t>(), label, level + 1, K, w.cand.as<uint32_t>(),
            &cnt->rem_count, w.longq.as<uint32_t>(), cnt);
        GRUS_LAUNCH_CHECK();
        launches += 3;
      } else {
        bfs::k_pull<<<grid_cap(ctx, div_up(W, bfs::PULL_UNROLL) * WAVE, BLOCK, 8), BLOCK, 0,
                      ctx.stream>>>(in, out.off, first_in, w.front[cur].as<uint64_t>(),
                                    w.visited.as<uint64_t>(), w.front[cur ^ 1].as<uint64_t>(),
                                    label, level + 1, K, w.longq.as<uint32_t>(), cnt, W);
        GRUS_LAUNCH_CHECK();
        launches++;
      }
      bfs::k_pull_long<<<grid_cap(ctx, V * WAVE, BLOCK, 8), BLOCK, 0, ctx.stream>>>(
          in, out.off, w.front[cur].as<uint64_t>(), w.visited.as<uint64_t>(),
          w.front[cur ^ 1].as<uint64_t>(), label, level + 1, K, w.longq.as<uint32_t>(), cnt);
      GRUS_LAUNCH_CHECK();
      launches++;
      have_bitmap = true;
      have_queue = false;
      pull_levels++;
    } else {
      if (!have_queue) {
        bitmap_to_queue(ctx, w.front[cur].as<uint64_t>(), W, w.q[cur].as<uint32_t>(),
                        &cnt->long_count);  // scratch counter, re-zeroed below
        GRUS_HIP(hipMemsetAsync(cnt, 0, sizeof(StepCounters), ctx.stream));
        launches += 1;
      }
      if (nf <= bfs::SL_F && mf <= bfs::SL_E) {
        // small levels: one workgroup runs them back to back on the device
        auto *sl = reinterpret_cast<bfs::SmallLevels *>(ctx.d_ctl.as<uint64_t>() + 8);
        bfs::k_small_levels<<<1, bfs::SL_T, 0, ctx.stream>>>(
            out, out.off, w.q[0].as<uint32_t>(), w.q[1].as<uint32_t>(), cur, (uint32_t)nf, mf,
            mu, dir == GRUS_DIR_AUTO ? alpha : 0.0, label, w.visited.as<uint64_t>(), level, sl);
        GRUS_LAUNCH_CHECK();
        launches++;
        ctx.publish_ctl(8 + sizeof(bfs::SmallLevels) / 8);
        bfs::SmallLevels r;
        std::memcpy(&r, ctx.h_ctl + 8, sizeof r);
        if (r.nf > V || r.levels == 0 || r.cur > 1)
          raise(GRUS_E_HIP, "bfs small levels reported an impossible state");
        if (trace) {
          auto now = std::chrono::steady_clock::now();
          fprintf(stderr, "bfs levels %u-%u in one launch: nf=%llu -> %u %.1f us\n", level,
                  level + r.levels - 1, (unsigned long long)nf, r.nf,
                  std::chrono::duration<double, std::micro>(now - t_level).count());
          t_level = now;
        }
        level += r.levels;
        cur = (int)r.cur;
        prev_nf = r.prev_nf;
        nf = r.nf;
        mf = r.mf;
        mu = r.mu;
        v_reached += r.v_add;
        e_reached += r.e_add;
        have_queue = true;  // the frontier is queue cur; no bitmap of it
        have_bitmap = false;
        continue;
      }
      bfs::job_t job{label, w.visited.as<uint64_t>(), out.off, level + 1};
      if (mf * bfs::STAMP_DIV > V) {
        // many edges: plain stores and one commit pass instead of an atomic
        // claim per discovered vertex; leaves the frontier as a bitmap
        if (++w.last_stamp == 0) {  // stamps wrapped: clear and restart at 1
          GRUS_HIP(hipMemsetAsync(w.stamp.p, 0, W * 64, ctx.stream));
          w.last_stamp = 1;
        }
        bfs::stamp_job_t sj{label, w.visited.as<uint64_t>(), w.stamp.as<uint8_t>(), level + 1,
                            w.last_stamp};
        if (nf <= PUSH_SMALL_F) {
          push_advance_small(ctx, out, w.q[cur].as<uint32_t>(), (uint32_t)nf, mf, sj, nullptr,
                             cnt);
          launches += 1;
        } else {
          push_scan_degrees(ctx, w.push, out.off, w.q[cur].as<uint32_t>(), (uint32_t)nf);
          push_advance(ctx, w.push, out, w.q[cur].as<uint32_t>(), (uint32_t)nf, mf, sj, nullptr,
                       cnt);
          launches += 3;
        }
        bfs::k_stamp_commit<<<grid_cap(ctx, W, BLOCK), BLOCK, 0, ctx.stream>>>(
            w.stamp.as<uint8_t>(), w.last_stamp, out.off, w.visited.as<uint64_t>(),
            w.front[cur ^ 1].as<uint64_t>(), W, V, cnt);
        GRUS_LAUNCH_CHECK();
        launches++;
        stamped = true;
      } else if (nf <= PUSH_SMALL_F) {  // one launch, the degree scan done in LDS
        push_advance_small(ctx, out, w.q[cur].as<uint32_t>(), (uint32_t)nf, mf, job,
                           w.q[cur ^ 1].as<uint32_t>(), cnt);
        launches += 1;
      } else {
        push_scan_degrees(ctx, w.push, out.off, w.q[cur].as<uint32_t>(), (uint32_t)nf);
        push_advance(ctx, w.push, out, w.q[cur].as<uint32_t>(), (uint32_t)nf, mf, job,
                     w.q[cur ^ 1].as<uint32_t>(), cnt);
        launches += 3;
      }
      have_queue = !stamped;  // a stamped step leaves the next frontier as a bitmap only
      have_bitmap = stamped;
    }
    ctx.publish_ctl(sizeof(StepCounters) / 8);  // read the step's counters and clear them
    const StepCounters *hc = reinterpret_cast<const StepCounters *>(ctx.h_ctl);
    if (trace) {
      auto now = std::chrono::steady_clock::now();
      fprintf(stderr, "bfs level %u %s nf=%llu mf=%llu -> next %u (long %u, candidates %u) %.1f us\n",
              level, pull ? (sparse ? "pull-sparse" : "pull") : "push", (unsigned long long)nf,
              (unsigned long long)mf, hc->next_count, hc->long_count, sparse ? hc->rem_count : 0u,
              std::chrono::duration<double, std::micro>(now - t_level).count());
      t_level = now;
    }
    prev_nf = nf;
    nf = hc->next_count;
    // the next step reads nf queue entries: a count past V would be a
    // kernel fault in the making, so it stops here
    if (nf > V) raise(GRUS_E_HIP, "bfs step reported more new vertices than the graph has");
    mf = hc->next_edges;
    mu = mu >= mf ? mu - mf : 0;
    v_reached += nf;
    e_reached += mf;
    cur ^= 1;
    level++;
  }
  GRUS_HIP(hipEventRecord(ctx.ev1, ctx.stream));
  GRUS_HIP(hipEventSynchronize(ctx.ev1));
  float ms = 0;
  GRUS_HIP(hipEventElapsedTime(&ms, ctx.ev0, ctx.ev1));
  if (st) {
    *st = grus_stats{};
    st->device_ms = ms;
    st->iterations = level;
    st->pull_iterations = pull_levels;
    st->vertices_reached = v_reached;
    st->edges_reached = e_reached;
    st->kernel_launches = launches;
  }
}

// Two-stage top-down BFS: BFS_2S_single_gpu (samples/bfs_2s/bfs_2s.cu:94-138)
// over Frontier_2S<BDF> and Subgraph<NORMAL> (include/frontier_2s.cuh:27-109,
// include/subgraph.cuh:46-165, include/kernel_2s.cuh:36-92).  Each level
// gathers the frontier's rows into a device buffer (Subgraph::build: there
// a scan of the degrees, then a host memcpy per row into managed memory and
// a prefetch) and expands them from there.  Here the gather is a kernel: the
// degree scan of the push operator gives every frontier row its place in the
// buffer, and a workgroup copies 1024 consecutive positions, reading rows
// from HBM or, for a capped graph, from pinned host memory over the link.
// The buffer holds stage_edges edges; a larger frontier is cut into pieces
// (a row may span two), and two buffers alternate: a copy stream gathers
// piece p+1 while the context stream expands piece p, ordered by events.
// A vertex enters a BFS frontier once, so the reference's "local" list (the
// previous subgraph's rows that stay active, frontier_2s.cuh:67-97) is
// always empty here and every row is gathered once.
void run_bfs_2s(Graph &g, uint32_t src, uint64_t stage_edges, grus_stats *st) {
  Context &ctx = *g.ctx;
  ctx.activate();
  if (src >= g.V) raise(GRUS_E_INVALID, "source vertex out of range");
  g.make_room(Graph::App::BFS);
  if (!g.bfs) g.bfs = std::make_shared<BfsWork>();
  BfsWork &w = *g.bfs;
  w.ensure(ctx, g.V);
  g.result.ensure(g.V * 4);
  const uint64_t V = g.V, W = g.words();
  const CsrView out = g.csr();
  // staging capacity: as asked (default 2^26 edges, 256 MB per buffer), no
  // more than the graph has, and within a device budget
  uint64_t C = stage_edges ? stage_edges : (64ull << 20);
  C = std::min<uint64_t>(C, std::max<uint64_t>(g.E, 1));
  if (const uint64_t budget = devmem_budget(ctx.device)) {
    const uint64_t held = devmem_live(ctx.device) + w.stage[0].bytes + w.stage[1].bytes;
    const uint64_t margin = 16ull << 20;
    const uint64_t room = budget > held + margin ? budget - held - margin : 0;
    C = std::min<uint64_t>(C, room / 8);
    if (C < 65536) raise(GRUS_E_BUDGET, "no room for the two-stage staging buffers");
  }
  if (w.stage[0].bytes != C * 4) {  // re-sized to this run's capacity
    w.stage[0].release();
    w.stage[1].release();
  }
  w.ensure_stage(C);
  uint32_t *label = g.result.as<uint32_t>();
  auto *cnt = reinterpret_cast<StepCounters *>(ctx.d_ctl.p);
  uint64_t src_deg = 0;
  {
    uint64_t o[2];
    GRUS_HIP(hipMemcpy(o, g.off.as<uint64_t>() + src, 16, hipMemcpyDeviceToHost));
    src_deg = o[1] - o[0];
  }
  uint64_t launches = 0;
  GRUS_HIP(hipMemsetAsync(cnt, 0, sizeof(StepCounters), ctx.stream));
  GRUS_HIP(hipEventRecord(ctx.ev0, ctx.stream));
  bfs::k_init<<<grid_cap(ctx, std::max(V, W), BLOCK), BLOCK, 0, ctx.stream>>>(
      label, V, w.visited.as<uint64_t>(), nullptr, W, src, w.q[0].as<uint32_t>());
  GRUS_LAUNCH_CHECK();
  launches++;
  int cur = 0;
  uint64_t nf = 1, mf = src_deg, v_reached = 1, e_reached = src_deg;
  uint32_t level = 0;
  while (nf) {
    if (level > V) raise(GRUS_E_HIP, "bfs did not converge within V levels");
    const uint32_t *q = w.q[cur].as<uint32_t>();
    bfs::job_t job{label, w.visited.as<uint64_t>(), out.off, level + 1};
    if (mf) {
      // places of the frontier's rows in the edge space (S[nf] = mf)
      push_scan_degrees(ctx, w.push, out.off, q, (uint32_t)nf);
      GRUS_HIP(hipEventRecord(w.ev_scan, ctx.stream));
      GRUS_HIP(hipStreamWaitEvent(w.copy_stream, w.ev_scan, 0));
      const uint64_t *S = w.push.scan_ptr();
      for (uint64_t k0 = 0, p = 0; k0 < mf; k0 += C, p++) {
        const int b = (int)(p & 1);
        const uint64_t k1 = std::min(mf, k0 + C);
        uint32_t *buf = w.stage[b].as<uint32_t>();
        // the copy stream fills buffer b once its previous piece is expanded
        GRUS_HIP(hipStreamWaitEvent(w.copy_stream, w.ev_free[b], 0));
        gather_rows(ctx, out, q, S, (uint32_t)nf, k0, k1, buf, nullptr, w.copy_stream);
        GRUS_HIP(hipEventRecord(w.ev_ready[b], w.copy_stream));
        GRUS_HIP(hipStreamWaitEvent(ctx.stream, w.ev_ready[b], 0));
        const CsrView stage{nullptr, buf, nullptr, nullptr, nullptr, ~0ull, V, k1 - k0};
        push_advance_staged(ctx, stage, q, S, (uint32_t)nf, k0, k1, job,
                            w.q[cur ^ 1].as<uint32_t>(), cnt, ctx.stream);
        GRUS_HIP(hipEventRecord(w.ev_free[b], ctx.stream));
        launches += 2;
      }
      launches += 2;
    }
    ctx.publish_ctl(sizeof(StepCounters) / 8);  // read the step's counters and clear them
    const StepCounters *hc = reinterpret_cast<const StepCounters *>(ctx.h_ctl);
    nf = hc->next_count;
    if (nf > V) raise(GRUS_E_HIP, "bfs step reported more new vertices than the graph has");
    mf = hc->next_edges;
    v_reached += nf;
    e_reached += mf;
    cur ^= 1;
    level++;
  }
  GRUS_HIP(hipEventRecord(ctx.ev1, ctx.stream));
  GRUS_HIP(hipEventSynchronize(ctx.ev1));
  float ms = 0;
  GRUS_HIP(hipEventElapsedTime(&ms, ctx.ev0, ctx.ev1));
  if (st) {
    *st = grus_stats{};
    st->device_ms = ms;
    st->iterations = level;
    st->vertices_reached = v_reached;
    st->edges_reached = e_reached;
    st->kernel_launches = launches;
    st->reserved[0] = C;  // staging capacity used, in edges
  }
}

}  // namespace mgg
