// advance.hpp -- frontier operators shared by the apps.
//
// Reference counterparts: push_kernel (include/kernel.cuh:42-123),
// flag_to_wl / wl_to_flag (src/worklist.cu:87-100) and the Frontier<BDF>
// (include/frontier.cuh:27-76).  The reference expands one 32-lane warp per
// frontier vertex, so a hub of degree 10^5 serializes in one warp and
// vertices of degree < 32 idle most lanes.  Here the push operator is
// edge-balanced: the frontier's out-degrees are prefix-summed, every
// workgroup takes a fixed-size chunk of the frontier's edge space, stages
// the owning queue entries in LDS and assigns one edge per lane (owner
// found by binary search in LDS), so any degree mix keeps all 64 lanes of
// every wave busy and all column reads coalesced.
//
// Functor contract kept from the reference (samples/bfs/bfs.cu:51-87):
//   bool Job::update(src, dst, edge_id, weight)  -- the app's updater
//   int  Job::generate(updated, dst)             -- the app's generator:
//        1 when dst joins the next frontier (deduplicated), 2 when it joins
//        the far queue (Job::kFar apps only), 0 otherwise
// The operator appends generated vertices to the next queue with one
// atomic per wave and counts the next frontier's size and out-edges (the
// inputs of the push/pull direction heuristic).
#pragma once

#include "device.hpp"
#include "graph.hpp"

namespace mgg {

// counters of one frontier step, in the context's device control block
struct StepCounters {
  uint32_t next_count;   // vertices appended to the next queue
  uint32_t long_count;   // bottom-up vertices handed to a cooperative scan
  uint64_t next_edges;   // sum of out-degrees of the next frontier
  uint32_t far_count;    // vertices appended to the far queue (SSSP)
  uint32_t min_far;      // minimum live far distance (SSSP threshold advance)
  uint32_t rem_count;    // BFS bottom-up: vertices still unvisited after the step
  uint32_t pad32;
  uint64_t pad[4];
};
static_assert(sizeof(StepCounters) == 64, "counter block layout");

__device__ __forceinline__ uint32_t edge_col(const CsrView &g, uint64_t e) {
  return e < g.split ? g.col[e] : g.col_far[e - g.split];
}
__device__ __forceinline__ uint32_t edge_w(const CsrView &g, uint64_t e) {
  if (!g.w) return 1u;
  return e < g.split ? g.w[e] : g.w_far[e - g.split];
}

// ---------------------------------------------------------------- helpers
// (host wrappers; kernels live in advance.hip)
// q <- every set bit of bm[0..W) in ascending order within each workgroup;
// *count (device) receives the number appended (must be zero on entry)
void bitmap_to_queue(Context &ctx, const uint64_t *bm, uint64_t W, uint32_t *q, uint32_t *count);
// the same for every clear bit of bm[0..W) (bits past the last vertex must be set)
void clear_bits_to_queue(Context &ctx, const uint64_t *bm, uint64_t W, uint32_t *q,
                         uint32_t *count);
// bm[0..W) <- bitmap of q[0..n)
void queue_to_bitmap(Context &ctx, const uint32_t *q, uint32_t n, uint64_t *bm, uint64_t W);
void fill_u32(Context &ctx, uint32_t *a, uint64_t n, uint32_t v);
void fill_u64(Context &ctx, uint64_t *a, uint64_t n, uint64_t v);
unsigned grid_cap(Context &ctx, uint64_t items, unsigned per_block, unsigned blocks_per_cu = 16);
// V_r = #{v : val[v] != INFINIT}, E_r = sum of out-degrees over them (synchronous)
void reached_stats(Context &ctx, const uint64_t *off, const uint32_t *val, uint64_t V,
                   uint64_t *vr, uint64_t *er);

constexpr int LB_CHUNK = 1024;          // edges per workgroup step (4 per lane)
constexpr int LB_WIN = LB_CHUNK + 64;   // queue entries staged in LDS
constexpr uint32_t LB_PER = LB_CHUNK / BLOCK;  // edges per thread per chunk
static_assert(LB_CHUNK % BLOCK == 0 && LB_PER <= 32, "a chunk's generate bits fit a u32 per thread");

// first index in [lo, hi) with a[idx] > key (a non-decreasing), hi if none
__device__ __forceinline__ uint32_t upper_bound_u64(const uint64_t *a, uint32_t lo, uint32_t hi,
                                                    uint64_t key) {
  while (lo < hi) {
    uint32_t mid = lo + ((hi - lo) >> 1);
    if (a[mid] <= key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Edge-balanced push over queue q[0..F) with S = exclusive scan of its
// out-degrees (S[F] = total), for the positions [k0, k1) of the frontier's
// edge space.  Job supplies update/generate/next-queue.  kStaged: the edges
// of positions [k0, k1) were gathered into g.col / g.w at k - k0 (a staged
// piece of a two-stage step); the edge id the updater sees is then k - k0.
template <class Job, bool kStaged = false>
__global__ __launch_bounds__(BLOCK) void k_push_lb(CsrView g, const uint32_t *q,
                                                   const uint64_t *S, uint32_t F, uint64_t k0,
                                                   uint64_t total, Job job, uint32_t *qn,
                                                   uint32_t *qfar, StepCounters *cnt) {
  __shared__ uint64_t sS[LB_WIN];
  __shared__ uint64_t sRow[LB_WIN];
  __shared__ uint32_t sQ[LB_WIN];
  __shared__ uint32_t s_i0, s_n;
  __shared__ uint32_t s_app[BLOCK / WAVE + 1];
  uint64_t my_edges = 0;
  for (uint64_t c0 = k0 + (uint64_t)blockIdx.x * LB_CHUNK; c0 < total;
       c0 += (uint64_t)gridDim.x * LB_CHUNK) {
    uint64_t c1 = c0 + LB_CHUNK < total ? c0 + LB_CHUNK : total;
    if (threadIdx.x == 0) {
      uint32_t i0 = upper_bound_u64(S, 0, F + 1, c0) - 1;
      uint32_t i1 = upper_bound_u64(S, i0, F + 1, c1 - 1) - 1;
      s_i0 = i0;
      s_n = i1 - i0 + 2;
    }
    __syncthreads();
    const uint32_t i0 = s_i0, n = s_n;
    const bool staged = n <= LB_WIN;
    if (staged) {
      for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) {
        sS[t] = S[i0 + t];
        if (i0 + t < F) {
          uint32_t u = q[i0 + t];
          sQ[t] = u;
          if (!kStaged) sRow[t] = g.off[u];
        }
      }
    }
    __syncthreads();
    // every edge of the chunk is updated first (LB_CHUNK / BLOCK per thread,
    // in coalesced rounds); the generated vertices then take their queue
    // slots with one reservation per queue for the whole chunk
    uint32_t dsts[LB_PER];
    uint32_t g_near = 0, g_far = 0;  // bit i: the thread's edge i generates into qn / qfar
#pragma unroll
    for (uint32_t i = 0; i < LB_PER; i++) {
      const uint64_t k = c0 + (uint64_t)i * BLOCK + threadIdx.x;
      dsts[i] = 0;
      if (k < c1) {
        uint32_t j, u;
        uint64_t base, row;
        if (staged) {
          j = upper_bound_u64(sS, 0, n, k) - 1;
          base = sS[j];
          u = sQ[j];
          row = kStaged ? 0 : sRow[j];
        } else {
          j = upper_bound_u64(S + i0, 0, n, k) - 1;
          base = S[i0 + j];
          u = q[i0 + j];
          row = kStaged ? 0 : g.off[u];
        }
        const uint64_t e = kStaged ? k - k0 : row + (k - base);
        const uint32_t dst = edge_col(g, e);
        uint32_t val = 0;  // what the updater wrote (a distance), for the generator
        bool upd = job.update(u, dst, e, edge_w(g, e), &val);
        const int gen = job.generate(upd, dst, val);
        dsts[i] = dst;
        if (gen == 1) g_near |= 1u << i;
        if (gen == 2) g_far |= 1u << i;
      }
    }
    uint32_t slot = block_reserve(&cnt->next_count, (uint32_t)__popc(g_near), s_app);
#pragma unroll
    for (uint32_t i = 0; i < LB_PER; i++) {
      if ((g_near >> i) & 1u) {
        qn[slot++] = dsts[i];
        my_edges += job.out_degree(dsts[i]);
      }
    }
    if (Job::kFar) {
      uint32_t fslot = block_reserve(&cnt->far_count, (uint32_t)__popc(g_far), s_app);
#pragma unroll
      for (uint32_t i = 0; i < LB_PER; i++)
        if ((g_far >> i) & 1u) qfar[fslot++] = dsts[i];
    }
    __syncthreads();
  }
  block_add_u64(my_edges, (unsigned long long *)&cnt->next_edges);
}

// The same push for a small frontier (F <= PUSH_SMALL_F) in one launch:
// every workgroup stages the whole queue, its row starts and degrees in LDS
// and prefix-sums the degrees itself, so no degree kernel and device-wide
// scan (three launches and their gaps) precede the expansion.  total (the
// frontier's out-degree sum, known to the caller) sizes the grid; the
// kernel walks the degree sum it computes itself.
constexpr uint32_t PUSH_SMALL_F = 1024;
template <class Job>
__global__ __launch_bounds__(BLOCK) void k_push_small(CsrView g, const uint32_t *q, uint32_t F,
                                                      uint64_t total, Job job, uint32_t *qn,
                                                      uint32_t *qfar, StepCounters *cnt) {
  __shared__ uint64_t sS[PUSH_SMALL_F + 1];  // exclusive prefix of the degrees
  __shared__ uint64_t sRow[PUSH_SMALL_F];
  __shared__ uint32_t sQ[PUSH_SMALL_F];
  __shared__ uint64_t s_wsum[BLOCK / WAVE];
  __shared__ uint32_t s_app[BLOCK / WAVE + 1];
  constexpr uint32_t PER = PUSH_SMALL_F / BLOCK;  // queue entries per thread
  const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
  // stage: thread t owns entries [t*PER, t*PER + PER)
  uint64_t d[PER];
  uint64_t mine = 0;
#pragma unroll
  for (int i = 0; i < (int)PER; i++) {
    const uint32_t t = threadIdx.x * PER + i;
    d[i] = 0;
    if (t < F) {
      const uint32_t u = q[t];
      const uint64_t a = g.off[u];
      sQ[t] = u;
      sRow[t] = a;
      d[i] = g.off[u + 1] - a;
    }
    mine += d[i];
  }
  // exclusive block scan of the per-thread sums (wave shuffles, then LDS)
  uint64_t x = mine;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const uint64_t y = __shfl_up(x, o);
    if ((int)lane >= o) x += y;
  }
  if (lane == WAVE - 1) s_wsum[wid] = x;
  __syncthreads();
  uint64_t before = 0, sum = 0;
  for (uint32_t w2 = 0; w2 < BLOCK / WAVE; w2++) {
    before += (w2 < wid) ? s_wsum[w2] : 0ull;
    sum += s_wsum[w2];
  }
  uint64_t run = before + x - mine;
#pragma unroll
  for (int i = 0; i < (int)PER; i++) {
    const uint32_t t = threadIdx.x * PER + i;
    if (t < F) sS[t] = run;
    run += d[i];
  }
  if (threadIdx.x == 0) sS[F] = sum;
  __syncthreads();
  // the edge space is the staged degree sum (total only sized the grid)
  uint64_t my_edges = 0;
  // every thread of the block runs the same trip count (block_append)
  for (uint64_t kb = (uint64_t)blockIdx.x * blockDim.x; kb < sum;
       kb += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = kb + threadIdx.x;
    int gen = 0;
    uint32_t dst = 0;
    if (k < sum) {
      const uint32_t j = upper_bound_u64(sS, 0, F + 1, k) - 1;
      const uint32_t u = sQ[j];
      const uint64_t e = sRow[j] + (k - sS[j]);
      dst = edge_col(g, e);
      uint32_t val = 0;
      const bool upd = job.update(u, dst, e, edge_w(g, e), &val);
      gen = job.generate(upd, dst, val);
    }
    const uint32_t slot = block_append(&cnt->next_count, gen == 1, s_app);
    if (gen == 1) {
      qn[slot] = dst;
      my_edges += job.out_degree(dst);
    }
    if (Job::kFar) {
      const uint32_t fslot = block_append(&cnt->far_count, gen == 2, s_app);
      if (gen == 2) qfar[fslot] = dst;
    }
  }
  block_add_u64(my_edges, (unsigned long long *)&cnt->next_edges);
}

// Host side of the push operator: degree scan + edge-balanced expansion
// (PushWork, the scan buffers, is declared in graph.hpp).

// w.scan[0..F] <- exclusive scan of outdeg(q[i]); w.scan[F] = total
void push_scan_degrees(Context &ctx, PushWork &w, const uint64_t *off, const uint32_t *q,
                       uint32_t F);
unsigned push_grid(Context &ctx, uint64_t total);

template <class Job>
void push_advance(Context &ctx, PushWork &w, const CsrView &g, const uint32_t *q, uint32_t F,
                  uint64_t total, const Job &job, uint32_t *qn, StepCounters *cnt,
                  uint32_t *qfar = nullptr) {
  if (F == 0 || total == 0) return;
  k_push_lb<Job><<<push_grid(ctx, total), BLOCK, 0, ctx.stream>>>(
      g, q, w.scan_ptr(), F, 0, total, job, qn, qfar, cnt);
  GRUS_LAUNCH_CHECK();
}

// Two-stage step (a staged piece): the frontier positions [k0, k1) whose
// edges were gathered into stage (stage.col / stage.w at k - k0).  The grid
// stays below a full chip (STAGED_PER_CU workgroups per CU) so that the
// gather of the next piece, on another stream, runs beside it.
#ifndef STAGED_PER_CU
#define STAGED_PER_CU 4
#endif
template <class Job>
void push_advance_staged(Context &ctx, const CsrView &stage, const uint32_t *q,
                         const uint64_t *S, uint32_t F, uint64_t k0, uint64_t k1, const Job &job,
                         uint32_t *qn, StepCounters *cnt, hipStream_t stream) {
  if (F == 0 || k1 <= k0) return;
  k_push_lb<Job, true><<<grid_cap(ctx, k1 - k0, LB_CHUNK, STAGED_PER_CU), BLOCK, 0, stream>>>(
      stage, q, S, F, k0, k1, job, qn, nullptr, cnt);
  GRUS_LAUNCH_CHECK();
}
// positions [k0, k1) of queue q's edge space (scan S) gathered from g into
// col / w (w may be null) at k - k0: the frontier's rows, wherever they live
// (HBM or the host part of a capped graph), copied to a device buffer
void gather_rows(Context &ctx, const CsrView &g, const uint32_t *q, const uint64_t *S, uint32_t F,
                 uint64_t k0, uint64_t k1, uint32_t *col, uint32_t *w, hipStream_t stream);

// a frontier of at most PUSH_SMALL_F vertices: one launch (k_push_small)
template <class Job>
void push_advance_small(Context &ctx, const CsrView &g, const uint32_t *q, uint32_t F,
                        uint64_t total, const Job &job, uint32_t *qn, StepCounters *cnt,
                        uint32_t *qfar = nullptr) {
  if (F == 0 || total == 0) return;
  if (F > PUSH_SMALL_F) raise(GRUS_E_INVALID, "push_advance_small: frontier too large");
  k_push_small<Job><<<grid_cap(ctx, total, BLOCK, 8), BLOCK, 0, ctx.stream>>>(
      g, q, F, total, job, qn, qfar, cnt);
  GRUS_LAUNCH_CHECK();
}

}  // namespace mgg
