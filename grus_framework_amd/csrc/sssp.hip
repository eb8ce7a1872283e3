// sssp.hip -- single-source shortest paths, near-far delta stepping.
//
// Reference: samples/sssp/sssp.cu (SSSPInit :18-23, job_t :24-49, updater
// :51-60, push driver :123-170).  The reference relaxes with a non-atomic
// check-then-store (sssp.cu:54-55), which can lose an update; here the
// relax is atomicMin, so the fixed point -- the unique shortest distances --
// is reached whatever the interleaving and the result is bit-exact.
//
// Schedule (Davidson et al., IPDPS'14 "near-far"): vertices whose tentative
// distance is below the threshold T form the near frontier and are relaxed
// by the edge-balanced push operator round after round; improved vertices
// at or beyond T go to a far queue.  When the near frontier empties, T
// advances to the bucket holding the smallest live far distance and the far
// queue is split.  Queue membership is deduplicated with per-vertex round
// stamps (u32), which never need clearing between rounds.
#include <algorithm>
#include <cstdlib>

#include "advance.hpp"
#include "apps.hpp"
#include "persist.hpp"
#include "window.hpp"

namespace mgg {

struct SsspWork {
  Buf q[2], far[2], near_stamp, far_stamp, pctl;
  Buf pad;               // persistent path: padded rows (see k_pad_rows)
  uint32_t pad_deg = 0;  // slots per padded row (0 = not built)
  Buf front_bm;               // pull rounds: frontier bitmap
  Buf win_row, head_min, tail_min;  // pull rounds: window table of the in-edges
  const void *win_for = nullptr;    // in-edge offsets the table was built on
  uint64_t nwin = 0;
  PushWork push;
  uint64_t cap_v = 0;
  uint32_t stamp = 0;  // last stamp handed out; stamps only grow
  double mean_w = -1;
  void ensure(Context &ctx, uint64_t V) {
    if (V <= cap_v) return;
    q[0].alloc((V + 1) * 4);
    q[1].alloc((V + 1) * 4);
    far[0].alloc((V + 1) * 4);
    far[1].alloc((V + 1) * 4);
    near_stamp.alloc((V + 1) * 4);
    far_stamp.alloc((V + 1) * 4);
    front_bm.alloc((V / 64 + 2) * 8);
    win_for = nullptr;
    fill_u32(ctx, near_stamp.as<uint32_t>(), V + 1, 0);
    fill_u32(ctx, far_stamp.as<uint32_t>(), V + 1, 0);
    push.ensure(V);
    cap_v = V;
    stamp = 0;
  }
  uint64_t wraps = 0;  // stamp restarts so far
  uint32_t next_stamp(Context &ctx, uint64_t V) {
    if (stamp >= 0xFFFFFF00u) {  // wrap: clear and restart
      fill_u32(ctx, near_stamp.as<uint32_t>(), V + 1, 0);
      fill_u32(ctx, far_stamp.as<uint32_t>(), V + 1, 0);
      stamp = 0;
      wraps++;
    }
    return ++stamp;
  }
};

namespace sssp {

// job_t of samples/sssp/sssp.cu:24-49 as the device-side job of one round
struct job_t {
  static constexpr bool kFar = true;
  uint32_t *dist;
  uint32_t *near_stamp;
  uint32_t *far_stamp;
  const uint64_t *out_off;
  uint32_t T;           // near threshold (exclusive)
  uint32_t near_id;     // stamp of the near queue being built
  uint32_t far_id;      // stamp of the current far epoch
  // updater (sssp.cu:51-60) with an atomic min.  Distances only decrease,
  // so plain (cached) reads can only be stale-high: relaxing with a stale
  // dist[src] can only miss an improvement whose own update re-queues src,
  // and a stale dist[dst] only lets through an atomic that turns out not to
  // improve.  The fixed point -- and the result -- are unchanged, and most
  // non-improving edges cost no device-scope atomic.
  __device__ __forceinline__ bool update(vtx_t src, vtx_t dst, eid_t, weight_t w,
                                         uint32_t *nd_out) const {
    const uint64_t nd = (uint64_t)dist[src] + w;
    if (nd >= INFINIT || nd >= dist[dst]) return false;  // distances are u32 below INFINIT
    *nd_out = (uint32_t)nd;
    return atomicMin(&dist[dst], (uint32_t)nd) > (uint32_t)nd;
  }
  // generator: an improved vertex joins the near queue (below T) or the far
  // queue, once per queue.  nd is the distance this thread wrote; a thread
  // that lowered it further classifies it again itself.  Stamps are unique
  // per queue, so a plain read that already shows this queue's stamp is
  // exact.
  __device__ __forceinline__ int generate(bool updated, vtx_t dst, uint32_t nd) const {
    if (!updated) return 0;
    if (nd < T) {
      if (near_stamp[dst] == near_id) return 0;
      return atomicExch(&near_stamp[dst], near_id) != near_id ? 1 : 0;
    }
    if (far_stamp[dst] == far_id) return 0;
    return atomicExch(&far_stamp[dst], far_id) != far_id ? 2 : 0;
  }
  __device__ __forceinline__ uint64_t out_degree(vtx_t v) const {
    return out_off[v + 1] - out_off[v];
  }
};

// ---- pull rounds (large frontiers) ----
// A pull round is a window reduction (window.hpp) over the in-edges: the
// value of edge (u, v) is dist[u] + w when u is in the frontier bitmap, the
// reduction is the minimum, and finishing row v lowers dist[v] when the
// minimum is smaller.  Only row v's finisher writes dist[v], so no atomics
// are needed on distances; reads of dist[u] may see values lowered in this
// same round, which are still lengths of real paths.  Improved vertices join
// the near queue (below T) or the far list (once per epoch, by stamp),
// staged in LDS by the window kernel and appended directly by the fixup.
struct PullMinOp {
  using T = uint32_t;
  CsrView in;
  const uint64_t *front;
  uint32_t *dist;
  const uint64_t *out_off;
  uint32_t T_thr;
  uint32_t *near_stamp, near_id, *far_stamp, far_id;
  uint32_t *head_min, *tail_min;
  uint32_t *qn, *qfar;
  StepCounters *cnt;
  // LDS staging (window kernel); null in the fixup
  uint32_t *s_near, *s_far, *s_nn, *s_nf;
  unsigned long long *s_edges;

  __device__ T identity() const { return 0xFFFFFFFFu; }
  __device__ T combine(T a, T b) const { return a < b ? a : b; }
  __device__ T value(uint64_t e, uint32_t u) const {
    if (!bm_test(front, u)) return 0xFFFFFFFFu;
    const uint64_t c = (uint64_t)dist[u] + edge_w(in, e);
    return c < INFINIT ? (uint32_t)c : 0xFFFFFFFFu;
  }
  __device__ void finish(uint32_t r, T best) const {
    if (best >= dist[r]) return;
    dist[r] = best;
    if (best < T_thr) {
      near_stamp[r] = near_id;
      const unsigned long long deg = out_off[r + 1] - out_off[r];
      if (s_near) {
        s_near[atomicAdd(s_nn, 1u)] = r;
        atomicAdd(s_edges, deg);
      } else {
        qn[atomicAdd(&cnt->next_count, 1u)] = r;
        atomicAdd((unsigned long long *)&cnt->next_edges, deg);
      }
    } else if (far_stamp[r] != far_id) {
      far_stamp[r] = far_id;
      if (s_far) s_far[atomicAdd(s_nf, 1u)] = r;
      else qfar[atomicAdd(&cnt->far_count, 1u)] = r;
    }
  }
  __device__ void empty(uint32_t) const {}
  __device__ void head(uint64_t k, T t) const { head_min[k] = t; }
  __device__ void tail(uint64_t k, T t) const { tail_min[k] = t; }
  __device__ T head_at(uint64_t k) const { return head_min[k]; }
  __device__ T tail_at(uint64_t k) const { return tail_min[k]; }
};

__global__ __launch_bounds__(BLOCK) void k_pull_min_window(CsrView in, const uint32_t *win_row,
                                                           PullMinOp op, uint64_t nwin) {
  __shared__ WindowShared sh;
  // a window finishes at most one row per edge (plus none for empty rows)
  __shared__ uint32_t s_near[WIN + 2], s_far[WIN + 2];
  __shared__ uint32_t s_nn, s_nf, s_bn, s_bf;
  __shared__ unsigned long long s_edges;
  const uint64_t k = blockIdx.x;
  if (k >= nwin) return;
  if (threadIdx.x == 0) {
    s_nn = s_nf = 0;
    s_edges = 0;
  }
  __syncthreads();
  PullMinOp o = op;
  o.s_near = s_near;
  o.s_far = s_far;
  o.s_nn = &s_nn;
  o.s_nf = &s_nf;
  o.s_edges = &s_edges;
  window_reduce(in, win_row, k, o, sh);
  __syncthreads();
  // one global append per list per workgroup
  if (threadIdx.x == 0) {
    s_bn = s_nn ? atomicAdd(&op.cnt->next_count, s_nn) : 0u;
    s_bf = s_nf ? atomicAdd(&op.cnt->far_count, s_nf) : 0u;
    if (s_edges) atomicAdd((unsigned long long *)&op.cnt->next_edges, s_edges);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < s_nn; i += BLOCK) op.qn[s_bn + i] = s_near[i];
  for (uint32_t i = threadIdx.x; i < s_nf; i += BLOCK) op.qfar[s_bf + i] = s_far[i];
}

__global__ void k_pull_min_fixup(CsrView in, const uint32_t *win_row, PullMinOp op,
                                 uint64_t nwin) {
  uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (k >= nwin) return;
  window_fixup(in, win_row, k, op);
}

// the first queue is {src}, stamped like every later queue
__global__ void k_init(uint32_t *dist, uint64_t V, uint32_t src, uint32_t *q0,
                       uint32_t *near_stamp, uint32_t id) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < V; i += stride) {
    dist[i] = (i == src) ? 0u : INFINIT;  // SSSPInit sssp.cu:21
    if (i == 0) {
      q0[0] = src;
      near_stamp[src] = id;
    }
  }
}

// smallest distance >= T among far entries
__global__ void k_far_min(const uint32_t *far, uint32_t n, const uint32_t *dist, uint32_t T,
                          uint32_t *out) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint32_t m = 0xFFFFFFFFu;
  for (; i < n; i += stride) {
    uint32_t d = dist[far[i]];
    if (d >= T && d < m) m = d;
  }
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t x = __shfl_xor(m, o);
    m = x < m ? x : m;
  }
  if (lane_id() == 0 && m != 0xFFFFFFFFu) atomicMin(out, m);
}

// split the far queue at the new threshold: [T_old, T) -> near, >= T -> far'
// (entries below T_old were already relaxed at their final near distance)
__global__ void k_far_split(const uint32_t *far, uint32_t n, const uint32_t *dist,
                            const uint64_t *out_off, uint32_t T_old, uint32_t T,
                            uint32_t *near_stamp, uint32_t near_id, uint32_t *far_stamp,
                            uint32_t far_id, uint32_t *qn, uint32_t *far_out,
                            StepCounters *cnt) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t first = blockIdx.x * (uint64_t)blockDim.x + (threadIdx.x & ~(uint64_t)(WAVE - 1));
  uint64_t my_edges = 0;
  for (uint64_t base = first; base < n; base += stride) {
    uint64_t i = base + lane_id();
    int to = 0;
    uint32_t v = 0;
    if (i < n) {
      v = far[i];
      uint32_t d = dist[v];
      if (d >= T_old) {
        if (d < T) to = atomicExch(&near_stamp[v], near_id) != near_id ? 1 : 0;
        else to = atomicExch(&far_stamp[v], far_id) != far_id ? 2 : 0;
      }
    }
    uint32_t s1 = wave_append(&cnt->next_count, to == 1);
    if (to == 1) {
      qn[s1] = v;
      my_edges += out_off[v + 1] - out_off[v];
    }
    uint32_t s2 = wave_append(&cnt->far_count, to == 2);
    if (to == 2) far_out[s2] = v;
  }
  block_add_u64(my_edges, (unsigned long long *)&cnt->next_edges);
}

__global__ void k_weight_sum(const uint32_t *w, uint64_t E, unsigned long long *out) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long s = 0;
  for (; i < E; i += stride) s += w[i];
  s = wave_sum(s);
  if (lane_id() == 0) atomicAdd(out, s);
}

// ---------------------------------------------------------------------------
// Persistent near-far for low-degree graphs (grids, road-like inputs): the
// whole traversal is one cooperative launch; rounds are separated by
// grid_sync instead of a return to the host.  A lane owns one frontier
// vertex and walks its (at most max_deg) edges; generated vertices are
// appended with one atomic per workgroup.  Counters live in a ring of 4
// slots: slot s&3 belongs to step s, workgroup 0 clears slot (s+2)&3 at the
// start of step s (no workgroup can still be reading it), and every
// workgroup derives identical loop state from the slots it reads after a
// barrier.
struct PCtl {
  GridSync gs;
  uint32_t ring[4][4];  // per step: next count, far count, min far distance, unused
  uint32_t rounds, stamp, status, passes;  // passes: local passes of workgroup 0
};

struct PParams {
  CsrView g;
  const uint2 *pad;  // padded rows of pad_deg slots {col, w}, or null (CSR)
  uint32_t pad_deg;
  uint32_t *dist, *q0, *q1, *far0, *far1, *near_stamp, *far_stamp;
  PCtl *ctl;
  uint64_t delta, round_cap, timeout;
  uint32_t stamp0, max_deg;
  uint32_t local_hops;  // extra in-workgroup relaxation passes per round
};

__device__ __forceinline__ void clear_slot(PCtl *c, uint32_t step) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    uint32_t *z = c->ring[(step + 2) & 3];
    st_agent(&z[0], 0);
    st_agent(&z[1], 0);
    st_agent(&z[2], 0xFFFFFFFFu);
  }
}

// LDS staging of one round's output: vertices generated by a workgroup are
// collected in LDS and published with one atomic per workgroup per round
constexpr uint32_t PQ_CAP = 2048;
// the ring of vertices a workgroup owes a relaxation within a round
constexpr uint32_t PL_CAP = 4096;
#ifndef RELAX_GROUP
#define RELAX_GROUP 4
#endif
static_assert(RELAX_GROUP == 2 || RELAX_GROUP == 4,
              "a padded row (4 or 8 slots) is read in whole groups of two-slot loads");

__device__ __forceinline__ void stage(uint32_t v, uint32_t *lq, uint32_t *lcount,
                                      uint32_t *gcount, uint32_t *gq, uint32_t goff) {
  uint32_t s = atomicAdd(lcount, 1u);
  if (s < PQ_CAP) {
    lq[s] = v;
  } else {  // overflow: publish directly
    uint32_t g = atomicAdd(gcount, 1u);
    st_agent(&gq[goff + g], v);
  }
}

__device__ __forceinline__ void publish(const uint32_t *lq, uint32_t lcount, uint32_t *gcount,
                                        uint32_t *gq, uint32_t goff, uint32_t *s_base) {
  uint32_t n = lcount < PQ_CAP ? lcount : PQ_CAP;
  if (threadIdx.x == 0) *s_base = n ? atomicAdd(gcount, n) : 0u;
  __syncthreads();
  const uint32_t base = *s_base;
  for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) st_agent(&gq[goff + base + k], lq[k]);
}

// Padded rows for low-degree graphs: row v holds D slots {col, w} at v * D
// (unused slots weigh INFINIT, so they never relax), so a relaxation reads
// its edges with no offset lookup first: one dependent load fewer per hop.
__global__ __launch_bounds__(BLOCK) void k_pad_rows(CsrView g, uint32_t D, uint2 *pad) {
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t v = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; v < g.V; v += stride) {
    const uint64_t e0 = g.off[v], d = g.off[v + 1] - e0;
    for (uint32_t i = 0; i < D; i++)
      pad[v * D + i] = i < d ? make_uint2(edge_col(g, e0 + i), edge_w(g, e0 + i))
                             : make_uint2(0u, INFINIT);
  }
}

__global__ __launch_bounds__(BLOCK) void k_persistent(PParams p) {
  __shared__ uint32_t s_near[PQ_CAP], s_far[PQ_CAP], s_loc[PL_CAP], s_locd[PL_CAP];
  __shared__ uint32_t s_nn, s_nf, s_base, s_hi, s_nl;
  __shared__ uint32_t s_scr[BLOCK / WAVE + 1];
  const CsrView g = p.g;
  unsigned int gen = 0;
  uint32_t step = 0, stamp = p.stamp0;
  uint32_t far_id = ++stamp;
  uint32_t nn = 1, nfar = 0;
  uint64_t T = p.delta, rounds = 0;
  uint32_t *q[2] = {p.q0, p.q1};
  uint32_t *far[2] = {p.far0, p.far1};
  int cur = 0, fcur = 0;
  bool ok = true;
  uint32_t passes = 0;
  // the smallest distance in the far list, known after a split that found
  // its bucket empty (until the next round adds far vertices)
  uint32_t fmin = 0xFFFFFFFFu;
  bool fmin_known = false;
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  while (ok) {
    uint32_t *C = p.ctl->ring[step & 3];
    clear_slot(p.ctl, step);
    if (nn > 0) {
      if (++rounds > p.round_cap) {
        ok = false;
        break;
      }
      const uint32_t near_id = ++stamp;
      const uint32_t Tk = T < 0xFFFFFFFFull ? (uint32_t)T : 0xFFFFFFFFu;
      if (threadIdx.x == 0) s_nn = s_nf = 0;
      __syncthreads();
      // An improved vertex below T is owed one more relaxation with its new
      // distance.  Its improver keeps it in the workgroup's LDS ring and
      // relaxes it in a later pass of this round, or, when the ring is full
      // or the round's pass budget is spent, publishes it to the next
      // round's queue (once per round: stamps).  A vertex improved again
      // later is owed again by that improver, so at the end of a bucket
      // every vertex below T has been relaxed with its final distance.
      // the current pass reads ring entries from pass_lo on: an append at
      // absolute position l is kept only while l < pass_lo + PL_CAP, so it
      // never lands on a slot still to be read
      uint32_t pass_lo = 0;
      auto owe = [&](uint32_t v, uint32_t dv) {
        const uint32_t l = atomicAdd(&s_nl, 1u);
        if (l < pass_lo + PL_CAP) {
          s_loc[l % PL_CAP] = v;
          s_locd[l % PL_CAP] = dv;  // the distance it is owed a relaxation with
        } else if (atomicExch(&p.near_stamp[v], near_id) != near_id) {
          stage(v, s_near, &s_nn, &C[0], q[cur ^ 1], 0);
        }
      };
      // A vertex improved to T or beyond belongs to the far list, once per
      // far epoch (stamps).  It is kept in LDS unstamped and the round's end
      // stamps all of them together, so no relaxation waits for a stamp
      // exchange (a returning device atomic; nearly every wave has a lane
      // with a far improvement).  An LDS overflow stamps at once.
      auto far_defer = [&](uint32_t v) {
        const uint32_t sl = atomicAdd(&s_nf, 1u);
        if (sl < PQ_CAP) {
          s_far[sl] = v;
        } else if (atomicExch(&p.far_stamp[v], far_id) != far_id) {
          st_agent(&far[fcur][nfar + atomicAdd(&C[1], 1u)], v);
        }
      };
      // the atomicMin of up to RELAX_GROUP edges issued together (the
      // updates of one row are independent), then their outcomes
      auto settle = [&](const uint32_t *v, const uint32_t *nd) {
        uint32_t old[RELAX_GROUP];
#pragma unroll
        for (int i = 0; i < RELAX_GROUP; i++)
          old[i] = nd[i] < INFINIT ? atomicMin(&p.dist[v[i]], nd[i]) : 0u;
#pragma unroll
        for (int i = 0; i < RELAX_GROUP; i++) {
          if (nd[i] >= INFINIT || old[i] <= nd[i]) continue;
          if (nd[i] < Tk) owe(v[i], nd[i]);
          else far_defer(v[i]);
        }
      };
      // relax u's out-edges with distance du (u's distance when it was
      // improved; a later improvement is owed by its own improver)
      auto relax = [&](uint32_t u, uint32_t du) {
        uint32_t v[RELAX_GROUP], nd[RELAX_GROUP];
        if (p.pad) {
          const uint4 *row = reinterpret_cast<const uint4 *>(p.pad + (uint64_t)u * p.pad_deg);
          for (uint32_t b = 0; b < p.pad_deg; b += RELAX_GROUP) {
#pragma unroll
            for (int i = 0; i < RELAX_GROUP; i += 2) {
              const uint4 x = row[(b + i) / 2];
              v[i] = x.x;
              v[i + 1] = x.z;
              const uint64_t a = (uint64_t)du + x.y, c = (uint64_t)du + x.w;
              nd[i] = a < INFINIT ? (uint32_t)a : INFINIT;
              nd[i + 1] = c < INFINIT ? (uint32_t)c : INFINIT;
            }
            settle(v, nd);
          }
          return;
        }
        const uint64_t e0 = g.off[u], e1 = g.off[u + 1];
        for (uint64_t eb = e0; eb < e1; eb += RELAX_GROUP) {
#pragma unroll
          for (int i = 0; i < RELAX_GROUP; i++) {
            v[i] = 0;
            nd[i] = INFINIT;
            if (eb + i < e1) {
              v[i] = edge_col(g, eb + i);
              const uint64_t x = (uint64_t)du + edge_w(g, eb + i);
              nd[i] = x < INFINIT ? (uint32_t)x : INFINIT;
            }
          }
          settle(v, nd);
        }
      };
      if (threadIdx.x == 0) s_nl = 0;
      __syncthreads();
      // the round's queue is dealt round-robin over the workgroups: a
      // workgroup keeps relaxing what it improves, so with the queue in
      // contiguous per-workgroup blocks a queue of a few thousand vertices
      // would leave every chain of the round to the first few workgroups
      for (uint64_t i = (uint64_t)threadIdx.x * gridDim.x + blockIdx.x; i < nn; i += stride) {
        const uint32_t u = ld_agent(&q[cur][i]);
        relax(u, ld_agent(&p.dist[u]));
      }
      // local passes: the workgroup relaxes the vertices it owes, so the
      // frontier advances many hops per grid barrier.  The pass count is
      // bounded and block-uniform: thread 0 snapshots the ring between two
      // barriers (others append while a pass runs), so every thread sees
      // the same bound and every workgroup reaches the grid barrier.
      for (uint32_t it = 0, lo = 0;; it++) {
        __syncthreads();  // the previous pass's appends done, s_hi read by all
        if (threadIdx.x == 0) {
          // appends that were kept all lie below pass_lo + PL_CAP
          const uint32_t n = s_nl < pass_lo + PL_CAP ? s_nl : pass_lo + PL_CAP;
          s_nl = n;
          s_hi = n;
        }
        __syncthreads();
        const uint32_t hi = s_hi;
        if (hi <= lo) break;
        if (it >= p.local_hops) {  // budget spent: what is still owed goes to the next round
          for (uint32_t k = lo + threadIdx.x; k < hi; k += BLOCK) {
            const uint32_t v = s_loc[k % PL_CAP];
            if (atomicExch(&p.near_stamp[v], near_id) != near_id)
              stage(v, s_near, &s_nn, &C[0], q[cur ^ 1], 0);
          }
          break;
        }
        pass_lo = lo;  // this pass reads [lo, hi)
        for (uint32_t k = lo + threadIdx.x; k < hi; k += BLOCK)
          relax(s_loc[k % PL_CAP], s_locd[k % PL_CAP]);
        lo = hi;
        passes++;
      }
      __syncthreads();
      publish(s_near, s_nn, &C[0], q[cur ^ 1], 0, &s_base);
      __syncthreads();
      {
        // the staged far vertices: every stamp exchange in flight together,
        // then the ones this workgroup claimed first are published
        constexpr uint32_t PER = (PQ_CAP + BLOCK - 1) / BLOCK;
        const uint32_t n = s_nf < PQ_CAP ? s_nf : PQ_CAP;
        uint32_t fv[PER], old[PER], keep = 0;
#pragma unroll
        for (uint32_t r = 0; r < PER; r++) {
          const uint32_t k = r * BLOCK + threadIdx.x;
          fv[r] = k < n ? s_far[k] : 0u;
          old[r] = k < n ? atomicExch(&p.far_stamp[fv[r]], far_id) : far_id;
        }
#pragma unroll
        for (uint32_t r = 0; r < PER; r++) keep += old[r] != far_id;
        uint32_t slot = block_reserve(&C[1], keep, s_scr);
#pragma unroll
        for (uint32_t r = 0; r < PER; r++)
          if (old[r] != far_id) st_agent(&far[fcur][nfar + slot++], fv[r]);
      }
      ok = grid_sync_atomic(&p.ctl->gs, gridDim.x, gen, p.timeout);
      step++;
      nn = ld_agent(&C[0]);
      nfar += ld_agent(&C[1]);
      fmin_known = false;  // the round may have added far vertices below it
      cur ^= 1;
    } else if (nfar > 0) {
      // Advance T by one bucket and split the far list in the same pass.
      // T is a multiple of delta and a far vertex lies at T or beyond, so
      // when the bucket [T, T + delta) holds a far vertex this is the bucket
      // of the smallest far distance.  The split also takes the minimum of
      // what stays far: when the bucket came out empty, the next split goes
      // straight to the bucket of that minimum (valid until a round adds
      // far vertices again).
      const uint32_t Tk = T < 0xFFFFFFFFull ? (uint32_t)T : 0xFFFFFFFFu;
      const uint64_t Tn = fmin_known ? ((uint64_t)fmin / p.delta + 1) * p.delta : T + p.delta;
      const uint32_t Tnk = Tn < 0xFFFFFFFFull ? (uint32_t)Tn : 0xFFFFFFFFu;
      far_id = ++stamp;
      const uint32_t near_id = ++stamp;
      if (threadIdx.x == 0) s_nn = s_nf = 0;
      __syncthreads();
      uint32_t m = 0xFFFFFFFFu;
      for (uint64_t i = (uint64_t)threadIdx.x * gridDim.x + blockIdx.x; i < nfar; i += stride) {
        const uint32_t v = ld_agent(&far[fcur][i]);
        const uint32_t d = ld_agent(&p.dist[v]);
        if (d < Tk) continue;  // relaxed in an earlier bucket at its final distance
        if (d < Tnk) {
          if (atomicExch(&p.near_stamp[v], near_id) != near_id)
            stage(v, s_near, &s_nn, &C[0], q[cur], 0);
        } else {
          m = d < m ? d : m;
          if (atomicExch(&p.far_stamp[v], far_id) != far_id)
            stage(v, s_far, &s_nf, &C[1], far[fcur ^ 1], 0);
        }
      }
      for (int o = 32; o > 0; o >>= 1) {
        const uint32_t x = __shfl_xor(m, o);
        m = x < m ? x : m;
      }
      if (lane_id() == 0 && m != 0xFFFFFFFFu) atomicMin(&C[2], m);
      __syncthreads();
      publish(s_near, s_nn, &C[0], q[cur], 0, &s_base);
      __syncthreads();
      publish(s_far, s_nf, &C[1], far[fcur ^ 1], 0, &s_base);
      ok = grid_sync_atomic(&p.ctl->gs, gridDim.x, gen, p.timeout);
      step++;
      nn = ld_agent(&C[0]);
      nfar = ld_agent(&C[1]);
      fmin = ld_agent(&C[2]);
      fmin_known = nn == 0;
      fcur ^= 1;
      T = Tn;
    } else {
      break;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    p.ctl->rounds = (uint32_t)rounds;
    p.ctl->stamp = stamp;
    p.ctl->status = ok ? 0u : 1u;
    p.ctl->passes = passes;
  }
}

}  // namespace sssp

namespace {
uint32_t pick_delta(Graph &g, SsspWork &w, bool persistent) {
  Context &ctx = *g.ctx;
  if (ctx.tuning.sssp_delta) return ctx.tuning.sssp_delta;
  if (!g.weighted) return 0xFFFFFFFFu;  // unit weights: one bucket (BFS order)
  if (w.mean_w < 0) {
    double mean = 0;
    // the HBM-resident edges are a fair sample in capped placement
    uint64_t n = g.placement == GRUS_PLACE_CAPPED ? g.split : g.E;
    if (n) {
      auto *acc = reinterpret_cast<unsigned long long *>(ctx.d_ctl.p) + 32;
      GRUS_HIP(hipMemsetAsync(acc, 0, 8, ctx.stream));
      sssp::k_weight_sum<<<grid_cap(ctx, n, BLOCK), BLOCK, 0, ctx.stream>>>(
          g.w.as<uint32_t>(), n, acc);
      GRUS_LAUNCH_CHECK();
      ctx.read_ctl(33);
      mean = (double)ctx.h_ctl[32] / (double)n;
    }
    w.mean_w = mean > 0 ? mean : 1.0;
  }
  double avg_deg = g.V ? (double)g.E / (double)g.V : 1.0;
  // the persistent path relaxes a bucket within its round, so its buckets
  // are wider (the 4096^2 grid, mean weight 500: widths 8000-16000 within
  // 4% of the best, 4004 20% slower, 24000 21% slower)
  const double span = persistent ? 80.0 : 32.0;
  double d = w.mean_w * std::max(2.0, span / std::max(avg_deg, 1.0));
  return (uint32_t)std::min(std::max(d, 1.0), 4.0e9);
}
}  // namespace

namespace {
// one cooperative launch for the whole traversal (low max degree)
void run_sssp_persistent(Graph &g, SsspWork &w, uint32_t src, uint64_t delta, uint64_t *rounds,
                         uint64_t *passes) {
  Context &ctx = *g.ctx;
  const uint64_t V = g.V;
  if (w.stamp >= 0xE0000000u) {
    fill_u32(ctx, w.near_stamp.as<uint32_t>(), V + 1, 0);
    fill_u32(ctx, w.far_stamp.as<uint32_t>(), V + 1, 0);
    w.stamp = 0;
  }
  w.pctl.ensure(sizeof(sssp::PCtl));
  sssp::PCtl init{};
  for (auto &slot : init.ring) slot[2] = 0xFFFFFFFFu;
  GRUS_HIP(hipMemcpyAsync(w.pctl.p, &init, sizeof init, hipMemcpyHostToDevice, ctx.stream));
  sssp::PParams p{};
  p.g = g.csr();
  p.dist = g.result.as<uint32_t>();
  p.q0 = w.q[0].as<uint32_t>();
  p.q1 = w.q[1].as<uint32_t>();
  p.far0 = w.far[0].as<uint32_t>();
  p.far1 = w.far[1].as<uint32_t>();
  p.near_stamp = w.near_stamp.as<uint32_t>();
  p.far_stamp = w.far_stamp.as<uint32_t>();
  p.ctl = w.pctl.as<sssp::PCtl>();
  p.delta = delta;
  p.round_cap = std::min<uint64_t>(64 * V + 1024, 0x08000000ull);
  p.timeout = ctx.wall_ticks_per_ms() * 20000;  // 20 s per barrier wait
  p.stamp0 = w.stamp;
  // local passes per round: a bucket's frontier mostly advances within one
  // round, and the round's barrier is paid once per bucket
  p.local_hops = ctx.tuning.sssp_local_hops ? ctx.tuning.sssp_local_hops : 256;
  p.max_deg = (uint32_t)std::max<uint64_t>(g.max_deg, 1);
  // padded rows where the maximum degree is at most 8 (built once per graph;
  // without room under a device budget the CSR rows are read instead)
  if (g.max_deg <= 8 && !w.pad_deg) {
    const uint32_t D = g.max_deg <= 4 ? 4 : 8;
    try {
      w.pad.alloc(V * D * 8);
      sssp::k_pad_rows<<<grid_cap(ctx, V, BLOCK), BLOCK, 0, ctx.stream>>>(g.csr(), D,
                                                                         w.pad.as<uint2>());
      GRUS_LAUNCH_CHECK();
      w.pad_deg = D;
    } catch (const Error &e) {
      if (e.code != GRUS_E_BUDGET) throw;
    }
  }
  p.pad = w.pad_deg ? w.pad.as<uint2>() : nullptr;
  p.pad_deg = w.pad_deg;
  int per_cu = 0;
  GRUS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void *>(sssp::k_persistent), BLOCK, 0));
  if (per_cu < 1) raise(GRUS_E_HIP, "persistent sssp kernel cannot be resident");
  // few workgroups keep the barrier cheap; the frontier of a high-diameter
  // graph is small (GRUS_PERSIST_BLOCKS overrides for experiments)
  // read once; the value is kept, not the environment's pointer (which a
  // later change of the environment may free)
  static const int env_blocks = [] {
    const char *e = getenv("GRUS_PERSIST_BLOCKS");
    return e ? atoi(e) : 0;
  }();
  // CUs / 2 measured best on the 4096^2 grid with the round's queue dealt
  // round-robin (64 / 128 / 256 workgroups: 40.5 / 37.1 / 40.1 ms at their
  // best bucket widths)
  unsigned want = env_blocks > 0 ? (unsigned)env_blocks : (unsigned)std::max(1, ctx.cu_count / 2);
  unsigned grid = std::max(1u, std::min(want, (unsigned)(ctx.cu_count * per_cu)));
  void *args[] = {&p};
  GRUS_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(sssp::k_persistent),
                                      dim3(grid), dim3(BLOCK), args, 0, ctx.stream));
  sssp::PCtl out;
  GRUS_HIP(hipMemcpyAsync(&out, w.pctl.p, sizeof out, hipMemcpyDeviceToHost, ctx.stream));
  ctx.sync();
  w.stamp = out.stamp;
  *rounds = out.rounds;
  *passes = out.passes;
  if (out.status) raise(GRUS_E_TIMEOUT, "persistent sssp aborted (barrier timeout or round cap)");
}
}  // namespace

void run_sssp(Graph &g, uint32_t src, grus_stats *st) {
  Context &ctx = *g.ctx;
  ctx.activate();
  if (src >= g.V) raise(GRUS_E_INVALID, "source vertex out of range");
  g.make_room(Graph::App::SSSP);
  if (!g.sssp) g.sssp = std::make_shared<SsspWork>();
  SsspWork &w = *g.sssp;
  w.ensure(ctx, g.V);
  g.result.ensure(g.V * 4);
  const uint64_t V = g.V;
  const CsrView out = g.csr();
  uint32_t *dist = g.result.as<uint32_t>();
  auto *cnt = reinterpret_cast<StepCounters *>(ctx.d_ctl.p);
  const StepCounters *hc = reinterpret_cast<const StepCounters *>(ctx.h_ctl);
  uint64_t src_deg;
  {
    uint64_t o[2];
    GRUS_HIP(hipMemcpy(o, g.off.as<uint64_t>() + src, 16, hipMemcpyDeviceToHost));
    src_deg = o[1] - o[0];
  }
  uint64_t launches = 0, rounds = 0, relaxed_vertices = 0, pull_rounds = 0, local_passes = 0;
  // low-degree graphs (grids, road-like): one persistent cooperative launch
  static const bool no_persist = getenv("GRUS_NO_PERSISTENT") != nullptr;
  // (not under a CU mask: a cooperative grid must be co-resident, and the
  // runtime's residency check does not know the mask; nor beside other
  // contexts' kernels)
  const bool persistent = g.max_deg <= 32 && !no_persist && !ctx.cu_masked && !ctx.shared_device;
  const uint64_t delta = pick_delta(g, w, persistent);
  // a round pulls when its frontier's out-edges exceed E / pull_div
  // (automatic: 16, measured best of 4..128 on RMAT s24; 0xFFFFFFFF: never)
  const uint32_t pull_div = ctx.tuning.sssp_pull_div ? ctx.tuning.sssp_pull_div : 16;
  const uint64_t pull_edges = pull_div == 0xFFFFFFFFu ? ~0ull : g.E / pull_div;
  const bool can_pull = g.symmetric || g.placement != GRUS_PLACE_CAPPED;  // needs the in-edges
  // pull rounds: binned push where the layout applies (built on first use
  // and kept with the graph), else the window reduction over the in-edges
  const uint32_t pull_layout = ctx.tuning.sssp_pull_layout;
  const bool binned_pull = pull_edges != ~0ull && pull_layout != 1 && !persistent &&
                           g.E > 0 && sssp_binned_ready(g);
  // a run starts far from the stamp wrap, so the stamps of the current
  // queue identify its vertices for the whole run (binned rounds read them
  // instead of a frontier bitmap)
  if (w.stamp >= 0xF0000000u) {
    fill_u32(ctx, w.near_stamp.as<uint32_t>(), V + 1, 0);
    fill_u32(ctx, w.far_stamp.as<uint32_t>(), V + 1, 0);
    w.stamp = 0;
  }
  uint32_t cur_id = w.next_stamp(ctx, V);  // stamp of the queue about to be relaxed
  const uint64_t wraps0 = w.wraps;
  GRUS_HIP(hipEventRecord(ctx.ev0, ctx.stream));
  sssp::k_init<<<grid_cap(ctx, V, BLOCK), BLOCK, 0, ctx.stream>>>(
      dist, V, src, w.q[0].as<uint32_t>(), w.near_stamp.as<uint32_t>(), cur_id);
  GRUS_LAUNCH_CHECK();
  launches++;
  if (persistent) {
    run_sssp_persistent(g, w, src, delta, &rounds, &local_passes);
    launches++;
  }
  int cur = 0, fcur = 0;
  uint64_t nn = persistent ? 0 : 1, mf = src_deg, nfar = 0;
  uint64_t T = delta;
  uint32_t far_id = w.next_stamp(ctx, V);
  const uint64_t round_cap = 64 * V + 1024;
  const uint64_t W = V / 64 + 1;
  for (;;) {
    while (nn) {
      if (++rounds > round_cap) raise(GRUS_E_HIP, "sssp did not converge");
      relaxed_vertices += nn;
      GRUS_HIP(hipMemsetAsync(cnt, 0, sizeof(StepCounters), ctx.stream));
      const uint32_t Tk = (uint32_t)std::min<uint64_t>(T, 0xFFFFFFFFull);
      const uint32_t near_id = w.next_stamp(ctx, V);
      if (binned_pull && mf > pull_edges) {
        // push over the binned layout: sequential streams instead of the
        // window's gathers of dist[u] (one cache line per in-edge).  The
        // frontier is the vertices stamped cur_id (a bitmap only after a
        // stamp wrap, which clears the stamps)
        BinnedMinArgs a{};
        if (w.wraps == wraps0) {
          a.front_stamp = w.near_stamp.as<uint32_t>();
          a.front_id = cur_id;
        } else {
          uint64_t *fbm = w.front_bm.as<uint64_t>();
          fill_u64(ctx, fbm, W, 0);
          queue_to_bitmap(ctx, w.q[cur].as<uint32_t>(), (uint32_t)nn, fbm, W);
          a.front = fbm;
        }
        a.dist = dist;
        a.out_off = out.off;
        a.T_thr = Tk;
        a.near_id = near_id;
        a.far_id = far_id;
        a.near_stamp = w.near_stamp.as<uint32_t>();
        a.far_stamp = w.far_stamp.as<uint32_t>();
        a.qn = w.q[cur ^ 1].as<uint32_t>();
        a.qfar = w.far[fcur].as<uint32_t>() + nfar;
        a.cnt = cnt;
        sssp_binned_round(g, a);
        launches += 5;
        pull_rounds++;
      } else if (can_pull && mf > pull_edges) {
        const CsrView in = g.csc();
        if (w.win_for != in.off) {
          w.nwin = in.E / WIN + 1;
          w.win_row.alloc((w.nwin + 1) * 4);
          w.head_min.alloc(w.nwin * 4);
          w.tail_min.alloc(w.nwin * 4);
          build_window_rows(ctx, in.off, V, w.nwin, w.win_row.as<uint32_t>());
          w.win_for = in.off;
        }
        uint64_t *fbm = w.front_bm.as<uint64_t>();
        fill_u64(ctx, fbm, W, 0);
        queue_to_bitmap(ctx, w.q[cur].as<uint32_t>(), (uint32_t)nn, fbm, W);
        sssp::PullMinOp op{};
        op.in = in;
        op.front = fbm;
        op.dist = dist;
        op.out_off = out.off;
        op.T_thr = Tk;
        op.near_stamp = w.near_stamp.as<uint32_t>();
        op.near_id = near_id;
        op.far_stamp = w.far_stamp.as<uint32_t>();
        op.far_id = far_id;
        op.head_min = w.head_min.as<uint32_t>();
        op.tail_min = w.tail_min.as<uint32_t>();
        op.qn = w.q[cur ^ 1].as<uint32_t>();
        op.qfar = w.far[fcur].as<uint32_t>() + nfar;
        op.cnt = cnt;
        sssp::k_pull_min_window<<<(unsigned)w.nwin, BLOCK, 0, ctx.stream>>>(
            in, w.win_row.as<uint32_t>(), op, w.nwin);
        GRUS_LAUNCH_CHECK();
        sssp::k_pull_min_fixup<<<(unsigned)div_up(w.nwin, BLOCK), BLOCK, 0, ctx.stream>>>(
            in, w.win_row.as<uint32_t>(), op, w.nwin);
        GRUS_LAUNCH_CHECK();
        launches += 4;
        pull_rounds++;
      } else {
        push_scan_degrees(ctx, w.push, out.off, w.q[cur].as<uint32_t>(), (uint32_t)nn);
        sssp::job_t job{dist,
                        w.near_stamp.as<uint32_t>(),
                        w.far_stamp.as<uint32_t>(),
                        out.off,
                        Tk,
                        near_id,
                        far_id};
        push_advance(ctx, w.push, out, w.q[cur].as<uint32_t>(), (uint32_t)nn, mf, job,
                     w.q[cur ^ 1].as<uint32_t>(), cnt, w.far[fcur].as<uint32_t>() + nfar);
        launches += 4;
      }
      ctx.read_ctl(3);
      nn = hc->next_count;
      mf = hc->next_edges;
      nfar += hc->far_count;
      cur ^= 1;
      cur_id = near_id;
    }
    if (nfar == 0) break;
    // advance the threshold to the bucket of the smallest live far distance
    GRUS_HIP(hipMemsetAsync(cnt, 0, sizeof(StepCounters), ctx.stream));
    GRUS_HIP(hipMemsetAsync(&cnt->min_far, 0xFF, 4, ctx.stream));
    uint32_t Tk = (uint32_t)std::min<uint64_t>(T, 0xFFFFFFFFull);
    sssp::k_far_min<<<grid_cap(ctx, nfar, BLOCK), BLOCK, 0, ctx.stream>>>(
        w.far[fcur].as<uint32_t>(), (uint32_t)nfar, dist, Tk, &cnt->min_far);
    GRUS_LAUNCH_CHECK();
    ctx.read_ctl(3);
    uint32_t mn = hc->min_far;
    if (mn == 0xFFFFFFFFu) break;  // every far entry was stale
    uint64_t T_new = ((uint64_t)mn / delta + 1) * delta;
    far_id = w.next_stamp(ctx, V);
    uint32_t near_id = w.next_stamp(ctx, V);
    GRUS_HIP(hipMemsetAsync(cnt, 0, sizeof(StepCounters), ctx.stream));
    sssp::k_far_split<<<grid_cap(ctx, nfar, BLOCK), BLOCK, 0, ctx.stream>>>(
        w.far[fcur].as<uint32_t>(), (uint32_t)nfar, dist, out.off, Tk,
        (uint32_t)std::min<uint64_t>(T_new, 0xFFFFFFFFull), w.near_stamp.as<uint32_t>(),
        near_id, w.far_stamp.as<uint32_t>(), far_id, w.q[cur].as<uint32_t>(),
        w.far[fcur ^ 1].as<uint32_t>(), cnt);
    GRUS_LAUNCH_CHECK();
    launches += 3;
    ctx.read_ctl(3);
    nn = hc->next_count;
    mf = hc->next_edges;
    nfar = hc->far_count;
    fcur ^= 1;
    T = T_new;
    cur_id = near_id;
  }
  GRUS_HIP(hipEventRecord(ctx.ev1, ctx.stream));
  GRUS_HIP(hipEventSynchronize(ctx.ev1));
  float ms = 0;
  GRUS_HIP(hipEventElapsedTime(&ms, ctx.ev0, ctx.ev1));
  if (st) {
    *st = grus_stats{};
    reached_stats(ctx, out.off, dist, V, &st->vertices_reached, &st->edges_reached);
    st->device_ms = ms;
    st->iterations = (uint32_t)rounds;
    st->pull_iterations = (uint32_t)pull_rounds;
    st->kernel_launches = launches;
    st->reserved[0] = relaxed_vertices;
    st->reserved[1] = delta;
    st->reserved[2] = local_passes;
  }
}

}  // namespace mgg
