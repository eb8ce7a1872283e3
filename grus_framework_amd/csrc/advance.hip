// advance.hip -- frontier conversion helpers and the push operator's scan.
//
// bitmap_to_queue replaces flag_to_wl (reference src/worklist.cu:87-93,
// one aggregated atomic per 32-lane group over a byte flag per vertex):
// here one lane owns a 64-vertex bitmap word, a wave prefix-sums the
// popcounts and one atomic per wave reserves the output range, so the
// output keeps ascending order inside each wave.  queue_to_bitmap replaces
// wl_to_flag (src/worklist.cu:94-100) with bit-level atomics.
#include <algorithm>

#include "advance.hpp"
#include "prims.hpp"

namespace mgg {

namespace {

__global__ void k_queue_degrees(const uint64_t *off, const uint32_t *q, uint32_t F,
                                uint64_t *deg) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i <= F; i += stride) {
    uint64_t d = 0;
    if (i < F) {
      uint32_t u = q[i];
      d = off[u + 1] - off[u];
    }
    deg[i] = d;
  }
}

__global__ void k_queue_to_bitmap(const uint32_t *q, uint32_t n, uint64_t *bm) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    uint32_t v = q[i];
    atomicOr((unsigned long long *)&bm[v >> 6], 1ull << (v & 63));
  }
}

// A workgroup owns `items` x BLOCK consecutive bitmap words (items <=
// B2Q_MAX_ITEMS; word i*BLOCK + t of the span goes to thread t, so every
// round of loads is coalesced).  It scans the popcounts of the whole span and
// reserves its queue slots with one atomic: the counter's single address is
// the serial point of the pass, so the fewer workgroups the better once the
// loads are in flight.
constexpr uint32_t B2Q_MAX_ITEMS = 8;

__global__ __launch_bounds__(BLOCK) void k_bitmap_to_queue(const uint64_t *bm, uint64_t W,
                                                            uint32_t items, uint32_t *q,
                                                            uint32_t *count, uint64_t flip) {
  __shared__ uint32_t s_wave[BLOCK / WAVE + 1];
  const uint64_t first = (uint64_t)blockIdx.x * items * BLOCK + threadIdx.x;
  uint64_t m[B2Q_MAX_ITEMS];
  uint32_t c = 0;
#pragma unroll
  for (uint32_t i = 0; i < B2Q_MAX_ITEMS; i++) {
    const uint64_t wd = first + (uint64_t)i * BLOCK;
    m[i] = (i < items && wd < W) ? bm[wd] ^ flip : 0ull;
    c += (uint32_t)__popcll(m[i]);
  }
  uint32_t x = c;  // inclusive wave scan of the popcounts
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    uint32_t y = __shfl_up(x, o);
    if ((int)lane_id() >= o) x += y;
  }
  const uint32_t wid = threadIdx.x / WAVE;
  if (lane_id() == WAVE - 1) s_wave[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (uint32_t i = 0; i < BLOCK / WAVE; i++) {
      uint32_t t = s_wave[i];
      s_wave[i] = s;
      s += t;
    }
    s_wave[BLOCK / WAVE] = s ? atomicAdd(count, s) : 0u;
  }
  __syncthreads();
  uint32_t at = s_wave[BLOCK / WAVE] + s_wave[wid] + (x - c);
#pragma unroll
  for (uint32_t i = 0; i < B2Q_MAX_ITEMS; i++) {
    uint64_t mm = m[i];
    const uint64_t wd = first + (uint64_t)i * BLOCK;
    while (mm) {
      int b = __ffsll((unsigned long long)mm) - 1;
      mm &= mm - 1;
      q[at++] = (uint32_t)(wd * 64 + b);
    }
  }
}

__global__ void k_fill_u32(uint32_t *a, uint64_t n, uint32_t v) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) a[i] = v;
}

__global__ void k_fill_u64(uint64_t *a, uint64_t n, uint64_t v) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) a[i] = v;
}

// reached vertices and their out-edges; one atomic pair per workgroup (a
// pair from every wave onto the same two addresses serialised at the L2:
// 0.4 ms at V = 2^24)
__global__ __launch_bounds__(BLOCK) void k_reached(const uint64_t *off, const uint32_t *val,
                                                   uint64_t V, unsigned long long *acc) {
  __shared__ unsigned long long s_n[BLOCK / WAVE], s_e[BLOCK / WAVE];
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long n = 0, e = 0;
  for (; i < V; i += stride) {
    if (val[i] != INFINIT) {
      n++;
      e += off[i + 1] - off[i];
    }
  }
  n = wave_sum(n);
  e = wave_sum(e);
  const uint32_t wid = threadIdx.x / WAVE;
  if (lane_id() == 0) {
    s_n[wid] = n;
    s_e[wid] = e;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t w = 1; w < BLOCK / WAVE; w++) {
      n += s_n[w];
      e += s_e[w];
    }
    if (n) atomicAdd(&acc[0], n);
    if (e) atomicAdd(&acc[1], e);
  }
}

}  // namespace

void reached_stats(Context &ctx, const uint64_t *off, const uint32_t *val, uint64_t V,
                   uint64_t *vr, uint64_t *er) {
  auto *acc = reinterpret_cast<unsigned long long *>(ctx.d_ctl.p) + 40;
  GRUS_HIP(hipMemsetAsync(acc, 0, 16, ctx.stream));
  if (V) {
    k_reached<<<grid_cap(ctx, V, BLOCK, 4), BLOCK, 0, ctx.stream>>>(off, val, V, acc);
    GRUS_LAUNCH_CHECK();
  }
  ctx.read_ctl(42);
  *vr = ctx.h_ctl[40];
  *er = ctx.h_ctl[41];
}

unsigned grid_cap(Context &ctx, uint64_t items, unsigned per_block, unsigned blocks_per_cu) {
  uint64_t g = div_up(items ? items : 1, per_block);
  uint64_t cap = (uint64_t)ctx.cu_count * blocks_per_cu;
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

unsigned push_grid(Context &ctx, uint64_t total) { return grid_cap(ctx, total, LB_CHUNK, 8); }

#ifndef GATHER_PER_CU
#define GATHER_PER_CU 2
#endif

// A workgroup copies LB_CHUNK consecutive positions: thread 0 finds the
// queue entries that own the chunk's ends, the owners' scan values and row
// starts are staged in LDS, and each thread copies LB_CHUNK / BLOCK edges in
// coalesced rounds (rows are contiguous, so consecutive positions of one
// row are consecutive reads).  A thread issues its LB_PER reads before its
// first write, so that many reads are in flight over the host link while
// the grid stays small (GATHER_PER_CU workgroups per CU) and the expansion
// of the previous piece, on another stream, finds room beside it.
__global__ __launch_bounds__(BLOCK) void k_gather_rows(CsrView g, const uint32_t *q,
                                                       const uint64_t *S, uint32_t F, uint64_t k0,
                                                       uint64_t k1, uint32_t *col, uint32_t *w) {
  __shared__ uint64_t sS[LB_WIN];
  __shared__ uint64_t sRow[LB_WIN];
  __shared__ uint32_t s_i0, s_n;
  for (uint64_t c0 = k0 + (uint64_t)blockIdx.x * LB_CHUNK; c0 < k1;
       c0 += (uint64_t)gridDim.x * LB_CHUNK) {
    const uint64_t c1 = c0 + LB_CHUNK < k1 ? c0 + LB_CHUNK : k1;
    if (threadIdx.x == 0) {
      const uint32_t i0 = upper_bound_u64(S, 0, F + 1, c0) - 1;
      const uint32_t i1 = upper_bound_u64(S, i0, F + 1, c1 - 1) - 1;
      s_i0 = i0;
      s_n = i1 - i0 + 2;
    }
    __syncthreads();
    const uint32_t i0 = s_i0, n = s_n;
    const bool staged = n <= LB_WIN;
    if (staged) {
      for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) {
        sS[t] = S[i0 + t];
        if (i0 + t < F) sRow[t] = g.off[q[i0 + t]];
      }
    }
    __syncthreads();
    uint64_t e[LB_PER];
    uint32_t c[LB_PER], wv[LB_PER];
#pragma unroll
    for (uint32_t r = 0; r < LB_PER; r++) {
      const uint64_t k = c0 + threadIdx.x + (uint64_t)r * BLOCK;
      e[r] = 0;
      if (k < c1) {
        if (staged) {
          const uint32_t j = upper_bound_u64(sS, 0, n, k) - 1;
          e[r] = sRow[j] + (k - sS[j]);
        } else {
          const uint32_t j = upper_bound_u64(S + i0, 0, n, k) - 1;
          e[r] = g.off[q[i0 + j]] + (k - S[i0 + j]);
        }
      }
    }
#pragma unroll
    for (uint32_t r = 0; r < LB_PER; r++) {
      const bool in = c0 + threadIdx.x + (uint64_t)r * BLOCK < c1;
      c[r] = in ? edge_col(g, e[r]) : 0;
      wv[r] = in && w ? edge_w(g, e[r]) : 0;
    }
#pragma unroll
    for (uint32_t r = 0; r < LB_PER; r++) {
      const uint64_t k = c0 + threadIdx.x + (uint64_t)r * BLOCK;
      if (k < c1) {
        col[k - k0] = c[r];
        if (w) w[k - k0] = wv[r];
      }
    }
    __syncthreads();  // sS / sRow reused by the next chunk
  }
}

void gather_rows(Context &ctx, const CsrView &g, const uint32_t *q, const uint64_t *S, uint32_t F,
                 uint64_t k0, uint64_t k1, uint32_t *col, uint32_t *w, hipStream_t stream) {
  if (F == 0 || k1 <= k0) return;
  k_gather_rows<<<grid_cap(ctx, k1 - k0, LB_CHUNK, GATHER_PER_CU), BLOCK, 0, stream>>>(
      g, q, S, F, k0, k1, col, w);
  GRUS_LAUNCH_CHECK();
}

void push_scan_degrees(Context &ctx, PushWork &w, const uint64_t *off, const uint32_t *q,
                       uint32_t F) {
  k_queue_degrees<<<grid_cap(ctx, (uint64_t)F + 1, BLOCK), BLOCK, 0, ctx.stream>>>(
      off, q, F, w.deg_ptr());
  GRUS_LAUNCH_CHECK();
  if (w.tmp) scan_excl_u64_with(ctx, w.deg_ptr(), w.scan_ptr(), (size_t)F + 1, w.tmp, w.tmp_bytes);
  else scan_excl_u64(ctx, w.deg_ptr(), w.scan_ptr(), (size_t)F + 1);
}

// about one workgroup per CU, up to B2Q_MAX_ITEMS words per thread
static void launch_bitmap_to_queue(Context &ctx, const uint64_t *bm, uint64_t W, uint32_t *q,
                                   uint32_t *count, uint64_t flip) {
  const uint64_t per_round = (uint64_t)ctx.cu_count * BLOCK;
  const uint32_t items = (uint32_t)std::min<uint64_t>(
      B2Q_MAX_ITEMS, std::max<uint64_t>(1, (W + per_round - 1) / per_round));
  const uint64_t span = (uint64_t)items * BLOCK;
  const unsigned grid = (unsigned)std::max<uint64_t>(1, (W + span - 1) / span);
  k_bitmap_to_queue<<<grid, BLOCK, 0, ctx.stream>>>(bm, W, items, q, count, flip);
}

void bitmap_to_queue(Context &ctx, const uint64_t *bm, uint64_t W, uint32_t *q, uint32_t *count) {
  launch_bitmap_to_queue(ctx, bm, W, q, count, 0ull);
  GRUS_LAUNCH_CHECK();
}

void clear_bits_to_queue(Context &ctx, const uint64_t *bm, uint64_t W, uint32_t *q,
                         uint32_t *count) {
  launch_bitmap_to_queue(ctx, bm, W, q, count, ~0ull);
  GRUS_LAUNCH_CHECK();
}

void queue_to_bitmap(Context &ctx, const uint32_t *q, uint32_t n, uint64_t *bm, uint64_t W) {
  fill_u64(ctx, bm, W, 0);
  if (n) {
    k_queue_to_bitmap<<<grid_cap(ctx, n, BLOCK), BLOCK, 0, ctx.stream>>>(q, n, bm);
    GRUS_LAUNCH_CHECK();
  }
}

void fill_u32(Context &ctx, uint32_t *a, uint64_t n, uint32_t v) {
  if (!n) return;
  k_fill_u32<<<grid_cap(ctx, n, BLOCK), BLOCK, 0, ctx.stream>>>(a, n, v);
  GRUS_LAUNCH_CHECK();
}

void fill_u64(Context &ctx, uint64_t *a, uint64_t n, uint64_t v) {
  if (!n) return;
  k_fill_u64<<<grid_cap(ctx, n, BLOCK), BLOCK, 0, ctx.stream>>>(a, n, v);
  GRUS_LAUNCH_CHECK();
}

}  // namespace mgg
