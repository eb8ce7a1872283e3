// device.hpp -- wave64 intrinsics and bitmap frontier primitives for gfx950.
//
// atomicAggInc (reference include/intrinsics.cuh:96-103) aggregates one
// atomicAdd per 32-lane coalesced group through cooperative_groups; here the
// same slot allocation is a 64-lane ballot + popcount + mbcnt, one atomic
// per wave, with the leader's base broadcast by a lane shuffle.
#pragma once

#include "common.hpp"

namespace mgg {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// XCD-aware workgroup order.  Workgroups are dispatched round-robin over
// the XCDs, so block b runs on XCD b % NXCD.  Giving XCD x the contiguous
// item range [x*per, (x+1)*per) keeps the working set of neighbouring items
// in that XCD's own L2 instead of spreading it over all eight.  gridDim.x
// must be a multiple of NXCD; the map is a bijection on [0, gridDim.x).
__device__ __forceinline__ uint64_t xcd_block() {
  const uint64_t per = gridDim.x / NXCD;
  return (uint64_t)(blockIdx.x % NXCD) * per + blockIdx.x / NXCD;
}

// number of set bits of m in lanes below this one
__device__ __forceinline__ uint32_t lane_rank(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Wave-aggregated append: every active lane of the wave must call it.
// Returns the slot for lanes with pred set, UINT32_MAX otherwise.
__device__ __forceinline__ uint32_t wave_append(uint32_t *count, bool pred) {
  uint64_t m = __ballot(pred);
  if (m == 0) return 0xFFFFFFFFu;
  int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if ((int)lane_id() == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  return pred ? base + lane_rank(m) : 0xFFFFFFFFu;
}

// Workgroup-aggregated append: every thread of the workgroup must call it
// (block-uniform control flow).  One atomic per workgroup; lanes get their
// slot in wave order, lane order within a wave.  `lds` holds >= waves + 1
// u32 of shared scratch.  Returns UINT32_MAX for threads with pred unset.
__device__ __forceinline__ uint32_t block_append(uint32_t *count, bool pred, uint32_t *lds) {
  const uint32_t nwaves = blockDim.x / WAVE;
  const uint32_t wid = threadIdx.x / WAVE;
  uint64_t m = __ballot(pred);
  if (lane_id() == 0) lds[wid] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (uint32_t i = 0; i < nwaves; i++) {
      uint32_t c = lds[i];
      lds[i] = s;
      s += c;
    }
    lds[nwaves] = s ? atomicAdd(count, s) : 0u;
  }
  __syncthreads();
  uint32_t slot = lds[nwaves] + lds[wid] + lane_rank(m);
  __syncthreads();  // lds reusable by the next call
  return pred ? slot : 0xFFFFFFFFu;
}

// Reserves n slots per thread at *count with one atomic per workgroup and
// returns the thread's first slot (a workgroup's slots are contiguous, in
// thread order).  Every thread of the workgroup must call it; lds holds
// blockDim.x / WAVE + 1 entries.
__device__ __forceinline__ uint32_t block_reserve(uint32_t *count, uint32_t n, uint32_t *lds) {
  const uint32_t nwaves = blockDim.x / WAVE;
  const uint32_t wid = threadIdx.x / WAVE;
  uint32_t x = n;  // inclusive wave scan
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if ((int)lane_id() >= o) x += y;
  }
  if (lane_id() == WAVE - 1) lds[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (uint32_t i = 0; i < nwaves; i++) {
      const uint32_t c = lds[i];
      lds[i] = s;
      s += c;
    }
    lds[nwaves] = s ? atomicAdd(count, s) : 0u;
  }
  __syncthreads();
  const uint32_t slot = lds[nwaves] + lds[wid] + (x - n);
  __syncthreads();  // lds reusable by the next call
  return slot;
}

template <class T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

// Adds every thread's (count, sum) to two device counters with one atomic
// pair per workgroup (skipped when the count is zero).  Every thread of the
// workgroup must call it, at the end of the kernel; blockDim.x <= 1024.
// An atomic pair from every wave lands on the same two addresses and
// serialises at the L2.
__device__ __forceinline__ void block_add_counts(uint32_t count, unsigned long long sum,
                                                 uint32_t *count_out,
                                                 unsigned long long *sum_out) {
  __shared__ uint32_t s_c[1024 / WAVE];
  __shared__ unsigned long long s_s[1024 / WAVE];
  count = wave_sum(count);
  sum = wave_sum(sum);
  const uint32_t wid = threadIdx.x / WAVE, nwaves = blockDim.x / WAVE;
  if (lane_id() == 0) {
    s_c[wid] = count;
    s_s[wid] = sum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t w = 1; w < nwaves; w++) {
      count += s_c[w];
      sum += s_s[w];
    }
    if (count) {
      atomicAdd(count_out, count);
      atomicAdd(sum_out, sum);
    }
  }
}

// The same for one 64-bit sum (skipped when zero).
__device__ __forceinline__ void block_add_u64(unsigned long long sum, unsigned long long *out) {
  __shared__ unsigned long long s_s[1024 / WAVE];
  sum = wave_sum(sum);
  const uint32_t wid = threadIdx.x / WAVE, nwaves = blockDim.x / WAVE;
  if (lane_id() == 0) s_s[wid] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t w = 1; w < nwaves; w++) sum += s_s[w];
    if (sum) atomicAdd(out, sum);
  }
}

__device__ __forceinline__ bool bm_test(const uint64_t *bm, uint32_t v) {
  return (bm[v >> 6] >> (v & 63)) & 1ull;
}

// Sets bit v; true when this call flipped it from 0 to 1.
__device__ __forceinline__ bool bm_claim(uint64_t *bm, uint32_t v) {
  unsigned long long bit = 1ull << (v & 63);
  unsigned long long old = atomicOr((unsigned long long *)&bm[v >> 6], bit);
  return (old & bit) == 0;
}

__device__ __forceinline__ uint32_t atomic_min_u32(uint32_t *p, uint32_t v) {
  return atomicMin(p, v);
}

// 64-bit splitmix64 output function (Steele, Lea, Flood 2014).
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace mgg
