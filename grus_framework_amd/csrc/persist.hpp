// persist.hpp -- grid-wide barrier for persistent (cooperative) kernels.
//
// The reference returns to the host after every level (three blocking
// device-to-host copies per level, src/worklist.cu:64-68, kernel.cuh:162).
// High-diameter inputs (a 4096x4096 grid has hop diameter 8190) then spend
// their time in launch and synchronisation latency.  A persistent kernel is
// launched cooperatively (all workgroups co-resident, which the runtime
// verifies) and separates rounds with this barrier instead.
//
// Every wait is bounded: a workgroup that waits longer than `timeout`
// wall-clock ticks sets the abort flag and leaves; all workgroups test the
// flag after each barrier and exit their loops, so the grid always drains
// and the host reports GRUS_E_TIMEOUT instead of hanging.
#pragma once

#include "device.hpp"

namespace mgg {

struct GridSync {
  unsigned int arrive;
  unsigned int gen;
  unsigned int abort;
  unsigned int pad;
};

// Returns false when the run was aborted (timeout).  All threads of every
// workgroup must call it the same number of times.
__device__ __forceinline__ bool grid_sync(GridSync *gs, unsigned int nblocks, unsigned int &gen,
                                          uint64_t timeout) {
  __shared__ unsigned int s_abort;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();  // release this workgroup's writes (agent scope)
    unsigned int g = gen;
    unsigned int a = __hip_atomic_fetch_add(&gs->arrive, 1u, __ATOMIC_ACQ_REL,
                                            __HIP_MEMORY_SCOPE_AGENT);
    if (a == nblocks - 1) {
      __hip_atomic_store(&gs->arrive, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&gs->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      uint64_t t0 = wall_clock64();
      while (__hip_atomic_load(&gs->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
        if (__hip_atomic_load(&gs->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        if (wall_clock64() - t0 > timeout) {
          __hip_atomic_store(&gs->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    gen = g + 1;
    __threadfence();  // acquire: later reads see the other workgroups' writes
    s_abort = __hip_atomic_load(&gs->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  return s_abort == 0;
}

// Barrier without cache maintenance.  Valid when every datum that crosses
// the barrier is written and read with agent-scope atomics (ld_agent,
// st_agent, atomicMin/Exch/Add), which are coherent across the XCDs' L2s;
// read-only graph arrays are read normally.  An agent-scope fence, by
// contrast, writes back and invalidates an XCD's whole L2 and costs tens of
// microseconds per round when every workgroup issues one.
//
// __syncthreads alone does not order them: for gfx950 it compiles to
// s_barrier without waiting on the vector memory counter, so a wave's
// agent-scope stores (the queue entries written by publish()) may still be
// in flight when thread 0's arrival completes the barrier for another
// workgroup.  Every wave therefore waits for all of its outstanding memory
// operations (s_waitcnt with every counter at zero) before the workgroup
// barrier; a completed agent-scope store is visible to the agent-scope
// loads other workgroups issue after the grid barrier.
__device__ __forceinline__ bool grid_sync_atomic(GridSync *gs, unsigned int nblocks,
                                                 unsigned int &gen, uint64_t timeout) {
  __shared__ unsigned int s_abort;
  __builtin_amdgcn_s_waitcnt(0);  // vmcnt = expcnt = lgkmcnt = 0
  __syncthreads();
  if (threadIdx.x == 0) {
    // monotonic arrival count (never reset): barrier number `gen` completes
    // when (gen + 1) * nblocks workgroups have arrived in total
    const unsigned int target = (gen + 1) * nblocks;
    __hip_atomic_fetch_add(&gs->arrive, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint64_t t0 = wall_clock64();
    while ((int)(__hip_atomic_load(&gs->arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) -
                 target) < 0) {
      if (__hip_atomic_load(&gs->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
      if (wall_clock64() - t0 > timeout) {
        __hip_atomic_store(&gs->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    gen++;
    s_abort = __hip_atomic_load(&gs->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  return s_abort == 0;
}

__device__ __forceinline__ uint32_t ld_agent(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_acquire(const uint32_t *p) { return ld_agent(p); }

}  // namespace mgg
