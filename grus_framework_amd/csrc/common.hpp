// common.hpp -- types, status codes and error plumbing shared by every
// translation unit of libgrus_hip.so.
//
// Types follow include/common.cuh:47-52 of the reference (vtx_t = weight_t =
// uint32, INFINIT = UINT_MAX-1) except that edge offsets are uint64: the
// reference's uint32 xadj (include/graph_loader.cuh:118) and uint32
// edge_id (include/kernel.cuh:32,54) overflow past 2^32 edges.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/grus_c.h"

namespace mgg {

using vtx_t = uint32_t;
using weight_t = uint32_t;
using eid_t = uint64_t;

constexpr uint32_t INFINIT = GRUS_INFINIT;
// "no neighbour": the first-neighbour arrays hold it for empty rows
#define GRUS_NO_VERTEX 0xFFFFFFFFu
constexpr int WAVE = 64;           // CDNA wavefront width
constexpr int BLOCK = 256;         // default workgroup: 4 waves
constexpr int NXCD = 8;            // MI355X: 8 XCDs, each with its own L2

// Error carried from the engine to the C-ABI, which turns it into a status
// code plus a thread-local message (the reference prints and exits instead,
// include/common.cuh:57-68).
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void raise(int code, const std::string &msg) { throw Error(code, msg); }

inline void hip_check(hipError_t e, const char *what, const char *file, int line) {
  if (e == hipSuccess) return;
  (void)hipGetLastError();  // clear the sticky non-fatal state for the next call
  int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? GRUS_E_NOMEM
             : (e == hipErrorNoDevice || e == hipErrorInvalidDevice)     ? GRUS_E_NODEVICE
                                                                           : GRUS_E_HIP;
  char buf[512];
  snprintf(buf, sizeof buf, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  throw Error(code, buf);
}
#define GRUS_HIP(call) ::mgg::hip_check((call), #call, __FILE__, __LINE__)
#define GRUS_LAUNCH_CHECK() ::mgg::hip_check(hipGetLastError(), "kernel launch", __FILE__, __LINE__)

inline uint64_t div_up(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

}  // namespace mgg
