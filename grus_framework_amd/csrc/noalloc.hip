// noalloc.hip -- the scratch area behind the allocation-free entry point.
//
// No reference counterpart.  Some callers cannot allocate device memory
// (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc).  grus_cc_noalloc gives them grus_cc's computation with every
// device buffer beyond the graph and the caller's output carved from this
// file-scope array, which each device that loads the library holds until the
// process ends.
//
// In host code `scratch` is a host-side stand-in, not a device address:
// passing it to a kernel would make the kernel fault.  scratch_setup() asks
// the runtime for the device address of the array once per device, with
// that device current (the library's init call, grus_init, runs it), and
// every later use goes through that address.  Calls on one device share the
// array and see the last call's contents, so whatever must start at zero is
// cleared by the call itself; two calls that overlap on one device race.
#include <mutex>

#include "apps.hpp"
#include "graph.hpp"

namespace mgg {

namespace {
constexpr uint64_t kScratchBytes = GRUS_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDev = 64;
std::mutex g_mu;
void *g_scratch[kMaxDev];  // device address of `scratch` on each device
}  // namespace

void scratch_setup(int device) {
  if (device < 0 || device >= kMaxDev) raise(GRUS_E_NODEVICE, "device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return;
  int cur = -1;
  GRUS_HIP(hipGetDevice(&cur));
  if (cur != device) raise(GRUS_E_HIP, "scratch_setup: the device must be current");
  void *p = nullptr;
  GRUS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) raise(GRUS_E_HIP, "no device address for the scratch area");
  g_scratch[device] = p;
}

void run_cc_noalloc(Graph &g, int method, uint32_t *labels_dev, grus_stats *st) {
  const int dev = g.ctx->device;
  void *p = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (dev >= 0 && dev < kMaxDev) p = g_scratch[dev];
  }
  if (!p) raise(GRUS_E_HIP, "the scratch area of this device is not set up");
  run_cc_in(g, method, labels_dev, p, kScratchBytes, st);
}

}  // namespace mgg
