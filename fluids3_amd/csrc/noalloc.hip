// noalloc.hip -- device scratch of fx_scan_device_noalloc, the entry point
// that allocates nothing (for callers that cannot call hipMalloc: code under
// HIP graph capture, tight serving loops).
//
// The array below holds the block totals of the scan's upper levels for up to
// FX_SCAN_NOALLOC_MAX elements (about 1 MB).  Each GPU the library runs on
// holds its own copy until the process ends.  In host code the array's name is
// only a stand-in: its device address comes from hipGetSymbolAddress, taken
// once per device with that device current (scratch_setup, called by
// fx_create) and kept per device.  Every call on one GPU shares that GPU's
// array, so calls that overlap on one GPU race.  Nothing in it must start at
// zero: each scan level writes every total the next level reads.
#include <hip/hip_runtime.h>

#include <mutex>

#include "fx_common.hpp"
#include "noalloc.hpp"

static __device__ __align__(256) uint32_t scratch[FX_SCAN_SCRATCH_WORDS];

namespace {
std::mutex g_mu;
void* g_ptr[FX_MAX_DEVICES] = {};
}  // namespace

int fx::scratch_setup(void** out) {
  int dev = 0;
  FX_HIP(hipGetDevice(&dev), "scratch_setup");
  if (dev < 0 || dev >= FX_MAX_DEVICES) return fail(FX_ERANGE, "scratch_setup: device %d beyond %d", dev, FX_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_ptr[dev] == nullptr) {
    void* p = nullptr;
    FX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch_setup");
    if (p == nullptr) return fail(FX_EHIP, "scratch_setup: no device address for the scratch array");
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return FX_OK;
}
