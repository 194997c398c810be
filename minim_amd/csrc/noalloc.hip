// noalloc.hip -- the allocation-free evaluation entry point's scratch memory.
//
// minim_potential_eval_async_noalloc reduces into this file-scope device
// array instead of the per-thread workspace (which is allocated on first use),
// so that callers who cannot allocate device memory (code under HIP graph
// capture, tight serving loops) can evaluate.  Each GPU holds its own copy of
// the array for the life of the process; its device address is looked up once
// per device (hipGetSymbolAddress, on the entry point's first call there) and
// that address, never the array's host-side name, is what kernels receive.
// Every call on one GPU shares the array: calls that overlap on one GPU race.
// Nothing in it needs to start at zero (every partial sum is written before it
// is read, and the final sums overwrite the result slots).
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "common.hpp"

namespace minim_amd {

// partial sums, then the kMaxDots result slots
constexpr size_t kNoAllocPartials = 32768;
constexpr size_t kNoAllocDoubles = kNoAllocPartials + kMaxDots;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {
double* g_scratch[kNoAllocMaxDevices] = {};
std::once_flag g_once[kNoAllocMaxDevices];
}  // namespace

// This device's copy of the array (the current device), looked up once.
double* noalloc_scratch(size_t* nPartials) {
  int dev = 0;
  MINIM_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) throw std::invalid_argument("noalloc: device index out of range");
  std::call_once(g_once[dev], [dev] {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      throw HipError(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch[dev] = static_cast<double*>(p);
  });
  if (!g_scratch[dev]) throw HipError("noalloc: scratch address unavailable on this device");
  *nPartials = kNoAllocPartials;
  return g_scratch[dev];
}

}  // namespace minim_amd
