// File-scope device scratch for the sampling entry point that allocates
// nothing (ticit_hip_sample_device_postselected_noalloc). It stages one
// postselected launch: keep flags, compaction tile offsets and the rows of
// every shot before the survivors move to the caller's buffers.
//
// Every GPU the library runs on holds its own copy of `scratch` until the
// process ends. All calls on one GPU share it, so calls that overlap on one
// GPU (other host threads, or other streams not ordered with each other)
// race on its contents.
#include "static_scratch.hpp"

#include <cstdint>
#include <map>
#include <mutex>

namespace ticit {

static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

hipError_t static_scratch_address(int device, void** out) {
  static std::mutex mu;
  static std::map<int, void*> addr;  // per device: each GPU has its own copy at its own address
  std::lock_guard<std::mutex> lock(mu);
  auto it = addr.find(device);
  if (it != addr.end()) {
    *out = it->second;
    return hipSuccess;
  }
  int current = -1;
  hipError_t e = hipGetDevice(&current);
  if (e != hipSuccess) return e;
  if (current != device) return hipErrorInvalidDevice;
  void* p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!p) return hipErrorInvalidSymbol;
  addr[device] = p;
  *out = p;
  return hipSuccess;
}

}  // namespace ticit
