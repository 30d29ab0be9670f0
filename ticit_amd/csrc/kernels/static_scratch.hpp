// Device scratch held by the library for the whole process: the staging of
// ticit_hip_sample_device_postselected_noalloc, which allocates nothing.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace ticit {

// Bytes of the file-scope device array (one copy per GPU the library runs on).
constexpr size_t kStaticScratchBytes = size_t(32) << 20;

// Address of that array on `device`, which must be the current device.
// Looked up once per device with hipGetSymbolAddress and kept; the array's
// host-side name is not a device address.
hipError_t static_scratch_address(int device, void** out);

}  // namespace ticit
