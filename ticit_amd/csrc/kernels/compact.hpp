// Survivor compaction of postselected rows (compact.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace ticit {

constexpr int kCompactTile = 2048;               // shots per block
constexpr uint32_t kCompactMaxRowBytes = 1u << 20;

// One set of rows: src [shots][row_bytes] -> dst [survivors][row_bytes].
struct CompactRows {
  const uint8_t* src;
  uint8_t* dst;
  uint32_t row_bytes;  // 0: skip the set
  uint32_t pad;
};

// uint64 words of tile offsets launch_compaction needs for `shots`.
size_t compaction_scratch_words(uint64_t shots);

// Counts keep[0..shots) != 0 into *accepted (device, may be null) and, for
// each of up to four row sets, copies the survivors' rows to the front of
// dst in shot order. tile_off: compaction_scratch_words(shots) device words.
hipError_t launch_compaction(const uint8_t* keep, uint64_t shots, uint64_t* tile_off, unsigned long long* accepted,
                             const CompactRows* sets, int nsets, hipStream_t stream);

// *dst = value, as a kernel on `stream` (no memset node in captured work).
hipError_t launch_store_u64(unsigned long long* dst, uint64_t value, hipStream_t stream);

}  // namespace ticit
