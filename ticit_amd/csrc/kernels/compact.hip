// Survivor compaction for postselected programs: the rows of the shots whose
// keep flag is 1 move to the front of the output, in shot order, and the
// number of survivors is written to a device word. The reference appends only
// the survivors' rows of every block (src/sampler/prepared.rs:881-1028,
// src/sampler/batch/postselect.rs:450-611 compacts dead shots out of a
// block); with one shot per stream the survivors' order is the shot order.
//
// Three passes over tiles of kCompactTile shots: per-tile survivor counts,
// one exclusive scan of the counts (which also yields the total), and a
// byte-parallel copy of each tile's surviving rows. Reads and writes are
// contiguous runs of row bytes, so the copy streams at HBM rate.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "compact.hpp"

namespace ticit {

namespace {

constexpr int kCompactThreads = 256;
constexpr int kFlagsPerThread = kCompactTile / kCompactThreads;  // 8

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds, uint32_t* total) {
  // Hillis-Steele over kCompactThreads values in LDS
  const int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (int off = 1; off < kCompactThreads; off <<= 1) {
    uint32_t add = t >= off ? lds[t - off] : 0u;
    __syncthreads();
    lds[t] += add;
    __syncthreads();
  }
  const uint32_t inclusive = lds[t];
  *total = lds[kCompactThreads - 1];
  __syncthreads();
  return inclusive - v;
}

__global__ __launch_bounds__(kCompactThreads) void keep_tile_count_kernel(const uint8_t* keep, uint64_t shots,
                                                                          uint64_t* tile_off) {
  __shared__ uint32_t lds[kCompactThreads];
  const uint64_t base = uint64_t(blockIdx.x) * kCompactTile + uint64_t(threadIdx.x) * kFlagsPerThread;
  uint32_t n = 0;
  for (int i = 0; i < kFlagsPerThread; ++i)
    if (base + i < shots) n += keep[base + i] != 0;
  uint32_t total = 0;
  (void)block_exclusive_scan(n, lds, &total);
  if (threadIdx.x == 0) tile_off[blockIdx.x] = total;
}

// One block: tile_off[0..ntiles) counts -> exclusive offsets; the total goes
// to tile_off[ntiles] and to *accepted (when given).
__global__ __launch_bounds__(kCompactThreads) void keep_tile_scan_kernel(uint64_t* tile_off, uint32_t ntiles,
                                                                         unsigned long long* accepted) {
  __shared__ uint64_t lds[kCompactThreads];
  __shared__ uint64_t carry;
  const int t = threadIdx.x;
  if (t == 0) carry = 0;
  __syncthreads();
  for (uint32_t c0 = 0; c0 < ntiles; c0 += kCompactThreads) {
    const uint32_t i = c0 + uint32_t(t);
    const uint64_t v = i < ntiles ? tile_off[i] : 0ull;
    lds[t] = v;
    __syncthreads();
    for (int off = 1; off < kCompactThreads; off <<= 1) {
      uint64_t add = t >= off ? lds[t - off] : 0ull;
      __syncthreads();
      lds[t] += add;
      __syncthreads();
    }
    if (i < ntiles) tile_off[i] = carry + lds[t] - v;
    __syncthreads();
    if (t == 0) carry += lds[kCompactThreads - 1];
    __syncthreads();
  }
  if (t == 0) {
    tile_off[ntiles] = carry;
    if (accepted) *accepted = carry;
  }
}

// grid (ntiles, nsets): set y copies its rows of this tile's survivors.
__global__ __launch_bounds__(kCompactThreads) void compact_rows_kernel(const uint8_t* keep, uint64_t shots,
                                                                       const uint64_t* tile_off,
                                                                       CompactRows sets0, CompactRows sets1,
                                                                       CompactRows sets2, CompactRows sets3) {
  __shared__ uint32_t lds[kCompactThreads];
  __shared__ uint16_t slot[kCompactTile];  // survivor index within the tile, 0xffff = discarded
  const CompactRows set = blockIdx.y == 0 ? sets0 : blockIdx.y == 1 ? sets1 : blockIdx.y == 2 ? sets2 : sets3;
  if (set.row_bytes == 0) return;  // uniform per block
  const uint64_t tile0 = uint64_t(blockIdx.x) * kCompactTile;
  const uint32_t tile_shots = uint32_t(shots - tile0 < uint64_t(kCompactTile) ? shots - tile0 : kCompactTile);
  const uint32_t first = uint32_t(threadIdx.x) * kFlagsPerThread;
  uint8_t flags[kFlagsPerThread];
  uint32_t n = 0;
  for (int i = 0; i < kFlagsPerThread; ++i) {
    flags[i] = first + i < tile_shots ? uint8_t(keep[tile0 + first + i] != 0) : uint8_t(0);
    n += flags[i];
  }
  uint32_t total = 0;
  uint32_t pre = block_exclusive_scan(n, lds, &total);
  for (int i = 0; i < kFlagsPerThread; ++i) {
    slot[first + i] = flags[i] ? uint16_t(pre) : uint16_t(0xffff);
    pre += flags[i];
  }
  __syncthreads();
  if (total == 0) return;  // uniform
  // row_bytes <= 2^20 (checked by the launcher): a tile spans < 2^31 bytes
  const uint32_t rb = set.row_bytes;
  uint8_t* dst = set.dst + tile_off[blockIdx.x] * uint64_t(rb);
  const uint8_t* src = set.src + tile0 * uint64_t(rb);
  const uint32_t nbytes = tile_shots * rb;
  for (uint32_t b = threadIdx.x; b < nbytes; b += kCompactThreads) {
    const uint32_t s = b / rb;
    const uint16_t q = slot[s];
    if (q != 0xffff) dst[uint64_t(q) * rb + (b - s * rb)] = src[b];
  }
}

__global__ void store_u64_kernel(unsigned long long* dst, unsigned long long v) { *dst = v; }

}  // namespace

hipError_t launch_store_u64(unsigned long long* dst, uint64_t value, hipStream_t stream) {
  hipLaunchKernelGGL(store_u64_kernel, dim3(1), dim3(1), 0, stream, dst, static_cast<unsigned long long>(value));
  return hipGetLastError();
}

size_t compaction_scratch_words(uint64_t shots) { return size_t((shots + kCompactTile - 1) / kCompactTile) + 1; }

hipError_t launch_compaction(const uint8_t* keep, uint64_t shots, uint64_t* tile_off, unsigned long long* accepted,
                             const CompactRows* sets, int nsets, hipStream_t stream) {
  if (shots == 0) return hipSuccess;
  const uint64_t ntiles64 = (shots + kCompactTile - 1) / kCompactTile;
  if (ntiles64 > 0x7fffffffull || nsets > 4) return hipErrorInvalidValue;
  for (int i = 0; i < nsets; ++i)
    if (sets[i].row_bytes > kCompactMaxRowBytes) return hipErrorInvalidValue;
  const uint32_t ntiles = uint32_t(ntiles64);
  hipLaunchKernelGGL(keep_tile_count_kernel, dim3(ntiles), dim3(kCompactThreads), 0, stream, keep, shots, tile_off);
  hipLaunchKernelGGL(keep_tile_scan_kernel, dim3(1), dim3(kCompactThreads), 0, stream, tile_off, ntiles, accepted);
  if (nsets > 0) {
    CompactRows s[4] = {};
    for (int i = 0; i < nsets; ++i) s[i] = sets[i];
    hipLaunchKernelGGL(compact_rows_kernel, dim3(ntiles, uint32_t(nsets)), dim3(kCompactThreads), 0, stream, keep,
                       shots, tile_off, s[0], s[1], s[2], s[3]);
  }
  return hipGetLastError();
}

}  // namespace ticit
