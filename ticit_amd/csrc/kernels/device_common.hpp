// Device helpers shared by the sampling kernels: SplitMix64 streams, the
// exogenous noise walk of one shot, and instruction-expression evaluation.
//
// Randomness contract: shot s of a sampling call with `seed` draws its noise
// from the stream block_seed(0x7eed0000, seed, s) and its measurement
// branches and dormant coins from block_seed(0x5eed1234, seed, s). That is
// exactly what ticit's prepared sampler does when run with one-shot chunks
// and one-shot blocks (src/sampler/prepared.rs:753-778), so every shot is an
// independent stream and results do not depend on the launch shape.
#pragma once

#include <hip/hip_runtime.h>

#include "device_abi.hpp"

namespace ticit {

// Program tables never change while a kernel runs. Read through the
// constant address space, a load at a wave-uniform index becomes a scalar
// load (SGPRs, the scalar cache) instead of one vector load per lane.
template <class T>
using ConstPtr = const __attribute__((address_space(4))) T*;
template <class T>
__device__ __forceinline__ ConstPtr<T> konst(const T* p) {
  return (ConstPtr<T>)p;
}

// A whole table entry from constant memory, word by word (scalar loads at a
// uniform address), into registers.
template <class T>
__device__ __forceinline__ T kload(ConstPtr<T> p) {
  static_assert(sizeof(T) % 4 == 0, "table entries are whole words");
  constexpr int N = int(sizeof(T) / 4);
  const auto w = (const __attribute__((address_space(4))) uint32_t*)p;
  uint32_t tmp[N];
#pragma unroll
  for (int k = 0; k < N; ++k) tmp[k] = w[k];
  T out;
  __builtin_memcpy(&out, tmp, sizeof(T));
  return out;
}

constexpr uint64_t kExoSeedBase = 0x7eed0000ull;
constexpr uint64_t kBranchSeedBase = 0x5eed1234ull;
constexpr double kInvSqrt2 = 0.70710678118654752440;

// src/random.rs:13-19
__device__ __forceinline__ uint64_t sm_next(uint64_t& s) {
  s += 0x9e3779b97f4a7c15ull;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// src/random.rs:22-24: 53 random bits times 2^-53
__device__ __forceinline__ double sm_uniform(uint64_t& s) {
  return double(sm_next(s) >> 11) * 0x1.0p-53;
}

// src/random.rs:77-80
// The sampling seed of a launch (plus its run-time offset, if any).
__device__ __forceinline__ uint64_t launch_seed(const DevLaunch& L) {
  return L.seed_offset ? L.seed + *L.seed_offset : L.seed;
}

__device__ __forceinline__ uint64_t seed_for(uint64_t base, uint64_t seed, uint64_t block) {
  return base ^ (0x9e3779b97f4a7c15ull * (block + 1)) ^ (0xbf58476d1ce4e5b9ull * (seed + 1));
}

// src/random.rs:64-73: one draw per gap, clamped to i32::MAX
__device__ __forceinline__ double gap_of(double u, double denom) {
  u = u < 2.2250738585072014e-308 ? 2.2250738585072014e-308 : u;
  double gap = floor(log(u) / denom);
  if (!isfinite(gap) || gap >= 2147483647.0) return 2147483647.0;
  return gap;
}

__device__ __forceinline__ double sample_gap(uint64_t& s, double denom) { return gap_of(sm_uniform(s), denom); }

// src/random.rs:40-50: the first row whose running sum reaches the uniform,
// else the last row. `cum` holds the running sums as the sequential walk
// forms them (host-summed in row order); they never decrease, so a binary
// search finds the same row.
__device__ __forceinline__ uint32_t categorical_row(uint64_t& s, const double* cum, uint32_t n) {
  const double r = sm_uniform(s);
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (r <= cum[mid]) hi = mid;
    else lo = mid + 1;
  }
  return lo < n ? lo : n - 1;
}

// or_low_probability_bits over a single shot (src/sampler/exogenous.rs:150-175):
// a miss costs one draw, a hit two (the second one ends the walk). A first
// uniform below `miss_below` is a miss without the logarithm.
__device__ __forceinline__ bool low_probability_bit(uint64_t& s, double denom, double miss_below) {
  const double u = sm_uniform(s);
  if (u < miss_below || gap_of(u, denom) >= 1.0) return false;
  (void)sample_gap(s, denom);
  return true;
}

// Walks one shot's exogenous noise in the reference family order
// (src/sampler/exogenous.rs:536-625) and calls flip(p) for every exogenous
// partial p an event changes (XOR semantics: a partial flipped twice is
// unchanged). Events reach their partials through the toggle lists.
template <class Flip>
__device__ void sample_exogenous(const DevProgram& P, uint64_t rng, Flip&& flip) {
  auto flip_list = [&](uint32_t k) {
    const uint32_t o1 = P.tl_off[k + 1];
    for (uint32_t j = P.tl_off[k]; j < o1; ++j) flip(P.tl_part[j]);
  };
  for (int i = 0; i < P.ncat; ++i) {
    const DevCat d = P.cats[i];
    uint32_t row = categorical_row(rng, P.f64pool + d.prob_off, d.nrows);
    flip_list(d.tl_base + row);
  }
  for (int i = 0; i < P.nrare; ++i) {
    const DevRare g = P.rares[i];
    if (!g.enabled) continue;
    const int64_t total = int64_t(g.nsets);
    int64_t draw = 0;
    const double u0 = sm_uniform(rng);
    if (u0 < g.none_below) continue;  // the first gap spans every set
    for (bool first = true;; first = false) {
      int64_t gap = int64_t(first ? gap_of(u0, g.denom) : sample_gap(rng, g.denom));
      if (gap >= total - draw) break;
      draw += gap;
      uint32_t set = uint32_t(draw);  // one shot per stream: draw / nsets == 0
      uint32_t er = categorical_row(rng, P.f64pool + g.ev_prob_off, g.nev);
      flip_list(g.tl_base + set * g.nev + er);
      draw += 1;
    }
  }
  for (int i = 0; i < P.nbern; ++i) {
    const DevBern b = P.berns[i];
    bool bit = false;
    switch (b.mode) {
      case kBernNone:
        continue;
      case kBernAll:
        bit = true;
        break;
      case kBernHalf:
        bit = sm_next(rng) & 1;
        break;
      case kBernLow:
        bit = low_probability_bit(rng, b.denom, b.miss_below);
        break;
      default: {  // kBernCoins: src/sampler/exogenous.rs:225-257
        uint64_t alive = sm_next(rng);
        uint64_t result = 0;
        for (int k = 6; k >= 0; --k) {
          uint64_t shoot = sm_next(rng);
          if ((b.top_bits >> k) & 1) result |= shoot & alive;
          alive &= ~shoot;
        }
        bit = result & 1;
        if (b.left_denom != 0.0) bit = low_probability_bit(rng, b.left_denom, b.left_miss_below) || bit;
        break;
      }
    }
    if (b.invert) bit = !bit;
    if (bit)
      for (uint32_t j = P.sym_off[b.cond - 1]; j < P.sym_off[b.cond]; ++j) flip(P.sym_part[j]);
  }
  for (int i = 0; i < P.nlowp; ++i) {
    const DevLowp g = P.lowps[i];
    if (!g.enabled) continue;
    const int64_t total = int64_t(g.nconds);
    int64_t draw = 0;
    const double u0 = sm_uniform(rng);
    if (u0 < g.none_below) continue;  // the first gap spans every condition
    for (bool first = true;; first = false) {
      int64_t gap = int64_t(first ? gap_of(u0, g.denom) : sample_gap(rng, g.denom));
      if (gap >= total - draw) break;
      draw += gap;
      flip_list(g.tl_base + uint32_t(draw));
      draw += 1;
    }
  }
}

__device__ __forceinline__ size_t insert_zero_bit(size_t packed, int bit) {
  size_t low = (size_t(1) << bit) - 1;
  return (packed & low) | ((packed & ~low) << 1);
}

__device__ __forceinline__ bool parity64(uint64_t v) { return __popcll(v) & 1; }

// PrecomputedActivePauliMeasurementKernel::diagonal_source (src/active.rs:347-356)
__device__ __forceinline__ size_t diagonal_source(const DevInstr& in, size_t packed, bool branch) {
  size_t wo = insert_zero_bit(packed, in.pivot);
  int par = in.zwop == 0 ? 0 : int(parity64(uint64_t(wo) & in.zwop));
  int pv = ((in.op & kFlagDiagPhase) ? 1 : 0) ^ par ^ int(branch);
  return wo | (size_t(pv) << in.pivot);
}

__device__ __forceinline__ double clamp01(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

}  // namespace ticit
