// C ABI of the hip sampling backend (include/ticit_hip.h): program upload,
// tier selection and the sampling entry points.
#include "../../../include/ticit_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <chrono>
#include <map>
#include <mutex>
#include <cstring>
#include <string>
#include <sys/mman.h>
#include <vector>

#include "../kernels/compact.hpp"
#include "../kernels/static_scratch.hpp"
#include "component_plan.hpp"
#include "handle.hpp"
#include "host_copy.hpp"
#include "pinning.hpp"

using namespace ticit;

namespace ticit {

thread_local std::string g_error;

ticit_hip_status fail(ticit_hip_status s, const std::string& msg) {
  g_error = msg;
  return s;
}

ticit_hip_status from_hip(hipError_t e, const char* what) {
  if (e == hipErrorOutOfMemory) return fail(TICIT_HIP_OUT_OF_DEVICE_MEMORY, std::string(what) + ": out of device memory");
  return fail(TICIT_HIP_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

// The lane tier's device memory regions for launch L: their count (one per
// resident wave, at most the launch's waves and 4 GiB in all; the kernel
// strides its wave blocks over them) and bytes each. Zero when the tables
// live in LDS.
hipError_t lane_region_plan(ticit_hip_program* h, const DevLaunch& L, uint64_t* slots, size_t* slot_bytes) {
  *slots = 0;
  *slot_bytes = 0;
  if (h->tier != kTierLane || h->lane.store == kLaneLds) return hipSuccess;
  if (h->lane_ws.max_slots == 0) {
    uint32_t lds = 0, blocks = 0;
    lane_tier_occupancy(h->tab.max_k, h->lane, 0, &lds, &blocks);
    int cus = 0;
    hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device);
    if (e != hipSuccess) return e;
    h->lane_ws.max_slots = std::max<uint64_t>(1, uint64_t(std::max<uint32_t>(blocks, 1)) * uint64_t(cus));
  }
  const size_t slot = lane_region_bytes(h->tab.max_k, lane_launch_layout(h->tab.max_k, h->lane, L));
  const uint64_t waves = std::max<uint64_t>(1, (L.shots + 63) / 64);
  uint64_t n = std::min<uint64_t>(waves, h->lane_ws.max_slots);
  n = std::max<uint64_t>(1, std::min<uint64_t>(n, (size_t(4) << 30) / slot));
  *slots = n;
  *slot_bytes = slot;
  return hipSuccess;
}

}  // namespace ticit

namespace {

// Widest active state the automatic tier choice gives the group tier.
constexpr int kGroupAutoMaxK = 10;
// Widest dense state the lane tier runs with its tables in device memory.
constexpr int kLaneGlobalMaxK = 8;

// Bit steps (records, detectors, coins) at least 256 and 16 times the
// amplitude steps.
bool bit_heavy_program(const HostTables& t) {
  size_t bits = 0, amps = 0;
  for (const DevInstr& in : t.instrs) {
    const uint32_t op = in.op & 0xff;
    if (op == kOpRecord || op == kOpDetector || op == kOpCoin) ++bits;
    else ++amps;
  }
  return bits >= 256 && bits >= 16 * amps;
}

// SamplerOptions::normalize_syndromes and expected_* (src/sampler/prepared.rs:111-116).
struct ReferenceRequest {
  ExpectedBits bits;
  bool normalize = false;
};

ticit_hip_status reference_rows(ticit_hip_program* h, std::vector<uint8_t>& det, std::vector<uint8_t>& obs);

ticit_hip_status create_from_program(Program prog, const ticit_hip_options* options, ticit_hip_program** out,
                                     const std::vector<MeasurementParity>* pins = nullptr,
                                     const ReferenceRequest* ref = nullptr) {
  // src/sampler/prepared.rs:515-550: the explicit bits and the noiseless
  // reference exclude each other; the reference obeys the same pins
  ExpectedBits expected;
  if (ref) {
    if (ref->normalize && (!ref->bits.detectors.empty() || !ref->bits.observables.empty()))
      return fail(TICIT_HIP_INVALID_INPUT,
                  "normalize_syndromes cannot be combined with expected_detectors or expected_observables");
    expected = ref->bits;
    if (ref->normalize && (prog.ndetectors > 0 || !prog.observables.empty())) {
      // the reference shot runs in FP64 on the planned layout, whatever
      // this sampler's precision and tier
      ticit_hip_options ro = options ? *options : ticit_hip_options{};
      ro.precision = 0;
      ro.tier = 0;
      ticit_hip_program* rh = nullptr;
      ticit_hip_status st = create_from_program(prog, &ro, &rh, pins, nullptr);
      if (st != TICIT_HIP_OK) return st;
      st = reference_rows(rh, expected.detectors, expected.observables);
      delete rh;
      if (st != TICIT_HIP_OK) return st;
    }
  }
  auto* h = new ticit_hip_program();
  h->prog = std::move(prog);
  if (options) h->opt = *options;
  // Execution layout: the program's own choice (use_active_components,
  // src/planner.rs:1482-1484) unless the caller forces one.
  if (h->opt.components > 2) {
    delete h;
    return fail(TICIT_HIP_INVALID_INPUT, "components must be 0 (planned), 1 (dense) or 2 (tensor components)");
  }
  bool use_components = h->opt.components == 2 || (h->opt.components == 0 && h->prog.use_components);
  if (h->opt.components != 0) {
    h->planner_source = 2;
  } else if (h->opt.planner_mode == 1) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    bool learned = false;
    if (learned_layout(h->prog, dev, &learned, &h->planner_p)) {
      use_components = learned;
      h->planner_source = 1;
    }
  }
  try {
    std::vector<ForcedBranch> forced;
    if (pins) forced = plan_pinned_measurements(h->prog, *pins);
    h->tab = lower_to_device(h->prog, use_components, &forced, &expected);
    h->expected = std::move(expected);
    if (pins) h->pins = *pins;
  } catch (const TicitError& e) {
    delete h;
    return fail(ticit_hip_status(e.kind), e.what());
  }
  const HostTables& t = h->tab;
  int lds_limit = 0;
  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device);
  if (e != hipSuccess) {
    delete h;
    return from_hip(e, "no HIP device");
  }
  DevProgram& d = h->dev;
#define UP(field, vec)                                     \
  do {                                                     \
    hipError_t ue = h->upload(vec, d.field);               \
    if (ue != hipSuccess) {                                \
      delete h;                                            \
      return from_hip(ue, "uploading program tables");     \
    }                                                      \
  } while (0)
  UP(instrs, t.instrs);
  UP(steps, t.steps);
  UP(lsteps, t.lsteps);
  UP(bsteps, t.bsteps);
  UP(resid, t.resid);
  UP(det_recs, t.det_recs);
  UP(post_dets, t.post_dets);
  UP(obs_off, t.obs_off);
  UP(obs_recs, t.obs_recs);
  UP(obs_flip, t.obs_flip);
  UP(sym_off, t.sym_off);
  UP(sym_part, t.sym_part);
  UP(tl_off, t.tl_off);
  UP(tl_part, t.tl_part);
  UP(exo_syms, t.exo_syms);
  UP(cats, t.cats);
  UP(rares, t.rares);
  UP(berns, t.berns);
  UP(lowps, t.lowps);
  UP(f64pool, t.f64pool);
  UP(u64pool, t.u64pool);
  UP(i32pool, t.i32pool);
#undef UP
  d.ninstr = t.ninstr;
  d.n = t.n;
  // component runs start from the explicit |0> components of the stream
  d.initial_k = t.components ? 0 : t.initial_k;
  d.max_k = t.max_k;
  d.nrec = t.nrec;
  d.ndet = t.ndet;
  d.nobs = t.nobs;
  d.nslots = t.nslots;
  d.npartials = t.npartials;
  d.nsym = t.nsym;
  d.nexo = int32_t(h->prog.nsymbols ? t.exo_syms.size() : 0);
  if (t.exo_syms.size() == 1 && t.exo_syms[0] == 0) d.nexo = 0;
  d.nmeasure = t.nmeasure;
  d.ncoin = t.ncoin;
  d.ncat = int32_t(t.cats.size());
  d.nrare = int32_t(t.rares.size());
  d.nbern = int32_t(t.berns.size());
  d.nlowp = int32_t(t.lowps.size());
  d.state_dim = t.state_dim;
  d.ncomp = int32_t(t.comp_off.size());
  d.max_local_k = t.max_local_k;
  d.nexp = t.nexp;
  d.nlsteps = t.nlsteps;
  d.npost = t.npost;
  e = hipMalloc(reinterpret_cast<void**>(&h->d_error), sizeof(int));
  if (e != hipSuccess) {
    delete h;
    return from_hip(e, "allocating the error flag");
  }
  h->allocations.push_back(h->d_error);

  // Tier choice: per-lane resident state up to k = 5, one shot per
  // workgroup up to width 10, the HBM-resident wide tier above.
  auto words = [](int bits) { return uint32_t((bits + 63) / 64); };
  h->lane = LaneLayout{words(t.npartials), words(t.nslots), words(t.nrec), words(t.ndet), 0,
                       uint32_t(t.npartials), uint32_t(t.nslots), uint32_t(t.nrec), uint32_t(t.ndet),
                       uint32_t(t.dets_once), 0};
  h->group = GroupLayout{h->lane.pw, h->lane.sw, h->lane.rw, h->lane.dw};
  uint32_t max_row = std::max<uint32_t>({uint32_t(t.nrec), uint32_t(t.ndet), uint32_t(t.nobs), 1});
  h->lane.stage_bytes = 64 * ((max_row + 15) / 16 * 16);
  h->wide = WideLayout{h->lane.pw, h->lane.sw, h->lane.rw, h->lane.dw,
                       h->lane.pw + h->lane.sw + h->lane.rw + h->lane.dw,
                       wide_partials_per_shot(t.max_local_k)};
  const bool lane_ok = !t.components && t.max_k <= 5 && lane_tier_lds_bytes(t.max_k, h->lane) <= size_t(lds_limit);
  // Tables too large for LDS: the lane tier keeps its bit tables (and, above
  // width 5, the amplitudes) in per-wave device memory regions, with the
  // same [item][lane] layout, so record and detector steps stay one word per
  // 64 shots. Dense states up to width 8 only: a lane walks its 2^k
  // amplitudes one by one.
  uint32_t lane_store = kLaneLds;
  if (!lane_ok && !t.components && t.max_k <= kLaneGlobalMaxK)
    lane_store = t.max_k <= 5 ? kLaneBitsGlobal : kLaneAllGlobal;
  const bool lane_global_ok = !lane_ok && !t.components && t.max_k <= kLaneGlobalMaxK;
  auto group_fits = [&](int tps) {
    return t.max_local_k <= group_max_local_k(tps) &&
           group_tier_lds_bytes(t.state_dim, h->group, tps) <= size_t(lds_limit);
  };
  // Default launch shape: the fewest threads per shot that hold the
  // program's widest component and fit the LDS (8, 16 or 32 threads pack
  // 8, 4 or 2 shots into a wave). The per-instruction interpreter work is
  // then spread over the most shots; measured per 1e5 shots on random
  // Clifford+T programs, the smallest fitting shape was the fastest or
  // within 12 % of it, where the former width rule (64 threads up to width
  // 7, 128 up to 10, 256 above) took 1.1-4.5x as long (width 6: 0.13 ms at
  // 8 threads against 0.58 ms at 64; 16 qubits, 12 T, dense width 10: 2.0 ms
  // at 64 against 2.24 ms at 128).
  //
  // The automatic choice stops the group tier at width 10. Widths 11 to 13
  // fit the LDS of a 256- or 512-thread shot, but a 32-128 KiB region
  // leaves one or two workgroups per CU, and the amplitude passes then wait
  // on LDS latency: on 30 random Clifford+T programs of widths 11 and 12,
  // 1e4-1e6 shots each, the wide tier took 0.39-1.14x the time of the best
  // group shape (median 0.67x; faster on 28). At width 10 and below the two
  // are within 2x either way and the group tier keeps the rows in LDS.
  int plain_tps = 256;
  for (int c : {8, 16, 32, 64, 128, 256, 512})
    if (group_fits(c)) {
      plain_tps = c;
      break;
    }
  int tps = h->opt.group_threads ? int(h->opt.group_threads) : plain_tps;
  if (tps != 8 && tps != 16 && tps != 32 && tps != 64 && tps != 128 && tps != 256 && tps != 512) {
    delete h;
    return fail(TICIT_HIP_INVALID_INPUT, "group_threads must be 8, 16, 32, 64, 128, 256 or 512");
  }
  h->group_tps = tps;
  const bool group_ok = group_fits(tps);
  uint32_t want = h->opt.tier;
  if (want == 0 && h->opt.precision == 1) want = kTierWide;  // the FP32 path lives in the wide tier
  // A program dominated by record, detector and coin steps (the CCZ
  // fixtures: k = 8, 22.9k and 61k steps, almost all of them bit steps)
  // runs in the lane tier with its tables in device memory, where a step is
  // one word operation for 64 shots; the group tier would spend a whole
  // workgroup per shot on it, the wide tier a thread per shot.
  h->bit_heavy = bit_heavy_program(t);
  if (want == 0) {
    if (lane_ok)
      want = kTierLane;
    else if (h->bit_heavy && lane_global_ok)
      want = kTierLane;
    else
      want = group_ok && t.max_local_k <= kGroupAutoMaxK && !h->bit_heavy ? kTierGroup : kTierWide;
  }
  if (want == kTierLane && !lane_ok && !lane_global_ok) {
    delete h;
    return fail(TICIT_HIP_UNSUPPORTED, "lane tier needs a dense max_k <= " + std::to_string(kLaneGlobalMaxK));
  }
  if (want == kTierLane) {
    h->lane.store = lane_ok ? kLaneLds : lane_store;
    // a caller's placement: 1 LDS, 2 bit tables in device memory, 3 all
    const uint32_t asked = h->opt.lane_store;
    if (asked > 3) {
      delete h;
      return fail(TICIT_HIP_INVALID_INPUT, "lane_store must be 0, 1, 2 or 3");
    }
    if (asked == 1 && !lane_ok) {
      delete h;
      return fail(TICIT_HIP_UNSUPPORTED, "the lane tier's tables do not fit LDS");
    }
    LaneLayout amps_only{};
    amps_only.store = kLaneBitsGlobal;
    if (asked == 2 && !(t.max_k <= 5 && !t.components && lane_tier_lds_bytes(t.max_k, amps_only) <= size_t(lds_limit))) {
      delete h;
      return fail(TICIT_HIP_UNSUPPORTED, "bit tables in device memory keep the amplitudes in LDS: width 5 at most");
    }
    if (asked == 3 && !(t.max_k > 5 && t.max_k <= kLaneGlobalMaxK && !t.components)) {
      delete h;
      return fail(TICIT_HIP_UNSUPPORTED, "all lane tables in device memory: dense widths 6 to 8");
    }
    if (asked) h->lane.store = asked - 1;
  }
  if (want == kTierGroup && !group_ok) {
    delete h;
    return fail(TICIT_HIP_UNSUPPORTED, "group tier state exceeds the LDS of one workgroup");
  }
  if (want == kTierWide) {
    if (t.max_local_k > 30 || t.state_dim > (uint32_t(1) << 31)) {
      delete h;
      return fail(TICIT_HIP_UNSUPPORTED, "active width " + std::to_string(t.max_local_k) + " exceeds the wide tier");
    }
    // Shots in flight: the reference's workspace policy, at most 16 GiB or
    // two thirds of free memory (src/gpu/mod.rs:680-711), capped by grid y.
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    size_t budget = std::min<size_t>(size_t(16) << 30, free_b / 3 * 2);
    size_t per = wide_bytes_per_shot(t.state_dim, h->opt.precision == 1, h->wide);
    size_t batch = per ? budget / per : 1;
    if (h->opt.wide_batch) batch = std::min<size_t>(batch, h->opt.wide_batch);
    batch = std::max<size_t>(1, std::min<size_t>(batch, 65535));
    h->wide_batch = uint32_t(batch);
  }
  h->tier = want;
  *out = h;
  return TICIT_HIP_OK;
}

// Grows the program's lane regions for launch L, after their previous user
// has finished. Nothing is allocated when they are large enough already.
hipError_t lane_workspace_prepare(ticit_hip_program* h, const DevLaunch& L) {
  uint64_t slots = 0;
  size_t slot = 0;
  hipError_t e = lane_region_plan(h, L, &slots, &slot);
  if (e != hipSuccess || slots == 0) return e;
  if (!h->lane_ws_done && (e = hipEventCreateWithFlags(&h->lane_ws_done, hipEventDisableTiming)) != hipSuccess)
    return e;
  const size_t need = size_t(slots) * slot;
  if (need <= h->lane_ws.bytes) return hipSuccess;
  if ((e = hipEventSynchronize(h->lane_ws_done)) != hipSuccess) return e;
  if (h->lane_ws.base) (void)hipFree(h->lane_ws.base);
  h->lane_ws.base = nullptr;
  h->lane_ws.bytes = 0;
  void* p = nullptr;
  if ((e = hipMalloc(&p, need)) != hipSuccess) return e;
  h->lane_ws.base = static_cast<unsigned char*>(p);
  h->lane_ws.bytes = need;
  return hipSuccess;
}

hipError_t launch(ticit_hip_program* h, const DevLaunch& L, hipStream_t stream, double* dump_re = nullptr,
                  double* dump_im = nullptr, int* dump_k = nullptr) {
  if (h->tier == kTierLane) {
    if (h->lane.store == kLaneLds) return launch_lane_tier(h->tab.max_k, h->dev, L, h->lane, stream);
    hipError_t e = lane_workspace_prepare(h, L);
    if (e != hipSuccess) return e;
    // calls sharing the regions may come on different streams: order them
    if ((e = hipStreamWaitEvent(stream, h->lane_ws_done, 0)) != hipSuccess) return e;
    if ((e = launch_lane_tier(h->tab.max_k, h->dev, L, h->lane, stream, &h->lane_ws)) != hipSuccess) return e;
    return hipEventRecord(h->lane_ws_done, stream);
  }
  if (h->tier == kTierGroup)
    return launch_group_tier(h->dev, L, h->group, stream, h->group_tps);
  uint32_t batch = uint32_t(std::min<uint64_t>(h->wide_batch, L.shots));
  // Two wide-tier workspaces per device, shared by every handle on it and
  // grown to the largest batch asked for; a launch of several shot batches
  // alternates them, the second on a library stream forked from and joined
  // to the caller's (the batch of 256 circuits: 2.43 -> 2.26 s with its
  // wide items in two sequences). Calls are serialized on them (host side
  // by the mutex, device side by their completion events).
  struct DeviceWide {
    WideWorkspace* ws = nullptr;
    WideSecond second{};
  };
  static std::mutex mu;
  static std::map<int, DeviceWide> pool;
  std::lock_guard<std::mutex> lock(mu);
  DeviceWide& d = pool[h->device];
  if (!d.ws) d.ws = wide_workspace_new();
  const WideSecond* second = nullptr;
  if (L.shots > batch && !dump_re) {
    if (!d.second.stream) {
      WideSecond q{wide_workspace_new(), nullptr, nullptr};
      if (hipStreamCreateWithFlags(&q.stream, hipStreamNonBlocking) == hipSuccess &&
          hipEventCreateWithFlags(&q.fork, hipEventDisableTiming) == hipSuccess) {
        d.second = q;
      } else {
        (void)hipGetLastError();
        if (q.stream) (void)hipStreamDestroy(q.stream);
        wide_workspace_free(q.ws);
      }
    }
    if (d.second.stream) second = &d.second;
  }
  return launch_wide_tier(h->dev, h->tab.instrs.data(), L, h->wide, batch, h->opt.precision == 1, d.ws, stream,
                          dump_re, dump_im, dump_k, true, second);
}

// One call's device buffers, carved from the program's arena: sizes are
// planned first, then one reservation (the arena grows only when a call
// needs more than any before it), then the pointers are assigned.
struct ArenaPlan {
  struct Slot {
    void** out;
    size_t bytes;
  };
  std::vector<Slot> slots;
  template <class T>
  void add(T** out, size_t count) {
    slots.push_back(Slot{reinterpret_cast<void**>(out), std::max<size_t>(count, 1) * sizeof(T)});
  }
  static size_t up(size_t b) { return (b + 255) / 256 * 256; }
  size_t total() const {
    size_t t = 0;
    for (const Slot& s : slots) t += up(s.bytes);
    return t;
  }
};

hipError_t arena_assign(ticit_hip_program* h, ArenaPlan& plan) {
  const size_t need = plan.total();
  if (need > h->arena_bytes) {
    // every earlier host-result call has finished with the arena
    if (h->arena) (void)hipFree(h->arena);
    h->arena = nullptr;
    h->arena_bytes = 0;
    hipError_t e = hipMalloc(&h->arena, need);
    if (e != hipSuccess) return e;
    h->arena_bytes = need;
  }
  uint8_t* p = static_cast<uint8_t*>(h->arena);
  for (const ArenaPlan::Slot& s : plan.slots) {
    *s.out = p;
    p += ArenaPlan::up(s.bytes);
  }
  return hipSuccess;
}

hipError_t pinned_reserve(ticit_hip_program* h, size_t bytes) {
  if (bytes <= h->pinned_bytes) return hipSuccess;
  if (h->pinned) (void)hipHostFree(h->pinned);
  h->pinned = nullptr;
  h->pinned_bytes = 0;
  hipError_t e = hipHostMalloc(&h->pinned, bytes, hipHostMallocDefault);
  if (e != hipSuccess) return e;
  h->pinned_bytes = bytes;
  return hipSuccess;
}

// Copies of device results into host buffers through the pinned staging:
// every piece is queued device-to-pinned, one wait, then each is copied to
// its destination.
struct CopyBack {
  struct Piece {
    void* dst;
    const void* src;
    size_t bytes;
  };
  std::vector<Piece> pieces;
  void add(void* dst, const void* src, size_t bytes) {
    if (bytes) pieces.push_back(Piece{dst, src, bytes});
  }
  hipError_t run(ticit_hip_program* h) {
    size_t total = 0;
    for (const Piece& p : pieces) total += ArenaPlan::up(p.bytes);
    if (!total) return hipSuccess;
    hipError_t e = pinned_reserve(h, total);
    if (e != hipSuccess) return e;
    uint8_t* stage = static_cast<uint8_t*>(h->pinned);
    size_t off = 0;
    for (const Piece& p : pieces) {
      if ((e = hipMemcpyAsync(stage + off, p.src, p.bytes, hipMemcpyDeviceToHost, nullptr)) != hipSuccess) return e;
      off += ArenaPlan::up(p.bytes);
    }
    if ((e = hipStreamSynchronize(nullptr)) != hipSuccess) return e;
    std::vector<HostCopy> out;
    off = 0;
    for (const Piece& p : pieces) {
      out.push_back(HostCopy{p.dst, stage + off, p.bytes});
      off += ArenaPlan::up(p.bytes);
    }
    host_copy(out.data(), out.size());
    return hipSuccess;
  }
};

constexpr int kDevErrAny = kDevErrImpossibleBranch | kDevErrBadProbability | kDevErrUnfairPin;

// Host result rows (freed with free() by ticit_hip_result_free): large ones
// start on a 2 MiB boundary with transparent huge pages requested, so the
// first touch of a row buffer faults once per 2 MiB rather than per 4 KiB.
void* host_rows(size_t bytes) {
  constexpr size_t kHuge = size_t(2) << 20;
  bytes = std::max<size_t>(bytes, 1);
  if (bytes >= 2 * kHuge) {
    void* p = nullptr;
    if (posix_memalign(&p, kHuge, bytes) == 0) {
      (void)madvise(p, bytes & ~size_t(4095), MADV_HUGEPAGE);
      return p;
    }
  }
  return malloc(bytes);
}

// Row bytes per chunk of a chunked host-result call (all row kinds
// together): small enough that the copies overlap the next chunks' kernels
// and the host copy of the previous chunk, large enough that a chunk fills
// the device.
constexpr size_t kChunkBytes = size_t(8) << 20;
constexpr uint64_t kMinChunkShots = 65536;

// Samples L in chunks and returns every chunk's rows into dst (full-size
// host buffers; dst[i] null = no such rows): kernels on the null stream,
// one event per chunk, device-to-pinned copies on the program's copy
// stream into two slots used in turn, host copies out of a slot once its
// copy event has fired. Ends with every queued operation finished, on
// error too.
hipError_t chunked_rows(ticit_hip_program* h, const DevLaunch& L, const uint8_t* const src[4],
                        const uint32_t rbs[4], uint8_t* const dst[4], uint64_t* observable_ones, int* derr,
                        float* ms) {
  const uint64_t shots = L.shots;
  size_t row_all = 0;
  for (int i = 0; i < 4; ++i)
    if (dst[i]) row_all += rbs[i];
  uint64_t chunk = std::max<uint64_t>(kChunkBytes / std::max<size_t>(row_all, 1), kMinChunkShots);
  chunk = std::min<uint64_t>((chunk + 63) / 64 * 64, shots);
  const uint64_t nchunks = (shots + chunk - 1) / chunk;
  size_t slot = 0;
  for (int i = 0; i < 4; ++i)
    if (dst[i]) slot += ArenaPlan::up(size_t(chunk) * rbs[i]);
  hipError_t e = pinned_reserve(h, 2 * slot);
  if (e != hipSuccess) return e;
  if (!h->copy_stream && (e = hipStreamCreateWithFlags(&h->copy_stream, hipStreamNonBlocking)) != hipSuccess)
    return e;
  for (hipEvent_t& ev : h->slot_done)
    if (!ev && (e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return e;
  while (h->chunk_done.size() < nchunks) {
    hipEvent_t ev = nullptr;
    if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return e;
    h->chunk_done.push_back(ev);
  }
  uint8_t* const stage = static_cast<uint8_t*>(h->pinned);
  auto queue_copy = [&](uint64_t c) -> hipError_t {
    const uint64_t first = c * chunk, n = std::min<uint64_t>(chunk, shots - first);
    hipError_t r = hipStreamWaitEvent(h->copy_stream, h->chunk_done[c], 0);
    size_t off = (c & 1) * slot;
    for (int i = 0; i < 4 && r == hipSuccess; ++i) {
      if (!dst[i]) continue;
      if (n * rbs[i])
        r = hipMemcpyAsync(stage + off, src[i] + first * rbs[i], n * rbs[i], hipMemcpyDeviceToHost, h->copy_stream);
      off += ArenaPlan::up(size_t(chunk) * rbs[i]);
    }
    return r == hipSuccess ? hipEventRecord(h->slot_done[c & 1], h->copy_stream) : r;
  };
  auto body = [&]() -> hipError_t {
    hipError_t r = hipEventRecord(h->ev_start, nullptr);
    for (uint64_t c = 0; c < nchunks && r == hipSuccess; ++c) {
      DevLaunch Lc = L;
      Lc.first_shot = L.first_shot + c * chunk;
      Lc.shots = std::min<uint64_t>(chunk, shots - c * chunk);
      Lc.meas = L.meas + c * chunk * L.meas_row;
      Lc.det = L.det + c * chunk * L.det_row;
      Lc.obs = L.obs + c * chunk * L.obs_row;
      if (L.expv) Lc.expv = L.expv + c * chunk * (rbs[3] / 8);
      r = launch(h, Lc, nullptr);
      if (r == hipSuccess) r = hipEventRecord(h->chunk_done[c], nullptr);
    }
    if (r == hipSuccess) r = hipEventRecord(h->ev_stop, nullptr);
    for (uint64_t c = 0; c < std::min<uint64_t>(2, nchunks) && r == hipSuccess; ++c) r = queue_copy(c);
    for (uint64_t c = 0; c < nchunks && r == hipSuccess; ++c) {
      if ((r = hipEventSynchronize(h->slot_done[c & 1])) != hipSuccess) break;
      const uint64_t first = c * chunk, n = std::min<uint64_t>(chunk, shots - first);
      size_t off = (c & 1) * slot;
      HostCopy pieces[4];
      size_t np = 0;
      for (int i = 0; i < 4; ++i) {
        if (!dst[i]) continue;
        pieces[np++] = HostCopy{dst[i] + first * rbs[i], stage + off, n * rbs[i]};
        off += ArenaPlan::up(size_t(chunk) * rbs[i]);
      }
      host_copy(pieces, np);
      if (c + 2 < nchunks) r = queue_copy(c + 2);
    }
    if (r != hipSuccess) return r;
    if ((r = hipEventSynchronize(h->ev_stop)) != hipSuccess) return r;
    (void)hipEventElapsedTime(ms, h->ev_start, h->ev_stop);
    if ((r = hipMemcpy(derr, h->d_error, sizeof(int), hipMemcpyDeviceToHost)) != hipSuccess) return r;
    return hipMemcpy(observable_ones, L.obs_ones, size_t(h->tab.nobs) * 8, hipMemcpyDeviceToHost);
  };
  e = body();
  if (e != hipSuccess) {
    // nothing queued may outlive the call: the next one reuses the buffers
    (void)hipStreamSynchronize(h->copy_stream);
    (void)hipStreamSynchronize(nullptr);
  }
  return e;
}

ticit_hip_status run_host(ticit_hip_program* h, uint64_t shots, uint64_t seed, int bit_packed, int keep_records,
                          bool injected, const uint64_t* exo, const double* uniforms, const uint64_t* coins,
                          double* probs, ticit_hip_result* out, double* state_re = nullptr,
                          double* state_im = nullptr, int* state_k = nullptr, bool reference = false) {
  auto t0 = std::chrono::steady_clock::now();
  std::memset(out, 0, sizeof(*out));
  const HostTables& t = h->tab;
  DevLaunch L{};
  L.shots = shots;
  L.seed = seed;
  L.first_shot = 0;
  L.bit_packed = bit_packed != 0;
  L.keep_records = keep_records != 0;
  L.injected = injected;
  L.want_probs = probs != nullptr;
  L.reference = reference;
  L.meas_row = h->row_bytes(uint64_t(t.nrec), bit_packed);
  L.det_row = h->row_bytes(uint64_t(t.ndet), bit_packed);
  L.obs_row = h->row_bytes(uint64_t(t.nobs), bit_packed);
  if (shots == 0) {
    out->bit_packed = bit_packed != 0;
    out->nmeasurements = uint64_t(t.nrec);
    out->ndetectors = uint64_t(t.ndet);
    out->nobservables = uint64_t(t.nobs);
    out->nexpvals = uint64_t(t.nexp);
    out->observable_ones = static_cast<uint64_t*>(calloc(std::max(1, t.nobs), 8));
    return TICIT_HIP_OK;
  }
  if (h->tier == kTierGroup && shots > 0x7fffffffull) return fail(TICIT_HIP_INVALID_INPUT, "too many shots for one call");
  std::lock_guard<std::mutex> lock(h->host_mu);
  hipError_t e = hipSuccess;
#define CK(expr, what)                             \
  do {                                             \
    e = (expr);                                    \
    if (e != hipSuccess) return from_hip(e, what); \
  } while (0)
  const size_t words = (shots + 63) / 64;
  const size_t nobs1 = size_t(std::max(1, t.nobs));
  const bool postselect = t.npost > 0 && !reference;  // the reference shot is never discarded
  if (state_re) {
    if (h->tier != kTierWide) return fail(TICIT_HIP_UNSUPPORTED, "state dumps are served by the wide tier");
    if (h->tab.components)
      return fail(TICIT_HIP_UNSUPPORTED, "state dumps need the dense layout (components = 1)");
  }
  // every device buffer of the call, carved from the program's arena
  double* d_expv = nullptr;
  uint8_t* d_keep = nullptr;
  uint64_t* d_exo = nullptr;
  double* d_uni = nullptr;
  uint64_t* d_coins = nullptr;
  double* d_probs = nullptr;
  double* d_state = nullptr;
  uint64_t* tile_off = nullptr;
  unsigned long long* d_acc = nullptr;
  uint8_t* survivors[4] = {nullptr, nullptr, nullptr, nullptr};
  const uint32_t rbs[4] = {L.meas_row, L.det_row, L.obs_row, uint32_t(t.nexp) * 8u};
  ArenaPlan plan;
  if (keep_records) {
    plan.add(&L.meas, size_t(shots) * L.meas_row);
    plan.add(&L.det, size_t(shots) * L.det_row);
    plan.add(&L.obs, size_t(shots) * L.obs_row);
    if (t.nexp > 0) plan.add(&d_expv, size_t(shots) * size_t(t.nexp));
  }
  plan.add(&L.obs_ones, nobs1);
  if (postselect) {
    plan.add(&d_keep, size_t(shots));
    plan.add(&tile_off, compaction_scratch_words(shots));
    plan.add(&d_acc, 1);
    if (keep_records)
      for (int i = 0; i < 4; ++i)
        if (rbs[i]) plan.add(&survivors[i], size_t(shots) * rbs[i]);
  }
  if (injected) {
    plan.add(&d_exo, size_t(std::max(1, t.nsym)) * words);
    plan.add(&d_uni, size_t(std::max(1, t.nmeasure)) * shots);
    plan.add(&d_coins, size_t(std::max(1, t.ncoin)) * words);
  }
  if (probs) plan.add(&d_probs, size_t(std::max(1, t.nmeasure)) * shots);
  if (state_re) plan.add(&d_state, size_t(2) << h->tab.max_k);
  CK(arena_assign(h, plan), "allocating device buffers");
  L.expv = d_expv;
  L.keep = d_keep;
  L.error = h->d_error;
  CK(hipMemsetAsync(L.obs_ones, 0, nobs1 * 8, nullptr), "clearing observable counts");
  CK(hipMemsetAsync(h->d_error, 0, sizeof(int), nullptr), "clearing the error flag");
  if (injected) {
    if (exo && t.nsym) CK(hipMemcpy(d_exo, exo, size_t(t.nsym) * words * 8, hipMemcpyHostToDevice), "copying exo");
    else CK(hipMemsetAsync(d_exo, 0, size_t(std::max(1, t.nsym)) * words * 8, nullptr), "clearing exo");
    if (uniforms && t.nmeasure)
      CK(hipMemcpy(d_uni, uniforms, size_t(t.nmeasure) * shots * 8, hipMemcpyHostToDevice), "copying uniforms");
    else CK(hipMemsetAsync(d_uni, 0, size_t(std::max(1, t.nmeasure)) * shots * 8, nullptr), "clearing uniforms");
    if (coins && t.ncoin) CK(hipMemcpy(d_coins, coins, size_t(t.ncoin) * words * 8, hipMemcpyHostToDevice), "copying coins");
    else CK(hipMemsetAsync(d_coins, 0, size_t(std::max(1, t.ncoin)) * words * 8, nullptr), "clearing coins");
    L.exo = d_exo;
    L.uniforms = d_uni;
    L.coins = d_coins;
  }
  L.probs = d_probs;
  if (!h->ev_start) CK(hipEventCreate(&h->ev_start), "creating events");
  if (!h->ev_stop) CK(hipEventCreate(&h->ev_stop), "creating events");

  out->shots = shots;
  out->accepted = shots;
  out->nexpvals = uint64_t(t.nexp);
  out->bit_packed = bit_packed != 0;
  out->nmeasurements = uint64_t(t.nrec);
  out->ndetectors = uint64_t(t.ndet);
  out->nobservables = uint64_t(t.nobs);
  int derr = 0;
  float ms = 0.f;
  // Lane- and group-tier rows come back in chunks: every chunk's kernel is
  // queued at once, and chunk i is copied to pinned staging on the copy
  // stream as soon as its kernel ends, while the host moves chunk i-1 into
  // the result rows.
  const bool chunked = keep_records && !postselect && !injected && !probs && !state_re && h->tier != kTierWide;
  if (chunked) {
    uint8_t* dst[4] = {nullptr, nullptr, nullptr, nullptr};
    const uint8_t* src[4] = {L.meas, L.det, L.obs, reinterpret_cast<const uint8_t*>(d_expv)};
    for (int i = 0; i < 4; ++i)
      if (i < 3 || t.nexp > 0) dst[i] = static_cast<uint8_t*>(host_rows(size_t(shots) * rbs[i]));
    out->measurements = dst[0];
    out->detectors = dst[1];
    out->observables = dst[2];
    out->exp_vals = reinterpret_cast<double*>(dst[3]);
    out->record_rows = shots;
    out->observable_ones = static_cast<uint64_t*>(calloc(nobs1, 8));
    e = chunked_rows(h, L, src, rbs, dst, out->observable_ones, &derr, &ms);
    if (e != hipSuccess) {
      ticit_hip_result_free(out);
      return from_hip(e, "sampling in chunks");
    }
  } else {
    CK(hipEventRecord(h->ev_start, nullptr), "recording event");
    int dump_k = 0;
    CK(launch(h, L, nullptr, d_state, d_state ? d_state + (size_t(1) << h->tab.max_k) : nullptr, &dump_k),
       "launching the sampling kernel");
    CK(hipEventRecord(h->ev_stop, nullptr), "recording event");
    CK(hipEventSynchronize(h->ev_stop), "running the sampling kernel");
    (void)hipEventElapsedTime(&ms, h->ev_start, h->ev_stop);
    CK(hipMemcpy(&derr, h->d_error, sizeof(int), hipMemcpyDeviceToHost), "reading the error flag");
    if (!(derr & kDevErrAny)) {
      // postselection: the kept shots' rows move to the front on the
      // device, in shot order (src/sampler/prepared.rs:881-1028 appends
      // only survivors), and only those are copied back
      uint8_t* src[4] = {L.meas, L.det, L.obs, reinterpret_cast<uint8_t*>(d_expv)};
      if (postselect) {
        CompactRows sets[4] = {};
        int nsets = 0, which[4] = {0, 0, 0, 0};
        for (int i = 0; i < 4; ++i) {
          if (!survivors[i]) continue;
          which[nsets] = i;
          sets[nsets++] = CompactRows{src[i], survivors[i], rbs[i], 0};
        }
        CK(launch_compaction(d_keep, shots, tile_off, d_acc, sets, nsets, nullptr), "compacting survivor rows");
        unsigned long long acc = 0;
        CK(hipMemcpy(&acc, d_acc, 8, hipMemcpyDeviceToHost), "reading the survivor count");
        out->accepted = acc;
        out->discarded = shots - acc;
        for (int j = 0; j < nsets; ++j) src[which[j]] = survivors[which[j]];
      }
      const uint64_t rows = out->accepted;
      CopyBack back;
      out->observable_ones = static_cast<uint64_t*>(calloc(nobs1, 8));
      back.add(out->observable_ones, L.obs_ones, size_t(t.nobs) * 8);
      if (keep_records) {
        out->record_rows = rows;
        uint8_t* dst[4] = {nullptr, nullptr, nullptr, nullptr};
        for (int i = 0; i < 4; ++i) {
          if (i == 3 && t.nexp == 0) continue;
          dst[i] = static_cast<uint8_t*>(host_rows(size_t(rows) * rbs[i]));
          back.add(dst[i], src[i], size_t(rows) * rbs[i]);
        }
        out->measurements = dst[0];
        out->detectors = dst[1];
        out->observables = dst[2];
        out->exp_vals = reinterpret_cast<double*>(dst[3]);
      }
      if (probs) back.add(probs, d_probs, size_t(t.nmeasure) * shots * 8);
      if (state_re) {
        const size_t nd = size_t(1) << dump_k;
        back.add(state_re, d_state, nd * 8);
        back.add(state_im, d_state + (size_t(1) << h->tab.max_k), nd * 8);
        *state_k = dump_k;
      }
      e = back.run(h);
      if (e != hipSuccess) {
        ticit_hip_result_free(out);
        return from_hip(e, "reading results");
      }
    }
  }
#undef CK
  if (derr & kDevErrAny) {
    ticit_hip_result_free(out);
    std::memset(out, 0, sizeof(*out));
    if (derr & kDevErrBadProbability) return fail(TICIT_HIP_INVALID_INPUT, "probability must be between 0 and 1");
    if (derr & kDevErrUnfairPin)
      return fail(TICIT_HIP_INVALID_INPUT,
                  "pinned measurement branch has a probability other than one half (not one half within 1e-9); "
                  "pinning it would bias sampling");
    return fail(TICIT_HIP_INVALID_INPUT, "sampled an impossible active measurement branch");
  }
  if (t.nobs > 0 && h->opt.observable < uint32_t(t.nobs)) out->logical_errors = out->observable_ones[h->opt.observable];
  out->kernel_s = double(ms) * 1e-3;
  out->sample_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return TICIT_HIP_OK;
}

// The noiseless reference sample (reference_sample_for_program,
// src/sampler/prepared.rs:355-405): one shot with every exogenous symbol 0
// and the branch stream started from state 1, as the reference's one-shot
// executor runs it; rows come back unpacked with the handle's own expected
// bits taken out again.
ticit_hip_status reference_rows(ticit_hip_program* h, std::vector<uint8_t>& det, std::vector<uint8_t>& obs) {
  const size_t nd = size_t(h->tab.ndet), no = size_t(h->tab.nobs);
  det.assign(nd, 0);
  obs.assign(no, 0);
  if (nd == 0 && no == 0) return TICIT_HIP_OK;
  if (h->opt.precision != 0) {
    // the reference runs in FP64: an FP64 twin with the same pins
    ticit_hip_options ro = h->opt;
    ro.precision = 0;
    ro.tier = 0;
    ticit_hip_program* rh = nullptr;
    ticit_hip_status st = create_from_program(h->prog, &ro, &rh, &h->pins, nullptr);
    if (st != TICIT_HIP_OK) return st;
    st = reference_rows(rh, det, obs);
    delete rh;
    return st;
  }
  ticit_hip_result r{};
  ticit_hip_status st = run_host(h, 1, 0, 0, 1, false, nullptr, nullptr, nullptr, nullptr, &r, nullptr, nullptr,
                                 nullptr, true);
  if (st == TICIT_HIP_OK) {
    for (size_t d = 0; d < nd; ++d)
      det[d] = uint8_t((r.detectors[d] != 0) != (!h->expected.detectors.empty() && h->expected.detectors[d] != 0));
    for (size_t o = 0; o < no; ++o)
      obs[o] = uint8_t((r.observables[o] != 0) != (!h->expected.observables.empty() && h->expected.observables[o] != 0));
  }
  ticit_hip_result_free(&r);
  return st;
}

// Staging layout of one postselected device launch: keep flags, compaction
// tile offsets, then the full rows of every shot (each region 256-byte
// aligned).
struct PostStaging {
  size_t keep_b = 0, off_b = 0, rows_b[4] = {0, 0, 0, 0};
  uint32_t rbs[4] = {0, 0, 0, 0};
  size_t total() const { return keep_b + off_b + rows_b[0] + rows_b[1] + rows_b[2] + rows_b[3]; }
};

DevLaunch post_launch_request(const ticit_hip_program* h, uint64_t shots, uint64_t seed, uint64_t first_shot,
                              int bit_packed, int keep_records, uint64_t* observable_ones, double* exp_vals,
                              PostStaging& stg) {
  const HostTables& t = h->tab;
  DevLaunch L{};
  L.shots = shots;
  L.seed = seed;
  L.first_shot = first_shot;
  L.bit_packed = bit_packed != 0;
  L.keep_records = keep_records != 0;
  L.meas_row = h->row_bytes(uint64_t(t.nrec), bit_packed);
  L.det_row = h->row_bytes(uint64_t(t.ndet), bit_packed);
  L.obs_row = h->row_bytes(uint64_t(t.nobs), bit_packed);
  L.obs_ones = reinterpret_cast<unsigned long long*>(observable_ones);
  L.error = h->d_error;
  const bool want_exp = keep_records && t.nexp > 0 && exp_vals;
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t n = size_t(shots);
  stg.keep_b = up(n);
  stg.off_b = up(compaction_scratch_words(shots) * 8);
  const uint32_t rbs[4] = {L.meas_row, L.det_row, L.obs_row, want_exp ? uint32_t(t.nexp) * 8u : 0u};
  for (int i = 0; i < 4; ++i) {
    stg.rbs[i] = rbs[i];
    stg.rows_b[i] = keep_records ? up(n * rbs[i]) : 0;
  }
  return L;
}

// Samples into the staging at `base` (stg.total() bytes of device memory),
// then moves the survivors' rows to the caller's buffers and their count to
// *acc, all on `st`.
ticit_hip_status post_launch_into(ticit_hip_program* h, DevLaunch& L, const PostStaging& stg, uint8_t* base,
                                  int keep_records, uint8_t* meas, uint8_t* det, uint8_t* obs, double* exp_vals,
                                  unsigned long long* acc, hipStream_t st) {
  L.keep = base;
  uint64_t* tile_off = reinterpret_cast<uint64_t*>(base + stg.keep_b);
  uint8_t* p = base + stg.keep_b + stg.off_b;
  uint8_t* stage[4] = {nullptr, nullptr, nullptr, nullptr};
  if (keep_records)
    for (int i = 0; i < 4; ++i) {
      stage[i] = stg.rows_b[i] ? p : nullptr;
      p += stg.rows_b[i];
    }
  L.meas = stage[0];
  L.det = stage[1];
  L.obs = stage[2];
  L.expv = reinterpret_cast<double*>(stage[3]);
  hipError_t e = launch(h, L, st);
  if (e != hipSuccess) return from_hip(e, "launching the sampling kernel");
  CompactRows sets[4] = {};
  int nsets = 0;
  uint8_t* dsts[4] = {meas, det, obs, reinterpret_cast<uint8_t*>(exp_vals)};
  if (keep_records)
    for (int i = 0; i < 4; ++i)
      if (stage[i] && stg.rbs[i]) sets[nsets++] = CompactRows{stage[i], dsts[i], stg.rbs[i], 0};
  e = launch_compaction(L.keep, L.shots, tile_off, acc, sets, nsets, st);
  if (e != hipSuccess) return from_hip(e, "compacting survivor rows");
  return TICIT_HIP_OK;
}

}  // namespace

extern "C" {

const char* ticit_hip_last_error(void) { return g_error.c_str(); }

ticit_hip_status ticit_hip_plan_text(const char* text, size_t len, uint64_t** blob, size_t* nwords) {
  return ticit_hip_plan_text_postselected(text, len, nullptr, 0, blob, nwords);
}

ticit_hip_status ticit_hip_plan_text_postselected(const char* text, size_t len, const uint8_t* mask,
                                                  size_t mask_len, uint64_t** blob, size_t* nwords) {
  if (!text || !blob || !nwords || (mask_len && !mask)) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  try {
    Program p = compile_circuit_text(std::string(text, len), nullptr,
                                     std::vector<uint8_t>(mask, mask + mask_len));
    std::vector<uint64_t> w = encode_program(p);
    *blob = static_cast<uint64_t*>(malloc(w.size() * 8));
    std::memcpy(*blob, w.data(), w.size() * 8);
    *nwords = w.size();
    return TICIT_HIP_OK;
  } catch (const TicitError& e) {
    return fail(ticit_hip_status(e.kind), e.what());
  } catch (const std::exception& e) {
    return fail(TICIT_HIP_INTERNAL, e.what());
  }
}

void ticit_hip_blob_free(uint64_t* blob) { free(blob); }

ticit_hip_status ticit_hip_plan_features(const uint64_t* blob, size_t nwords, double* out, size_t nout) {
  if (!blob || (!out && nout)) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  Program p;
  std::string err;
  if (!decode_program(blob, nwords, p, err)) return fail(TICIT_HIP_INVALID_INPUT, err);
  try {
    double f[TICIT_HIP_NFEATURES] = {};
    HostTables dense = lower_to_device(p, false);
    ComponentPlan plan = plan_components(p);
    int nrot = 0, nprom = 0;
    for (const Instr& in : p.instrs) {
      nrot += in.op == Op::Rotate;
      nprom += in.op == Op::Promote;
    }
    auto packed = [](int n) { return double((n + 7) / 8); };
    f[TICIT_HIP_FEATURE_QUBITS] = p.n;
    f[TICIT_HIP_FEATURE_INITIAL_K] = p.initial_k;
    f[TICIT_HIP_FEATURE_MAX_K] = p.max_k;
    f[TICIT_HIP_FEATURE_SYMBOLS] = p.nsymbols;
    f[TICIT_HIP_FEATURE_RECORDS] = p.nrecords;
    f[TICIT_HIP_FEATURE_DETECTORS] = p.ndetectors;
    f[TICIT_HIP_FEATURE_INSTRUCTIONS] = double(p.instrs.size());
    f[TICIT_HIP_FEATURE_MEASUREMENTS] = dense.nmeasure;
    f[TICIT_HIP_FEATURE_ROTATIONS] = nrot;
    f[TICIT_HIP_FEATURE_PROMOTIONS] = nprom;
    f[TICIT_HIP_FEATURE_QUANTUM_OPS] = double(dense.quantum_ops);
    f[TICIT_HIP_FEATURE_EXPRESSION_TERMS] = double(dense.expression_terms);
    f[TICIT_HIP_FEATURE_SELECTED] = p.use_components ? 1.0 : 0.0;
    f[TICIT_HIP_FEATURE_PLANNED] = plan.planned ? 1.0 : 0.0;
    f[TICIT_HIP_FEATURE_DENSE_WORK] = plan.dense_work;
    f[TICIT_HIP_FEATURE_COMPONENT_WORK] = plan.component_work;
    f[TICIT_HIP_FEATURE_DENSE_PEAK] = double(plan.dense_peak);
    f[TICIT_HIP_FEATURE_PEAK_LIVE] = double(plan.peak_live);
    f[TICIT_HIP_FEATURE_ALLOCATED] = double(plan.allocated);
    f[TICIT_HIP_FEATURE_MAX_COMPONENT_K] = plan.max_component_k;
    f[TICIT_HIP_FEATURE_MERGES] = plan.merges;
    f[TICIT_HIP_FEATURE_DENSE_AMP_TRAFFIC] = dense.amp_traffic;
    if (plan.planned) f[TICIT_HIP_FEATURE_COMPONENT_AMP_TRAFFIC] = lower_to_device(p, true).amp_traffic;
    f[TICIT_HIP_FEATURE_OUTPUT_BYTES] = packed(p.nrecords) + packed(p.ndetectors) + packed(int(p.observables.size()));
    std::memcpy(out, f, std::min<size_t>(nout, TICIT_HIP_NFEATURES) * sizeof(double));
    return TICIT_HIP_OK;
  } catch (const TicitError& e) {
    return fail(ticit_hip_status(e.kind), e.what());
  } catch (const std::exception& e) {
    return fail(TICIT_HIP_INTERNAL, e.what());
  }
}

ticit_hip_status ticit_hip_program_create(const uint64_t* blob, size_t nwords, const ticit_hip_options* options,
                                          ticit_hip_program** out) {
  if (!blob || !out) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  Program p;
  std::string err;
  if (!decode_program(blob, nwords, p, err)) return fail(TICIT_HIP_INVALID_INPUT, err);
  try {
    return create_from_program(std::move(p), options, out);
  } catch (const std::exception& e) {
    return fail(TICIT_HIP_INTERNAL, e.what());
  }
}

ticit_hip_status ticit_hip_program_create_pinned(const uint64_t* blob, size_t nwords,
                                                const ticit_hip_options* options, const ticit_hip_parity* pins,
                                                size_t npins, ticit_hip_program** out) {
  return ticit_hip_program_create_ex(blob, nwords, options, pins, npins, nullptr, out);
}

ticit_hip_status ticit_hip_program_create_ex(const uint64_t* blob, size_t nwords, const ticit_hip_options* options,
                                            const ticit_hip_parity* pins, size_t npins,
                                            const ticit_hip_reference_bits* reference, ticit_hip_program** out) {
  if (!blob || !out || (npins && !pins)) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  ReferenceRequest rq;
  if (reference) {
    if ((reference->ndetectors && !reference->detectors) || (reference->nobservables && !reference->observables))
      return fail(TICIT_HIP_INVALID_INPUT, "null argument");
    rq.bits.detectors.assign(reference->detectors, reference->detectors + reference->ndetectors);
    rq.bits.observables.assign(reference->observables, reference->observables + reference->nobservables);
    rq.normalize = reference->normalize_syndromes != 0;
  }
  std::vector<MeasurementParity> ps(npins);
  for (size_t i = 0; i < npins; ++i) {
    if (pins[i].nrecords && !pins[i].records) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
    ps[i].records.assign(pins[i].records, pins[i].records + pins[i].nrecords);
    ps[i].value = pins[i].value != 0;
  }
  Program p;
  std::string err;
  if (!decode_program(blob, nwords, p, err)) return fail(TICIT_HIP_INVALID_INPUT, err);
  try {
    return create_from_program(std::move(p), options, out, &ps, reference ? &rq : nullptr);
  } catch (const std::exception& e) {
    return fail(TICIT_HIP_INTERNAL, e.what());
  }
}

ticit_hip_status ticit_hip_compile_text(const char* text, size_t len, const ticit_hip_options* options,
                                        ticit_hip_program** out) {
  if (!text || !out) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  try {
    Program p = compile_circuit_text(std::string(text, len));
    return create_from_program(std::move(p), options, out);
  } catch (const TicitError& e) {
    return fail(ticit_hip_status(e.kind), e.what());
  } catch (const std::exception& e) {
    return fail(TICIT_HIP_INTERNAL, e.what());
  }
}

void ticit_hip_program_destroy(ticit_hip_program* program) { delete program; }

ticit_hip_status ticit_hip_reference_sample(ticit_hip_program* h, uint8_t* detectors, uint8_t* observables) {
  if (!h || (h->tab.ndet > 0 && !detectors) || (h->tab.nobs > 0 && !observables))
    return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  std::vector<uint8_t> d, o;
  try {
    ticit_hip_status st = reference_rows(h, d, o);
    if (st != TICIT_HIP_OK) return st;
  } catch (const std::exception& e) {
    return fail(TICIT_HIP_INTERNAL, e.what());
  }
  if (!d.empty()) std::memcpy(detectors, d.data(), d.size());
  if (!o.empty()) std::memcpy(observables, o.data(), o.size());
  return TICIT_HIP_OK;
}

ticit_hip_status ticit_hip_program_info_get(const ticit_hip_program* h, ticit_hip_program_info* out) {
  if (!h || !out) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  std::memset(out, 0, sizeof(*out));
  out->qubits = uint64_t(h->prog.n);
  out->measurement_records = uint64_t(h->prog.nrecords);
  out->detectors = uint64_t(h->prog.ndetectors);
  out->observables = uint64_t(h->prog.observables.size());
  out->max_active_qubits = uint64_t(h->prog.max_k);
  out->symbols = uint64_t(h->prog.nsymbols);
  out->instructions = uint64_t(h->prog.instrs.size());
  out->measure_instructions = uint64_t(h->tab.nmeasure);
  out->coin_instructions = uint64_t(h->tab.ncoin);
  out->tier = h->tier;
  out->precision = h->opt.precision;
  out->wide_batch = h->wide_batch;
  auto packed = [](int n) { return double((n + 7) / 8); };
  out->output_bytes_per_shot = packed(h->tab.nrec) + packed(h->tab.ndet) + packed(h->tab.nobs);
  out->state_bytes_per_shot = h->tab.amp_traffic * (h->opt.precision == 1 ? 8.0 : 16.0);
  out->components = h->tab.components ? 1 : 0;
  out->planner_source = h->planner_source;
  out->table_words[0] = h->lane.pw;
  out->table_words[1] = h->lane.sw;
  out->table_words[2] = h->lane.rw;
  out->table_words[3] = h->lane.dw;
  out->expectation_values = uint64_t(h->tab.nexp);
  out->postselected_detectors = uint64_t(h->tab.npost);
  if (h->tier == kTierLane) {
    uint32_t row = std::max({h->row_bytes(uint64_t(h->tab.nrec), true), h->row_bytes(uint64_t(h->tab.ndet), true),
                             h->row_bytes(uint64_t(h->tab.nobs), true)});
    lane_tier_occupancy(h->tab.max_k, h->lane, row, &out->lds_bytes_per_block, &out->blocks_per_cu);
    out->lane_store = h->lane.store;
    LaneLayout y = h->lane;
    y.stage_bytes = 64 * ((row + 7) / 8 * 8);
    out->lane_region_bytes = uint32_t(lane_region_bytes(h->tab.max_k, y));
  } else if (h->tier == kTierGroup) {
    out->lds_bytes_per_block = uint32_t(group_tier_lds_bytes(h->tab.state_dim, h->group, h->group_tps));
  }
  out->planner_p_components = h->planner_p;
  out->state_dim = h->tab.state_dim;
  out->component_count = h->tab.comp_off.size();
  out->max_local_k = h->tab.max_local_k;
  out->algorithmic_bytes_per_shot =
      out->output_bytes_per_shot + (h->tier == kTierWide ? out->state_bytes_per_shot : 0.0);
  return TICIT_HIP_OK;
}

ticit_hip_status ticit_hip_sample(ticit_hip_program* h, uint64_t shots, uint64_t seed, int bit_packed,
                                  int keep_records, ticit_hip_result* out) {
  if (!h || !out) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  return run_host(h, shots, seed, bit_packed, keep_records, false, nullptr, nullptr, nullptr, nullptr, out);
}

ticit_hip_status ticit_hip_sample_injected(ticit_hip_program* h, uint64_t shots, const uint64_t* exo,
                                           const double* uniforms, const uint64_t* coins, int bit_packed,
                                           ticit_hip_result* out, double* probs) {
  if (!h || !out) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  return run_host(h, shots, 0, bit_packed, 1, true, exo, uniforms, coins, probs, out);
}

ticit_hip_status ticit_hip_final_state(ticit_hip_program* h, uint64_t shots, const uint64_t* exo,
                                       const double* uniforms, const uint64_t* coins, double* re, double* im,
                                       int* k) {
  if (!h || !re || !im || !k) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  ticit_hip_result r;
  ticit_hip_status s = run_host(h, shots, 0, 1, 1, true, exo, uniforms, coins, nullptr, &r, re, im, k);
  if (s == TICIT_HIP_OK) ticit_hip_result_free(&r);
  return s;
}

ticit_hip_status ticit_hip_sample_device(ticit_hip_program* h, uint64_t shots, uint64_t seed, uint64_t first_shot,
                                         int bit_packed, int keep_records, uint8_t* meas, uint8_t* det,
                                         uint8_t* obs, uint64_t* observable_ones, double* exp_vals,
                                         void* stream) {
  if (!h || !observable_ones) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  if (keep_records && (!meas || !det || !obs)) return fail(TICIT_HIP_INVALID_INPUT, "null row buffer");
  if (h->tab.npost > 0)
    return fail(TICIT_HIP_UNSUPPORTED,
                "postselected programs return their surviving rows through ticit_hip_sample or "
                "ticit_hip_sample_device_postselected");
  if (shots == 0) return TICIT_HIP_OK;
  const HostTables& t = h->tab;
  DevLaunch L{};
  L.shots = shots;
  L.seed = seed;
  L.first_shot = first_shot;
  L.bit_packed = bit_packed != 0;
  L.keep_records = keep_records != 0;
  L.meas_row = h->row_bytes(uint64_t(t.nrec), bit_packed);
  L.det_row = h->row_bytes(uint64_t(t.ndet), bit_packed);
  L.obs_row = h->row_bytes(uint64_t(t.nobs), bit_packed);
  L.meas = meas;
  L.det = det;
  L.obs = obs;
  L.obs_ones = reinterpret_cast<unsigned long long*>(observable_ones);
  L.expv = keep_records && t.nexp > 0 ? exp_vals : nullptr;
  L.error = h->d_error;
  hipError_t e = launch(h, L, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return from_hip(e, "launching the sampling kernel");
  return TICIT_HIP_OK;
}

ticit_hip_status ticit_hip_sample_device_postselected(ticit_hip_program* h, uint64_t shots, uint64_t seed,
                                                     uint64_t first_shot, int bit_packed, int keep_records,
                                                     uint8_t* meas, uint8_t* det, uint8_t* obs,
                                                     uint64_t* observable_ones, double* exp_vals,
                                                     uint64_t* accepted, void* stream) {
  if (!h || !observable_ones || !accepted) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  if (keep_records && (!meas || !det || !obs)) return fail(TICIT_HIP_INVALID_INPUT, "null row buffer");
  hipStream_t st = static_cast<hipStream_t>(stream);
  auto* acc = reinterpret_cast<unsigned long long*>(accepted);
  const HostTables& t = h->tab;
  if (t.npost == 0 || shots == 0) {
    ticit_hip_status s0 = ticit_hip_sample_device(h, shots, seed, first_shot, bit_packed, keep_records, meas, det,
                                                  obs, observable_ones, exp_vals, stream);
    if (s0 != TICIT_HIP_OK) return s0;
    hipError_t e = launch_store_u64(acc, shots, st);
    return e == hipSuccess ? TICIT_HIP_OK : from_hip(e, "writing the survivor count");
  }
  PostStaging stg;
  DevLaunch L = post_launch_request(h, shots, seed, first_shot, bit_packed, keep_records, observable_ones, exp_vals,
                                    stg);
  std::lock_guard<std::mutex> lock(h->post_mu);
  const size_t need = stg.total();
  hipError_t e = hipSuccess;
  if (!h->post_done) {
    e = hipEventCreateWithFlags(&h->post_done, hipEventDisableTiming);
    if (e != hipSuccess) return from_hip(e, "creating the staging event");
  }
  if (need > h->post_bytes) {
    // the previous call may still read the old buffer
    (void)hipEventSynchronize(h->post_done);
    if (h->post_buf) (void)hipFree(h->post_buf);
    h->post_buf = nullptr;
    h->post_bytes = 0;
    e = hipMalloc(&h->post_buf, need);
    if (e != hipSuccess) return from_hip(e, "allocating the postselection staging");
    h->post_bytes = need;
  } else if (h->post_pending) {
    e = hipStreamWaitEvent(st, h->post_done, 0);
    if (e != hipSuccess) return from_hip(e, "ordering the staging reuse");
  }
  ticit_hip_status s1 = post_launch_into(h, L, stg, static_cast<uint8_t*>(h->post_buf), keep_records, meas, det, obs,
                                         exp_vals, acc, st);
  if (s1 != TICIT_HIP_OK) return s1;
  e = hipEventRecord(h->post_done, st);
  if (e != hipSuccess) return from_hip(e, "recording the staging event");
  h->post_pending = true;
  return TICIT_HIP_OK;
}

ticit_hip_status ticit_hip_sample_device_postselected_noalloc(ticit_hip_program* h, uint64_t shots, uint64_t seed,
                                                             uint64_t first_shot, int bit_packed, int keep_records,
                                                             uint8_t* meas, uint8_t* det, uint8_t* obs,
                                                             uint64_t* observable_ones, double* exp_vals,
                                                             uint64_t* accepted, void* stream) {
  if (!h || !observable_ones || !accepted) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  if (keep_records && (!meas || !det || !obs)) return fail(TICIT_HIP_INVALID_INPUT, "null row buffer");
  if (h->tier == kTierWide)
    return fail(TICIT_HIP_UNSUPPORTED,
                "the wide tier keeps its per-shot states in a workspace the library allocates; "
                "ticit_hip_sample_device_postselected_noalloc serves lane- and group-tier programs");
  if (h->tier == kTierLane && h->lane.store != kLaneLds)
    return fail(TICIT_HIP_UNSUPPORTED,
                "this lane-tier program keeps its tables in device memory the library allocates; "
                "ticit_hip_sample_device_postselected_noalloc serves programs whose tables fit LDS");
  int current = -1;
  hipError_t e = hipGetDevice(&current);
  if (e != hipSuccess) return from_hip(e, "reading the current device");
  if (current != h->device) return fail(TICIT_HIP_INVALID_INPUT, "the program's device must be the current device");
  hipStream_t st = static_cast<hipStream_t>(stream);
  auto* acc = reinterpret_cast<unsigned long long*>(accepted);
  const HostTables& t = h->tab;
  if (t.npost == 0 || shots == 0) {
    // lane and group launches into caller buffers allocate nothing
    ticit_hip_status s0 = ticit_hip_sample_device(h, shots, seed, first_shot, bit_packed, keep_records, meas, det,
                                                  obs, observable_ones, exp_vals, stream);
    if (s0 != TICIT_HIP_OK) return s0;
    e = launch_store_u64(acc, shots, st);
    return e == hipSuccess ? TICIT_HIP_OK : from_hip(e, "writing the survivor count");
  }
  PostStaging stg;
  DevLaunch L = post_launch_request(h, shots, seed, first_shot, bit_packed, keep_records, observable_ones, exp_vals,
                                    stg);
  const size_t need = stg.total();
  if (need > kStaticScratchBytes)
    return fail(TICIT_HIP_INVALID_INPUT,
                "this call stages " + std::to_string(need) + " bytes, more than the " +
                    std::to_string(kStaticScratchBytes) +
                    "-byte static scratch; split the shots into calls with first_shot");
  void* base = nullptr;
  e = static_scratch_address(h->device, &base);
  if (e != hipSuccess) return from_hip(e, "looking up the static scratch");
  // every staged byte the compaction reads is written by this launch first
  // (keep flags and rows of every shot, tile offsets by the count pass), so
  // nothing left by an earlier call needs clearing
  return post_launch_into(h, L, stg, static_cast<uint8_t*>(base), keep_records, meas, det, obs, exp_vals, acc, st);
}

void ticit_hip_result_free(ticit_hip_result* r) {
  if (!r) return;
  free(r->measurements);
  free(r->detectors);
  free(r->observables);
  free(r->observable_ones);
  free(r->exp_vals);
  r->measurements = r->detectors = r->observables = nullptr;
  r->observable_ones = nullptr;
  r->exp_vals = nullptr;
}

int ticit_hip_device_query(int64_t* props) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return 0;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  int v[6] = {0, 0, 0, 0, 0, 0};
  (void)hipDeviceGetAttribute(&v[0], hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  (void)hipDeviceGetAttribute(&v[1], hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipDeviceGetAttribute(&v[2], hipDeviceAttributeClockRate, dev);
  (void)hipDeviceGetAttribute(&v[3], hipDeviceAttributeWarpSize, dev);
  (void)hipDeviceGetAttribute(&v[4], hipDeviceAttributeMemoryClockRate, dev);
  (void)hipDeviceGetAttribute(&v[5], hipDeviceAttributeMemoryBusWidth, dev);
  for (int i = 0; i < 6; ++i) props[i] = v[i];
  return n;
}

}  // extern "C"
