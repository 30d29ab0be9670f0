// Batched sampling of many programs (ticit_hip_batch_*): hipGraphs captured
// once whose replay runs every program of the batch.
//
// Programs are grouped by the tier they run on:
//   * group tier: buckets of programs with similar per-shot LDS footprints,
//     each bucket one persistent kernel that pulls (program, shot block)
//     work units from an atomic counter; threads per shot (8 to 512: 8, 4
//     or 2 shots per wave below 64, else one shot per workgroup) are picked
//     per bucket by timing each candidate on a sample of the bucket's
//     shots;
//   * lane tier: one launch per program;
//   * wide tier: each program's kernel sequence, split over two more graphs
//     replayed on two more streams (balanced by estimated HBM bytes).
// The graphs are captured once at creation and replayed by every launch; a
// device word added to every program's seed lets replays draw fresh shots.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../kernels/compact.hpp"
#include "handle.hpp"

using namespace ticit;

struct ticit_hip_batch {
  struct Bucket {
    std::vector<uint32_t> items;
    uint32_t size_class = 0;    // power-of-two class of the per-shot LDS bytes
    int width_class = 0;        // 0: widths up to 10; 1: 11-12 (128+ threads); 2: 13 (512)
    uint32_t region = 0;        // LDS bytes per shot: the largest of the bucket
    int max_local_k = 0;
    std::vector<int> candidates;
    int tps = 256;
    int grid = 0;
    double tuned_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // sample time per candidate (8 ... 512, by log2 - 3)
    DevProgram* d_progs = nullptr;
    DevLaunch* d_launches = nullptr;
    GroupLayout* d_layouts = nullptr;
    uint64_t* d_unit_off = nullptr;
    uint64_t total_units = 0;
  };
  int device = 0;
  int bit_packed = 1, keep_records = 1;
  int group_grid_percent = 100;  // persistent group grid, percent of full occupancy
  int wide_cu_percent = 0;        // > 0: CU-masked streams split the device between wide and group work
  std::vector<ticit_hip_program*> progs;
  std::vector<DevLaunch> launches;
  std::vector<uint32_t> lane_items, wide_items;
  std::vector<Bucket> buckets;
  // Postselected items: the kernels write full rows and keep flags into
  // batch-owned staging; compaction kernels captured after the item's work
  // move the survivors to the caller's buffers and write its count.
  struct PostItem {
    uint8_t* keep = nullptr;
    uint64_t* tile_off = nullptr;
    unsigned long long* accepted = nullptr;
    CompactRows sets[4] = {};
    int nsets = 0;
  };
  std::map<uint32_t, PostItem> post;
  std::vector<void*> allocs;
  unsigned long long* d_counters = nullptr;  // one per bucket
  uint64_t* d_seed_offset = nullptr;
  int* d_error = nullptr;
  WideWorkspace* ws = nullptr;
  // Wide items run in kWideSequences sequences on as many streams, split
  // by their estimated HBM bytes, each with its own workspace: one
  // sequence's small, latency-bound kernels then overlap another's. The
  // first is wide_items on `side`; these are the others.
  struct WideSeq {
    std::vector<uint32_t> items;
    WideWorkspace* ws = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t join = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<WideSeq> more;
  // lane items with tables in device memory share these regions: they run
  // one after another on `main`
  LaneWorkspace lane_ws{};
  hipStream_t main = nullptr, side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr, join_main = nullptr;
  hipGraph_t graph = nullptr;        // lane and group work
  hipGraphExec_t exec = nullptr;
  hipGraph_t wide_graph = nullptr;   // wide work, replayed on `side` beside `graph`
  hipGraphExec_t wide_exec = nullptr;
  hipStream_t last = nullptr;
  bool split = false;  // main/side are CU-masked halves of the device
  int group_cus = 0;   // CUs the group/lane stream may use when split

  ~ticit_hip_batch() {
    if (last) (void)hipStreamSynchronize(last);
    if (side) (void)hipStreamSynchronize(side);
    for (auto& q : more)
      if (q.stream) (void)hipStreamSynchronize(q.stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (wide_exec) (void)hipGraphExecDestroy(wide_exec);
    if (wide_graph) (void)hipGraphDestroy(wide_graph);
    for (auto& q : more) {
      if (q.exec) (void)hipGraphExecDestroy(q.exec);
      if (q.graph) (void)hipGraphDestroy(q.graph);
      if (q.join) (void)hipEventDestroy(q.join);
      if (q.stream) (void)hipStreamDestroy(q.stream);
      wide_workspace_free(q.ws);
    }
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    if (join_main) (void)hipEventDestroy(join_main);
    if (main) (void)hipStreamDestroy(main);
    if (side) (void)hipStreamDestroy(side);
    wide_workspace_free(ws);
    for (void* p : allocs) (void)hipFree(p);
  }

  template <class T>
  hipError_t alloc(T** p, size_t n) {
    void* d = nullptr;
    hipError_t e = hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(T));
    if (e == hipSuccess) {
      allocs.push_back(d);
      *p = static_cast<T*>(d);
    }
    return e;
  }
  template <class T>
  hipError_t upload(T** p, const std::vector<T>& v) {
    hipError_t e = alloc(p, v.size());
    if (e == hipSuccess && !v.empty()) e = hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    return e;
  }
};

namespace {

int tps_slot(int tps) {
  int s = 0;
  while ((8 << s) < tps) ++s;
  return s;
}

// Shots up to which a batch runs a width-11 to 13 program in the group
// tier when the automatic choice gave it the wide tier. A lone launch of
// such a program is faster wide, but the batch's wide sequence is its
// critical path while the group stream idles, and a small program's shots
// pack into a persistent group bucket: on 256 random circuits (14 of them
// moved) the batch took 0.97x the time, on 512 corpus programs of 2e4 shots
// (88 moved) 0.94-0.97x; adding width 13 (512 threads per shot) took the
// latter to 0.94x again (28 more moved) and left the former unchanged.
// With the group tier's narrower default shapes the 256 circuits (1e4 to
// 9.8e5 shots each) gain further up to 4e5 shots: 2.27-2.32 s at 5e4,
// 2.23 s at 2e5, 2.17-2.20 s at 4e5 (13 more moved), and 2.88 s at 1e6,
// where the group stream becomes the critical path.
constexpr uint64_t kBatchGroupMaxShots = 400000;

// Wide sequences of a batch, each on its own stream (C2 and C4 with two:
// 0.93x and 0.94x the time of one).
constexpr size_t kWideSequences = 2;

bool group_in_batch(const ticit_hip_program* h, uint64_t shots, int lds_limit) {
  const HostTables& t = h->tab;
  auto fits = [&](int tps) {
    return t.max_local_k <= group_max_local_k(tps) && group_tier_lds_bytes(t.state_dim, h->group, tps) <= size_t(lds_limit);
  };
  return h->opt.tier == 0 && h->opt.precision == 0 && !h->bit_heavy && shots <= kBatchGroupMaxShots &&
         (fits(256) || fits(512));
}

// Tuned threads-per-shot by bucket signature, kept for the process.
std::mutex g_tune_mu;
std::map<std::tuple<int, uint32_t, int, int>, int> g_tuned;

uint32_t round_region(size_t bytes) {
  uint32_t r = 256;
  while (r < bytes) r <<= 1;
  return r;
}

// LDS bytes per shot of a bucket at `tps` threads per shot (each wave of a
// shot keeps its own bit tables, so the region grows with the waves).
uint32_t bucket_region(const ticit_hip_batch& b, const ticit_hip_batch::Bucket& k, int tps) {
  size_t r = 0;
  for (uint32_t i : k.items) {
    const ticit_hip_program* h = b.progs[i];
    r = std::max(r, group_region_bytes(h->tab.state_dim, h->group, tps));
  }
  return uint32_t(r);
}

hipError_t unit_table(const ticit_hip_batch& b, const ticit_hip_batch::Bucket& k, int tps,
                      const std::vector<uint64_t>& shots, std::vector<uint64_t>& off, uint64_t& total) {
  const int spu = group_shots_per_unit(tps);
  off.clear();
  total = 0;
  for (size_t j = 0; j < k.items.size(); ++j) {
    off.push_back(total);
    total += (shots[j] + spu - 1) / spu;
  }
  (void)b;
  return hipSuccess;
}

hipError_t run_bucket(ticit_hip_batch& b, ticit_hip_batch::Bucket& k, int tps, const DevLaunch* d_launches,
                      const uint64_t* d_unit_off, uint64_t total_units, unsigned long long* counter,
                      hipStream_t stream) {
  const uint32_t region = bucket_region(b, k, tps);
  int occ = group_batch_occupancy(tps, region, k.max_local_k);
  int cus = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b.device);
  if (b.split && b.group_cus > 0) cus = b.group_cus;
  int grid = std::max(1, occ) * std::max(1, cus);
  if (b.group_grid_percent > 0 && b.group_grid_percent < 100)
    grid = std::max(cus, grid * b.group_grid_percent / 100);
  k.grid = grid;
  return launch_group_batch(k.d_progs, d_launches, k.d_layouts, d_unit_off, uint32_t(k.items.size()), total_units,
                            counter, region, tps, k.max_local_k, grid, stream);
}

// Times every candidate on up to `sample` shots per program (more when the
// bucket's shots would not fill the grid) and keeps the fastest. Rows of
// the sample land in the real output buffers (the batch
// run overwrites them); counts and errors go to scratch words.
hipError_t autotune(ticit_hip_batch& b, ticit_hip_batch::Bucket& k, uint64_t sample) {
  if (k.candidates.size() == 1) {
    k.tps = k.candidates[0];
    return hipSuccess;
  }
  auto key = std::make_tuple(k.max_local_k, k.region, int(b.keep_records), b.group_grid_percent);
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end() &&
        std::find(k.candidates.begin(), k.candidates.end(), it->second) != k.candidates.end()) {
      k.tps = it->second;
      return hipSuccess;
    }
  }
  hipError_t e = hipSuccess;
  // Every candidate is timed on the same shots, at least twice what its
  // resident grid holds: a sample smaller than one wave of the grid times
  // per-shot latency, which favours wide shapes (16 threads per shot took
  // 0.59 ms per 1e5 shots where 128 took 2.4 ms, yet 4096 shots timed 128
  // as the faster).
  {
    int cus = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b.device);
    if (b.split && b.group_cus > 0) cus = b.group_cus;
    uint64_t need = 0, most = 0;
    for (int tps : k.candidates) {
      const uint64_t occ = uint64_t(std::max(1, group_batch_occupancy(tps, bucket_region(b, k, tps), k.max_local_k)));
      need = std::max(need, 2 * occ * uint64_t(std::max(1, cus)) * uint64_t(group_shots_per_block(tps)));
    }
    for (uint32_t i : k.items) most = std::max(most, b.launches[i].shots);
    auto covered = [&](uint64_t s) {
      uint64_t n = 0;
      for (uint32_t i : k.items) n += std::min<uint64_t>(b.launches[i].shots, s);
      return n;
    };
    while (sample < most && covered(sample) < need) sample *= 2;
  }
  int max_obs = 0;
  for (uint32_t i : k.items) max_obs = std::max(max_obs, b.progs[i]->tab.nobs);
  unsigned long long* scratch = nullptr;  // [0]: counter, [1]: error word, [8..]: counts
  if ((e = b.alloc(&scratch, size_t(max_obs) + 16)) != hipSuccess) return e;
  std::vector<DevLaunch> ls;
  std::vector<uint64_t> shots;
  for (uint32_t i : k.items) {
    DevLaunch L = b.launches[i];
    L.shots = std::min<uint64_t>(L.shots, sample);
    L.obs_ones = scratch + 8;
    L.error = reinterpret_cast<int*>(scratch + 1);
    L.seed_offset = nullptr;
    ls.push_back(L);
    shots.push_back(L.shots);
  }
  DevLaunch* d_ls = nullptr;
  if ((e = b.upload(&d_ls, ls)) != hipSuccess) return e;
  hipEvent_t t0, t1;
  (void)hipEventCreate(&t0);
  (void)hipEventCreate(&t1);
  double best = 1e300;
  for (int tps : k.candidates) {
    std::vector<uint64_t> off;
    uint64_t total = 0;
    unit_table(b, k, tps, shots, off, total);
    uint64_t* d_off = nullptr;
    if ((e = b.upload(&d_off, off)) != hipSuccess) break;
    float ms = 0.f, sum = 0.f;
    for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
      (void)hipEventRecord(t0, b.main);
      e = run_bucket(b, k, tps, d_ls, d_off, total, scratch, b.main);
      (void)hipEventRecord(t1, b.main);
      if (e == hipSuccess) e = hipEventSynchronize(t1);
      if (e == hipSuccess) (void)hipEventElapsedTime(&ms, t0, t1);
      if (rep == 1) sum = ms;  // the first run warms code and caches
    }
    if (e != hipSuccess) break;
    k.tuned_ms[tps_slot(tps)] = sum;
    if (sum < best) {
      best = sum;
      k.tps = tps;
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (e == hipSuccess) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tuned[key] = k.tps;
  }
  return e;
}

}  // namespace

extern "C" {

ticit_hip_status ticit_hip_batch_create(const ticit_hip_batch_item* items, size_t count,
                                        const ticit_hip_batch_options* options, ticit_hip_batch** out) {
  if (!out || (count && !items)) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  ticit_hip_batch_options opt{};
  if (options) opt = *options;
  auto* b = new ticit_hip_batch();
  b->bit_packed = opt.bit_packed ? 1 : 0;
  b->keep_records = opt.keep_records ? 1 : 0;
  if (opt.group_grid_percent) b->group_grid_percent = int(std::min<uint32_t>(opt.group_grid_percent, 100));
  if (opt.wide_cu_percent) b->wide_cu_percent = int(std::min<uint32_t>(opt.wide_cu_percent, 95));
  hipError_t e = hipGetDevice(&b->device);
  int lds_limit = 0;
  if (e == hipSuccess) e = hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, b->device);
  auto bail_hip = [&](hipError_t err, const char* what) {
    delete b;
    return from_hip(err, what);
  };
  if (e != hipSuccess) return bail_hip(e, "no HIP device");
  if ((e = b->alloc(&b->d_seed_offset, 1)) != hipSuccess) return bail_hip(e, "allocating batch words");
  if ((e = b->alloc(&b->d_error, 1)) != hipSuccess) return bail_hip(e, "allocating batch words");

  for (size_t i = 0; i < count; ++i) {
    ticit_hip_program* h = items[i].program;
    if (!h) {
      delete b;
      return fail(TICIT_HIP_INVALID_INPUT, "batch item " + std::to_string(i) + " has no program");
    }
    if (h->device != b->device) {
      delete b;
      return fail(TICIT_HIP_INVALID_INPUT, "batch programs must live on the current device");
    }
    if (h->tab.npost > 0 && items[i].shots && !items[i].accepted) {
      delete b;
      return fail(TICIT_HIP_INVALID_INPUT,
                  "batch item " + std::to_string(i) + " is postselected and needs an accepted word");
    }
    const HostTables& t = h->tab;
    DevLaunch L{};
    L.shots = items[i].shots;
    L.seed = items[i].seed;
    L.bit_packed = b->bit_packed;
    L.keep_records = b->keep_records;
    L.meas = items[i].measurements;
    L.det = items[i].detectors;
    L.obs = items[i].observables;
    L.obs_ones = reinterpret_cast<unsigned long long*>(items[i].observable_ones);
    L.expv = b->keep_records && t.nexp > 0 ? items[i].exp_vals : nullptr;
    L.error = b->d_error;
    L.seed_offset = b->d_seed_offset;
    L.meas_row = h->row_bytes(uint64_t(t.nrec), b->bit_packed);
    L.det_row = h->row_bytes(uint64_t(t.ndet), b->bit_packed);
    L.obs_row = h->row_bytes(uint64_t(t.nobs), b->bit_packed);
    if (L.shots && ((b->keep_records && ((L.meas_row && !L.meas) || (L.det_row && !L.det) ||
                                         (L.obs_row && !L.obs))) ||
                    (t.nobs && !L.obs_ones))) {
      delete b;
      return fail(TICIT_HIP_INVALID_INPUT, "batch item " + std::to_string(i) + " lacks an output buffer");
    }
    if (t.npost > 0 && L.shots) {
      // full rows and keep flags go to staging; the captured compaction
      // fills the caller's buffers
      ticit_hip_batch::PostItem pi;
      pi.accepted = reinterpret_cast<unsigned long long*>(items[i].accepted);
      if ((e = b->alloc(&pi.keep, size_t(L.shots))) != hipSuccess) return bail_hip(e, "allocating staging");
      if ((e = b->alloc(&pi.tile_off, compaction_scratch_words(L.shots))) != hipSuccess)
        return bail_hip(e, "allocating staging");
      L.keep = pi.keep;
      if (b->keep_records) {
        uint8_t** slots[4] = {&L.meas, &L.det, &L.obs, reinterpret_cast<uint8_t**>(&L.expv)};
        const uint32_t rbs[4] = {L.meas_row, L.det_row, L.obs_row, L.expv ? uint32_t(t.nexp) * 8u : 0u};
        for (int q = 0; q < 4; ++q) {
          if (!rbs[q] || !*slots[q]) continue;
          uint8_t* stage = nullptr;
          if ((e = b->alloc(&stage, size_t(L.shots) * rbs[q])) != hipSuccess) return bail_hip(e, "allocating staging");
          pi.sets[pi.nsets++] = CompactRows{stage, *slots[q], rbs[q], 0};
          *slots[q] = stage;
        }
      }
      b->post[uint32_t(i)] = pi;
    }
    b->progs.push_back(h);
    b->launches.push_back(L);
    if (!L.shots) continue;
    if (h->tier == kTierLane) {
      b->lane_items.push_back(uint32_t(i));
    } else if (h->tier == kTierWide && !group_in_batch(h, L.shots, lds_limit)) {
      b->wide_items.push_back(uint32_t(i));
    } else {
      const uint32_t region = uint32_t(group_region_bytes(t.state_dim, h->group, 64));
      const uint32_t cls = round_region(region);
      // Buckets also split by the threads per shot a width needs, so every
      // item of a bucket fits the shape its widest item forces: an item
      // fits at the shape it was created for (or, moved from the wide
      // tier, at 256 or, width 13, 512), and a narrower shape only shrinks its
      // region.
      const int wcls = t.max_local_k > group_max_local_k(256) ? 2 : (t.max_local_k > group_max_local_k(64) ? 1 : 0);
      auto it = std::find_if(b->buckets.begin(), b->buckets.end(),
                             [&](const ticit_hip_batch::Bucket& k) {
                               return k.size_class == cls && k.width_class == wcls;
                             });
      if (it == b->buckets.end()) {
        b->buckets.push_back(ticit_hip_batch::Bucket{});
        it = b->buckets.end() - 1;
        it->size_class = cls;
        it->width_class = wcls;
      }
      it->region = std::max(it->region, region);
      it->items.push_back(uint32_t(i));
      it->max_local_k = std::max(it->max_local_k, t.max_local_k);
    }
  }

  if (b->wide_cu_percent > 0 && !b->wide_items.empty()) {
    // Spatial split: the wide (HBM-streaming) work gets every CU whose
    // index falls in the first `wide_cu_percent` of each 20-CU stride, the
    // group and lane work the rest, so both run at once on their own CUs.
    int cus = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
    std::vector<uint32_t> wm(size_t((cus + 31) / 32), 0), gm(wm.size(), 0);
    for (int c = 0; c < cus; ++c) {
      const bool wide = (c % 20) < (b->wide_cu_percent * 20 + 50) / 100;
      (wide ? wm : gm)[size_t(c / 32)] |= 1u << (c % 32);
      b->group_cus += wide ? 0 : 1;
    }
    if ((e = hipExtStreamCreateWithCUMask(&b->side, uint32_t(wm.size()), wm.data())) != hipSuccess)
      return bail_hip(e, "CU-masked streams");
    if ((e = hipExtStreamCreateWithCUMask(&b->main, uint32_t(gm.size()), gm.data())) != hipSuccess)
      return bail_hip(e, "CU-masked streams");
    b->split = true;
  } else {
    if ((e = hipStreamCreateWithFlags(&b->main, hipStreamNonBlocking)) != hipSuccess) return bail_hip(e, "streams");
    if ((e = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking)) != hipSuccess) return bail_hip(e, "streams");
  }
  if ((e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming)) != hipSuccess) return bail_hip(e, "events");
  if ((e = hipEventCreateWithFlags(&b->join, hipEventDisableTiming)) != hipSuccess) return bail_hip(e, "events");
  if ((e = hipEventCreateWithFlags(&b->join_main, hipEventDisableTiming)) != hipSuccess) return bail_hip(e, "events");
  if (!b->buckets.empty() && (e = b->alloc(&b->d_counters, b->buckets.size())) != hipSuccess)
    return bail_hip(e, "allocating counters");

  // Group buckets: candidates, tables, tuning.
  for (auto& k : b->buckets) {
    for (int tps : {8, 16, 32, 64, 128, 256, 512})
      if (k.max_local_k <= group_max_local_k(tps) &&
          64 + size_t(group_shots_per_block(tps)) * bucket_region(*b, k, tps) <= size_t(lds_limit))
        k.candidates.push_back(tps);
    if (k.candidates.empty()) {
      delete b;
      return fail(TICIT_HIP_INTERNAL, "a group-tier program no longer fits the workgroup");
    }
    std::vector<DevProgram> ps;
    std::vector<DevLaunch> ls;
    std::vector<GroupLayout> ys;
    for (uint32_t i : k.items) {
      ps.push_back(b->progs[i]->dev);
      ls.push_back(b->launches[i]);
      ys.push_back(b->progs[i]->group);
    }
    if ((e = b->upload(&k.d_progs, ps)) != hipSuccess) return bail_hip(e, "uploading batch tables");
    if ((e = b->upload(&k.d_launches, ls)) != hipSuccess) return bail_hip(e, "uploading batch tables");
    if ((e = b->upload(&k.d_layouts, ys)) != hipSuccess) return bail_hip(e, "uploading batch tables");
    k.tps = k.candidates.back();
    if (opt.autotune && (e = autotune(*b, k, opt.autotune_shots ? opt.autotune_shots : 8192)) != hipSuccess)
      return bail_hip(e, "autotuning");
    std::vector<uint64_t> shots, off;
    for (uint32_t i : k.items) shots.push_back(b->launches[i].shots);
    unit_table(*b, k, k.tps, shots, off, k.total_units);
    if ((e = b->upload(&k.d_unit_off, off)) != hipSuccess) return bail_hip(e, "uploading batch tables");
  }

  // Lane items with tables in device memory share one set of regions,
  // sized for the largest before capture.
  {
    size_t need = 0;
    for (uint32_t i : b->lane_items) {
      uint64_t slots = 0;
      size_t slot = 0;
      if ((e = lane_region_plan(b->progs[i], b->launches[i], &slots, &slot)) != hipSuccess)
        return bail_hip(e, "planning lane regions");
      need = std::max(need, size_t(slots) * slot);
      b->lane_ws.max_slots = std::max(b->lane_ws.max_slots, slots);
    }
    if (need) {
      if ((e = b->alloc(&b->lane_ws.base, need)) != hipSuccess) return bail_hip(e, "allocating lane regions");
      b->lane_ws.bytes = need;
    }
  }

  // Wide items: two sequences balanced by their estimated HBM bytes
  // (largest first, each to the lighter sequence), each with a workspace
  // sized before capture. Where the second workspace cannot be allocated
  // every item runs in the first sequence.
  auto prepare = [&](WideWorkspace* w, const std::vector<uint32_t>& items) {
    for (uint32_t i : items) {
      ticit_hip_program* h = b->progs[i];
      uint32_t batch = uint32_t(std::min<uint64_t>(h->wide_batch, b->launches[i].shots));
      hipError_t r = wide_workspace_prepare(w, batch, h->tab.state_dim, h->opt.precision == 1, h->wide);
      if (r != hipSuccess) return r;
    }
    return hipSuccess;
  };
  if (!b->split && b->wide_items.size() >= 2) {
    std::vector<std::pair<double, uint32_t>> cost;
    for (uint32_t i : b->wide_items) {
      const ticit_hip_program* h = b->progs[i];
      cost.push_back({h->tab.amp_traffic * (h->opt.precision == 1 ? 8.0 : 16.0) * double(b->launches[i].shots), i});
    }
    std::stable_sort(cost.begin(), cost.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
    const size_t nseq = std::min<size_t>(kWideSequences, cost.size());
    std::vector<std::vector<uint32_t>> seq(nseq);
    std::vector<double> load(nseq, 0.0);
    for (const auto& [c, i] : cost) {
      const size_t t = size_t(std::min_element(load.begin(), load.end()) - load.begin());
      seq[t].push_back(i);
      load[t] += c;
    }
    for (auto& q : seq) std::sort(q.begin(), q.end());  // program order within a sequence
    b->wide_items = seq[0];
    for (size_t t = 1; t < nseq; ++t) {
      ticit_hip_batch::WideSeq q;
      q.ws = wide_workspace_new();
      if (prepare(q.ws, seq[t]) == hipSuccess && hipStreamCreateWithFlags(&q.stream, hipStreamNonBlocking) == hipSuccess &&
          hipEventCreateWithFlags(&q.join, hipEventDisableTiming) == hipSuccess) {
        q.items = seq[t];
        b->more.push_back(q);
      } else {
        (void)hipGetLastError();  // the failed allocation's error, not a sticky one
        if (q.stream) (void)hipStreamDestroy(q.stream);
        wide_workspace_free(q.ws);
        b->wide_items.insert(b->wide_items.end(), seq[t].begin(), seq[t].end());
        std::sort(b->wide_items.begin(), b->wide_items.end());
      }
    }
  }
  if (!b->wide_items.empty()) {
    b->ws = wide_workspace_new();
    if ((e = prepare(b->ws, b->wide_items)) != hipSuccess) return bail_hip(e, "allocating the wide workspace");
  }

  // Capture: the lane and group work into one graph, the wide work into
  // another. A launch replays them on two streams (two hardware queues), so
  // the HBM-streaming wide kernels run beside the LDS-bound group kernels.
  hipError_t ce = hipSuccess;
  auto step = [&](hipError_t r) {
    if (ce == hipSuccess) ce = r;
  };
  if ((e = hipStreamBeginCapture(b->main, hipStreamCaptureModeRelaxed)) != hipSuccess)
    return bail_hip(e, "beginning graph capture");
  for (size_t j = 0; j < b->buckets.size(); ++j) {
    auto& k = b->buckets[j];
    step(run_bucket(*b, k, k.tps, k.d_launches, k.d_unit_off, k.total_units, b->d_counters + j, b->main));
  }
  for (uint32_t i : b->lane_items) {
    ticit_hip_program* h = b->progs[i];
    step(launch_lane_tier(h->tab.max_k, h->dev, b->launches[i], h->lane, b->main, &b->lane_ws));
  }
  auto compact_item = [&](uint32_t i, hipStream_t st) {
    auto it = b->post.find(i);
    if (it == b->post.end()) return;
    const auto& pi = it->second;
    step(launch_compaction(pi.keep, b->launches[i].shots, pi.tile_off, pi.accepted, pi.sets, pi.nsets, st));
  };
  for (const auto& k : b->buckets)
    for (uint32_t i : k.items) compact_item(i, b->main);
  for (uint32_t i : b->lane_items) compact_item(i, b->main);
  e = hipStreamEndCapture(b->main, &b->graph);
  if (ce != hipSuccess) return bail_hip(ce, "capturing the batch");
  if (e != hipSuccess) return bail_hip(e, "ending graph capture");
  if ((e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0)) != hipSuccess)
    return bail_hip(e, "instantiating the batch graph");
  if (!b->wide_items.empty()) {
    if ((e = hipStreamBeginCapture(b->side, hipStreamCaptureModeRelaxed)) != hipSuccess)
      return bail_hip(e, "beginning graph capture");
    for (uint32_t i : b->wide_items) {
      ticit_hip_program* h = b->progs[i];
      uint32_t batch = uint32_t(std::min<uint64_t>(h->wide_batch, b->launches[i].shots));
      step(launch_wide_tier(h->dev, h->tab.instrs.data(), b->launches[i], h->wide, batch, h->opt.precision == 1,
                            b->ws, b->side, nullptr, nullptr, nullptr, false));
      compact_item(i, b->side);
    }
    e = hipStreamEndCapture(b->side, &b->wide_graph);
    if (ce != hipSuccess) return bail_hip(ce, "capturing the batch");
    if (e != hipSuccess) return bail_hip(e, "ending graph capture");
    if ((e = hipGraphInstantiate(&b->wide_exec, b->wide_graph, nullptr, nullptr, 0)) != hipSuccess)
      return bail_hip(e, "instantiating the batch graph");
  }
  for (auto& q : b->more) {
    if ((e = hipStreamBeginCapture(q.stream, hipStreamCaptureModeRelaxed)) != hipSuccess)
      return bail_hip(e, "beginning graph capture");
    for (uint32_t i : q.items) {
      ticit_hip_program* h = b->progs[i];
      uint32_t batch = uint32_t(std::min<uint64_t>(h->wide_batch, b->launches[i].shots));
      step(launch_wide_tier(h->dev, h->tab.instrs.data(), b->launches[i], h->wide, batch, h->opt.precision == 1,
                            q.ws, q.stream, nullptr, nullptr, nullptr, false));
      compact_item(i, q.stream);
    }
    e = hipStreamEndCapture(q.stream, &q.graph);
    if (ce != hipSuccess) return bail_hip(ce, "capturing the batch");
    if (e != hipSuccess) return bail_hip(e, "ending graph capture");
    if ((e = hipGraphInstantiate(&q.exec, q.graph, nullptr, nullptr, 0)) != hipSuccess)
      return bail_hip(e, "instantiating the batch graph");
  }
  *out = b;
  return TICIT_HIP_OK;
}

ticit_hip_status ticit_hip_batch_launch(ticit_hip_batch* b, uint64_t seed_offset, void* stream) {
  if (!b) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  // clears the error word and sets this replay's seed offset, in stream order
  hipError_t e = launch_batch_prologue(b->d_error, b->d_seed_offset, seed_offset, s);
  if (e != hipSuccess) return from_hip(e, "launching the batch prologue");
  if (b->wide_exec) {
    // fork: the side stream starts after the prologue and everything the
    // caller enqueued before; join: the caller's stream waits for it
    if ((e = hipEventRecord(b->fork, s)) != hipSuccess) return from_hip(e, "forking the batch");
    if ((e = hipStreamWaitEvent(b->side, b->fork, 0)) != hipSuccess) return from_hip(e, "forking the batch");
    if ((e = hipGraphLaunch(b->wide_exec, b->side)) != hipSuccess) return from_hip(e, "launching the batch graph");
    if ((e = hipEventRecord(b->join, b->side)) != hipSuccess) return from_hip(e, "joining the batch");
  }
  for (auto& q : b->more) {  // the first sequence exists whenever these do, and recorded `fork`
    if ((e = hipStreamWaitEvent(q.stream, b->fork, 0)) != hipSuccess) return from_hip(e, "forking the batch");
    if ((e = hipGraphLaunch(q.exec, q.stream)) != hipSuccess) return from_hip(e, "launching the batch graph");
    if ((e = hipEventRecord(q.join, q.stream)) != hipSuccess) return from_hip(e, "joining the batch");
  }
  if (b->split) {
    // the group and lane graph runs on its CU-masked stream beside the wide one
    if ((e = hipStreamWaitEvent(b->main, b->fork, 0)) != hipSuccess) return from_hip(e, "forking the batch");
    if ((e = hipGraphLaunch(b->exec, b->main)) != hipSuccess) return from_hip(e, "launching the batch graph");
    if ((e = hipEventRecord(b->join_main, b->main)) != hipSuccess) return from_hip(e, "joining the batch");
    if ((e = hipStreamWaitEvent(s, b->join_main, 0)) != hipSuccess) return from_hip(e, "joining the batch");
  } else if ((e = hipGraphLaunch(b->exec, s)) != hipSuccess) {
    return from_hip(e, "launching the batch graph");
  }
  if (b->wide_exec && (e = hipStreamWaitEvent(s, b->join, 0)) != hipSuccess) return from_hip(e, "joining the batch");
  for (auto& q : b->more)
    if ((e = hipStreamWaitEvent(s, q.join, 0)) != hipSuccess) return from_hip(e, "joining the batch");
  b->last = s;
  return TICIT_HIP_OK;
}

ticit_hip_status ticit_hip_batch_wait(ticit_hip_batch* b) {
  if (!b) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  hipError_t e = hipStreamSynchronize(b->last);
  if (e != hipSuccess) return from_hip(e, "running the batch");
  int err = 0;
  if ((e = hipMemcpy(&err, b->d_error, sizeof(int), hipMemcpyDeviceToHost)) != hipSuccess)
    return from_hip(e, "reading the batch error flag");
  if (err & 2) return fail(TICIT_HIP_INVALID_INPUT, "probability must be between 0 and 1");
  if (err & 4)
    return fail(TICIT_HIP_INVALID_INPUT, "pinned measurement branch is not one half; pinning it would bias sampling");
  if (err & 1) return fail(TICIT_HIP_INTERNAL, "sampled an impossible active measurement branch");
  return TICIT_HIP_OK;
}

ticit_hip_status ticit_hip_batch_info_get(const ticit_hip_batch* b, ticit_hip_batch_info* out) {
  if (!b || !out) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  *out = ticit_hip_batch_info{};
  out->items = b->progs.size();
  out->lane_items = b->lane_items.size();
  out->wide_items = b->wide_items.size();
  for (const auto& q : b->more) out->wide_items += q.items.size();
  out->group_buckets = b->buckets.size();
  size_t n = std::min<size_t>(b->buckets.size(), sizeof(out->bucket_tps) / sizeof(out->bucket_tps[0]));
  for (size_t j = 0; j < n; ++j) {
    const auto& k = b->buckets[j];
    out->bucket_tps[j] = uint32_t(k.tps);
    out->bucket_items[j] = uint32_t(k.items.size());
    out->bucket_region_bytes[j] = bucket_region(*b, k, k.tps);
    out->bucket_grid[j] = uint32_t(k.grid);
    for (int c = 0; c < 8; ++c) out->bucket_tuned_ms[j][c] = k.tuned_ms[c];
  }
  return TICIT_HIP_OK;
}

void ticit_hip_batch_destroy(ticit_hip_batch* b) { delete b; }

ticit_hip_status ticit_hip_sample_batch(ticit_hip_program* const* programs, const uint64_t* shots,
                                        const uint64_t* seeds, size_t count, int bit_packed, int keep_records,
                                        ticit_hip_result* results) {
  if (!programs || !shots || !seeds || !results) return fail(TICIT_HIP_INVALID_INPUT, "null argument");
  auto t0 = std::chrono::steady_clock::now();
  std::vector<void*> bufs;
  auto release = [&]() {
    for (void* p : bufs) (void)hipFree(p);
  };
  auto dev = [&](size_t bytes, void** p) {
    hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 8));
    if (e == hipSuccess) bufs.push_back(*p);
    return e;
  };
  std::vector<ticit_hip_batch_item> items(count);
  for (size_t i = 0; i < count; ++i) {
    ticit_hip_program* h = programs[i];
    if (!h) {
      release();
      return fail(TICIT_HIP_INVALID_INPUT, "batch item " + std::to_string(i) + " has no program");
    }
    const HostTables& t = h->tab;
    items[i].program = h;
    items[i].shots = shots[i];
    items[i].seed = seeds[i];
    void* p = nullptr;
    hipError_t e = hipSuccess;
    if (keep_records) {
      if ((e = dev(shots[i] * h->row_bytes(uint64_t(t.nrec), bit_packed), &p)) == hipSuccess)
        items[i].measurements = static_cast<uint8_t*>(p);
      if (e == hipSuccess && (e = dev(shots[i] * h->row_bytes(uint64_t(t.ndet), bit_packed), &p)) == hipSuccess)
        items[i].detectors = static_cast<uint8_t*>(p);
      if (e == hipSuccess && (e = dev(shots[i] * h->row_bytes(uint64_t(t.nobs), bit_packed), &p)) == hipSuccess)
        items[i].observables = static_cast<uint8_t*>(p);
      if (e == hipSuccess && t.nexp > 0 && (e = dev(shots[i] * size_t(t.nexp) * 8, &p)) == hipSuccess)
        items[i].exp_vals = static_cast<double*>(p);
    }
    if (e == hipSuccess && (e = dev(size_t(std::max(1, t.nobs)) * 8, &p)) == hipSuccess) {
      items[i].observable_ones = static_cast<uint64_t*>(p);
      e = hipMemset(p, 0, size_t(std::max(1, t.nobs)) * 8);
    }
    if (e == hipSuccess && t.npost > 0 && (e = dev(8, &p)) == hipSuccess) items[i].accepted = static_cast<uint64_t*>(p);
    if (e != hipSuccess) {
      release();
      return from_hip(e, "allocating batch outputs");
    }
  }
  ticit_hip_batch_options opt{};
  opt.bit_packed = bit_packed;
  opt.keep_records = keep_records;
  opt.autotune = 1;
  ticit_hip_batch* b = nullptr;
  ticit_hip_status st = ticit_hip_batch_create(items.data(), count, &opt, &b);
  if (st != TICIT_HIP_OK) {
    release();
    return st;
  }
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  (void)hipEventCreate(&ev0);
  (void)hipEventCreate(&ev1);
  (void)hipEventRecord(ev0, b->main);
  st = ticit_hip_batch_launch(b, 0, b->main);
  (void)hipEventRecord(ev1, b->main);
  if (st == TICIT_HIP_OK) st = ticit_hip_batch_wait(b);
  float ms = 0.f;
  if (st == TICIT_HIP_OK) (void)hipEventElapsedTime(&ms, ev0, ev1);
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  ticit_hip_batch_destroy(b);
  for (size_t i = 0; i < count && st == TICIT_HIP_OK; ++i) {
    ticit_hip_program* h = programs[i];
    const HostTables& t = h->tab;
    ticit_hip_result& r = results[i];
    r = ticit_hip_result{};
    r.shots = r.accepted = shots[i];
    r.bit_packed = bit_packed != 0;
    r.nmeasurements = uint64_t(t.nrec);
    r.ndetectors = uint64_t(t.ndet);
    r.nobservables = uint64_t(t.nobs);
    r.nexpvals = uint64_t(t.nexp);
    r.observable_ones = static_cast<uint64_t*>(calloc(std::max(1, t.nobs), 8));
    hipError_t e = hipMemcpy(r.observable_ones, items[i].observable_ones, size_t(t.nobs) * 8, hipMemcpyDeviceToHost);
    if (t.nobs > 0 && h->opt.observable < uint32_t(t.nobs)) r.logical_errors = r.observable_ones[h->opt.observable];
    uint64_t rows = shots[i];
    if (e == hipSuccess && items[i].accepted && shots[i]) {
      e = hipMemcpy(&rows, items[i].accepted, 8, hipMemcpyDeviceToHost);
      r.accepted = rows;
      r.discarded = shots[i] - rows;
    }
    if (keep_records) {
      r.record_rows = rows;
      size_t mb = rows * h->row_bytes(uint64_t(t.nrec), bit_packed);
      size_t db = rows * h->row_bytes(uint64_t(t.ndet), bit_packed);
      size_t ob = rows * h->row_bytes(uint64_t(t.nobs), bit_packed);
      r.measurements = static_cast<uint8_t*>(malloc(std::max<size_t>(1, mb)));
      r.detectors = static_cast<uint8_t*>(malloc(std::max<size_t>(1, db)));
      r.observables = static_cast<uint8_t*>(malloc(std::max<size_t>(1, ob)));
      if (e == hipSuccess && mb) e = hipMemcpy(r.measurements, items[i].measurements, mb, hipMemcpyDeviceToHost);
      if (e == hipSuccess && db) e = hipMemcpy(r.detectors, items[i].detectors, db, hipMemcpyDeviceToHost);
      if (e == hipSuccess && ob) e = hipMemcpy(r.observables, items[i].observables, ob, hipMemcpyDeviceToHost);
      if (t.nexp > 0) {
        size_t eb = rows * size_t(t.nexp) * 8;
        r.exp_vals = static_cast<double*>(malloc(std::max<size_t>(8, eb)));
        if (e == hipSuccess && eb) e = hipMemcpy(r.exp_vals, items[i].exp_vals, eb, hipMemcpyDeviceToHost);
      }
    }
    r.kernel_s = double(ms) * 1e-3;
    r.sample_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (e != hipSuccess) st = from_hip(e, "reading batch results");
  }
  release();
  return st;
}

}  // extern "C"
