// The program handle behind the C ABI and the device entry points shared by
// the single-program and batched paths.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/ticit_hip.h"
#include "device_program.hpp"
#include "frontend.hpp"
#include "pinning.hpp"

namespace ticit {

size_t lane_tier_lds_bytes(int kmax, const LaneLayout& y);
size_t lane_region_bytes(int kmax, const LaneLayout& y);
LaneLayout lane_launch_layout(int kmax, const LaneLayout& Yp, const DevLaunch& L);
void lane_tier_occupancy(int kmax, const LaneLayout& Y, uint32_t row_bytes, uint32_t* lds, uint32_t* blocks);
size_t group_tier_lds_bytes(uint32_t state_dim, const GroupLayout& y, int tps);
size_t group_region_bytes(uint32_t state_dim, const GroupLayout& y, int tps);
int group_max_local_k(int tps);
int group_shots_per_block(int tps);
int group_shots_per_unit(int tps);
int group_batch_occupancy(int tps, uint32_t region_bytes, int max_local_k);
// `ws`: the device memory regions when Y.store is not kLaneLds.
hipError_t launch_lane_tier(int kmax, const DevProgram& P, const DevLaunch& L, const LaneLayout& Y,
                            hipStream_t stream, const LaneWorkspace* ws = nullptr);
hipError_t launch_group_tier(const DevProgram& P, const DevLaunch& L, const GroupLayout& Y, hipStream_t stream,
                             int tps);
hipError_t launch_group_batch(const DevProgram* progs, const DevLaunch* launches, const GroupLayout* layouts,
                              const uint64_t* unit_off, uint32_t nitems, uint64_t total_units,
                              unsigned long long* counter, uint32_t region_bytes, int tps, int max_local_k, int grid,
                              hipStream_t stream);
hipError_t launch_batch_prologue(int* error, uint64_t* seed_offset, uint64_t value, hipStream_t stream);
size_t wide_bytes_per_shot(uint32_t state_dim, int fp32, const WideLayout& y);
uint32_t wide_partials_per_shot(int max_local_k);
struct WideWorkspace;
WideWorkspace* wide_workspace_new();
void wide_workspace_free(WideWorkspace* ws);
hipError_t wide_workspace_prepare(WideWorkspace* ws, uint32_t batch, uint32_t state_dim, int fp32,
                                  const WideLayout& Y);
// `ordered` = order the run after the workspace's previous user with an
// event (off inside graph capture, where the graph orders the work).
hipError_t launch_wide_tier(const DevProgram& P, const DevInstr* host_instrs, const DevLaunch& L,
                            const WideLayout& Y, uint32_t batch, int fp32, WideWorkspace* ws, hipStream_t stream,
                            double* dump_re, double* dump_im, int* dump_k, bool ordered = true,
                            const WideSecond* second = nullptr);

constexpr uint32_t kTierLane = 1, kTierGroup = 2, kTierWide = 3;

}  // namespace ticit

struct ticit_hip_program;

namespace ticit {

// Lane-tier programs with tables in device memory: the regions a launch
// uses (count and bytes each; zero otherwise).
hipError_t lane_region_plan(ticit_hip_program* h, const DevLaunch& L, uint64_t* slots, size_t* slot_bytes);

// Learned layout decision (planner.cpp); false = keep the heuristic's.
bool learned_layout(const Program& p, int device, bool* use_components, float* p_components);

ticit_hip_status fail(ticit_hip_status s, const std::string& msg);
ticit_hip_status from_hip(hipError_t e, const char* what);

}  // namespace ticit

struct ticit_hip_program {
  ticit::Program prog;
  ticit::HostTables tab;
  ticit::DevProgram dev{};
  std::vector<void*> allocations;
  ticit_hip_options opt{};
  uint32_t tier = 0;
  int device = 0;
  int* d_error = nullptr;
  ticit::LaneLayout lane{};
  ticit::GroupLayout group{};
  ticit::WideLayout wide{};
  int group_tps = 256;  // group tier threads per shot of the plain launch
  bool bit_heavy = false;  // record/detector/coin steps dominate the program
  // lane tier with tables in device memory: the wave regions, grown to the
  // largest launch and reused in stream order through `lane_ws_done`
  ticit::LaneWorkspace lane_ws{};
  hipEvent_t lane_ws_done = nullptr;
  uint32_t wide_batch = 0;
  int32_t planner_source = 0;
  float planner_p = -1.f;
  uint32_t meas_row_packed = 0, det_row_packed = 0, obs_row_packed = 0;
  // expected bits folded into the lowered tables (empty: none) and the
  // pinned parities the program was built with
  ticit::ExpectedBits expected;
  std::vector<ticit::MeasurementParity> pins;

  // Staging of postselected device launches (full rows, keep flags, tile
  // offsets): grown to the largest call, reused in stream order through
  // `post_done`.
  void* post_buf = nullptr;
  size_t post_bytes = 0;
  hipEvent_t post_done = nullptr;
  bool post_pending = false;
  std::mutex post_mu;  // host threads sharing the program take turns on the staging

  // Device memory of the synchronous host-result calls (ticit_hip_sample,
  // the injected and reference runs): rows, counts, flags, injected inputs
  // and compaction scratch, grown to the largest call and reused, since a
  // call has finished with it before it returns. Rows come back through
  // pinned host staging, also kept.
  void* arena = nullptr;
  size_t arena_bytes = 0;
  void* pinned = nullptr;
  size_t pinned_bytes = 0;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  // chunked row return: copies run on their own stream behind per-chunk
  // kernel events, into two pinned slots used in turn
  hipStream_t copy_stream = nullptr;
  std::vector<hipEvent_t> chunk_done;
  hipEvent_t slot_done[2] = {nullptr, nullptr};
  std::mutex host_mu;  // host threads sharing the program take turns on these

  ~ticit_hip_program() {
    if (post_done) {
      (void)hipEventSynchronize(post_done);
      (void)hipEventDestroy(post_done);
    }
    if (post_buf) (void)hipFree(post_buf);
    if (lane_ws_done) {
      (void)hipEventSynchronize(lane_ws_done);
      (void)hipEventDestroy(lane_ws_done);
    }
    if (lane_ws.base) (void)hipFree(lane_ws.base);
    if (arena) (void)hipFree(arena);
    if (pinned) (void)hipHostFree(pinned);
    if (copy_stream) {
      (void)hipStreamSynchronize(copy_stream);
      (void)hipStreamDestroy(copy_stream);
    }
    if (ev_start) (void)hipEventDestroy(ev_start);
    if (ev_stop) (void)hipEventDestroy(ev_stop);
    for (hipEvent_t ev : chunk_done) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : slot_done)
      if (ev) (void)hipEventDestroy(ev);
    for (void* p : allocations) (void)hipFree(p);
  }

  template <class T>
  hipError_t upload(const std::vector<T>& v, const T*& out) {
    void* d = nullptr;
    size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
    hipError_t e = hipMalloc(&d, bytes);
    if (e != hipSuccess) return e;
    allocations.push_back(d);
    if (!v.empty()) {
      e = hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
      if (e != hipSuccess) return e;
    }
    out = static_cast<const T*>(d);
    return hipSuccess;
  }

  uint32_t row_bytes(uint64_t cols, bool packed) const { return uint32_t(packed ? (cols + 7) / 8 : cols); }
};
