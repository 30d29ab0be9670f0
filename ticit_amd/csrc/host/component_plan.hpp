// Tensor-component planning of the active state (ticit's
// build_active_component_plan, src/component_plan.rs:277-522).
//
// The planner replays the instruction stream tracking only which active
// coordinates have been linked. Each connected group is a component with its
// own small amplitude vector; an instruction whose support spans several
// components first merges them (Kronecker product into the largest one).
// Every step is static — identical for all shots — so the whole plan lowers
// to fixed offsets and widths.
#pragma once

#include <cstdint>
#include <vector>

#include "program.hpp"

namespace ticit {

enum class StepKind : uint8_t { None, IgnoredPhase, Rotation, Promotion, Measurement };

struct ComponentStep {
  StepKind kind = StepKind::None;
  int32_t comp = -1;             // component the operation runs on
  std::vector<int32_t> sources;  // components folded into `comp` first, in order
  Instr local;                   // the operation in the component's bit order
  bool deactivate = false;       // measurement consumed the last coordinate
};

struct ComponentPlan {
  bool planned = false;   // max_k >= 8 and no expectation probes
  bool selected = false;  // the cost model's four gates passed
  int32_t initial_components = 0;
  std::vector<int32_t> comp_max_k;  // capacity (log2) of every component
  std::vector<ComponentStep> steps;  // one per instruction when planned
  double dense_work = 0.0;
  double component_work = 0.0;
  uint64_t dense_peak = 0;
  uint64_t peak_live = 0;
  uint64_t allocated = 0;
  int32_t quantum_ops = 0;
  int32_t merges = 0;
  int32_t max_component_k = 0;
};

// Throws TicitError (kind 4) when the stream is inconsistent.
ComponentPlan plan_components(const Program& p);

}  // namespace ticit
