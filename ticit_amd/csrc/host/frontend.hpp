// Circuit front end: .ticit text -> planned Program.
//
// Mirrors ticit's Circuit::compile path (src/circuit/mod.rs:75-77,274-327):
// parse, lower into a Clifford frame plus symbolic Pauli frame
// (src/circuit/lowering.rs:786), optimize the pending queue
// (src/pending_optimizer.rs:217), plan it into six opcodes
// (src/planner.rs:1504), splice detectors (src/circuit/mod.rs:224-271) and
// freeze the exogenous sampling plan (src/planner.rs:1409-1487).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "program.hpp"

namespace ticit {

// Error kinds follow TicitError (src/errors.rs:9-57).
struct TicitError : std::runtime_error {
  int kind;  // 1 invalid input, 2 parse, 3 unsupported, 4 internal
  TicitError(int k, const std::string& m) : std::runtime_error(m), kind(k) {}
};

struct CompileStats {
  int ncircuit_instructions = 0;
  int npending = 0;
  int fused_rotations = 0;
  int cancelled_rotations = 0;
  int measurement_left_swaps = 0;
};

// Parses and plans `text`. `postselection_mask` is accepted for interface
// parity; detectors are never postselected by this backend yet.
// postselection_mask: flat detector flags, nonzero postselects the detector
// (SamplerOptions::postselection_mask, src/sampler/prepared.rs:108-109);
// missing entries are zero.
Program compile_circuit_text(const std::string& text, CompileStats* stats = nullptr,
                             const std::vector<uint8_t>& postselection_mask = {});

}  // namespace ticit
