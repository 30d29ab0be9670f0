// Pinned measurement parities (ticit's plan_pinned_measurements,
// src/sampler/pinning.rs:227-289): each constraint fixes the XOR of a set of
// measurement records in the noiseless circuit. Records expand into XORs of
// measurement-branch symbols (exogenous noise symbols are zero in the
// noiseless circuit); the rows are reduced by Gaussian elimination over
// GF(2), each taking as its pivot the branch drawn last, so when that
// branch's instruction runs every other term of its row already holds a
// value and the branch is a plain XOR of assigned branch symbols instead of
// a draw.
#pragma once

#include <cstdint>
#include <vector>

#include "program.hpp"

namespace ticit {

struct MeasurementParity {
  std::vector<uint64_t> records;  // zero-based record indices
  bool value = false;
};

struct ForcedBranch {
  int32_t instruction = 0;  // index into Program::instrs (a Measure or Coin)
  SymBool plan;             // constant XOR branch symbols assigned earlier
};

// Throws TicitError (invalid input) for a record the program never writes
// or a constraint the noiseless circuit already decides the other way.
std::vector<ForcedBranch> plan_pinned_measurements(const Program& p, const std::vector<MeasurementParity>& pins);

}  // namespace ticit
