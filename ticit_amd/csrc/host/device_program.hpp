// Lowering of a planned Program into the flat tables the HIP kernels read.
#pragma once

#include <string>
#include <vector>

#include "../kernels/device_abi.hpp"
#include "program.hpp"

namespace ticit {

struct HostTables {
  std::vector<DevInstr> instrs;
  std::vector<DevStep> steps;  // compact view, one per instr
  // Lane-tier schedule: the steps in order, except that records wait until
  // an in-order step reads a slot one of them assigns (then they run as
  // dependency levels, 64 side by side) and detectors run last, together.
  // A dormant coin splits into its draw, merged with the draws of the
  // coins next to it into one step that keeps their program order (and
  // runs before any later draw), and its outcome record, deferred like the
  // others. Every runtime slot is assigned once and never reused, so a
  // deferred step reads the values it would have read in order.
  std::vector<DevStep> lsteps, bsteps;
  int32_t nlsteps = 0;  // lsteps without the padding entry
  int32_t npost = 0;    // post_dets without the padding entry
  std::vector<uint32_t> resid, det_recs, obs_off, obs_recs, sym_off, sym_part;
  std::vector<uint32_t> tl_off, tl_part;  // noise-event toggle lists
  std::vector<uint32_t> post_dets;  // 0-based postselected detectors
  std::vector<uint8_t> obs_flip;    // [max(1, nobs)] inverted observables
  std::vector<int32_t> exo_syms;
  std::vector<DevCat> cats;
  std::vector<DevRare> rares;
  std::vector<DevBern> berns;
  std::vector<DevLowp> lowps;
  std::vector<double> f64pool;
  std::vector<uint64_t> u64pool;
  std::vector<int32_t> i32pool;
  int n = 0, initial_k = 0, max_k = 0, nrec = 0, ndet = 0, nobs = 0, nslots = 0;
  bool dets_once = false;  // each detector id has exactly one Detector instruction
  int npartials = 0, nsym = 0, nmeasure = 0, ncoin = 0, nexp = 0;
  // planning features used by the tier planner
  size_t expression_terms = 0;
  size_t quantum_ops = 0;
  double dense_work = 0.0;  // sum over amplitude instructions of 2^k
  // amplitudes moved per shot when the state streams from memory: a rotation
  // reads and writes 2^k, a promotion reads 2^k and writes 2^(k+1), a
  // nondiagonal measurement reads 2^k twice (probability, projection) and
  // writes 2^(k-1); a diagonal one reads only one half twice and writes it
  // (3 * 2^(k-1)); a merge reads both factors and writes their product
  double amp_traffic = 0.0;
  // execution layout: one component spanning the state (dense) or the
  // planner's tensor components
  bool components = false;
  uint32_t state_dim = 0;
  std::vector<uint32_t> comp_off;  // per component
  int32_t max_local_k = 0;
  int32_t ninstr = 0;  // device instructions (merges included, ignored phases dropped)
};

// Validates the program (every residual symbol assigned before use, active
// widths consistent with the opcode sequence, ids in range) and builds the
// tables. Throws TicitError on failure.
// With components = true the instruction stream is lowered through the
// tensor-component plan (throws if the program has none).
// forced: pinned branches (pinning.hpp), by instruction index.
struct ForcedBranch;

// Expected detector and observable bits (SamplerOptions::expected_detectors
// and expected_observables, src/sampler/prepared.rs:111-116): a nonzero
// entry inverts that column in every row, in postselection and in the
// observable counts (src/sampler/prepared.rs:909-974,
// src/sampler/batch/postselect.rs:631-660). Empty: no inversion.
struct ExpectedBits {
  std::vector<uint8_t> detectors, observables;
};

HostTables lower_to_device(const Program& p, bool components = false,
                           const std::vector<ForcedBranch>* forced = nullptr,
                           const ExpectedBits* expected = nullptr);

}  // namespace ticit
