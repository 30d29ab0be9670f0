// Lowering of a planned Program into the flat tables the HIP kernels read.
#pragma once

#include <string>
#include <vector>

#include "../kernels/device_abi.hpp"
#include "program.hpp"

namespace ticit {

struct HostTables {
  std::vector<DevInstr> instrs;
  std::vector<uint32_t> resid, det_recs, obs_off, obs_recs, sym_off, sym_part;
  std::vector<int32_t> exo_syms;
  std::vector<DevCat> cats;
  std::vector<DevRare> rares;
  std::vector<DevBern> berns;
  std::vector<DevLowp> lowps;
  std::vector<double> f64pool;
  std::vector<uint64_t> u64pool;
  std::vector<int32_t> i32pool;
  int n = 0, initial_k = 0, max_k = 0, nrec = 0, ndet = 0, nobs = 0, nslots = 0;
  int npartials = 0, nsym = 0, nmeasure = 0, ncoin = 0;
  // planning features used by the tier planner
  size_t expression_terms = 0;
  size_t quantum_ops = 0;
  double dense_work = 0.0;  // sum over amplitude instructions of 2^k
};

// Validates the program (every residual symbol assigned before use, active
// widths consistent with the opcode sequence, ids in range) and builds the
// tables. Throws TicitError on failure.
HostTables lower_to_device(const Program& p);

}  // namespace ticit
