// Planned-program data model shared by the compiler front end, the device
// lowering and the plan-blob codec.
//
// A program is the six-opcode instruction stream of ticit's
// `FactoredInstructionProgram` (src/factored.rs:284-292, 866-886) plus the
// exogenous sampling plan that names every noise symbol
// (src/planner.rs:1441-1480) and the observable record lists
// (src/sampler/prepared.rs:241-266).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace ticit {

// constant XOR s_{c1} XOR ... ; conditions are 1-based, ascending, unique
// (src/symbolic.rs:24-48).
struct SymBool {
  bool constant = false;
  std::vector<int32_t> conds;
  int32_t max_condition() const { return conds.empty() ? 0 : conds.back(); }
  bool operator==(const SymBool& o) const { return constant == o.constant && conds == o.conds; }
  bool operator!=(const SymBool& o) const { return !(*this == o); }
};

SymBool sym_xor(const SymBool& a, const SymBool& b);
SymBool sym_xor_const(const SymBool& a, bool v);
SymBool sym_var(int32_t c);
void normalize_conditions(std::vector<int32_t>& c);

enum class Op : uint8_t {
  Rotate = 0,    // ApplyPrecomputedActivePauliRotation
  Promote = 1,   // PromoteDormantRotation
  Record = 2,    // RecordMeasurement
  Detector = 3,  // RecordDetector
  Measure = 4,   // MeasurePrecomputedActivePauli
  Coin = 5,      // IntroduceDormantMeasurementBranch
};

// One planned instruction. Rotation fields follow
// PrecomputedActivePauliRotationKernel (src/active.rs:162-230), measurement
// fields PrecomputedActivePauliMeasurementKernel (src/active.rs:242-377).
struct Instr {
  Op op = Op::Record;
  int32_t nq = 0;          // active width the kernel acts on (k)
  uint64_t xmask = 0;      // active X bits
  uint64_t zmask = 0;      // active Z bits
  int32_t phase = 0;       // coefficient i^phase of the active body
  double angle = 0.0;      // kernel angle phi of exp(-i phi P)
  int32_t pivot = 0;       // measurement pivot coordinate
  bool diagonal = false;   // measurement of a Z-only body
  int32_t diag_phase_bit = 0;
  int32_t branch = 0;      // minted branch symbol (Measure/Coin), 0 = none
  int32_t record = 0;      // 1-based record slot, 0 = none
  int32_t record_condition = 0;  // symbol assigned the outcome, 0 = none
  int32_t detector = 0;    // 1-based detector id (Detector)
  int32_t exp_val = -1;    // expectation slot, -1 = none
  bool postselect = false; // Detector: a firing discards the shot (src/factored.rs:227)
  std::vector<int32_t> det_records;  // records XORed by a detector
  SymBool expr;            // sign (Rotate/Promote) or outcome (others)
};

struct CategoricalDist {
  int32_t nbits = 0;
  std::vector<int32_t> conds;
  std::vector<uint64_t> assignments;  // row -> packed bits over conds
  std::vector<double> probs;
};

struct RareGroup {
  double event_p = 0.0;
  int32_t nbits = 0;
  std::vector<std::vector<int32_t>> sets;  // one condition list per member
  std::vector<uint64_t> assignments;
  std::vector<double> probs;
  std::vector<int32_t> event_rows;
  std::vector<double> event_probs;
};

struct BernGroup {
  double p = 0.0;
  std::vector<int32_t> conds;
};

struct Program {
  int32_t n = 0;
  int32_t initial_k = 0;
  int32_t max_k = 0;
  int32_t nsymbols = 0;
  int32_t nrecords = 0;
  int32_t ndetectors = 0;
  int32_t nexpvals = 0;
  // FactoredInstructionProgram::use_active_components (src/factored.rs:885):
  // the planner selected the tensor-component execution of the active state
  bool use_components = false;
  // TICK-separated layers of the source circuit (0 = unknown): the depth
  // feature of the learned planner, carried in the blob's flags word
  int32_t circuit_layers = 0;
  std::vector<Instr> instrs;
  std::vector<CategoricalDist> categorical;
  std::vector<RareGroup> rare;
  std::vector<int32_t> bern_conds;
  std::vector<double> bern_probs;
  std::vector<BernGroup> lowp;
  // observables[i] = flattened record list whose XOR is observable i
  std::vector<std::vector<int32_t>> observables;
};

// Binary plan blob: a little-endian stream of u64 words, documented in
// include/ticit_plan.h.
std::vector<uint64_t> encode_program(const Program& p);
bool decode_program(const uint64_t* words, size_t nwords, Program& out, std::string& err);

}  // namespace ticit
