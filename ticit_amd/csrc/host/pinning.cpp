// GF(2) solver for pinned measurement parities; see pinning.hpp.
#include "pinning.hpp"

#include <map>
#include <set>
#include <string>
#include <unordered_map>

#include "frontend.hpp"

namespace ticit {

namespace {

// How a symbol gets its value during execution (src/sampler/pinning.rs:64-75):
// a measurement branch drawn at an instruction (a free variable), or a record
// condition assigned from an instruction's outcome. Any other symbol is
// exogenous noise, zero in the noiseless circuit.
struct Symbols {
  const Program& p;
  std::unordered_map<int32_t, int32_t> branch_at;   // branch symbol -> instruction
  std::unordered_map<int32_t, int32_t> derived_at;  // record condition -> instruction
  std::unordered_map<int32_t, int32_t> record_at;   // 1-based record -> instruction
  std::unordered_map<int32_t, SymBool> expansion;   // symbol -> XOR of branch symbols

  explicit Symbols(const Program& prog) : p(prog) {
    for (size_t i = 0; i < p.instrs.size(); ++i) {
      const Instr& in = p.instrs[i];
      if (in.exp_val >= 0) continue;  // probes assign nothing
      if ((in.op == Op::Measure || in.op == Op::Coin) && in.branch > 0) branch_at[in.branch] = int32_t(i);
      if (in.record_condition > 0 && !branch_at.count(in.record_condition))
        derived_at.emplace(in.record_condition, int32_t(i));
      if (in.record > 0) record_at[in.record] = int32_t(i);
    }
  }

  // Expansion of one symbol into branch symbols, iteratively (records chain
  // through their outcome expressions).
  const SymBool& expand_symbol(int32_t symbol) {
    std::vector<int32_t> stack{symbol};
    std::set<int32_t> in_progress;
    while (!stack.empty()) {
      const int32_t top = stack.back();
      if (expansion.count(top)) {
        in_progress.erase(top);
        stack.pop_back();
        continue;
      }
      if (branch_at.count(top)) {
        expansion[top] = sym_var(top);
      } else {
        auto d = derived_at.find(top);
        if (d == derived_at.end()) {
          expansion[top] = SymBool{};  // exogenous: zero without noise
        } else {
          in_progress.insert(top);
          const SymBool& outcome = p.instrs[size_t(d->second)].expr;
          bool pending = false;
          for (int32_t c : outcome.conds) {
            if (expansion.count(c)) continue;
            if (in_progress.count(c)) throw TicitError(1, "measurement record expression is cyclic");
            stack.push_back(c);
            pending = true;
          }
          if (pending) continue;
          SymBool v;
          v.constant = outcome.constant;
          for (int32_t c : outcome.conds) v = sym_xor(v, expansion.at(c));
          expansion[top] = v;
        }
      }
      in_progress.erase(top);
      stack.pop_back();
    }
    return expansion.at(symbol);
  }

  SymBool record_expansion(uint64_t record) {
    auto it = record_at.find(int32_t(record + 1));
    if (record >= uint64_t(INT32_MAX) || it == record_at.end())
      throw TicitError(1, "pinned measurement record " + std::to_string(record) + " is not written by this circuit");
    const SymBool& outcome = p.instrs[size_t(it->second)].expr;
    SymBool v;
    v.constant = outcome.constant;
    for (int32_t c : outcome.conds) v = sym_xor(v, expand_symbol(c));
    return v;
  }

  // The branch drawn last among `conds`: (instruction, symbol), or (-1, 0).
  std::pair<int32_t, int32_t> last_drawn(const std::vector<int32_t>& conds) const {
    std::pair<int32_t, int32_t> best{-1, 0};
    for (int32_t c : conds) {
      auto it = branch_at.find(c);
      if (it != branch_at.end() && std::make_pair(it->second, c) > best) best = {it->second, c};
    }
    return best;
  }
};

}  // namespace

std::vector<ForcedBranch> plan_pinned_measurements(const Program& p, const std::vector<MeasurementParity>& pins) {
  std::vector<ForcedBranch> out;
  if (pins.empty()) return out;
  Symbols sym(p);
  std::map<int32_t, std::pair<int32_t, SymBool>> pivot_rows;  // instruction -> (pivot, row)
  for (const MeasurementParity& pin : pins) {
    SymBool row;
    row.constant = pin.value;
    for (uint64_t r : pin.records) row = sym_xor(row, sym.record_expansion(r));
    for (;;) {
      auto [instruction, pivot] = sym.last_drawn(row.conds);
      if (instruction < 0) {
        if (row.constant) {
          std::string recs;
          for (size_t i = 0; i < pin.records.size(); ++i) recs += (i ? ", " : "") + std::to_string(pin.records[i]);
          throw TicitError(1, "pinned measurement parity over [" + recs + "] is deterministic and cannot be " +
                                  std::string(pin.value ? "1" : "0"));
        }
        break;
      }
      auto it = pivot_rows.find(instruction);
      if (it != pivot_rows.end()) {
        row = sym_xor(row, it->second.second);
      } else {
        pivot_rows.emplace(instruction, std::make_pair(pivot, row));
        break;
      }
    }
  }
  for (const auto& [instruction, pr] : pivot_rows) {
    ForcedBranch f;
    f.instruction = instruction;
    f.plan = sym_xor(pr.second, sym_var(pr.first));  // the pivot takes the rest of its row
    out.push_back(std::move(f));
  }
  return out;
}

}  // namespace ticit
