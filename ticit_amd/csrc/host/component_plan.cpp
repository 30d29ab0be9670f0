// Tensor-component planner; see component_plan.hpp.
//
// Cost model and gates follow src/component_plan.rs:17-42, 282-298: dense
// work is 2^k per rotation and 2 * 2^k per promotion or measurement;
// component work is 2 * 2^k_merged per merge, 2^k_c for the operation and a
// dispatch charge of 32 per operation (2 for a promotion's fresh pair).
#include "component_plan.hpp"

#include <algorithm>
#include <cmath>
#include <string>

#include "frontend.hpp"

namespace ticit {
namespace {

constexpr double kDispatch = 32.0;
constexpr double kMinSaved = 8192.0;
constexpr double kRatio = 1.8;
constexpr double kAllocRatio = 1.25;

class Planner {
 public:
  explicit Planner(const Program& p) : prog_(p) {}

  ComponentPlan run() {
    ComponentPlan& out = out_;
    out.initial_components = prog_.initial_k;
    out.dense_peak = uint64_t(1) << prog_.max_k;
    if (prog_.max_k < 8 || prog_.nexpvals != 0) return out;
    out.planned = true;
    out.steps.resize(prog_.instrs.size());
    for (int32_t c = 0; c < prog_.initial_k; ++c) add_coordinate();
    note_live();
    for (size_t i = 0; i < prog_.instrs.size(); ++i) step(prog_.instrs[i], out.steps[i]);
    for (int32_t mk : out.comp_max_k) {
      out.allocated += uint64_t(1) << mk;
      out.max_component_k = std::max(out.max_component_k, mk);
    }
    bool sel = prog_.max_k >= 8 && out.quantum_ops >= 4;
    sel = sel && out.dense_work >= kRatio * out.component_work;
    sel = sel && out.dense_work - out.component_work >= kMinSaved;
    sel = sel && double(out.allocated) <= kAllocRatio * double(out.dense_peak);
    out.selected = sel;  // NaN work fails every >= above
    return out;
  }

 private:
  struct Group {
    bool live = true;
    std::vector<int32_t> coords;  // local bit j = coords[j]
  };

  const Program& prog_;
  ComponentPlan out_;
  std::vector<Group> groups_;
  std::vector<int32_t> order_;  // active position -> coordinate
  std::vector<int32_t> owner_;  // coordinate -> group, -1 once measured away

  [[noreturn]] static void bad(const std::string& m) { throw TicitError(4, m); }

  int32_t add_coordinate() {
    int32_t coord = int32_t(owner_.size());
    groups_.push_back(Group{true, {coord}});
    owner_.push_back(int32_t(groups_.size()) - 1);
    order_.push_back(coord);
    out_.comp_max_k.push_back(1);
    return int32_t(groups_.size()) - 1;
  }

  void note_live() {
    uint64_t live = 0;
    for (const Group& g : groups_)
      if (g.live && !g.coords.empty()) live += uint64_t(1) << g.coords.size();
    out_.peak_live = std::max(out_.peak_live, live);
  }

  int32_t position_of(int32_t coord) const {
    auto it = std::find(order_.begin(), order_.end(), coord);
    return it == order_.end() ? -1 : int32_t(it - order_.begin());
  }

  // groups an action's support reaches, first-touched order; an identity
  // action is pinned to the first live coordinate's group
  std::vector<int32_t> reached(const Instr& in) const {
    std::vector<int32_t> r;
    const uint64_t support = in.xmask | in.zmask;
    for (size_t pos = 0; pos < order_.size(); ++pos) {
      if (!((support >> pos) & 1)) continue;
      int32_t g = owner_[size_t(order_[pos])];
      if (g < 0) bad("component planner found an active coordinate without a component");
      if (std::find(r.begin(), r.end(), g) == r.end()) r.push_back(g);
    }
    if (r.empty() && !order_.empty()) r.push_back(owner_[size_t(order_[0])]);
    return r;
  }

  // largest group wins; ties go to the lowest id
  int32_t pick_target(const std::vector<int32_t>& r) const {
    int32_t best = r.front();
    for (int32_t g : r) {
      size_t a = groups_[size_t(best)].coords.size(), b = groups_[size_t(g)].coords.size();
      if (b > a || (b == a && g < best)) best = g;
    }
    return best;
  }

  // folds every other reached group into `target` (ascending ids)
  void fold(int32_t target, std::vector<int32_t> r, ComponentStep& st) {
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
    Group& t = groups_[size_t(target)];
    if (!t.live) bad("component merge target is inactive");
    for (int32_t g : r) {
      if (g == target) continue;
      Group& s = groups_[size_t(g)];
      if (!s.live) bad("component merge source is inactive");
      for (int32_t c : s.coords) owner_[size_t(c)] = target;
      t.coords.insert(t.coords.end(), s.coords.begin(), s.coords.end());
      s.coords.clear();
      s.live = false;
      out_.component_work += 2.0 * std::ldexp(1.0, int(t.coords.size()));
      st.sources.push_back(g);
      ++out_.merges;
    }
    out_.comp_max_k[size_t(target)] = std::max(out_.comp_max_k[size_t(target)], int32_t(t.coords.size()));
  }

  Instr localize(const Instr& in, int32_t g) const {
    Instr l = in;
    const auto& cs = groups_[size_t(g)].coords;
    l.nq = int32_t(cs.size());
    l.xmask = 0;
    l.zmask = 0;
    for (size_t j = 0; j < cs.size(); ++j) {
      int32_t pos = position_of(cs[j]);
      if (pos < 0) bad("component planner tried to remap an inactive coordinate");
      l.xmask |= ((in.xmask >> pos) & 1) << j;
      l.zmask |= ((in.zmask >> pos) & 1) << j;
    }
    return l;
  }

  void step(const Instr& in, ComponentStep& st) {
    const int32_t k = int32_t(order_.size());
    const double dense_dim = std::ldexp(1.0, k);
    switch (in.op) {
      case Op::Rotate: {
        if (in.nq != k) bad("component planner saw a rotation with the wrong active width");
        ++out_.quantum_ops;
        out_.dense_work += dense_dim;
        auto r = reached(in);
        if (r.empty()) {
          st.kind = StepKind::IgnoredPhase;  // a phase on the k == 0 scalar
          return;
        }
        int32_t g = pick_target(r);
        fold(g, r, st);
        st.kind = StepKind::Rotation;
        st.comp = g;
        st.local = localize(in, g);
        out_.component_work += std::ldexp(1.0, st.local.nq) + kDispatch;
        note_live();
        return;
      }
      case Op::Promote: {
        ++out_.quantum_ops;
        out_.dense_work += 2.0 * dense_dim;
        st.kind = StepKind::Promotion;
        st.comp = add_coordinate();
        st.local = in;
        st.local.nq = 0;
        out_.component_work += 2.0 + kDispatch;
        note_live();
        return;
      }
      case Op::Measure: {
        if (in.exp_val >= 0) return;  // expectation probes read the dense state
        if (in.nq != k) bad("component planner saw a measurement with the wrong active width");
        ++out_.quantum_ops;
        out_.dense_work += 2.0 * dense_dim;
        auto r = reached(in);
        if (r.empty()) bad("active measurement has no component support");
        int32_t g = pick_target(r);
        fold(g, r, st);
        note_live();
        if (in.pivot < 0 || in.pivot >= k) bad("component planner saw an invalid measurement pivot");
        const int32_t pcoord = order_[size_t(in.pivot)];
        Instr l = localize(in, g);
        auto& cs = groups_[size_t(g)].coords;
        auto it = std::find(cs.begin(), cs.end(), pcoord);
        if (it == cs.end()) bad("component measurement pivot is not in its target component");
        // src/active.rs:290-339: the kernel around an explicit local pivot
        l.pivot = int32_t(it - cs.begin());
        if (l.xmask) {
          if (!((l.xmask >> l.pivot) & 1)) bad("nondiagonal active measurement pivot must have an X component");
          l.diagonal = false;
        } else {
          if (!((l.zmask >> l.pivot) & 1)) bad("diagonal active measurement pivot must have a Z component");
          l.diagonal = true;
        }
        out_.component_work += 2.0 * std::ldexp(1.0, l.nq) + kDispatch;
        cs.erase(it);
        order_.erase(order_.begin() + in.pivot);
        owner_[size_t(pcoord)] = -1;
        st.kind = StepKind::Measurement;
        st.comp = g;
        st.local = l;
        st.deactivate = cs.empty();
        if (st.deactivate) groups_[size_t(g)].live = false;
        note_live();
        return;
      }
      default:
        return;
    }
  }
};

}  // namespace

ComponentPlan plan_components(const Program& p) { return Planner(p).run(); }

}  // namespace ticit
