// Program -> flat device tables, with the static checks ticit's batch
// executor makes at run time (src/sampler/batch/symbols.rs:14-128,
// src/sampler/batch/active.rs:272-276, 339-343, 739-748) moved to upload time.
#include "device_program.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <set>
#include <unordered_map>

#include "component_plan.hpp"
#include "frontend.hpp"
#include "pinning.hpp"

namespace ticit {

namespace {

const double kInvSqrt2 = 0.70710678118654752440;

struct Cx {
  double re, im;
};

Cx phase_factor(int ph) {
  switch (ph & 3) {
    case 0: return Cx{1.0, 0.0};
    case 1: return Cx{0.0, 1.0};
    case 2: return Cx{-1.0, 0.0};
    default: return Cx{0.0, -1.0};
  }
}

double gap_denominator(double p) {
  if (!(p > 0.0 && p < 1.0)) throw TicitError(1, "geometric gap probability must be in (0, 1)");
  return std::log1p(-p);
}

// Largest uniform below which floor(ln(u) / denom) >= trials is certain
// (no event among `trials`): (1-p)^trials shrunk by a relative 1e-9. The
// kernels' ln and division err by a few ulps, i.e. a relative 1e-15 of a
// gap below 700 / |denom|, while the shrink moves the exact gap up by
// 1e-9 / |denom|; beyond trials * |denom| = 700 the bound would approach
// the DBL_MIN clamp of the uniform, so the shortcut is off there.
double no_event_below(double trials, double denom) {
  const double x = trials * denom;
  if (!(denom < 0.0) || !(x > -700.0)) return 0.0;
  return std::exp(x) * (1.0 - 1e-9);
}

void check_probability(double p) {
  if (!(p >= 0.0 && p <= 1.0)) throw TicitError(1, "probability must be between 0 and 1");
}

// The lane tier's schedule (a list scheduler over the compact steps): a
// record is deferred with level 1 + the highest level among the pending
// records whose slots it reads; an in-order step that reads a pending
// record's slot first flushes every pending record, level by level (a
// level's records read only earlier levels and in-order steps, so they run
// side by side); detectors read only records and slots and run last.
void build_lane_schedule(HostTables& t) {
  t.lsteps.clear();
  t.bsteps.clear();
  std::vector<std::pair<int, DevStep>> pending;  // (level, record)
  std::unordered_map<int32_t, int> slot_level;   // slot -> level of its pending record
  std::vector<DevStep> dets;
  std::vector<DevStep> coin_run;  // coin draws not yet emitted
  auto emit_coins = [&]() {
    if (coin_run.empty()) return;
    DevStep b{};
    b.op = kOpCoinBatch;
    b.partial = -1;
    b.a = int32_t(t.bsteps.size());
    b.b = int32_t(coin_run.size());
    t.bsteps.insert(t.bsteps.end(), coin_run.begin(), coin_run.end());
    t.lsteps.push_back(b);
    coin_run.clear();
  };
  auto flush = [&]() {
    emit_coins();  // pending records may read the coins' branch slots
    if (pending.empty()) return;
    std::stable_sort(pending.begin(), pending.end(),
                     [](const std::pair<int, DevStep>& x, const std::pair<int, DevStep>& y) { return x.first < y.first; });
    size_t i = 0;
    while (i < pending.size()) {
      size_t j = i;
      while (j < pending.size() && pending[j].first == pending[i].first) ++j;
      DevStep b{};
      b.op = kOpRecBatch;
      b.partial = -1;
      b.a = int32_t(t.bsteps.size());
      b.b = int32_t(j - i);
      for (size_t q = i; q < j; ++q) t.bsteps.push_back(pending[q].second);
      t.lsteps.push_back(b);
      i = j;
    }
    pending.clear();
    slot_level.clear();
  };
  auto reads_pending = [&](const DevStep& s, int* level) {
    bool any = false;
    *level = 0;
    for (uint32_t j = 0; j < s.resid_len; ++j) {
      auto it = slot_level.find(int32_t(t.resid[s.resid_off + j]));
      if (it != slot_level.end()) {
        any = true;
        *level = std::max(*level, it->second);
      }
    }
    return any;
  };
  for (const DevStep& s : t.steps) {
    const uint32_t op = s.op & 0xff;
    int level = 0;
    if (op == kOpRecord && !(s.op & kFlagExpVal)) {
      reads_pending(s, &level);
      pending.emplace_back(level + 1, s);
      if (s.b >= 0) slot_level[s.b] = level + 1;
      continue;
    }
    if (op == kOpDetector) {
      dets.push_back(s);
      continue;
    }
    if (op == kOpCoin && !(s.op & (kFlagExpVal | kFlagForced))) {
      // a dormant coin: its draw joins the run of consecutive draws, its
      // outcome record is deferred like any record (it reads the branch
      // slot the draw writes, which every flush emits first)
      coin_run.push_back(s);
      DevStep r = s;
      r.op = (s.op & ~0xffu) | kOpRecord;
      reads_pending(r, &level);
      pending.emplace_back(level + 1, r);
      if (r.b >= 0) slot_level[r.b] = level + 1;
      continue;
    }
    // draws stay in program order, and the step may read a coin's branch
    emit_coins();
    if (reads_pending(s, &level)) flush();
    t.lsteps.push_back(s);
  }
  flush();
  if (!dets.empty()) {
    DevStep b{};
    b.op = kOpDetBatch;
    b.partial = -1;
    b.a = int32_t(t.bsteps.size());
    b.b = int32_t(dets.size());
    t.bsteps.insert(t.bsteps.end(), dets.begin(), dets.end());
    t.lsteps.push_back(b);
  }
  // Batch entries carry short id lists inline, saving each lane a dependent
  // global load (the coin draws keep their fields).
  for (const DevStep& l : t.lsteps) {
    const uint32_t lop = l.op & 0xff;
    if (lop != kOpRecBatch && lop != kOpDetBatch) continue;
    for (int32_t j = l.a; j < l.a + l.b; ++j) {
      DevStep& q = t.bsteps[size_t(j)];
      if (lop == kOpDetBatch && (q.op & kFlagDetRecords)) {
        if (q.c < 0 || q.c > 3) continue;
        const int32_t off = q.b;
        q.partial = q.c > 0 ? int32_t(t.det_recs[size_t(off)]) : 0;
        q.resid_off = q.c > 1 ? t.det_recs[size_t(off) + 1] : 0u;
        q.full = q.c > 2 ? int32_t(t.det_recs[size_t(off) + 2]) : 0;
      } else {
        if (q.resid_len > 2) continue;
        const uint32_t off = q.resid_off;
        q.c = q.resid_len > 0 ? int32_t(t.resid[off]) : 0;
        q.full = q.resid_len > 1 ? int32_t(t.resid[off + 1]) : 0;
      }
      q.op |= kFlagInline;
    }
  }
  t.nlsteps = int32_t(t.lsteps.size());
  if (t.lsteps.empty()) t.lsteps.push_back(DevStep{});
  if (t.bsteps.empty()) t.bsteps.push_back(DevStep{});
}

}  // namespace

HostTables lower_to_device(const Program& p, bool components, const std::vector<ForcedBranch>* forced,
                           const ExpectedBits* expected) {
  if (expected && !expected->detectors.empty() && expected->detectors.size() != size_t(p.ndetectors))
    throw TicitError(1, "expected_detectors has length " + std::to_string(expected->detectors.size()) +
                            ", expected " + std::to_string(p.ndetectors));
  if (expected && !expected->observables.empty() && expected->observables.size() != p.observables.size())
    throw TicitError(1, "expected_observables has length " + std::to_string(expected->observables.size()) +
                            ", expected " + std::to_string(p.observables.size()));
  std::map<int32_t, const SymBool*> forced_at;
  if (forced)
    for (const ForcedBranch& f : *forced) forced_at[f.instruction] = &f.plan;
  HostTables t;
  t.n = p.n;
  t.initial_k = p.initial_k;
  t.max_k = p.max_k;
  t.nrec = p.nrecords;
  t.nexp = p.nexpvals;
  t.ndet = p.ndetectors;
  t.nobs = int(p.observables.size());
  t.nsym = p.nsymbols;
  if (p.max_k >= 62) throw TicitError(3, "active qubit count is too large for machine basis indices");

  // Exogenous symbol set (src/sampler/exogenous.rs:283-298).
  std::vector<char> exo(size_t(p.nsymbols) + 1, 0);
  auto mark = [&](int32_t c) {
    if (c <= 0 || c > p.nsymbols) throw TicitError(1, "exogenous condition exceeds program symbol table");
    exo[c] = 1;
  };
  for (const auto& d : p.categorical) for (int32_t c : d.conds) mark(c);
  for (const auto& g : p.rare) for (const auto& s : g.sets) for (int32_t c : s) mark(c);
  for (int32_t c : p.bern_conds) mark(c);
  for (const auto& g : p.lowp) for (int32_t c : g.conds) mark(c);
  for (int32_t c = 1; c <= p.nsymbols; ++c)
    if (exo[c]) t.exo_syms.push_back(c);

  // Noise plan.
  for (const auto& d : p.categorical) {
    DevCat dc{};
    dc.nbits = uint32_t(d.nbits);
    dc.nrows = uint32_t(d.probs.size());
    dc.prob_off = uint32_t(t.f64pool.size());
    dc.assign_off = uint32_t(t.u64pool.size());
    dc.cond_off = uint32_t(t.i32pool.size());
    if (d.probs.empty() || d.conds.size() != size_t(d.nbits))
      throw TicitError(1, "invalid categorical symbolic distribution");
    // cumulative sums in row order: the kernels binary-search the first
    // row whose running sum reaches the uniform, which is the row the
    // sequential walk stops at (src/random.rs:40-50)
    double cum = 0.0;
    for (size_t r = 0; r < d.probs.size(); ++r) {
      cum += d.probs[r];
      t.f64pool.push_back(cum);
      t.u64pool.push_back(d.assignments[r]);
    }
    for (int32_t c : d.conds) t.i32pool.push_back(c);
    t.cats.push_back(dc);
  }
  for (const auto& g : p.rare) {
    DevRare dr{};
    dr.enabled = g.event_p > 0.0 && !g.sets.empty();
    dr.denom = dr.enabled ? gap_denominator(g.event_p) : 0.0;
    dr.nsets = uint32_t(g.sets.size());
    dr.none_below = dr.enabled ? no_event_below(double(dr.nsets), dr.denom) : 0.0;
    dr.nbits = uint32_t(g.nbits);
    dr.nev = uint32_t(g.event_rows.size());
    dr.cond_off = uint32_t(t.i32pool.size());
    for (const auto& s : g.sets) {
      if (s.size() != size_t(g.nbits)) throw TicitError(1, "rare categorical set has the wrong width");
      for (int32_t c : s) t.i32pool.push_back(c);
    }
    dr.assign_off = uint32_t(t.u64pool.size());
    for (uint64_t a : g.assignments) t.u64pool.push_back(a);
    dr.ev_row_off = uint32_t(t.i32pool.size());
    for (int32_t r : g.event_rows) {
      if (r < 0 || size_t(r) >= g.assignments.size()) throw TicitError(1, "rare categorical event row out of range");
      t.i32pool.push_back(r);
    }
    dr.ev_prob_off = uint32_t(t.f64pool.size());
    double cum = 0.0;
    for (double q : g.event_probs) {
      cum += q;
      t.f64pool.push_back(cum);
    }
    if (dr.enabled && dr.nev == 0) throw TicitError(1, "rare categorical group has no event rows");
    t.rares.push_back(dr);
  }
  for (size_t i = 0; i < p.bern_conds.size(); ++i) {
    DevBern b{};
    b.cond = p.bern_conds[i];
    double pr = p.bern_probs[i];
    check_probability(pr);
    if (pr <= 0.0) {
      b.mode = kBernNone;
    } else if (pr >= 1.0) {
      b.mode = kBernAll;
    } else {
      b.invert = pr > 0.5;
      double q = b.invert ? 1.0 - pr : pr;
      if (q <= 0.0) {
        b.mode = b.invert ? kBernAll : kBernNone;
        b.invert = 0;
      } else if (q == 0.5) {
        b.mode = kBernHalf;
      } else if (q < 0.02) {
        b.mode = kBernLow;
        b.denom = gap_denominator(q);
        b.miss_below = no_event_below(1.0, b.denom);
      } else {
        b.mode = kBernCoins;
        uint64_t raw = uint64_t(q * 256.0);
        uint64_t top = raw < 128 ? raw : 127;
        double pt = double(top) / 256.0;
        b.top_bits = uint32_t(top);
        double left = q - pt;
        b.left_denom = left > 0.0 ? gap_denominator(left / (1.0 - pt)) : 0.0;
        b.left_miss_below = left > 0.0 ? no_event_below(1.0, b.left_denom) : 0.0;
      }
    }
    t.berns.push_back(b);
  }
  for (const auto& g : p.lowp) {
    DevLowp l{};
    l.enabled = g.p > 0.0 && !g.conds.empty();
    l.denom = l.enabled ? gap_denominator(g.p) : 0.0;
    l.nconds = uint32_t(g.conds.size());
    l.none_below = l.enabled ? no_event_below(double(l.nconds), l.denom) : 0.0;
    l.cond_off = uint32_t(t.i32pool.size());
    for (int32_t c : g.conds) t.i32pool.push_back(c);
    t.lowps.push_back(l);
  }

  // Instructions, exogenous partials and runtime slots.
  std::map<std::vector<int32_t>, int32_t> partial_index;
  std::vector<std::vector<int32_t>> partials;
  std::vector<int32_t> slot_of(size_t(p.nsymbols) + 1, -1);
  std::vector<char> assigned(size_t(p.nsymbols) + 1, 0);
  auto slot_for = [&](int32_t c) {
    if (c <= 0 || c > p.nsymbols) throw TicitError(1, "symbolic condition exceeds batch executor symbol table");
    if (exo[c]) throw TicitError(4, "an exogenous symbol cannot be assigned at run time");
    if (slot_of[c] < 0) slot_of[c] = t.nslots++;
    return slot_of[c];
  };
  auto assign = [&](int32_t c) {
    int32_t s = slot_for(c);
    if (assigned[c]) throw TicitError(4, "a runtime symbol is assigned twice");
    assigned[c] = 1;
    return s;
  };
  // Execution layout: dense (one component at offset 0) or the planner's
  // tensor components (src/sampler/batch/runtime.rs:388-552).
  ComponentPlan plan;
  std::vector<int32_t> width;  // current width of every component
  t.components = components;
  if (components) {
    plan = plan_components(p);
    if (!plan.planned) throw TicitError(4, "program does not contain an executable active component plan");
    uint32_t off = 0;
    for (int32_t mk : plan.comp_max_k) {
      t.comp_off.push_back(off);
      off += uint32_t(1) << mk;
      t.max_local_k = std::max(t.max_local_k, mk);
    }
    t.state_dim = off;
    width.assign(plan.comp_max_k.size(), 0);
  } else {
    t.comp_off.push_back(0);
    t.state_dim = uint32_t(1) << p.max_k;
    t.max_local_k = p.max_k;
  }
  auto blank = [](uint32_t op) {
    DevInstr d{};
    d.op = op;
    d.partial = -1;
    d.record = -1;
    d.rc_slot = -1;
    d.branch_slot = -1;
    d.detector = -1;
    d.ordinal = -1;
    d.src_comp = -1;
    return d;
  };
  // src/active.rs:193-229
  auto fill_rotation = [&](DevInstr& d, const Instr& in) {
    d.c = std::cos(in.angle);
    d.s = std::sin(in.angle);
    Cx e = phase_factor(in.phase);
    double a = 0.0, b = -d.s;  // (0 - i sin) * i^phase, in num-complex's order
    d.me_re = a * e.re - b * e.im;
    d.me_im = a * e.im + b * e.re;
    d.xmask = in.xmask;
    d.zmask = in.zmask;
    if (in.xmask == 0) {
      d.op |= kFlagDiagonal;
    } else {
      if (in.zmask == 0) d.op |= kFlagUniformImag;
      d.pivot = 63 - __builtin_clzll(in.xmask);
    }
    if (__builtin_popcountll(in.xmask & in.zmask) & 1) d.op |= kFlagXzOdd;
  };
  // src/active.rs:290-339
  auto fill_measure = [&](DevInstr& d, const Instr& in) {
    d.xmask = in.xmask;
    d.zmask = in.zmask;
    d.pivot = in.pivot;
    if (in.diagonal) {
      d.op |= kFlagDiagonal;
      if (in.diag_phase_bit) d.op |= kFlagDiagPhase;
      d.zwop = in.zmask & ~(1ull << in.pivot);
    }
    Cx e = phase_factor(in.phase);
    d.c1_re = e.re * kInvSqrt2;
    d.c1_im = -e.im * kInvSqrt2;
  };
  // src/sampler/batch/mod.rs:597-680: every source of a step folded into its
  // component, each fold its own device instruction
  // Traffic of a merge chain as the wide tier runs it: the factors are read
  // (once per pass that uses them) and the merged state is written only when
  // the next operation is not a measurement of it.
  double chain_reads = 0.0;
  bool chain = false;
  auto emit_merges = [&](const ComponentStep& st) {
    if (width[size_t(st.comp)] == 0) throw TicitError(4, "active component merge target is inactive");
    chain = !st.sources.empty();
    chain_reads = chain ? std::ldexp(1.0, width[size_t(st.comp)]) : 0.0;
    for (int32_t src : st.sources) {
      if (width[size_t(src)] == 0) throw TicitError(4, "active component merge source is inactive");
      DevInstr m = blank(kOpMerge);
      m.comp = st.comp;
      m.comp_off = t.comp_off[size_t(st.comp)];
      m.nq = width[size_t(st.comp)];
      m.src_comp = src;
      m.src_off = t.comp_off[size_t(src)];
      m.src_k = width[size_t(src)];
      if (m.nq + m.src_k > plan.comp_max_k[size_t(st.comp)])
        throw TicitError(4, "active component merge target has insufficient storage");
      chain_reads += std::ldexp(1.0, m.src_k);
      width[size_t(st.comp)] += m.src_k;
      width[size_t(src)] = 0;
      t.instrs.push_back(m);
    }
  };
  if (components) {
    // src/sampler/batch/mod.rs:366-402: initial components hold |0>
    for (int32_t c = 0; c < plan.initial_components; ++c) {
      DevInstr z = blank(kOpPromote | kFlagFresh | kFlagZeroState);
      z.comp = c;
      z.comp_off = t.comp_off[size_t(c)];
      z.op |= kFlagConstant;  // never evaluated: a fixed state
      width[size_t(c)] = 1;
      t.instrs.push_back(z);
    }
  }

  int k = p.initial_k;
  for (size_t ii = 0; ii < p.instrs.size(); ++ii) {
    const Instr& in = p.instrs[ii];
    const ComponentStep* st = components ? &plan.steps[ii] : nullptr;
    DevInstr d = blank(uint32_t(in.op));
    d.nq = in.nq;
    if (in.record < 0 || in.record > p.nrecords) throw TicitError(1, "measurement record id out of range");
    if (in.record > 0) d.record = in.record - 1;
    if (in.expr.constant) d.op |= kFlagConstant;
    switch (in.op) {
      case Op::Rotate: {
        if (in.nq != k) throw TicitError(4, "rotation kernel dimension does not match batch active state");
        t.quantum_ops++;
        t.dense_work += std::ldexp(1.0, k);
        if (st) {
          if (st->kind == StepKind::IgnoredPhase) continue;  // a phase on the k == 0 scalar
          emit_merges(*st);
          if (st->local.nq != width[size_t(st->comp)])
            throw TicitError(4, "component rotation width does not match its component");
          fill_rotation(d, st->local);
          d.nq = st->local.nq;
          d.comp = st->comp;
          d.comp_off = t.comp_off[size_t(st->comp)];
          if (chain) t.amp_traffic += chain_reads + std::ldexp(1.0, d.nq);  // one pass writes the product
        } else {
          fill_rotation(d, in);
        }
        t.amp_traffic += std::ldexp(2.0, d.nq);
        break;
      }
      case Op::Promote:
        if (k >= p.n) throw TicitError(4, "cannot promote a dormant qubit when none remain");
        if (in.nq != k) throw TicitError(4, "promotion width does not match batch active state");
        d.c = std::cos(in.angle);
        d.s = std::sin(in.angle);
        t.quantum_ops++;
        t.dense_work += std::ldexp(3.0, k);
        if (st) {
          if (width[size_t(st->comp)] != 0) throw TicitError(4, "active component promotion target is unavailable");
          d.op |= kFlagFresh;
          d.nq = 0;
          d.comp = st->comp;
          d.comp_off = t.comp_off[size_t(st->comp)];
          width[size_t(st->comp)] = 1;
          t.amp_traffic += 2.0;
        } else {
          t.amp_traffic += std::ldexp(3.0, k);
        }
        ++k;
        break;
      case Op::Record:
        break;
      case Op::Detector:
        if (in.detector <= 0 || in.detector > p.ndetectors) throw TicitError(1, "detector id out of range");
        d.detector = in.detector - 1;
        if (in.postselect) {
          d.op |= kFlagPostselect;
          t.post_dets.push_back(uint32_t(d.detector));
        }
        if (!in.det_records.empty()) {
          // the value is the records' XOR; the constant flag carries only
          // the expected bit below
          d.op = (d.op & ~kFlagConstant) | kFlagDetRecords;
          d.det_off = uint32_t(t.det_recs.size());
          d.det_len = uint32_t(in.det_records.size());
          for (int32_t r : in.det_records) {
            if (r <= 0 || r > p.nrecords)
              throw TicitError(1, "detector references an out-of-range measurement record");
            t.det_recs.push_back(uint32_t(r - 1));
          }
        }
        if (expected && !expected->detectors.empty() && expected->detectors[size_t(d.detector)] != 0)
          d.op ^= kFlagConstant;
        break;
      case Op::Measure: {
        if (k == 0) throw TicitError(4, "cannot measure an active Pauli when k == 0");
        if (in.nq != k) throw TicitError(4, "measurement kernel dimension does not match batch active state");
        t.quantum_ops++;
        t.dense_work += std::ldexp(2.5, k);
        if (st) {
          if (st->kind != StepKind::Measurement) throw TicitError(4, "component measurement step does not match its instruction");
          emit_merges(*st);
          if (st->local.nq != width[size_t(st->comp)])
            throw TicitError(4, "component measurement width does not match its component");
          fill_measure(d, st->local);
          d.nq = st->local.nq;
          d.comp = st->comp;
          d.comp_off = t.comp_off[size_t(st->comp)];
          width[size_t(st->comp)] -= 1;
          if (st->deactivate && width[size_t(st->comp)] != 0)
            throw TicitError(4, "active component measurement did not remove its last coordinate");
        } else {
          fill_measure(d, in);
        }
        if (in.exp_val >= 0) {
          // src/sampler/batch/active.rs:818-850: the probability pass only;
          // no draw, no collapse, the width stays
          if (st) throw TicitError(4, "expectation probes need the dense layout");
          t.amp_traffic += std::ldexp((d.op & kFlagDiagonal) ? 0.5 : 1.0, d.nq);
          d.ordinal = t.nmeasure++;
          break;
        }
        if (st && chain) {
          // measured straight from the factors: they are read by the
          // probability and the projection passes, only the projected half
          // is written
          t.amp_traffic += 2.0 * chain_reads + std::ldexp(0.5, d.nq);
        } else {
          t.amp_traffic += std::ldexp((d.op & kFlagDiagonal) ? 1.5 : 2.5, d.nq);
        }
        d.ordinal = t.nmeasure++;
        d.branch_slot = assign(in.branch);
        --k;
        break;
      }
      case Op::Coin:
        d.ordinal = t.ncoin++;
        if (in.exp_val < 0) d.branch_slot = assign(in.branch);
        break;
    }
    // a pinned branch (src/sampler/pinning.rs:281-288): XOR of branch symbols
    // assigned before it, plus a constant
    auto fit = forced_at.find(int32_t(ii));
    if (fit != forced_at.end()) {
      if ((in.op != Op::Measure && in.op != Op::Coin) || in.exp_val >= 0)
        throw TicitError(4, "a pinned branch must belong to a measurement");
      d.op |= kFlagForced;
      if (fit->second->constant) d.op |= kFlagForcedOne;
      d.forced_off = uint32_t(t.resid.size());
      for (int32_t c : fit->second->conds) {
        if (c <= 0 || c > p.nsymbols || exo[c] || !assigned[c])
          throw TicitError(4, "a pinned branch reads a symbol not assigned before it");
        t.resid.push_back(uint32_t(slot_of[c]));
      }
      d.forced_len = uint32_t(t.resid.size()) - d.forced_off;
    }
    if (in.exp_val >= 0) {
      // src/sampler/batch/runtime.rs:46-75,311-356: a probe writes its
      // expectation slot instead of a record
      if (in.op != Op::Record && in.op != Op::Measure && in.op != Op::Coin)
        throw TicitError(1, "expectation probe on an instruction that cannot carry one");
      if (in.exp_val >= p.nexpvals) throw TicitError(1, "expectation value index is out of range");
      if (in.record != 0 || in.record_condition != 0)
        throw TicitError(1, "expectation probes take no measurement record");
      d.op |= kFlagExpVal;
      d.record = in.exp_val;
    }
    bool evaluates = !(in.op == Op::Detector && !in.det_records.empty());
    if (evaluates) {
      std::vector<int32_t> ex;
      d.resid_off = uint32_t(t.resid.size());
      for (int32_t c : in.expr.conds) {
        if (c <= 0 || c > p.nsymbols) throw TicitError(1, "symbolic condition exceeds batch executor symbol table");
        if (exo[c]) {
          ex.push_back(c);
        } else {
          if (!assigned[c]) throw TicitError(1, "symbolic condition expression has no concrete batch value");
          t.resid.push_back(uint32_t(slot_of[c]));
        }
      }
      d.resid_len = uint32_t(t.resid.size()) - d.resid_off;
      t.expression_terms += in.expr.conds.size();
      if (!ex.empty()) {
        auto it = partial_index.find(ex);
        if (it == partial_index.end()) {
          it = partial_index.emplace(ex, int32_t(partials.size())).first;
          partials.push_back(ex);
        }
        d.partial = it->second;
      }
    }
    if (in.record_condition > 0) d.rc_slot = assign(in.record_condition);
    t.instrs.push_back(d);
  }
  t.npartials = int(partials.size());

  // symbol -> partial CSR
  std::vector<std::vector<uint32_t>> by_sym(size_t(p.nsymbols));
  for (size_t e = 0; e < partials.size(); ++e)
    for (int32_t c : partials[e]) by_sym[size_t(c - 1)].push_back(uint32_t(e));
  t.sym_off.push_back(0);
  for (const auto& v : by_sym) {
    t.sym_part.insert(t.sym_part.end(), v.begin(), v.end());
    t.sym_off.push_back(uint32_t(t.sym_part.size()));
  }

  // Toggle lists: the partials one noise event flips, with partials named
  // by two of its symbols cancelled, so an event costs one list walk.
  auto begin_lists = [&]() {
    t.tl_off.push_back(uint32_t(t.tl_part.size()));
    return uint32_t(t.tl_off.size() - 1);
  };
  auto push_list = [&](const std::vector<int32_t>& syms) {
    std::vector<uint32_t> parts;
    for (int32_t c : syms) {
      if (c <= 0 || c > p.nsymbols) throw TicitError(1, "noise condition out of range");
      parts.insert(parts.end(), by_sym[size_t(c - 1)].begin(), by_sym[size_t(c - 1)].end());
    }
    std::sort(parts.begin(), parts.end());
    for (size_t i = 0; i < parts.size();) {
      size_t j = i;
      while (j < parts.size() && parts[j] == parts[i]) ++j;
      if ((j - i) & 1) t.tl_part.push_back(parts[i]);
      i = j;
    }
    t.tl_off.push_back(uint32_t(t.tl_part.size()));
  };
  auto set_bits = [](const std::vector<int32_t>& conds, uint64_t a) {
    std::vector<int32_t> out;
    for (size_t b = 0; b < conds.size(); ++b)
      if ((a >> b) & 1) out.push_back(conds[b]);
    return out;
  };
  for (size_t i = 0; i < p.categorical.size(); ++i) {
    const auto& d = p.categorical[i];
    t.cats[i].tl_base = begin_lists();
    for (uint64_t a : d.assignments) push_list(set_bits(d.conds, a));
  }
  for (size_t i = 0; i < p.rare.size(); ++i) {
    const auto& g = p.rare[i];
    t.rares[i].tl_base = begin_lists();
    for (const auto& set : g.sets)
      for (int32_t row : g.event_rows) push_list(set_bits(set, g.assignments[size_t(row)]));
  }
  for (size_t i = 0; i < p.lowp.size(); ++i) {
    t.lowps[i].tl_base = begin_lists();
    for (int32_t c : p.lowp[i].conds) push_list({c});
  }

  t.obs_off.push_back(0);
  for (const auto& ob : p.observables) {
    for (int32_t r : ob) {
      if (r <= 0 || r > p.nrecords) throw TicitError(1, "observable references an out-of-range measurement record");
      t.obs_recs.push_back(uint32_t(r - 1));
    }
    t.obs_off.push_back(uint32_t(t.obs_recs.size()));
  }
  t.obs_flip.assign(std::max<size_t>(1, p.observables.size()), 0);
  if (expected && !expected->observables.empty())
    for (size_t o = 0; o < p.observables.size(); ++o) t.obs_flip[o] = expected->observables[o] != 0;
  t.ninstr = int32_t(t.instrs.size());
  for (size_t i = 0; i < t.instrs.size(); ++i) {
    const DevInstr& d = t.instrs[i];
    DevStep s{};
    s.op = d.op;
    s.partial = d.partial;
    s.resid_off = d.resid_off;
    s.resid_len = d.resid_len;
    s.full = int32_t(i);
    switch (d.op & 0xff) {
      case kOpDetector:
        s.a = d.detector;
        s.b = int32_t(d.det_off);
        s.c = int32_t(d.det_len);
        break;
      case kOpCoin:
        s.full = (d.op & kFlagForced) ? int32_t(i) : d.ordinal;
        s.a = d.record;
        s.b = d.rc_slot;
        s.c = d.branch_slot;
        break;
      default:
        s.a = d.record;
        s.b = d.rc_slot;
        s.c = d.branch_slot;
        break;
    }
    t.steps.push_back(s);
  }
  build_lane_schedule(t);
  {
    std::vector<int> writes(size_t(std::max(0, t.ndet)), 0);
    bool ok = true;
    for (const DevInstr& d : t.instrs)
      if ((d.op & 0xff) == kOpDetector) {
        if (d.detector < 0 || d.detector >= t.ndet || writes[size_t(d.detector)]++) ok = false;
      }
    for (int w : writes) ok = ok && w == 1;
    t.dets_once = ok;
  }
  // keep every table non-empty so device pointers are always valid
  if (t.resid.empty()) t.resid.push_back(0);
  if (t.det_recs.empty()) t.det_recs.push_back(0);
  t.npost = int32_t(t.post_dets.size());
  if (t.post_dets.empty()) t.post_dets.push_back(0);
  if (t.obs_recs.empty()) t.obs_recs.push_back(0);
  if (t.sym_part.empty()) t.sym_part.push_back(0);
  if (t.tl_off.empty()) t.tl_off.push_back(0);
  if (t.tl_part.empty()) t.tl_part.push_back(0);
  if (t.exo_syms.empty()) t.exo_syms.push_back(0);
  if (t.f64pool.empty()) t.f64pool.push_back(0);
  if (t.u64pool.empty()) t.u64pool.push_back(0);
  if (t.i32pool.empty()) t.i32pool.push_back(0);
  if (t.instrs.empty()) t.instrs.push_back(DevInstr{});
  if (t.steps.empty()) t.steps.push_back(DevStep{});
  return t;
}

}  // namespace ticit
