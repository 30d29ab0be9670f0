// Plan-blob codec: a flat little-endian u64 stream (layout in
// include/ticit_plan.h). The blob is the boundary format between a program
// producer (this library's front end, or a Rust host serializing ticit's
// FactoredInstructionProgram) and the sampling backend.
#include <cstring>

#include "program.hpp"
#include "../../../include/ticit_plan.h"

namespace ticit {

namespace {

uint64_t dbits(double d) {
  uint64_t u;
  std::memcpy(&u, &d, 8);
  return u;
}

double bitsd(uint64_t u) {
  double d;
  std::memcpy(&d, &u, 8);
  return d;
}

struct Reader {
  const uint64_t* w;
  size_t n;
  size_t pos = 0;
  bool ok = true;
  uint64_t next() {
    if (pos >= n) {
      ok = false;
      return 0;
    }
    return w[pos++];
  }
  // a count that must fit in what remains of the blob
  size_t count(size_t per_item = 1) {
    uint64_t c = next();
    if (!ok || per_item == 0 || c > (n - pos) / per_item + 1) {
      ok = false;
      return 0;
    }
    return size_t(c);
  }
  int32_t i32() { return int32_t(uint32_t(next())); }
};

}  // namespace

std::vector<uint64_t> encode_program(const Program& p) {
  std::vector<uint64_t> o;
  o.push_back(TICIT_PLAN_MAGIC);
  o.push_back(TICIT_PLAN_VERSION);
  o.push_back(uint64_t(p.n));
  o.push_back(uint64_t(p.initial_k));
  o.push_back(uint64_t(p.max_k));
  o.push_back(uint64_t(p.nsymbols));
  o.push_back(uint64_t(p.nrecords));
  o.push_back(uint64_t(p.ndetectors));
  o.push_back(uint64_t(p.nexpvals));
  o.push_back((p.use_components ? TICIT_PLAN_FLAG_COMPONENTS : 0) | (uint64_t(uint32_t(p.circuit_layers)) << 32));
  o.push_back(uint64_t(p.instrs.size()));
  for (const auto& in : p.instrs) {
    uint64_t w0 = uint64_t(uint8_t(in.op)) | (uint64_t(in.diagonal) << 8) |
                  (uint64_t(in.diag_phase_bit & 1) << 9) | (uint64_t(in.expr.constant) << 10) | (uint64_t(in.postselect) << 11) |
                  (uint64_t(uint32_t(in.nq) & 0xffff) << 16) | (uint64_t(uint32_t(in.pivot) & 0xffff) << 32) |
                  (uint64_t(uint32_t(in.phase) & 3) << 48);
    o.push_back(w0);
    o.push_back(in.xmask);
    o.push_back(in.zmask);
    o.push_back(dbits(in.angle));
    o.push_back(uint64_t(uint32_t(in.branch)) | (uint64_t(uint32_t(in.record)) << 32));
    o.push_back(uint64_t(uint32_t(in.record_condition)) | (uint64_t(uint32_t(in.detector)) << 32));
    o.push_back(uint64_t(uint32_t(in.exp_val)));
    o.push_back(uint64_t(in.expr.conds.size()));
    for (int32_t c : in.expr.conds) o.push_back(uint64_t(uint32_t(c)));
    o.push_back(uint64_t(in.det_records.size()));
    for (int32_t r : in.det_records) o.push_back(uint64_t(uint32_t(r)));
  }
  o.push_back(uint64_t(p.categorical.size()));
  for (const auto& d : p.categorical) {
    o.push_back(uint64_t(d.nbits));
    o.push_back(uint64_t(d.conds.size()));
    for (int32_t c : d.conds) o.push_back(uint64_t(uint32_t(c)));
    o.push_back(uint64_t(d.probs.size()));
    for (size_t r = 0; r < d.probs.size(); ++r) {
      o.push_back(d.assignments[r]);
      o.push_back(dbits(d.probs[r]));
    }
  }
  o.push_back(uint64_t(p.rare.size()));
  for (const auto& g : p.rare) {
    o.push_back(dbits(g.event_p));
    o.push_back(uint64_t(g.nbits));
    o.push_back(uint64_t(g.sets.size()));
    for (const auto& s : g.sets) {
      o.push_back(uint64_t(s.size()));
      for (int32_t c : s) o.push_back(uint64_t(uint32_t(c)));
    }
    o.push_back(uint64_t(g.probs.size()));
    for (size_t r = 0; r < g.probs.size(); ++r) {
      o.push_back(g.assignments[r]);
      o.push_back(dbits(g.probs[r]));
    }
    o.push_back(uint64_t(g.event_rows.size()));
    for (size_t r = 0; r < g.event_rows.size(); ++r) {
      o.push_back(uint64_t(uint32_t(g.event_rows[r])));
      o.push_back(dbits(g.event_probs[r]));
    }
  }
  o.push_back(uint64_t(p.bern_conds.size()));
  for (size_t i = 0; i < p.bern_conds.size(); ++i) {
    o.push_back(uint64_t(uint32_t(p.bern_conds[i])));
    o.push_back(dbits(p.bern_probs[i]));
  }
  o.push_back(uint64_t(p.lowp.size()));
  for (const auto& g : p.lowp) {
    o.push_back(dbits(g.p));
    o.push_back(uint64_t(g.conds.size()));
    for (int32_t c : g.conds) o.push_back(uint64_t(uint32_t(c)));
  }
  o.push_back(uint64_t(p.observables.size()));
  for (const auto& ob : p.observables) {
    o.push_back(uint64_t(ob.size()));
    for (int32_t r : ob) o.push_back(uint64_t(uint32_t(r)));
  }
  o.push_back(TICIT_PLAN_MAGIC);
  return o;
}

bool decode_program(const uint64_t* words, size_t nwords, Program& p, std::string& err) {
  Reader r{words, nwords};
  if (r.next() != TICIT_PLAN_MAGIC) {
    err = "plan blob has a bad magic word";
    return false;
  }
  const uint64_t version = r.next();
  if (version != 1 && version != TICIT_PLAN_VERSION) {
    err = "plan blob has an unsupported version";
    return false;
  }
  p = Program{};
  p.n = r.i32();
  p.initial_k = r.i32();
  p.max_k = r.i32();
  p.nsymbols = r.i32();
  p.nrecords = r.i32();
  p.ndetectors = r.i32();
  p.nexpvals = r.i32();
  if (version >= 2) {
    const uint64_t flags = r.next();
    p.use_components = (flags & TICIT_PLAN_FLAG_COMPONENTS) != 0;
    p.circuit_layers = int32_t(flags >> 32);
  }
  size_t ninstr = r.count(8);
  p.instrs.resize(ninstr);
  for (size_t i = 0; i < ninstr && r.ok; ++i) {
    Instr& in = p.instrs[i];
    uint64_t w0 = r.next();
    in.op = Op(w0 & 0xff);
    in.diagonal = (w0 >> 8) & 1;
    in.diag_phase_bit = int32_t((w0 >> 9) & 1);
    in.expr.constant = (w0 >> 10) & 1;
    in.postselect = (w0 >> 11) & 1;
    in.nq = int32_t((w0 >> 16) & 0xffff);
    in.pivot = int32_t((w0 >> 32) & 0xffff);
    in.phase = int32_t((w0 >> 48) & 3);
    in.xmask = r.next();
    in.zmask = r.next();
    in.angle = bitsd(r.next());
    uint64_t w4 = r.next();
    in.branch = int32_t(uint32_t(w4));
    in.record = int32_t(uint32_t(w4 >> 32));
    uint64_t w5 = r.next();
    in.record_condition = int32_t(uint32_t(w5));
    in.detector = int32_t(uint32_t(w5 >> 32));
    in.exp_val = r.i32();
    size_t nc = r.count();
    in.expr.conds.resize(nc);
    for (size_t c = 0; c < nc; ++c) in.expr.conds[c] = r.i32();
    size_t nd = r.count();
    in.det_records.resize(nd);
    for (size_t c = 0; c < nd; ++c) in.det_records[c] = r.i32();
    if (uint8_t(in.op) > 5) r.ok = false;
  }
  size_t ncat = r.count(3);
  p.categorical.resize(ncat);
  for (size_t i = 0; i < ncat && r.ok; ++i) {
    auto& d = p.categorical[i];
    d.nbits = r.i32();
    size_t nc = r.count();
    d.conds.resize(nc);
    for (auto& c : d.conds) c = r.i32();
    size_t nrow = r.count(2);
    d.assignments.resize(nrow);
    d.probs.resize(nrow);
    for (size_t k = 0; k < nrow; ++k) {
      d.assignments[k] = r.next();
      d.probs[k] = bitsd(r.next());
    }
  }
  size_t nrare = r.count(4);
  p.rare.resize(nrare);
  for (size_t i = 0; i < nrare && r.ok; ++i) {
    auto& g = p.rare[i];
    g.event_p = bitsd(r.next());
    g.nbits = r.i32();
    size_t ns = r.count();
    g.sets.resize(ns);
    for (auto& s : g.sets) {
      size_t nc = r.count();
      s.resize(nc);
      for (auto& c : s) c = r.i32();
    }
    size_t nrow = r.count(2);
    g.assignments.resize(nrow);
    g.probs.resize(nrow);
    for (size_t k = 0; k < nrow; ++k) {
      g.assignments[k] = r.next();
      g.probs[k] = bitsd(r.next());
    }
    size_t nev = r.count(2);
    g.event_rows.resize(nev);
    g.event_probs.resize(nev);
    for (size_t k = 0; k < nev; ++k) {
      g.event_rows[k] = r.i32();
      g.event_probs[k] = bitsd(r.next());
    }
  }
  size_t nb = r.count(2);
  p.bern_conds.resize(nb);
  p.bern_probs.resize(nb);
  for (size_t i = 0; i < nb && r.ok; ++i) {
    p.bern_conds[i] = r.i32();
    p.bern_probs[i] = bitsd(r.next());
  }
  size_t nl = r.count(2);
  p.lowp.resize(nl);
  for (size_t i = 0; i < nl && r.ok; ++i) {
    p.lowp[i].p = bitsd(r.next());
    size_t nc = r.count();
    p.lowp[i].conds.resize(nc);
    for (auto& c : p.lowp[i].conds) c = r.i32();
  }
  size_t nobs = r.count();
  p.observables.resize(nobs);
  for (size_t i = 0; i < nobs && r.ok; ++i) {
    size_t nrec = r.count();
    p.observables[i].resize(nrec);
    for (auto& c : p.observables[i]) c = r.i32();
  }
  if (r.next() != TICIT_PLAN_MAGIC || !r.ok) {
    err = "plan blob is truncated or malformed";
    return false;
  }
  return true;
}

}  // namespace ticit
