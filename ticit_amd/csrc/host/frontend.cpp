// Circuit front end: parser, lowering, pending-queue optimizer, planner and
// program freezing. Each stage cites the ticit function whose behaviour it
// follows; the code itself is written for this library.
#include "frontend.hpp"

#include "component_plan.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <sstream>
#include <unordered_map>

namespace ticit {

// ============================================================================
// Symbolic booleans (src/symbolic.rs:24-146, src/bits.rs:120-142)
// ============================================================================

void normalize_conditions(std::vector<int32_t>& c) {
  std::sort(c.begin(), c.end());
  size_t w = 0, r = 0;
  while (r < c.size()) {
    size_t e = r + 1;
    while (e < c.size() && c[e] == c[r]) ++e;
    if ((e - r) & 1) c[w++] = c[r];
    r = e;
  }
  c.resize(w);
}

SymBool sym_xor(const SymBool& a, const SymBool& b) {
  SymBool out;
  out.constant = a.constant != b.constant;
  out.conds.reserve(a.conds.size() + b.conds.size());
  size_t i = 0, j = 0;
  while (i < a.conds.size() && j < b.conds.size()) {
    if (a.conds[i] < b.conds[j]) out.conds.push_back(a.conds[i++]);
    else if (a.conds[i] > b.conds[j]) out.conds.push_back(b.conds[j++]);
    else { ++i; ++j; }
  }
  out.conds.insert(out.conds.end(), a.conds.begin() + i, a.conds.end());
  out.conds.insert(out.conds.end(), b.conds.begin() + j, b.conds.end());
  return out;
}

SymBool sym_xor_const(const SymBool& a, bool v) {
  SymBool out = a;
  out.constant = a.constant != v;
  return out;
}

SymBool sym_var(int32_t c) {
  SymBool out;
  out.conds.push_back(c);
  return out;
}

namespace {

inline int word_of(int32_t condition) { return (condition - 1) >> 6; }

// ============================================================================
// Pauli strings: i^phase * prod_q X_q^x Z_q^z (src/pauli.rs:10-16)
// ============================================================================

struct Pauli {
  int n = 0;
  std::vector<uint64_t> x, z;
  int phase = 0;
  Pauli() = default;
  explicit Pauli(int nq) : n(nq), x((nq + 63) / 64, 0), z((nq + 63) / 64, 0) {}
  bool xb(int q) const { return (x[q >> 6] >> (q & 63)) & 1; }
  bool zb(int q) const { return (z[q >> 6] >> (q & 63)) & 1; }
  void setx(int q, bool v) {
    uint64_t m = 1ull << (q & 63);
    if (v) x[q >> 6] |= m; else x[q >> 6] &= ~m;
  }
  void setz(int q, bool v) {
    uint64_t m = 1ull << (q & 63);
    if (v) z[q >> 6] |= m; else z[q >> 6] &= ~m;
  }
  void set_phase(int p) { phase = ((p % 4) + 4) % 4; }
  int ycount() const {
    int c = 0;
    for (size_t i = 0; i < x.size(); ++i) c += __builtin_popcountll(x[i] & z[i]);
    return c;
  }
  bool nonidentity() const {
    for (size_t i = 0; i < x.size(); ++i)
      if (x[i] | z[i]) return true;
    return false;
  }
  bool same_body(const Pauli& o) const { return x == o.x && z == o.z; }
  bool operator==(const Pauli& o) const { return n == o.n && phase == o.phase && x == o.x && z == o.z; }
};

// out <- out * row, carrying the i^2 from row's X crossing out's Z.
inline void mul_into(Pauli& out, const Pauli& row) {
  uint64_t carry = 0;
  for (size_t i = 0; i < out.x.size(); ++i) {
    carry += __builtin_popcountll(out.z[i] & row.x[i]);
    out.x[i] ^= row.x[i];
    out.z[i] ^= row.z[i];
  }
  out.set_phase(out.phase + row.phase + 2 * int(carry & 1));
}

inline Pauli pmul(const Pauli& a, const Pauli& b) {
  Pauli out = a;
  mul_into(out, b);
  return out;
}

inline bool anticommutes(const Pauli& a, const Pauli& b) {
  uint64_t acc = 0;
  for (size_t i = 0; i < a.x.size(); ++i) acc ^= (a.x[i] & b.z[i]) ^ (a.z[i] & b.x[i]);
  return __builtin_popcountll(acc) & 1;
}

inline bool hermitian(const Pauli& p) { return ((p.phase - p.ycount()) & 1) == 0; }

// src/pauli.rs:457-466
inline bool phase_sign(const Pauli& p) {
  int v = ((p.phase - p.ycount()) % 4 + 4) % 4;
  if (v == 0) return false;
  if (v == 2) return true;
  throw TicitError(1, "Pauli measurement requires a Hermitian Pauli with real coefficient");
}

inline Pauli pX(int n, int q) { Pauli p(n); p.setx(q, true); return p; }
inline Pauli pZ(int n, int q) { Pauli p(n); p.setz(q, true); return p; }
inline Pauli pY(int n, int q) { Pauli p(n); p.setx(q, true); p.setz(q, true); p.phase = 1; return p; }

inline Pauli positive_hermitian(Pauli p) {
  if (!hermitian(p)) throw TicitError(4, "Pauli frame row requires a Hermitian Pauli body");
  p.set_phase(p.ycount());
  return p;
}

// ============================================================================
// Clifford tableau: rows[q] = U^dag X_q U, rows[n+q] = U^dag Z_q U
// (src/frames.rs:292-572)
// ============================================================================

struct Clifford {
  int n = 0;
  std::vector<Pauli> rows;
  Clifford() = default;
  explicit Clifford(int nq) : n(nq), rows(2 * nq, Pauli(nq)) {
    for (int q = 0; q < nq; ++q) {
      rows[q].setx(q, true);
      rows[nq + q].setz(q, true);
    }
  }
  static Clifford from_rows(int nq, std::vector<Pauli> xr, std::vector<Pauli> zr) {
    Clifford c;
    c.n = nq;
    c.rows = std::move(xr);
    for (auto& r : zr) c.rows.push_back(std::move(r));
    return c;
  }
  // U^dag P U, multiplying rows in ascending-q, X-before-Z order.
  Pauli preimage(const Pauli& p) const {
    Pauli out(n);
    out.phase = p.phase;
    for (size_t w = 0; w < p.x.size(); ++w) {
      uint64_t bits = p.x[w] | p.z[w];
      while (bits) {
        int b = __builtin_ctzll(bits);
        int q = int(w) * 64 + b;
        if ((p.x[w] >> b) & 1) mul_into(out, rows[q]);
        if ((p.z[w] >> b) & 1) mul_into(out, rows[n + q]);
        bits &= bits - 1;
      }
    }
    return out;
  }
  // Rows whose body is not the identity tableau's (X_q, Z_q). For the others
  // the anticommutation tests of coordinates() are p's own bits, so a frame
  // that moves few rows costs few tests per Pauli.
  mutable std::vector<int> moved;
  // the moved rows' nonzero words (word, x, z), moved_off[i] .. [i + 1]:
  // anticommutation tests read only those
  struct RowWord {
    uint32_t w;
    uint64_t x, z;
  };
  mutable std::vector<RowWord> moved_words;
  mutable std::vector<uint32_t> moved_off;
  // qubits whose X or Z row differs from the identity's in body or phase
  mutable std::vector<uint64_t> touched;
  // touched qubits and the supports of the moved rows: coordinates() leaves
  // every Pauli outside them unchanged
  mutable std::vector<uint64_t> footprint;
  mutable bool moved_ready = false;
  bool anticommutes_moved(const Pauli& p, size_t i) const {
    uint32_t c = 0;
    for (uint32_t j = moved_off[i]; j < moved_off[i + 1]; ++j) {
      const RowWord& rw = moved_words[j];
      c += uint32_t(__builtin_popcountll(p.x[rw.w] & rw.z) + __builtin_popcountll(p.z[rw.w] & rw.x));
    }
    return c & 1;
  }
  void index_moved() const {
    moved.clear();
    moved_words.clear();
    moved_off.assign(1, 0);
    touched.assign(size_t((n + 63) / 64), 0);
    for (int r = 0; r < 2 * n; ++r) {
      const Pauli& row = rows[size_t(r)];
      const int q = r < n ? r : r - n;
      bool trivial = true;
      for (size_t w = 0; w < row.x.size() && trivial; ++w) {
        const uint64_t bit = (size_t(q) >> 6) == w ? (1ull << (q & 63)) : 0ull;
        trivial = row.x[w] == (r < n ? bit : 0ull) && row.z[w] == (r < n ? 0ull : bit);
      }
      if (!trivial) {
        moved.push_back(r);
        for (size_t w = 0; w < row.x.size(); ++w)
          if (row.x[w] | row.z[w]) moved_words.push_back(RowWord{uint32_t(w), row.x[w], row.z[w]});
        moved_off.push_back(uint32_t(moved_words.size()));
      }
      if (!trivial || row.phase != 0) touched[size_t(q) >> 6] |= 1ull << (q & 63);
    }
    footprint = touched;
    for (const RowWord& rw : moved_words) footprint[rw.w] |= rw.x | rw.z;
    moved_ready = true;
  }
  // Decomposition of p over the rows: preimage(result) == p exactly. Bit q
  // of the result's z (x) part is whether p anticommutes with row q (n + q).
  Pauli coordinates(const Pauli& p) const {
    if (!moved_ready) index_moved();
    Pauli out(n);
    out.x = p.x;  // anticommutes(p, Z_q) = x_q(p)
    out.z = p.z;  // anticommutes(p, X_q) = z_q(p)
    for (size_t i = 0; i < moved.size(); ++i) {
      const int r = moved[i];
      const bool a = anticommutes_moved(p, i);
      if (r < n) out.setz(r, a);
      else out.setx(r - n, a);
    }
    // Unchanged bits over qubits whose rows are all the identity's: the
    // preimage is p itself, phase included (X_q Z_q multiply to the XZ form
    // with phase 0).
    bool same = true;
    for (size_t w = 0; w < out.x.size() && same; ++w)
      same = out.x[w] == p.x[w] && out.z[w] == p.z[w] && ((p.x[w] | p.z[w]) & touched[w]) == 0;
    if (same) return p;
    Pauli rec = preimage(out);
    if (!rec.same_body(p)) throw TicitError(4, "frame rows do not span the Pauli body");
    out.set_phase(p.phase - rec.phase);
    return out;
  }
  // p <- coordinates(p), without a copy when p comes out unchanged (no
  // moved row flips one of its bits and its support meets no moved row).
  void coordinates_into(Pauli& p) const {
    if (!moved_ready) index_moved();
    bool change = false;
    for (size_t w = 0; w < p.x.size() && !change; ++w) change = ((p.x[w] | p.z[w]) & touched[w]) != 0;
    for (size_t i = 0; i < moved.size() && !change; ++i) {
      const int r = moved[i];
      const bool cur = r < n ? p.zb(r) : p.xb(r - n);
      change = anticommutes_moved(p, i) != cur;
    }
    if (change) p = coordinates(p);
  }
  void swap_rows(int a, int b) {
    moved_ready = false;
    std::swap(rows[a], rows[b]);
  }
  void add_phase(int r, int d) { rows[r].set_phase(rows[r].phase + d); }
  void mul_rows(int dst, int lhs, int rhs, int extra) {
    moved_ready = false;
    Pauli o = pmul(rows[lhs], rows[rhs]);
    o.set_phase(o.phase + extra);
    rows[dst] = std::move(o);
  }
  // Left-multiplication gates U <- G U (src/frames.rs:613-951).
  void h(int q) { swap_rows(q, n + q); }
  void s(int q) { mul_rows(q, q, n + q, 3); }
  void sdg(int q) { mul_rows(q, q, n + q, 1); }
  void sqrt_x(int q) { swap_rows(q, n + q); mul_rows(q, q, n + q, 3); swap_rows(q, n + q); }
  void sqrt_x_dag(int q) { swap_rows(q, n + q); mul_rows(q, q, n + q, 1); swap_rows(q, n + q); }
  void sqrt_y(int q) { swap_rows(q, n + q); add_phase(n + q, 2); }
  void sqrt_y_dag(int q) { swap_rows(q, n + q); add_phase(q, 2); }
  void gx(int q) { add_phase(n + q, 2); }
  void gy(int q) { add_phase(q, 2); add_phase(n + q, 2); }
  void gz(int q) { add_phase(q, 2); }
  void cx(int c, int t) { mul_rows(c, c, t, 0); mul_rows(n + t, n + c, n + t, 0); }
  void cy(int c, int t) {
    int xc = c, zc = n + c, xt = t, zt = n + t;
    mul_rows(xc, xc, zc, 3);
    mul_rows(xc, xc, zt, 0);
    mul_rows(xt, zc, xt, 0);
    mul_rows(xc, xc, xt, 0);
    mul_rows(zt, zc, zt, 0);
  }
  void cz(int a, int b) { mul_rows(a, a, n + b, 0); mul_rows(b, n + a, b, 0); }
  void swap(int a, int b) { swap_rows(a, b); swap_rows(n + a, n + b); }
};

// ============================================================================
// Symbol context (src/symbolic.rs:203-331)
// ============================================================================

struct Context {
  int32_t next = 1;
  std::map<int32_t, double> bern;  // ascending condition order is a contract
  std::vector<CategoricalDist> cats;
  void bump(int32_t c) { next = std::max(next, c + 1); }
  int32_t fresh() { return next++; }
  static double check_p(double p) {
    if (!(p >= 0.0 && p <= 1.0)) throw TicitError(1, "probability must be between 0 and 1");
    return p;
  }
  int32_t fresh_bern(double p) {
    check_p(p);
    int32_t c = fresh();
    bern[c] = p;
    return c;
  }
  std::vector<int32_t> fresh_cat(int nbits, const std::vector<uint64_t>& assignments,
                                 const std::vector<double>& probs) {
    double total = 0.0;
    for (double p : probs) total += check_p(p);
    if (std::fabs(total - 1.0) > 1e-12)
      throw TicitError(1, "categorical symbolic distribution probabilities must sum to 1");
    CategoricalDist d;
    d.nbits = nbits;
    for (int i = 0; i < nbits; ++i) d.conds.push_back(fresh());
    d.assignments = assignments;
    d.probs = probs;
    cats.push_back(d);
    return d.conds;
  }
};

// ============================================================================
// Symbolic Pauli frame (src/frames.rs:71-232)
// ============================================================================

struct ActiveFrame {
  int n = 0;
  std::vector<int32_t> conds;
  // transposed occupancy: 64 terms per block, [block * n + q]
  std::vector<uint64_t> xblk, zblk;
  explicit ActiveFrame(int nq = 0) : n(nq) {}
  void add(const Pauli& p, int32_t c) {
    size_t t = conds.size();
    if ((t & 63) == 0) {
      xblk.resize(xblk.size() + n, 0);
      zblk.resize(zblk.size() + n, 0);
    }
    size_t base = (t >> 6) * n;
    uint64_t m = 1ull << (t & 63);
    for (int q = 0; q < n; ++q) {
      if (p.xb(q)) xblk[base + q] |= m;
      if (p.zb(q)) zblk[base + q] |= m;
    }
    conds.push_back(c);
  }
  SymBool conjugate(const Pauli& p) const {
    std::vector<int> xs, zs;
    for (int q = 0; q < n; ++q) {
      if (p.xb(q)) xs.push_back(q);
      if (p.zb(q)) zs.push_back(q);
    }
    std::vector<int32_t> hit;
    size_t blocks = (conds.size() + 63) / 64;
    for (size_t b = 0; b < blocks; ++b) {
      uint64_t anti = 0;
      size_t base = b * n;
      for (int q : xs) anti ^= zblk[base + q];
      for (int q : zs) anti ^= xblk[base + q];
      while (anti) {
        size_t t = b * 64 + __builtin_ctzll(anti);
        if (t < conds.size()) hit.push_back(conds[t]);
        anti &= anti - 1;
      }
    }
    SymBool out;
    normalize_conditions(hit);
    out.conds = std::move(hit);
    return out;
  }
};

// ============================================================================
// Circuit IR and parser (src/circuit/parser.rs, src/circuit/ir.rs)
// ============================================================================

enum class K {
  Tick, H, S, SDag, SqrtX, SqrtXDag, SqrtY, SqrtYDag, X, Y, Z, CX, CY, CZ, Swap,
  T, TDag, Rot, MZ, MX, MY, MRZ, MRX, MRY, RZ, RX, RY, MPP, XErr, YErr, ZErr, Dep1, Dep2,
  FeedbackX, FeedbackY, FeedbackZ, ExpVal,
  PauliChannel,    // PAULI_CHANNEL_1/2/3 and DEPOLARIZE3: `arity` targets per call, `probs`
  ProductChannel,  // a closed E / ELSE_CORRELATED_ERROR chain: absolute `probs` per product
  Herald,          // HERALDED_ERASE / HERALDED_PAULI_CHANNEL_1: probs (I, X, Y, Z) after a herald
  MPad,            // MPAD: mtargets hold the literal pad values
};

struct MTarget {
  int qubit = 0;
  bool inverted = false;
};

struct Product {
  Pauli pauli;
  bool inverted = false;
};

struct CInstr {
  K kind = K::Tick;
  int line = 0;
  std::vector<int> qubits;
  std::vector<MTarget> mtargets;
  std::vector<Product> products;
  double probability = 0.0;
  double angle = 0.0;
  int feedback_record = 0;  // classically controlled Pauli: record id
  int exp_base = 0;         // EXP_VAL: slot of the first product
  int arity = 1;            // PauliChannel: targets per channel call
  std::vector<double> probs;  // channel probabilities (see K)
};

struct CDetector {
  std::vector<int> records;
  size_t after_instruction = 0;
  bool discard = false;  // DISCARD: postselected (src/circuit/parser.rs:1361-1374)
};

struct CObservable {
  int index = 0;
  std::vector<int> records;
};

struct Circuit {
  int nqubits = 0;
  int nrecords = 0;
  int ticks = 0;  // TICK lines, REPEAT bodies expanded
  int nexpvals = 0;  // EXP_VAL slots, one per product (src/circuit/parser.rs:1477-1487)
  std::vector<CInstr> instrs;
  std::vector<CDetector> detectors;
  std::vector<CObservable> observables;
};

struct RawLine {
  std::string name;
  std::vector<double> args;
  std::vector<std::string> targets;
  int line = 0;
};

[[noreturn]] void parse_error(int line, const std::string& msg) {
  throw TicitError(2, "line " + std::to_string(line) + ": " + msg);
}

std::string upper(std::string s) {
  for (auto& c : s) c = char(std::toupper(static_cast<unsigned char>(c)));
  return s;
}

// Marker line that closes an open correlated-error chain (no instruction
// name can contain a space).
const char* const kFlushChain = "END OF BLOCK";

// Reads lines into a flat list, expanding REPEAT blocks.
void read_block(const std::vector<std::string>& lines, size_t& pos, int depth,
                std::vector<RawLine>& out) {
  while (pos < lines.size()) {
    int lineno = int(pos) + 1;
    std::string s = lines[pos++];
    size_t hash = s.find('#');
    if (hash != std::string::npos) s = s.substr(0, hash);
    size_t a = 0;
    while (a < s.size() && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
    size_t b = s.size();
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    s = s.substr(a, b - a);
    if (s.empty()) continue;
    if (s == "}") {
      if (depth == 0) parse_error(lineno, "unmatched '}'");
      return;
    }
    // mnemonic
    size_t i = 0;
    while (i < s.size() && (std::isalnum(static_cast<unsigned char>(s[i])) || s[i] == '_')) ++i;
    RawLine rl;
    rl.line = lineno;
    rl.name = upper(s.substr(0, i));
    if (rl.name.empty()) parse_error(lineno, "expected an instruction name");
    if (i < s.size() && s[i] == '[') {  // tag
      size_t close = s.find(']', i);
      if (close == std::string::npos) parse_error(lineno, "unterminated instruction tag");
      i = close + 1;
    }
    if (i < s.size() && s[i] == '(') {
      size_t close = s.find(')', i);
      if (close == std::string::npos) parse_error(lineno, "unterminated argument list");
      std::string inner = s.substr(i + 1, close - i - 1);
      std::stringstream ss(inner);
      std::string tok;
      while (std::getline(ss, tok, ',')) {
        char* end = nullptr;
        const char* cs = tok.c_str();
        while (*cs && std::isspace(static_cast<unsigned char>(*cs))) ++cs;
        if (!*cs) continue;
        double v = std::strtod(cs, &end);
        if (end == cs) parse_error(lineno, "malformed numeric argument");
        rl.args.push_back(v);
      }
      i = close + 1;
    }
    std::string rest = s.substr(i);
    if (rl.name == "REPEAT") {
      std::stringstream ss(rest);
      long count = -1;
      std::string brace;
      ss >> count >> brace;
      if (count < 0 || brace != "{") parse_error(lineno, "malformed REPEAT");
      size_t body_start = pos;
      std::vector<RawLine> body;
      read_block(lines, pos, depth + 1, body);
      (void)body_start;
      // a correlated-error chain never straddles a REPEAT block or one of
      // its repetitions (src/circuit/parser.rs:1590-1610)
      RawLine flush;
      flush.name = kFlushChain;
      flush.line = lineno;
      out.push_back(flush);
      for (long r = 0; r < count; ++r) {
        out.insert(out.end(), body.begin(), body.end());
        out.push_back(flush);
      }
      continue;
    }
    // MPP-style products may be written "X0 * Z1": glue '*' to neighbours.
    std::string glued;
    for (size_t c = 0; c < rest.size(); ++c) {
      if (rest[c] == '*') {
        while (!glued.empty() && glued.back() == ' ') glued.pop_back();
        glued.push_back('*');
        ++c;
        while (c < rest.size() && std::isspace(static_cast<unsigned char>(rest[c]))) ++c;
        --c;
      } else {
        glued.push_back(std::isspace(static_cast<unsigned char>(rest[c])) ? ' ' : rest[c]);
      }
    }
    std::stringstream ts(glued);
    std::string tok;
    while (ts >> tok) rl.targets.push_back(tok);
    out.push_back(std::move(rl));
  }
  if (depth != 0) parse_error(int(lines.size()), "unterminated REPEAT block");
}

bool parse_int(const std::string& s, long& v) {
  if (s.empty()) return false;
  char* end = nullptr;
  v = std::strtol(s.c_str(), &end, 10);
  return end && *end == '\0';
}

int qubit_target(const RawLine& rl, const std::string& t) {
  long v;
  if (!parse_int(t, v) || v < 0) parse_error(rl.line, "malformed qubit target '" + t + "'");
  return int(v);
}

Circuit parse_circuit(const std::string& text) {
  std::vector<std::string> lines;
  {
    std::stringstream ss(text);
    std::string l;
    while (std::getline(ss, l)) lines.push_back(l);
  }
  std::vector<RawLine> raw;
  size_t pos = 0;
  read_block(lines, pos, 0, raw);

  // Qubit count: one past the largest qubit index named anywhere.
  int maxq = -1;
  for (const auto& rl : raw) {
    if (rl.name == "DETECTOR" || rl.name == "OBSERVABLE_INCLUDE" || rl.name == "DISCARD" ||
        rl.name == "SHIFT_COORDS" || rl.name == "TICK" || rl.name == "MPAD" || rl.name == kFlushChain)
      continue;
    for (const auto& t : rl.targets) {
      std::string body = t;
      std::stringstream fs(body);
      std::string f;
      while (std::getline(fs, f, '*')) {
        if (!f.empty() && f[0] == '!') f = f.substr(1);
        if (!f.empty() && (f[0] == 'X' || f[0] == 'Y' || f[0] == 'Z' || f[0] == 'x' ||
                           f[0] == 'y' || f[0] == 'z'))
          f = f.substr(1);
        long v;
        if (parse_int(f, v) && v >= 0) maxq = std::max(maxq, int(v));
      }
    }
  }
  Circuit c;
  c.nqubits = maxq + 1;
  const int n = c.nqubits;

  auto need_no_args = [](const RawLine& rl) {
    if (!rl.args.empty()) parse_error(rl.line, rl.name + " takes no arguments");
  };
  auto one_prob = [](const RawLine& rl) {
    if (rl.args.size() != 1) parse_error(rl.line, rl.name + " expects one probability");
    double p = rl.args[0];
    if (!(p >= 0.0 && p <= 1.0)) parse_error(rl.line, "probability must be between 0 and 1");
    return p;
  };
  auto opt_prob = [](const RawLine& rl) {
    if (rl.args.empty()) return 0.0;
    if (rl.args.size() != 1) parse_error(rl.line, rl.name + " expects at most one probability");
    double p = rl.args[0];
    if (!(p >= 0.0 && p <= 1.0)) parse_error(rl.line, "probability must be between 0 and 1");
    return p;
  };
  auto qubits_of = [&](const RawLine& rl) {
    std::vector<int> qs;
    for (const auto& t : rl.targets) qs.push_back(qubit_target(rl, t));
    return qs;
  };
  auto records_of = [&](const RawLine& rl) {
    std::vector<int> recs;
    for (const auto& t : rl.targets) {
      if (t.size() < 6 || t.compare(0, 4, "rec[") != 0 || t.back() != ']')
        parse_error(rl.line, "malformed record target '" + t + "'");
      long v;
      if (!parse_int(t.substr(4, t.size() - 5), v) || v >= 0)
        parse_error(rl.line, "record targets must be negative lookbacks");
      long r = long(c.nrecords) + 1 + v;
      if (r <= 0) parse_error(rl.line, "record lookback exceeds the measurement record");
      recs.push_back(int(r));
    }
    return recs;
  };
  auto products_of = [&](const RawLine& rl) {
    std::vector<Product> prods;
    for (const auto& t : rl.targets) {
      Product pr;
      pr.pauli = Pauli(n);
      std::stringstream fs(t);
      std::string f;
      while (std::getline(fs, f, '*')) {
        if (f.empty()) parse_error(rl.line, "malformed Pauli product");
        if (f[0] == '!') { pr.inverted = !pr.inverted; f = f.substr(1); }
        if (f.empty()) parse_error(rl.line, "malformed Pauli product");
        char axis = char(std::toupper(static_cast<unsigned char>(f[0])));
        long q;
        if (!parse_int(f.substr(1), q) || q < 0) parse_error(rl.line, "malformed Pauli target");
        Pauli fac = axis == 'X' ? pX(n, int(q)) : axis == 'Y' ? pY(n, int(q)) : axis == 'Z' ? pZ(n, int(q)) : Pauli();
        if (fac.n == 0) parse_error(rl.line, "unknown Pauli axis");
        mul_into(pr.pauli, fac);
      }
      prods.push_back(pr);
    }
    return prods;
  };

  static const std::map<std::string, K> k1 = {
      {"H", K::H}, {"H_XZ", K::H}, {"S", K::S}, {"SQRT_Z", K::S}, {"S_DAG", K::SDag},
      {"SQRT_Z_DAG", K::SDag}, {"SQRT_X", K::SqrtX}, {"SQRT_X_DAG", K::SqrtXDag},
      {"SQRT_Y", K::SqrtY}, {"SQRT_Y_DAG", K::SqrtYDag}, {"X", K::X}, {"Y", K::Y}, {"Z", K::Z},
  };
  static const std::map<std::string, K> k2 = {
      {"CX", K::CX}, {"CNOT", K::CX}, {"ZCX", K::CX}, {"CY", K::CY}, {"ZCY", K::CY},
      {"CZ", K::CZ}, {"ZCZ", K::CZ}, {"SWAP", K::Swap},
  };

  // E / ELSE_CORRELATED_ERROR chains (src/circuit/parser.rs:987-1021): the
  // alternatives accumulate with absolute probabilities p_i * prod_{j<i}
  // (1 - p_j) and close into one channel at the next other instruction.
  std::vector<Product> chain_products;
  std::vector<double> chain_probs;
  double chain_remaining = 1.0;
  auto flush_chain = [&](int line) {
    chain_remaining = 1.0;
    if (chain_products.empty()) return;
    CInstr e;
    e.kind = K::ProductChannel;
    e.line = line;
    e.products = std::move(chain_products);
    e.probs = std::move(chain_probs);
    chain_products.clear();
    chain_probs.clear();
    c.instrs.push_back(std::move(e));
  };
  auto required_prob = [](const RawLine& rl) {
    if (rl.args.size() != 1) parse_error(rl.line, "operation expects one probability argument");
    double p = rl.args[0];
    if (!(p >= 0.0 && p <= 1.0)) parse_error(rl.line, "probability must be between 0 and 1");
    return p;
  };
  auto check_prob_list = [](const RawLine& rl, const std::string& msg) {
    double total = 0.0;
    for (double p : rl.args) {
      if (!(p >= 0.0 && p <= 1.0)) parse_error(rl.line, "probability must be between 0 and 1");
      total += p;
    }
    if (total > 1.0 + 1e-12) parse_error(rl.line, msg);
  };

  for (const auto& rl : raw) {
    const std::string& op = rl.name;
    CInstr ci;
    ci.line = rl.line;
    if (op == "E" || op == "CORRELATED_ERROR" || op == "ELSE_CORRELATED_ERROR") {
      if (op != "ELSE_CORRELATED_ERROR") {
        flush_chain(rl.line);
      } else if (chain_products.empty()) {
        parse_error(rl.line, "ELSE_CORRELATED_ERROR must follow CORRELATED_ERROR or ELSE_CORRELATED_ERROR");
      }
      const double p = required_prob(rl);
      // the implicit product of the targets, one Pauli factor each
      Product pr;
      pr.pauli = Pauli(n);
      for (const auto& t : rl.targets) {
        if (t.find('*') != std::string::npos)
          parse_error(rl.line, "implicit Pauli products do not use combiner targets");
        RawLine one = rl;
        one.targets = {t};
        Product f = products_of(one)[0];
        mul_into(pr.pauli, f.pauli);
        pr.inverted = pr.inverted != f.inverted;
      }
      chain_products.push_back(pr);
      chain_probs.push_back(chain_remaining * p);
      chain_remaining *= 1.0 - p;
      continue;
    }
    flush_chain(rl.line);
    if (op == kFlushChain) continue;
    if (op == "TICK") {
      ++c.ticks;
      continue;
    }
    if (op == "QUBIT_COORDS" || op == "SHIFT_COORDS" || op == "I") {
      continue;
    } else if (k1.count(op)) {
      need_no_args(rl);
      ci.kind = k1.at(op);
      ci.qubits = qubits_of(rl);
    } else if (k2.count(op) && std::any_of(rl.targets.begin(), rl.targets.end(), [](const std::string& t) {
                 return t.compare(0, 4, "rec[") == 0;
               })) {
      // Classically controlled Paulis (src/circuit/parser.rs:1229-1305): one
      // instruction per (record, qubit) pair.
      need_no_args(rl);
      if (rl.targets.size() % 2) parse_error(rl.line, "controlled two-qubit gate requires paired targets");
      K kind = k2.at(op);
      for (size_t i = 0; i < rl.targets.size(); i += 2) {
        std::string a = rl.targets[i], b = rl.targets[i + 1];
        bool ra = a.compare(0, 4, "rec[") == 0, rb = b.compare(0, 4, "rec[") == 0;
        CInstr fi;
        fi.line = rl.line;
        if (kind == K::CZ && (ra || rb)) {
          if (!ra) std::swap(a, b);
          fi.kind = K::FeedbackZ;
        } else if ((kind == K::CX || kind == K::CY) && ra) {
          fi.kind = kind == K::CY ? K::FeedbackY : K::FeedbackX;
        } else if (ra || rb) {
          parse_error(rl.line, "unsupported classically controlled gate");
        } else {
          fi.kind = kind;
          fi.qubits = {qubit_target(rl, a), qubit_target(rl, b)};
          c.instrs.push_back(std::move(fi));
          continue;
        }
        RawLine one = rl;
        one.targets = {a};
        fi.feedback_record = records_of(one)[0];
        fi.qubits = {qubit_target(rl, b)};
        c.instrs.push_back(std::move(fi));
      }
      continue;
    } else if (k2.count(op)) {
      need_no_args(rl);
      ci.kind = k2.at(op);
      ci.qubits = qubits_of(rl);
      if (ci.qubits.size() % 2) parse_error(rl.line, "two-qubit Clifford requires paired targets");
      for (size_t i = 0; i < ci.qubits.size(); i += 2)
        if (ci.qubits[i] == ci.qubits[i + 1]) parse_error(rl.line, "two-qubit gate requires distinct qubits");
    } else if (op == "T" || op == "T_DAG") {
      need_no_args(rl);
      ci.kind = op == "T" ? K::T : K::TDag;
      ci.qubits = qubits_of(rl);
    } else if (op == "R_X" || op == "R_Y" || op == "R_Z") {
      if (rl.args.size() != 1) parse_error(rl.line, op + " expects one angle argument");
      ci.kind = K::Rot;
      ci.angle = rl.args[0] * M_PI / 2.0;  // half turns -> kernel angle
      for (int q : qubits_of(rl)) {
        Product pr;
        pr.pauli = op == "R_X" ? pX(n, q) : op == "R_Y" ? pY(n, q) : pZ(n, q);
        ci.products.push_back(pr);
      }
    } else if (op == "M" || op == "MZ" || op == "MX" || op == "MY" || op == "MR" || op == "MRZ" ||
               op == "MRX" || op == "MRY") {
      ci.probability = opt_prob(rl);
      ci.kind = (op == "MX") ? K::MX : (op == "MY") ? K::MY : (op == "M" || op == "MZ") ? K::MZ
              : (op == "MRX") ? K::MRX : (op == "MRY") ? K::MRY : K::MRZ;
      for (const auto& t : rl.targets) {
        MTarget mt;
        std::string s = t;
        if (!s.empty() && s[0] == '!') { mt.inverted = true; s = s.substr(1); }
        mt.qubit = qubit_target(rl, s);
        ci.mtargets.push_back(mt);
      }
      c.nrecords += int(ci.mtargets.size());
    } else if (op == "R" || op == "RZ" || op == "RX" || op == "RY") {
      need_no_args(rl);
      ci.kind = op == "RX" ? K::RX : op == "RY" ? K::RY : K::RZ;
      ci.qubits = qubits_of(rl);
    } else if (op == "MPP") {
      ci.kind = K::MPP;
      ci.probability = opt_prob(rl);
      ci.products = products_of(rl);
      c.nrecords += int(ci.products.size());
    } else if (op == "EXP_VAL") {
      // src/circuit/parser.rs:1477-1487: one slot per product, no records
      need_no_args(rl);
      ci.kind = K::ExpVal;
      ci.products = products_of(rl);
      if (ci.products.empty()) parse_error(rl.line, "EXP_VAL expects at least one Pauli product");
      ci.exp_base = c.nexpvals;
      c.nexpvals += int(ci.products.size());
    } else if (op == "X_ERROR" || op == "Y_ERROR" || op == "Z_ERROR") {
      ci.kind = op == "X_ERROR" ? K::XErr : op == "Y_ERROR" ? K::YErr : K::ZErr;
      ci.probability = one_prob(rl);
      ci.qubits = qubits_of(rl);
    } else if (op == "MXX" || op == "MYY" || op == "MZZ") {
      // src/circuit/parser.rs:1488-1497,741-790: consecutive target pairs,
      // the product's sign the XOR of the pair's '!' flags
      ci.kind = K::MPP;
      ci.probability = opt_prob(rl);
      if (rl.targets.size() % 2) parse_error(rl.line, "pair measurement requires paired targets");
      for (size_t i = 0; i < rl.targets.size(); i += 2) {
        std::string a = rl.targets[i], b = rl.targets[i + 1];
        bool ia = false, ib = false;
        if (!a.empty() && a[0] == '!') { ia = true; a = a.substr(1); }
        if (!b.empty() && b[0] == '!') { ib = true; b = b.substr(1); }
        const int qa = qubit_target(rl, a), qb = qubit_target(rl, b);
        if (qa == qb) parse_error(rl.line, "two-qubit Pauli operation requires distinct qubits");
        auto axis = [&](int q) { return op == "MXX" ? pX(n, q) : op == "MYY" ? pY(n, q) : pZ(n, q); };
        Product pr;
        pr.pauli = axis(qa);
        mul_into(pr.pauli, axis(qb));
        pr.inverted = ia != ib;
        ci.products.push_back(pr);
      }
      c.nrecords += int(ci.products.size());
    } else if (op == "SPP" || op == "SPP_DAG") {
      // src/circuit/parser.rs:1498-1502: kernel angle +-pi/4
      need_no_args(rl);
      ci.kind = K::Rot;
      ci.angle = op == "SPP" ? M_PI / 4.0 : -M_PI / 4.0;
      ci.products = products_of(rl);
    } else if (op == "R_PAULI") {
      if (rl.args.size() != 1) parse_error(rl.line, "R_PAULI expects one angle argument");
      ci.kind = K::Rot;
      ci.angle = rl.args[0] * M_PI / 2.0;  // half turns -> kernel angle
      ci.products = products_of(rl);
    } else if (op == "I_ERROR" || op == "II_ERROR") {
      // validated, then nothing happens (src/circuit/parser.rs:1509-1521)
      check_prob_list(rl, op + " probabilities must be disjoint and sum to at most 1");
      if (op == "II_ERROR" && qubits_of(rl).size() % 2) parse_error(rl.line, "II_ERROR requires paired targets");
      (void)qubits_of(rl);
      continue;
    } else if (op == "PAULI_CHANNEL_1" || op == "PAULI_CHANNEL_2" || op == "PAULI_CHANNEL_3") {
      // src/circuit/parser.rs:1529-1549: 4^n - 1 probabilities
      const int arity = op == "PAULI_CHANNEL_1" ? 1 : op == "PAULI_CHANNEL_2" ? 2 : 3;
      if (int(rl.args.size()) != (1 << (2 * arity)) - 1)
        parse_error(rl.line, "PAULI_CHANNEL probability count does not match gate arity");
      ci.kind = K::PauliChannel;
      ci.arity = arity;
      ci.probs = rl.args;
      ci.qubits = qubits_of(rl);
    } else if (op == "DEPOLARIZE3") {
      ci.kind = K::PauliChannel;
      ci.arity = 3;
      const double p = one_prob(rl);
      ci.probs.assign(63, p / 63.0);
      ci.qubits = qubits_of(rl);
    } else if (op == "HERALDED_ERASE" || op == "HERALDED_PAULI_CHANNEL_1") {
      // src/circuit/parser.rs:1551-1572; one herald record per target
      ci.kind = K::Herald;
      if (op == "HERALDED_ERASE") {
        const double p = required_prob(rl) / 4.0;
        ci.probs.assign(4, p);
      } else {
        if (rl.args.size() != 4) parse_error(rl.line, "HERALDED_PAULI_CHANNEL_1 expects four probabilities");
        for (double p : rl.args)
          if (!(p >= 0.0 && p <= 1.0)) parse_error(rl.line, "probability must be between 0 and 1");
        ci.probs = rl.args;
      }
      ci.qubits = qubits_of(rl);
      c.nrecords += int(ci.qubits.size());
    } else if (op == "MPAD") {
      ci.kind = K::MPad;
      ci.probability = opt_prob(rl);
      for (const auto& t : rl.targets) {
        MTarget mt;
        std::string s = t;
        if (!s.empty() && s[0] == '!') { mt.inverted = true; s = s.substr(1); }
        mt.qubit = qubit_target(rl, s);
        if (mt.qubit != 0 && mt.qubit != 1) parse_error(rl.line, "MPAD targets must be 0 or 1");
        ci.mtargets.push_back(mt);
      }
      c.nrecords += int(ci.mtargets.size());
    } else if (op == "DEPOLARIZE1") {
      ci.kind = K::Dep1;
      ci.probability = one_prob(rl);
      ci.qubits = qubits_of(rl);
    } else if (op == "DEPOLARIZE2") {
      ci.kind = K::Dep2;
      ci.probability = one_prob(rl);
      ci.qubits = qubits_of(rl);
      if (ci.qubits.size() % 2) parse_error(rl.line, "DEPOLARIZE2 requires paired targets");
    } else if (op == "DETECTOR" || op == "DISCARD") {
      CDetector d;
      d.discard = op == "DISCARD";
      d.records = records_of(rl);
      d.after_instruction = c.instrs.size();
      c.detectors.push_back(d);
      continue;
    } else if (op == "OBSERVABLE_INCLUDE") {
      if (rl.args.size() != 1 || rl.args[0] < 0 || rl.args[0] != std::floor(rl.args[0]))
        parse_error(rl.line, "OBSERVABLE_INCLUDE expects one nonnegative integer index");
      CObservable o;
      o.index = int(rl.args[0]);
      o.records = records_of(rl);
      c.observables.push_back(o);
      continue;
    } else {
      throw TicitError(3, "line " + std::to_string(rl.line) + ": unsupported instruction '" + op +
                              "' for the hip backend");
    }
    c.instrs.push_back(std::move(ci));
  }
  flush_chain(raw.empty() ? 0 : raw.back().line);
  return c;
}

// ============================================================================
// Lowering into a frame-factored state (src/circuit/lowering.rs)
// ============================================================================

struct PendingOp {
  int kind = 0;  // 0 rotation, 1 measurement, 2 classical record
  Pauli pauli;
  SymBool sign;
  double angle = 0.0;
  int32_t record = 0;
  int32_t record_condition = 0;
  int32_t exp_val = -1;  // expectation probe slot (kind 1), -1 = a measurement
  SymBool outcome;
  int32_t max_condition() const {
    if (kind == 0) return sign.max_condition();
    if (kind == 1) return std::max(sign.max_condition(), record_condition);
    return std::max(outcome.max_condition(), record_condition);
  }
};

struct Lowering {
  int n;
  Clifford cliff;
  ActiveFrame frame;
  Context ctx;
  std::vector<PendingOp> pending;
  std::vector<SymBool> records;  // record symbol per 1-based record

  explicit Lowering(int nq) : n(nq), cliff(nq), frame(nq) {}

  // src/factored.rs:507-514
  std::pair<Pauli, SymBool> prepare(const Pauli& p) {
    Pauli pre = cliff.preimage(p);
    SymBool s = frame.conjugate(pre);
    return {pre, s};
  }
  // src/factored.rs:517-536
  void apply_pauli(const Pauli& p, int32_t cond) {
    ctx.bump(cond);
    Pauli pre = cliff.preimage(p);
    if (pre.nonidentity()) {
      frame.add(pre, cond);
      ctx.bump(cond);
    }
  }
  // src/factored.rs:540-555
  void apply_pauli_symbolic(const Pauli& p, const SymBool& c) {
    if (c.constant) apply_pauli(p, ctx.fresh_bern(1.0));
    for (int32_t id : c.conds) apply_pauli(p, id);
  }
  void rotation(const Pauli& p, double angle) {
    PendingOp op;
    op.kind = 0;
    op.angle = angle;
    auto pr = prepare(p);
    op.pauli = pr.first;
    op.sign = pr.second;
    pending.push_back(std::move(op));
  }
  void measurement(const Pauli& p, const SymBool& sign, int32_t record, int32_t rc) {
    auto pr = prepare(p);
    PendingOp op;
    op.kind = 1;
    op.pauli = pr.first;
    op.sign = sym_xor(pr.second, sign);
    op.record = record;
    op.record_condition = rc;
    pending.push_back(std::move(op));
  }
  // src/factored.rs:604-622: a probe queues a record-less measurement
  void expectation(const Pauli& p, int32_t slot) {
    auto pr = prepare(p);
    PendingOp op;
    op.kind = 1;
    op.pauli = pr.first;
    op.sign = pr.second;
    op.exp_val = slot;
    pending.push_back(std::move(op));
  }
  // src/factored.rs:624-640: a record whose value is a known expression
  void classical_record(const SymBool& outcome, int32_t record) {
    ctx.bump(outcome.max_condition());
    PendingOp op;
    op.kind = 2;
    op.outcome = outcome;
    op.record = record;
    pending.push_back(std::move(op));
  }
  // src/circuit/lowering.rs:120-128: axis code 1 = X, 2 = Y, 3 = Z
  static void axis_bits(int code, bool& x, bool& z) {
    x = code == 1 || code == 2;
    z = code == 2 || code == 3;
  }
  // PAULI_CHANNEL_n on `qs` (src/circuit/lowering.rs:195-242): a 4^n-way
  // categorical over (x, z) bit pairs, target i's bits at 2i, 2i + 1, its
  // axis code read from the high end of the case number.
  void pauli_channel(const std::vector<int>& qs, const std::vector<double>& probs) {
    for (size_t i = 0; i < qs.size(); ++i)
      for (size_t j = i + 1; j < qs.size(); ++j)
        if (qs[i] == qs[j]) throw TicitError(1, "Pauli channel targets must be distinct");
    const int nq = int(qs.size());
    const size_t cases = size_t(1) << (2 * nq);
    if (probs.size() != cases - 1) throw TicitError(1, "Pauli channel probability count does not match target arity");
    double total = 0.0;
    for (double p : probs) total += Context::check_p(p);
    if (total > 1.0 + 1e-12) throw TicitError(1, "Pauli channel probabilities sum to more than 1");
    std::vector<uint64_t> assign{0};
    std::vector<double> dist{std::max(0.0, 1.0 - total)};
    for (size_t code = 1; code < cases; ++code) {
      uint64_t a = 0;
      size_t rem = code;
      for (int pos = nq - 1; pos >= 0; --pos) {
        bool x, z;
        axis_bits(int(rem & 3), x, z);
        rem >>= 2;
        if (x) a |= uint64_t(1) << (2 * pos);
        if (z) a |= uint64_t(1) << (2 * pos + 1);
      }
      assign.push_back(a);
      dist.push_back(probs[code - 1]);
    }
    auto bits = ctx.fresh_cat(2 * nq, assign, dist);
    for (int i = 0; i < nq; ++i) {
      apply_pauli_symbolic(pX(n, qs[size_t(i)]), sym_var(bits[size_t(2 * i)]));
      apply_pauli_symbolic(pZ(n, qs[size_t(i)]), sym_var(bits[size_t(2 * i + 1)]));
    }
  }
  SymBool readout_sign(bool inverted, double p) {
    SymBool s;
    s.constant = inverted;
    if (p != 0.0) s = sym_xor(s, sym_var(ctx.fresh_bern(p)));
    return s;
  }
  std::pair<int32_t, int32_t> reserve_record() {
    int32_t c = ctx.fresh();
    records.push_back(sym_var(c));
    return {int32_t(records.size()), c};
  }

  void apply(const CInstr& ci) {
    switch (ci.kind) {
      case K::Tick: return;
      case K::H: for (int q : ci.qubits) cliff.h(q); return;
      case K::S: for (int q : ci.qubits) cliff.s(q); return;
      case K::SDag: for (int q : ci.qubits) cliff.sdg(q); return;
      case K::SqrtX: for (int q : ci.qubits) cliff.sqrt_x(q); return;
      case K::SqrtXDag: for (int q : ci.qubits) cliff.sqrt_x_dag(q); return;
      case K::SqrtY: for (int q : ci.qubits) cliff.sqrt_y(q); return;
      case K::SqrtYDag: for (int q : ci.qubits) cliff.sqrt_y_dag(q); return;
      case K::X: for (int q : ci.qubits) cliff.gx(q); return;
      case K::Y: for (int q : ci.qubits) cliff.gy(q); return;
      case K::Z: for (int q : ci.qubits) cliff.gz(q); return;
      case K::CX: for (size_t i = 0; i < ci.qubits.size(); i += 2) cliff.cx(ci.qubits[i], ci.qubits[i + 1]); return;
      case K::CY: for (size_t i = 0; i < ci.qubits.size(); i += 2) cliff.cy(ci.qubits[i], ci.qubits[i + 1]); return;
      case K::CZ: for (size_t i = 0; i < ci.qubits.size(); i += 2) cliff.cz(ci.qubits[i], ci.qubits[i + 1]); return;
      case K::Swap: for (size_t i = 0; i < ci.qubits.size(); i += 2) cliff.swap(ci.qubits[i], ci.qubits[i + 1]); return;
      case K::T:
      case K::TDag: {
        // T = exp(-i pi/8 Z) up to global phase (src/circuit/lowering.rs:573-584)
        double a = ci.kind == K::T ? M_PI / 8.0 : -M_PI / 8.0;
        for (int q : ci.qubits) rotation(pZ(n, q), a);
        return;
      }
      case K::Rot:
        for (const auto& pr : ci.products) {
          Pauli p = pr.pauli;
          if (pr.inverted) p.set_phase(p.phase + 2);
          rotation(p, ci.angle);
        }
        return;
      case K::MZ:
      case K::MX:
      case K::MY:
        for (const auto& t : ci.mtargets) {
          SymBool sign = readout_sign(t.inverted, ci.probability);
          auto rr = reserve_record();
          Pauli p = ci.kind == K::MX ? pX(n, t.qubit) : ci.kind == K::MY ? pY(n, t.qubit) : pZ(n, t.qubit);
          measurement(p, sign, rr.first, rr.second);
        }
        return;
      case K::MRZ:
      case K::MRX:
      case K::MRY:
        // src/circuit/lowering.rs:400-422
        for (const auto& t : ci.mtargets) {
          Pauli m = ci.kind == K::MRX ? pX(n, t.qubit) : ci.kind == K::MRY ? pY(n, t.qubit) : pZ(n, t.qubit);
          Pauli corr = ci.kind == K::MRX ? pZ(n, t.qubit) : pX(n, t.qubit);
          SymBool sign = readout_sign(t.inverted, ci.probability);
          auto rr = reserve_record();
          measurement(m, sign, rr.first, rr.second);
          apply_pauli_symbolic(corr, sym_xor(sym_var(rr.second), sign));
        }
        return;
      case K::RZ:
      case K::RX:
      case K::RY:
        // src/circuit/lowering.rs:384-398
        for (int q : ci.qubits) {
          Pauli m = ci.kind == K::RX ? pX(n, q) : ci.kind == K::RY ? pY(n, q) : pZ(n, q);
          Pauli corr = ci.kind == K::RX ? pZ(n, q) : pX(n, q);
          int32_t cond = ctx.fresh();
          measurement(m, SymBool{}, 0, cond);
          apply_pauli(corr, cond);
        }
        return;
      case K::MPP:
        for (const auto& pr : ci.products) {
          SymBool sign = readout_sign(pr.inverted, ci.probability);
          auto rr = reserve_record();
          measurement(pr.pauli, sign, rr.first, rr.second);
        }
        return;
      case K::ExpVal:
        // src/circuit/lowering.rs:676-685; an inverted product is -P
        for (size_t i = 0; i < ci.products.size(); ++i) {
          Pauli p = ci.products[i].pauli;
          if (ci.products[i].inverted) p.set_phase(p.phase + 2);
          expectation(p, int32_t(ci.exp_base + int(i)));
        }
        return;
      case K::XErr:
      case K::YErr:
      case K::ZErr:
        for (int q : ci.qubits) {
          int32_t flip = ctx.fresh_bern(Context::check_p(ci.probability));
          Pauli p = ci.kind == K::XErr ? pX(n, q) : ci.kind == K::YErr ? pY(n, q) : pZ(n, q);
          apply_pauli_symbolic(p, sym_var(flip));
        }
        return;
      case K::FeedbackX:
      case K::FeedbackY:
      case K::FeedbackZ: {
        // src/circuit/lowering.rs:370-382
        int q = ci.qubits[0];
        Pauli p = ci.kind == K::FeedbackX ? pX(n, q) : ci.kind == K::FeedbackY ? pY(n, q) : pZ(n, q);
        if (ci.feedback_record <= 0 || ci.feedback_record > int(records.size()))
          throw TicitError(1, "measurement record target out of range");
        apply_pauli_symbolic(p, records[ci.feedback_record - 1]);
        return;
      }
      case K::PauliChannel: {
        if (ci.qubits.size() % size_t(ci.arity))
          throw TicitError(1, ci.arity == 2 ? "PAULI_CHANNEL_2 requires paired targets"
                                            : "Pauli channel requires whole groups of targets");
        for (size_t i = 0; i < ci.qubits.size(); i += size_t(ci.arity))
          pauli_channel(std::vector<int>(ci.qubits.begin() + long(i), ci.qubits.begin() + long(i) + ci.arity),
                        ci.probs);
        return;
      }
      case K::ProductChannel: {
        // src/circuit/lowering.rs:254-292: one alternative per product
        const size_t m = ci.products.size();
        double total = 0.0;
        for (double p : ci.probs) total += Context::check_p(p);
        if (total > 1.0 + 1e-12) throw TicitError(1, "Pauli product channel probabilities sum to more than 1");
        if (m > 64) throw TicitError(3, "correlated-error chains of more than 64 alternatives are not supported");
        std::vector<uint64_t> assign{0};
        std::vector<double> dist{std::max(0.0, 1.0 - total)};
        for (size_t i = 0; i < m; ++i) {
          assign.push_back(uint64_t(1) << i);
          dist.push_back(ci.probs[i]);
        }
        auto bits = ctx.fresh_cat(int(m), assign, dist);
        for (size_t i = 0; i < m; ++i) {
          Pauli pp = ci.products[i].pauli;
          if (ci.products[i].inverted) pp.set_phase(pp.phase + 2);
          apply_pauli_symbolic(pp, sym_var(bits[i]));
        }
        return;
      }
      case K::Herald: {
        // src/circuit/lowering.rs:294-340: (herald, x, z) rows 000, 100,
        // 110, 111, 101 = no event, herald + I, + X, + Y, + Z; the herald
        // bit is a record
        double total = 0.0;
        for (double p : ci.probs) total += Context::check_p(p);
        if (total > 1.0 + 1e-12) throw TicitError(1, "heralded channel probabilities sum to more than 1");
        for (int q : ci.qubits) {
          auto bits = ctx.fresh_cat(3, {0, 1, 3, 7, 5},
                                    {1.0 - total, ci.probs[0], ci.probs[1], ci.probs[2], ci.probs[3]});
          records.push_back(sym_var(bits[0]));
          classical_record(sym_var(bits[0]), int32_t(records.size()));
          apply_pauli_symbolic(pX(n, q), sym_var(bits[1]));
          apply_pauli_symbolic(pZ(n, q), sym_var(bits[2]));
        }
        return;
      }
      case K::MPad:
        // src/circuit/lowering.rs:343-366: the literal pad value, optionally
        // inverted and noisy
        for (const auto& t : ci.mtargets) {
          SymBool v;
          v.constant = (t.qubit != 0) != t.inverted;
          if (ci.probability != 0.0) v = sym_xor(v, sym_var(ctx.fresh_bern(Context::check_p(ci.probability))));
          records.push_back(v);
          classical_record(v, int32_t(records.size()));
        }
        return;
      case K::Dep1:
        // src/circuit/lowering.rs:143-159: bits (x, z), rows I, Z, X, Y
        for (int q : ci.qubits) {
          double p = Context::check_p(ci.probability);
          double third = p / 3.0;
          auto bits = ctx.fresh_cat(2, {0, 2, 1, 3}, {1.0 - p, third, third, third});
          apply_pauli_symbolic(pX(n, q), sym_var(bits[0]));
          apply_pauli_symbolic(pZ(n, q), sym_var(bits[1]));
        }
        return;
      case K::Dep2:
        // src/circuit/lowering.rs:161-193: bits (x1, z1, x2, z2)
        for (size_t i = 0; i < ci.qubits.size(); i += 2) {
          double p = Context::check_p(ci.probability);
          std::vector<uint64_t> assign{0};
          for (int x1 = 0; x1 < 2; ++x1)
            for (int z1 = 0; z1 < 2; ++z1)
              for (int x2 = 0; x2 < 2; ++x2)
                for (int z2 = 0; z2 < 2; ++z2)
                  if (x1 || z1 || x2 || z2)
                    assign.push_back(uint64_t(x1) | uint64_t(z1) << 1 | uint64_t(x2) << 2 | uint64_t(z2) << 3);
          std::vector<double> probs(16, p / 15.0);
          probs[0] = 1.0 - p;
          auto bits = ctx.fresh_cat(4, assign, probs);
          int q1 = ci.qubits[i], q2 = ci.qubits[i + 1];
          apply_pauli_symbolic(pX(n, q1), sym_var(bits[0]));
          apply_pauli_symbolic(pZ(n, q1), sym_var(bits[1]));
          apply_pauli_symbolic(pX(n, q2), sym_var(bits[2]));
          apply_pauli_symbolic(pZ(n, q2), sym_var(bits[3]));
        }
        return;
    }
  }
};

// ============================================================================
// Pending-queue peephole optimizer (src/pending_optimizer.rs)
// ============================================================================

struct Canon {
  Pauli body;
  SymBool sign;
};

Canon canonical_rotation(const PendingOp& r) {
  if (!hermitian(r.pauli)) throw TicitError(1, "pending Pauli rotation must have a Hermitian generator");
  Canon c;
  c.body = r.pauli;
  c.sign = sym_xor_const(r.sign, phase_sign(r.pauli));
  c.body.set_phase(c.body.ycount());
  return c;
}

bool rotation_can_cross(const Canon& rot, const PendingOp& op) {
  if (op.kind == 2) return true;
  return !anticommutes(rot.body, op.pauli);
}

void fuse_commuting_rotations(std::vector<PendingOp>& ops, CompileStats& st) {
  std::vector<char> deleted(ops.size(), 0);
  for (size_t i = 0; i < ops.size(); ++i) {
    if (deleted[i] || ops[i].kind != 0) continue;
    PendingOp earlier = ops[i];
    Canon a = canonical_rotation(earlier);
    for (size_t j = i + 1; j < ops.size(); ++j) {
      if (deleted[j]) continue;
      if (!rotation_can_cross(a, ops[j])) break;
      if (ops[j].kind != 0) continue;
      Canon b = canonical_rotation(ops[j]);
      if (!a.body.same_body(b.body) || a.sign.conds != b.sign.conds) continue;
      double dir = a.sign.constant == b.sign.constant ? 1.0 : -1.0;
      double angle = ops[j].angle + dir * earlier.angle;
      deleted[i] = 1;
      st.fused_rotations++;
      if (angle == 0.0) {
        deleted[j] = 1;
        st.cancelled_rotations++;
      } else {
        PendingOp fused;
        fused.kind = 0;
        fused.angle = angle;
        fused.pauli = b.body;
        fused.sign = b.sign;
        ops[j] = fused;
      }
      break;
    }
  }
  std::vector<PendingOp> kept;
  kept.reserve(ops.size());
  for (size_t i = 0; i < ops.size(); ++i)
    if (!deleted[i]) kept.push_back(std::move(ops[i]));
  ops = std::move(kept);
}

void move_measurements_earlier(std::vector<PendingOp>& ops, CompileStats& st, bool allow_all) {
  for (size_t i = 1; i < ops.size(); ++i) {
    if (ops[i].kind != 1) continue;
    Pauli body = ops[i].pauli;
    size_t target = i;
    for (size_t cursor = i; cursor >= 1; --cursor) {
      const PendingOp& prev = ops[cursor - 1];
      if (prev.kind != 0) break;
      if (anticommutes(body, prev.pauli)) break;
      if (allow_all) {
        target = cursor - 1;
        continue;
      }
      if (body.same_body(prev.pauli)) {
        target = cursor - 1;
        break;
      }
    }
    size_t cur = i;
    while (cur > target) {
      std::swap(ops[cur - 1], ops[cur]);
      --cur;
      st.measurement_left_swaps++;
    }
  }
}

// Returns prefix remap: new position of each preserved prefix.
std::vector<int> optimize_pending(std::vector<PendingOp>& pending, const std::vector<size_t>& preserved,
                                  CompileStats& st) {
  size_t count = pending.size();
  std::vector<int> remap(count + 1, -1);
  // src/pending_optimizer.rs:230-242: expectation probes pin the schedule
  if (std::any_of(pending.begin(), pending.end(), [](const PendingOp& op) { return op.exp_val >= 0; })) {
    for (size_t i = 0; i <= count; ++i) remap[i] = int(i);
    return remap;
  }
  remap[0] = 0;
  std::vector<size_t> ends;
  for (size_t p : preserved) {
    if (p > count) throw TicitError(4, "preserved pending-operation prefix is out of range");
    if (p > 0) ends.push_back(p);
  }
  ends.push_back(count);
  std::sort(ends.begin(), ends.end());
  ends.erase(std::unique(ends.begin(), ends.end()), ends.end());
  std::vector<PendingOp> out;
  out.reserve(count);
  size_t start = 0;
  for (size_t e : ends) {
    std::vector<PendingOp> seg(pending.begin() + start, pending.begin() + e);
    int before = st.fused_rotations;
    fuse_commuting_rotations(seg, st);
    move_measurements_earlier(seg, st, st.fused_rotations != before);
    for (auto& op : seg) out.push_back(std::move(op));
    remap[e] = int(out.size());
    start = e;
  }
  pending = std::move(out);
  return remap;
}

// ============================================================================
// Planner (src/planner.rs)
// ============================================================================

size_t word_cost(const SymBool& e) {
  size_t count = 0;
  int prev = 0;
  for (int32_t c : e.conds) {
    int w = word_of(c);
    if (count == 0 || w != prev) {
      ++count;
      prev = w;
    }
  }
  return count;
}

bool lower_cost(const SymBool& cand, const SymBool& cur) {
  size_t a = word_cost(cand), b = word_cost(cur);
  if (a != b) return a < b;
  return cand.conds.size() < cur.conds.size();
}

// src/planner.rs:185-230
SymBool reduce_once(const SymBool& expr, const SymBool& rel) {
  if (rel.conds.empty() && !rel.constant) return expr;
  size_t l = 0, r = 0, size = 0, words = 0;
  int prev = 0;
  bool have = false, overlap = false;
  auto add = [&](int32_t c) {
    ++size;
    int w = word_of(c);
    if (!have || w != prev) {
      ++words;
      prev = w;
      have = true;
    }
  };
  while (l < expr.conds.size() || r < rel.conds.size()) {
    if (r == rel.conds.size() || (l < expr.conds.size() && expr.conds[l] < rel.conds[r])) add(expr.conds[l++]);
    else if (l == expr.conds.size() || rel.conds[r] < expr.conds[l]) add(rel.conds[r++]);
    else { overlap = true; ++l; ++r; }
  }
  if (!overlap) return expr;
  size_t cur = word_cost(expr);
  if (words > cur || (words == cur && size >= expr.conds.size())) return expr;
  return sym_xor(expr, rel);
}

struct RelationReducer {
  std::vector<SymBool> rels;
  std::unordered_map<int32_t, std::vector<size_t>> index;
  std::unordered_map<int32_t, bool> fixed;
  void add(const SymBool& rel) {
    if (rel.conds.empty()) return;
    if (rel.conds.size() == 1) {
      fixed[rel.conds[0]] = rel.constant;
      return;
    }
    size_t i = rels.size();
    for (int32_t c : rel.conds) index[c].push_back(i);
    rels.push_back(rel);
  }
  SymBool reduce(SymBool e) const {
    for (;;) {
      bool changed = false;
      if (!fixed.empty()) {
        std::vector<int32_t> rem;
        rem.reserve(e.conds.size());
        for (int32_t c : e.conds) {
          auto it = fixed.find(c);
          if (it != fixed.end()) {
            e.constant ^= it->second;
            changed = true;
          } else {
            rem.push_back(c);
          }
        }
        e.conds = std::move(rem);
      }
      if (!e.conds.empty() && !rels.empty()) {
        std::vector<size_t> cand;
        for (int32_t c : e.conds) {
          auto it = index.find(c);
          if (it != index.end()) cand.insert(cand.end(), it->second.begin(), it->second.end());
        }
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        for (size_t r : cand) {
          SymBool red = reduce_once(e, rels[r]);
          if (red != e) {
            e = red;
            changed = true;
          }
        }
      }
      if (!changed) return e;
    }
  }
};

int32_t instr_max_condition(const Instr& in) {
  int32_t m = std::max(in.branch, in.expr.max_condition());
  return std::max(m, in.record_condition);
}

bool instr_has_sign(const Instr& in) { return in.op == Op::Rotate || in.op == Op::Promote; }

// src/planner.rs:1271-1289
void reduce_program_expressions(std::vector<Instr>& instrs) {
  RelationReducer red;
  for (auto& in : instrs) {
    in.expr = red.reduce(in.expr);
    if (!instr_has_sign(in) && in.record_condition != 0 && in.op != Op::Detector) {
      SymBool rel = sym_xor(sym_var(in.record_condition), in.expr);
      if (!rel.conds.empty() || rel.constant) red.add(rel);
    }
  }
}

struct Planner {
  int n;
  int k = 0;
  int max_k = 0;
  Context ctx;
  std::vector<PendingOp> queue;  // operations still to plan
  size_t head = 0;
  std::vector<Instr> instrs;
  std::vector<int> prefix_idx;
  std::unordered_map<int32_t, SymBool> subs;
  std::vector<SymBool> rels;
  std::vector<std::vector<int>> rel_words;
  std::unordered_map<int32_t, std::vector<size_t>> rel_index;
  int32_t next_record = 1;

  void set_k(int nk) {
    if (nk > n) throw TicitError(4, "active qubit count exceeds total qubit count");
    k = nk;
    max_k = std::max(max_k, nk);
  }

  void push(Instr in) {
    ctx.bump(instr_max_condition(in));
    instrs.push_back(std::move(in));
  }

  int32_t take_record(int32_t record, int32_t rc) {
    if (record == 0 && rc != 0) return 0;
    if (record == 0) return next_record++;
    next_record = std::max(next_record, record + 1);
    return record;
  }

  Pauli project_active(const Pauli& p) const {
    Pauli out(k);
    int y = 0;
    for (int q = 0; q < k; ++q) {
      bool xb = p.xb(q), zb = p.zb(q);
      out.setx(q, xb);
      out.setz(q, zb);
      if (xb && zb) ++y;
    }
    out.set_phase(y);
    return out;
  }

  Pauli embed_active(const Pauli& body) const {
    Pauli out(n);
    for (int q = 0; q < body.n; ++q) {
      out.setx(q, body.xb(q));
      out.setz(q, body.zb(q));
    }
    out.set_phase(body.phase);
    return out;
  }

  int highest_dormant_x(const Pauli& p) const {
    for (int q = n - 1; q >= k; --q)
      if (p.xb(q)) return q - k;
    return -1;
  }

  void substitute(SymBool& e) const {
    if (e.conds.empty() || subs.empty()) return;
    std::vector<int32_t> expanded = e.conds;
    for (;;) {
      bool changed = false;
      std::vector<int32_t> next;
      next.reserve(expanded.size());
      for (int32_t c : expanded) {
        auto it = subs.find(c);
        if (it != subs.end()) {
          changed = true;
          e.constant ^= it->second.constant;
          next.insert(next.end(), it->second.conds.begin(), it->second.conds.end());
        } else {
          next.push_back(c);
        }
      }
      normalize_conditions(next);
      expanded = std::move(next);
      if (!changed) break;
    }
    e.conds = std::move(expanded);
  }

  // src/planner.rs:517-573
  void reduce_pending(SymBool& e) const {
    substitute(e);
    if (e.conds.empty() || rels.empty()) return;
    std::vector<size_t> cand;
    for (int32_t c : e.conds) {
      auto it = rel_index.find(c);
      if (it != rel_index.end()) cand.insert(cand.end(), it->second.begin(), it->second.end());
    }
    std::sort(cand.begin(), cand.end());
    std::vector<int> ewords;
    for (int32_t c : e.conds) {
      int w = word_of(c);
      if (ewords.empty() || ewords.back() != w) ewords.push_back(w);
    }
    size_t s = 0;
    while (s < cand.size()) {
      size_t t = s + 1;
      while (t < cand.size() && cand[t] == cand[s]) ++t;
      size_t ri = cand[s];
      size_t overlap = t - s;
      const SymBool& rel = rels[ri];
      bool can = 2 * overlap > rel.conds.size();
      if (!can) {
        size_t ew = 0, shared = 0, fresh = 0;
        for (int w : rel_words[ri]) {
          while (ew < ewords.size() && ewords[ew] < w) ++ew;
          if (ew < ewords.size() && ewords[ew] == w) ++shared;
          else ++fresh;
        }
        can = shared > fresh;
      }
      if (can) e = reduce_once(e, rel);
      s = t;
    }
  }

  // src/planner.rs:600-637
  void note_measurement_relation(int32_t rc, const SymBool& outcome) {
    if (rc == 0) return;
    SymBool red = outcome;
    substitute(red);
    SymBool rel = sym_xor(sym_var(rc), red);
    ctx.bump(rel.max_condition());
    bool self_ref = std::binary_search(red.conds.begin(), red.conds.end(), rc);
    if (!self_ref && (red.conds.size() <= 1 || lower_cost(red, sym_var(rc)))) {
      subs[rc] = red;
      return;
    }
    size_t ri = rels.size();
    std::vector<int> words;
    for (int32_t c : rel.conds) {
      rel_index[c].push_back(ri);
      int w = word_of(c);
      if (words.empty() || words.back() != w) words.push_back(w);
    }
    rels.push_back(rel);
    rel_words.push_back(words);
  }

  // Queued operations by the qubits their Pauli acts on. A frame or a
  // pushed Pauli can only change operations that share a qubit with it, so
  // only those are visited. Entries go stale when an operation's Pauli
  // moves (it is indexed again under its new support) or leaves the queue;
  // visits re-check the operation itself.
  std::vector<std::vector<uint32_t>> by_qubit;
  std::vector<uint32_t> seen;
  uint32_t visit = 0;
  std::vector<uint32_t> hits;

  void index_op(uint32_t i) {
    const Pauli& p = queue[i].pauli;
    for (size_t w = 0; w < p.x.size(); ++w)
      for (uint64_t b = p.x[w] | p.z[w]; b; b &= b - 1) by_qubit[w * 64 + size_t(__builtin_ctzll(b))].push_back(i);
  }
  void build_index() {
    by_qubit.assign(size_t(n), {});
    seen.assign(queue.size(), 0);
    for (size_t i = head; i < queue.size(); ++i)
      if (queue[i].kind != 2) index_op(uint32_t(i));
  }
  // Queued (not yet planned) operations indexed under a qubit of `mask`,
  // each once, in ascending queue order.
  const std::vector<uint32_t>& ops_on(const std::vector<uint64_t>& mask) {
    ++visit;
    hits.clear();
    for (size_t w = 0; w < mask.size(); ++w)
      for (uint64_t b = mask[w]; b; b &= b - 1) {
        auto& list = by_qubit[w * 64 + size_t(__builtin_ctzll(b))];
        size_t keep = 0;
        for (uint32_t i : list) {
          if (i < head) continue;  // planned: dropped from the index
          list[keep++] = i;
          if (seen[i] == visit) continue;
          seen[i] = visit;
          hits.push_back(i);
        }
        list.resize(keep);
      }
    std::sort(hits.begin(), hits.end());
    return hits;
  }

  void push_through(const Pauli& p, const SymBool& sign) {
    ctx.bump(sign.max_condition());
    std::vector<uint64_t> support(p.x.size());
    for (size_t w = 0; w < p.x.size(); ++w) support[w] = p.x[w] | p.z[w];
    for (uint32_t i : ops_on(support)) {
      PendingOp& op = queue[i];
      if (op.kind == 2) continue;
      if (anticommutes(p, op.pauli)) op.sign = sym_xor(op.sign, sign);
    }
  }

  void transform_queue(const Clifford& f) {
    if (!f.moved_ready) f.index_moved();
    const std::vector<uint32_t> cand = ops_on(f.footprint);
    for (uint32_t i : cand) {
      PendingOp& op = queue[i];
      if (op.kind == 2) continue;
      const Pauli before = op.pauli;
      f.coordinates_into(op.pauli);
      bool moved_support = false;
      for (size_t w = 0; w < before.x.size() && !moved_support; ++w)
        moved_support = ((op.pauli.x[w] | op.pauli.z[w]) & ~(before.x[w] | before.z[w])) != 0;
      if (moved_support) index_op(i);
    }
  }

  static Pauli fix_row(const Pauli& row, const Pauli& installed, const Pauli& stab) {
    if (anticommutes(row, installed)) {
      Pauli prod = pmul(row, stab);
      prod.set_phase(prod.ycount());
      return prod;
    }
    return row;
  }

  // src/planner.rs:697-751
  Clifford promotion_frame(const Pauli& rot, int picked) const {
    int old_k = k, picked_q = old_k + picked;
    Pauli stab = pZ(n, picked_q);
    Pauli px = positive_hermitian(rot);
    if (!anticommutes(px, stab))
      throw TicitError(4, "dormant rotation promotion requires an anti-commuting fixed stabilizer");
    std::vector<Pauli> xr, zr;
    for (int q = 0; q < old_k; ++q) {
      zr.push_back(fix_row(pZ(n, q), px, stab));
      xr.push_back(fix_row(pX(n, q), px, stab));
    }
    zr.push_back(stab);
    xr.push_back(px);
    for (int q = old_k; q < n; ++q) {
      if (q == picked_q) continue;
      zr.push_back(fix_row(pZ(n, q), px, stab));
      xr.push_back(fix_row(pX(n, q), px, stab));
    }
    return Clifford::from_rows(n, std::move(xr), std::move(zr));
  }

  // src/planner.rs:758-792
  Clifford replacement_frame(const Pauli& measured, int picked) const {
    int picked_q = k + picked;
    Pauli old_stab = pZ(n, picked_q);
    Pauli new_stab = positive_hermitian(measured);
    if (!anticommutes(new_stab, old_stab))
      throw TicitError(4, "dormant measurement replacement requires an anti-commuting fixed stabilizer");
    std::vector<Pauli> xr, zr;
    for (int q = 0; q < n; ++q) {
      if (q == picked_q) {
        xr.push_back(old_stab);
        zr.push_back(new_stab);
      } else {
        zr.push_back(fix_row(pZ(n, q), new_stab, old_stab));
        xr.push_back(fix_row(pX(n, q), new_stab, old_stab));
      }
    }
    return Clifford::from_rows(n, std::move(xr), std::move(zr));
  }

  // src/planner.rs:799-847
  Clifford measurement_frame(const Pauli& body, const Instr& kern) const {
    int pivot = kern.pivot;
    Pauli measured = embed_active(body);
    Pauli fixed_x = kern.diagonal ? pX(n, pivot) : pZ(n, pivot);
    std::vector<Pauli> xr, zr;
    for (int q = 0; q < k; ++q) {
      if (q == pivot) continue;
      Pauli zrow = pZ(n, q), xrow = pX(n, q);
      if (kern.diagonal) {
        if (body.zb(q)) xrow = pmul(xrow, pX(n, pivot));
      } else {
        if (body.xb(q)) zrow = pmul(zrow, pZ(n, pivot));
        if (body.zb(q)) xrow = pmul(xrow, pZ(n, pivot));
      }
      xr.push_back(xrow);
      zr.push_back(zrow);
    }
    xr.push_back(fixed_x);
    zr.push_back(measured);
    for (int q = k; q < n; ++q) {
      xr.push_back(pX(n, q));
      zr.push_back(pZ(n, q));
    }
    return Clifford::from_rows(n, std::move(xr), std::move(zr));
  }

  static void fill_action(Instr& in, const Pauli& body) {
    if (body.n >= 63) throw TicitError(3, "Pauli string has too many qubits for active-state basis indexing");
    if (!hermitian(body)) throw TicitError(4, "Pauli rotation requires P^2 == I");
    in.nq = body.n;
    in.xmask = body.x.empty() ? 0 : body.x[0];
    in.zmask = body.z.empty() ? 0 : body.z[0];
    in.phase = body.phase;
  }

  // src/active.rs:279-345
  static void fill_measure_kernel(Instr& in, const Pauli& body) {
    fill_action(in, body);
    if (in.nq == 0) throw TicitError(4, "cannot build an active measurement kernel for k == 0");
    if (in.xmask) {
      in.pivot = 63 - __builtin_clzll(in.xmask);
      in.diagonal = false;
    } else if (in.zmask) {
      in.pivot = 63 - __builtin_clzll(in.zmask);
      in.diagonal = true;
      if (in.phase == 2) in.diag_phase_bit = 1;
      else if (in.phase == 0) in.diag_phase_bit = 0;
      else throw TicitError(4, "diagonal active measurement Pauli must have real eigenvalues");
    } else {
      throw TicitError(4, "cannot build an active measurement kernel for identity Pauli");
    }
  }

  void process_rotation(PendingOp op) {
    ctx.bump(op.sign.max_condition());
    int picked = highest_dormant_x(op.pauli);
    if (picked < 0) {
      Pauli body = project_active(op.pauli);
      SymBool sign = sym_xor_const(op.sign, phase_sign(op.pauli));
      if (!body.nonidentity()) return;  // a global phase
      if (!hermitian(body)) throw TicitError(4, "active rotation Pauli must square to identity");
      Instr in;
      in.op = Op::Rotate;
      fill_action(in, body);
      in.angle = op.angle;
      in.expr = sign;
      push(std::move(in));
      return;
    }
    int old_k = k;
    Clifford f = promotion_frame(op.pauli, picked);
    op.pauli = f.coordinates(op.pauli);
    transform_queue(f);
    if (!op.pauli.same_body(pX(n, old_k)))
      throw TicitError(4, "dormant rotation tableau reduction did not expose promoted X");
    Instr in;
    in.op = Op::Promote;
    in.nq = old_k;
    in.angle = op.angle;
    in.expr = sym_xor_const(op.sign, phase_sign(op.pauli));
    push(std::move(in));
    set_k(old_k + 1);
  }

  void process_measurement(PendingOp op) {
    ctx.bump(std::max(op.sign.max_condition(), op.record_condition));
    Pauli body = project_active(op.pauli);
    int picked = highest_dormant_x(op.pauli);
    if (picked >= 0 && op.exp_val >= 0) {
      // src/planner.rs:1017-1028: an unbiased probe reads 0, no state change
      Instr in;
      in.op = Op::Coin;
      in.exp_val = op.exp_val;
      push(std::move(in));
      return;
    }
    if (picked >= 0) {
      int picked_q = k + picked;
      Clifford f = replacement_frame(op.pauli, picked);
      op.pauli = f.coordinates(op.pauli);
      transform_queue(f);
      if (!op.pauli.same_body(pZ(n, picked_q)))
        throw TicitError(4, "dormant measurement tableau reduction did not expose fixed Z");
      SymBool base = sym_xor_const(op.sign, phase_sign(op.pauli));
      int32_t branch = ctx.fresh();
      push_through(pX(n, picked_q), sym_var(branch));
      SymBool outcome = sym_xor(base, sym_var(branch));
      Instr in;
      in.op = Op::Coin;
      in.branch = branch;
      in.expr = outcome;
      in.record = take_record(op.record, op.record_condition);
      in.record_condition = op.record_condition;
      push(std::move(in));
      note_measurement_relation(op.record_condition, outcome);
      return;
    }
    SymBool base = sym_xor_const(op.sign, phase_sign(op.pauli));
    if (!body.nonidentity()) {
      Instr in;
      in.op = Op::Record;
      in.expr = base;
      in.exp_val = op.exp_val;
      // src/planner.rs:96-108: a probe takes no record slot
      in.record = op.exp_val >= 0 ? 0 : take_record(op.record, op.record_condition);
      in.record_condition = op.record_condition;
      push(std::move(in));
      note_measurement_relation(op.record_condition, base);
      return;
    }
    if (op.exp_val >= 0) {
      // src/planner.rs:1101-1124: non-destructive, no branch, no basis change
      if (!hermitian(body)) throw TicitError(4, "active expectation Pauli must square to identity");
      Instr in;
      in.op = Op::Measure;
      fill_measure_kernel(in, body);
      in.expr = base;
      in.exp_val = op.exp_val;
      push(std::move(in));
      return;
    }
    if (!hermitian(body)) throw TicitError(4, "active measurement Pauli must square to identity");
    Instr in;
    in.op = Op::Measure;
    fill_measure_kernel(in, body);
    Clifford f = measurement_frame(body, in);
    transform_queue(f);
    int32_t branch = ctx.fresh();
    push_through(pX(n, k - 1), sym_var(branch));
    SymBool outcome = sym_xor(base, sym_var(branch));
    in.branch = branch;
    in.expr = outcome;
    in.record = take_record(op.record, op.record_condition);
    in.record_condition = op.record_condition;
    push(std::move(in));
    note_measurement_relation(op.record_condition, outcome);
    set_k(k - 1);
  }

  void process_classical(PendingOp op) {
    ctx.bump(std::max(op.outcome.max_condition(), op.record_condition));
    Instr in;
    in.op = Op::Record;
    in.expr = op.outcome;
    in.record = take_record(op.record, op.record_condition);
    in.record_condition = op.record_condition;
    push(std::move(in));
  }

  void run() {
    while (head < queue.size()) {
      if (prefix_idx.empty()) prefix_idx.push_back(int(instrs.size()));
      PendingOp op = queue[head];
      ++head;  // the planned operation leaves the queue before its effects land
      if (op.kind == 2) reduce_pending(op.outcome);
      else reduce_pending(op.sign);
      if (op.kind == 0) process_rotation(std::move(op));
      else if (op.kind == 1) process_measurement(std::move(op));
      else process_classical(std::move(op));
      prefix_idx.push_back(int(instrs.size()));
    }
  }
};

// src/planner.rs:1308-1336
bool rare_info(const CategoricalDist& d, RareGroup& g) {
  size_t false_row = d.assignments.size();
  uint64_t mask = d.nbits >= 64 ? ~0ull : ((1ull << d.nbits) - 1);
  for (size_t r = 0; r < d.assignments.size(); ++r)
    if ((d.assignments[r] & mask) == 0) {
      false_row = r;
      break;
    }
  if (false_row == d.assignments.size()) return false;
  double ev = 1.0 - d.probs[false_row];
  if (!(ev < 0.02)) return false;
  g.event_p = ev;
  g.nbits = d.nbits;
  g.assignments = d.assignments;
  g.probs = d.probs;
  if (ev > 0.0) {
    double inv = 1.0 / ev;
    for (size_t r = 0; r < d.probs.size(); ++r) {
      if (r == false_row || d.probs[r] <= 0.0) continue;
      g.event_rows.push_back(int32_t(r));
      g.event_probs.push_back(d.probs[r] * inv);
    }
  }
  return true;
}

// src/planner.rs:1409-1487 (component planning is not run: the dense path is
// always selected).
Program freeze(int n, int initial_k, int max_k, std::vector<Instr> instrs, Context& ctx) {
  if (initial_k > n || max_k > n || initial_k > max_k)
    throw TicitError(4, "invalid factored instruction program dimensions");
  reduce_program_expressions(instrs);
  Program p;
  p.n = n;
  p.initial_k = initial_k;
  p.max_k = max_k;
  int32_t rec = 0, det = 0, ev = 0;
  for (auto& in : instrs) {
    ctx.bump(instr_max_condition(in));
    rec = std::max(rec, in.record);
    det = std::max(det, in.detector);
    if (in.exp_val >= 0) ev = std::max(ev, in.exp_val + 1);
  }
  p.instrs = std::move(instrs);
  p.nsymbols = std::max(0, ctx.next - 1);
  p.nrecords = rec;
  p.ndetectors = det;
  p.nexpvals = ev;
  const double kLow = 0.02;
  for (const auto& d : ctx.cats) {
    RareGroup g;
    if (rare_info(d, g)) {
      bool merged = false;
      for (auto& e : p.rare) {
        if (e.event_p == g.event_p && e.nbits == g.nbits && e.assignments == g.assignments &&
            e.probs == g.probs && e.event_rows == g.event_rows && e.event_probs == g.event_probs) {
          e.sets.push_back(d.conds);
          merged = true;
          break;
        }
      }
      if (!merged) {
        g.sets.push_back(d.conds);
        p.rare.push_back(std::move(g));
      }
    } else {
      p.categorical.push_back(d);
    }
  }
  for (const auto& kv : ctx.bern) {
    if (kv.second < kLow) {
      bool merged = false;
      for (auto& g : p.lowp)
        if (g.p == kv.second) {
          g.conds.push_back(kv.first);
          merged = true;
          break;
        }
      if (!merged) p.lowp.push_back(BernGroup{kv.second, {kv.first}});
    } else {
      p.bern_conds.push_back(kv.first);
      p.bern_probs.push_back(kv.second);
    }
  }
  // src/planner.rs:1482-1484: the component cost model decides the execution
  p.use_components = plan_components(p).selected;
  return p;
}

}  // namespace

Program compile_circuit_text(const std::string& text, CompileStats* stats_out,
                             const std::vector<uint8_t>& postselection_mask) {
  CompileStats st;
  Circuit c = parse_circuit(text);
  st.ncircuit_instructions = int(c.instrs.size());
  Lowering low(c.nqubits);
  std::vector<size_t> counts{0};
  for (const auto& ci : c.instrs) {
    low.apply(ci);
    counts.push_back(low.pending.size());
  }
  if (int(low.records.size()) != c.nrecords) throw TicitError(4, "circuit measurement record count mismatch");
  st.npending = int(low.pending.size());

  // Detector anchors are pending-operation prefixes (src/circuit/mod.rs:145-164).
  std::vector<size_t> det_prefix;
  for (const auto& d : c.detectors) {
    if (d.after_instruction >= counts.size()) throw TicitError(4, "detector source position is out of range");
    det_prefix.push_back(counts[d.after_instruction]);
  }
  for (const auto& op : low.pending) low.ctx.bump(op.max_condition());
  std::vector<int> remap = optimize_pending(low.pending, det_prefix, st);
  for (auto& pfx : det_prefix) {
    int r = remap[pfx];
    if (r < 0) throw TicitError(4, "detector pending-operation prefix was not preserved by optimization");
    pfx = size_t(r);
  }

  Planner pl;
  pl.n = c.nqubits;
  pl.ctx = low.ctx;
  pl.queue = std::move(low.pending);
  pl.build_index();
  pl.run();
  Program planned = freeze(pl.n, 0, pl.max_k, std::move(pl.instrs), pl.ctx);

  // Splice one detector per declaration at its planned checkpoint
  // (src/circuit/mod.rs:224-271), then refreeze.
  std::vector<Instr> instrs;
  if (!c.detectors.empty()) {
    std::vector<std::vector<Instr>> events(planned.instrs.size() + 1);
    for (size_t i = 0; i < c.detectors.size(); ++i) {
      Instr d;
      d.op = Op::Detector;
      d.detector = int32_t(i + 1);
      // src/circuit/mod.rs:236-243: DISCARD declarations union the mask
      d.postselect = c.detectors[i].discard ||
                     (i < postselection_mask.size() && postselection_mask[i] != 0);
      for (int r : c.detectors[i].records) {
        if (r <= 0 || r > int(low.records.size()))
          throw TicitError(1, "detector references an out-of-range measurement record");
        d.expr = sym_xor(d.expr, low.records[r - 1]);
        d.det_records.push_back(r);
      }
      size_t pfx = det_prefix[i];
      size_t checkpoint = 0;
      if (!(pfx == 0 && pl.prefix_idx.empty())) {
        if (pfx >= pl.prefix_idx.size()) throw TicitError(4, "detector pending-operation position is out of range");
        checkpoint = size_t(pl.prefix_idx[pfx]);
      }
      events[checkpoint].push_back(std::move(d));
    }
    for (auto& e : events[0]) instrs.push_back(std::move(e));
    for (size_t i = 0; i < planned.instrs.size(); ++i) {
      instrs.push_back(std::move(planned.instrs[i]));
      for (auto& e : events[i + 1]) instrs.push_back(std::move(e));
    }
    planned = freeze(pl.n, 0, planned.max_k, std::move(instrs), pl.ctx);
  }

  int nobs = 0;
  for (const auto& o : c.observables) nobs = std::max(nobs, o.index + 1);
  planned.observables.assign(nobs, {});
  for (const auto& o : c.observables)
    for (int r : o.records) planned.observables[o.index].push_back(r);
  planned.circuit_layers = c.ticks + 1;
  if (stats_out) *stats_out = st;
  return planned;
}

}  // namespace ticit
