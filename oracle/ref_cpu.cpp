// ORACLE — test infrastructure only. Nothing in the product links or calls
// this file; tests, the smoke check and bench.py's CPU-baseline leg load it
// as the checker.
//
// A CPU restatement of ticit's compiled sampler, consuming the plan blob of
// include/ticit_plan.h:
//   * prepared.rs sample_impl (src/sampler/prepared.rs:692-859): chunks of
//     `chunk_shots`, exogenous seed block_seed(0x7eed0000, seed, chunk),
//     blocks of `batch_size` shots with branch seed
//     block_seed(0x5eed1234, seed, chunk*blocks_per_chunk + block),
//     chunk-strided worker threads concatenated in worker order;
//   * exogenous presampling (src/sampler/exogenous.rs:150-641);
//   * the shot-major batch executor (src/sampler/batch/runtime.rs:293-382,
//     batch/active.rs:267-728, batch/symbols.rs:84-437);
//   * the scalar dense kernels (src/contiguous.rs:93-177, 1050-1113,
//     1149-1320, 1582-1733, 2013-2044) whose evaluation order is the
//     reference's contract (src/contiguous.rs:10-15). Build with
//     -ffp-contract=off so no product is fused.
//   * row assembly (src/sampler/prepared.rs:881-1028).
//
// An injected mode replaces every random draw with caller-supplied values:
// exogenous symbol bit-planes, one uniform per (measure instruction, shot)
// and one coin bit per (coin instruction, shot).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>
#include <chrono>

#include "../include/ticit_plan.h"

namespace {

thread_local std::string g_err;
thread_local bool g_fp32_storage = false;
int g_component_mode = -1;  // read by worker threads: not thread-local
bool g_exact_sums = false;  // Neumaier-compensated probability sums (accuracy studies)
// keep_records (src/sampler/prepared.rs:104-142): false samples counts only,
// emitting no rows (the reference's counts-only benchmark mode). Read by
// worker threads: not thread-local.
bool g_keep_records = true;
// Pinned branches (src/sampler/pinning.rs ForcedBranch): instruction index ->
// the XOR plan its branch takes instead of a draw.
struct Forced {
  bool constant = false;
  std::vector<int32_t> conds;
};
std::map<size_t, Forced> g_forced;
// SamplerOptions::expected_detectors / expected_observables
// (src/sampler/prepared.rs:111-116): nonzero inverts that column; empty
// inverts nothing. Read by worker threads: not thread-local.
std::vector<uint8_t> g_expected_det, g_expected_obs;

struct OracleError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// ---------------------------------------------------------------- program ----

struct Expr {
  bool constant = false;
  std::vector<int32_t> conds;
};

struct Instr {
  int op = 0;
  int nq = 0;
  uint64_t xmask = 0, zmask = 0;
  int phase = 0;
  double angle = 0;
  int pivot = 0;
  bool diagonal = false;
  int diag_phase_bit = 0;
  int32_t branch = 0, record = 0, record_condition = 0, detector = 0, exp_val = -1;
  bool postselect = false;  // a firing detector discards the shot (src/factored.rs:227)
  std::vector<int32_t> det_records;
  Expr expr;
};

struct Cat {
  int nbits = 0;
  std::vector<int32_t> conds;
  std::vector<uint64_t> assign;
  std::vector<double> probs;
};

struct Rare {
  double event_p = 0;
  int nbits = 0;
  std::vector<std::vector<int32_t>> sets;
  std::vector<uint64_t> assign;
  std::vector<double> probs;
  std::vector<int32_t> event_rows;
  std::vector<double> event_probs;
};

struct LowP {
  double p = 0;
  std::vector<int32_t> conds;
};

struct Program {
  int n = 0, initial_k = 0, max_k = 0, nsymbols = 0, nrecords = 0, ndetectors = 0, nexpvals = 0;
  bool use_components = false;  // FactoredInstructionProgram::use_active_components
  std::vector<Instr> instrs;
  std::vector<Cat> cats;
  std::vector<Rare> rare;
  std::vector<int32_t> bern_conds;
  std::vector<double> bern_probs;
  std::vector<LowP> lowp;
  std::vector<std::vector<int32_t>> observables;
};

double bitsd(uint64_t u) {
  double d;
  std::memcpy(&d, &u, 8);
  return d;
}

struct Rd {
  const uint64_t* w;
  size_t n, pos = 0;
  uint64_t next() {
    if (pos >= n) throw OracleError("plan blob is truncated");
    return w[pos++];
  }
  int32_t i32() { return int32_t(uint32_t(next())); }
  size_t cnt() {
    uint64_t c = next();
    if (c > n) throw OracleError("plan blob count is out of range");
    return size_t(c);
  }
};

Program decode(const uint64_t* w, size_t n) {
  Rd r{w, n};
  if (r.next() != TICIT_PLAN_MAGIC) throw OracleError("bad plan blob header");
  const uint64_t version = r.next();
  if (version != 1 && version != TICIT_PLAN_VERSION) throw OracleError("bad plan blob header");
  Program p;
  p.n = r.i32();
  p.initial_k = r.i32();
  p.max_k = r.i32();
  p.nsymbols = r.i32();
  p.nrecords = r.i32();
  p.ndetectors = r.i32();
  p.nexpvals = r.i32();
  if (version >= 2) p.use_components = (r.next() & TICIT_PLAN_FLAG_COMPONENTS) != 0;
  p.instrs.resize(r.cnt());
  for (auto& in : p.instrs) {
    uint64_t w0 = r.next();
    in.op = int(w0 & 0xff);
    in.diagonal = (w0 >> 8) & 1;
    in.diag_phase_bit = int((w0 >> 9) & 1);
    in.expr.constant = (w0 >> 10) & 1;
    in.postselect = (w0 >> 11) & 1;
    in.nq = int((w0 >> 16) & 0xffff);
    in.pivot = int((w0 >> 32) & 0xffff);
    in.phase = int((w0 >> 48) & 3);
    in.xmask = r.next();
    in.zmask = r.next();
    in.angle = bitsd(r.next());
    uint64_t w4 = r.next(), w5 = r.next();
    in.branch = int32_t(uint32_t(w4));
    in.record = int32_t(uint32_t(w4 >> 32));
    in.record_condition = int32_t(uint32_t(w5));
    in.detector = int32_t(uint32_t(w5 >> 32));
    in.exp_val = r.i32();
    in.expr.conds.resize(r.cnt());
    for (auto& c : in.expr.conds) c = r.i32();
    in.det_records.resize(r.cnt());
    for (auto& c : in.det_records) c = r.i32();
  }
  p.cats.resize(r.cnt());
  for (auto& d : p.cats) {
    d.nbits = r.i32();
    d.conds.resize(r.cnt());
    for (auto& c : d.conds) c = r.i32();
    size_t rows = r.cnt();
    for (size_t k = 0; k < rows; ++k) {
      d.assign.push_back(r.next());
      d.probs.push_back(bitsd(r.next()));
    }
  }
  p.rare.resize(r.cnt());
  for (auto& g : p.rare) {
    g.event_p = bitsd(r.next());
    g.nbits = r.i32();
    g.sets.resize(r.cnt());
    for (auto& s : g.sets) {
      s.resize(r.cnt());
      for (auto& c : s) c = r.i32();
    }
    size_t rows = r.cnt();
    for (size_t k = 0; k < rows; ++k) {
      g.assign.push_back(r.next());
      g.probs.push_back(bitsd(r.next()));
    }
    size_t ne = r.cnt();
    for (size_t k = 0; k < ne; ++k) {
      g.event_rows.push_back(r.i32());
      g.event_probs.push_back(bitsd(r.next()));
    }
  }
  size_t nb = r.cnt();
  for (size_t k = 0; k < nb; ++k) {
    p.bern_conds.push_back(r.i32());
    p.bern_probs.push_back(bitsd(r.next()));
  }
  p.lowp.resize(r.cnt());
  for (auto& g : p.lowp) {
    g.p = bitsd(r.next());
    g.conds.resize(r.cnt());
    for (auto& c : g.conds) c = r.i32();
  }
  p.observables.resize(r.cnt());
  for (auto& o : p.observables) {
    o.resize(r.cnt());
    for (auto& c : o) c = r.i32();
  }
  if (r.next() != TICIT_PLAN_MAGIC) throw OracleError("bad plan blob trailer");
  return p;
}

// ----------------------------------------------------------------- random ----
// src/random.rs:13-80

uint64_t next_u64(uint64_t& s) {
  s += 0x9e3779b97f4a7c15ull;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

double rand_float(uint64_t& s) { return double(next_u64(s) >> 11) * 0x1.0p-53; }

bool check_prob(double p) { return p >= 0.0 && p <= 1.0; }

bool sample_bernoulli(uint64_t& s, double p) {
  if (!check_prob(p)) throw OracleError("probability must be between 0 and 1");
  if (p <= 0.0) return false;
  if (p >= 1.0) return true;
  return rand_float(s) < p;
}

size_t sample_categorical_row(uint64_t& s, const std::vector<double>& probs) {
  double r = rand_float(s);
  double cum = 0.0;
  for (size_t i = 0; i < probs.size(); ++i) {
    cum += probs[i];
    if (r <= cum) return i;
  }
  return probs.size() - 1;
}

double gap_denominator(double p) {
  if (!(p > 0.0 && p < 1.0)) throw OracleError("geometric gap probability must be in (0, 1)");
  return std::log1p(-p);
}

double sample_gap(uint64_t& s, double denom) {
  double u = rand_float(s);
  if (u < 2.2250738585072014e-308) u = 2.2250738585072014e-308;  // f64::MIN_POSITIVE
  double gap = std::floor(std::log(u) / denom);
  if (!std::isfinite(gap) || gap >= 2147483647.0) return 2147483647.0;
  return gap;
}

uint64_t block_seed(uint64_t base, uint64_t seed, uint64_t block) {
  return base ^ (0x9e3779b97f4a7c15ull * (block + 1)) ^ (0xbf58476d1ce4e5b9ull * (seed + 1));
}

const uint64_t kExoBase = 0x7eed0000ull;
const uint64_t kBranchBase = 0x5eed1234ull;

uint64_t low_mask(int64_t nbits) {
  if (nbits <= 0) return 0;
  if (nbits >= 64) return ~0ull;
  return (1ull << nbits) - 1;
}

uint64_t live_mask(size_t shots, size_t word) { return low_mask(int64_t(shots) - int64_t(word << 6)); }

// ------------------------------------------------------------- exogenous ----
// Symbol-major table: vals[(c - 1) * words + w]. Every family writes a fully
// zeroed table; sparse/dense storage choices of the reference do not change
// any value, so one dense table serves all of them.

struct ExoTable {
  size_t shots = 0, words = 0;
  std::vector<uint64_t> vals;
  void toggle(int32_t c, size_t shot) { vals[size_t(c - 1) * words + (shot >> 6)] ^= 1ull << (shot & 63); }
  bool get(int32_t c, size_t shot) const { return (vals[size_t(c - 1) * words + (shot >> 6)] >> (shot & 63)) & 1; }
};

void or_low_probability_bits(uint64_t* row, size_t words, uint64_t& rng, double p, size_t shots) {
  if (p <= 0.0 || shots == 0) return;
  double denom = gap_denominator(p);
  size_t draw = 0;
  for (;;) {
    size_t gap = size_t(sample_gap(rng, denom));
    if (gap >= shots - draw) return;
    draw += gap;
    size_t w = draw >> 6;
    if (w >= words) return;
    row[w] |= 1ull << (draw & 63);
    draw += 1;
  }
}

// src/sampler/exogenous.rs:185-266
void biased_bits(uint64_t* row, size_t words, uint64_t& rng, double probability, size_t shots) {
  if (shots == 0 || words == 0 || probability <= 0.0) return;
  if (probability >= 1.0) {
    for (size_t w = 0; w < words; ++w) row[w] = live_mask(shots, w);
    return;
  }
  bool invert = probability > 0.5;
  double p = invert ? 1.0 - probability : probability;
  if (p <= 0.0) {
    for (size_t w = 0; w < words; ++w) row[w] = invert ? live_mask(shots, w) : 0;
    return;
  }
  if (p == 0.5) {
    for (size_t w = 0; w < words; ++w) row[w] = next_u64(rng) & live_mask(shots, w);
  } else if (p < 0.02) {
    or_low_probability_bits(row, words, rng, p, shots);
  } else {
    const double buckets = 256.0;
    uint64_t raw = uint64_t(p * buckets);
    uint64_t half = 128;
    uint64_t top = raw < half ? raw : half - 1;
    double pt = double(top) / buckets;
    for (size_t w = 0; w < words; ++w) {
      uint64_t alive = next_u64(rng);
      uint64_t result = 0;
      for (int bit = 6; bit >= 0; --bit) {
        uint64_t shoot = next_u64(rng);
        if ((top >> bit) & 1) result |= shoot & alive;
        alive &= ~shoot;
      }
      row[w] = result & live_mask(shots, w);
    }
    double left = p - pt;
    if (left > 0.0) or_low_probability_bits(row, words, rng, left / (1.0 - pt), shots);
  }
  if (invert)
    for (size_t w = 0; w < words; ++w) row[w] = ~row[w] & live_mask(shots, w);
}

// src/sampler/exogenous.rs:490-641
void presample(ExoTable& t, const Program& p, size_t shots, uint64_t seed) {
  t.shots = shots;
  t.words = (shots + 63) / 64;
  t.vals.assign(size_t(p.nsymbols) * t.words, 0);
  uint64_t rng = seed;
  for (const auto& d : p.cats) {
    for (size_t s = 0; s < shots; ++s) {
      size_t row = sample_categorical_row(rng, d.probs);
      for (size_t b = 0; b < d.conds.size(); ++b)
        if ((d.assign[row] >> b) & 1) t.toggle(d.conds[b], s);
    }
  }
  for (const auto& g : p.rare) {
    size_t nsets = g.sets.size();
    if (g.event_p <= 0.0 || nsets == 0) continue;
    int64_t total = int64_t(shots) * int64_t(nsets);
    double denom = gap_denominator(g.event_p);
    int64_t draw = 0;
    for (;;) {
      int64_t gap = int64_t(sample_gap(rng, denom));
      if (gap >= total - draw) break;
      draw += gap;
      size_t shot = size_t(draw / int64_t(nsets));
      size_t set = size_t(draw % int64_t(nsets));
      size_t row = size_t(g.event_rows[sample_categorical_row(rng, g.event_probs)]);
      const auto& conds = g.sets[set];
      for (size_t b = 0; b < conds.size(); ++b)
        if ((g.assign[row] >> b) & 1) t.toggle(conds[b], shot);
      draw += 1;
    }
  }
  for (size_t i = 0; i < p.bern_conds.size(); ++i) {
    double pr = p.bern_probs[i];
    if (!check_prob(pr)) throw OracleError("probability must be between 0 and 1");
    if (pr <= 0.0) continue;
    biased_bits(&t.vals[size_t(p.bern_conds[i] - 1) * t.words], t.words, rng, pr, shots);
  }
  for (const auto& g : p.lowp) {
    size_t nc = g.conds.size();
    if (g.p <= 0.0 || nc == 0) continue;
    int64_t total = int64_t(shots) * int64_t(nc);
    double denom = gap_denominator(g.p);
    int64_t draw = 0;
    for (;;) {
      int64_t gap = int64_t(sample_gap(rng, denom));
      if (gap >= total - draw) break;
      draw += gap;
      t.toggle(g.conds[size_t(draw % int64_t(nc))], size_t(draw / int64_t(nc)));
      draw += 1;
    }
  }
}

// ----------------------------------------------------------- dense kernels ----

const double kInvSqrt2 = 0.70710678118654752440;  // f64::consts::FRAC_1_SQRT_2

struct Cx {
  double re, im;
};

Cx cmul(Cx a, Cx b) { return Cx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

Cx phase_factor(int ph) {
  switch (ph & 3) {
    case 0: return Cx{1.0, 0.0};
    case 1: return Cx{0.0, 1.0};
    case 2: return Cx{-1.0, 0.0};
    default: return Cx{0.0, -1.0};
  }
}

inline size_t insert_zero_bit(size_t packed, int bit) {
  size_t low = (size_t(1) << bit) - 1;
  return (packed & low) | ((packed & ~low) << 1);
}

inline bool parity(uint64_t v) { return __builtin_popcountll(v) & 1; }

// Kernel constants derived exactly as src/active.rs:193-230 and 311-316 do.
struct RotK {
  uint64_t xmask, zmask;
  bool diagonal, uniform_imag, xz_odd;
  int pair_bit;
  size_t pair_count;
  double c, s;
  Cx minus_even;
};

RotK rot_kernel(const Instr& in) {
  RotK k{};
  k.xmask = in.xmask;
  k.zmask = in.zmask;
  k.c = std::cos(in.angle);
  k.s = std::sin(in.angle);
  k.minus_even = cmul(Cx{0.0, -k.s}, phase_factor(in.phase));
  k.diagonal = in.xmask == 0;
  k.xz_odd = parity(in.xmask & in.zmask);
  if (!k.diagonal) {
    k.uniform_imag = in.zmask == 0;
    k.pair_bit = 63 - __builtin_clzll(in.xmask);
    k.pair_count = (size_t(1) << in.nq) >> 1;
  }
  return k;
}

Cx coefficient(const RotK& k, size_t source, bool sign) {
  bool odd = parity(uint64_t(source) & k.zmask);
  if (sign != odd) return Cx{-k.minus_even.re, -k.minus_even.im};
  return k.minus_even;
}

// src/contiguous.rs:93-177 (scalar) and 1050-1086
void rotate(double* re, double* im, size_t dim, const RotK& k, bool sign) {
  double c = k.c;
  if (k.diagonal) {
    for (size_t b = 0; b < dim; ++b) {
      Cx co = coefficient(k, b, sign);
      double fr = c + co.re;
      double fi = co.im;
      double r = re[b], i = im[b];
      re[b] = fr * r - fi * i;
      im[b] = fr * i + fi * r;
    }
    return;
  }
  if (k.uniform_imag) {
    double q = coefficient(k, 0, sign).im;
    size_t sel = size_t(1) << k.pair_bit, step = sel << 1;
    for (size_t block = 0; block < dim; block += step) {
      for (size_t off = 0; off < sel; ++off) {
        size_t i0 = block + off, i1 = i0 ^ size_t(k.xmask);
        double r0 = re[i0], m0 = im[i0], r1 = re[i1], m1 = im[i1];
        re[i0] = c * r0 - q * m1;
        im[i0] = c * m0 + q * r1;
        re[i1] = c * r1 - q * m0;
        im[i1] = c * m1 + q * r0;
      }
    }
    return;
  }
  for (size_t idx = 0; idx < k.pair_count; ++idx) {
    size_t left = insert_zero_bit(idx, k.pair_bit), right = left ^ size_t(k.xmask);
    bool lodd = parity(uint64_t(left) & k.zmask);
    double ld = (sign != lodd) ? -1.0 : 1.0;
    double rd = k.xz_odd ? -ld : ld;
    double lre = ld * k.minus_even.re, lim = ld * k.minus_even.im;
    double rre = rd * k.minus_even.re, rim = rd * k.minus_even.im;
    double r0 = re[left], i0 = im[left], r1 = re[right], i1 = im[right];
    re[left] = c * r0 + rre * r1 - rim * i1;
    im[left] = c * i0 + rre * i1 + rim * r1;
    re[right] = c * r1 + lre * r0 - lim * i0;
    im[right] = c * i1 + lre * i0 + lim * r0;
  }
}

// src/contiguous.rs:1088-1113
void promote(double* re, double* im, size_t dim, double c, double q) {
  for (size_t i = 0; i < dim; ++i) {
    double r = re[i], m = im[i];
    re[i] = c * r;
    im[i] = c * m;
    re[dim + i] = -q * m;
    im[dim + i] = q * r;
  }
}

struct MeasK {
  uint64_t xmask, zmask, z_without_pivot;
  int pivot, diag_phase_bit;
  bool diagonal;
  size_t out_dim;
  Cx coef1_even;
};

MeasK meas_kernel(const Instr& in) {
  MeasK k{};
  k.xmask = in.xmask;
  k.zmask = in.zmask;
  k.pivot = in.pivot;
  k.diagonal = in.diagonal;
  k.diag_phase_bit = in.diag_phase_bit;
  k.out_dim = (size_t(1) << in.nq) >> 1;
  k.z_without_pivot = in.diagonal ? (in.zmask & ~(1ull << in.pivot)) : 0;
  Cx e = phase_factor(in.phase);
  k.coef1_even = Cx{e.re * kInvSqrt2, -e.im * kInvSqrt2};
  return k;
}

size_t diagonal_source(const MeasK& k, size_t packed, bool branch) {
  size_t wo = insert_zero_bit(packed, k.pivot);
  int par = k.z_without_pivot == 0 ? 0 : int(parity(uint64_t(wo) & k.z_without_pivot));
  int pv = k.diag_phase_bit ^ par ^ int(branch);
  return wo | (size_t(pv) << k.pivot);
}

double clamp01(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

// src/contiguous.rs:1149-1320 scalar paths, branch = true
// Neumaier's compensated summation (Neumaier 1974), used only to measure
// how far a plain sum is from the exact one.
struct CompSum {
  double s = 0.0, c = 0.0;
  void add(double x) {
    double t = s + x;
    c += std::fabs(s) >= std::fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  double value() const { return s + c; }
};

double prob_sum(const double* re, const double* im, const MeasK& k);

double prob_true(const double* re, const double* im, const MeasK& k) { return clamp01(prob_sum(re, im, k)); }

// FP32 storage emulation of the device's FP32 path: P(true) relative to the
// stored state's own norm (sum of |a|^2 over the 2 * out_dim amplitudes of
// the measured body, in index order).
double stored_norm(const double* re, const double* im, const MeasK& k) {
  double n = 0.0;
  for (size_t i = 0; i < 2 * k.out_dim; ++i) n += re[i] * re[i] + im[i] * im[i];
  return n;
}

double prob_sum(const double* re, const double* im, const MeasK& k) {
  if (g_exact_sums) {
    CompSum acc;
    for (size_t idx = 0; idx < k.out_dim; ++idx) {
      if (k.diagonal) {
        size_t s = diagonal_source(k, idx, true);
        acc.add(re[s] * re[s] + im[s] * im[s]);
      } else {
        size_t s0 = insert_zero_bit(idx, k.pivot), s1 = s0 ^ size_t(k.xmask);
        bool odd = parity(uint64_t(s0) & k.zmask);
        double dir = (true != odd) ? -1.0 : 1.0;
        double c1r = dir * k.coef1_even.re, c1i = dir * k.coef1_even.im;
        double ar = kInvSqrt2 * re[s0] + c1r * re[s1] - c1i * im[s1];
        double ai = kInvSqrt2 * im[s0] + c1r * im[s1] + c1i * re[s1];
        acc.add(ar * ar + ai * ai);
      }
    }
    return acc.value();
  }
  double p = 0.0;
  if (k.diagonal) {
    for (size_t idx = 0; idx < k.out_dim; ++idx) {
      size_t s = diagonal_source(k, idx, true);
      p += re[s] * re[s] + im[s] * im[s];
    }
    return p;
  }
  for (size_t idx = 0; idx < k.out_dim; ++idx) {
    size_t s0 = insert_zero_bit(idx, k.pivot), s1 = s0 ^ size_t(k.xmask);
    bool odd = parity(uint64_t(s0) & k.zmask);
    double dir = (true != odd) ? -1.0 : 1.0;
    double c1r = dir * k.coef1_even.re, c1i = dir * k.coef1_even.im;
    double ar = kInvSqrt2 * re[s0] + c1r * re[s1] - c1i * im[s1];
    double ai = kInvSqrt2 * im[s0] + c1r * im[s1] + c1i * re[s1];
    p += ar * ar + ai * ai;
  }
  return p;
}

// src/contiguous.rs:1617-1733, 2013-2044
void project(double* re, double* im, double* sre, double* sim, const MeasK& k, bool branch, double inv) {
  if (k.diagonal) {
    for (size_t idx = 0; idx < k.out_dim; ++idx) {
      size_t s = diagonal_source(k, idx, branch);
      re[idx] = re[s] * inv;
      im[idx] = im[s] * inv;
    }
    return;
  }
  for (size_t idx = 0; idx < k.out_dim; ++idx) {
    size_t s0 = insert_zero_bit(idx, k.pivot), s1 = s0 ^ size_t(k.xmask);
    bool odd = parity(uint64_t(s0) & k.zmask);
    double dir = (branch != odd) ? -1.0 : 1.0;
    double c1r = dir * k.coef1_even.re, c1i = dir * k.coef1_even.im;
    double ar = kInvSqrt2 * re[s0] + c1r * re[s1] - c1i * im[s1];
    double ai = kInvSqrt2 * im[s0] + c1r * im[s1] + c1i * re[s1];
    sre[idx] = ar * inv;
    sim[idx] = ai * inv;
  }
  std::memcpy(re, sre, k.out_dim * sizeof(double));
  std::memcpy(im, sim, k.out_dim * sizeof(double));
}

// ------------------------------------------------------ component plan ----
// src/component_plan.rs:277-522: simulate which active coordinates an
// instruction links, keep each connected group as its own small vector and
// merge groups only when an instruction spans them.

enum CompKind { kCompNone = 0, kCompIgnoredPhase = 1, kCompRotation = 2, kCompPromotion = 3, kCompMeasurement = 4 };

struct CompStep {
  int kind = kCompNone;
  int comp = 0;
  size_t merge_off = 0, merge_count = 0;
  Instr local;  // kernel in the component's local bit order
  bool deactivate = false;
};

struct CompPlan {
  bool selected = false;
  int initial_components = 0;
  int component_count = 0;
  std::vector<int> comp_max_k;
  std::vector<CompStep> steps;  // one per instruction, empty when not planned
  std::vector<int> merges;
  double dense_work = 0.0, comp_work = 0.0;
  size_t dense_peak = 0, peak_live = 0, allocated = 0;
  int quantum_ops = 0;
};

// src/component_plan.rs:36-42
constexpr double kCompDispatchWork = 32.0;
constexpr double kCompMinSaved = 8192.0;
constexpr double kCompWorkRatio = 1.8;
constexpr double kCompAllocRatio = 1.25;

CompPlan build_component_plan(const Program& p) {
  struct PComp {
    bool active;
    std::vector<int> coords;
  };
  CompPlan plan;
  plan.initial_components = p.initial_k;
  plan.dense_peak = size_t(1) << p.max_k;
  if (p.max_k < 8 || p.nexpvals != 0) return plan;
  plan.steps.assign(p.instrs.size(), CompStep{});
  std::vector<PComp> comps;
  std::vector<int> order;       // active position -> coordinate
  std::vector<int> coord_comp;  // coordinate -> component, -1 = none
  std::vector<int> coord_pos;   // coordinate -> active position, -1 = none
  for (int c = 0; c < p.initial_k; ++c) {
    comps.push_back(PComp{true, {c}});
    order.push_back(c);
    coord_comp.push_back(int(comps.size()) - 1);
    coord_pos.push_back(c);
    plan.comp_max_k.push_back(1);
  }
  auto live_dim = [&]() {
    size_t t = 0;
    for (const auto& c : comps)
      if (c.active && !c.coords.empty()) t += size_t(1) << c.coords.size();
    return t;
  };
  auto update_peak = [&]() { plan.peak_live = std::max(plan.peak_live, live_dim()); };
  auto refresh = [&]() {
    std::fill(coord_pos.begin(), coord_pos.end(), -1);
    for (size_t i = 0; i < order.size(); ++i) coord_pos[size_t(order[i])] = int(i);
  };
  auto touched = [&](const Instr& in) {
    std::vector<int> t;
    uint64_t support = in.xmask | in.zmask;
    for (size_t pos = 0; pos < order.size(); ++pos) {
      if (!((support >> pos) & 1)) continue;
      int c = coord_comp[size_t(order[pos])];
      if (c < 0) throw OracleError("component planner found an active coordinate without a component");
      if (std::find(t.begin(), t.end(), c) == t.end()) t.push_back(c);
    }
    // an identity action is pinned to the first live component
    if (t.empty() && !order.empty()) t.push_back(coord_comp[size_t(order[0])]);
    return t;
  };
  auto target_of = [&](const std::vector<int>& t) {
    int best = t[0];
    for (size_t i = 1; i < t.size(); ++i) {
      int cand = t[i];
      size_t bs = comps[size_t(best)].coords.size(), cs = comps[size_t(cand)].coords.size();
      bool take = bs != cs ? bs < cs : best > cand;
      if (take) best = cand;
    }
    return best;
  };
  auto ordered = [&](int target, std::vector<int> t) {
    std::sort(t.begin(), t.end());
    t.erase(std::unique(t.begin(), t.end()), t.end());
    t.erase(std::remove(t.begin(), t.end(), target), t.end());
    t.insert(t.begin(), target);
    return t;
  };
  auto merge = [&](const std::vector<int>& src) {
    double work = 0.0;
    int target = src[0];
    size_t tk = comps[size_t(target)].coords.size();
    for (size_t i = 1; i < src.size(); ++i) {
      PComp& s = comps[size_t(src[i])];
      if (!s.active) throw OracleError("component merge source is inactive");
      tk += s.coords.size();
      work += 2.0 * std::ldexp(1.0, int(tk));
      for (int c : s.coords) coord_comp[size_t(c)] = target;
      auto moved = std::move(s.coords);
      s.coords.clear();
      comps[size_t(target)].coords.insert(comps[size_t(target)].coords.end(), moved.begin(), moved.end());
      s.active = false;
    }
    plan.comp_max_k[size_t(target)] = std::max(plan.comp_max_k[size_t(target)], int(tk));
    return work;
  };
  auto remap = [&](const Instr& in, int comp) {
    Instr l = in;
    const auto& cc = comps[size_t(comp)].coords;
    l.nq = int(cc.size());
    l.xmask = l.zmask = 0;
    for (size_t lp = 0; lp < cc.size(); ++lp) {
      int gp = coord_pos[size_t(cc[lp])];
      if (gp < 0) throw OracleError("component planner tried to remap an inactive coordinate");
      if ((in.xmask >> gp) & 1) l.xmask |= 1ull << lp;
      if ((in.zmask >> gp) & 1) l.zmask |= 1ull << lp;
    }
    return l;
  };
  auto push_merges = [&](CompStep& st, const std::vector<int>& src) {
    st.merge_off = plan.merges.size();
    st.merge_count = src.size();
    plan.merges.insert(plan.merges.end(), src.begin(), src.end());
  };

  for (size_t i = 0; i < p.instrs.size(); ++i) {
    const Instr& in = p.instrs[i];
    const int gk = int(order.size());
    refresh();
    const double dense_dim = std::ldexp(1.0, gk);
    CompStep& st = plan.steps[i];
    if (in.op == TICIT_OP_ROTATE) {
      if (in.nq != gk) throw OracleError("component planner saw a rotation with the wrong active width");
      plan.quantum_ops++;
      plan.dense_work += dense_dim;
      auto t = touched(in);
      if (t.empty()) {
        st.kind = kCompIgnoredPhase;
        continue;
      }
      int target = target_of(t);
      auto src = ordered(target, t);
      plan.comp_work += merge(src);
      push_merges(st, src);
      st.kind = kCompRotation;
      st.comp = target;
      st.local = remap(in, target);
      plan.comp_work += std::ldexp(1.0, int(comps[size_t(target)].coords.size())) + kCompDispatchWork;
      update_peak();
    } else if (in.op == TICIT_OP_PROMOTE) {
      plan.quantum_ops++;
      plan.dense_work += 2.0 * dense_dim;
      int coord = int(coord_comp.size());
      int comp = int(comps.size());
      comps.push_back(PComp{true, {coord}});
      order.push_back(coord);
      coord_comp.push_back(comp);
      coord_pos.push_back(int(order.size()) - 1);
      plan.comp_max_k.push_back(1);
      st.kind = kCompPromotion;
      st.comp = comp;
      plan.comp_work += 2.0 + kCompDispatchWork;
      update_peak();
    } else if (in.op == TICIT_OP_MEASURE) {
      if (in.exp_val >= 0) continue;  // expectation probes stay dense
      if (in.nq != gk) throw OracleError("component planner saw a measurement with the wrong active width");
      plan.quantum_ops++;
      plan.dense_work += 2.0 * dense_dim;
      auto t = touched(in);
      if (t.empty()) throw OracleError("active measurement has no component support");
      int target = target_of(t);
      auto src = ordered(target, t);
      plan.comp_work += merge(src);
      update_peak();
      push_merges(st, src);
      refresh();
      Instr l = remap(in, target);
      if (in.pivot >= gk) throw OracleError("component planner saw an invalid measurement pivot");
      int pcoord = order[size_t(in.pivot)];
      auto& cc = comps[size_t(target)].coords;
      auto it = std::find(cc.begin(), cc.end(), pcoord);
      if (it == cc.end()) throw OracleError("component measurement pivot is not in its target component");
      int lpivot = int(it - cc.begin());
      // src/active.rs:290-339 with the explicit local pivot
      l.pivot = lpivot;
      if (l.xmask) {
        if (!((l.xmask >> lpivot) & 1)) throw OracleError("nondiagonal active measurement pivot must have an X component");
        l.diagonal = false;
      } else {
        if (!((l.zmask >> lpivot) & 1)) throw OracleError("diagonal active measurement pivot must have a Z component");
        l.diagonal = true;
      }
      plan.comp_work += 2.0 * std::ldexp(1.0, int(cc.size())) + kCompDispatchWork;
      cc.erase(it);
      order.erase(order.begin() + in.pivot);
      coord_comp[size_t(pcoord)] = -1;
      st.kind = kCompMeasurement;
      st.comp = target;
      st.local = l;
      st.deactivate = cc.empty();
      if (st.deactivate) comps[size_t(target)].active = false;
      update_peak();
    }
  }
  plan.component_count = int(comps.size());
  for (int mk : plan.comp_max_k) plan.allocated += size_t(1) << mk;
  // src/component_plan.rs:282-298 (negated comparisons reject NaN)
  bool sel = p.max_k >= 8 && plan.quantum_ops >= 4;
  if (sel && !(plan.dense_work >= kCompWorkRatio * plan.comp_work)) sel = false;
  if (sel && !(plan.dense_work - plan.comp_work >= kCompMinSaved)) sel = false;
  if (sel && !(double(plan.allocated) <= kCompAllocRatio * double(plan.dense_peak))) sel = false;
  plan.selected = sel;
  return plan;
}

// src/sampler/batch/mod.rs:597-680: target (x) source in place, walking
// the source basis downwards so no unread target entry is overwritten.
void merge_into(double* tre, double* tim, int tk, const double* sre, const double* sim, int sk) {
  size_t td = size_t(1) << tk, sd = size_t(1) << sk;
  for (size_t sb = sd; sb-- > 0;) {
    double sr = sre[sb], si = sim[sb];
    size_t base = sb << tk;
    for (size_t tb = 0; tb < td; ++tb) {
      double tr = tre[tb], ti = tim[tb];
      tre[base + tb] = tr * sr - ti * si;
      tim[base + tb] = tr * si + ti * sr;
    }
  }
}

// --------------------------------------------------------------- executor ----

// Where the random decisions of one block come from.
struct Source {
  bool injected = false;
  uint64_t rng = 0;  // block stream (reference mode)
  // injected mode
  const double* uniforms = nullptr;  // [measure_ordinal][shots_total]
  const uint64_t* coins = nullptr;   // [coin_ordinal][words_total]
  size_t shots_total = 0;
  size_t first_shot = 0;  // global index of the block's shot 0
};

struct Outputs {
  std::vector<uint8_t> meas, det, obs;
  std::vector<uint64_t> obs_ones;
  uint64_t shots = 0, accepted = 0, discarded = 0, logical = 0;
  std::vector<double> expv;         // [shot][nexpvals] expectation values
  std::vector<double> probs;        // optional per (measure ordinal, global shot)
  std::vector<double> final_re, final_im;  // optional final state of shot 0
  int final_k = 0;
};

struct Runner {
  const Program& p;
  bool bit_packed;
  int selected_observable;
  std::vector<size_t> measure_ordinal, coin_ordinal;
  std::vector<RotK> rot;
  std::vector<MeasK> meas;
  std::vector<char> exogenous;
  std::vector<double> re, im, sre, sim;
  bool fp32 = false;  // emulate FP32 amplitude storage (precision studies)
  // the reference shot (src/sampler/prepared.rs:355-405) runs without the
  // expected bits and without postselection
  bool apply_expected = true;
  bool postselect = true;
  // tensor-component execution (src/sampler/batch/runtime.rs:388-552)
  CompPlan cplan;
  bool comp = false;
  std::vector<size_t> comp_off;
  size_t comp_stride = 0;
  std::vector<RotK> crot;
  std::vector<MeasK> cmeas;
  explicit Runner(const Program& prog, bool packed, int sel, int component_mode = -1)
      : p(prog), bit_packed(packed), selected_observable(sel) {
    cplan = build_component_plan(p);
    comp = component_mode < 0 ? p.use_components : component_mode != 0;
    if (comp && cplan.steps.empty()) throw OracleError("program does not contain an executable active component plan");
    if (comp) {
      for (int mk : cplan.comp_max_k) {
        comp_off.push_back(comp_stride);
        comp_stride += size_t(1) << mk;
      }
      crot.resize(p.instrs.size());
      cmeas.resize(p.instrs.size());
      for (size_t i = 0; i < p.instrs.size(); ++i) {
        const CompStep& st = cplan.steps[i];
        if (st.kind == kCompRotation) crot[i] = rot_kernel(st.local);
        if (st.kind == kCompMeasurement) cmeas[i] = meas_kernel(st.local);
      }
    }
    exogenous.assign(size_t(p.nsymbols) + 1, 0);
    for (const auto& d : p.cats) for (int32_t c : d.conds) exogenous[c] = 1;
    for (const auto& g : p.rare) for (const auto& st : g.sets) for (int32_t c : st) exogenous[c] = 1;
    for (int32_t c : p.bern_conds) exogenous[c] = 1;
    for (const auto& g : p.lowp) for (int32_t c : g.conds) exogenous[c] = 1;
    size_t m = 0, c = 0;
    measure_ordinal.assign(p.instrs.size(), 0);
    coin_ordinal.assign(p.instrs.size(), 0);
    rot.resize(p.instrs.size());
    meas.resize(p.instrs.size());
    for (size_t i = 0; i < p.instrs.size(); ++i) {
      const Instr& in = p.instrs[i];
      if (in.op == TICIT_OP_MEASURE) {
        measure_ordinal[i] = m++;
        meas[i] = meas_kernel(in);
      }
      if (in.op == TICIT_OP_COIN) coin_ordinal[i] = c++;
      if (in.op == TICIT_OP_ROTATE) rot[i] = rot_kernel(in);
    }
  }

  void round_state(size_t base, size_t n) {
    if (!fp32) return;
    for (size_t a = 0; a < n; ++a) {
      re[base + a] = double(float(re[base + a]));
      im[base + a] = double(float(im[base + a]));
    }
  }

  // One block of `shots` shots whose chunk-local offset is `off` into `exo`.
  void run_block(const ExoTable& exo, size_t off, size_t shots, Source& src, Outputs& out, double* probs_out,
                 bool want_final) {
    const size_t words = (shots + 63) / 64;
    const size_t stride = comp ? comp_stride : size_t(1) << p.max_k;
    if (re.size() < stride * shots) {
      re.resize(stride * shots);
      im.resize(stride * shots);
      sre.resize(stride);
      sim.resize(stride);
    }
    std::vector<int> ck;  // component widths (identical for every shot)
    if (comp) {
      // src/sampler/batch/mod.rs:366-402: initial components hold |0>
      ck.assign(cplan.comp_max_k.size(), 0);
      for (int c = 0; c < cplan.initial_components; ++c) {
        ck[size_t(c)] = 1;
        for (size_t s = 0; s < shots; ++s) {
          size_t b = s * stride + comp_off[size_t(c)];
          re[b] = 1.0;
          im[b] = 0.0;
          re[b + 1] = 0.0;
          im[b + 1] = 0.0;
        }
      }
    } else {
      // src/sampler/batch/mod.rs:424-436: only the first 2^initial_k entries
      // are cleared; promotion fully writes every newly exposed half.
      for (size_t s = 0; s < shots; ++s) {
        size_t d0 = size_t(1) << p.initial_k;
        std::fill(re.begin() + s * stride, re.begin() + s * stride + d0, 0.0);
        std::fill(im.begin() + s * stride, im.begin() + s * stride + d0, 0.0);
        re[s * stride] = 1.0;
      }
    }
    // src/sampler/batch/mod.rs:597-680 for every shot of the block
    auto do_merges = [&](const CompStep& st, bool round_merged = true) {
      if (st.merge_count == 0) throw OracleError("active component instruction has an invalid merge range");
      int t = cplan.merges[st.merge_off];
      if (t != st.comp || ck[size_t(t)] == 0) throw OracleError("active component merge has an invalid target");
      for (size_t m = 1; m < st.merge_count; ++m) {
        int src = cplan.merges[st.merge_off + m];
        if (src == t || ck[size_t(src)] == 0) throw OracleError("active component merge has an invalid source");
        if (ck[size_t(t)] + ck[size_t(src)] > cplan.comp_max_k[size_t(t)])
          throw OracleError("active component merge target has insufficient storage");
        for (size_t s = 0; s < shots; ++s) {
          size_t bt = s * stride + comp_off[size_t(t)], bs = s * stride + comp_off[size_t(src)];
          merge_into(&re[bt], &im[bt], ck[size_t(t)], &re[bs], &im[bs], ck[size_t(src)]);
          if (round_merged) round_state(bt, size_t(1) << (ck[size_t(t)] + ck[size_t(src)]));
        }
        ck[size_t(t)] += ck[size_t(src)];
        ck[size_t(src)] = 0;
      }
    };
    int k = p.initial_k;
    int ndormant = p.n - p.initial_k;
    std::vector<uint64_t> sym(size_t(p.nsymbols) * words, 0);
    std::vector<char> assigned(size_t(p.nsymbols) + 1, 0);
    std::vector<uint64_t> rec(size_t(p.nrecords) * words, 0), det(size_t(p.ndetectors) * words, 0);
    // src/sampler/batch/postselect.rs: a fired postselected detector marks
    // its shot dead; dead shots leave no rows and no counts
    std::vector<uint64_t> dead(words, 0);
    std::vector<uint64_t> bits(words);
    // src/sampler/batch/runtime.rs:46-75: expectation slots per shot
    const size_t nexp = size_t(p.nexpvals);
    std::vector<double> ev(shots * nexp, 0.0);
    auto exp_slot = [&](int32_t slot) {
      if (slot < 0 || size_t(slot) >= nexp) throw OracleError("expectation value index is out of range");
      return size_t(slot);
    };
    // src/sampler/batch/runtime.rs:97-118: a block reads its slice of the
    // chunk-wide exogenous rows starting at chunk-local shot `off`.
    const size_t sh = off & 63, w0 = off >> 6;
    auto exo_word = [&](int32_t c, size_t w) {
      const uint64_t* row = &exo.vals[size_t(c - 1) * exo.words];
      size_t src = w0 + w;
      uint64_t v = src < exo.words ? row[src] >> sh : 0;
      if (sh && src + 1 < exo.words) v |= row[src + 1] << (64 - sh);
      return v & live_mask(shots, w);
    };

    // Blocks start on a word boundary of the chunk's rows unless the batch
    // is not a multiple of 64 shots; then a row slice is a plain word range.
    const bool aligned = sh == 0 && w0 + words <= exo.words;
    auto eval = [&](const Expr& e) {
      uint64_t* out = bits.data();
      for (size_t w = 0; w < words; ++w) out[w] = e.constant ? ~0ull : 0;
      for (int32_t c : e.conds) {
        if (c <= 0 || c > p.nsymbols) throw OracleError("symbolic condition exceeds batch executor symbol table");
        if (exogenous[c]) {
          if (aligned) {
            const uint64_t* row = &exo.vals[size_t(c - 1) * exo.words + w0];
            for (size_t w = 0; w < words; ++w) out[w] ^= row[w];
          } else {
            for (size_t w = 0; w < words; ++w) out[w] ^= exo_word(c, w);
          }
        } else {
          if (!assigned[c]) throw OracleError("symbolic condition expression has no concrete batch value");
          const uint64_t* row = &sym[size_t(c - 1) * words];
          for (size_t w = 0; w < words; ++w) out[w] ^= row[w];
        }
      }
      for (size_t w = 0; w < words; ++w) out[w] &= live_mask(shots, w);
    };
    auto assign = [&](int32_t c, const std::vector<uint64_t>& v) {
      if (c <= 0 || c > p.nsymbols) throw OracleError("symbolic condition exceeds batch executor symbol table");
      if (assigned[c]) {
        for (size_t w = 0; w < words; ++w)
          if ((sym[size_t(c - 1) * words + w] ^ v[w]) & live_mask(shots, w))
            throw OracleError("symbolic condition was assigned inconsistent concrete batch values");
        return;
      }
      assigned[c] = 1;
      for (size_t w = 0; w < words; ++w) sym[size_t(c - 1) * words + w] = v[w] & live_mask(shots, w);
    };
    auto write_record = [&](int32_t record, int32_t rc, const std::vector<uint64_t>& v) {
      if (record > 0) {
        if (record > p.nrecords) throw OracleError("measurement record id out of range");
        for (size_t w = 0; w < words; ++w) rec[size_t(record - 1) * words + w] = v[w] & live_mask(shots, w);
      }
      if (rc > 0) assign(rc, v);
    };
    auto bit = [](const std::vector<uint64_t>& v, size_t s) { return (v[s >> 6] >> (s & 63)) & 1; };

    for (size_t i = 0; i < p.instrs.size(); ++i) {
      const Instr& in = p.instrs[i];
      switch (in.op) {
        case TICIT_OP_ROTATE: {
          if (in.nq != k) throw OracleError("rotation kernel dimension does not match batch active state");
          if (comp) {
            const CompStep& st = cplan.steps[i];
            if (st.kind == kCompIgnoredPhase) break;
            eval(in.expr);
            do_merges(st);
            RotK rk = crot[i];
            if (fp32) {
              rk.c = double(float(rk.c));
              rk.minus_even = Cx{double(float(rk.minus_even.re)), double(float(rk.minus_even.im))};
            }
            size_t dim = size_t(1) << ck[size_t(st.comp)];
            for (size_t s = 0; s < shots; ++s) {
              size_t b = s * stride + comp_off[size_t(st.comp)];
              rotate(&re[b], &im[b], dim, rk, bit(bits, s));
              round_state(b, dim);
            }
            break;
          }
          eval(in.expr);
          size_t dim = size_t(1) << k;
          RotK rk = rot[i];
          if (fp32) {
            rk.c = double(float(rk.c));
            rk.minus_even = Cx{double(float(rk.minus_even.re)), double(float(rk.minus_even.im))};
          }
          for (size_t s = 0; s < shots; ++s) {
            rotate(&re[s * stride], &im[s * stride], dim, rk, bit(bits, s));
            round_state(s * stride, dim);
          }
          break;
        }
        case TICIT_OP_PROMOTE: {
          if (ndormant == 0) throw OracleError("cannot promote a dormant qubit when none remain");
          eval(in.expr);
          if (comp) {
            // src/sampler/batch/runtime.rs:511-550
            const CompStep& st = cplan.steps[i];
            if (ck[size_t(st.comp)] != 0 || cplan.comp_max_k[size_t(st.comp)] < 1)
              throw OracleError("active component promotion target is unavailable");
            double c = std::cos(in.angle), sn = std::sin(in.angle);
            if (fp32) {
              c = double(float(c));
              sn = double(float(sn));
            }
            ck[size_t(st.comp)] = 1;
            for (size_t s = 0; s < shots; ++s) {
              size_t b = s * stride + comp_off[size_t(st.comp)];
              re[b] = c;
              im[b] = 0.0;
              re[b + 1] = 0.0;
              im[b + 1] = bit(bits, s) ? sn : -sn;
            }
            ++k;
            --ndormant;
            break;
          }
          size_t dim = size_t(1) << k;
          double c = std::cos(in.angle), sn = std::sin(in.angle);
          if (fp32) {
            c = double(float(c));
            sn = double(float(sn));
          }
          for (size_t s = 0; s < shots; ++s) {
            promote(&re[s * stride], &im[s * stride], dim, c, bit(bits, s) ? sn : -sn);
            round_state(s * stride, 2 * dim);
          }
          ++k;
          --ndormant;
          break;
        }
        case TICIT_OP_RECORD:
          eval(in.expr);
          if (in.exp_val >= 0) {
            // src/sampler/batch/runtime.rs:46-62: a deterministic probe is +-1
            size_t e = exp_slot(in.exp_val);
            for (size_t s = 0; s < shots; ++s) ev[s * nexp + e] = bit(bits, s) ? -1.0 : 1.0;
            break;
          }
          write_record(in.record, in.record_condition, bits);
          break;
        case TICIT_OP_DETECTOR: {
          if (!in.det_records.empty()) {
            std::fill(bits.begin(), bits.end(), 0);
            for (int32_t r : in.det_records) {
              if (r <= 0 || r > p.nrecords) throw OracleError("detector references an out-of-range measurement record");
              for (size_t w = 0; w < words; ++w) bits[w] ^= rec[size_t(r - 1) * words + w];
            }
          } else {
            eval(in.expr);
          }
          if (in.detector <= 0 || in.detector > p.ndetectors) throw OracleError("detector id out of range");
          // src/sampler/prepared.rs:926-936 XORs the expected bit into the
          // row; src/sampler/batch/postselect.rs:631-660 discards a shot
          // whose value differs from it
          if (apply_expected && !g_expected_det.empty() && g_expected_det[size_t(in.detector - 1)])
            for (size_t w = 0; w < words; ++w) bits[w] = ~bits[w];
          for (size_t w = 0; w < words; ++w) det[size_t(in.detector - 1) * words + w] = bits[w] & live_mask(shots, w);
          if (in.postselect && postselect)
            for (size_t w = 0; w < words; ++w) dead[w] |= bits[w] & live_mask(shots, w);
          break;
        }
        case TICIT_OP_MEASURE: {
          if (k == 0) throw OracleError("cannot measure an active Pauli when k == 0");
          const auto forced = g_forced.find(i);
          if (in.nq != k) throw OracleError("measurement kernel dimension does not match batch active state");
          if (in.exp_val >= 0) {
            // src/sampler/batch/active.rs:818-850: +-(1 - 2 P(true)), no
            // draw, no collapse (probe programs always run dense)
            if (comp) throw OracleError("expectation probes need the dense layout");
            size_t e = exp_slot(in.exp_val);
            eval(in.expr);
            for (size_t s = 0; s < shots; ++s) {
              const double* r = &re[s * stride];
              const double* m = &im[s * stride];
              double pt = clamp01(prob_sum(r, m, meas[i]) / (fp32 ? stored_norm(r, m, meas[i]) : 1.0));
              if (probs_out) probs_out[measure_ordinal[i] * src.shots_total + src.first_shot + s] = pt;
              double sign = bit(bits, s) ? -1.0 : 1.0;
              ev[s * nexp + e] = sign * (1.0 - 2.0 * pt);
            }
            break;
          }
          size_t coff = 0;
          // FP32 storage emulation: a merged state measured straight away is
          // never stored, so probability and projection read the unrounded
          // products; only the projected half is rounded (as stored)
          bool round_late = false;
          if (comp) {
            const CompStep& st = cplan.steps[i];
            if (st.kind != kCompMeasurement) throw OracleError("component measurement step does not match its instruction");
            round_late = fp32 && st.merge_count > 1;
            do_merges(st, !round_late);
            coff = comp_off[size_t(st.comp)];
          }
          const MeasK& mk = comp ? cmeas[i] : meas[i];
          std::vector<uint64_t> br(words, 0);
          for (size_t s = 0; s < shots; ++s) {
            double* r = &re[s * stride + coff];
            double* m = &im[s * stride + coff];
            const double nrm = fp32 ? stored_norm(r, m, mk) : 1.0;
            double pt = clamp01(prob_sum(r, m, mk) / nrm);
            if (probs_out) probs_out[measure_ordinal[i] * src.shots_total + src.first_shot + s] = pt;
            bool b;
            if (forced != g_forced.end()) {
              // src/sampler/batch/active.rs:538-554: only a fair coin is pinned
              if (std::fabs(pt - 0.5) > 1e-9)
                throw OracleError("pinned measurement branch has probability " + std::to_string(pt) +
                                  ", not one half; pinning it would bias sampling");
              bool f = forced->second.constant;
              for (int32_t c : forced->second.conds) {
                if (c <= 0 || c > p.nsymbols || exogenous[c] || !assigned[c])
                  throw OracleError("pinned branch reads an unassigned branch symbol");
                f ^= bool((sym[size_t(c - 1) * words + (s >> 6)] >> (s & 63)) & 1);
              }
              b = f;
            } else if (src.injected) {
              double u = src.uniforms[measure_ordinal[i] * src.shots_total + src.first_shot + s];
              b = pt <= 0.0 ? false : (pt >= 1.0 ? true : u < pt);
            } else {
              b = sample_bernoulli(src.rng, pt);
            }
            if (b) br[s >> 6] |= 1ull << (s & 63);
            double ps = b ? pt : 1.0 - pt;
            if (ps <= 0.0) throw OracleError("sampled an impossible active measurement branch");
            project(r, m, sre.data(), sim.data(), mk, b, 1.0 / std::sqrt(ps * nrm));
            round_state(s * stride + coff, mk.out_dim);
          }
          if (comp) {
            const CompStep& st = cplan.steps[i];
            ck[size_t(st.comp)] -= 1;
            if (st.deactivate && ck[size_t(st.comp)] != 0)
              throw OracleError("active component measurement did not remove its last coordinate");
          }
          --k;
          ++ndormant;
          assign(in.branch, br);
          eval(in.expr);
          write_record(in.record, in.record_condition, bits);
          break;
        }
        case TICIT_OP_COIN: {
          const auto forced = g_forced.find(i);
          if (forced != g_forced.end() && in.exp_val < 0) {
            // a pinned dormant branch copies its plan, no draw
            // (src/sampler/batch/runtime.rs:353-367)
            Expr fe;
            fe.constant = forced->second.constant;
            fe.conds = forced->second.conds;
            eval(fe);
            std::vector<uint64_t> br(bits.begin(), bits.end());
            assign(in.branch, br);
            eval(in.expr);
            write_record(in.record, in.record_condition, bits);
            break;
          }
          if (in.exp_val >= 0) {
            // src/sampler/batch/runtime.rs:353-356: an unbiased probe is 0
            size_t e = exp_slot(in.exp_val);
            for (size_t s = 0; s < shots; ++s) ev[s * nexp + e] = 0.0;
            break;
          }
          std::vector<uint64_t> br(words, 0);
          if (src.injected) {
            size_t wt = (src.shots_total + 63) / 64;
            for (size_t s = 0; s < shots; ++s) {
              size_t g = src.first_shot + s;
              if ((src.coins[coin_ordinal[i] * wt + (g >> 6)] >> (g & 63)) & 1) br[s >> 6] |= 1ull << (s & 63);
            }
          } else {
            for (size_t w = 0; w < words; ++w) br[w] = next_u64(src.rng) & live_mask(shots, w);
          }
          assign(in.branch, br);
          eval(in.expr);
          write_record(in.record, in.record_condition, bits);
          break;
        }
        default:
          throw OracleError("unknown opcode");
      }
    }
    if (want_final) {
      if (comp) throw OracleError("the final state of a component-mode run is not assembled");
      out.final_k = k;
      size_t dim = size_t(1) << k;
      out.final_re.assign(re.begin(), re.begin() + dim);
      out.final_im.assign(im.begin(), im.begin() + dim);
    }
    // src/sampler/prepared.rs:881-1028
    size_t nobs = p.observables.size();
    std::vector<uint64_t> ow(nobs * words, 0);
    for (size_t o = 0; o < nobs; ++o)
      for (int32_t r : p.observables[o]) {
        if (r <= 0 || r > p.nrecords) throw OracleError("observable references an out-of-range measurement record");
        for (size_t w = 0; w < words; ++w) ow[o * words + w] ^= rec[size_t(r - 1) * words + w];
      }
    // src/sampler/prepared.rs:965-974: an expected observable of 1 flips
    // the column before counting
    if (apply_expected && !g_expected_obs.empty())
      for (size_t o = 0; o < nobs; ++o)
        if (g_expected_obs[o])
          for (size_t w = 0; w < words; ++w) ow[o * words + w] ^= live_mask(shots, w);
    std::vector<uint64_t> keep(words);
    uint64_t nkeep = 0;
    for (size_t w = 0; w < words; ++w) {
      keep[w] = live_mask(shots, w) & ~dead[w];
      nkeep += __builtin_popcountll(keep[w]);
    }
    out.accepted += nkeep;
    out.discarded += shots - nkeep;
    out.shots += shots;
    for (size_t o = 0; o < nobs; ++o) {
      uint64_t ones = 0;
      for (size_t w = 0; w < words; ++w) ones += __builtin_popcountll(ow[o * words + w] & keep[w]);
      out.obs_ones[o] += ones;
      if (int(o) == selected_observable) out.logical += ones;
    }
    if (!g_keep_records) return;
    append_rows(out.meas, rec, size_t(p.nrecords), words, shots, keep);
    append_rows(out.det, det, size_t(p.ndetectors), words, shots, keep);
    append_rows(out.obs, ow, nobs, words, shots, keep);
    for (size_t s = 0; s < shots; ++s)
      if ((keep[s >> 6] >> (s & 63)) & 1) out.expv.insert(out.expv.end(), ev.begin() + s * nexp, ev.begin() + (s + 1) * nexp);
  }

  // Rows of the shots whose keep bit is set, in shot order.
  void append_rows(std::vector<uint8_t>& out, const std::vector<uint64_t>& cols, size_t ncols, size_t words,
                   size_t rows, const std::vector<uint64_t>& keep) {
    size_t oc = bit_packed ? (ncols + 7) / 8 : ncols;
    for (size_t s = 0; s < rows; ++s) {
      if (!((keep[s >> 6] >> (s & 63)) & 1)) continue;
      size_t base = out.size();
      out.resize(base + oc, 0);
      uint8_t* row = &out[base];
      for (size_t c = 0; c < ncols; ++c) {
        bool b = (cols[c * words + (s >> 6)] >> (s & 63)) & 1;
        if (bit_packed) {
          if (b) row[c / 8] |= uint8_t(1u << (c % 8));
        } else {
          row[c] = uint8_t(b);
        }
      }
    }
  }
};

// src/sampler/prepared.rs:523-540
void check_expected(const Program& p) {
  if (!g_expected_det.empty() && g_expected_det.size() != size_t(p.ndetectors))
    throw OracleError("expected_detectors has length " + std::to_string(g_expected_det.size()) + ", expected " +
                      std::to_string(p.ndetectors));
  if (!g_expected_obs.empty() && g_expected_obs.size() != p.observables.size())
    throw OracleError("expected_observables has length " + std::to_string(g_expected_obs.size()) + ", expected " +
                      std::to_string(p.observables.size()));
}

size_t default_batch(int max_k) {
  if (max_k >= 62) throw OracleError("active qubit count is too large for machine basis indices");
  size_t dim = size_t(1) << max_k;
  size_t b = 32768 / dim;
  if (b < 1) b = 1;
  return b < 2048 ? b : 2048;
}

void copy_out(const Outputs& o, uint8_t* meas, uint8_t* det, uint8_t* obs, uint64_t* counts, uint64_t* ones,
              double* exp_vals) {
  if (exp_vals && !o.expv.empty()) std::memcpy(exp_vals, o.expv.data(), o.expv.size() * sizeof(double));
  if (meas && !o.meas.empty()) std::memcpy(meas, o.meas.data(), o.meas.size());
  if (det && !o.det.empty()) std::memcpy(det, o.det.data(), o.det.size());
  if (obs && !o.obs.empty()) std::memcpy(obs, o.obs.data(), o.obs.size());
  if (counts) {
    counts[0] = o.shots;
    counts[1] = o.discarded;
    counts[2] = o.accepted;
    counts[3] = o.logical;
  }
  if (ones)
    for (size_t i = 0; i < o.obs_ones.size(); ++i) ones[i] = o.obs_ones[i];
}

}  // namespace

extern "C" {

const char* oracle_last_error(void) { return g_err.c_str(); }

// Injected runs on this thread emulate FP32 amplitude storage when set.
void oracle_set_fp32_storage(int on) { g_fp32_storage = on != 0; }

// -1: execute as the blob's use_active_components flag says; 0: dense;
// 1: tensor components (the program must have a component plan).
void oracle_set_component_mode(int mode) { g_component_mode = mode < 0 ? -1 : (mode ? 1 : 0); }

// 1: probability sums with Neumaier compensation instead of the reference's
// plain scalar order, to measure the summation error of either side.
void oracle_set_exact_sums(int on) { g_exact_sums = on != 0; }

// keep_records = false: counts only, no rows written (the row buffers of
// oracle_sample are then not touched).
void oracle_set_keep_records(int on) { g_keep_records = on != 0; }

// Pinned branches for the following runs: spec = repeated (instruction,
// constant, ncond, conds...); n = 0 clears.
void oracle_set_forced(const int64_t* spec, size_t n) {
  g_forced.clear();
  size_t i = 0;
  while (i + 3 <= n) {
    Forced f;
    size_t instr = size_t(spec[i]);
    f.constant = spec[i + 1] != 0;
    size_t nc = size_t(spec[i + 2]);
    for (size_t j = 0; j < nc && i + 3 + j < n; ++j) f.conds.push_back(int32_t(spec[i + 3 + j]));
    g_forced[instr] = f;
    i += 3 + nc;
  }
}

// The component plan of src/component_plan.rs:301-522.
// summary[0..8] = selected, initial_components, component_count,
//   dense_peak_dimension, component_peak_live_dimension,
//   component_allocated_dimension, quantum ops, merges, planned (steps exist)
// work[0..1] = estimated dense work, estimated component work
// comp_max_k[component_count]; steps[ninstr][6] = kind, component,
//   merge offset, merge count, local pivot, deactivate_after;
// merges[nmerges]. Arrays are filled up to `cap` entries each.
int oracle_component_plan(const uint64_t* blob, size_t nwords, int64_t* summary, double* work, int64_t* comp_max_k,
                          int64_t* steps, int64_t* merges, size_t cap) {
  try {
    Program p = decode(blob, nwords);
    CompPlan c = build_component_plan(p);
    int64_t v[9] = {c.selected, c.initial_components, c.component_count, int64_t(c.dense_peak),
                    int64_t(c.peak_live), int64_t(c.allocated), c.quantum_ops, int64_t(c.merges.size()),
                    !c.steps.empty()};
    std::memcpy(summary, v, sizeof v);
    work[0] = c.dense_work;
    work[1] = c.comp_work;
    for (size_t i = 0; i < c.comp_max_k.size() && i < cap; ++i) comp_max_k[i] = c.comp_max_k[i];
    for (size_t i = 0; i < c.steps.size() && i < cap; ++i) {
      const CompStep& st = c.steps[i];
      int64_t row[6] = {st.kind, st.comp, int64_t(st.merge_off), int64_t(st.merge_count), st.local.pivot,
                        st.deactivate};
      std::memcpy(steps + 6 * i, row, sizeof row);
    }
    for (size_t i = 0; i < c.merges.size() && i < cap; ++i) merges[i] = c.merges[i];
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// info[0..10] = n, initial_k, max_k, nsymbols, nrecords, ndetectors,
// nobservables, ninstr, nmeasure, ncoin, nexpvals
int oracle_info(const uint64_t* blob, size_t nwords, int64_t* info) {
  try {
    Program p = decode(blob, nwords);
    int64_t nm = 0, nc = 0;
    for (const auto& in : p.instrs) {
      nm += in.op == TICIT_OP_MEASURE;
      nc += in.op == TICIT_OP_COIN;
    }
    int64_t v[11] = {p.n, p.initial_k, p.max_k, p.nsymbols, p.nrecords, p.ndetectors,
                     int64_t(p.observables.size()), int64_t(p.instrs.size()), nm, nc, p.nexpvals};
    std::memcpy(info, v, sizeof v);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int64_t oracle_default_batch(int max_k) {
  try {
    return int64_t(default_batch(max_k));
  } catch (...) {
    return -1;
  }
}

// Reference-stream sampling: `chunk_shots`/`batch_size` of 0 select the
// reference defaults. Rows are concatenated in worker order, exactly as
// src/sampler/prepared.rs:834-852 does. timing[0] = wall seconds of the
// sampling call; exp_vals (optional) receives the expectation values
// [shots][nexpvals] in the same row order (src/sampler/prepared.rs:1020-1025).
int oracle_sample(const uint64_t* blob, size_t nwords, uint64_t shots, uint64_t seed, uint64_t chunk_shots,
                  uint64_t batch_size, int threads, int bit_packed, int observable, uint8_t* meas, uint8_t* det,
                  uint8_t* obs, uint64_t* counts, uint64_t* observable_ones, double* timing, double* exp_vals) {
  try {
    Program p = decode(blob, nwords);
    check_expected(p);
    size_t batch = batch_size ? size_t(batch_size) : default_batch(p.max_k);
    size_t chunk = chunk_shots ? size_t(chunk_shots) : (batch > 2048 ? batch : 2048);
    // The reference compacts dead shots out of a block before every draw
    // (src/sampler/batch/postselect.rs:1-5), which moves later draws of a
    // multi-shot block; this restatement keeps dead shots in place, so it
    // runs postselection only where no branch is drawn or one shot per block.
    bool postselect = false, draws = false;
    for (const auto& in : p.instrs) {
      postselect |= in.op == TICIT_OP_DETECTOR && in.postselect;
      draws |= in.op == TICIT_OP_MEASURE || in.op == TICIT_OP_COIN;
    }
    if (postselect && draws && batch != 1)
      throw OracleError("postselection with branch draws is restated at one shot per block");
    uint64_t nchunks = (shots + chunk - 1) / chunk;
    uint64_t bpc = (chunk + batch - 1) / batch;
    if (threads < 1) threads = 1;
    size_t active = size_t(nchunks < uint64_t(threads) ? (nchunks ? nchunks : 1) : uint64_t(threads));
    std::vector<Outputs> outs(active);
    std::vector<std::string> errs(active);
    auto t0 = std::chrono::steady_clock::now();
    auto worker = [&](size_t wid) {
      try {
        Runner run(p, bit_packed != 0, observable, g_component_mode);
        Outputs& o = outs[wid];
        o.obs_ones.assign(p.observables.size(), 0);
        ExoTable exo;
        for (uint64_t c = wid; c < nchunks; c += active) {
          size_t here = size_t(std::min<uint64_t>(chunk, shots - c * chunk));
          presample(exo, p, here, block_seed(kExoBase, seed, c));
          size_t local = 0;
          uint64_t lb = 0;
          while (local < here) {
            size_t blk = std::min(batch, here - local);
            Source src;
            src.rng = block_seed(kBranchBase, seed, c * bpc + lb);
            run.run_block(exo, local, blk, src, o, nullptr, false);
            local += batch;
            ++lb;
          }
        }
      } catch (const std::exception& e) {
        errs[wid] = e.what();
      }
    };
    if (active == 1) {
      worker(0);
    } else {
      std::vector<std::thread> ts;
      for (size_t w = 0; w < active; ++w) ts.emplace_back(worker, w);
      for (auto& t : ts) t.join();
    }
    double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (const auto& e : errs)
      if (!e.empty()) throw OracleError(e);
    Outputs all;
    all.obs_ones.assign(p.observables.size(), 0);
    for (auto& o : outs) {
      all.meas.insert(all.meas.end(), o.meas.begin(), o.meas.end());
      all.det.insert(all.det.end(), o.det.begin(), o.det.end());
      all.obs.insert(all.obs.end(), o.obs.begin(), o.obs.end());
      all.expv.insert(all.expv.end(), o.expv.begin(), o.expv.end());
      all.shots += o.shots;
      all.accepted += o.accepted;
      all.discarded += o.discarded;
      all.logical += o.logical;
      for (size_t i = 0; i < o.obs_ones.size(); ++i) all.obs_ones[i] += o.obs_ones[i];
    }
    copy_out(all, meas, det, obs, counts, observable_ones, exp_vals);
    if (timing) timing[0] = wall;
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// Injected sampling: exo = symbol-major planes [nsymbols][ceil(shots/64)]
// (only exogenous symbols are read), uniforms = [nmeasure][shots], coins =
// [ncoin][ceil(shots/64)]. probs (optional) receives P(branch = 1) per
// (measure ordinal, shot); final_re/final_im (optional, length 2^max_k)
// receive shot 0's final active state and final_k its width; exp_vals
// (optional) receives [shots][nexpvals].
int oracle_sample_injected(const uint64_t* blob, size_t nwords, uint64_t shots, const uint64_t* exo_planes,
                           const double* uniforms, const uint64_t* coins, int bit_packed, int observable,
                           uint8_t* meas, uint8_t* det, uint8_t* obs, uint64_t* counts, uint64_t* observable_ones,
                           double* probs, double* final_re, double* final_im, int* final_k, double* exp_vals) {
  try {
    Program p = decode(blob, nwords);
    check_expected(p);
    Runner run(p, bit_packed != 0, observable, g_component_mode);
    run.fp32 = g_fp32_storage;
    Outputs o;
    o.obs_ones.assign(p.observables.size(), 0);
    ExoTable exo;
    exo.shots = shots;
    exo.words = (shots + 63) / 64;
    exo.vals.assign(size_t(p.nsymbols) * exo.words, 0);
    if (exo_planes) std::memcpy(exo.vals.data(), exo_planes, exo.vals.size() * 8);
    // one shot per block keeps every shot's arithmetic independent
    for (size_t s = 0; s < shots; ++s) {
      Source src;
      src.injected = true;
      src.uniforms = uniforms;
      src.coins = coins;
      src.shots_total = shots;
      src.first_shot = s;
      run.run_block(exo, s, 1, src, o, probs, s == 0 && final_re != nullptr);
    }
    copy_out(o, meas, det, obs, counts, observable_ones, exp_vals);
    if (final_re && shots > 0) {
      std::memcpy(final_re, o.final_re.data(), o.final_re.size() * 8);
      std::memcpy(final_im, o.final_im.data(), o.final_im.size() * 8);
      if (final_k) *final_k = o.final_k;
    }
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// reference_sample_for_program (src/sampler/prepared.rs:355-405): one
// noiseless shot (every exogenous symbol 0) whose branches come from the
// stream the reference's one-shot executor starts at state 1
// (BatchFactoredExecutorState::new(program, 1, 1), src/sampler/batch/mod.rs:
// 451-465), obeying the pinned branches, without postselection or expected
// bits. det/obs receive ndetectors / nobservables bytes (0 / 1).
int oracle_reference_sample(const uint64_t* blob, size_t nwords, uint8_t* det, uint8_t* obs) {
  try {
    Program p = decode(blob, nwords);
    const size_t nobs = p.observables.size();
    if (p.ndetectors == 0 && nobs == 0) return 0;
    Runner run(p, false, 0, g_component_mode);
    run.apply_expected = false;
    run.postselect = false;
    Outputs o;
    o.obs_ones.assign(nobs, 0);
    ExoTable exo;
    exo.shots = 1;
    exo.words = 1;
    exo.vals.assign(size_t(p.nsymbols), 0);
    Source src;
    src.rng = 1;
    run.run_block(exo, 0, 1, src, o, nullptr, false);
    if (p.ndetectors) std::memcpy(det, o.det.data(), size_t(p.ndetectors));
    if (nobs) std::memcpy(obs, o.obs.data(), nobs);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// Expected bits for the following runs (nonzero inverts); n = 0 clears.
void oracle_set_expected(const uint8_t* det, size_t ndet, const uint8_t* obs, size_t nobs) {
  g_expected_det.assign(det, det + (det ? ndet : 0));
  g_expected_obs.assign(obs, obs + (obs ? nobs : 0));
}

// Exogenous presample of one chunk, exposed so tests can pin the noise
// stream on its own: planes [nsymbols][ceil(shots/64)].
int oracle_presample(const uint64_t* blob, size_t nwords, uint64_t shots, uint64_t stream_seed, uint64_t* planes) {
  try {
    Program p = decode(blob, nwords);
    ExoTable exo;
    presample(exo, p, size_t(shots), stream_seed);
    std::memcpy(planes, exo.vals.data(), exo.vals.size() * 8);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

uint64_t oracle_splitmix_next(uint64_t* state) { return next_u64(*state); }
double oracle_rand_float(uint64_t* state) { return rand_float(*state); }
uint64_t oracle_block_seed(uint64_t base, uint64_t seed, uint64_t block) { return block_seed(base, seed, block); }
uint64_t oracle_categorical_row(uint64_t* state, const double* probs, size_t n) {
  std::vector<double> v(probs, probs + n);
  return sample_categorical_row(*state, v);
}
int oracle_bernoulli(uint64_t* state, double p, int* out) {
  try {
    *out = sample_bernoulli(*state, p);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}
size_t oracle_insert_zero_bit(size_t packed, int bit) { return insert_zero_bit(packed, bit); }

}  // extern "C"
