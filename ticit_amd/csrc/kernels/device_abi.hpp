// Flat device tables for one planned program. Shared by the host lowering
// (host/device_program.cpp) and the HIP kernels; plain structs only.
//
// Instruction semantics follow ticit's six opcodes (src/factored.rs:284-292);
// the precomputed constants are those of PrecomputedActivePauliRotationKernel
// (src/active.rs:162-230) and PrecomputedActivePauliMeasurementKernel
// (src/active.rs:242-377), computed once on the host so device and CPU use
// bit-identical cos/sin values.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ticit {

enum : uint32_t {
  kOpRotate = 0,
  kOpPromote = 1,
  kOpRecord = 2,
  kOpDetector = 3,
  kOpMeasure = 4,
  kOpCoin = 5,
  kOpMerge = 6,  // device only: component src folded into component comp
  // lane-tier schedule only: steps bsteps[a, a + b) evaluated side by side,
  // one per lane (records of one dependency level, or the detectors)
  kOpRecBatch = 7,
  kOpDetBatch = 8,
  // lane-tier schedule only: the draws of consecutive dormant coins
  // bsteps[a, a + b), in program order (their records join the batches)
  kOpCoinBatch = 9,
};

enum : uint32_t {
  kFlagDiagonal = 1u << 8,      // rotation with xmask == 0, or diagonal measurement
  kFlagUniformImag = 1u << 9,   // rotation pairs with zmask == 0
  kFlagXzOdd = 1u << 10,        // popcount(xmask & zmask) odd
  kFlagConstant = 1u << 11,     // expression constant; a detector XORs it into its
                                // value however it is evaluated (an expected
                                // detector bit of 1, src/sampler/prepared.rs:909-941)
  kFlagDiagPhase = 1u << 12,    // diagonal measurement of a -1 coefficient body
  kFlagDetRecords = 1u << 13,   // detector reads record bits, not its expression
  kFlagFresh = 1u << 14,        // promotion that starts a new one-qubit component
  kFlagZeroState = 1u << 15,    // fresh component in |0> (initial components)
  kFlagPostselect = 1u << 17,   // detector whose firing discards the shot (src/factored.rs:227)
  kFlagForced = 1u << 18,       // pinned Measure/Coin: the branch is forced_const XOR the
                                // slots forced_off..+forced_len, not drawn
                                // (src/sampler/pinning.rs, src/sampler/batch/active.rs:535-554)
  kFlagForcedOne = 1u << 19,    // forced_const
  kFlagInline = 1u << 20,       // lane-tier batch entry with its short id list inline:
                                // a record-list detector's <= 3 record ids in
                                // partial, resid_off, full (count in c); otherwise
                                // <= 2 runtime-slot ids in c, full (count in resid_len)
  kFlagExpVal = 1u << 16,       // expectation probe (Record/Measure/Coin): `record`
                                // holds the expectation slot; no record, branch or
                                // collapse (src/sampler/batch/runtime.rs:311-356)
};

struct alignas(16) DevInstr {
  uint32_t op;          // opcode | flags
  int32_t nq;           // active width the instruction acts on
  int32_t partial;      // exogenous partial index, -1 = none
  uint32_t resid_off;   // runtime-slot list of the expression residual
  uint32_t resid_len;
  int32_t record;       // 0-based record written, -1 = none
  int32_t rc_slot;      // runtime slot assigned the outcome, -1 = none
  int32_t branch_slot;  // runtime slot of the minted branch symbol, -1 = none
  int32_t detector;     // 0-based detector written, -1 = none
  uint32_t det_off;     // detector record list (0-based)
  uint32_t det_len;
  int32_t pivot;        // measurement pivot / rotation pair bit
  int32_t ordinal;      // Measure: measurement ordinal; Coin: coin ordinal
  uint32_t comp_off;    // amplitude offset of the component acted on (0 dense)
  int32_t comp;         // component id (0 dense)
  int32_t src_comp;     // Merge: component folded in
  uint32_t src_off;     // Merge: its amplitude offset
  int32_t src_k;        // Merge: its width
  uint32_t forced_off;   // pinned Measure/Coin: runtime-slot list of the forced branch
  uint32_t forced_len;
  uint64_t xmask;
  uint64_t zmask;
  uint64_t zwop;        // z mask without the pivot bit (diagonal measurement)
  double c;             // cos(angle)
  double s;             // sin(angle)
  double me_re;         // rotation coefficient -i sin(phi) i^phase
  double me_im;
  double c1_re;         // measurement partner weight conj(i^phase)/sqrt(2)
  double c1_im;
};

// Compact view of one instruction for the resident tiers' interpreter loop:
// everything a record, detector, coin or branch update reads, in 32 bytes,
// so the stream the loop walks stays in the scalar data cache; amplitude
// instructions read their full DevInstr through `full`.
struct alignas(16) DevStep {
  uint32_t op;          // opcode | flags (as DevInstr::op)
  int32_t partial;      // exogenous partial index, -1 = none
  uint32_t resid_off;   // runtime-slot list of the expression residual
  uint32_t resid_len;
  int32_t a;            // Record/Measure/Coin: record (-1 none); Detector: detector
  int32_t b;            // Record/Measure/Coin: rc_slot (-1 none); Detector: det_off
  int32_t c;            // Measure/Coin: branch_slot; Detector: det_len
  int32_t full;         // index of the DevInstr (Coin: its ordinal, unless pinned)
};

// Exogenous noise plan (src/factored.rs:444-462, src/planner.rs:1455-1480).
// Noise events flip partials through toggle lists: entry k of a family
// covers partials tl_part[tl_off[tl_base + k] .. tl_off[tl_base + k + 1]),
// the XOR of the partial lists of the symbols the event sets.
struct DevCat {       // categorical distribution drawn once per shot
  uint32_t nbits, nrows;
  uint32_t prob_off;  // f64 pool: cumulative probabilities, summed in row order
  uint32_t assign_off;  // u64 pool
  uint32_t cond_off;  // i32 pool
  uint32_t tl_base;   // toggle list of row r: entry r
};

// Geometric-gap walks: a first uniform below `none_below` is a gap of at
// least the walk's trial count (no event), decided without the logarithm;
// 0 disables the shortcut. The host sets it to (1-p)^trials * (1 - 1e-9),
// far enough below the exact boundary that the rounded ln(u)/denom the
// kernels would compute lands on the same side.
struct DevRare {      // rare categorical group walked by geometric gaps
  double denom;       // ln(1 - event_p)
  double none_below;  // first-draw shortcut over nsets trials
  uint32_t enabled;   // event_p > 0
  uint32_t nsets, nbits, nev;
  uint32_t cond_off;  // i32 pool: nsets * nbits conditions
  uint32_t assign_off;  // u64 pool: one assignment per row
  uint32_t ev_row_off;  // i32 pool
  uint32_t ev_prob_off;  // f64 pool: cumulative event probabilities
  uint32_t tl_base;      // toggle list of (set s, event e): entry s * nev + e
};

enum : uint32_t {
  kBernNone = 0,   // p <= 0: never set
  kBernAll = 1,    // p >= 1: always set
  kBernHalf = 2,   // fair coin word
  kBernLow = 3,    // geometric skip
  kBernCoins = 4,  // 8 coin flips plus geometric leftover
};

struct DevBern {      // src/sampler/exogenous.rs:185-266 at one shot per stream
  int32_t cond;
  uint32_t mode;
  uint32_t invert;
  uint32_t top_bits;
  double denom;       // kBernLow: ln(1 - p)
  double left_denom;  // kBernCoins leftover: ln(1 - p'), 0 when no leftover
  double miss_below;       // shortcut of the kBernLow draw over one trial
  double left_miss_below;  // shortcut of the leftover draw
};

struct DevLowp {      // low-probability Bernoulli group
  double denom;
  double none_below;  // first-draw shortcut over nconds trials
  uint32_t enabled;
  uint32_t nconds;
  uint32_t cond_off;
  uint32_t tl_base;   // toggle list of condition j: entry j
};

// Device pointers of one uploaded program.
struct DevProgram {
  const DevInstr* instrs;
  const DevStep* steps;       // [ninstr] compact view of instrs
  const DevStep* lsteps;      // [nlsteps] lane-tier schedule: steps and batch steps
  const DevStep* bsteps;      // records and detectors the batch steps cover
  const uint32_t* resid;      // runtime slot ids
  const uint32_t* det_recs;   // 0-based record ids
  const uint32_t* post_dets;  // [npost] 0-based postselected detectors
  const uint32_t* obs_off;    // nobs + 1 offsets into obs_recs
  const uint32_t* obs_recs;
  const uint8_t* obs_flip;    // [nobs] 1: the observable's bit is inverted (an expected
                              // observable bit of 1, src/sampler/prepared.rs:965-974)
  const uint32_t* sym_off;    // nsym + 1 offsets: symbol (c - 1) -> partial ids
  const uint32_t* sym_part;
  const uint32_t* tl_off;     // noise-event toggle lists (see DevCat)
  const uint32_t* tl_part;
  const int32_t* exo_syms;    // every exogenous symbol id (injected mode scan)
  const DevCat* cats;
  const DevRare* rares;
  const DevBern* berns;
  const DevLowp* lowps;
  const double* f64pool;
  const uint64_t* u64pool;
  const int32_t* i32pool;
  int32_t ninstr, n, initial_k, max_k;
  int32_t nrec, ndet, nobs, nslots;
  int32_t npartials, nsym, nexo, nmeasure;
  int32_t ncoin, ncat, nrare, nbern;
  int32_t nlowp;
  uint32_t state_dim;   // amplitudes per shot: 2^max_k dense, sum of component capacities otherwise
  int32_t ncomp;        // components (1 dense)
  int32_t max_local_k;  // widest component (max_k dense)
  int32_t nexp;         // expectation slots per shot
  int32_t nlsteps;
  int32_t npost;        // postselected detectors (0: every shot is kept)
};

// Per-launch sampling request.
struct DevLaunch {
  uint64_t shots;
  uint64_t seed;
  uint64_t first_shot;     // global index of shot 0 of this launch
  int32_t bit_packed;
  int32_t keep_records;
  int32_t injected;
  int32_t want_probs;
  // outputs (device)
  uint8_t* meas;           // [shots][meas_row]
  uint8_t* det;
  uint8_t* obs;
  unsigned long long* obs_ones;  // [nobs]
  int* error;              // nonzero on a device-side failure
  double* probs;           // optional [nmeasure][shots]
  // injected inputs (device)
  const uint64_t* exo;     // [nsym][ceil(shots/64)]
  const double* uniforms;  // [nmeasure][shots]
  const uint64_t* coins;   // [ncoin][ceil(shots/64)]
  uint32_t meas_row, det_row, obs_row;
  // noiseless reference shot (src/sampler/prepared.rs:355-405): every
  // exogenous symbol is 0 and the branch stream starts from state 1
  uint32_t reference;
  // optional device word added to `seed` at run time, so a captured graph
  // can be replayed with fresh seeds
  const uint64_t* seed_offset;
  // optional [shots][nexp] expectation values (written when non-null)
  double* expv;
  // postselection: keep[shot] = 1 when no postselected detector fired
  // (written when non-null); observable counts only count kept shots
  uint8_t* keep;
};

// LDS table sizes of the two resident tiers (words of bits per shot).
struct LaneLayout {
  uint32_t pw, sw, rw, dw;  // partial, runtime-slot, record, detector words per shot
  uint32_t stage_bytes;     // per-wave row staging
  // bit-sliced tables of the lane tier: one 64-bit word per item, bit l =
  // the shot of lane l
  uint32_t np, ns, nr, nd;  // partials, runtime slots, records, detectors
  // every detector is written exactly once, by the schedule's last step:
  // its table may share the amplitudes' LDS, dead by then
  uint32_t dets_once;
  uint32_t dets_in_amps;    // per launch: the detector table lives there
  uint32_t store;           // kLaneLds, kLaneBitsGlobal or kLaneAllGlobal
};

// Where a lane-tier program keeps its per-wave tables: all in LDS; the bit
// tables and row staging in device memory with the amplitudes in LDS; or
// everything in device memory (states of 2^6 to 2^8 amplitudes per lane).
enum : uint32_t { kLaneLds = 0, kLaneBitsGlobal = 1, kLaneAllGlobal = 2 };

// Device memory of the lane tier's tables outside LDS: `bytes` split into
// regions of one wave each, at most `max_slots` of them in use (the
// resident waves of the chip).
struct LaneWorkspace {
  unsigned char* base;
  size_t bytes;
  uint64_t max_slots;
};

// A second workspace and stream for a plain wide launch's odd shot batches,
// and the event that forks that stream from the caller's (host side).
struct WideWorkspace;
struct WideSecond {
  WideWorkspace* ws;
  hipStream_t stream;
  hipEvent_t fork;
};

struct GroupLayout {
  uint32_t pw, sw, rw, dw;
};

struct WideLayout {
  uint32_t pw, sw, rw, dw;  // words per shot
  uint32_t words;           // pw + sw + rw + dw
  uint32_t max_tiles;       // probability partials per shot
};

enum : int {
  kDevErrImpossibleBranch = 1,
  kDevErrBadProbability = 2,
  kDevErrUnfairPin = 4,  // a pinned active branch whose P(true) is not 1/2 within 1e-9
};

// Pinning a branch conditions the shot on it; only a fair coin may be pinned
// (src/sampler/batch/active.rs:525,547).
constexpr double kFairBranchTolerance = 1e-9;

}  // namespace ticit
