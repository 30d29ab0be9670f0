// HIP sampling kernels for gfx950.
//
// Two resident tiers share one instruction interpreter:
//   * lane tier (max_k <= 5): one shot per lane, 64 shots per wave. Each
//     lane owns a column of the LDS tables laid out [row][lane], so every
//     uniform-index access is one conflict-free 64-lane LDS transaction. A
//     lane's shot runs the scalar evaluation order of ticit's contiguous
//     kernels (src/contiguous.rs:93-177, 1088-1113, 1582-1606, 2013-2044),
//     so amplitudes and probabilities are bit-identical to the CPU oracle.
//   * group tier (group_tier.hip): states up to the LDS of a workgroup,
//     64-256 threads per shot.
// Both fuse the exogenous-noise walk (frame), expression evaluation,
// amplitude kernels, per-shot branch sampling, record/detector/observable
// XOR and row packing into a single launch.
#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace ticit {

// --------------------------------------------------------------------------
// Lane tier
// --------------------------------------------------------------------------

// 64 x 64 bit transpose across a wave: lane j holds row j on entry; on exit
// lane l holds column l (bit j = bit l of row j). Six butterfly stages, each
// swapping the off-diagonal s x s blocks between lanes l and l ^ s.
__device__ __forceinline__ uint64_t transpose64(uint64_t x, int lane) {
  const uint64_t mask[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                            0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
  for (int i = 0, sh = 32; i < 6; ++i, sh >>= 1) {
    const uint32_t lo = uint32_t(__shfl_xor(int(uint32_t(x)), sh, 64));
    const uint32_t hi = uint32_t(__shfl_xor(int(uint32_t(x >> 32)), sh, 64));
    const uint64_t p = (uint64_t(hi) << 32) | lo;
    x = (lane & sh) ? ((x & ~mask[i]) | ((p & ~mask[i]) >> sh)) : ((x & mask[i]) | ((p & mask[i]) << sh));
  }
  return x;
}

// A pair of amplitudes of one wave's shot: indices i0 (pivot bit clear)
// and i1 = i0 ^ xmask.
struct AmpPair {
  double2 v0, v1;
  int i0, i1;
};

// Runs use(i, load(i)) for i in [0, n), n a power of two, in index order.
// With B > 1 and n >= B it issues the loads of B consecutive indices before
// their uses, so one memory latency covers B elements; use(i) must not
// write what load(j) reads for a later j of the same batch. No guards inside
// a batch, so the batch stays straight-line code.
template <int B, class Load, class Use>
__device__ __forceinline__ void amp_loop(int n, Load load, Use use) {
  if (B > 1 && n >= B) {
#pragma nounroll
    for (int i0 = 0; i0 < n; i0 += B) {
      decltype(load(0)) v[B];
#pragma unroll
      for (int u = 0; u < B; ++u) v[u] = load(i0 + u);
#pragma unroll
      for (int u = 0; u < B; ++u) use(i0 + u, v[u]);
    }
  } else {
    for (int i = 0; i < n; ++i) use(i, load(i));
  }
}

// Runs the 64 shots [wb * 64, wb * 64 + 64) of launch L on one wave.
// STORE (kLaneLds, kLaneBitsGlobal, kLaneAllGlobal) says where the tables
// live: all in LDS; the bit tables and row staging in the wave's device
// memory region `region`, the amplitudes in LDS; or everything in the
// region. The region layout is the LDS layout (tables [item][lane]), so
// every uniform-index access is one coalesced 64-lane transaction there as
// well, and the wave's memory operations on one address complete in issue
// order in both spaces.
template <int KMAX, int STORE>
__device__ __forceinline__ void lane_wave(const DevProgram& P, const DevLaunch& L, const LaneLayout& Y, uint64_t wb,
                                          unsigned char* smem, unsigned char* region) {
  constexpr int DIM = 1 << KMAX;
  constexpr int SDIM = KMAX > 0 ? DIM / 2 : 0;
  // the collapse's output half: registers up to four amplitudes, LDS above
  constexpr bool kRegScratch = KMAX <= 3;
  constexpr int SDIM_LDS = kRegScratch ? 0 : SDIM;
  // With the amplitudes in device memory each amplitude loop issues a batch
  // of independent loads before it computes and stores, so one memory
  // latency covers the batch; in LDS the loops stay one element at a time.
  constexpr bool kBatchAmps = STORE == kLaneAllGlobal;
  constexpr int kAmpBatch = kBatchAmps ? 4 : 1, kPairBatch = kBatchAmps ? 2 : 1;
  static_assert(!kBatchAmps || !kRegScratch, "the batched collapse stages through the scratch half");
  const int lane = threadIdx.x;
  const uint64_t local = wb * 64 + lane;
  const bool live = local < L.shots;
  const uint64_t shot = L.first_shot + local;

  unsigned char* const abase = STORE == kLaneAllGlobal ? region : smem;
  double2* amp = reinterpret_cast<double2*>(abase);  // [DIM][64]
  double2* scr = amp + DIM * 64;                     // [SDIM_LDS][64]
  // Bit-sliced tables: word i holds item i of all 64 shots (bit l = lane l),
  // so a record, detector or expression update is one uniform word
  // operation for the whole wave.
  uint64_t* Pw = STORE == kLaneLds ? reinterpret_cast<uint64_t*>(scr + SDIM_LDS * 64)
                                   : reinterpret_cast<uint64_t*>(region + (STORE == kLaneAllGlobal
                                                                               ? size_t(DIM + SDIM_LDS) * 64 * 16
                                                                               : 0));  // exogenous partials
  uint64_t* Sw = Pw + Y.np;  // runtime slots
  uint64_t* Rw = Sw + Y.ns;  // records
  // The amplitudes are dead once the schedule's last step (the detectors)
  // runs: in LDS the detector table and the row staging share their space
  // when they fit (the host decides, lane_tier_lds_bytes counts the same
  // way). Outside LDS both follow the records.
  const uint32_t amp_bytes = uint32_t((DIM + SDIM_LDS) * 64 * 16);
  const bool stage_in_amps = STORE == kLaneLds && Y.stage_bytes <= amp_bytes;
  const bool dets_in_amps = STORE == kLaneLds && Y.dets_in_amps;
  uint64_t* Dw = dets_in_amps ? reinterpret_cast<uint64_t*>(smem + (stage_in_amps ? Y.stage_bytes : 0u))
                              : Rw + Y.nr;  // detectors
  uint8_t* stage = stage_in_amps ? smem : reinterpret_cast<uint8_t*>(Rw + Y.nr + (dets_in_amps ? 0u : Y.nd));

  const uint32_t nzero = Y.np + Y.ns + Y.nr + (dets_in_amps ? 0u : Y.nd);
  for (uint32_t w = lane; w < nzero; w += 64) Pw[w] = 0;
  __syncthreads();

  // Frame: noise events toggle the exogenous partial bits they feed.
  if (L.injected) {
    const uint64_t words = (L.shots + 63) / 64;
    for (int i = 0; i < P.nexo; ++i) {
      const int32_t c = P.exo_syms[i];
      const uint64_t wv = L.exo[uint64_t(c - 1) * words + wb];
      if (wv)
        for (uint32_t j = P.sym_off[c - 1]; j < P.sym_off[c]; ++j) Pw[P.sym_part[j]] ^= wv;  // same value in every lane
      __syncthreads();
    }
  } else if (!L.reference) {
    const uint64_t me = 1ull << lane;
    sample_exogenous(P, seed_for(kExoSeedBase, launch_seed(L), shot),
                     [&](uint32_t part) { atomicXor(&Pw[part], me); });
  }
  __syncthreads();

  for (int a = 0; a < DIM; ++a) amp[a * 64 + lane] = make_double2(a == 0 ? 1.0 : 0.0, 0.0);

  // expression value of all 64 shots
  auto eval = [&](const DevStep& st) -> uint64_t {
    uint64_t v = (st.op & kFlagConstant) ? ~0ull : 0ull;
    if (st.partial >= 0) v ^= Pw[st.partial];
    for (uint32_t j = 0; j < st.resid_len; ++j) v ^= Sw[konst(P.resid)[st.resid_off + j]];
    return v;
  };
  // Every lane stores the same word; the wave's LDS operations complete in
  // issue order, so later reads by any lane see it.
  auto write_outcome = [&](const DevStep& st, uint64_t v) {
    if (st.a >= 0) Rw[st.a] = v;
    if (st.b >= 0) Sw[st.b] = v;
  };

  // expectation probes (src/sampler/batch/runtime.rs:46-75): one value per
  // shot and slot, [shots][nexp] in global memory
  double* expv = (L.expv && live) ? L.expv + local * uint64_t(P.nexp) : nullptr;

  int k = P.initial_k;
  uint64_t brng = L.reference ? 1ull : seed_for(kBranchSeedBase, launch_seed(L), shot);
  int err = 0;
  // Program tables indexed uniformly are read through the constant address
  // space (scalar loads into SGPRs, no vector round trip per step); C1 took
  // 0.245 ms per 1e6 shots against 0.252 ms with vector loads.
  for (int i = 0; i < P.nlsteps; ++i) {
    // the lane schedule drives the loop; amplitude work reads the full record
    const DevStep st = kload(konst(P.lsteps) + i);
    const uint32_t op = st.op & 0xff;
    if (op == kOpCoinBatch) {
      // consecutive dormant coins (src/sampler/batch/mod.rs:254-274): one
      // draw per coin from every shot's stream, in program order
      for (int32_t j = st.a; j < st.a + st.b; ++j) {
        const DevStep q = kload(konst(P.bsteps) + j);
        uint64_t coins;
        if (L.injected) {
          const uint64_t words = (L.shots + 63) / 64;
          coins = L.coins[uint64_t(q.full) * words + wb];
        } else {
          coins = __ballot(sm_next(brng) & 1);
        }
        Sw[q.c] = coins;
      }
      continue;
    }
    if (op == kOpRecBatch || op == kOpDetBatch) {
      // independent records (one dependency level) or the detectors, one
      // step per lane; each lane's step updates the words of all 64 shots
      for (int32_t base = st.a; base < st.a + st.b; base += 64) {
        const int32_t j = base + lane;
        if (j < st.a + st.b) {
          const DevStep q = P.bsteps[j];
          uint64_t v = (q.op & kFlagConstant) ? ~0ull : 0ull;
          if (op == kOpDetBatch && (q.op & kFlagDetRecords)) {
            if (q.op & kFlagInline) {
              if (q.c > 0) v ^= Rw[q.partial];
              if (q.c > 1) v ^= Rw[q.resid_off];
              if (q.c > 2) v ^= Rw[q.full];
            } else {
              for (int32_t r = 0; r < q.c; ++r) v ^= Rw[P.det_recs[q.b + r]];
            }
          } else {
            if (q.partial >= 0) v ^= Pw[q.partial];
            if (q.op & kFlagInline) {
              if (q.resid_len > 0) v ^= Sw[q.c];
              if (q.resid_len > 1) v ^= Sw[q.full];
            } else {
              for (uint32_t r = 0; r < q.resid_len; ++r) v ^= Sw[P.resid[q.resid_off + r]];
            }
          }
          if (op == kOpDetBatch) {
            Dw[q.a] = v;
          } else {
            if (q.a >= 0) Rw[q.a] = v;
            if (q.b >= 0) Sw[q.b] = v;
          }
        }
      }
      __syncthreads();  // the next level reads these words
      continue;
    }
    if (op == kOpRecord) {
      if (st.op & kFlagExpVal) {
        if (expv) expv[st.a] = ((eval(st) >> lane) & 1) ? -1.0 : 1.0;
        continue;
      }
      write_outcome(st, eval(st));
      continue;
    }
    if (op == kOpCoin) {
      if (st.op & kFlagExpVal) {  // an unbiased probe reads 0 and draws nothing
        if (expv) expv[st.a] = 0.0;
        continue;
      }
      uint64_t coins;
      if (st.op & kFlagForced) {  // pinned: the XOR of earlier branches, no draw
        const DevInstr& fin = P.instrs[st.full];
        coins = (fin.op & kFlagForcedOne) ? ~0ull : 0ull;
        for (uint32_t j = 0; j < fin.forced_len; ++j) coins ^= Sw[konst(P.resid)[fin.forced_off + j]];
      } else if (L.injected) {
        const uint64_t words = (L.shots + 63) / 64;
        coins = L.coins[uint64_t(st.full) * words + wb];
      } else {
        coins = __ballot(sm_next(brng) & 1);  // one word per shot's stream
      }
      Sw[st.c] = coins;
      write_outcome(st, eval(st));
      continue;
    }
    // With the amplitudes in device memory the compiler cannot tell their
    // stores from the instruction table, and would re-read the instruction
    // (and wait for every store in flight) inside each amplitude loop: take
    // a private copy there, by scalar loads. (In the LDS layout the fields
    // are read where used; a whole copy made C1 3 % slower.)
    const DevInstr& in_table = P.instrs[st.full];
    DevInstr in_copy;
    if constexpr (kBatchAmps) in_copy = kload(konst(P.instrs) + st.full);
    const DevInstr& in = kBatchAmps ? in_copy : in_table;
    auto load_pair = [&](int idx) {
      AmpPair a;
      a.i0 = int(insert_zero_bit(size_t(idx), in.pivot));
      a.i1 = a.i0 ^ int(in.xmask);
      a.v0 = amp[a.i0 * 64 + lane];
      a.v1 = amp[a.i1 * 64 + lane];
      return a;
    };
    if (op == kOpRotate) {
      const bool sign = (eval(st) >> lane) & 1;
      const int dim = 1 << k;
      const double c = in.c;
      if (in.op & kFlagDiagonal) {
        amp_loop<kAmpBatch>(dim, [&](int b) { return amp[b * 64 + lane]; }, [&](int b, double2 v) {
          bool flip = sign != parity64(uint64_t(b) & in.zmask);
          double cr = flip ? -in.me_re : in.me_re;
          double ci = flip ? -in.me_im : in.me_im;
          double fr = c + cr, fi = ci;
          amp[b * 64 + lane] = make_double2(fr * v.x - fi * v.y, fr * v.y + fi * v.x);
        });
      } else if (in.op & kFlagUniformImag) {
        const double q = sign ? -in.me_im : in.me_im;
        if constexpr (kBatchAmps) {
          // the pairs are disjoint: a batch of them is loaded, then written
          amp_loop<kPairBatch>(dim >> 1, load_pair, [&](int, AmpPair a) {
            amp[a.i0 * 64 + lane] = make_double2(c * a.v0.x - q * a.v1.y, c * a.v0.y + q * a.v1.x);
            amp[a.i1 * 64 + lane] = make_double2(c * a.v1.x - q * a.v0.y, c * a.v1.y + q * a.v0.x);
          });
        } else {
          const int sel = 1 << in.pivot, step = sel << 1;
          for (int block = 0; block < dim; block += step) {
            for (int off = 0; off < sel; ++off) {
              int i0 = block + off, i1 = i0 ^ int(in.xmask);
              double2 a0 = amp[i0 * 64 + lane], a1 = amp[i1 * 64 + lane];
              amp[i0 * 64 + lane] = make_double2(c * a0.x - q * a1.y, c * a0.y + q * a1.x);
              amp[i1 * 64 + lane] = make_double2(c * a1.x - q * a0.y, c * a1.y + q * a0.x);
            }
          }
        }
      } else {
        const bool xz_odd = in.op & kFlagXzOdd;
        amp_loop<kPairBatch>(dim >> 1, load_pair, [&](int, AmpPair a) {
          bool lodd = parity64(uint64_t(a.i0) & in.zmask);
          double ld = (sign != lodd) ? -1.0 : 1.0;
          double rd = xz_odd ? -ld : ld;
          double lre = ld * in.me_re, lim = ld * in.me_im;
          double rre = rd * in.me_re, rim = rd * in.me_im;
          amp[a.i0 * 64 + lane] = make_double2(c * a.v0.x + rre * a.v1.x - rim * a.v1.y,
                                               c * a.v0.y + rre * a.v1.y + rim * a.v1.x);
          amp[a.i1 * 64 + lane] = make_double2(c * a.v1.x + lre * a.v0.x - lim * a.v0.y,
                                               c * a.v1.y + lre * a.v0.y + lim * a.v0.x);
        });
      }
    } else if (op == kOpPromote) {
      const bool sign = (eval(st) >> lane) & 1;
      const int dim = 1 << k;
      const double c = in.c, q = sign ? in.s : -in.s;
      // the new half [dim, 2 dim) is written, never read, by the loop
      amp_loop<kAmpBatch>(dim, [&](int b) { return amp[b * 64 + lane]; }, [&](int b, double2 v) {
        amp[b * 64 + lane] = make_double2(c * v.x, c * v.y);
        amp[(dim + b) * 64 + lane] = make_double2(-q * v.y, q * v.x);
      });
      ++k;
    } else if (op == kOpMeasure) {
      const int out_dim = (1 << k) >> 1;
      const bool diag = in.op & kFlagDiagonal;
      double pt = 0.0;  // P(true), accumulated in index order
      if (diag) {
        amp_loop<kAmpBatch>(
            out_dim, [&](int i) { return amp[int(diagonal_source(in, size_t(i), true)) * 64 + lane]; },
            [&](int, double2 v) { pt += v.x * v.x + v.y * v.y; });
      } else {
        amp_loop<kPairBatch>(out_dim, load_pair, [&](int, AmpPair a) {
          double dir = parity64(uint64_t(a.i0) & in.zmask) ? 1.0 : -1.0;  // branch = true
          double c1r = dir * in.c1_re, c1i = dir * in.c1_im;
          double ar = kInvSqrt2 * a.v0.x + c1r * a.v1.x - c1i * a.v1.y;
          double ai = kInvSqrt2 * a.v0.y + c1r * a.v1.y + c1i * a.v1.x;
          pt += ar * ar + ai * ai;
        });
      }
      pt = clamp01(pt);
      if (L.want_probs && live) L.probs[uint64_t(in.ordinal) * L.shots + local] = pt;
      if (st.op & kFlagExpVal) {
        // src/sampler/batch/active.rs:818-850: +-(1 - 2 P(true)), no collapse
        const double sign = ((eval(st) >> lane) & 1) ? -1.0 : 1.0;
        if (expv) expv[st.a] = sign * (1.0 - 2.0 * pt);
        continue;
      }
      bool branch = false;
      if (in.op & kFlagForced) {
        // pinned: the XOR of earlier branches, no draw; only a fair coin
        uint64_t f = (in.op & kFlagForcedOne) ? ~0ull : 0ull;
        for (uint32_t j = 0; j < in.forced_len; ++j) f ^= Sw[konst(P.resid)[in.forced_off + j]];
        branch = (f >> lane) & 1;
        if (live && fabs(pt - 0.5) > kFairBranchTolerance) err |= kDevErrUnfairPin;
      } else if (L.injected) {
        if (live) {
          double u = L.uniforms[uint64_t(in.ordinal) * L.shots + local];
          branch = pt <= 0.0 ? false : (pt >= 1.0 ? true : u < pt);
        }
      } else if (!(pt >= 0.0 && pt <= 1.0)) {
        err |= kDevErrBadProbability;
      } else {
        branch = pt <= 0.0 ? false : (pt >= 1.0 ? true : sm_uniform(brng) < pt);
      }
      double ps = branch ? pt : 1.0 - pt;
      if (!(ps > 0.0)) {
        err |= kDevErrImpossibleBranch;
        ps = 1.0;
      }
      const double inv = 1.0 / sqrt(ps);
      if (diag) {
        // in place: the source of output i is at or above i, and a batch
        // loads all its sources before it writes
        amp_loop<kAmpBatch>(
            out_dim, [&](int i) { return amp[int(diagonal_source(in, size_t(i), branch)) * 64 + lane]; },
            [&](int i, double2 v) { amp[i * 64 + lane] = make_double2(v.x * inv, v.y * inv); });
      } else {
        auto mixed_of = [&](const AmpPair& a) {
          bool odd = parity64(uint64_t(a.i0) & in.zmask);
          double dir = (branch != odd) ? -1.0 : 1.0;
          double c1r = dir * in.c1_re, c1i = dir * in.c1_im;
          double ar = kInvSqrt2 * a.v0.x + c1r * a.v1.x - c1i * a.v1.y;
          double ai = kInvSqrt2 * a.v0.y + c1r * a.v1.y + c1i * a.v1.x;
          return make_double2(ar * inv, ai * inv);
        };
        auto mixed = [&](int idx) { return mixed_of(load_pair(idx)); };
        if constexpr (kBatchAmps) {
          // an output's partner may lie below it: outputs go to the scratch
          // half, then are copied down
          amp_loop<kPairBatch>(out_dim, load_pair,
                               [&](int i, AmpPair a) { scr[i * 64 + lane] = mixed_of(a); });
          amp_loop<kAmpBatch>(out_dim, [&](int i) { return scr[i * 64 + lane]; },
                              [&](int i, double2 v) { amp[i * 64 + lane] = v; });
        } else if constexpr (kRegScratch) {
          double2 out[SDIM > 0 ? SDIM : 1];
#pragma unroll
          for (int idx = 0; idx < SDIM; ++idx)
            if (idx < out_dim) out[idx] = mixed(idx);
#pragma unroll
          for (int idx = 0; idx < SDIM; ++idx)
            if (idx < out_dim) amp[idx * 64 + lane] = out[idx];
        } else {
          for (int idx = 0; idx < out_dim; ++idx) scr[idx * 64 + lane] = mixed(idx);
          for (int idx = 0; idx < out_dim; ++idx) amp[idx * 64 + lane] = scr[idx * 64 + lane];
        }
      }
      --k;
      Sw[st.c] = __ballot(branch);
      write_outcome(st, eval(st));
    }
  }
  if (err && live) atomicOr(L.error, err);

  // Observables (src/sampler/prepared.rs:881-907) and their popcounts;
  // shots a postselected detector fired on are discarded
  // (src/sampler/batch/postselect.rs) and not counted.
  const uint32_t nlive = uint32_t(min<uint64_t>(64, L.shots - wb * 64));
  uint64_t live_mask = nlive >= 64 ? ~0ull : ((1ull << nlive) - 1);
  if (P.npost > 0) {
    uint64_t dead = 0;
    for (int32_t d = 0; d < P.npost; ++d) dead |= Dw[konst(P.post_dets)[d]];
    live_mask &= ~dead;
    if (L.keep && live) L.keep[local] = uint8_t((live_mask >> lane) & 1);
  }
  auto obs_word = [&](uint32_t o) {
    uint64_t v = konst(P.obs_flip)[o] ? ~0ull : 0ull;
    for (uint32_t j = konst(P.obs_off)[o]; j < konst(P.obs_off)[o + 1]; ++j) v ^= Rw[konst(P.obs_recs)[j]];
    return v;
  };
  for (int o = 0; o < P.nobs; ++o) {
    unsigned long long ones = __popcll(obs_word(uint32_t(o)) & live_mask);
    if (lane == 0 && ones) atomicAdd(&L.obs_ones[o], ones);
  }
  if (!L.keep_records) return;

  // Row packing (src/sampler/prepared.rs:909-941). Columns go 64 at a time:
  // lane j loads column c0 + j's word (bit l = shot l), a 64 x 64 bit
  // transpose across the wave leaves shot l's bits of those 64 columns in
  // lane l (bit j = column c0 + j: eight packed row bytes, little bit
  // order), and the rows are staged contiguously in LDS so the wave writes
  // its block of rows as 8-byte words.
  auto emit = [&](uint8_t* out, uint32_t row_bytes, auto&& word_at, uint32_t ncols) {
    if (row_bytes == 0) return;
    uint8_t* row = stage + size_t(lane) * row_bytes;
    for (uint32_t c0 = 0; c0 < ncols; c0 += 64) {
      const uint32_t c = c0 + uint32_t(lane);
      const uint64_t x = transpose64(c < ncols ? word_at(c) : 0ull, lane);
      if (L.bit_packed) {
        const uint32_t b0 = c0 / 8;
        for (uint32_t q = 0; q < 8 && b0 + q < row_bytes; ++q) row[b0 + q] = uint8_t(x >> (8 * q));
      } else {
        for (uint32_t q = 0; q < 64 && c0 + q < ncols; ++q) row[c0 + q] = uint8_t((x >> q) & 1);
      }
    }
    __syncthreads();
    const size_t total = size_t(nlive) * row_bytes;
    uint8_t* dst = out + (wb * 64) * row_bytes;
    if (!(total & 7) && !(reinterpret_cast<uintptr_t>(dst) & 7)) {
      const uint2* src8 = reinterpret_cast<const uint2*>(stage);
      uint2* dst8 = reinterpret_cast<uint2*>(dst);
      for (size_t b = lane; b < total / 8; b += 64) dst8[b] = src8[b];
    } else {
      for (size_t b = lane; b < total; b += 64) dst[b] = stage[b];
    }
    __syncthreads();
  };
  emit(L.meas, L.meas_row, [&](uint32_t c) { return Rw[c]; }, uint32_t(P.nrec));
  emit(L.det, L.det_row, [&](uint32_t c) { return Dw[c]; }, uint32_t(P.ndet));
  emit(L.obs, L.obs_row, obs_word, uint32_t(P.nobs));
}

// One wave per 64 shots. With the tables in LDS the grid covers the launch;
// with tables in device memory the grid is the workspace's slot count, and
// each workgroup runs the wave blocks blockIdx.x, blockIdx.x + gridDim.x,
// ... in its own region (every wave leaves the loop after the launch's last
// block).
template <int KMAX, int STORE>
// With the amplitudes in device memory the kernel is held to 128 VGPRs (four
// waves per SIMD, without spills at four-element batches): its waves wait
// on their regions, and a fourth wave per SIMD took the CCZ d05 fixture from
// 3.28e7 to 3.49e7 shots/s against three waves at 165 VGPRs.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(STORE == kLaneAllGlobal ? 4 : 1))) void
lane_tier_kernel(DevProgram P, DevLaunch L, LaneLayout Y, unsigned char* ws, size_t slot_bytes) {
  extern __shared__ __align__(16) unsigned char smem[];
  if constexpr (STORE == kLaneLds) {
    lane_wave<KMAX, STORE>(P, L, Y, blockIdx.x, smem, nullptr);
  } else {
    unsigned char* region = ws + size_t(blockIdx.x) * slot_bytes;
    const uint64_t nwb = (L.shots + 63) / 64;
    for (uint64_t wb = blockIdx.x; wb < nwb; wb += gridDim.x) {
      lane_wave<KMAX, STORE>(P, L, Y, wb, smem, region);
      __syncthreads();  // the region's last readers before the next block clears it
    }
  }
}

// --------------------------------------------------------------------------
// Host launchers
// --------------------------------------------------------------------------

static size_t lane_amp_bytes(int kmax) {
  size_t dim = size_t(1) << kmax;
  size_t sdim = kmax > 3 ? dim / 2 : 0;  // widths up to 3 collapse in registers
  return (dim + sdim) * 64 * 16;
}

// Whether the detector table fits the amplitudes' LDS beside the row staging.
static uint32_t lane_dets_in_amps(int kmax, const LaneLayout& y) {
  const size_t amps = lane_amp_bytes(kmax);
  const size_t staged = y.stage_bytes <= amps ? y.stage_bytes : 0;
  return uint32_t(y.dets_once && y.nd > 0 && staged % 8 == 0 && staged + size_t(y.nd) * 8 <= amps);
}

size_t lane_tier_lds_bytes(int kmax, const LaneLayout& y) {
  if (y.store == kLaneBitsGlobal) return lane_amp_bytes(kmax);
  if (y.store == kLaneAllGlobal) return 0;
  const size_t amps = lane_amp_bytes(kmax);
  const size_t dets = lane_dets_in_amps(kmax, y) ? 0 : size_t(y.nd) * 8;
  return amps + size_t(y.np + y.ns + y.nr) * 8 + dets + (y.stage_bytes <= amps ? 0 : y.stage_bytes);
}

// Device memory of one wave's region (tables outside LDS): the bit tables,
// the row staging and, for kLaneAllGlobal, the amplitudes ahead of them.
size_t lane_region_bytes(int kmax, const LaneLayout& y) {
  if (y.store == kLaneLds) return 0;
  size_t b = size_t(y.np + y.ns + y.nr + y.nd) * 8 + y.stage_bytes;
  if (y.store == kLaneAllGlobal) b += lane_amp_bytes(kmax);
  return (b + 255) / 256 * 256;
}

// The layout of one launch: row staging sized for its rows (packed rows are
// 8x smaller than the program's worst case), so in LDS a wave takes less
// space and more waves fit a CU.
LaneLayout lane_launch_layout(int kmax, const LaneLayout& Yp, const DevLaunch& L) {
  LaneLayout Y = Yp;
  uint32_t row = L.keep_records ? (L.meas_row > L.det_row ? L.meas_row : L.det_row) : 0;
  row = row > L.obs_row ? row : L.obs_row;
  Y.stage_bytes = 64 * ((row + 7) / 8 * 8);
  Y.dets_in_amps = Y.store == kLaneLds ? lane_dets_in_amps(kmax, Y) : 0;
  return Y;
}

template <int K, int STORE>
static hipError_t launch_lane_k(const DevProgram& P, const DevLaunch& L, const LaneLayout& Yp,
                                const LaneWorkspace* ws, hipStream_t stream) {
  const LaneLayout Y = lane_launch_layout(K, Yp, L);
  const size_t lds = lane_tier_lds_bytes(K, Y);
  const void* fn = reinterpret_cast<const void*>(&lane_tier_kernel<K, STORE>);
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
  if (e != hipSuccess) return e;
  const uint64_t nwb = (L.shots + 63) / 64;
  if constexpr (STORE == kLaneLds) {
    hipLaunchKernelGGL((lane_tier_kernel<K, STORE>), dim3(unsigned(nwb)), dim3(64), lds, stream, P, L, Y, nullptr,
                       size_t(0));
  } else {
    const size_t slot = lane_region_bytes(K, Y);
    if (!ws || !ws->base || slot == 0) return hipErrorInvalidValue;
    uint64_t slots = ws->bytes / slot;
    if (ws->max_slots && slots > ws->max_slots) slots = ws->max_slots;
    if (slots > nwb) slots = nwb;
    if (slots == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL((lane_tier_kernel<K, STORE>), dim3(unsigned(slots)), dim3(64), lds, stream, P, L, Y,
                       ws->base, slot);
  }
  return hipGetLastError();
}

// LDS bytes and resident workgroups per CU of a lane-tier launch with
// packed rows of `row_bytes`.
template <int K, int STORE>
static void lane_occupancy_k(const LaneLayout& Yp, uint32_t row_bytes, uint32_t* lds, uint32_t* blocks) {
  LaneLayout Y = Yp;
  Y.stage_bytes = 64 * ((row_bytes + 7) / 8 * 8);
  Y.dets_in_amps = Y.store == kLaneLds ? lane_dets_in_amps(K, Y) : 0;
  *lds = uint32_t(lane_tier_lds_bytes(K, Y));
  int b = 0;
  const void* fn = reinterpret_cast<const void*>(&lane_tier_kernel<K, STORE>);
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(*lds)) == hipSuccess)
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, fn, 64, *lds);
  *blocks = uint32_t(b);
}

template <int STORE>
static void lane_occupancy_store(int kmax, const LaneLayout& Y, uint32_t row_bytes, uint32_t* lds, uint32_t* blocks) {
  switch (kmax) {
    case 0: return lane_occupancy_k<0, STORE>(Y, row_bytes, lds, blocks);
    case 1: return lane_occupancy_k<1, STORE>(Y, row_bytes, lds, blocks);
    case 2: return lane_occupancy_k<2, STORE>(Y, row_bytes, lds, blocks);
    case 3: return lane_occupancy_k<3, STORE>(Y, row_bytes, lds, blocks);
    case 4: return lane_occupancy_k<4, STORE>(Y, row_bytes, lds, blocks);
    case 5: return lane_occupancy_k<5, STORE>(Y, row_bytes, lds, blocks);
    default: return;
  }
}

void lane_tier_occupancy(int kmax, const LaneLayout& Y, uint32_t row_bytes, uint32_t* lds, uint32_t* blocks) {
  *lds = *blocks = 0;
  if (Y.store == kLaneLds) return lane_occupancy_store<kLaneLds>(kmax, Y, row_bytes, lds, blocks);
  if (Y.store == kLaneBitsGlobal) return lane_occupancy_store<kLaneBitsGlobal>(kmax, Y, row_bytes, lds, blocks);
  switch (kmax) {
    case 6: return lane_occupancy_k<6, kLaneAllGlobal>(Y, row_bytes, lds, blocks);
    case 7: return lane_occupancy_k<7, kLaneAllGlobal>(Y, row_bytes, lds, blocks);
    case 8: return lane_occupancy_k<8, kLaneAllGlobal>(Y, row_bytes, lds, blocks);
    default: return;
  }
}

template <int STORE>
static hipError_t launch_lane_store(int kmax, const DevProgram& P, const DevLaunch& L, const LaneLayout& Y,
                                    const LaneWorkspace* ws, hipStream_t stream) {
  switch (kmax) {
    case 0: return launch_lane_k<0, STORE>(P, L, Y, ws, stream);
    case 1: return launch_lane_k<1, STORE>(P, L, Y, ws, stream);
    case 2: return launch_lane_k<2, STORE>(P, L, Y, ws, stream);
    case 3: return launch_lane_k<3, STORE>(P, L, Y, ws, stream);
    case 4: return launch_lane_k<4, STORE>(P, L, Y, ws, stream);
    case 5: return launch_lane_k<5, STORE>(P, L, Y, ws, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_lane_tier(int kmax, const DevProgram& P, const DevLaunch& L, const LaneLayout& Y,
                            hipStream_t stream, const LaneWorkspace* ws) {
  if (Y.store == kLaneLds) return launch_lane_store<kLaneLds>(kmax, P, L, Y, nullptr, stream);
  if (Y.store == kLaneBitsGlobal) return launch_lane_store<kLaneBitsGlobal>(kmax, P, L, Y, ws, stream);
  switch (kmax) {
    case 6: return launch_lane_k<6, kLaneAllGlobal>(P, L, Y, ws, stream);
    case 7: return launch_lane_k<7, kLaneAllGlobal>(P, L, Y, ws, stream);
    case 8: return launch_lane_k<8, kLaneAllGlobal>(P, L, Y, ws, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ticit
