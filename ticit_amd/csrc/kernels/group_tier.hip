// Group tier: every shot's active state (split-complex FP64) and bit tables
// resident in LDS, its amplitude work spread over TPS threads. TPS of 64 to
// 512 is one shot per workgroup (a 64-thread shot is one wave, so its
// barriers never wait on another shot); TPS of 8, 16 or 32 packs 8, 4 or 2
// shots into one wave, reduced with shuffles inside each lane group, which
// spreads the interpreter's per-instruction work over several shots for
// small states. A plain launch takes the fewest threads that hold its
// widest component; 512 holds states of 2^13 (128 KiB FP64).
// Projections go through registers, PROJ outputs per thread (kProjPer at
// most), so a shot group of TPS threads handles local widths up to
// log2(2 * PROJ * TPS). Each launch picks the smallest PROJ that covers its
// program: the staging registers set the kernel's VGPR count, and with it
// how many waves a SIMD holds.
//
// Two entry points share the body: a plain grid over one program's shots,
// and a persistent kernel that pulls (program, shot block) work units of a
// batch of programs from an atomic counter.
//
// Arithmetic follows ticit's contiguous kernels (src/contiguous.rs:93-177,
// 1088-1113, 1582-1606, 2013-2044); a probability is the pairwise sum of
// its terms in index order, the same tree at every TPS.
#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace ticit {

// Workgroup size for shots of TPS threads: one wave for sub-wave shots,
// else one shot per workgroup.
constexpr int group_block(int tps) { return tps < 64 ? 64 : tps; }
// Shots one work unit of the persistent batch kernel runs in sequence: as
// many as a 256-thread group would hold, so the unit counter is taken once
// per 256 threads' worth of shots.
constexpr int group_unit_shots(int tps) { return tps < 256 ? 256 / tps : 1; }
constexpr int kProjPer = 8;
constexpr int kProjNarrow = 2;
// Copies of a shot's runtime-slot, record and detector bits: one per wave of
// the shot, so bit-only steps need no workgroup barrier (every lane of a
// wave writes the same value to its own copy; a wave's LDS operations
// complete in order). Exogenous partial bits are written once, before the
// first barrier, and shared.
constexpr int group_bit_copies(int tps) { return tps < 64 ? 1 : tps / 64; }

// One DPP lane move of a double (both halves the same pattern).
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double read_lane(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// Sum over each group of N consecutive lanes (N = 8, 16, 32 or 64), left in
// every lane of the group, in a fixed order: pairs of lanes combined by DPP
// moves inside rows of 16 (xor 1, xor 2, mirror of 8, mirror of 16), then
// rows by a shuffle (N = 32) or, for a whole wave, the four row sums read
// as scalars and added as (r0 + r1) + (r2 + r3). IEEE addition commutes, so
// both lanes of a pair hold the same value at every step.
template <int N>
__device__ __forceinline__ double group_lane_sum(double v) {
  v += dpp_move<0xB1>(v);   // quad_perm(1, 0, 3, 2)
  v += dpp_move<0x4E>(v);   // quad_perm(2, 3, 0, 1)
  v += dpp_move<0x141>(v);  // row_half_mirror
  if constexpr (N >= 16) v += dpp_move<0x140>(v);  // row_mirror
  if constexpr (N == 32) v += __shfl_xor(v, 16, 64);
  if constexpr (N == 64) v = (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
  return v;
}

// Index arithmetic of a group-tier state (at most 2^13 amplitudes per
// component): 32-bit forms of insert_zero_bit and parity.
__device__ __forceinline__ uint32_t ins0(uint32_t v, int bit) {
  const uint32_t low = (1u << bit) - 1u;
  return (v & low) | ((v & ~low) << 1);
}
__device__ __forceinline__ bool par32(uint32_t v) { return __popc(v) & 1; }

// Orders the shot group's LDS accesses. A group of at most 64 threads is
// one wave: its lanes run in lockstep and its LDS operations complete in
// order, so only the compiler must not move accesses across this point.
template <int TPS>
__device__ __forceinline__ void group_sync() {
  if constexpr (TPS <= 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Bytes of LDS one shot needs.
__host__ __device__ inline size_t group_shot_bytes(uint32_t state_dim, const GroupLayout& y, int tps) {
  return size_t(state_dim) * 16 + size_t(y.pw + uint32_t(group_bit_copies(tps)) * (y.sw + y.rw + y.dw)) * 8;
}

// Runs shots [first_local, first_local + group_block(TPS) / TPS) of launch L
// on the workgroup (one shot unless TPS < 64).
template <int TPS, int PROJ>
__device__ void group_shots(const DevProgram& P, const DevLaunch& L, const GroupLayout& Y, uint64_t first_local,
                            unsigned char* smem, size_t region_bytes) {
  constexpr int SPB = group_block(TPS) / TPS;
  constexpr int WPG = TPS / 64;  // waves per shot group (0: a group inside one wave)
  static_assert(SPB * WPG * sizeof(double) <= 64, "reduction slots exceed the reserved 64 bytes");
  const int tid = threadIdx.x;
  const int g = tid / TPS;       // shot group
  const int t = tid % TPS;       // thread within the group
  const uint64_t local = first_local + uint64_t(g);
  const bool valid = local < L.shots;
  const uint64_t shot = L.first_shot + local;
  const uint32_t DIM = P.state_dim;
  double* red = reinterpret_cast<double*>(smem);  // [SPB * WPG]
  unsigned char* region = smem + 64 + size_t(g) * region_bytes;
  double* const re0 = reinterpret_cast<double*>(region);
  double* const im0 = re0 + DIM;
  constexpr int COPIES = group_bit_copies(TPS);
  const uint32_t bit_words = Y.sw + Y.rw + Y.dw;
  uint64_t* pbits = reinterpret_cast<uint64_t*>(im0 + DIM);
  uint64_t* sbits = pbits + Y.pw + uint32_t(COPIES > 1 ? t / 64 : 0) * bit_words;  // this wave's copy
  uint64_t* rbits = sbits + Y.sw;
  uint64_t* dbits = rbits + Y.rw;
  // the first lane of each wave (of each shot group below 64 threads)
  // writes that wave's copy; the wave's later reads follow in issue order
  const bool bit_writer = (t & ((TPS < 64 ? TPS : 64) - 1)) == 0;

  for (uint32_t w = t; w < Y.pw + uint32_t(COPIES) * bit_words; w += TPS) pbits[w] = 0;
  const int dim0 = 1 << P.initial_k;
  for (int a = t; a < dim0; a += TPS) {
    re0[a] = a == 0 ? 1.0 : 0.0;
    im0[a] = 0.0;
  }
  group_sync<TPS>();
  if (t == 0 && valid) {
    auto toggle = [&](int32_t c) {
      for (uint32_t j = P.sym_off[c - 1]; j < P.sym_off[c]; ++j) {
        uint32_t part = P.sym_part[j];
        pbits[part >> 6] ^= 1ull << (part & 63);
      }
    };
    if (L.injected) {
      const uint64_t words = (L.shots + 63) / 64;
      for (int i = 0; i < P.nexo; ++i) {
        int32_t c = P.exo_syms[i];
        if ((L.exo[uint64_t(c - 1) * words + (local >> 6)] >> (local & 63)) & 1) toggle(c);
      }
    } else if (!L.reference) {
      sample_exogenous(P, seed_for(kExoSeedBase, launch_seed(L), shot),
                       [&](uint32_t part) { pbits[part >> 6] ^= 1ull << (part & 63); });
    }
  }
  group_sync<TPS>();

  auto bit_of = [&](const uint64_t* tab, uint32_t idx) -> bool { return (tab[idx >> 6] >> (idx & 63)) & 1; };
  // an LDS atomic without return: no read-modify-write round trip
  auto set_bit = [&](uint64_t* tab, uint32_t idx, bool v) {
    const unsigned long long m = 1ull << (idx & 63);
    unsigned long long* w = reinterpret_cast<unsigned long long*>(tab + (idx >> 6));
    if (v) atomicOr(w, m);
    else atomicAnd(w, ~m);
  };
  // program tables, read with scalar loads where the index is uniform
  const auto steps = konst(P.steps);
  const auto instrs = konst(P.instrs);
  const auto resid = konst(P.resid);
  const auto det_recs = konst(P.det_recs);
  auto eval = [&](const DevStep& st) -> bool {
    bool b = (st.op & kFlagConstant) != 0;
    if (st.partial >= 0) b ^= bit_of(pbits, uint32_t(st.partial));
    for (uint32_t j = 0; j < st.resid_len; ++j) b ^= bit_of(sbits, resid[st.resid_off + j]);
    return b;
  };

  // identical in the group
  uint64_t brng = L.reference ? 1ull : seed_for(kBranchSeedBase, launch_seed(L), shot);
  int err = 0;
  // expectation probes (src/sampler/batch/runtime.rs:46-75): thread 0 of the
  // shot writes its [nexp] row
  double* expv = (L.expv && valid && t == 0) ? L.expv + local * uint64_t(P.nexp) : nullptr;
  for (int i = 0; i < P.ninstr; ++i) {
    // the compact step drives the loop; amplitude work reads the full record
    const DevStep st = kload(steps + i);
    const uint32_t op = st.op & 0xff;
    const DevInstr in = kload(instrs + (op == kOpCoin ? 0 : st.full));
    // every amplitude instruction acts on one component: 2^nq amplitudes at
    // comp_off (the whole state, offset 0, when the state is dense)
    double* re = re0 + in.comp_off;
    double* im = im0 + in.comp_off;
    if (op == kOpRotate) {
      const bool sign = eval(st);
      const int dim = 1 << in.nq;
      const double c = in.c;
      const uint32_t xm = uint32_t(in.xmask), zm = uint32_t(in.zmask);
      if (in.op & kFlagDiagonal) {
        for (int b = t; b < dim; b += TPS) {
          bool flip = sign != par32(uint32_t(b) & zm);
          double fr = c + (flip ? -in.me_re : in.me_re), fi = flip ? -in.me_im : in.me_im;
          double r = re[b], m = im[b];
          re[b] = fr * r - fi * m;
          im[b] = fr * m + fi * r;
        }
      } else {
        const bool uni = in.op & kFlagUniformImag;
        const bool xz_odd = in.op & kFlagXzOdd;
        const double q = sign ? -in.me_im : in.me_im;
        for (int idx = t; idx < (dim >> 1); idx += TPS) {
          const uint32_t i0 = ins0(uint32_t(idx), in.pivot), i1 = i0 ^ xm;
          double r0 = re[i0], m0 = im[i0], r1 = re[i1], m1 = im[i1];
          if (uni) {
            re[i0] = c * r0 - q * m1;
            im[i0] = c * m0 + q * r1;
            re[i1] = c * r1 - q * m0;
            im[i1] = c * m1 + q * r0;
          } else {
            bool lodd = par32(i0 & zm);
            double ld = (sign != lodd) ? -1.0 : 1.0;
            double rd = xz_odd ? -ld : ld;
            double lre = ld * in.me_re, lim = ld * in.me_im, rre = rd * in.me_re, rim = rd * in.me_im;
            re[i0] = c * r0 + rre * r1 - rim * m1;
            im[i0] = c * m0 + rre * m1 + rim * r1;
            re[i1] = c * r1 + lre * r0 - lim * m0;
            im[i1] = c * m1 + lre * m0 + lim * r0;
          }
        }
      }
      group_sync<TPS>();
    } else if (op == kOpPromote) {
      const bool sign = eval(st);
      const double c = in.c, q = sign ? in.s : -in.s;
      if (in.op & kFlagFresh) {
        // src/sampler/batch/runtime.rs:511-550: a new one-qubit component
        if (t == 0) {
          const bool zero = in.op & kFlagZeroState;
          re[0] = zero ? 1.0 : c;
          im[0] = 0.0;
          re[1] = 0.0;
          im[1] = zero ? 0.0 : q;
        }
      } else {
        const int dim = 1 << in.nq;
        for (int b = t; b < dim; b += TPS) {
          double r = re[b], m = im[b];
          re[b] = c * r;
          im[b] = c * m;
          re[dim + b] = -q * m;
          im[dim + b] = q * r;
        }
      }
      group_sync<TPS>();
    } else if (op == kOpMerge) {
      // src/sampler/batch/mod.rs:597-680: target (x) source into the target's
      // storage. Outputs with a non-zero source index lie above the target's
      // live 2^nq entries and go first; the in-place zero-index slice after
      // a barrier.
      const int tk = in.nq, td = 1 << in.nq, total = 1 << (in.nq + in.src_k);
      const double* sr = re0 + in.src_off;
      const double* si = im0 + in.src_off;
      for (int o = td + t; o < total; o += TPS) {
        const int tb = o & (td - 1), sb = o >> tk;
        const double a = re[tb], b = im[tb], cr = sr[sb], ci = si[sb];
        re[o] = a * cr - b * ci;
        im[o] = a * ci + b * cr;
      }
      group_sync<TPS>();
      for (int o = t; o < td; o += TPS) {
        const double a = re[o], b = im[o], cr = sr[0], ci = si[0];
        re[o] = a * cr - b * ci;
        im[o] = a * ci + b * cr;
      }
      group_sync<TPS>();
    } else if (op == kOpRecord || op == kOpDetector) {
      bool b = false;
      if (op == kOpDetector && (st.op & kFlagDetRecords)) {
        b = (st.op & kFlagConstant) != 0;
        for (int32_t j = 0; j < st.c; ++j) b ^= bit_of(rbits, det_recs[st.b + j]);
      } else {
        b = eval(st);
      }
      // One lane per wave writes its wave's copy; later readers are lanes
      // of the same wave.
      if (st.op & kFlagExpVal) {
        if (expv) expv[st.a] = b ? -1.0 : 1.0;
      } else if (!bit_writer) {
      } else if (op == kOpDetector) {
        set_bit(dbits, uint32_t(st.a), b);
      } else {
        if (st.a >= 0) set_bit(rbits, uint32_t(st.a), b);
        if (st.b >= 0) set_bit(sbits, uint32_t(st.b), b);
      }
      // A fired postselected detector discards the shot: nothing later can
      // reach its outputs (src/sampler/batch/postselect.rs removes it from
      // the block). With one shot per workgroup `b` is the same in every
      // thread, so the whole group leaves the loop at this step; packed
      // sub-wave shots run on. The noiseless reference shot records every
      // detector (src/sampler/prepared.rs:355-405 runs it without
      // postselection).
      if constexpr (TPS >= 64) {
        if (b && (st.op & kFlagPostselect) && op == kOpDetector && !L.reference) break;
      }
    } else if (op == kOpCoin && (st.op & kFlagExpVal)) {
      if (expv) expv[st.a] = 0.0;  // an unbiased probe reads 0 and draws nothing
    } else if (op == kOpMeasure || op == kOpCoin) {
      bool branch = false;
      if (op == kOpMeasure) {
        const int out_dim = (1 << in.nq) >> 1;
        const bool diag = in.op & kFlagDiagonal;
        const int piv = in.pivot;
        const uint32_t xm = uint32_t(in.xmask), zm = uint32_t(in.zmask), zw = uint32_t(in.zwop);
        const uint32_t dph = (in.op & kFlagDiagPhase) ? 1u : 0u;
        // diagonal_source (src/active.rs:347-356) in 32 bits
        auto dsrc = [&](uint32_t idx, bool br) {
          const uint32_t wo = ins0(idx, piv);
          return wo | ((dph ^ uint32_t(par32(wo & zw)) ^ uint32_t(br)) << piv);
        };
        // P(true) is the pairwise sum of the out_dim terms in index order:
        // thread t sums its block of `per` consecutive terms pairwise, then
        // the group's lanes combine pairwise (group_lane_sum), then its
        // waves. The tree is the same at every TPS, so a shot's rows do not
        // depend on the launch shape (an exact 0 or 1 decides whether a
        // uniform is drawn, and the rounding of the sum decides that).
        double part = 0.0;
        {
          auto term = [&](int idx) -> double {
            if (diag) {
              const uint32_t s = dsrc(uint32_t(idx), true);
              return re[s] * re[s] + im[s] * im[s];
            }
            const uint32_t s0 = ins0(uint32_t(idx), piv), s1 = s0 ^ xm;
            double dir = par32(s0 & zm) ? 1.0 : -1.0;
            double c1r = dir * in.c1_re, c1i = dir * in.c1_im;
            double ar = kInvSqrt2 * re[s0] + c1r * re[s1] - c1i * im[s1];
            double ai = kInvSqrt2 * im[s0] + c1r * im[s1] + c1i * re[s1];
            return ar * ar + ai * ai;
          };
          // out_dim is a power of two of at most PROJ * TPS (the launch's
          // staging bound), so `per` is a power of two of at most PROJ
          const int per = out_dim > TPS ? out_dim / TPS : 1;
          const int base = t * per;
          double v[PROJ];
#pragma unroll
          for (int j = 0; j < PROJ; ++j) v[j] = (j < per && base + j < out_dim) ? term(base + j) : 0.0;
          // absent terms are +0.0, and x + 0.0 is x: the tree over `per`
          // terms is unchanged by the padding
#pragma unroll
          for (int w = 1; w < PROJ; w <<= 1)
#pragma unroll
            for (int j = 0; j + w < PROJ; j += 2 * w) v[j] = v[j] + v[j + w];
          part = v[0];
        }
        double pt = 0.0;
        if constexpr (TPS <= 64) {
          // lanes [g * TPS, (g + 1) * TPS) of one wave: the group's sum in
          // each of them
          pt = group_lane_sum<TPS>(part);
        } else {
          part = group_lane_sum<64>(part);
          if ((tid & 63) == 0) red[tid >> 6] = part;
          __syncthreads();
          double ws[WPG > 0 ? WPG : 1];
#pragma unroll
          for (int w = 0; w < WPG; ++w) ws[w] = red[g * WPG + w];
#pragma unroll
          for (int s = 1; s < WPG; s <<= 1)
#pragma unroll
            for (int w = 0; w + s < WPG; w += 2 * s) ws[w] = ws[w] + ws[w + s];
          pt = ws[0];
        }
        pt = clamp01(pt);
        if (st.op & kFlagExpVal) {
          // src/sampler/batch/active.rs:818-850: +-(1 - 2 P(true)), no collapse
          if (valid && L.want_probs && t == 0) L.probs[uint64_t(in.ordinal) * L.shots + local] = pt;
          if (expv) expv[st.a] = (eval(st) ? -1.0 : 1.0) * (1.0 - 2.0 * pt);
          group_sync<TPS>();  // every thread has read `red` before the next reduction
          continue;
        }
        if (valid) {
          if (L.want_probs && t == 0) L.probs[uint64_t(in.ordinal) * L.shots + local] = pt;
          if (in.op & kFlagForced) {
            // pinned: the XOR of earlier branches, no draw; only a fair coin
            bool f = (in.op & kFlagForcedOne) != 0;
            for (uint32_t j = 0; j < in.forced_len; ++j) f ^= bit_of(sbits, resid[in.forced_off + j]);
            branch = f;
            if (fabs(pt - 0.5) > kFairBranchTolerance) err |= kDevErrUnfairPin;
          } else if (L.injected) {
            double u = L.uniforms[uint64_t(in.ordinal) * L.shots + local];
            branch = pt <= 0.0 ? false : (pt >= 1.0 ? true : u < pt);
          } else if (!(pt >= 0.0 && pt <= 1.0)) {
            err |= kDevErrBadProbability;
          } else {
            branch = pt <= 0.0 ? false : (pt >= 1.0 ? true : sm_uniform(brng) < pt);
          }
        }
        double ps = branch ? pt : 1.0 - pt;
        if (!(ps > 0.0)) {
          if (valid) err |= kDevErrImpossibleBranch;
          ps = 1.0;
        }
        const double inv = 1.0 / sqrt(ps);
        // staged outputs j = 0, 1, ... while j * TPS < out_dim (a uniform
        // bound: the loops stop at the state's size, not at PROJ)
        double nr[PROJ], ni[PROJ];
#pragma unroll
        for (int j = 0; j < PROJ; ++j) {
          if (j * TPS >= out_dim) break;
          const int idx = t + j * TPS;
          nr[j] = 0.0;
          ni[j] = 0.0;
          if (idx < out_dim) {
            if (diag) {
              const uint32_t s = dsrc(uint32_t(idx), branch);
              nr[j] = re[s] * inv;
              ni[j] = im[s] * inv;
            } else {
              const uint32_t s0 = ins0(uint32_t(idx), piv), s1 = s0 ^ xm;
              bool odd = par32(s0 & zm);
              double dir = (branch != odd) ? -1.0 : 1.0;
              double c1r = dir * in.c1_re, c1i = dir * in.c1_im;
              double ar = kInvSqrt2 * re[s0] + c1r * re[s1] - c1i * im[s1];
              double ai = kInvSqrt2 * im[s0] + c1r * im[s1] + c1i * re[s1];
              nr[j] = ar * inv;
              ni[j] = ai * inv;
            }
          }
        }
        group_sync<TPS>();
#pragma unroll
        for (int j = 0; j < PROJ; ++j) {
          if (j * TPS >= out_dim) break;
          const int idx = t + j * TPS;
          if (idx < out_dim) {
            re[idx] = nr[j];
            im[idx] = ni[j];
          }
        }
      } else if (valid) {
        if (st.op & kFlagForced) {  // pinned: the XOR of earlier branches, no draw
          const uint32_t fop = instrs[st.full].op, foff = instrs[st.full].forced_off;
          const uint32_t flen = instrs[st.full].forced_len;
          bool f = (fop & kFlagForcedOne) != 0;
          for (uint32_t j = 0; j < flen; ++j) f ^= bit_of(sbits, resid[foff + j]);
          branch = f;
        } else if (L.injected) {
          const uint64_t words = (L.shots + 63) / 64;
          branch = (L.coins[uint64_t(st.full) * words + (local >> 6)] >> (local & 63)) & 1;
        } else {
          branch = sm_next(brng) & 1;
        }
      }
      // `branch` is the same in every thread of the shot: one lane per wave
      // writes the wave's copy
      if (bit_writer) set_bit(sbits, uint32_t(st.c), branch);
      bool b = eval(st);
      if (bit_writer) {
        if (st.a >= 0) set_bit(rbits, uint32_t(st.a), b);
        if (st.b >= 0) set_bit(sbits, uint32_t(st.b), b);
      }
      if (op == kOpMeasure) group_sync<TPS>();  // the collapsed state before its next reader
    }
  }
  if (!valid) return;
  if (err && t == 0) atomicOr(L.error, err);
  // postselection (src/sampler/batch/postselect.rs): a fired postselected
  // detector discards the shot, which is then not counted
  bool kept = true;
  for (int32_t d = 0; d < P.npost; ++d) kept = kept && !bit_of(dbits, P.post_dets[d]);
  if (L.keep && t == 0 && P.npost > 0) L.keep[local] = uint8_t(kept);

  for (int o = t; o < P.nobs && kept; o += TPS) {
    bool b = P.obs_flip[o] != 0;
    for (uint32_t j = P.obs_off[o]; j < P.obs_off[o + 1]; ++j) b ^= bit_of(rbits, P.obs_recs[j]);
    if (b) atomicAdd(&L.obs_ones[o], 1ull);
  }
  if (!L.keep_records) return;
  auto emit = [&](uint8_t* out, uint32_t row_bytes, auto&& bit_at, uint32_t ncols) {
    uint8_t* dst = out + local * row_bytes;
    for (uint32_t byte = t; byte < row_bytes; byte += TPS) {
      if (L.bit_packed) {
        uint32_t v = 0;
        for (uint32_t bb = 0; bb < 8; ++bb) {
          uint32_t col = byte * 8 + bb;
          if (col < ncols && bit_at(col)) v |= 1u << bb;
        }
        dst[byte] = uint8_t(v);
      } else {
        dst[byte] = uint8_t(bit_at(byte));
      }
    }
  };
  emit(L.meas, L.meas_row, [&](uint32_t c) { return bit_of(rbits, c); }, uint32_t(P.nrec));
  emit(L.det, L.det_row, [&](uint32_t c) { return bit_of(dbits, c); }, uint32_t(P.ndet));
  emit(
      L.obs, L.obs_row,
      [&](uint32_t o) {
        bool b = P.obs_flip[o] != 0;
        for (uint32_t j = P.obs_off[o]; j < P.obs_off[o + 1]; ++j) b ^= bit_of(rbits, P.obs_recs[j]);
        return b;
      },
      uint32_t(P.nobs));
}

// Narrow staging fits 80 VGPRs without spilling (six waves per SIMD), the
// full staging 96 (five).
template <int TPS, int PROJ>
__global__ __launch_bounds__(group_block(TPS)) __attribute__((amdgpu_waves_per_eu(PROJ == kProjNarrow ? 6 : 5)))
void group_tier_kernel(DevProgram P, DevLaunch L, GroupLayout Y, uint32_t region_bytes) {
  extern __shared__ __align__(16) unsigned char smem[];
  group_shots<TPS, PROJ>(P, L, Y, uint64_t(blockIdx.x) * (group_block(TPS) / TPS), smem, region_bytes);
}

// Persistent kernel over a batch: unit u of item i covers shots
// [(u - unit_off[i]) * SPU, ...) of that item, run one after another;
// workgroups take units from *counter until all `total_units` are done.
template <int TPS, int PROJ>
__global__ __launch_bounds__(group_block(TPS)) __attribute__((amdgpu_waves_per_eu(4))) void group_batch_kernel(
    const DevProgram* __restrict__ progs, const DevLaunch* __restrict__ launches, const GroupLayout* __restrict__ layouts,
                                                                  const uint64_t* unit_off, uint32_t nitems,
                                                                  uint64_t total_units, unsigned long long* counter,
                                                                  uint32_t region_bytes) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ unsigned long long unit_sh;
  __shared__ uint32_t item_sh;
  for (;;) {
    if (threadIdx.x == 0) {
      unsigned long long u = atomicAdd(counter, 1ull);
      uint32_t lo = 0, hi = nitems;  // last item with unit_off <= u
      while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2;
        if (unit_off[mid] <= u) lo = mid;
        else hi = mid;
      }
      unit_sh = u;
      item_sh = lo;
    }
    __syncthreads();
    // the same unit in every lane: make that visible to the compiler so the
    // per-item tables load through the scalar unit into SGPRs
    const uint64_t u = (uint64_t(__builtin_amdgcn_readfirstlane(uint32_t(unit_sh >> 32))) << 32) |
                       uint64_t(__builtin_amdgcn_readfirstlane(uint32_t(unit_sh)));
    const uint32_t item = __builtin_amdgcn_readfirstlane(item_sh);
    __syncthreads();
    if (u >= total_units) return;  // uniform: every thread read the same unit
    // fields are read where used (scalar loads), not copied into registers
    const DevProgram& P = progs[item];
    const DevLaunch& L = launches[item];
    const GroupLayout& Y = layouts[item];
    constexpr int SPU = group_unit_shots(TPS);
    constexpr int SPB = group_block(TPS) / TPS;
    const uint64_t first = (u - unit_off[item]) * SPU;
    for (int q = 0; q < SPU; q += SPB) {
      if (first + uint64_t(q) >= L.shots) break;  // uniform: one item per unit
      group_shots<TPS, PROJ>(P, L, Y, first + uint64_t(q), smem, region_bytes);
      __syncthreads();
    }
  }
}

__global__ void reset_counter_kernel(unsigned long long* counter) { *counter = 0ull; }

// Run before every replay of a batch graph: clears the error word and sets
// the seed offset the graph's launches read.
__global__ void batch_prologue_kernel(int* error, uint64_t* seed_offset, uint64_t value) {
  *error = 0;
  *seed_offset = value;
}

hipError_t launch_batch_prologue(int* error, uint64_t* seed_offset, uint64_t value, hipStream_t stream) {
  hipLaunchKernelGGL(batch_prologue_kernel, dim3(1), dim3(1), 0, stream, error, seed_offset, value);
  return hipGetLastError();
}

// --------------------------------------------------------------------------
// Host side
// --------------------------------------------------------------------------

size_t group_region_bytes(uint32_t state_dim, const GroupLayout& y, int tps) {
  return (group_shot_bytes(state_dim, y, tps) + 15) / 16 * 16;
}

int group_shots_per_block(int tps) { return group_block(tps) / tps; }
int group_shots_per_unit(int tps) { return group_unit_shots(tps); }

size_t group_tier_lds_bytes(uint32_t state_dim, const GroupLayout& y, int tps) {
  return 64 + size_t(group_shots_per_block(tps)) * group_region_bytes(state_dim, y, tps);
}

// Largest local width a shot group of `tps` threads projects through registers.
int group_max_local_k(int tps) {
  int k = 0;
  while ((size_t(1) << (k + 1)) <= size_t(2) * kProjPer * size_t(tps)) ++k;
  return k;
}

// Whether `max_local_k` fits the narrow projection staging at `tps`.
static bool group_narrow(int tps, int max_local_k) {
  return max_local_k <= 1 || (size_t(1) << (max_local_k - 1)) <= size_t(kProjNarrow) * size_t(tps);
}

template <int TPS, int PROJ>
static hipError_t launch_group_tp(const DevProgram& P, const DevLaunch& L, const GroupLayout& Y, hipStream_t stream) {
  size_t region = group_region_bytes(P.state_dim, Y, TPS);
  constexpr int SPB = group_block(TPS) / TPS;
  size_t lds = 64 + size_t(SPB) * region;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&group_tier_kernel<TPS, PROJ>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
  if (e != hipSuccess) return e;
  dim3 grid(unsigned((L.shots + SPB - 1) / SPB));
  hipLaunchKernelGGL((group_tier_kernel<TPS, PROJ>), grid, dim3(group_block(TPS)), lds, stream, P, L, Y,
                     uint32_t(region));
  return hipGetLastError();
}

template <int TPS>
static hipError_t launch_group_t(const DevProgram& P, const DevLaunch& L, const GroupLayout& Y, hipStream_t stream) {
  return group_narrow(TPS, P.max_local_k) ? launch_group_tp<TPS, kProjNarrow>(P, L, Y, stream)
                                          : launch_group_tp<TPS, kProjPer>(P, L, Y, stream);
}

hipError_t launch_group_tier(const DevProgram& P, const DevLaunch& L, const GroupLayout& Y, hipStream_t stream,
                             int tps) {
  if (P.max_local_k > group_max_local_k(tps)) return hipErrorInvalidValue;
  if (L.shots == 0) return hipSuccess;
  switch (tps) {
    case 8: return launch_group_t<8>(P, L, Y, stream);
    case 16: return launch_group_t<16>(P, L, Y, stream);
    case 32: return launch_group_t<32>(P, L, Y, stream);
    case 64: return launch_group_t<64>(P, L, Y, stream);
    case 128: return launch_group_t<128>(P, L, Y, stream);
    case 256: return launch_group_t<256>(P, L, Y, stream);
    case 512: return launch_group_t<512>(P, L, Y, stream);
    default: return hipErrorInvalidValue;
  }
}

template <int TPS, int PROJ>
static hipError_t launch_group_batch_tp(const DevProgram* progs, const DevLaunch* launches,
                                        const GroupLayout* layouts, const uint64_t* unit_off, uint32_t nitems,
                                        uint64_t total_units, unsigned long long* counter, uint32_t region_bytes,
                                        int grid, hipStream_t stream) {
  size_t lds = 64 + size_t(group_block(TPS) / TPS) * region_bytes;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&group_batch_kernel<TPS, PROJ>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, stream, counter);
  hipLaunchKernelGGL((group_batch_kernel<TPS, PROJ>), dim3(unsigned(grid)), dim3(group_block(TPS)), lds, stream,
                     progs, launches, layouts, unit_off, nitems, total_units, counter, region_bytes);
  return hipGetLastError();
}

template <int TPS>
static hipError_t launch_group_batch_t(const DevProgram* progs, const DevLaunch* launches, const GroupLayout* layouts,
                                       const uint64_t* unit_off, uint32_t nitems, uint64_t total_units,
                                       unsigned long long* counter, uint32_t region_bytes, int max_local_k, int grid,
                                       hipStream_t stream) {
  return group_narrow(TPS, max_local_k)
             ? launch_group_batch_tp<TPS, kProjNarrow>(progs, launches, layouts, unit_off, nitems, total_units,
                                                       counter, region_bytes, grid, stream)
             : launch_group_batch_tp<TPS, kProjPer>(progs, launches, layouts, unit_off, nitems, total_units, counter,
                                                    region_bytes, grid, stream);
}

hipError_t launch_group_batch(const DevProgram* progs, const DevLaunch* launches, const GroupLayout* layouts,
                              const uint64_t* unit_off, uint32_t nitems, uint64_t total_units,
                              unsigned long long* counter, uint32_t region_bytes, int tps, int max_local_k, int grid,
                              hipStream_t stream) {
  switch (tps) {
    case 8:
      return launch_group_batch_t<8>(progs, launches, layouts, unit_off, nitems, total_units, counter, region_bytes,
                                     max_local_k, grid, stream);
    case 16:
      return launch_group_batch_t<16>(progs, launches, layouts, unit_off, nitems, total_units, counter, region_bytes,
                                      max_local_k, grid, stream);
    case 32:
      return launch_group_batch_t<32>(progs, launches, layouts, unit_off, nitems, total_units, counter, region_bytes,
                                      max_local_k, grid, stream);
    case 64:
      return launch_group_batch_t<64>(progs, launches, layouts, unit_off, nitems, total_units, counter, region_bytes,
                                      max_local_k, grid, stream);
    case 128:
      return launch_group_batch_t<128>(progs, launches, layouts, unit_off, nitems, total_units, counter,
                                       region_bytes, max_local_k, grid, stream);
    case 256:
      return launch_group_batch_t<256>(progs, launches, layouts, unit_off, nitems, total_units, counter,
                                       region_bytes, max_local_k, grid, stream);
    case 512:
      return launch_group_batch_t<512>(progs, launches, layouts, unit_off, nitems, total_units, counter,
                                       region_bytes, max_local_k, grid, stream);
    default:
      return hipErrorInvalidValue;
  }
}

template <int PROJ>
static const void* group_batch_fn(int tps) {
  return tps == 8     ? reinterpret_cast<const void*>(&group_batch_kernel<8, PROJ>)
         : tps == 16  ? reinterpret_cast<const void*>(&group_batch_kernel<16, PROJ>)
         : tps == 32  ? reinterpret_cast<const void*>(&group_batch_kernel<32, PROJ>)
         : tps == 64  ? reinterpret_cast<const void*>(&group_batch_kernel<64, PROJ>)
         : tps == 128 ? reinterpret_cast<const void*>(&group_batch_kernel<128, PROJ>)
         : tps == 256 ? reinterpret_cast<const void*>(&group_batch_kernel<256, PROJ>)
                      : reinterpret_cast<const void*>(&group_batch_kernel<512, PROJ>);
}

// Workgroups of the persistent kernel that fit on the device at once.
int group_batch_occupancy(int tps, uint32_t region_bytes, int max_local_k) {
  size_t lds = 64 + size_t(group_shots_per_block(tps)) * region_bytes;
  int blocks = 0;
  const void* fn = group_narrow(tps, max_local_k) ? group_batch_fn<kProjNarrow>(tps) : group_batch_fn<kProjPer>(tps);
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fn, group_block(tps), lds) != hipSuccess) return 0;
  return blocks;
}

}  // namespace ticit
