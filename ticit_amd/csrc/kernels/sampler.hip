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

template <int KMAX>
__global__ __launch_bounds__(64) void lane_tier_kernel(DevProgram P, DevLaunch L, LaneLayout Y) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int DIM = 1 << KMAX;
  constexpr int SDIM = KMAX > 0 ? DIM / 2 : 0;
  const int lane = threadIdx.x;
  const uint64_t local = uint64_t(blockIdx.x) * 64 + lane;
  const bool live = local < L.shots;
  const uint64_t shot = L.first_shot + local;

  double2* amp = reinterpret_cast<double2*>(smem);  // [DIM][64]
  double2* scr = amp + DIM * 64;                    // [SDIM][64]
  uint64_t* pbits = reinterpret_cast<uint64_t*>(scr + SDIM * 64);
  uint64_t* sbits = pbits + Y.pw * 64;
  uint64_t* rbits = sbits + Y.sw * 64;
  uint64_t* dbits = rbits + Y.rw * 64;
  uint8_t* stage = reinterpret_cast<uint8_t*>(dbits + Y.dw * 64);

  for (uint32_t w = 0; w < Y.pw; ++w) pbits[w * 64 + lane] = 0;
  for (uint32_t w = 0; w < Y.sw; ++w) sbits[w * 64 + lane] = 0;
  for (uint32_t w = 0; w < Y.rw; ++w) rbits[w * 64 + lane] = 0;
  for (uint32_t w = 0; w < Y.dw; ++w) dbits[w * 64 + lane] = 0;

  // Frame: noise events toggle the exogenous partial bits they feed.
  auto toggle = [&](int32_t c) {
    uint32_t b = P.sym_off[c - 1], e = P.sym_off[c];
    for (uint32_t j = b; j < e; ++j) {
      uint32_t part = P.sym_part[j];
      pbits[(part >> 6) * 64 + lane] ^= 1ull << (part & 63);
    }
  };
  if (L.injected) {
    const uint64_t words = (L.shots + 63) / 64;
    for (int i = 0; i < P.nexo; ++i) {
      int32_t c = P.exo_syms[i];
      uint64_t wv = L.exo[uint64_t(c - 1) * words + blockIdx.x];
      if ((wv >> lane) & 1) toggle(c);
    }
  } else {
    sample_exogenous(P, seed_for(kExoSeedBase, launch_seed(L), shot), toggle);
  }

  for (int a = 0; a < DIM; ++a) amp[a * 64 + lane] = make_double2(a == 0 ? 1.0 : 0.0, 0.0);

  auto bit_of = [&](const uint64_t* tab, uint32_t idx) -> bool {
    return (tab[(idx >> 6) * 64 + lane] >> (idx & 63)) & 1;
  };
  auto set_bit = [&](uint64_t* tab, uint32_t idx, bool v) {
    uint64_t m = 1ull << (idx & 63);
    uint64_t& w = tab[(idx >> 6) * 64 + lane];
    w = v ? (w | m) : (w & ~m);
  };
  auto eval = [&](const DevInstr& in) -> bool {
    bool b = (in.op & kFlagConstant) != 0;
    if (in.partial >= 0) b ^= bit_of(pbits, uint32_t(in.partial));
    for (uint32_t j = 0; j < in.resid_len; ++j) b ^= bit_of(sbits, P.resid[in.resid_off + j]);
    return b;
  };
  auto write_outcome = [&](const DevInstr& in, bool b) {
    if (in.record >= 0) set_bit(rbits, uint32_t(in.record), b);
    if (in.rc_slot >= 0) set_bit(sbits, uint32_t(in.rc_slot), b);
  };

  int k = P.initial_k;
  uint64_t brng = seed_for(kBranchSeedBase, launch_seed(L), shot);
  int err = 0;
  for (int i = 0; i < P.ninstr; ++i) {
    const DevInstr in = P.instrs[i];
    const uint32_t op = in.op & 0xff;
    if (op == kOpRotate) {
      const bool sign = eval(in);
      const int dim = 1 << k;
      const double c = in.c;
      if (in.op & kFlagDiagonal) {
        for (int b = 0; b < dim; ++b) {
          bool flip = sign != parity64(uint64_t(b) & in.zmask);
          double cr = flip ? -in.me_re : in.me_re;
          double ci = flip ? -in.me_im : in.me_im;
          double fr = c + cr, fi = ci;
          double2 v = amp[b * 64 + lane];
          amp[b * 64 + lane] = make_double2(fr * v.x - fi * v.y, fr * v.y + fi * v.x);
        }
      } else if (in.op & kFlagUniformImag) {
        const double q = sign ? -in.me_im : in.me_im;
        const int sel = 1 << in.pivot, step = sel << 1;
        for (int block = 0; block < dim; block += step) {
          for (int off = 0; off < sel; ++off) {
            int i0 = block + off, i1 = i0 ^ int(in.xmask);
            double2 a0 = amp[i0 * 64 + lane], a1 = amp[i1 * 64 + lane];
            amp[i0 * 64 + lane] = make_double2(c * a0.x - q * a1.y, c * a0.y + q * a1.x);
            amp[i1 * 64 + lane] = make_double2(c * a1.x - q * a0.y, c * a1.y + q * a0.x);
          }
        }
      } else {
        const int pairs = dim >> 1;
        const bool xz_odd = in.op & kFlagXzOdd;
        for (int idx = 0; idx < pairs; ++idx) {
          int left = int(insert_zero_bit(size_t(idx), in.pivot)), right = left ^ int(in.xmask);
          bool lodd = parity64(uint64_t(left) & in.zmask);
          double ld = (sign != lodd) ? -1.0 : 1.0;
          double rd = xz_odd ? -ld : ld;
          double lre = ld * in.me_re, lim = ld * in.me_im;
          double rre = rd * in.me_re, rim = rd * in.me_im;
          double2 a0 = amp[left * 64 + lane], a1 = amp[right * 64 + lane];
          amp[left * 64 + lane] =
              make_double2(c * a0.x + rre * a1.x - rim * a1.y, c * a0.y + rre * a1.y + rim * a1.x);
          amp[right * 64 + lane] =
              make_double2(c * a1.x + lre * a0.x - lim * a0.y, c * a1.y + lre * a0.y + lim * a0.x);
        }
      }
    } else if (op == kOpPromote) {
      const bool sign = eval(in);
      const int dim = 1 << k;
      const double c = in.c, q = sign ? in.s : -in.s;
      for (int b = 0; b < dim; ++b) {
        double2 v = amp[b * 64 + lane];
        amp[b * 64 + lane] = make_double2(c * v.x, c * v.y);
        amp[(dim + b) * 64 + lane] = make_double2(-q * v.y, q * v.x);
      }
      ++k;
    } else if (op == kOpRecord) {
      write_outcome(in, eval(in));
    } else if (op == kOpDetector) {
      bool b = false;
      if (in.op & kFlagDetRecords) {
        for (uint32_t j = 0; j < in.det_len; ++j) b ^= bit_of(rbits, P.det_recs[in.det_off + j]);
      } else {
        b = eval(in);
      }
      set_bit(dbits, uint32_t(in.detector), b);
    } else if (op == kOpMeasure) {
      const int out_dim = (1 << k) >> 1;
      const bool diag = in.op & kFlagDiagonal;
      double pt = 0.0;
      if (diag) {
        for (int idx = 0; idx < out_dim; ++idx) {
          double2 v = amp[int(diagonal_source(in, size_t(idx), true)) * 64 + lane];
          pt += v.x * v.x + v.y * v.y;
        }
      } else {
        for (int idx = 0; idx < out_dim; ++idx) {
          int s0 = int(insert_zero_bit(size_t(idx), in.pivot)), s1 = s0 ^ int(in.xmask);
          double dir = parity64(uint64_t(s0) & in.zmask) ? 1.0 : -1.0;  // branch = true
          double c1r = dir * in.c1_re, c1i = dir * in.c1_im;
          double2 v0 = amp[s0 * 64 + lane], v1 = amp[s1 * 64 + lane];
          double ar = kInvSqrt2 * v0.x + c1r * v1.x - c1i * v1.y;
          double ai = kInvSqrt2 * v0.y + c1r * v1.y + c1i * v1.x;
          pt += ar * ar + ai * ai;
        }
      }
      pt = clamp01(pt);
      if (L.want_probs && live) L.probs[uint64_t(in.ordinal) * L.shots + local] = pt;
      bool branch;
      if (L.injected) {
        double u = L.uniforms[uint64_t(in.ordinal) * L.shots + local];
        branch = pt <= 0.0 ? false : (pt >= 1.0 ? true : u < pt);
      } else if (!(pt >= 0.0 && pt <= 1.0)) {
        err |= kDevErrBadProbability;
        branch = false;
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
        for (int idx = 0; idx < out_dim; ++idx) {
          double2 v = amp[int(diagonal_source(in, size_t(idx), branch)) * 64 + lane];
          amp[idx * 64 + lane] = make_double2(v.x * inv, v.y * inv);
        }
      } else {
        for (int idx = 0; idx < out_dim; ++idx) {
          int s0 = int(insert_zero_bit(size_t(idx), in.pivot)), s1 = s0 ^ int(in.xmask);
          bool odd = parity64(uint64_t(s0) & in.zmask);
          double dir = (branch != odd) ? -1.0 : 1.0;
          double c1r = dir * in.c1_re, c1i = dir * in.c1_im;
          double2 v0 = amp[s0 * 64 + lane], v1 = amp[s1 * 64 + lane];
          double ar = kInvSqrt2 * v0.x + c1r * v1.x - c1i * v1.y;
          double ai = kInvSqrt2 * v0.y + c1r * v1.y + c1i * v1.x;
          scr[idx * 64 + lane] = make_double2(ar * inv, ai * inv);
        }
        for (int idx = 0; idx < out_dim; ++idx) amp[idx * 64 + lane] = scr[idx * 64 + lane];
      }
      --k;
      set_bit(sbits, uint32_t(in.branch_slot), branch);
      write_outcome(in, eval(in));
    } else {  // kOpCoin
      bool coin;
      if (L.injected) {
        const uint64_t words = (L.shots + 63) / 64;
        coin = (L.coins[uint64_t(in.ordinal) * words + blockIdx.x] >> lane) & 1;
      } else {
        coin = sm_next(brng) & 1;  // one word per 64 shots, one live shot
      }
      set_bit(sbits, uint32_t(in.branch_slot), coin);
      write_outcome(in, eval(in));
    }
  }
  if (err && live) atomicOr(L.error, err);

  // Observables (src/sampler/prepared.rs:881-907) and their popcounts.
  const uint32_t nlive = uint32_t(min<uint64_t>(64, L.shots - uint64_t(blockIdx.x) * 64));
  uint8_t* row = stage + size_t(lane) * (Y.stage_bytes / 64);
  for (int o = 0; o < P.nobs; ++o) {
    bool b = false;
    for (uint32_t j = P.obs_off[o]; j < P.obs_off[o + 1]; ++j) b ^= bit_of(rbits, P.obs_recs[j]);
    unsigned long long ones = __popcll(__ballot(b && live));
    if (lane == 0 && ones) atomicAdd(&L.obs_ones[o], ones);
  }
  if (!L.keep_records) return;

  // Row packing (src/sampler/prepared.rs:909-941): a lane's record words
  // are its packed row in little-endian byte order; rows are staged in LDS
  // so the wave writes one contiguous, 16-byte-vectorized block per table.
  auto emit = [&](uint8_t* out, uint32_t row_bytes, auto&& bit_at, uint32_t ncols) {
    if (row_bytes == 0) return;
    if (L.bit_packed) {
      for (uint32_t byte = 0; byte < row_bytes; ++byte) {
        uint32_t v = 0;
        for (uint32_t bb = 0; bb < 8; ++bb) {
          uint32_t col = byte * 8 + bb;
          if (col < ncols && bit_at(col)) v |= 1u << bb;
        }
        row[byte] = uint8_t(v);
      }
    } else {
      for (uint32_t col = 0; col < ncols; ++col) row[col] = uint8_t(bit_at(col));
    }
    __syncthreads();
    const size_t total = size_t(nlive) * row_bytes;
    uint8_t* dst = out + (uint64_t(blockIdx.x) * 64) * row_bytes;
    const uint32_t pitch = Y.stage_bytes / 64;
    for (size_t b = lane; b < total; b += 64) {
      size_t r = b / row_bytes, cidx = b - r * row_bytes;
      dst[b] = stage[r * pitch + cidx];
    }
    __syncthreads();
  };
  emit(L.meas, L.meas_row, [&](uint32_t c) { return bit_of(rbits, c); }, uint32_t(P.nrec));
  emit(L.det, L.det_row, [&](uint32_t c) { return bit_of(dbits, c); }, uint32_t(P.ndet));
  emit(
      L.obs, L.obs_row,
      [&](uint32_t o) {
        bool b = false;
        for (uint32_t j = P.obs_off[o]; j < P.obs_off[o + 1]; ++j) b ^= bit_of(rbits, P.obs_recs[j]);
        return b;
      },
      uint32_t(P.nobs));
}

// --------------------------------------------------------------------------
// Host launchers
// --------------------------------------------------------------------------

size_t lane_tier_lds_bytes(int kmax, const LaneLayout& y) {
  size_t dim = size_t(1) << kmax;
  size_t sdim = kmax > 0 ? dim / 2 : 0;
  return (dim + sdim) * 64 * 16 + size_t(y.pw + y.sw + y.rw + y.dw) * 64 * 8 + y.stage_bytes;
}

template <int K>
static hipError_t launch_lane_k(const DevProgram& P, const DevLaunch& L, const LaneLayout& Y, hipStream_t stream) {
  size_t lds = lane_tier_lds_bytes(K, Y);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&lane_tier_kernel<K>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
  if (e != hipSuccess) return e;
  dim3 grid(unsigned((L.shots + 63) / 64));
  hipLaunchKernelGGL(lane_tier_kernel<K>, grid, dim3(64), lds, stream, P, L, Y);
  return hipGetLastError();
}

hipError_t launch_lane_tier(int kmax, const DevProgram& P, const DevLaunch& L, const LaneLayout& Y,
                            hipStream_t stream) {
  switch (kmax) {
    case 0: return launch_lane_k<0>(P, L, Y, stream);
    case 1: return launch_lane_k<1>(P, L, Y, stream);
    case 2: return launch_lane_k<2>(P, L, Y, stream);
    case 3: return launch_lane_k<3>(P, L, Y, stream);
    case 4: return launch_lane_k<4>(P, L, Y, stream);
    case 5: return launch_lane_k<5>(P, L, Y, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ticit
