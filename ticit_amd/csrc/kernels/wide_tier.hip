// Wide tier: active states too large for one workgroup's LDS (k > 12).
//
// A batch of S shots is in flight; shot j's split-complex state occupies
// re[j * DIM .. ], im[j * DIM .. ] in HBM (DIM = 2^max_k) with a ping-pong
// partner for projections. Per shot, a row of bit tables (partials, runtime
// slots, records, detectors) and a control record (branch stream, last
// branch, inverse norm) live in global memory. The host walks the program:
//   * runs of Record/Detector/Coin instructions -> one per-shot bit kernel;
//   * Rotate / Promote -> one streaming kernel over (tile, shot);
//   * Measure -> FP64 partial sums per (tile, shot), a per-shot decision
//     kernel (fixed-order sum, clamp, draw, 1/sqrt) and a projection kernel
//     into the partner buffer.
// Amplitudes are T = double (FP64) or T = float (FP32 amplitude updates with
// FP64 probability sums, renormalisation and collapse). Arithmetic follows
// ticit's contiguous kernels (src/contiguous.rs:93-177, 1088-1113,
// 1582-1606, 2013-2044).
#include <hip/hip_runtime.h>

#include <type_traits>
#include <vector>

#include "device_common.hpp"

namespace ticit {


struct WideCtl {
  uint64_t brng;
  double invnorm;
  int32_t branch;
  int32_t err;
  // a postselected detector fired: the shot is discarded and every later
  // kernel skips it (src/sampler/batch/postselect.rs drops it from the block)
  int32_t dead;
  int32_t pad;
};

constexpr int kWideThreads = 256;
constexpr int kWidePer = 8;                              // amplitudes per thread
constexpr int kWideTile = kWideThreads * kWidePer;       // amplitudes per block
constexpr int kWidePairs = kWidePer / 2;                 // adjacent pairs per thread

// Two adjacent amplitudes in one access (16 bytes FP64, 8 bytes FP32).
template <class T>
struct Pair2;
template <>
struct Pair2<double> {
  using V = double2;
};
template <>
struct Pair2<float> {
  using V = float2;
};
template <class T>
__device__ __forceinline__ typename Pair2<T>::V ld2(const T* p) {
  return *reinterpret_cast<const typename Pair2<T>::V*>(p);
}
template <class T>
__device__ __forceinline__ void st2(T* p, T a, T b) {
  typename Pair2<T>::V v;
  v.x = a;
  v.y = b;
  *reinterpret_cast<typename Pair2<T>::V*>(p) = v;
}
// pair p of a (tile, shot) block: elements 2p, 2p + 1
__device__ __forceinline__ size_t pair_index(int u) {
  return size_t(blockIdx.x) * (kWideTile / 2) + size_t(u) * kWideThreads + threadIdx.x;
}

__device__ __forceinline__ bool wbit(const uint64_t* t, uint32_t i) { return (t[i >> 6] >> (i & 63)) & 1; }
__device__ __forceinline__ void wset(uint64_t* t, uint32_t i, bool v) {
  uint64_t m = 1ull << (i & 63);
  t[i >> 6] = v ? (t[i >> 6] | m) : (t[i >> 6] & ~m);
}

struct WideShot {
  uint64_t* p;
  uint64_t* s;
  uint64_t* r;
  uint64_t* d;
};

__device__ __forceinline__ WideShot wide_shot(uint64_t* bits, const WideLayout& Y, uint32_t j) {
  WideShot w;
  w.p = bits + size_t(j) * Y.words;
  w.s = w.p + Y.pw;
  w.r = w.s + Y.sw;
  w.d = w.r + Y.rw;
  return w;
}

__device__ __forceinline__ bool wide_eval(const DevProgram& P, const DevInstr& in, const WideShot& w) {
  bool b = (in.op & kFlagConstant) != 0;
  if (in.partial >= 0) b ^= wbit(w.p, uint32_t(in.partial));
  for (uint32_t j = 0; j < in.resid_len; ++j) b ^= wbit(w.s, P.resid[in.resid_off + j]);
  return b;
}

// Per-shot setup: clear tables, walk the noise, seed the branch stream,
// reset the state to |0>.
template <class T>
__global__ __launch_bounds__(kWideThreads) void wide_init_kernel(DevProgram P, DevLaunch L, WideLayout Y,
                                                                 uint64_t batch_start, uint32_t nshots,
                                                                 uint64_t* bits, WideCtl* ctl, T* re, T* im,
                                                                 size_t dim) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nshots) return;
  const uint64_t local = batch_start + j;
  const uint64_t shot = L.first_shot + local;
  WideShot w = wide_shot(bits, Y, j);
  for (uint32_t k = 0; k < Y.words; ++k) w.p[k] = 0;
  auto toggle = [&](int32_t c) {
    for (uint32_t q = P.sym_off[c - 1]; q < P.sym_off[c]; ++q) {
      uint32_t part = P.sym_part[q];
      w.p[part >> 6] ^= 1ull << (part & 63);
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
                     [&](uint32_t part) { w.p[part >> 6] ^= 1ull << (part & 63); });
  }
  ctl[j].brng = L.reference ? 1ull : seed_for(kBranchSeedBase, launch_seed(L), shot);
  ctl[j].err = 0;
  ctl[j].branch = 0;
  ctl[j].invnorm = 1.0;
  ctl[j].dead = 0;
  const size_t d0 = size_t(1) << P.initial_k;
  for (size_t a = 0; a < d0; ++a) {
    re[j * dim + a] = a == 0 ? T(1) : T(0);
    im[j * dim + a] = T(0);
  }
}

// A run of Record / Detector / Coin instructions for every shot.
__global__ __launch_bounds__(kWideThreads) void wide_bits_kernel(DevProgram P, DevLaunch L, WideLayout Y,
                                                                 uint64_t batch_start, uint32_t nshots,
                                                                 uint64_t* bits, WideCtl* ctl, int first,
                                                                 int last) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nshots || ctl[j].dead) return;
  const uint64_t local = batch_start + j;
  WideShot w = wide_shot(bits, Y, j);
  uint64_t brng = ctl[j].brng;
  double* expv = L.expv ? L.expv + local * uint64_t(P.nexp) : nullptr;
  for (int i = first; i < last; ++i) {
    const DevInstr in = P.instrs[i];
    const uint32_t op = in.op & 0xff;
    if (in.op & kFlagExpVal) {
      // src/sampler/batch/runtime.rs:46-75,311-356: a deterministic probe is
      // +-1, an unbiased one 0 (no draw)
      if (expv) expv[in.record] = op == kOpCoin ? 0.0 : (wide_eval(P, in, w) ? -1.0 : 1.0);
      continue;
    }
    if (op == kOpRecord) {
      bool b = wide_eval(P, in, w);
      if (in.record >= 0) wset(w.r, uint32_t(in.record), b);
      if (in.rc_slot >= 0) wset(w.s, uint32_t(in.rc_slot), b);
    } else if (op == kOpDetector) {
      bool b = false;
      if (in.op & kFlagDetRecords) {
        b = (in.op & kFlagConstant) != 0;
        for (uint32_t q = 0; q < in.det_len; ++q) b ^= wbit(w.r, P.det_recs[in.det_off + q]);
      } else {
        b = wide_eval(P, in, w);
      }
      wset(w.d, uint32_t(in.detector), b);
      // the noiseless reference shot records every detector
      // (src/sampler/prepared.rs:355-405 runs it without postselection)
      if (b && (in.op & kFlagPostselect) && !L.reference) {
        ctl[j].dead = 1;
        break;
      }
    } else if (op == kOpCoin) {
      bool coin;
      if (in.op & kFlagForced) {  // pinned: the XOR of earlier branches, no draw
        coin = (in.op & kFlagForcedOne) != 0;
        for (uint32_t q = 0; q < in.forced_len; ++q) coin ^= wbit(w.s, P.resid[in.forced_off + q]);
      } else if (L.injected) {
        const uint64_t words = (L.shots + 63) / 64;
        coin = (L.coins[uint64_t(in.ordinal) * words + (local >> 6)] >> (local & 63)) & 1;
      } else {
        coin = sm_next(brng) & 1;
      }
      wset(w.s, uint32_t(in.branch_slot), coin);
      bool b = wide_eval(P, in, w);
      if (in.record >= 0) wset(w.r, uint32_t(in.record), b);
      if (in.rc_slot >= 0) wset(w.s, uint32_t(in.rc_slot), b);
    }
  }
  ctl[j].brng = brng;
}

template <class T>
__global__ __launch_bounds__(kWideThreads) void wide_rotate_kernel(DevProgram P, WideLayout Y, int instr,
                                                                   uint64_t* bits, const WideCtl* ctl, T* re, T* im,
                                                                   size_t dim, int k) {
  const uint32_t j = blockIdx.y;
  if (ctl[j].dead) return;
  const DevInstr in = P.instrs[instr];
  const bool sign = wide_eval(P, in, wide_shot(bits, Y, j));
  T* R = re + size_t(j) * dim + in.comp_off;
  T* I = im + size_t(j) * dim + in.comp_off;
  const size_t n = size_t(1) << k;
  const T c = T(in.c);
  if (in.op & kFlagDiagonal) {
    const T mr = T(in.me_re), mi = T(in.me_im);
    if (n >= 2 && !(dim & 1)) {
#pragma unroll
      for (int u = 0; u < kWidePairs; ++u) {
        const size_t b = 2 * pair_index(u);
        if (b >= n) break;
        const bool f0 = sign != parity64(uint64_t(b) & in.zmask);
        const bool f1 = f0 != bool(in.zmask & 1);
        const T fr0 = c + (f0 ? -mr : mr), fi0 = f0 ? -mi : mi;
        const T fr1 = c + (f1 ? -mr : mr), fi1 = f1 ? -mi : mi;
        const auto r = ld2(R + b), m = ld2(I + b);
        st2(R + b, fr0 * r.x - fi0 * m.x, fr1 * r.y - fi1 * m.y);
        st2(I + b, fr0 * m.x + fi0 * r.x, fr1 * m.y + fi1 * r.y);
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < kWidePer; ++q) {
      size_t b = size_t(blockIdx.x) * kWideTile + size_t(q) * kWideThreads + threadIdx.x;
      if (b >= n) break;
      bool flip = sign != parity64(uint64_t(b) & in.zmask);
      T fr = c + (flip ? -mr : mr), fi = flip ? -mi : mi;
      T r = R[b], m = I[b];
      R[b] = fr * r - fi * m;
      I[b] = fr * m + fi * r;
    }
    return;
  }
  const bool uni = in.op & kFlagUniformImag;
  const bool xz_odd = in.op & kFlagXzOdd;
  const T q = T(sign ? -in.me_im : in.me_im);
  const T mr = T(in.me_re), mi = T(in.me_im);
  if (in.pivot >= 1 && !(dim & 1)) {
    // pairs (idx, idx + 1) have adjacent left sources; the right sources are
    // the aligned pair at i1 & ~1, swapped when xmask has bit 0
    const bool swap = in.xmask & 1;
#pragma unroll
    for (int u = 0; u < kWidePairs; ++u) {
      const size_t idx = 2 * pair_index(u);
      if (idx >= (n >> 1)) break;
      const size_t i0 = insert_zero_bit(idx, in.pivot), i1 = (i0 ^ size_t(in.xmask)) & ~size_t(1);
      const auto a0 = ld2(R + i0), b0 = ld2(I + i0);
      auto a1 = ld2(R + i1), b1 = ld2(I + i1);
      if (swap) {
        T t = a1.x; a1.x = a1.y; a1.y = t;
        t = b1.x; b1.x = b1.y; b1.y = t;
      }
      T o0r[2], o0i[2], o1r[2], o1i[2];
      const T r0s[2] = {a0.x, a0.y}, m0s[2] = {b0.x, b0.y}, r1s[2] = {a1.x, a1.y}, m1s[2] = {b1.x, b1.y};
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const T r0 = r0s[e], m0 = m0s[e], r1 = r1s[e], m1 = m1s[e];
        if (uni) {
          o0r[e] = c * r0 - q * m1;
          o0i[e] = c * m0 + q * r1;
          o1r[e] = c * r1 - q * m0;
          o1i[e] = c * m1 + q * r0;
        } else {
          const bool lodd = parity64(uint64_t(i0 + size_t(e)) & in.zmask);
          const T ld = (sign != lodd) ? T(-1) : T(1);
          const T rd = xz_odd ? -ld : ld;
          const T lre = ld * mr, lim = ld * mi, rre = rd * mr, rim = rd * mi;
          o0r[e] = c * r0 + rre * r1 - rim * m1;
          o0i[e] = c * m0 + rre * m1 + rim * r1;
          o1r[e] = c * r1 + lre * r0 - lim * m0;
          o1i[e] = c * m1 + lre * m0 + lim * r0;
        }
      }
      st2(R + i0, o0r[0], o0r[1]);
      st2(I + i0, o0i[0], o0i[1]);
      if (swap) {
        st2(R + i1, o1r[1], o1r[0]);
        st2(I + i1, o1i[1], o1i[0]);
      } else {
        st2(R + i1, o1r[0], o1r[1]);
        st2(I + i1, o1i[0], o1i[1]);
      }
    }
    return;
  }
#pragma unroll
  for (int u = 0; u < kWidePer; ++u) {
    size_t idx = size_t(blockIdx.x) * kWideTile + size_t(u) * kWideThreads + threadIdx.x;
    if (idx >= (n >> 1)) break;
    size_t i0 = insert_zero_bit(idx, in.pivot), i1 = i0 ^ size_t(in.xmask);
    T r0 = R[i0], m0 = I[i0], r1 = R[i1], m1 = I[i1];
    if (uni) {
      R[i0] = c * r0 - q * m1;
      I[i0] = c * m0 + q * r1;
      R[i1] = c * r1 - q * m0;
      I[i1] = c * m1 + q * r0;
    } else {
      bool lodd = parity64(uint64_t(i0) & in.zmask);
      T ld = (sign != lodd) ? T(-1) : T(1);
      T rd = xz_odd ? -ld : ld;
      T lre = ld * mr, lim = ld * mi, rre = rd * mr, rim = rd * mi;
      R[i0] = c * r0 + rre * r1 - rim * m1;
      I[i0] = c * m0 + rre * m1 + rim * r1;
      R[i1] = c * r1 + lre * r0 - lim * m0;
      I[i1] = c * m1 + lre * m0 + lim * r0;
    }
  }
}

template <class T>
__global__ __launch_bounds__(kWideThreads) void wide_promote_kernel(DevProgram P, WideLayout Y, int instr,
                                                                    uint64_t* bits, const WideCtl* ctl, T* re, T* im,
                                                                    size_t dim, int k) {
  const uint32_t j = blockIdx.y;
  if (ctl[j].dead) return;
  const DevInstr in = P.instrs[instr];
  const bool sign = wide_eval(P, in, wide_shot(bits, Y, j));
  T* R = re + size_t(j) * dim + in.comp_off;
  T* I = im + size_t(j) * dim + in.comp_off;
  const size_t n = size_t(1) << k;
  const T c = T(in.c), q = T(sign ? in.s : -in.s);
  if (n >= 2 && !(dim & 1)) {
#pragma unroll
    for (int u = 0; u < kWidePairs; ++u) {
      const size_t b = 2 * pair_index(u);
      if (b >= n) break;
      const auto r = ld2(R + b), m = ld2(I + b);
      st2(R + b, c * r.x, c * r.y);
      st2(I + b, c * m.x, c * m.y);
      st2(R + n + b, -q * m.x, -q * m.y);
      st2(I + n + b, q * r.x, q * r.y);
    }
    return;
  }
#pragma unroll
  for (int u = 0; u < kWidePer; ++u) {
    size_t b = size_t(blockIdx.x) * kWideTile + size_t(u) * kWideThreads + threadIdx.x;
    if (b >= n) break;
    T r = R[b], m = I[b];
    R[b] = c * r;
    I[b] = c * m;
    R[n + b] = -q * m;
    I[n + b] = q * r;
  }
}

// Consecutive promotions of one component, widths k0, k0 + 1, ...,
// k0 + m - 1 (m <= kWideMaxPromoteRun), in one pass into the partner
// buffer. Promotion j extends the state by the one-qubit factor
// (c_j, i q_j), so output b takes input b mod 2^k0 and, for j = 0, 1, ...
// in turn, the factor's arithmetic of wide_promote_kernel in T (c_j times
// both parts on bit k0 + j clear; (-q_j m, q_j r) on it set): the values of
// the promotions run one after another, with one read of the old state and
// one write of the new instead of a read and two writes per promotion.
constexpr int kWideMaxPromoteRun = 16;

template <class T>
__global__ __launch_bounds__(kWideThreads) void wide_promote_run_kernel(DevProgram P, WideLayout Y, int instr, int m,
                                                                        uint64_t* bits, const WideCtl* ctl,
                                                                        const T* re, const T* im, T* ore, T* oim,
                                                                        size_t dim, int k0) {
  const uint32_t j = blockIdx.y;
  if (ctl[j].dead) return;
  __shared__ T cs[kWideMaxPromoteRun], qs[kWideMaxPromoteRun];
  if (int(threadIdx.x) < m) {
    const DevInstr in = P.instrs[instr + int(threadIdx.x)];
    const bool sign = wide_eval(P, in, wide_shot(bits, Y, j));
    cs[threadIdx.x] = T(in.c);
    qs[threadIdx.x] = T(sign ? in.s : -in.s);
  }
  __syncthreads();
  const size_t off = size_t(j) * dim + P.instrs[instr].comp_off;
  const T* R = re + off;
  const T* I = im + off;
  T* OR = ore + off;
  T* OI = oim + off;
  const size_t n0 = size_t(1) << k0, total = n0 << m;
#pragma unroll
  for (int u = 0; u < kWidePer; ++u) {
    const size_t b = size_t(blockIdx.x) * kWideTile + size_t(u) * kWideThreads + threadIdx.x;
    if (b >= total) break;
    T r = R[b & (n0 - 1)], mi = I[b & (n0 - 1)];
    for (int q = 0; q < m; ++q) {
      if ((b >> (k0 + q)) & 1) {
        const T nr = -qs[q] * mi, ni = qs[q] * r;
        r = nr;
        mi = ni;
      } else {
        r = cs[q] * r;
        mi = cs[q] * mi;
      }
    }
    OR[b] = r;
    OI[b] = mi;
  }
}

// A promotion that starts a new one-qubit component, or an initial |0>
// component (src/sampler/batch/runtime.rs:511-550, batch/mod.rs:376-402).
template <class T>
__global__ __launch_bounds__(kWideThreads) void wide_fresh_kernel(DevProgram P, WideLayout Y, int instr,
                                                                  uint32_t nshots, uint64_t* bits,
                                                                  const WideCtl* ctl, T* re, T* im, size_t dim) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nshots || ctl[j].dead) return;
  const DevInstr in = P.instrs[instr];
  const bool zero = in.op & kFlagZeroState;
  const bool sign = zero ? false : wide_eval(P, in, wide_shot(bits, Y, j));
  T* R = re + size_t(j) * dim + in.comp_off;
  T* I = im + size_t(j) * dim + in.comp_off;
  R[0] = zero ? T(1) : T(in.c);
  I[0] = T(0);
  R[1] = T(0);
  I[1] = zero ? T(0) : T(sign ? in.s : -in.s);
}

// target (x) source of src/sampler/batch/mod.rs:597-680, written into the
// target's partner buffer: out[(sb << tk) + tb] = t[tb] * s[sb].
template <class T>
__global__ __launch_bounds__(kWideThreads) void wide_merge_kernel(DevProgram P, int instr, const WideCtl* ctl,
                                                                  const T* tre, const T* tim, const T* sre,
                                                                  const T* sim, T* ore, T* oim, size_t dim) {
  const uint32_t j = blockIdx.y;
  if (ctl[j].dead) return;
  const DevInstr in = P.instrs[instr];
  const size_t tk = size_t(in.nq), td = size_t(1) << tk, total = size_t(1) << (tk + size_t(in.src_k));
  const T* TR = tre + size_t(j) * dim + in.comp_off;
  const T* TI = tim + size_t(j) * dim + in.comp_off;
  const T* SR = sre + size_t(j) * dim + in.src_off;
  const T* SI = sim + size_t(j) * dim + in.src_off;
  T* OR = ore + size_t(j) * dim + in.comp_off;
  T* OI = oim + size_t(j) * dim + in.comp_off;
  if (!(dim & 1)) {  // td >= 2: outputs (o, o + 1) share the source entry
#pragma unroll
    for (int u = 0; u < kWidePairs; ++u) {
      const size_t o = 2 * pair_index(u);
      if (o >= total) break;
      const size_t tb = o & (td - 1), sb = o >> tk;
      const auto a = ld2(TR + tb), b = ld2(TI + tb);
      const T cr = SR[sb], ci = SI[sb];
      st2(OR + o, a.x * cr - b.x * ci, a.y * cr - b.y * ci);
      st2(OI + o, a.x * ci + b.x * cr, a.y * ci + b.y * cr);
    }
    return;
  }
#pragma unroll
  for (int u = 0; u < kWidePer; ++u) {
    size_t o = size_t(blockIdx.x) * kWideTile + size_t(u) * kWideThreads + threadIdx.x;
    if (o >= total) break;
    size_t tb = o & (td - 1), sb = o >> tk;
    T a = TR[tb], b = TI[tb], cr = SR[sb], ci = SI[sb];
    OR[o] = a * cr - b * ci;
    OI[o] = a * ci + b * cr;
  }
}

// Two block sums at once (their reduction slots are separate).
__device__ __forceinline__ void block_sum2(double& a, double& b, double (*red)[kWideThreads / 64]) {
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = a;
    red[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  a = 0.0;
  b = 0.0;
  for (int w = 0; w < kWideThreads / 64; ++w) {
    a += red[0][w];
    b += red[1][w];
  }
}

// Whether a measurement's outputs (idx, idx + 1), idx even, read adjacent,
// aligned source pairs: the pivot is not bit 0 and, for a diagonal body,
// the parity mask does not include bit 0 (so both outputs pick the same
// half); the state dimension is even so pairs stay aligned.
__device__ __forceinline__ bool pairs_ok(const DevInstr& in, size_t out_dim, size_t dim) {
  if (out_dim < 2 || (dim & 1) || in.pivot < 1) return false;
  return !(in.op & kFlagDiagonal) || !(in.zwop & 1);
}

// FP64 partial sums of P(branch = 1) per (tile, shot). With `npart` (the
// FP32 path) also the partial sums of the component's norm, over every
// amplitude it holds, so the decision renormalizes by the stored state's
// own norm instead of assuming it is 1.
template <class T>
__global__ __launch_bounds__(kWideThreads) void wide_prob_kernel(DevProgram P, int instr, const WideCtl* ctl,
                                                                 const T* re, const T* im, size_t dim, int k,
                                                                 double* part, double* npart, uint32_t max_tiles) {
  __shared__ double red[2][kWideThreads / 64];
  const uint32_t j = blockIdx.y;
  if (ctl[j].dead) return;
  const DevInstr in = P.instrs[instr];
  const T* R = re + size_t(j) * dim + in.comp_off;
  const T* I = im + size_t(j) * dim + in.comp_off;
  const size_t out_dim = (size_t(1) << k) >> 1;
  const bool diag = in.op & kFlagDiagonal;
  const bool norm = npart != nullptr;
  double acc = 0.0, nacc = 0.0;
  auto sq = [](double r, double m) { return r * r + m * m; };
  if (pairs_ok(in, out_dim, dim)) {
#pragma unroll
    for (int u = 0; u < kWidePairs; ++u) {
      const size_t idx = 2 * pair_index(u);
      if (idx >= out_dim) break;
      if (diag) {
        const size_t s0 = diagonal_source(in, idx, true);  // s0 + 1 is the source of idx + 1
        const auto r = ld2(R + s0), m = ld2(I + s0);
        acc += double(r.x) * double(r.x) + double(m.x) * double(m.x);
        acc += double(r.y) * double(r.y) + double(m.y) * double(m.y);
        if (norm) {  // the other half: the pivot bit flipped (pivot >= 1 keeps the pair)
          const size_t c0 = s0 ^ (size_t(1) << in.pivot);
          const auto rc = ld2(R + c0), mc = ld2(I + c0);
          nacc += sq(double(r.x), double(m.x)) + sq(double(r.y), double(m.y)) + sq(double(rc.x), double(mc.x)) +
                  sq(double(rc.y), double(mc.y));
        }
      } else {
        const size_t s0 = insert_zero_bit(idx, in.pivot), s1 = (s0 ^ size_t(in.xmask)) & ~size_t(1);
        const auto a0 = ld2(R + s0), b0 = ld2(I + s0), a1 = ld2(R + s1), b1 = ld2(I + s1);
        if (norm)  // s0 and s0 ^ xmask cover every amplitude once
          nacc += sq(double(a0.x), double(b0.x)) + sq(double(a0.y), double(b0.y)) + sq(double(a1.x), double(b1.x)) +
                  sq(double(a1.y), double(b1.y));
        const bool swap = in.xmask & 1;
        const double r0[2] = {double(a0.x), double(a0.y)}, m0[2] = {double(b0.x), double(b0.y)};
        const double r1[2] = {double(swap ? a1.y : a1.x), double(swap ? a1.x : a1.y)};
        const double m1[2] = {double(swap ? b1.y : b1.x), double(swap ? b1.x : b1.y)};
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const double dir = parity64(uint64_t(s0 + size_t(e)) & in.zmask) ? 1.0 : -1.0;
          const double c1r = dir * in.c1_re, c1i = dir * in.c1_im;
          const double ar = kInvSqrt2 * r0[e] + c1r * r1[e] - c1i * m1[e];
          const double ai = kInvSqrt2 * m0[e] + c1r * m1[e] + c1i * r1[e];
          acc += ar * ar + ai * ai;
        }
      }
    }
    block_sum2(acc, nacc, red);
    if (threadIdx.x == 0) {
      part[size_t(j) * max_tiles + blockIdx.x] = acc;
      if (norm) npart[size_t(j) * max_tiles + blockIdx.x] = nacc;
    }
    return;
  }
#pragma unroll
  for (int u = 0; u < kWidePer; ++u) {
    size_t idx = size_t(blockIdx.x) * kWideTile + size_t(u) * kWideThreads + threadIdx.x;
    if (idx >= out_dim) break;
    if (diag) {
      size_t s = diagonal_source(in, idx, true);
      double r = double(R[s]), m = double(I[s]);
      acc += r * r + m * m;
      if (norm) {
        const size_t c = diagonal_source(in, idx, false);
        nacc += sq(r, m) + sq(double(R[c]), double(I[c]));
      }
    } else {
      size_t s0 = insert_zero_bit(idx, in.pivot), s1 = s0 ^ size_t(in.xmask);
      double dir = parity64(uint64_t(s0) & in.zmask) ? 1.0 : -1.0;
      double c1r = dir * in.c1_re, c1i = dir * in.c1_im;
      double ar = kInvSqrt2 * double(R[s0]) + c1r * double(R[s1]) - c1i * double(I[s1]);
      double ai = kInvSqrt2 * double(I[s0]) + c1r * double(I[s1]) + c1i * double(R[s1]);
      acc += ar * ar + ai * ai;
      if (norm) nacc += sq(double(R[s0]), double(I[s0])) + sq(double(R[s1]), double(I[s1]));
    }
  }
  block_sum2(acc, nacc, red);
  if (threadIdx.x == 0) {
    part[size_t(j) * max_tiles + blockIdx.x] = acc;
    if (norm) npart[size_t(j) * max_tiles + blockIdx.x] = nacc;
  }
}

// Per shot, given the summed P(branch = 1) and the measured component's
// norm `nrm`: P(true) = pt / nrm clamped, draw, 1/sqrt(P(branch) * nrm),
// branch symbol, outcome, record (or the expectation of a probe). The FP64
// path passes nrm = 1, which leaves pt and P(branch) bit for bit as the
// reference forms them; the FP32 path passes the stored state's norm, so
// the collapse renormalizes the state to 1 and rounding accumulated in the
// stored amplitudes does not grow through the 1 / P(branch) scaling.
__device__ void wide_decide_shot(const DevProgram& P, const DevLaunch& L, const WideLayout& Y, uint32_t j,
                                 uint64_t local, const DevInstr& in, double pt, double nrm, uint64_t* bits,
                                 WideCtl* ctl) {
  if (ctl[j].dead) return;
  pt = clamp01(pt / nrm);
  if (L.want_probs) L.probs[uint64_t(in.ordinal) * L.shots + local] = pt;
  if (in.op & kFlagExpVal) {
    // src/sampler/batch/active.rs:818-850: +-(1 - 2 P(true)), no draw, no
    // collapse
    const double sign = wide_eval(P, in, wide_shot(bits, Y, j)) ? -1.0 : 1.0;
    if (L.expv) L.expv[local * uint64_t(P.nexp) + uint32_t(in.record)] = sign * (1.0 - 2.0 * pt);
    return;
  }
  WideCtl c = ctl[j];
  bool branch = false;
  if (in.op & kFlagForced) {
    // pinned: the XOR of earlier branches, no draw; only a fair coin
    const WideShot w0 = wide_shot(bits, Y, j);
    bool f = (in.op & kFlagForcedOne) != 0;
    for (uint32_t q = 0; q < in.forced_len; ++q) f ^= wbit(w0.s, P.resid[in.forced_off + q]);
    branch = f;
    if (fabs(pt - 0.5) > kFairBranchTolerance) c.err |= kDevErrUnfairPin;
  } else if (L.injected) {
    double u = L.uniforms[uint64_t(in.ordinal) * L.shots + local];
    branch = pt <= 0.0 ? false : (pt >= 1.0 ? true : u < pt);
  } else if (!(pt >= 0.0 && pt <= 1.0)) {
    c.err |= kDevErrBadProbability;
  } else {
    branch = pt <= 0.0 ? false : (pt >= 1.0 ? true : sm_uniform(c.brng) < pt);
  }
  double ps = branch ? pt : 1.0 - pt;
  if (!(ps > 0.0)) {
    c.err |= kDevErrImpossibleBranch;
    ps = 1.0;
  }
  c.invnorm = 1.0 / sqrt(ps * nrm);
  c.branch = branch;
  ctl[j] = c;
  WideShot w = wide_shot(bits, Y, j);
  wset(w.s, uint32_t(in.branch_slot), branch);
  bool b = wide_eval(P, in, w);
  if (in.record >= 0) wset(w.r, uint32_t(in.record), b);
  if (in.rc_slot >= 0) wset(w.s, uint32_t(in.rc_slot), b);
}

// Per shot: fixed-order sum of the tile partials, then the decision.
__global__ __launch_bounds__(kWideThreads) void wide_decide_kernel(DevProgram P, DevLaunch L, WideLayout Y,
                                                                   uint64_t batch_start, uint32_t nshots,
                                                                   int instr, uint64_t* bits, WideCtl* ctl,
                                                                   const double* part, const double* npart,
                                                                   uint32_t ntiles, uint32_t max_tiles) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nshots) return;
  const DevInstr in = P.instrs[instr];
  double pt = 0.0, nrm = npart ? 0.0 : 1.0;
  for (uint32_t t = 0; t < ntiles; ++t) pt += part[size_t(j) * max_tiles + t];
  if (npart)
    for (uint32_t t = 0; t < ntiles; ++t) nrm += npart[size_t(j) * max_tiles + t];
  wide_decide_shot(P, L, Y, j, batch_start + j, in, pt, nrm, bits, ctl);
}

// Collapse into the partner buffer (the renormalising multiply in FP64).
template <class T>
__global__ __launch_bounds__(kWideThreads) void wide_project_kernel(DevProgram P, int instr, const WideCtl* ctl,
                                                                    const T* re, const T* im, T* ore, T* oim,
                                                                    size_t dim, int k) {
  const uint32_t j = blockIdx.y;
  if (ctl[j].dead) return;
  const DevInstr in = P.instrs[instr];
  const bool branch = ctl[j].branch;
  const double inv = ctl[j].invnorm;
  const T* R = re + size_t(j) * dim + in.comp_off;
  const T* I = im + size_t(j) * dim + in.comp_off;
  T* OR = ore + size_t(j) * dim + in.comp_off;
  T* OI = oim + size_t(j) * dim + in.comp_off;
  const size_t out_dim = (size_t(1) << k) >> 1;
  const bool diag = in.op & kFlagDiagonal;
  if (pairs_ok(in, out_dim, dim)) {
#pragma unroll
    for (int u = 0; u < kWidePairs; ++u) {
      const size_t idx = 2 * pair_index(u);
      if (idx >= out_dim) break;
      if (diag) {
        const size_t s0 = diagonal_source(in, idx, branch);  // s0 + 1 feeds idx + 1
        const auto r = ld2(R + s0), m = ld2(I + s0);
        st2(OR + idx, T(double(r.x) * inv), T(double(r.y) * inv));
        st2(OI + idx, T(double(m.x) * inv), T(double(m.y) * inv));
      } else {
        const size_t s0 = insert_zero_bit(idx, in.pivot), s1 = (s0 ^ size_t(in.xmask)) & ~size_t(1);
        const auto a0 = ld2(R + s0), b0 = ld2(I + s0), a1 = ld2(R + s1), b1 = ld2(I + s1);
        const bool swap = in.xmask & 1;
        const double r0[2] = {double(a0.x), double(a0.y)}, m0[2] = {double(b0.x), double(b0.y)};
        const double r1[2] = {double(swap ? a1.y : a1.x), double(swap ? a1.x : a1.y)};
        const double m1[2] = {double(swap ? b1.y : b1.x), double(swap ? b1.x : b1.y)};
        T outr[2], outi[2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const bool odd = parity64(uint64_t(s0 + size_t(e)) & in.zmask);
          const double dir = (branch != odd) ? -1.0 : 1.0;
          const double c1r = dir * in.c1_re, c1i = dir * in.c1_im;
          const double ar = kInvSqrt2 * r0[e] + c1r * r1[e] - c1i * m1[e];
          const double ai = kInvSqrt2 * m0[e] + c1r * m1[e] + c1i * r1[e];
          outr[e] = T(ar * inv);
          outi[e] = T(ai * inv);
        }
        st2(OR + idx, outr[0], outr[1]);
        st2(OI + idx, outi[0], outi[1]);
      }
    }
    return;
  }
#pragma unroll
  for (int u = 0; u < kWidePer; ++u) {
    size_t idx = size_t(blockIdx.x) * kWideTile + size_t(u) * kWideThreads + threadIdx.x;
    if (idx >= out_dim) break;
    if (diag) {
      size_t s = diagonal_source(in, idx, branch);
      OR[idx] = T(double(R[s]) * inv);
      OI[idx] = T(double(I[s]) * inv);
    } else {
      size_t s0 = insert_zero_bit(idx, in.pivot), s1 = s0 ^ size_t(in.xmask);
      bool odd = parity64(uint64_t(s0) & in.zmask);
      double dir = (branch != odd) ? -1.0 : 1.0;
      double c1r = dir * in.c1_re, c1i = dir * in.c1_im;
      double ar = kInvSqrt2 * double(R[s0]) + c1r * double(R[s1]) - c1i * double(I[s1]);
      double ai = kInvSqrt2 * double(I[s0]) + c1r * double(I[s1]) + c1i * double(R[s1]);
      OR[idx] = T(ar * inv);
      OI[idx] = T(ai * inv);
    }
  }
}

// Collapse of measurement `instr` fused with the probability pass of the
// measurement `next` that follows it on the same component (width k - 1):
// block (b, shot) covers next's output pairs [b * kWideTile / 2, ...); for
// each it forms the two collapsed amplitudes next pairs up (s0', s0' ^ x'
// for a nondiagonal body, both halves' sources for a diagonal one), stores
// them, and adds next's P(branch = 1) term from the stored values, so the
// collapsed state is not read back for the next probability.
template <class T>
__global__ __launch_bounds__(kWideThreads) void wide_project_measure_kernel(DevProgram P, int instr, int next,
                                                                            const WideCtl* ctl, const T* re,
                                                                            const T* im, T* ore, T* oim,
                                                                            size_t dim, int k, double* part,
                                                                            double* npart, uint32_t max_tiles) {
  __shared__ double red[2][kWideThreads / 64];
  const uint32_t j = blockIdx.y;
  if (ctl[j].dead) return;
  const DevInstr in = P.instrs[instr];
  const DevInstr nx = P.instrs[next];
  const bool branch = ctl[j].branch;
  const double inv = ctl[j].invnorm;
  const T* R = re + size_t(j) * dim + in.comp_off;
  const T* I = im + size_t(j) * dim + in.comp_off;
  T* OR = ore + size_t(j) * dim + in.comp_off;
  T* OI = oim + size_t(j) * dim + in.comp_off;
  const bool diag = in.op & kFlagDiagonal, ndiag = nx.op & kFlagDiagonal;
  // collapsed amplitude idx of `in`, rounded as stored
  auto collapsed = [&](size_t idx, T& r, T& m) {
    if (diag) {
      const size_t s = diagonal_source(in, idx, branch);
      r = T(double(R[s]) * inv);
      m = T(double(I[s]) * inv);
    } else {
      const size_t s0 = insert_zero_bit(idx, in.pivot), s1 = s0 ^ size_t(in.xmask);
      const bool odd = parity64(uint64_t(s0) & in.zmask);
      const double dir = (branch != odd) ? -1.0 : 1.0;
      const double c1r = dir * in.c1_re, c1i = dir * in.c1_im;
      const double ar = kInvSqrt2 * double(R[s0]) + c1r * double(R[s1]) - c1i * double(I[s1]);
      const double ai = kInvSqrt2 * double(I[s0]) + c1r * double(I[s1]) + c1i * double(R[s1]);
      r = T(ar * inv);
      m = T(ai * inv);
    }
  };
  const size_t pairs = ((size_t(1) << k) >> 1) >> 1;  // next's output half
  double acc = 0.0, nacc = 0.0;
#pragma unroll
  for (int u = 0; u < kWidePairs; ++u) {
    const size_t q = size_t(blockIdx.x) * (kWideTile / 2) + size_t(u) * kWideThreads + threadIdx.x;
    if (q >= pairs) break;
    size_t a, b;  // a: the amplitude next's P(true) reads (nondiagonal: with b)
    if (ndiag) {
      a = diagonal_source(nx, q, true);
      b = diagonal_source(nx, q, false);
    } else {
      a = insert_zero_bit(q, nx.pivot);
      b = a ^ size_t(nx.xmask);
    }
    T ra, ma, rb, mb;
    collapsed(a, ra, ma);
    collapsed(b, rb, mb);
    OR[a] = ra;
    OI[a] = ma;
    OR[b] = rb;
    OI[b] = mb;
    // a and b over every q cover the collapsed state once: its norm, from
    // the stored values
    if (npart) nacc += double(ra) * double(ra) + double(ma) * double(ma) + double(rb) * double(rb) +
                       double(mb) * double(mb);
    if (ndiag) {
      acc += double(ra) * double(ra) + double(ma) * double(ma);
    } else {
      const double dir = parity64(uint64_t(a) & nx.zmask) ? 1.0 : -1.0;
      const double c1r = dir * nx.c1_re, c1i = dir * nx.c1_im;
      const double ar = kInvSqrt2 * double(ra) + c1r * double(rb) - c1i * double(mb);
      const double ai = kInvSqrt2 * double(ma) + c1r * double(mb) + c1i * double(rb);
      acc += ar * ar + ai * ai;
    }
  }
  block_sum2(acc, nacc, red);
  if (threadIdx.x == 0) {
    part[size_t(j) * max_tiles + blockIdx.x] = acc;
    if (npart) npart[size_t(j) * max_tiles + blockIdx.x] = nacc;
  }
}

// ---- Factored components -------------------------------------------------
// A component about to be merged from factors f0 (target) and f1..fn-1
// (sources, in merge order): amplitude x is
// ((f0[x0] * f1[x1]) * f2[x2]) ..., x_f = (x >> shift_f) & mask_f, each
// product in the order and formula of a merge (an FP64 state's values
// equal the materialized state's bit for bit). A merge chain followed by a
// measurement of the merged component never writes the merged state: its
// probability comes from the factors' expectations, its projection from
// the factors (which stay in cache), and only the projected half is
// written.
constexpr int kMaxFactors = 6;

template <class T>
struct Factors {
  int n;
  const T* re[kMaxFactors];  // shot 0 base + component offset
  const T* im[kMaxFactors];
  int shift[kMaxFactors];
  uint32_t mask[kMaxFactors];
  int tile0[kMaxFactors + 1];  // expectation pass: first block of each factor
};

// Tile table of a factor chain for the expectation pass and its decision.
struct FactorTiles {
  int n;
  int tile0[kMaxFactors + 1];
};

// A = T forms the product as a merge pass stores it; A = double keeps a
// never-stored merged amplitude unrounded (FP32 storage rounds only what
// it writes).
template <class T, class A>
__device__ __forceinline__ void factor_amp(const Factors<T>& F, size_t shot, size_t x, A& r, A& m) {
  const size_t x0 = x & F.mask[0];
  r = A(F.re[0][shot + x0]);
  m = A(F.im[0][shot + x0]);
  for (int f = 1; f < F.n; ++f) {
    const size_t xf = (x >> F.shift[f]) & F.mask[f];
    const A cr = A(F.re[f][shot + xf]), ci = A(F.im[f][shot + xf]);
    const A nr = r * cr - m * ci, ni = r * ci + m * cr;
    r = nr;
    m = ni;
  }
}

// The merged amplitudes at x and x + 1 (x even): they share every factor
// but the first, whose two entries are one aligned pair; each product in
// the order of factor_amp.
template <class T>
__device__ __forceinline__ void factor_amp2(const Factors<T>& F, size_t shot, size_t x, double& r0, double& m0,
                                            double& r1, double& m1) {
  const size_t x0 = x & F.mask[0];
  const auto a = ld2(F.re[0] + shot + x0), b = ld2(F.im[0] + shot + x0);
  r0 = double(a.x);
  m0 = double(b.x);
  r1 = double(a.y);
  m1 = double(b.y);
  for (int f = 1; f < F.n; ++f) {
    const size_t xf = (x >> F.shift[f]) & F.mask[f];
    const double cr = double(F.re[f][shot + xf]), ci = double(F.im[f][shot + xf]);
    const double n0r = r0 * cr - m0 * ci, n0i = r0 * ci + m0 * cr;
    const double n1r = r1 * cr - m1 * ci, n1i = r1 * ci + m1 * cr;
    r0 = n0r;
    m0 = n0i;
    r1 = n1r;
    m1 = n1i;
  }
}

// The merged state written out (a chain of merges in one pass).
template <class T>
__global__ __launch_bounds__(kWideThreads) void wide_multimerge_kernel(Factors<T> F, const WideCtl* ctl, T* ore, T* oim,
                                                                       size_t dim, uint32_t out_off, int width) {
  if (ctl[blockIdx.y].dead) return;
  const size_t shot = size_t(blockIdx.y) * dim;
  const size_t total = size_t(1) << width;
  T* OR = ore + shot + out_off;
  T* OI = oim + shot + out_off;
#pragma unroll
  for (int u = 0; u < kWidePer; ++u) {
    const size_t o = size_t(blockIdx.x) * kWideTile + size_t(u) * kWideThreads + threadIdx.x;
    if (o >= total) break;
    T r, m;
    factor_amp<T, T>(F, shot, o, r, m);
    OR[o] = r;
    OI[o] = m;
  }
}

// P(branch = 1) of a measurement of a product state |f0> (x) |f1> (x) ...
// from its factors alone. With the measured body's masks x, z and phase
// i^ph, every measurement kernel's sum over the output half equals
//   P(true) = (N - Re(i^-ph T)) / 2,  N = prod_f N_f,  T = prod_f T_f,
//   N_f = sum_b |f[b]|^2,  T_f = sum_b (-1)^|b & z_f| conj(f[b]) f[b ^ x_f]
// (x_f, z_f: the masks' bits in factor f's index field; a diagonal body has
// x = 0 and i^-ph = (-1)^diag_phase), so the pass reads sum_f 2^k_f
// amplitudes instead of gathering 2^W products. Block (b, shot) covers
// kWideTile entries of one factor and writes (N, Re T, Im T) partials.
template <class T>
__global__ __launch_bounds__(kWideThreads) void wide_expect_factored_kernel(DevProgram P, int instr, Factors<T> F,
                                                                            const WideCtl* ctl, size_t dim,
                                                                            double* part, uint32_t nshots) {
  __shared__ double red[3][kWideThreads / 64];
  const uint32_t j = blockIdx.y;
  if (ctl[j].dead) return;
  const size_t shot = size_t(j) * dim;
  const DevInstr in = P.instrs[instr];
  int f = 0;
  while (f + 1 < F.n && int(blockIdx.x) >= F.tile0[f + 1]) ++f;
  const size_t base = size_t(int(blockIdx.x) - F.tile0[f]) * kWideTile;
  const size_t fmask = F.mask[f];
  const uint64_t xf = (in.xmask >> F.shift[f]) & fmask, zf = (in.zmask >> F.shift[f]) & fmask;
  const T* R = F.re[f] + shot;
  const T* I = F.im[f] + shot;
  double n = 0.0, tr = 0.0, ti = 0.0;
  if (fmask < 64 * kWidePer) {
    // A factor of at most 512 entries is one wave's work: the other waves
    // leave at once (the branch is the same in the whole block) and the sums
    // reduce by shuffles alone, with no barrier.
    if (threadIdx.x >= 64) return;
#pragma unroll
    for (int u = 0; u < kWidePer; ++u) {
      const size_t b = size_t(u) * 64 + threadIdx.x;
      if (b > fmask) break;
      const double r = double(R[b]), m = double(I[b]);
      const double r2 = double(R[b ^ xf]), m2 = double(I[b ^ xf]);
      const double sg = parity64(uint64_t(b) & zf) ? -1.0 : 1.0;
      n += r * r + m * m;
      tr += sg * (r * r2 + m * m2);
      ti += sg * (r * m2 - m * r2);
    }
    for (int off = 32; off > 0; off >>= 1) {
      n += __shfl_xor(n, off, 64);
      tr += __shfl_xor(tr, off, 64);
      ti += __shfl_xor(ti, off, 64);
    }
    if (threadIdx.x == 0) {
      part[(size_t(3) * blockIdx.x) * nshots + j] = n;
      part[(size_t(3) * blockIdx.x + 1) * nshots + j] = tr;
      part[(size_t(3) * blockIdx.x + 2) * nshots + j] = ti;
    }
    return;
  }
  if (xf == 0) {
#pragma unroll
    for (int u = 0; u < kWidePer; ++u) {
      const size_t b = base + size_t(u) * kWideThreads + threadIdx.x;
      if (b > fmask) break;
      const double r = double(R[b]), m = double(I[b]);
      const double sg = parity64(uint64_t(b) & zf) ? -1.0 : 1.0;
      n += r * r + m * m;
      tr += sg * (r * r + m * m);
      ti += sg * (r * m - m * r);
    }
  } else {
    // The entries pair up as (b, b ^ x_f), b with the lowest bit of x_f
    // clear: one thread reads both once and adds both terms, so the pass
    // reads each amplitude once (a thread per entry read each twice). The
    // block's tile of kWideTile entries is kWideTile / 2 pairs.
    const int low = __builtin_ctzll(xf);
    const size_t pbase = base >> 1, npairs = (fmask + 1) >> 1;
#pragma unroll
    for (int u = 0; u < kWidePer / 2; ++u) {
      const size_t q = pbase + size_t(u) * kWideThreads + threadIdx.x;
      if (q >= npairs) break;
      const size_t b0 = insert_zero_bit(q, low), b1 = b0 ^ xf;
      const double r0 = double(R[b0]), m0 = double(I[b0]);
      const double r1 = double(R[b1]), m1 = double(I[b1]);
      const double s0 = parity64(uint64_t(b0) & zf) ? -1.0 : 1.0;
      const double s1 = parity64(uint64_t(b1) & zf) ? -1.0 : 1.0;
      n += r0 * r0 + m0 * m0;
      n += r1 * r1 + m1 * m1;
      tr += s0 * (r0 * r1 + m0 * m1);
      tr += s1 * (r1 * r0 + m1 * m0);
      ti += s0 * (r0 * m1 - m0 * r1);
      ti += s1 * (r1 * m0 - m1 * r0);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    n += __shfl_xor(n, off, 64);
    tr += __shfl_xor(tr, off, 64);
    ti += __shfl_xor(ti, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = n;
    red[1][threadIdx.x >> 6] = tr;
    red[2][threadIdx.x >> 6] = ti;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double t = 0.0;
    for (int w = 0; w < kWideThreads / 64; ++w) t += red[threadIdx.x][w];
    // tile-major, shots fastest (stride nshots, the decision kernel's
    // stride): the per-shot decision reads these coalesced;
    // 3 * tiles * nshots <= max_tiles * batch (checked at launch)
    part[(size_t(3) * blockIdx.x + threadIdx.x) * nshots + j] = t;
  }
}

// Per shot: the factors' partials in fixed order, P(true) as above, then
// the decision.
__global__ __launch_bounds__(kWideThreads) void wide_decide_factored_kernel(DevProgram P, DevLaunch L,
                                                                            WideLayout Y, uint64_t batch_start,
                                                                            uint32_t nshots, int instr,
                                                                            uint64_t* bits, WideCtl* ctl,
                                                                            const double* part, FactorTiles ft,
                                                                            int normalize) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nshots) return;
  const DevInstr in = P.instrs[instr];
  // partial (tile t, component c) of shot j sits at (3t + c) * nshots + j
  const double* pj = part + j;
  const size_t step = nshots;
  double n = 1.0, tr = 1.0, ti = 0.0;
  for (int f = 0; f < ft.n; ++f) {
    double nf = 0.0, rf = 0.0, jf = 0.0;
    for (int t = ft.tile0[f]; t < ft.tile0[f + 1]; ++t) {
      nf += pj[(size_t(3) * t) * step];
      rf += pj[(size_t(3) * t + 1) * step];
      jf += pj[(size_t(3) * t + 2) * step];
    }
    n *= nf;
    const double r2 = tr * rf - ti * jf, i2 = tr * jf + ti * rf;
    tr = r2;
    ti = i2;
  }
  // i^-ph: a diagonal body's phase is +-1; otherwise conj(i^ph) is
  // sqrt(2) * c1, one of +-1, +-i
  double er = 1.0, ei = 0.0;
  if (in.op & kFlagDiagonal) {
    er = (in.op & kFlagDiagPhase) ? -1.0 : 1.0;
  } else {
    er = in.c1_re > 0.35 ? 1.0 : (in.c1_re < -0.35 ? -1.0 : 0.0);
    ei = in.c1_im > 0.35 ? 1.0 : (in.c1_im < -0.35 ? -1.0 : 0.0);
  }
  const double pt = 0.5 * (n - (er * tr - ei * ti));
  // the FP32 path renormalizes by the factors' product norm n
  wide_decide_shot(P, L, Y, j, batch_start + j, in, pt, normalize ? n : 1.0, bits, ctl);
}

// The collapse from the factors into the output buffer (as wide_project_kernel).
template <class T>
__global__ __launch_bounds__(kWideThreads) void wide_project_factored_kernel(DevProgram P, int instr,
                                                                             const WideCtl* ctl, Factors<T> F,
                                                                             T* ore, T* oim, size_t dim, int k) {
  const uint32_t j = blockIdx.y;
  if (ctl[j].dead) return;
  const size_t shot = size_t(j) * dim;
  const DevInstr in = P.instrs[instr];
  const bool branch = ctl[j].branch;
  const double inv = ctl[j].invnorm;
  T* OR = ore + shot + in.comp_off;
  T* OI = oim + shot + in.comp_off;
  const size_t out_dim = (size_t(1) << k) >> 1;
  const bool diag = in.op & kFlagDiagonal;
  if (pairs_ok(in, out_dim, dim) && F.mask[0] >= 1) {
    // outputs (idx, idx + 1) from source pairs (s, s + 1): the first factor
    // read as aligned pairs, the others shared, the outputs stored as pairs
    const bool swap = in.xmask & 1;
#pragma unroll
    for (int u = 0; u < kWidePairs; ++u) {
      const size_t idx = 2 * pair_index(u);
      if (idx >= out_dim) break;
      if (diag) {
        const size_t s0 = diagonal_source(in, idx, branch);  // s0 + 1 feeds idx + 1
        double r0, m0, r1, m1;
        factor_amp2(F, shot, s0, r0, m0, r1, m1);
        st2(OR + idx, T(r0 * inv), T(r1 * inv));
        st2(OI + idx, T(m0 * inv), T(m1 * inv));
      } else {
        const size_t s0 = insert_zero_bit(idx, in.pivot), s1 = (s0 ^ size_t(in.xmask)) & ~size_t(1);
        double ar[2], ai[2], br[2], bi[2];
        factor_amp2(F, shot, s0, ar[0], ai[0], ar[1], ai[1]);
        factor_amp2(F, shot, s1, br[0], bi[0], br[1], bi[1]);
        double orr[2], oii[2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const double r1 = br[e ^ int(swap)], m1 = bi[e ^ int(swap)];
          const bool odd = parity64(uint64_t(s0 + size_t(e)) & in.zmask);
          const double dir = (branch != odd) ? -1.0 : 1.0;
          const double c1r = dir * in.c1_re, c1i = dir * in.c1_im;
          orr[e] = (kInvSqrt2 * ar[e] + c1r * r1 - c1i * m1) * inv;
          oii[e] = (kInvSqrt2 * ai[e] + c1r * m1 + c1i * r1) * inv;
        }
        st2(OR + idx, T(orr[0]), T(orr[1]));
        st2(OI + idx, T(oii[0]), T(oii[1]));
      }
    }
    return;
  }
#pragma unroll
  for (int u = 0; u < kWidePer; ++u) {
    const size_t idx = size_t(blockIdx.x) * kWideTile + size_t(u) * kWideThreads + threadIdx.x;
    if (idx >= out_dim) break;
    double r, m;
    if (diag) {
      factor_amp(F, shot, diagonal_source(in, idx, branch), r, m);
      OR[idx] = T(r * inv);
      OI[idx] = T(m * inv);
    } else {
      const size_t s0 = insert_zero_bit(idx, in.pivot), s1 = s0 ^ size_t(in.xmask);
      double r1, m1;
      factor_amp(F, shot, s0, r, m);
      factor_amp(F, shot, s1, r1, m1);
      const bool odd = parity64(uint64_t(s0) & in.zmask);
      const double dir = (branch != odd) ? -1.0 : 1.0;
      const double c1r = dir * in.c1_re, c1i = dir * in.c1_im;
      const double ar = kInvSqrt2 * double(r) + c1r * double(r1) - c1i * double(m1);
      const double ai = kInvSqrt2 * double(m) + c1r * double(m1) + c1i * double(r1);
      OR[idx] = T(ar * inv);
      OI[idx] = T(ai * inv);
    }
  }
}

// The collapse from the factors (wide_project_factored_kernel's values) fused
// with the probability of the next measurement of the same component, as
// wide_project_measure_kernel fuses them for a stored state: each thread
// writes a pair (a, b) of the next measurement's sources and sums that
// measurement's terms from the stored values.
template <class T>
__global__ __launch_bounds__(kWideThreads) void wide_project_factored_measure_kernel(
    DevProgram P, int instr, int next, const WideCtl* ctl, Factors<T> F, T* ore, T* oim, size_t dim, int k,
    double* part, double* npart, uint32_t max_tiles) {
  __shared__ double red[2][kWideThreads / 64];
  const uint32_t j = blockIdx.y;
  if (ctl[j].dead) return;
  const size_t shot = size_t(j) * dim;
  const DevInstr in = P.instrs[instr];
  const DevInstr nx = P.instrs[next];
  const bool branch = ctl[j].branch;
  const double inv = ctl[j].invnorm;
  T* OR = ore + shot + in.comp_off;
  T* OI = oim + shot + in.comp_off;
  const bool diag = in.op & kFlagDiagonal, ndiag = nx.op & kFlagDiagonal;
  // collapsed amplitude idx of `in` from the factors, rounded as stored
  auto collapsed = [&](size_t idx, T& r, T& m) {
    double rr, mm;
    if (diag) {
      factor_amp(F, shot, diagonal_source(in, idx, branch), rr, mm);
      r = T(rr * inv);
      m = T(mm * inv);
    } else {
      const size_t s0 = insert_zero_bit(idx, in.pivot), s1 = s0 ^ size_t(in.xmask);
      double r1, m1;
      factor_amp(F, shot, s0, rr, mm);
      factor_amp(F, shot, s1, r1, m1);
      const bool odd = parity64(uint64_t(s0) & in.zmask);
      const double dir = (branch != odd) ? -1.0 : 1.0;
      const double c1r = dir * in.c1_re, c1i = dir * in.c1_im;
      const double ar = kInvSqrt2 * rr + c1r * r1 - c1i * m1;
      const double ai = kInvSqrt2 * mm + c1r * m1 + c1i * r1;
      r = T(ar * inv);
      m = T(ai * inv);
    }
  };
  const size_t pairs = ((size_t(1) << k) >> 1) >> 1;  // next's output half
  double acc = 0.0, nacc = 0.0;
#pragma unroll
  for (int u = 0; u < kWidePairs; ++u) {
    const size_t q = size_t(blockIdx.x) * (kWideTile / 2) + size_t(u) * kWideThreads + threadIdx.x;
    if (q >= pairs) break;
    size_t a, b;  // a: the amplitude next's P(true) reads (nondiagonal: with b)
    if (ndiag) {
      a = diagonal_source(nx, q, true);
      b = diagonal_source(nx, q, false);
    } else {
      a = insert_zero_bit(q, nx.pivot);
      b = a ^ size_t(nx.xmask);
    }
    T ra, ma, rb, mb;
    collapsed(a, ra, ma);
    collapsed(b, rb, mb);
    OR[a] = ra;
    OI[a] = ma;
    OR[b] = rb;
    OI[b] = mb;
    if (npart) nacc += double(ra) * double(ra) + double(ma) * double(ma) + double(rb) * double(rb) +
                       double(mb) * double(mb);
    if (ndiag) {
      acc += double(ra) * double(ra) + double(ma) * double(ma);
    } else {
      const double dir = parity64(uint64_t(a) & nx.zmask) ? 1.0 : -1.0;
      const double c1r = dir * nx.c1_re, c1i = dir * nx.c1_im;
      const double ar = kInvSqrt2 * double(ra) + c1r * double(rb) - c1i * double(mb);
      const double ai = kInvSqrt2 * double(ma) + c1r * double(mb) + c1i * double(rb);
      acc += ar * ar + ai * ai;
    }
  }
  block_sum2(acc, nacc, red);
  if (threadIdx.x == 0) {
    part[size_t(j) * max_tiles + blockIdx.x] = acc;
    if (npart) npart[size_t(j) * max_tiles + blockIdx.x] = nacc;
  }
}

// Observables, counts, error flags and packed rows.
__global__ __launch_bounds__(kWideThreads) void wide_finish_kernel(DevProgram P, DevLaunch L, WideLayout Y,
                                                                   uint64_t batch_start, uint32_t nshots,
                                                                   uint64_t* bits, const WideCtl* ctl) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nshots) return;
  const uint64_t local = batch_start + j;
  WideShot w = wide_shot(bits, Y, j);
  if (ctl[j].err) atomicOr(L.error, ctl[j].err);
  auto obs_bit = [&](uint32_t o) {
    bool b = P.obs_flip[o] != 0;
    for (uint32_t q = P.obs_off[o]; q < P.obs_off[o + 1]; ++q) b ^= wbit(w.r, P.obs_recs[q]);
    return b;
  };
  // postselection (src/sampler/batch/postselect.rs): a fired postselected
  // detector discards the shot, which is then not counted
  bool kept = true;
  for (int32_t d = 0; d < P.npost; ++d) kept = kept && !wbit(w.d, P.post_dets[d]);
  if (L.keep && P.npost > 0) L.keep[local] = uint8_t(kept);
  for (int o = 0; o < P.nobs && kept; ++o)
    if (obs_bit(uint32_t(o))) atomicAdd(&L.obs_ones[o], 1ull);
  if (!L.keep_records) return;
  auto emit = [&](uint8_t* out, uint32_t row_bytes, auto&& bit_at, uint32_t ncols) {
    uint8_t* dst = out + local * row_bytes;
    for (uint32_t byte = 0; byte < row_bytes; ++byte) {
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
  emit(L.meas, L.meas_row, [&](uint32_t c) { return wbit(w.r, c); }, uint32_t(P.nrec));
  emit(L.det, L.det_row, [&](uint32_t c) { return wbit(w.d, c); }, uint32_t(P.ndet));
  emit(L.obs, L.obs_row, obs_bit, uint32_t(P.nobs));
}

// Copies shot 0's state (first 2^k amplitudes) out as FP64, for precision
// checks against the oracle.
template <class T>
__global__ void wide_dump_kernel(const T* re, const T* im, size_t n, double* ore, double* oim) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) {
    ore[i] = double(re[i]);
    oim[i] = double(im[i]);
  }
}

// Device workspace of one program handle, kept across calls and grown only
// when a call needs more: re-allocating gigabytes per call costs far more
// than the sampling (fresh pages are mapped on first touch).
struct WideWorkspace {
  void* re[2] = {nullptr, nullptr};
  void* im[2] = {nullptr, nullptr};
  uint64_t* bits = nullptr;
  WideCtl* ctl = nullptr;
  double* part = nullptr;
  size_t amp_bytes = 0;   // bytes of each of re[0..1], im[0..1]
  size_t bit_bytes = 0, ctl_bytes = 0, part_bytes = 0;
  hipEvent_t done = nullptr;  // end of the last call's work
};

WideWorkspace* wide_workspace_new() { return new WideWorkspace(); }

void wide_workspace_free(WideWorkspace* ws) {
  if (!ws) return;
  if (ws->done) (void)hipEventSynchronize(ws->done);
  for (int b = 0; b < 2; ++b) {
    (void)hipFree(ws->re[b]);
    (void)hipFree(ws->im[b]);
  }
  (void)hipFree(ws->bits);
  (void)hipFree(ws->ctl);
  (void)hipFree(ws->part);
  if (ws->done) (void)hipEventDestroy(ws->done);
  delete ws;
}

static hipError_t grow(void** p, size_t& have, size_t need) {
  if (need <= have) return hipSuccess;
  (void)hipFree(*p);
  *p = nullptr;
  have = 0;
  hipError_t e = hipMalloc(p, need);
  if (e == hipSuccess) have = need;
  return e;
}

static hipError_t wide_workspace_reserve(WideWorkspace& ws, size_t batch, size_t dim, size_t elem,
                                         const WideLayout& Y) {
  if (!ws.done) {
    hipError_t e = hipEventCreateWithFlags(&ws.done, hipEventDisableTiming);
    if (e != hipSuccess) return e;
  }
  const size_t amp = batch * dim * elem;
  if (amp > ws.amp_bytes) {
    // an earlier call may still be reading the buffers
    hipError_t e = hipEventSynchronize(ws.done);
    if (e != hipSuccess) return e;
    for (int b = 0; b < 2; ++b) {
      size_t h = ws.amp_bytes;
      if ((e = grow(&ws.re[b], h, amp)) != hipSuccess) return e;
      h = ws.amp_bytes;
      if ((e = grow(&ws.im[b], h, amp)) != hipSuccess) return e;
    }
    ws.amp_bytes = amp;
  }
  hipError_t e = hipSuccess;
  // probability partials, then (FP32 path) norm partials
  const size_t bb = batch * Y.words * 8 + 8, cb = batch * sizeof(WideCtl), pb = 2 * batch * Y.max_tiles * 8 + 8;
  if (bb > ws.bit_bytes || cb > ws.ctl_bytes || pb > ws.part_bytes) {
    if ((e = hipEventSynchronize(ws.done)) != hipSuccess) return e;
    if ((e = grow(reinterpret_cast<void**>(&ws.bits), ws.bit_bytes, bb)) != hipSuccess) return e;
    if ((e = grow(reinterpret_cast<void**>(&ws.ctl), ws.ctl_bytes, cb)) != hipSuccess) return e;
    if ((e = grow(reinterpret_cast<void**>(&ws.part), ws.part_bytes, pb)) != hipSuccess) return e;
  }
  return hipSuccess;
}

// Probability partials per shot: one per tile of a measurement's output
// half, or three per tile of a factor chain's expectation pass.
uint32_t wide_partials_per_shot(int max_local_k) {
  const size_t full = size_t(1) << max_local_k;
  const size_t plain = (full / 2 + kWideTile - 1) / kWideTile + 1;
  const size_t factored = 3 * ((full + kWideTile - 1) / kWideTile + kMaxFactors);
  return uint32_t(std::max(plain, factored));
}

size_t wide_bytes_per_shot(uint32_t state_dim, int fp32, const WideLayout& y) {
  size_t dim = state_dim;
  size_t t = fp32 ? 4 : 8;
  return 4 * dim * t + size_t(y.words) * 8 + sizeof(WideCtl) + 2 * size_t(y.max_tiles) * 8;
}

// Enqueues the whole run on `stream` without synchronizing; `instrs` is the
// host copy of P.instrs.
template <class T>
static hipError_t run_wide(const DevProgram& P, const DevInstr* instrs, const DevLaunch& L, const WideLayout& Y,
                           uint32_t batch, WideWorkspace& ws0, hipStream_t stream0, double* dump_re, double* dump_im,
                           int* dump_k, bool ordered, const WideSecond* second) {
  const size_t dim = P.state_dim;
  hipError_t e = wide_workspace_reserve(ws0, batch, dim, sizeof(T), Y);
  if (e != hipSuccess) return e;
#define WCK(x)             \
  do {                     \
    e = (x);               \
    if (e != hipSuccess) { \
      return e;            \
    }                      \
  } while (0)
  // Shot batches alternate between two workspaces on two streams when a
  // second pair is given and the launch has more than one batch: one
  // batch's small, latency-bound kernels then overlap the other's streaming
  // ones. Where the second workspace cannot grow, every batch stays on the
  // first.
  bool two = second && second->ws && second->stream && L.shots > batch && !dump_re;
  if (two && wide_workspace_reserve(*second->ws, batch, dim, sizeof(T), Y) != hipSuccess) {
    (void)hipGetLastError();  // the failed allocation's error, not a sticky one
    two = false;
  }
  // calls sharing a workspace may come on different streams: order them
  if (ordered) WCK(hipStreamWaitEvent(stream0, ws0.done, 0));
  if (two) {
    // the second stream starts after the caller's earlier work
    WCK(hipEventRecord(second->fork, stream0));
    WCK(hipStreamWaitEvent(second->stream, second->fork, 0));
    if (ordered) WCK(hipStreamWaitEvent(second->stream, second->ws->done, 0));
  }
  const unsigned per_shot_blocks = 256;
  uint64_t bi = 0;
  for (uint64_t start = 0; start < L.shots; start += batch, ++bi) {
    WideWorkspace& ws = (two && (bi & 1)) ? *second->ws : ws0;
    const hipStream_t stream = (two && (bi & 1)) ? second->stream : stream0;
    // FP32 amplitudes: every measurement also sums the measured component's
    // norm (after the probability partials) and renormalizes by it
    double* npart = std::is_same<T, float>::value ? ws.part + size_t(batch) * Y.max_tiles : nullptr;
    uint32_t n = uint32_t(std::min<uint64_t>(batch, L.shots - start));
    dim3 sg((n + per_shot_blocks - 1) / per_shot_blocks), sb(per_shot_blocks);
    // which of the two buffers holds each component's live amplitudes
    std::vector<int> cur(size_t(std::max(P.ncomp, 1)), 0);
    auto RE = [&](int b) { return static_cast<T*>(ws.re[b]); };
    auto IM = [&](int b) { return static_cast<T*>(ws.im[b]); };
    hipLaunchKernelGGL(wide_init_kernel<T>, sg, sb, 0, stream, P, L, Y, start, n, ws.bits, ws.ctl, RE(0), IM(0),
                       dim);
    WCK(hipGetLastError());
    int k = P.initial_k;  // dense width, for the final-state dump
    unsigned fused_tiles = 0;  // > 0: the next measurement's partials are in ws.part
    // The first instruction at or after e that is not a record, detector or
    // coin: bit steps leave the amplitudes and the partials alone, so a
    // collapse can sum the probability of a measurement behind them.
    auto next_amplitude_step = [&](int e) -> const DevInstr* {
      while (e < P.ninstr) {
        const uint32_t o = instrs[e].op & 0xff;
        if (o != kOpRecord && o != kOpDetector && o != kOpCoin) return &instrs[e];
        ++e;
      }
      return nullptr;
    };
    int i = 0;
    while (i < P.ninstr) {
      const DevInstr& in = instrs[i];
      const uint32_t op = in.op & 0xff;
      if (op == kOpRecord || op == kOpDetector || op == kOpCoin) {
        int e2 = i;
        while (e2 < P.ninstr) {
          uint32_t o2 = instrs[e2].op & 0xff;
          if (o2 != kOpRecord && o2 != kOpDetector && o2 != kOpCoin) break;
          ++e2;
        }
        hipLaunchKernelGGL(wide_bits_kernel, sg, sb, 0, stream, P, L, Y, start, n, ws.bits, ws.ctl, i, e2);
        WCK(hipGetLastError());
        i = e2;
        continue;
      }
      const int cb = cur[size_t(in.comp)];
      T* R = RE(cb);
      T* I = IM(cb);
      const int w = in.nq;  // width of the component acted on
      if (op == kOpRotate) {
        size_t work = in.op & kFlagDiagonal ? (size_t(1) << w) : (size_t(1) << w) >> 1;
        dim3 g(unsigned((work + kWideTile - 1) / kWideTile), n);
        hipLaunchKernelGGL(wide_rotate_kernel<T>, g, dim3(kWideThreads), 0, stream, P, Y, i, ws.bits, ws.ctl, R, I,
                           dim, w);
        WCK(hipGetLastError());
      } else if (op == kOpPromote) {
        // consecutive promotions of this component at widths w, w + 1, ...
        int run = 1;
        while (!(in.op & kFlagFresh) && i + run < P.ninstr && run < kWideMaxPromoteRun) {
          const DevInstr& nx = instrs[i + run];
          if ((nx.op & 0xff) != kOpPromote || (nx.op & kFlagFresh) || nx.comp != in.comp ||
              nx.comp_off != in.comp_off || nx.nq != w + run)
            break;
          ++run;
        }
        if (in.op & kFlagFresh) {
          hipLaunchKernelGGL(wide_fresh_kernel<T>, sg, sb, 0, stream, P, Y, i, n, ws.bits, ws.ctl, R, I, dim);
        } else if (run > 1) {
          const size_t total = size_t(1) << (w + run);
          dim3 g(unsigned((total + kWideTile - 1) / kWideTile), n);
          hipLaunchKernelGGL(wide_promote_run_kernel<T>, g, dim3(kWideThreads), 0, stream, P, Y, i, run, ws.bits,
                             ws.ctl, static_cast<const T*>(R), static_cast<const T*>(I), RE(cb ^ 1), IM(cb ^ 1),
                             dim, w);
          WCK(hipGetLastError());
          cur[size_t(in.comp)] ^= 1;
          k += run;
          i += run;
          continue;
        } else {
          dim3 g(unsigned(((size_t(1) << w) + kWideTile - 1) / kWideTile), n);
          hipLaunchKernelGGL(wide_promote_kernel<T>, g, dim3(kWideThreads), 0, stream, P, Y, i, ws.bits, ws.ctl, R,
                             I, dim, w);
          ++k;
        }
        WCK(hipGetLastError());
      } else if (op == kOpMerge) {
        // the chain of merges into this component, as factors
        Factors<T> F{};
        F.n = 1;
        F.re[0] = R + in.comp_off;
        F.im[0] = I + in.comp_off;
        F.shift[0] = 0;
        F.mask[0] = (uint32_t(1) << w) - 1;
        int W = w, e2 = i;
        while (e2 < P.ninstr && F.n < kMaxFactors) {
          const DevInstr& m = instrs[e2];
          if ((m.op & 0xff) != kOpMerge || m.comp != in.comp || m.nq != W) break;
          const int sb2 = cur[size_t(m.src_comp)];
          F.re[F.n] = RE(sb2) + m.src_off;
          F.im[F.n] = IM(sb2) + m.src_off;
          F.shift[F.n] = W;
          F.mask[F.n] = (uint32_t(1) << m.src_k) - 1;
          W += m.src_k;
          ++F.n;
          ++e2;
        }
        const DevInstr* next = e2 < P.ninstr ? &instrs[e2] : nullptr;
        if (next && (next->op & 0xff) == kOpMeasure && next->comp == in.comp && next->nq == W) {
          // merge-and-measure: never materialize the merged state; the
          // probability comes from the factors' expectations
          FactorTiles ft{};
          ft.n = F.n;
          for (int f = 0; f < F.n; ++f)
            ft.tile0[f + 1] = ft.tile0[f] + int((size_t(F.mask[f]) + kWideTile) / kWideTile);
          for (int f = 0; f <= F.n; ++f) F.tile0[f] = ft.tile0[f];
          if (uint32_t(3 * ft.tile0[F.n]) > Y.max_tiles) return hipErrorInvalidValue;
          hipLaunchKernelGGL(wide_expect_factored_kernel<T>, dim3(unsigned(ft.tile0[F.n]), n), dim3(kWideThreads), 0,
                             stream, P, e2, F, ws.ctl, dim, ws.part, n);
          WCK(hipGetLastError());
          hipLaunchKernelGGL(wide_decide_factored_kernel, sg, sb, 0, stream, P, L, Y, start, n, e2, ws.bits, ws.ctl,
                             ws.part, ft, int(npart != nullptr));
          WCK(hipGetLastError());
          size_t out_dim = (size_t(1) << W) >> 1;
          unsigned tiles = unsigned((out_dim + kWideTile - 1) / kWideTile);
          // the next measurement of the merged component, if it follows at
          // once, has its probability summed by the collapse
          const DevInstr* after = next_amplitude_step(e2 + 1);
          const size_t fpairs = out_dim >> 1;
          const unsigned ftiles = unsigned((fpairs + kWideTile / 2 - 1) / (kWideTile / 2));
          if (after && (after->op & 0xff) == kOpMeasure && !(after->op & kFlagExpVal) && after->comp == in.comp &&
              after->nq == W - 1 && fpairs > 0 && ftiles <= Y.max_tiles) {
            hipLaunchKernelGGL(wide_project_factored_measure_kernel<T>, dim3(ftiles, n), dim3(kWideThreads), 0,
                               stream, P, e2, int(after - instrs), ws.ctl, F, RE(cb ^ 1), IM(cb ^ 1), dim, W, ws.part,
                               npart, Y.max_tiles);
            fused_tiles = ftiles;
          } else {
            dim3 g(tiles, n);
            hipLaunchKernelGGL(wide_project_factored_kernel<T>, g, dim3(kWideThreads), 0, stream, P, e2, ws.ctl, F,
                               RE(cb ^ 1), IM(cb ^ 1), dim, W);
          }
          WCK(hipGetLastError());
          cur[size_t(in.comp)] ^= 1;
          i = e2 + 1;
          continue;
        }
        if (F.n == 2) {
          const int sbuf = cur[size_t(in.src_comp)];
          size_t total = size_t(1) << W;
          dim3 g(unsigned((total + kWideTile - 1) / kWideTile), n);
          hipLaunchKernelGGL(wide_merge_kernel<T>, g, dim3(kWideThreads), 0, stream, P, i, ws.ctl, R, I, RE(sbuf),
                             IM(sbuf), RE(cb ^ 1), IM(cb ^ 1), dim);
        } else {
          size_t total = size_t(1) << W;
          dim3 g(unsigned((total + kWideTile - 1) / kWideTile), n);
          hipLaunchKernelGGL(wide_multimerge_kernel<T>, g, dim3(kWideThreads), 0, stream, F, ws.ctl, RE(cb ^ 1),
                             IM(cb ^ 1), dim, in.comp_off, W);
        }
        WCK(hipGetLastError());
        cur[size_t(in.comp)] ^= 1;
        i = e2;
        continue;
      } else {  // Measure
        size_t out_dim = (size_t(1) << w) >> 1;
        unsigned tiles = unsigned((out_dim + kWideTile - 1) / kWideTile);
        dim3 g(tiles, n);
        if (fused_tiles) {  // the previous collapse already summed this probability
          tiles = fused_tiles;
          fused_tiles = 0;
        } else {
          hipLaunchKernelGGL(wide_prob_kernel<T>, g, dim3(kWideThreads), 0, stream, P, i, ws.ctl, R, I, dim, w,
                             ws.part, npart, Y.max_tiles);
          WCK(hipGetLastError());
        }
        hipLaunchKernelGGL(wide_decide_kernel, sg, sb, 0, stream, P, L, Y, start, n, i, ws.bits, ws.ctl, ws.part,
                           npart, tiles, Y.max_tiles);
        WCK(hipGetLastError());
        if (in.op & kFlagExpVal) {  // a probe reads the state and leaves it
          ++i;
          continue;
        }
        const DevInstr* nx = next_amplitude_step(i + 1);
        const size_t pairs = out_dim >> 1;
        const unsigned ftiles = unsigned((pairs + kWideTile / 2 - 1) / (kWideTile / 2));
        if (nx && (nx->op & 0xff) == kOpMeasure && !(nx->op & kFlagExpVal) && nx->comp == in.comp &&
            nx->nq == w - 1 && pairs > 0 && ftiles <= Y.max_tiles) {
          hipLaunchKernelGGL(wide_project_measure_kernel<T>, dim3(ftiles, n), dim3(kWideThreads), 0, stream, P, i,
                             int(nx - instrs), ws.ctl, R, I, RE(cb ^ 1), IM(cb ^ 1), dim, w, ws.part, npart,
                             Y.max_tiles);
          fused_tiles = ftiles;
        } else {
          hipLaunchKernelGGL(wide_project_kernel<T>, g, dim3(kWideThreads), 0, stream, P, i, ws.ctl, R, I,
                             RE(cb ^ 1), IM(cb ^ 1), dim, w);
        }
        WCK(hipGetLastError());
        cur[size_t(in.comp)] ^= 1;
        --k;
      }
      ++i;
    }
    hipLaunchKernelGGL(wide_finish_kernel, sg, sb, 0, stream, P, L, Y, start, n, ws.bits, ws.ctl);
    WCK(hipGetLastError());
    if (dump_re && start == 0) {
      if (P.ncomp != 1) return hipErrorInvalidValue;  // component states are not assembled
      size_t nd = size_t(1) << k;
      hipLaunchKernelGGL(wide_dump_kernel<T>, dim3(unsigned((nd + 255) / 256)), dim3(256), 0, stream,
                         static_cast<const T*>(RE(cur[0])), static_cast<const T*>(IM(cur[0])), nd, dump_re,
                         dump_im);
      WCK(hipGetLastError());
      *dump_k = k;
    }
  }
  if (two) {
    // join: the caller's stream waits for the second stream's batches
    WCK(hipEventRecord(second->ws->done, second->stream));
    WCK(hipStreamWaitEvent(stream0, second->ws->done, 0));
  }
  if (ordered) WCK(hipEventRecord(ws0.done, stream0));
#undef WCK
  return hipSuccess;
}

hipError_t launch_wide_tier(const DevProgram& P, const DevInstr* host_instrs, const DevLaunch& L,
                            const WideLayout& Y, uint32_t batch, int fp32, WideWorkspace* ws, hipStream_t stream,
                            double* dump_re, double* dump_im, int* dump_k, bool ordered, const WideSecond* second) {
  if (!ws) return hipErrorInvalidValue;
  if (fp32)
    return run_wide<float>(P, host_instrs, L, Y, batch, *ws, stream, dump_re, dump_im, dump_k, ordered, second);
  return run_wide<double>(P, host_instrs, L, Y, batch, *ws, stream, dump_re, dump_im, dump_k, ordered, second);
}

// Grows the workspace for `batch` shots of a program ahead of a graph
// capture (no allocation may happen while capturing).
hipError_t wide_workspace_prepare(WideWorkspace* ws, uint32_t batch, uint32_t state_dim, int fp32,
                                  const WideLayout& Y) {
  if (!ws) return hipErrorInvalidValue;
  return wide_workspace_reserve(*ws, batch, state_dim, fp32 ? sizeof(float) : sizeof(double), Y);
}

}  // namespace ticit
