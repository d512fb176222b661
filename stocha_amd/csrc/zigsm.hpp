// zigsm.hpp -- the ziggurat as a per-lane state machine fed one draw at a
// time.
//
// A wavefront runs 64 paths in lockstep.  With the plain per-normal loop
// (zig_normal64/32), a lane that leaves the fast path (wedge: 1.4% of
// attempts, layer 255: 0.4%) makes the whole wave execute the extra wedge draw
// and the retry; with 64 lanes that happens on ~69% of normals, so the wave
// pays ~2 PCG draws per normal.  Here every call consumes exactly the one u64
// it is given, and each lane decides what that draw is from its own state:
//   0 attempt  -> fast accept (normal done), wedge pending, or tail pending
//   2 wedge    -> the draw is the wedge uniform: accept, or back to attempt
//   1, 3 tail  -> the draws are the tail sampler's (u1, u2) pair
// A lane's draws are consumed in exactly the order of normal.rs:18-67, so
// values are unchanged; only the wave's schedule differs.  Since a call never
// draws by itself, the caller's generator step is unconditional and can
// overlap the LDS table reads.
//
// FP32 form: the fast-path test is the integer test (identical to FP64); the
// wedge test is evaluated in FP32 and decided there unless |y - e| is within
// a guard band of 4e-6, where the FP64 test of the reference is recomputed
// from the stored draws (the FP32 evaluation errs by < 2e-6 absolute: y from
// 24 bits of the uniform and FP32 table values, e = exp2(-x^2/2 log2 e) with
// x from the top 32 bits of m), so every accept/reject decision is the FP64
// one and both forms consume the same draws.
#pragma once
#include <type_traits>

#include "rng.hpp"

namespace stocha {

template <bool kFp32>
struct ZigFeed {
  using Real = typename std::conditional<kFp32, float, double>::type;
  using Tables = typename std::conditional<kFp32, ZigLds32, ZigLds64>::type;
  uint32_t state = 0;  // 0 attempt, 2 wedge uniform next, 1 / 3 tail pair
  uint32_t idx = 0;    // table index (layer | sign << 8) of a pending wedge / tail
  uint64_t ua = 0;     // wedge: the attempt draw; tail: the pair's first draw
  Real xw = 0;         // wedge: the attempt's (signed) x

  // The next draw goes to the slow block (a wedge uniform or the second draw
  // of a tail pair): states 2 and 3.
  __device__ __forceinline__ bool pending() const { return state >= 2u; }

  // Consumes draw u when `active` (state unchanged otherwise).  Straight-line
  // on every lane: the table read, the fast test and value, and the cheap
  // transitions as selects (a missed attempt records its draw and x; the
  // first draw of a tail pair is recorded).  Only lanes whose draw must be
  // evaluated -- a wedge uniform or the second draw of a tail pair -- take
  // the divergent block (feed_eval).  `out` is written on every call and
  // holds a normal only when the call returns true.
  __device__ __forceinline__ bool feed(uint64_t u, const Tables* __restrict__ z, Real& out,
                                       bool active = true) {
    const uint32_t i = (uint32_t)u & 0x1ffu;
    const uint32_t hi = (uint32_t)(u >> 32);
    uint32_t th;
    if constexpr (kFp32) {
      const ZigFast32 f = z->fast32[i];
      th = f.th;
      out = hi32_to_f32(u) * f.xs;
    } else {
      const ZigFast f = z->fast[i];
      th = f.th;
      out = ((double)(u >> 11) * 0x1p-53) * f.x;
    }
    // (bitwise on bools: no short-circuit control flow)
    const bool at = active & (state == 0);
    const bool take = at & (hi < th);
    // a tie on the high word (one draw in 2^32) is booked as a miss here and
    // settled with the full threshold in feed_eval
    const bool tie = at & (hi == th);
    const bool slow = (active & pending()) | tie;
    const bool miss = at & !take;
    const bool tail1 = active & (state == 1u);
    ua = (miss | tail1) ? u : ua;
    idx = miss ? i : idx;
    xw = miss ? out : xw;
    state = miss ? ((i & 0xffu) == 0 ? 1u : 2u) : (tail1 ? 3u : state);
    if (slow) return feed_eval(u, z, out, tie);
    return take;
  }

  __device__ __forceinline__ bool feed_eval(uint64_t u, const Tables* __restrict__ z, Real& out,
                                            bool tie) {
    if (tie) {  // exact fast-path test (normal.rs:26-29); out holds x
      if ((u >> 11) < kZigT[idx & 0xffu]) {
        state = 0;
        return true;
      }
      return false;  // a real miss, already booked
    }
    if (state == 2u) {  // wedge (normal.rs:36-42): u is the uniform
      state = 0;
      const uint32_t l = idx & 0xffu;
      out = xw;
      if constexpr (kFp32) {
        const float2 yd = z->ydy32[l];
        const float y = fmaf(top24_to_f32(u) * 0x1p-24f, yd.y, yd.x);
        const float e = __builtin_amdgcn_exp2f(-0.72134752044448170f * xw * xw);
        const float diff = y - e;
        if (fabsf(diff) > 4e-6f) return diff < 0.f;
        const double x64 = ((double)(ua >> 11) * 0x1p-53) * kZigX[l];
        const double y64 = kZigY[l] + ((double)(u >> 11) * 0x1p-53) * (kZigY[l + 1] - kZigY[l]);
        return y64 < exp(-0.5 * x64 * x64);
      } else {
        const double y0 = z->y[l], y1 = z->y[l + 1];
        const double y = y0 + ((double)(u >> 11) * 0x1p-53) * (y1 - y0);
        return y < exp(-0.5 * xw * xw);
      }
    }
    state = 1;  // tail (normal.rs:54-67): u is the pair's second draw; on a
                // rejection the next draw starts a new pair
    Real t;
    bool ok;
    if constexpr (kFp32) {
      ok = zig_tail_pair32(ua, u, &t);
    } else {
      ok = zig_tail_pair64(ua, u, &t);
    }
    if (!ok) return false;
    state = 0;
    out = (idx & 256u) ? -t : t;
    return true;
  }
};

// ---------------------------------------------------------------------------
// The same sampler for a whole wavefront at once (the queue schedule): the
// per-lane state is held as lane masks -- wave-uniform 64-bit words, one bit
// per lane -- so every state transition is scalar-unit work and the VALU
// only computes the draw's values and compares (whose ballots are free: a
// compare's result already is a lane mask).  The rarely needed work (a tied
// fast test, the draw after a missed attempt, a tail pair's draws) runs
// under uniform branches, each only when some lane needs it, with its outcome
// ballotted back into the masks.  Wave-collective: every lane of the
// wavefront calls feed together; `act` names the lanes that consume a draw.
// A lane's draws, values and accept/reject decisions are exactly ZigFeed's.
__device__ __forceinline__ bool lane_in(uint64_t mask) {
  return __builtin_amdgcn_inverse_ballot_w64(mask);
}
__device__ __forceinline__ uint64_t lanes(bool c) { return __builtin_amdgcn_ballot_w64(c); }

template <bool kFp32>
struct ZigWave {
  using Real = typename std::conditional<kFp32, float, double>::type;
  using Tables = typename std::conditional<kFp32, ZigLds32, ZigLds64>::type;
  // The states after a missed attempt are told apart only when the lane is
  // served (in the slow block): the fast part neither tests the layer of a
  // miss nor keeps its sign.
  uint64_t missed = 0;  // lanes whose next draw follows the missed attempt ua:
                        // the wedge uniform (layer > 0) or a tail pair's first
  uint64_t tail1 = 0;   // lanes whose next draw starts a new tail pair
  uint64_t tail2 = 0;   // lanes whose next draw is the second of the pair at ua
  uint64_t neg = 0;     // sign of the layer-0 attempt that led each tail
  uint64_t ua = 0;      // per lane: the missed attempt or the pair's first draw

  __device__ __forceinline__ uint64_t pending_mask() const { return missed | tail1 | tail2; }

  // The fast-path table entry of draw u (layer and sign: bits 0..8).
  using Fast = typename std::conditional<kFp32, ZigFast32, ZigFast>::type;
  __device__ static __forceinline__ Fast fast_entry(const Tables* __restrict__ z, uint64_t u) {
    if constexpr (kFp32) {
      return z->fast32[(uint32_t)u & 0x1ffu];
    } else {
      return z->fast[(uint32_t)u & 0x1ffu];
    }
  }

  // Consumes draw u on the lanes of `act`; returns the lanes whose draw
  // completed a normal (in out).  `out` is written on every lane.
  __device__ __forceinline__ uint64_t feed(uint64_t u, const Tables* __restrict__ z, Real& out,
                                           uint64_t act) {
    return feed_with(u, fast_entry(z, u), z, out, act);
  }
  // The same with u's table entry f already read.
  __device__ __forceinline__ uint64_t feed_with(uint64_t u, const Fast f,
                                                const Tables* __restrict__ z, Real& out,
                                                uint64_t act) {
    const uint32_t hi = (uint32_t)(u >> 32);
    const uint32_t th = f.th;
    if constexpr (kFp32) {
      out = hi32_to_f32(u) * f.xs;  // the table folds the sign bit (bit 8) into X
    } else {
      out = ((double)(u >> 11) * 0x1p-53) * f.x;
    }
    const uint64_t lt = lanes(hi < th), eq = lanes(hi == th);
    const uint64_t held = missed | tail1 | tail2;
    const uint64_t at = act & ~held;
    const uint64_t take = at & lt;
    // a tie on the high word (one draw in 2^32) is booked as a miss and
    // settled with the full threshold in the slow block
    const uint64_t tie = at & eq, miss = at & ~lt;
    const uint64_t served = act & held;  // their states are updated in the slow block
    ua = lane_in(miss) ? u : ua;
    missed |= miss;
    uint64_t got = take;
    if ((tie | served) != 0) got |= slow(u, z, out, tie, served);
    return got;
  }

  __device__ __forceinline__ uint64_t slow(uint64_t u, const Tables* __restrict__ z, Real& out,
                                           uint64_t tie, uint64_t served) {
    uint64_t got = 0;
    // (this draw's misses, ties included, are not among the served lanes)
    const uint64_t ml = served & missed, t1 = served & tail1, t2 = served & tail2;
    missed &= ~ml;
    tail1 &= ~t1;
    tail2 &= ~t2;
    if (tie != 0) {  // exact fast-path test (normal.rs:26-29); out holds x
      const uint64_t ok = tie & lanes((u >> 11) < kZigT[u & 0xffu]);
      missed &= ~ok;
      got |= ok;
    }
    if (t2 != 0) {  // tail (normal.rs:54-67): u is the pair's second draw
      Real t;
      bool ok;
      if constexpr (kFp32) {
        ok = zig_tail_pair32(ua, u, &t);
      } else {
        ok = zig_tail_pair64(ua, u, &t);
      }
      const uint64_t okm = t2 & lanes(ok);
      tail1 |= t2 & ~okm;  // rejected: the next draw starts a new pair
      out = lane_in(okm) ? (lane_in(neg) ? -t : t) : out;
      got |= okm;
    }
    if (ml != 0) {
      const uint32_t ia = (uint32_t)ua & 0x1ffu, l = ia & 0xffu;
      const uint64_t l0 = ml & lanes(l == 0u);
      const uint64_t wl = ml & ~l0;
      if (wl != 0) {  // wedge (normal.rs:36-42): u is the uniform, ua the attempt
        // (every lane evaluates; indices are masked into the tables)
        Real xw;
        uint64_t acc;
        if constexpr (kFp32) {
          xw = hi32_to_f32(ua) * z->fast32[ia].xs;  // the attempt's value
          const float2 yd = z->ydy32[l];
          const float y = fmaf(top24_to_f32(u) * 0x1p-24f, yd.y, yd.x);
          const float e = __builtin_amdgcn_exp2f(-0.72134752044448170f * xw * xw);
          const float diff = y - e;
          acc = lanes(diff < 0.f);
          const uint64_t close = wl & ~lanes(fabsf(diff) > 4e-6f);
          if (close != 0) {  // inside the guard band: the FP64 test decides
            const double x64 = ((double)(ua >> 11) * 0x1p-53) * kZigX[l];
            const double y64 =
                kZigY[l] + ((double)(u >> 11) * 0x1p-53) * (kZigY[l + 1] - kZigY[l]);
            acc = (acc & ~close) | (close & lanes(y64 < exp(-0.5 * x64 * x64)));
          }
        } else {
          xw = ((double)(ua >> 11) * 0x1p-53) * z->fast[ia].x;
          const double y0 = z->y[l], y1 = z->y[l + 1];
          const double y = y0 + ((double)(u >> 11) * 0x1p-53) * (y1 - y0);
          acc = lanes(y < exp(-0.5 * xw * xw));
        }
        out = lane_in(wl) ? xw : out;
        got |= wl & acc;
      }
      if (l0 != 0) {  // layer 0 (normal.rs:54): u is the first of a tail pair;
                      // the attempt's sign (bit 8) signs the tail value
        neg = (neg & ~l0) | (l0 & lanes((ia & 0x100u) != 0u));
        ua = lane_in(l0) ? u : ua;
        tail2 |= l0;
      }
    }
    if (t1 != 0) {  // the first draw of a new pair
      ua = lane_in(t1) ? u : ua;
      tail2 |= t1;
    }
    return got;
  }
};

}  // namespace stocha
