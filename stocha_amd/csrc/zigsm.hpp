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
  using Tables = typename std::conditional<kFp32, ZigLds32, ZigLds>::type;
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
      if ((u >> 11) < z->t[idx & 0xffu]) {
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
        const double x64 = ((double)(ua >> 11) * 0x1p-53) * z->x[l];
        const double y64 = z->y[l] + ((double)(u >> 11) * 0x1p-53) * (z->y[l + 1] - z->y[l]);
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

}  // namespace stocha
