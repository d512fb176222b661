// zigsm.hpp -- the ziggurat as a per-lane state machine fed one draw at a
// time.
//
// A wavefront runs 64 paths in lockstep.  With the plain per-normal loop
// (zig_normal64/32), a lane that leaves the fast path (wedge: 1.4% of
// attempts, layer 255: 0.4%) makes the whole wave execute the extra wedge draw
// and the retry; with 64 lanes that happens on ~69% of normals, so the wave
// pays ~2 PCG draws per normal.  Here every loop iteration draws exactly one
// u64 on every active lane, and each lane decides what that draw is from its
// own state:
//   attempt  -> fast accept (normal done), tail (rare, resolved in place), or
//               wedge pending (x kept)
//   wedge    -> the draw is the wedge uniform: accept (normal done) or back to
//               attempt
// A lane's draws are consumed in exactly the order of normal.rs:18-67, so
// values are unchanged; only the wave's schedule differs.  The steppers layer
// their own per-step grammar on top (Merton's jump uniform is a draw too).
#pragma once
#include "rng.hpp"

namespace stocha {

template <bool kFp32>
struct ZigFeed {
  using Real = typename std::conditional<kFp32, float, double>::type;
  bool wedge = false;  // next draw is the wedge uniform of a pending candidate
  bool negw = false;
  uint32_t layer = 0;
  double xw = 0.0;  // FP64 candidate (the wedge decision needs it exactly)

  // Feed draw u (already taken from g).  Returns true when a normal is
  // complete and stores it in out.  g is only used by the tail sampler,
  // which draws its own pairs.
  __device__ __forceinline__ bool feed(uint64_t u, Pcg& g, const U128 inc,
                                       const ZigLds* __restrict__ z, Real& out) {
    if (!wedge) {
      const uint32_t l = (uint32_t)u & 0xffu;
      const uint64_t m = u >> 11;
      const ZigFast f = z->fast[l];
      const bool neg = (u >> 8) & 1ull;
      if (m < f.t) {
        Real x;
        if constexpr (kFp32)
          x = (float)(uint32_t)(u >> 32) * ((float)f.x * 0x1p-32f);
        else
          x = ((double)m * 0x1p-53) * f.x;
        out = neg ? -x : x;
        return true;
      }
      if (l == 0) {
        const double tl = zig_tail(g, inc);
        out = neg ? -(Real)tl : (Real)tl;
        return true;
      }
      wedge = true;
      negw = neg;
      layer = l;
      xw = ((double)m * 0x1p-53) * f.x;
      return false;
    }
    wedge = false;
    const double y0 = z->y[layer], y1 = z->y[layer + 1];
    const double y = y0 + ((double)(u >> 11) * 0x1p-53) * (y1 - y0);
    bool accept;
    if constexpr (kFp32) {
      // hardware exp2 in FP32; exact FP64 exp only near the boundary
      const float e = __builtin_amdgcn_exp2f((float)(-0.5 * xw * xw) * 1.44269504088896341f);
      const double diff = y - (double)e;
      if (fabs(diff) > 1e-5 * (double)e)
        accept = diff < 0.0;
      else
        accept = y < exp(-0.5 * xw * xw);
    } else {
      accept = y < exp(-0.5 * xw * xw);
    }
    if (accept) {
      out = negw ? -(Real)xw : (Real)xw;
      return true;
    }
    return false;
  }
};

}  // namespace stocha
