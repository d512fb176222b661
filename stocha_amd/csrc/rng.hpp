// rng.hpp -- PCG64DXSM block-split streams and ziggurat normals for gfx950.
//
// Stream semantics follow src/prng/pcg64.rs:23-43 (rand_pcg 0.10.2
// Lcg128CmDxsm64): a 128-bit LCG  s <- s*M + inc  with the 64-bit "cheap"
// multiplier M, DXSM output taken from the pre-step state.  Path i of a
// launch starts at the seed's state advanced by i*block steps
// (gbm.rs:55-59 and the other steppers).
//
// Device layout: the 128-bit state lives in four VGPRs (lo/hi u64); the
// increment and the jump table are uniform over the launch and are read with
// scalar loads.  The 64x64 products are built by the compiler from
// v_mad_u64_u32 / v_mul_hi_u32 / v_mul_lo_u32.
//
// Ziggurat semantics follow src/dist/normal.rs:18-67.  The fast-path test
// fl(fl(m*2^-53)*X[l]) < X[l+1] is replaced by the equivalent integer test
// m < T[l] (T generated with the tables, see gen_ziggurat_tables.py), so the
// FP32 form consumes the stream draw-for-draw like the FP64 form; its wedge
// test runs in FP32 with a guard band and falls back to the exact FP64 test
// near the boundary, and the tail is always FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zig_tables.h"

namespace stocha {

constexpr uint64_t kPcgMul = 0xda942042e4dd58b5ull;

struct U128 {
  uint64_t lo, hi;
};

__host__ __device__ __forceinline__ U128 add128(U128 a, U128 b) {
  U128 r;
  r.lo = a.lo + b.lo;
  r.hi = a.hi + b.hi + (r.lo < a.lo ? 1ull : 0ull);
  return r;
}

__device__ __forceinline__ U128 mul128(U128 a, U128 b) {
  U128 r;
  r.lo = a.lo * b.lo;
  r.hi = __umul64hi(a.lo, b.lo) + a.lo * b.hi + a.hi * b.lo;
  return r;
}

// One jump entry: the affine map s -> A*s + C for a fixed number of steps.
struct Jump {
  U128 a, c;
};

// Per-launch stream description, filled on the host (stream_setup()).
//   base : state of path 0 (Pcg64Dxsm::new(seed), before any draw)
//   inc  : LCG increment ((stream << 1) | 1)
//   jump : jump[k] advances by block * 2^k steps
constexpr int kJumpBits = 48;
struct StreamDesc {
  U128 base;
  U128 inc;
  Jump jump[kJumpBits];
};

struct Pcg {
  U128 s;

  __device__ __forceinline__ uint64_t next(const U128 inc) {
    // DXSM output of the pre-step state (rand_pcg output_dxsm)
    uint64_t hi = s.hi;
    uint64_t lo = s.lo | 1ull;
    hi ^= hi >> 32;
    hi *= kPcgMul;
    hi ^= hi >> 48;
    hi *= lo;
    // LCG step: s = s*M + inc (mod 2^128), M < 2^64
    U128 n;
    n.lo = s.lo * kPcgMul;
    n.hi = __umul64hi(s.lo, kPcgMul) + s.hi * kPcgMul;
    s = add128(n, inc);
    return hi;
  }

  // pcg64.rs:37-39: (u >> 11) * 2^-53, exact
  __device__ __forceinline__ double next_f64(const U128 inc) {
    return (double)(next(inc) >> 11) * 0x1p-53;
  }
};

// State of path `path`: apply jump[k] for every set bit k of path.
// The maps are powers of one LCG map, so they commute and the order of
// application does not matter.
__device__ __forceinline__ U128 stream_at(const StreamDesc* __restrict__ d, uint64_t path) {
  U128 s = d->base;
#pragma unroll 1
  for (int k = 0; k < kJumpBits; ++k) {
    if ((path >> k) == 0) break;
    if ((path >> k) & 1ull) {
      const Jump j = d->jump[k];
      s = add128(mul128(j.a, s), j.c);
    }
  }
  return s;
}

// ---------------------------------------------------------------------------
// Ziggurat tables staged in LDS.  Fast-path entry packs the integer
// threshold and X[l] so one 16-byte LDS read serves the common case.
// ---------------------------------------------------------------------------
struct ZigFast {
  uint64_t t;  // accept iff (u >> 11) < t
  double x;    // X[l]
};

struct ZigLds {
  ZigFast fast[256];
  double x[257];
  double y[257];
};

__constant__ const double kZigX[257] = {ZIG_X_VALUES};
__constant__ const double kZigY[257] = {ZIG_Y_VALUES};
__constant__ const uint64_t kZigT[256] = {ZIG_T_VALUES};
constexpr double kZigR = ZIG_R_VALUE;

// Cooperative fill; caller must __syncthreads() afterwards.
__device__ __forceinline__ void zig_load(ZigLds* z) {
  for (int i = threadIdx.x; i < 257; i += blockDim.x) {
    if (i < 256) {
      z->fast[i].t = kZigT[i];
      z->fast[i].x = kZigX[i];
    }
    z->x[i] = kZigX[i];
    z->y[i] = kZigY[i];
  }
}

// normal.rs:54-67 (always FP64: both forms consume the same draws).
// Kept out of line (it runs for ~0.02% of attempts); the generator state goes
// in and comes back by value so the caller's state stays in registers.
struct TailOut {
  double x;
  U128 s;
};
__device__ __noinline__ TailOut zig_tail_impl(U128 s, const U128 inc) {
  Pcg g{s};
  for (;;) {
    double u1 = g.next_f64(inc);
    double u2 = g.next_f64(inc);
    if (u1 <= 0.0 || u2 <= 0.0) continue;
    double x = -log(u1) / kZigR;
    if (-2.0 * log(u2) >= x * x) return TailOut{x + kZigR, g.s};
  }
}
__device__ __forceinline__ double zig_tail(Pcg& g, const U128 inc) {
  TailOut t = zig_tail_impl(g.s, inc);
  g.s = t.s;
  return t.x;
}

// normal.rs:18-44, FP64 (bit-exact with the reference up to the libm exp/log
// used in the wedge and tail).
__device__ __forceinline__ double zig_normal64(Pcg& g, const U128 inc, const ZigLds* __restrict__ z) {
  for (;;) {
    const uint64_t u = g.next(inc);
    const uint32_t layer = (uint32_t)u & 0xffu;
    const uint64_t m = u >> 11;
    const ZigFast f = z->fast[layer];
    const double x = ((double)m * 0x1p-53) * f.x;
    const bool neg = (u >> 8) & 1ull;
    if (m < f.t) return neg ? -x : x;
    if (layer == 0) {
      double tl = zig_tail(g, inc);
      return neg ? -tl : tl;
    }
    const double y = z->y[layer] + g.next_f64(inc) * (z->y[layer + 1] - z->y[layer]);
    if (y < exp(-0.5 * x * x)) return neg ? -x : x;
  }
}

// FP32 form.  Same draws as zig_normal64; the returned value is the FP32
// rounding of the top 32 bits of m times X[l] (relative error < 2^-23).
// The wedge decision uses FP64 y and x (cheap: full-rate FP64 mul/add) and a
// hardware exp2 in FP32; when |y - e| is within 1e-5*e of the boundary it
// recomputes exp in FP64, so the accept/reject decision equals the FP64
// form's (the FP32 exp error is below 1e-6 relative for |x| <= R).
__device__ __forceinline__ float zig_normal32(Pcg& g, const U128 inc, const ZigLds* __restrict__ z) {
  for (;;) {
    const uint64_t u = g.next(inc);
    const uint32_t layer = (uint32_t)u & 0xffu;
    const uint64_t m = u >> 11;
    const ZigFast f = z->fast[layer];
    const bool neg = (u >> 8) & 1ull;
    if (m < f.t) {
      // m = u >> 11 has 53 bits; its top 32 are u >> 32.
      float xf = (float)(uint32_t)(u >> 32) * ((float)f.x * 0x1p-32f);
      return neg ? -xf : xf;
    }
    if (layer == 0) {
      double tl = zig_tail(g, inc);
      return neg ? -(float)tl : (float)tl;
    }
    const double x = ((double)m * 0x1p-53) * f.x;
    const double y = z->y[layer] + g.next_f64(inc) * (z->y[layer + 1] - z->y[layer]);
    const float arg = (float)(-0.5 * x * x);
    const float e = __builtin_amdgcn_exp2f(arg * 1.44269504088896341f);
    const double diff = y - (double)e;
    bool accept;
    if (fabs(diff) > 1e-5 * (double)e) {
      accept = diff < 0.0;
    } else {
      accept = y < exp(-0.5 * x * x);
    }
    if (accept) return neg ? -(float)x : (float)x;
  }
}

}  // namespace stocha
