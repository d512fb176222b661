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

  // DXSM output of the current (pre-step) state (rand_pcg output_dxsm)
  __device__ __forceinline__ uint64_t output() const {
    uint64_t hi = s.hi;
    const uint64_t lo = s.lo | 1ull;
    hi ^= hi >> 32;
    hi *= kPcgMul;
    hi ^= hi >> 48;
    hi *= lo;
    return hi;
  }
  // LCG step: s*M + inc (mod 2^128), M < 2^64, on 32-bit limbs so that the
  // 32x32+64 multiply-adds (v_mad_u64_u32) and the carry chain map directly:
  // lo*M = t + (u << 32) + ... needs 4 multiply-adds, hi*M mod 2^64 needs
  // one more and two 32-bit products.
  __device__ __forceinline__ U128 stepped(const U128 inc) const {
    constexpr uint32_t m0 = (uint32_t)kPcgMul, m1 = (uint32_t)(kPcgMul >> 32);
    const uint32_t l0 = (uint32_t)s.lo, l1 = (uint32_t)(s.lo >> 32);
    const uint32_t h0 = (uint32_t)s.hi, h1 = (uint32_t)(s.hi >> 32);
    const uint64_t t = (uint64_t)l0 * m0;
    const uint64_t u = (uint64_t)l0 * m1 + (t >> 32);
    const uint64_t w = (uint64_t)l1 * m0 + (uint32_t)u;
    // bits 64..127 of lo*M (no overflow: (2^32-1)^2 + 2 (2^32-1) < 2^64)
    const uint64_t x = (uint64_t)l1 * m1 + ((u >> 32) + (w >> 32));
    const uint64_t a = (uint64_t)h0 * m0 + x;  // wraps mod 2^64, as wanted
    const uint32_t cross = h0 * m1 + h1 * m0;
    uint32_t c0, c1, c2;
    const uint32_t n0 = __builtin_addc((uint32_t)inc.lo, (uint32_t)t, 0u, &c0);
    const uint32_t n1 = __builtin_addc((uint32_t)(inc.lo >> 32), (uint32_t)w, c0, &c1);
    const uint32_t n2 = __builtin_addc((uint32_t)inc.hi, (uint32_t)a, c1, &c2);
    const uint32_t n3 = (uint32_t)(a >> 32) + cross + (uint32_t)(inc.hi >> 32) + c2;
    U128 n;
    n.lo = ((uint64_t)n1 << 32) | n0;
    n.hi = ((uint64_t)n3 << 32) | n2;
    return n;
  }
  __device__ __forceinline__ uint64_t next(const U128 inc) {
    const uint64_t o = output();
    s = stepped(inc);
    return o;
  }

  // The DXSM output of the current state, and the LCG step applied in place
  // on the lanes of `act` only (the others keep their state), for the queue
  // schedule's fill loop.  The output is written on 32-bit limbs so that
  // h1 * m0, a term of both the output's first multiply and the step's, is
  // formed once.  The step s * M + inc (mod 2^128) is
  //   A = l0 m0 + inc_lo  (carry cA at 2^64)      D = l0 m1 + l1 m0  (carry cD at 2^96)
  //   Q = l1 m1 + h0 m0 + inc_hi (mod 2^64)       x = h0 m1 + h1 m0  (mod 2^32)
  //   word 0 = A.lo, word 1 = A.hi + D.lo (carry e1), word 2 = Q.lo + D.hi + cA + e1,
  //   word 3 = Q.hi + x + cD + the carries of word 2,
  // in gfx950 code: the multiply-adds take the increment's 64-bit halves and
  // each other's whole results as addends and hand their carries out in
  // scalar registers (no partial product is moved into a register pair), and
  // the four result words are written into the state under an execution mask
  // of the `act` lanes (so no selects commit the new state).  A VALU write of
  // a scalar register needs two wait states before a VALU reads it as a
  // carry; the instruction order provides them.  inc_lo / inc_hi: the
  // increment's halves in VGPR pairs (the multiply-adds already read the
  // multiplier from a scalar register).  The execution mask is restored
  // before the block ends; the block sets SCC (s_and_saveexec), and says so.
  __device__ __forceinline__ uint64_t draw_step(const uint64_t inc_lo, const uint64_t inc_hi,
                                                const uint64_t act) {
    const uint32_t h1m0 = (uint32_t)(s.hi >> 32) * (uint32_t)kPcgMul;
    const uint64_t out = dxsm(h1m0);
    step_masked(inc_lo, inc_hi, act, h1m0);
    return out;
  }

  // DXSM (rand_pcg output_dxsm): hi ^= hi >> 32; hi *= M; hi ^= hi >> 48;
  // hi *= lo | 1, on 32-bit limbs; h1m0 = (s.hi >> 32) * M.lo (mod 2^32).
  __device__ __forceinline__ uint64_t dxsm(const uint32_t h1m0) const {
    constexpr uint32_t m0 = (uint32_t)kPcgMul, m1 = (uint32_t)(kPcgMul >> 32);
    const uint32_t l0 = (uint32_t)s.lo, l1 = (uint32_t)(s.lo >> 32);
    const uint32_t h0 = (uint32_t)s.hi, h1 = (uint32_t)(s.hi >> 32);
    const uint32_t x0 = h0 ^ h1;
    const uint64_t p = (uint64_t)x0 * m0;
    const uint32_t y1 = (uint32_t)(p >> 32) + x0 * m1 + h1m0;
    const uint32_t y0 = (uint32_t)p ^ (y1 >> 16);
    const uint32_t o0 = l0 | 1u;
    const uint64_t q = (uint64_t)y0 * o0;
    return ((uint64_t)((uint32_t)(q >> 32) + y0 * l1 + y1 * o0) << 32) | (uint32_t)q;
  }

  // The LCG step s * M + inc on the lanes of `act` (see draw_step);
  // h1m0 = (s.hi >> 32) * M.lo of the current state.
  __device__ __forceinline__ void step_masked(const uint64_t inc_lo, const uint64_t inc_hi,
                                              const uint64_t act, const uint32_t h1m0) {
    constexpr uint32_t m0 = (uint32_t)kPcgMul, m1 = (uint32_t)(kPcgMul >> 32);
    uint32_t l0 = (uint32_t)s.lo, l1 = (uint32_t)(s.lo >> 32);
    uint32_t h0 = (uint32_t)s.hi, h1 = (uint32_t)(s.hi >> 32);
    uint64_t A, D, Q, cA, cD, dm;
    uint32_t cr;
    asm volatile(
        "v_mad_u64_u32 %[A], %[cA], %[l0], %[m0], %[il]\n\t"
        "v_mad_u64_u32 %[D], %[dm], %[l0], %[m1], 0\n\t"
        "v_mad_u64_u32 %[Q], %[dm], %[h0], %[m0], %[ih]\n\t"
        "v_mul_lo_u32 %[cr], %[h0], %[m1]\n\t"
        "v_mad_u64_u32 %[D], %[cD], %[l1], %[m0], %[D]\n\t"
        "v_mad_u64_u32 %[Q], %[dm], %[l1], %[m1], %[Q]"
        : [A] "=&v"(A), [D] "=&v"(D), [Q] "=&v"(Q), [cr] "=&v"(cr), [cA] "=&s"(cA),
          [cD] "=&s"(cD), [dm] "=&s"(dm)
        : [l0] "v"(l0), [l1] "v"(l1), [h0] "v"(h0), [m0] "s"(m0), [m1] "s"(m1),
          [il] "v"(inc_lo), [ih] "v"(inc_hi));
    // the four words into the state on the act lanes (cA and cD were written
    // at least four instructions before their first read)
    uint64_t e1, f1, f2, dm2, save;
    uint32_t t;
    asm volatile(
        "s_and_saveexec_b64 %[save], %[act]\n\t"
        "v_mov_b32 %[l0], %[alo]\n\t"
        "v_add_co_u32_e64 %[l1], %[e1], %[ahi], %[dlo]\n\t"
        "v_addc_co_u32_e64 %[h0], %[f1], %[qlo], %[dhi], %[cA]\n\t"
        "v_add3_u32 %[t], %[qhi], %[cr], %[h1m0]\n\t"
        "v_addc_co_u32_e64 %[h0], %[f2], %[h0], 0, %[e1]\n\t"
        "v_addc_co_u32_e64 %[t], %[dm], %[t], 0, %[cD]\n\t"
        "v_addc_co_u32_e64 %[t], %[dm], %[t], 0, %[f1]\n\t"
        "v_addc_co_u32_e64 %[h1], %[dm], %[t], 0, %[f2]\n\t"
        "s_mov_b64 exec, %[save]"
        : [l0] "+v"(l0), [l1] "+v"(l1), [h0] "+v"(h0), [h1] "+v"(h1), [t] "=&v"(t),
          [e1] "=&s"(e1), [f1] "=&s"(f1), [f2] "=&s"(f2), [dm] "=&s"(dm2), [save] "=&s"(save)
        : [alo] "v"((uint32_t)A), [ahi] "v"((uint32_t)(A >> 32)), [dlo] "v"((uint32_t)D),
          [dhi] "v"((uint32_t)(D >> 32)), [qlo] "v"((uint32_t)Q),
          [qhi] "v"((uint32_t)(Q >> 32)), [cr] "v"(cr), [h1m0] "v"(h1m0), [cA] "s"(cA),
          [cD] "s"(cD), [act] "s"(act)
        : "scc");  // s_and_saveexec sets SCC
    (void)dm;
    (void)dm2;
    s.lo = ((uint64_t)l1 << 32) | l0;
    s.hi = ((uint64_t)h1 << 32) | h0;
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

// The affine map of the state over `paths` consecutive streams (block steps
// each), composed from the jump table: a thread that walks paths i, i + paths,
// ... moves its path's first state with one multiply-add per path instead of
// stream_at's walk over the bits of the path index.  `paths` is uniform.
__device__ __forceinline__ Jump stream_stride(const StreamDesc* __restrict__ d, uint64_t paths) {
  Jump r{U128{1, 0}, U128{0, 0}};
#pragma unroll 1
  for (int k = 0; k < kJumpBits; ++k) {
    if ((paths >> k) == 0) break;
    if ((paths >> k) & 1ull) {
      const Jump j = d->jump[k];
      r.a = mul128(j.a, r.a);
      r.c = add128(mul128(j.a, r.c), j.c);
    }
  }
  return r;
}

// u32 -> f32 of the high word of a draw.  LLVM folds uitofp(trunc(x >> 32))
// back into a 64-bit conversion (six instructions on gfx950); the empty asm
// makes the 32-bit operand opaque so a single v_cvt_f32_u32 is emitted.  (No
// memory access is involved.)
__device__ __forceinline__ float hi32_to_f32(uint64_t u) {
  float f;
  asm("v_cvt_f32_u32 %0, %1" : "=v"(f) : "v"((uint32_t)(u >> 32)));
  return f;
}
__device__ __forceinline__ float top24_to_f32(uint64_t u) {  // (u >> 40) as float, exact
  uint32_t h = (uint32_t)(u >> 40);
  asm("" : "+v"(h));
  return (float)h;
}

// ---------------------------------------------------------------------------
// Ziggurat tables staged in LDS, indexed by the low 9 bits of a draw
// (layer = u & 255, sign = bit 8, normal.rs:20-24): entry i holds layer
// i & 255 with the sign of i folded into X, so the value needs no separate
// negation ((-a) * b == -(a * b) exactly in IEEE arithmetic).
//
// Fast-path test (normal.rs:26-29): (u >> 11) < T[l]  <=>  u < T[l] << 11
// (T < 2^53, no overflow).  With th = T >> 21, the high word of u decides it
// unless it equals th: hi < th accepts, hi > th rejects, and the tie (one
// draw in 2^32) is settled from the full T in the slow block.
// ---------------------------------------------------------------------------
struct ZigFast {
  uint32_t th;  // T[l] >> 21
  uint32_t pad;
  double x;  // +-X[l]
};

// FP32 form: x = hi32(u) * (X[l] * 2^-32) in FP32 (the top 32 of the 53 bits
// of m = u >> 11), one 8-byte entry.
struct ZigFast32 {
  uint32_t th;  // T[l] >> 21
  float xs;     // +-X[l] * 2^-32, rounded to FP32
};

// What the FP32 queue samplers read per draw (6 KB): the fast path and the
// FP32 wedge.  The exact threshold T and the FP64 X and Y, read only on a tied
// fast test or inside the FP32 wedge's guard band (a few draws in a million),
// come from constant memory (kZigT, kZigX, kZigY), which keeps the block's LDS
// small enough for seven resident blocks per CU.
struct ZigLds32 {
  ZigFast32 fast32[512];
  float2 ydy32[256];  // FP32 wedge: (Y[l], Y[l+1] - Y[l])
};
// What the FP64 queue samplers read (10 KB): the FP64 fast path and Y for the
// FP64 wedge.
struct ZigLds64 {
  ZigFast fast[512];
  double y[257];
};
// Everything (18 KB), for kernels that also run the loop-form samplers, whose
// fast test uses T.
struct ZigLds : ZigLds32, ZigLds64 {
  uint64_t t[256];  // T[l], exact fast-path threshold on m = u >> 11
};

__constant__ const double kZigX[257] = {ZIG_X_VALUES};
__constant__ const double kZigY[257] = {ZIG_Y_VALUES};
__constant__ const uint64_t kZigT[256] = {ZIG_T_VALUES};
constexpr double kZigR = ZIG_R_VALUE;

// Cooperative fill; caller must __syncthreads() afterwards.
__device__ __forceinline__ void zig_load(ZigLds32* z) {
  for (int i = threadIdx.x; i < 512; i += blockDim.x) {
    const int l = i & 255;
    const double xv = (i & 256) ? -kZigX[l] : kZigX[l];
    const uint32_t th = (uint32_t)(kZigT[l] >> 21);
    z->fast32[i].th = th;
    z->fast32[i].xs = (float)(xv * 0x1p-32);
    if (i < 256) z->ydy32[i] = make_float2((float)kZigY[i], (float)(kZigY[i + 1] - kZigY[i]));
  }
}
__device__ __forceinline__ void zig_load(ZigLds64* z) {
  for (int i = threadIdx.x; i < 512; i += blockDim.x) {
    const int l = i & 255;
    z->fast[i].th = (uint32_t)(kZigT[l] >> 21);
    z->fast[i].pad = 0u;
    z->fast[i].x = (i & 256) ? -kZigX[l] : kZigX[l];
    if (i < 257) z->y[i] = kZigY[i];
  }
}
__device__ __forceinline__ void zig_load(ZigLds* z) {
  zig_load(static_cast<ZigLds32*>(z));
  zig_load(static_cast<ZigLds64*>(z));
  for (int i = threadIdx.x; i < 256; i += blockDim.x) z->t[i] = kZigT[i];
}

// next_f64 of a draw (pcg64.rs:37-39) rounded to FP32, from its bits: the
// top 32 and the next 21 bits of m = u >> 11, so the result is within
// 2^-24 relative of m * 2^-53 for every m > 0 (also tiny m).
__device__ __forceinline__ float u53_to_f32(uint64_t u) {
  uint32_t lo = ((uint32_t)u) >> 11;
  asm("" : "+v"(lo));
  return fmaf((float)lo, 0x1p-53f, hi32_to_f32(u) * 0x1p-32f);
}

// normal.rs:54-67, one (u1, u2) pair from draws a, b: returns whether the pair
// is accepted and then the tail value x + R (x = -ln(u1) / R) in *t.
// The accept test -2 ln(u2) >= x^2 is evaluated with the hardware log2 in
// FP32 and decided there unless |d| = |-2 ln u2 - x^2| is inside a guard band
// of 1e-5 (1 + |terms|); there the FP64 test is recomputed (the FP32 error is
// below 2e-6 (1 + |terms|): u rounded to FP32, log2 within a few ulp), so the
// decision -- and the draws consumed -- are always the FP64 ones.  A zero
// uniform rejects the pair (normal.rs:59-61).
__device__ __forceinline__ bool zig_tail_test(uint64_t a, uint64_t b, float* l1_out) {
  const float l1 = __builtin_amdgcn_logf(u53_to_f32(a));  // log2; u >= 2^-53 is normal
  const float l2 = __builtin_amdgcn_logf(u53_to_f32(b));
  *l1_out = l1;
  const float x = l1 * (float)(-0.69314718055994531 / kZigR);
  const float aa = l2 * -1.3862943611198906f;
  const float x2 = x * x;
  const float d = aa - x2;
  if (fabsf(d) > 1e-5f * (1.f + aa + x2)) return d >= 0.f;
  const double u1 = (double)(a >> 11) * 0x1p-53, u2 = (double)(b >> 11) * 0x1p-53;
  const double xd = -log(u1) / kZigR;
  return -2.0 * log(u2) >= xd * xd;
}
__device__ __forceinline__ bool zig_tail_pair64(uint64_t a, uint64_t b, double* t) {
  if ((a >> 11) == 0 || (b >> 11) == 0) return false;
  float l1;
  if (!zig_tail_test(a, b, &l1)) return false;
  *t = -log((double)(a >> 11) * 0x1p-53) / kZigR + kZigR;
  return true;
}
// FP32 form: same decision; the value from the FP32 log (within ~2 ulp of the
// FP64 value rounded to FP32).
__device__ __forceinline__ bool zig_tail_pair32(uint64_t a, uint64_t b, float* t) {
  if ((a >> 11) == 0 || (b >> 11) == 0) return false;
  float l1;
  if (!zig_tail_test(a, b, &l1)) return false;
  *t = fmaf(l1, (float)(-0.69314718055994531 / kZigR), (float)kZigR);
  return true;
}

// The tail loop for the loop-form samplers.  Kept out of line (it runs for
// 1/256 of attempts); the generator state goes in and comes back by value so
// the caller's state stays in registers.
template <class Real>
struct TailOut {
  Real x;
  U128 s;
};
template <class Real>
__device__ __noinline__ TailOut<Real> zig_tail_impl(U128 s, const U128 inc) {
  Pcg g{s};
  for (;;) {
    const uint64_t a = g.next(inc);
    const uint64_t b = g.next(inc);
    Real t;
    bool ok;
    if constexpr (sizeof(Real) == 4) {
      ok = zig_tail_pair32(a, b, &t);
    } else {
      ok = zig_tail_pair64(a, b, &t);
    }
    if (ok) return TailOut<Real>{t, g.s};
  }
}
template <class Real>
__device__ __forceinline__ Real zig_tail(Pcg& g, const U128 inc) {
  const TailOut<Real> t = zig_tail_impl<Real>(g.s, inc);
  g.s = t.s;
  return t.x;
}

// normal.rs:18-44, FP64 (bit-exact with the reference up to the libm exp/log
// used in the wedge and tail).
__device__ __forceinline__ double zig_normal64(Pcg& g, const U128 inc, const ZigLds* __restrict__ z) {
  for (;;) {
    const uint64_t u = g.next(inc);
    const uint32_t idx = (uint32_t)u & 0x1ffu, layer = idx & 0xffu;
    const uint64_t m = u >> 11;
    const double x = ((double)m * 0x1p-53) * z->fast[idx].x;  // signed
    if (m < z->t[layer]) return x;
    if (layer == 0) {
      const double tl = zig_tail<double>(g, inc);
      return (idx & 256u) ? -tl : tl;
    }
    const double y = z->y[layer] + g.next_f64(inc) * (z->y[layer + 1] - z->y[layer]);
    if (y < exp(-0.5 * x * x)) return x;
  }
}

// FP32 form.  Same draws as zig_normal64; the fast-path value is the top 32
// bits of m times X[l] in FP32 (absolute error <= 2^-32 * X[l] < 1e-9 from
// the dropped mantissa bits, plus FP32 rounding).
// The wedge decision uses FP64 y and x (cheap: full-rate FP64 mul/add) and a
// hardware exp2 in FP32; when |y - e| is within 1e-5*e of the boundary it
// recomputes exp in FP64, so the accept/reject decision equals the FP64
// form's (the FP32 exp error is below 1e-6 relative for |x| <= R).
__device__ __forceinline__ float zig_normal32(Pcg& g, const U128 inc, const ZigLds32* __restrict__ z) {
  for (;;) {
    const uint64_t u = g.next(inc);
    const uint32_t idx = (uint32_t)u & 0x1ffu, layer = idx & 0xffu;
    const uint64_t m = u >> 11;
    const float xf = hi32_to_f32(u) * z->fast32[idx].xs;  // signed
    if (m < kZigT[layer]) return xf;
    if (layer == 0) {
      const float tl = zig_tail<float>(g, inc);
      return (idx & 256u) ? -tl : tl;
    }
    const double x = ((double)m * 0x1p-53) * kZigX[layer];
    const double y = kZigY[layer] + g.next_f64(inc) * (kZigY[layer + 1] - kZigY[layer]);
    const float arg = (float)(-0.5 * x * x);
    const float e = __builtin_amdgcn_exp2f(arg * 1.44269504088896341f);
    const double diff = y - (double)e;
    bool accept;
    if (fabs(diff) > 1e-5 * (double)e) {
      accept = diff < 0.0;
    } else {
      accept = y < exp(-0.5 * x * x);
    }
    if (accept) return xf;  // same FP32 value as the fast path and the feed schedule
  }
}

}  // namespace stocha
