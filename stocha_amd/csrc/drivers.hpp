// drivers.hpp -- per-model draw-feed drivers: each consumes one u64 per
// call and advances the path through the model's per-step grammar of draws
// (GBM: one normal; Heston: two normals, heston.rs:60-61; Merton: a normal,
// one uniform, and a jump normal when the uniform is below lambda*dt,
// jump_diffusion.rs:75-85).  feed() returns true when the last step is done;
// the model's terminal() then gives S_T.
#pragma once
#include "models.hpp"
#include "zigsm.hpp"

namespace stocha {

template <class M>
struct GbmDrv {
  using Real = typename ZigFeed<M::kFp32>::Real;
  M m;
  ZigFeed<M::kFp32> zf;
  uint32_t done = 0;
  template <class C>
  __device__ __forceinline__ explicit GbmDrv(const C& c) : m(c, false) {}
  __device__ __forceinline__ bool feed(uint64_t u, Pcg& g, const U128 inc, const ZigLds* z,
                                       uint32_t steps) {
    Real x;
    if (zf.feed(u, g, inc, z, x)) {
      m.apply(x);
      return ++done == steps;
    }
    return false;
  }
};

template <class M>
struct HestonDrv {
  using Real = typename ZigFeed<M::kFp32>::Real;
  M m;
  ZigFeed<M::kFp32> zf;
  uint32_t done = 0;
  bool have = false;
  Real dw1 = 0;
  template <class C>
  __device__ __forceinline__ explicit HestonDrv(const C& c) : m(c, false) {}
  __device__ __forceinline__ bool feed(uint64_t u, Pcg& g, const U128 inc, const ZigLds* z,
                                       uint32_t steps) {
    Real x;
    if (zf.feed(u, g, inc, z, x)) {
      if (!have) {
        dw1 = x;
        have = true;
        return false;
      }
      have = false;
      m.apply(dw1, x);
      return ++done == steps;
    }
    return false;
  }
};

// Merton's per-step grammar: normal z, then one uniform (always), then a
// jump normal when the uniform is below lambda*dt (jump_diffusion.rs:75-85).
template <class M>
struct MertonDrv {
  using Real = typename ZigFeed<M::kFp32>::Real;
  M m;
  ZigFeed<M::kFp32> zf;
  uint32_t done = 0;
  int phase = 0;  // 0: drawing z, 1: next draw is the jump uniform, 2: drawing z_j
  Real zz = 0;
  template <class C>
  __device__ __forceinline__ explicit MertonDrv(const C& c) : m(c, false) {}
  __device__ __forceinline__ bool feed(uint64_t u, Pcg& g, const U128 inc, const ZigLds* z,
                                       uint32_t steps) {
    if (phase == 1) {
      const double uf = (double)(u >> 11) * 0x1p-53;
      if (uf < m.c.jump_prob) {
        phase = 2;
        return false;
      }
      phase = 0;
      m.apply(zz, false, (Real)0);
      return ++done == steps;
    }
    Real x;
    if (zf.feed(u, g, inc, z, x)) {
      if (phase == 0) {
        zz = x;
        phase = 1;
        return false;
      }
      phase = 0;
      m.apply(zz, true, x);
      return ++done == steps;
    }
    return false;
  }
};

}  // namespace stocha
