// drivers.hpp -- the queue schedule: per-model step-input sources fed one u64
// per call, following each model's per-step grammar of draws (GBM: one
// normal; Heston: two normals, heston.rs:60-61; Merton: a normal, one
// uniform, and a jump normal when the uniform is below lambda*dt,
// jump_diffusion.rs:75-85), and queue_path, which runs a whole wavefront of
// paths through them.
#pragma once
#include "models.hpp"
#include "zigsm.hpp"

namespace stocha {

// Block barrier ordering LDS only: unlike __syncthreads it does not wait for
// this wave's outstanding global loads and stores, so memory operations
// issued earlier stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Orders this wavefront's LDS accesses across its lanes: the LDS performs a
// wavefront's operations in issue order, so a read issued after another
// lane's write of the same wavefront sees it; only the compiler must keep
// the order (nothing is emitted).  For tiles that one wavefront writes and
// reads back, with no block barrier.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront", "local");
  asm volatile("" ::: "memory");
}

// Slow-block deferral of a source (Src::kDefer): 0 -- none (every waiting
// lane is served at once); 1..64 -- waiting lanes are served once kDefer of
// them wait or one of them still lacks entries; kDeferUntilNeeded -- only once
// one of them still lacks entries.
constexpr uint32_t kDeferUntilNeeded = 65;

// Step-input sources for queue_path.  A source turns the lanes' draws, one
// per lane per call, into queue entries; Src::kPerStep entries make one
// step's input.  Sources are wave-collective: feed(u, z, rec, act) is called
// by every lane of the wavefront together, consumes u on the lanes of the
// lane mask `act` (the others keep their state), and returns the lane mask of
// lanes whose rec holds a new entry.  pending_mask(): the lanes whose next
// draw goes to the ziggurat's slow block.
//
// GBM (gbm.rs:60-66) and Heston Euler (heston.rs:59-71) take only normals, so the
// entries are single normals: 1 per GBM step, 2 per Heston step (dw1 first,
// then the independent dz).
template <bool kFp32, uint32_t kPer>
struct NormalSrc {
  using Real = typename ZigWave<kFp32>::Real;
  using Tables = typename ZigWave<kFp32>::Tables;
  using Rec = Real;
  static constexpr uint32_t kPerStep = kPer;
  // (a count rule -- serve once 8 wait -- took 1.3% fewer draw iterations but
  // 3 more scalar instructions per iteration: 1.5% slower on the headline)
  static constexpr uint32_t kDefer = kDeferUntilNeeded;
  static constexpr int kWaves = kFp32 ? 7 : 5;  // occupancy target of the fused kernel
  ZigWave<kFp32> zf;
  template <class C>
  __device__ __forceinline__ explicit NormalSrc(const C&) {}
  __device__ __forceinline__ uint64_t pending_mask() const { return zf.pending_mask(); }
  __device__ __forceinline__ uint64_t feed(uint64_t u, const Tables* z, Rec& r, uint64_t act) {
    return zf.feed(u, z, r, act);
  }
  // (queue_fill_ahead: u's fast-table entry read by the caller)
  using Wave = ZigWave<kFp32>;
  using Fast = typename Wave::Fast;
  __device__ __forceinline__ uint64_t feed_with(uint64_t u, const Fast f, const Tables* z, Rec& r,
                                                uint64_t act) {
    return zf.feed_with(u, f, z, r, act);
  }
  template <uint32_t kSlots, int kCols, class M>
  __device__ static __forceinline__ void apply(M& m, const Rec (*q)[kCols], uint32_t head) {
    if constexpr (kPer == 1) {
      m.apply(q[head][threadIdx.x]);
    } else {
      // head advances by 2 from 0 in a ring of an even size, so it is even
      // and head + 1 needs no wrap (one address, two reads at fixed offsets)
      static_assert(kSlots % 2 == 0, "ring size");
      m.apply(q[head][threadIdx.x], q[head + 1][threadIdx.x]);
    }
  }
};
template <bool kFp32>
using GbmSrc = NormalSrc<kFp32, 1>;
template <bool kFp32>
using HestonSrc = NormalSrc<kFp32, 2>;

// Merton: normal z, one uniform, a jump normal when the uniform is below
// lambda*dt (jump_diffusion.rs:75-85); one entry per step carrying z and the
// jump return (0 without a jump).
template <bool kFp32, class M>
struct MertonSrc {
  using Real = typename ZigWave<kFp32>::Real;
  using Tables = typename ZigWave<kFp32>::Tables;
  struct Rec {
    Real z, jr;
  };
  static constexpr uint32_t kPerStep = 1;
  // no slow-block deferral: with a step's draws split between the ziggurat
  // and the uniform, waiting lanes cost more draws than the shared pass saves
  static constexpr uint32_t kDefer = 0;
  // FP32: 72 VGPRs, 23,040 B of LDS -- 7 waves per SIMD (1.10 -> 0.96 ms at
  // 1M x 252); FP64: 91 VGPRs, 27,152 B -- 5 waves
  static constexpr int kWaves = kFp32 ? 7 : 5;
  ZigWave<kFp32> zf;
  uint64_t uni_next = 0;   // lanes whose next draw is the uniform
  uint64_t jump_next = 0;  // lanes whose next normal is the jump size
  Real zz = 0;
  uint64_t t53;  // next_f64() < lambda*dt  <=>  (u >> 11) < t53
  M jm;          // for jump_return()
  template <class C>
  __device__ __forceinline__ explicit MertonSrc(const C& c) : t53(c.jump_t53), jm(c, false) {}
  __device__ __forceinline__ uint64_t pending_mask() const { return zf.pending_mask(); }
  // The phase transitions are mask operations; rec is written on every lane.
  __device__ __forceinline__ uint64_t feed(uint64_t u, const Tables* z, Rec& r, uint64_t act) {
    Real x;
    const uint64_t uni = act & uni_next;
    const uint64_t got = zf.feed(u, z, x, act & ~uni);
    const uint64_t jump = lanes((u >> 11) < t53);  // the uniform draw (pcg64.rs:37-39)
    const uint64_t got0 = got & ~jump_next;  // the diffusion normal z
    const uint64_t got2 = got & jump_next;   // the jump normal
    r.z = zz;
    r.jr = lane_in(got2) ? jm.jump_return(x) : (Real)0;
    zz = lane_in(got0) ? x : zz;
    uni_next = (uni_next & ~uni) | got0;
    jump_next = (jump_next & ~got2) | (uni & jump);
    return (uni & ~jump) | got2;
  }
  template <uint32_t kSlots, int kCols, class MM>
  __device__ static __forceinline__ void apply(MM& m, const Rec (*q)[kCols], uint32_t head) {
    const Rec r = q[head][threadIdx.x];
    m.apply_jr(r.z, r.jr);
  }
};

// Heston QE (heston.rs:108-159): a step draws the variance sample -- a normal
// when psi <= 1.5, else one uniform -- and then the asset normal z1.  Which
// kind the first draw is depends on V(t), so the source runs the variance
// recursion itself (it needs only the draws) and queues (V(t+dt), z1) per
// step; the model does the log-price update.  FP64, the reference's
// arithmetic; the per-step setup (m, s2, psi and the branch constants for
// the next variance draw) runs when a variance sample completes.
template <class M>
struct HestonQeSrc {
  using Real = double;
  using Tables = ZigLds64;
  struct Rec {
    double vn, z1;
  };
  static constexpr uint32_t kPerStep = 1;
  static constexpr uint32_t kDefer = 0;
  static constexpr int kWaves = 4;
  ZigFeed<false> zf;  // the per-lane sampler (its slow path is the rarer one here)
  uint32_t phase = 0;  // 0: variance draw next, 1: asset normal next
  bool quad = true;    // the pending variance draw is a normal (quadratic branch)
  double qa = 0, qb = 0;  // quadratic: a, b; exponential: p, beta
  double vn = 0;          // V(t+dt) of the step being assembled
  HestonQeConst c;
  template <class C>
  __device__ __forceinline__ explicit HestonQeSrc(const C& k) : c(k) {
    setup(k.v0);
  }
  // heston.rs:109-130: the branch for the variance draw out of V(t) = v
  __device__ __forceinline__ void setup(double v) {
    const double v_plus = fmax(v, 0.0);
    double m = c.m_const + v_plus * c.e_kdt;
    m = fmax(m, 1e-15);
    const double s2 = v_plus * c.xi * c.xi * c.e_kdt * c.one_m_e / c.kappa + c.s2_const;
    const double psi = s2 / (m * m);
    quad = psi <= 1.5;
    if (quad) {
      const double two_psi = 2.0 / psi;
      double b2 = two_psi - 1.0 + sqrt(two_psi * (two_psi - 1.0));
      b2 = fmax(b2, 0.0);
      qa = m / (1.0 + b2);
      qb = sqrt(b2);
    } else {
      const double p = (psi - 1.0) / (psi + 1.0);
      qa = p;
      qb = (1.0 - p) / m;
    }
  }
  __device__ __forceinline__ uint64_t pending_mask() const { return lanes(zf.pending()); }
  __device__ __forceinline__ uint64_t feed(uint64_t u, const Tables* z, Rec& r, uint64_t act) {
    return lanes(feed_lane(u, z, r, lane_in(act)));
  }
  __device__ __forceinline__ bool feed_lane(uint64_t u, const Tables* z, Rec& r, bool active) {
    double x;
    const bool uni = active & (phase == 0u) & !quad;
    const bool got = zf.feed(u, z, x, active & !uni);
    const bool var_done = uni | (got & (phase == 0u));
    const bool out = got & (phase == 1u);
    r.vn = vn;
    r.z1 = x;
    if (var_done) {
      double v;
      if (uni) {
        const double uf = (double)(u >> 11) * 0x1p-53;  // next_f64 (pcg64.rs:37-39)
        v = uf <= qa ? 0.0 : log((1.0 - qa) / (1.0 - uf)) / qb;
      } else {
        v = qa * (qb + x) * (qb + x);
      }
      vn = fmax(v, 0.0);
      setup(vn);  // the next step's branch
    }
    phase = var_done ? 1u : (out ? 0u : phase);
    return out;
  }
  template <uint32_t kSlots, int kCols, class MM>
  __device__ static __forceinline__ void apply(MM& m, const Rec (*q)[kCols], uint32_t head) {
    const Rec r = q[head][threadIdx.x];
    m.apply_qe(r.vn, r.z1);
  }
};

// One path simulated with the queue schedule.  Each lane feeds its draws into
// src and appends the entries to its column of q, a ring of kSlots (a power of
// two) holding at most kSlots - 1 entries; before each step the wave draws
// until every lane holds a step's input (lanes that have one keep drawing
// ahead while they have room), then applies the step on all lanes together.
// The fill loop is branch-free apart from the ziggurat's slow block: every
// lane computes the output and the next generator state, a lane with a full
// ring keeps its old state, and the entry is stored unconditionally into the
// first free slot (one slot is always free, so no live entry is overwritten).
// hook(step, model) runs after every step (path-storing callers write the
// value there; it may contain block barriers, since every wave of the block
// runs the same steps).  Must be called by whole wavefronts (every lane, same
// steps).  mirror: the model negates its normals (antithetic rows).
struct NoHook {
  template <class M>
  __device__ __forceinline__ void operator()(uint32_t, const M&) const {}
};

// The stream increment as two 64-bit halves in VGPR pairs for the whole
// loop: they are addends of Pcg::draw_step's multiply-adds, whose other
// scalar operand is the multiplier (gfx9 allows one scalar operand per VALU
// instruction).
struct LcgInc {
  uint64_t lo, hi;
};
__device__ __forceinline__ LcgInc vgpr_increment(const U128 inc) {
  LcgInc r{inc.lo, inc.hi};
  asm("" : "+v"(r.lo));
  asm("" : "+v"(r.hi));
  return r;
}

// c + 1 on the lanes of mask m, c elsewhere: one add with the lane mask as
// the carry-in (a select of 0/1 and an add otherwise).
__device__ __forceinline__ uint32_t add_lanes(uint32_t c, uint64_t m) {
  uint32_t r;
  uint64_t carry;
  asm("v_addc_co_u32_e64 %0, %1, 0, %2, %3" : "=v"(r), "=s"(carry) : "v"(c), "s"(m));
  return r;
}

// The lanes that draw in this fill iteration: `room` (lanes with a free ring
// slot) without the waiting lanes, unless one of them still needs entries or
// (kDefer <= 64) kDefer of them wait (go): room & (go ? ~0 : ~waiting).
// Four or seven scalar instructions (the compiler's form of the same
// expression took 14 with the count rule).
template <uint32_t kDefer>
__device__ __forceinline__ uint64_t defer_act(uint64_t room, uint64_t waiting, uint64_t need) {
  static_assert(kDefer >= 1 && kDefer <= kDeferUntilNeeded, "deferral count");
  uint64_t act, g, t;
  if constexpr (kDefer == kDeferUntilNeeded) {
    asm("s_and_b64 %[t], %[need], %[w]\n\t"
        "s_cselect_b64 %[g], -1, 0\n\t"
        "s_orn2_b64 %[g], %[g], %[w]\n\t"
        "s_and_b64 %[act], %[room], %[g]"
        : [act] "=s"(act), [g] "=&s"(g), [t] "=&s"(t)
        : [w] "s"(waiting), [need] "s"(need), [room] "s"(room)
        : "scc");
  } else {
    uint32_t n;
    asm("s_bcnt1_i32_b64 %[n], %[w]\n\t"
        "s_cmp_gt_u32 %[n], %[km1]\n\t"
        "s_cselect_b64 %[g], -1, 0\n\t"
        "s_and_b64 %[t], %[need], %[w]\n\t"
        "s_cselect_b64 %[g], -1, %[g]\n\t"
        "s_orn2_b64 %[g], %[g], %[w]\n\t"
        "s_and_b64 %[act], %[room], %[g]"
        : [act] "=s"(act), [g] "=&s"(g), [t] "=&s"(t), [n] "=&s"(n)
        : [w] "s"(waiting), [need] "s"(need), [room] "s"(room), [km1] "n"(kDefer - 1)
        : "scc");
  }
  return act;
}

// One fill of the per-lane rings q (kSlots entries per lane, the lane's
// `cnt` entries starting at `head`): draws until every lane holds at least
// `need` entries.  Bottom-tested: one draw is always taken (by lanes with
// room), which keeps the loop-carried state in place across the test.
// Waiting lanes (their next draw goes to the slow block) are served by the
// source's deferral rule (kDefer above); the rule always serves them once one
// of them lacks entries, so a fill ends: a lane that lacks entries either
// draws or, waiting, gets the waiting lanes served.
template <class Src, uint32_t kSlots, int kCols>
__device__ __forceinline__ void queue_fill(Src& src, Pcg& g, const LcgInc& incv,
                                           const typename Src::Tables* zig,
                                           typename Src::Rec (*q)[kCols], uint32_t head,
                                           uint32_t& cnt, uint32_t need_n) {
  static_assert((kSlots & (kSlots - 1)) == 0, "ring size");
  constexpr uint32_t kCap = kSlots - 1;
  using Rec = typename Src::Rec;
  // lanes that still lack the entries asked for (a wave-uniform lane mask)
  uint64_t need = lanes(cnt < need_n);
  do {
    // A lane whose next draw goes to the ziggurat's slow block holds it
    // until the rule above serves the waiting lanes, so that one pass
    // through the slow block serves many lanes (without the wait, 69% of a
    // wave's draws would take it).  The lane's own draws are consumed in the
    // same order either way.
    uint64_t act = lanes(cnt < kCap);
    if constexpr (Src::kDefer > 0) act = defer_act<Src::kDefer>(act, src.pending_mask(), need);
    const uint64_t u = g.draw_step(incv.lo, incv.hi, act);  // steps the act lanes
    Rec r;
    const uint64_t got = src.feed(u, zig, r, act);
    q[(head + cnt) & (kSlots - 1)][threadIdx.x] = r;
    cnt = add_lanes(cnt, got);
    need = lanes(cnt < need_n);
  } while (need != 0);
}

// The generator one draw ahead, for kernels with few waves per SIMD, where
// the fast-table read of each draw would stall the fill loop: u is the DXSM
// output of g's state (the lane's next draw), f its fast-table entry (read
// one fill iteration before it is used), h1m0 the product (s.hi >> 32) * M.lo
// that the output and the next step share.  A lane's draws are the same
// values in the same order as with queue_fill.
template <class Src>
struct DrawAhead {
  Pcg g;
  uint64_t u;
  uint32_t h1m0;
  typename Src::Fast f;
  __device__ __forceinline__ DrawAhead(const U128 s, const typename Src::Tables* zig) : g{s} {
    h1m0 = (uint32_t)(g.s.hi >> 32) * (uint32_t)kPcgMul;
    u = g.dxsm(h1m0);
    f = Src::Wave::fast_entry(zig, u);
  }
};

// queue_fill with the generator one draw ahead: each iteration steps the
// active lanes' states, forms the next draw and issues its table read, then
// feeds the current draw with the entry read one iteration earlier.
// Inactive lanes keep their state, so their next draw (and entry) is
// recomputed unchanged.
template <class Src, uint32_t kSlots, int kCols>
__device__ __forceinline__ void queue_fill_ahead(Src& src, DrawAhead<Src>& da,
                                                 const LcgInc& incv,
                                                 const typename Src::Tables* zig,
                                                 typename Src::Rec (*q)[kCols], uint32_t head,
                                                 uint32_t& cnt, uint32_t need_n) {
  static_assert((kSlots & (kSlots - 1)) == 0, "ring size");
  constexpr uint32_t kCap = kSlots - 1;
  using Rec = typename Src::Rec;
  uint64_t need = lanes(cnt < need_n);
  do {
    uint64_t act = lanes(cnt < kCap);
    if constexpr (Src::kDefer > 0) act = defer_act<Src::kDefer>(act, src.pending_mask(), need);
    const uint64_t u = da.u;
    const typename Src::Fast f = da.f;
    da.g.step_masked(incv.lo, incv.hi, act, da.h1m0);
    da.h1m0 = (uint32_t)(da.g.s.hi >> 32) * (uint32_t)kPcgMul;
    da.u = da.g.dxsm(da.h1m0);
    da.f = Src::Wave::fast_entry(zig, da.u);
    Rec r;
    const uint64_t got = src.feed_with(u, f, zig, r, act);
    q[(head + cnt) & (kSlots - 1)][threadIdx.x] = r;
    cnt = add_lanes(cnt, got);
    need = lanes(cnt < need_n);
  } while (need != 0);
}

// The loop itself, on a caller-owned model (for models whose state the
// caller reads afterwards).
template <class Src, uint32_t kSlots, int kCols, class Model, class Hook = NoHook>
__device__ __forceinline__ void queue_run(Src& src, Model& m, Pcg& g, const U128 inc,
                                          uint32_t steps, const typename Src::Tables* zig,
                                          typename Src::Rec (*q)[kCols], Hook hook = Hook()) {
  static_assert((kSlots & (kSlots - 1)) == 0 && kSlots > Src::kPerStep, "ring size");
  constexpr uint32_t kNeed = Src::kPerStep;
  const LcgInc incv = vgpr_increment(inc);
  uint32_t head = 0, cnt = 0;
  for (uint32_t step = 0; step < steps; ++step) {
    queue_fill<Src, kSlots, kCols>(src, g, incv, zig, q, head, cnt, kNeed);
    Src::template apply<kSlots, kCols>(m, q, head);
    head = (head + kNeed) & (kSlots - 1);
    cnt -= kNeed;
    hook(step, m);
  }
}

// queue_run with the generator one draw ahead (queue_fill_ahead).
template <class Src, uint32_t kSlots, int kCols, class Model, class Hook = NoHook>
__device__ __forceinline__ void queue_run_ahead(Src& src, Model& m, const U128 s0, const U128 inc,
                                                uint32_t steps, const typename Src::Tables* zig,
                                                typename Src::Rec (*q)[kCols],
                                                Hook hook = Hook()) {
  static_assert((kSlots & (kSlots - 1)) == 0 && kSlots > Src::kPerStep, "ring size");
  constexpr uint32_t kNeed = Src::kPerStep;
  const LcgInc incv = vgpr_increment(inc);
  DrawAhead<Src> da(s0, zig);
  uint32_t head = 0, cnt = 0;
  for (uint32_t step = 0; step < steps; ++step) {
    queue_fill_ahead<Src, kSlots, kCols>(src, da, incv, zig, q, head, cnt, kNeed);
    Src::template apply<kSlots, kCols>(m, q, head);
    head = (head + kNeed) & (kSlots - 1);
    cnt -= kNeed;
    hook(step, m);
  }
}

template <class Src, class Model, uint32_t kSlots, int kCols, class Const, class Hook = NoHook>
__device__ __forceinline__ double queue_path(const Const& c, Pcg& g, const U128 inc,
                                             uint32_t steps,
                                             const typename Src::Tables* zig,
                                             typename Src::Rec (*q)[kCols], bool mirror = false,
                                             Hook hook = Hook()) {
  Src src(c);
  Model m(c, mirror);
  queue_run<Src, kSlots, kCols>(src, m, g, inc, steps, zig, q, hook);
  return m.terminal();
}

}  // namespace stocha
