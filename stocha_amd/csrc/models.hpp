// models.hpp -- per-path steppers for GBM, Heston (full-truncation Euler)
// and Merton jump-diffusion, in the reference's evaluation order (FP64) and in
// a log-space FP32 form.
//
// FP64 forms restate, operation for operation (this file is compiled with
// -ffp-contract=off, so a*b+c is two roundings as in Rust):
//   GBM     gbm.rs:43-45 (constants), :65-73 (step)
//   Heston  heston.rs:45-47 (constants), :59-71 (step)
//   Merton  jump_diffusion.rs:49-64 (constants), :74-89 (step)
// The host computes the per-launch constants (drift, diffusion, ...) with the
// same expressions, so device and reference agree bit-for-bit up to the libm
// exp used per step.
//
// FP32 forms step the log-price: log S += increment, with the increment
// computed in FP32 (explicit fmaf), and exponentiate once when S is needed.
// They draw exactly the same normals/uniforms, in the same order, as the FP64
// forms (see rng.hpp), so their error against the FP64 result is rounding
// only.
//
// Each model separates the arithmetic of one step (apply*, given its normals)
// from how the normals are drawn: step() draws with the per-normal ziggurat
// loop; the state-machine drivers in fused.hip feed draws one at a time.
#pragma once
#include "rng.hpp"

namespace stocha {

enum ModelKind : int { kGbm = 0, kHeston = 1, kMerton = 2, kHestonQe = 3, kHullWhite = 4 };

struct GbmConst {
  double s0, drift, diffusion;
  float drift_f, diffusion_f;
};

struct HestonConst {
  double s0, v0, mu, kappa, theta, xi, rho, rho_comp, dt, sqrt_dt;
  // FP32 log-space constants
  float v0_f, rho_f, rho_comp_f, kappa_dt_f, theta_f, xi_sqdt_f, mu_dt_f, half_dt_f, sqrt_dt_f;
};

struct MertonConst {
  double s0, drift, diffusion, jump_prob, mu_j, sigma_j;
  // next_f64() < jump_prob  <=>  (u >> 11) < jump_t53: next_f64 is m * 2^-53
  // exactly, so the comparison is m < jump_prob * 2^53 (an exact scaling),
  // i.e. m < ceil(jump_prob * 2^53), clamped to [0, 2^53].
  uint64_t jump_t53;
  float drift_f, diffusion_f, mu_j_f, sigma_j_f;
};

// Heston QE (heston.rs:91-166).  The loop-invariant parts of the
// reference's per-step expressions, evaluated once on the host with the same
// operations in the same order.
struct HestonQeConst {
  double s0, v0, kappa, xi;
  double e_kdt;       // exp(-kappa dt)
  double one_m_e;     // 1 - e_kdt
  double m_const;     // k1 * (1 - e_kdt) / kappa, k1 = kappa theta
  double s2_const;    // k1 xi xi (1 - e_kdt)^2 / (2 kappa kappa)
  double k0s, k1s, k2s, k3s, k4s;  // Andersen's log-price coefficients
};

// Hull-White one-factor, exact transition (hull_white.rs:26-52); s0 is r0.
struct HullWhiteConst {
  double s0, e, mean_reversion, vol;
};

// Stream budget per path (reference block sizes; QE keeps Euler's,
// heston.rs:97; Hull-White hull_white.rs:38).
__host__ __device__ inline uint64_t block_size(int model, uint64_t steps) {
  return model == kGbm                            ? (steps + 1024) * 2
         : model == kHeston || model == kHestonQe ? (steps + 1024) * 6
         : model == kHullWhite                    ? (steps + 1024) * 4
                                                  : (steps + 1024) * 8;
}

// ---------------------------------------------------------------- FP64 ----

struct Gbm64 {
  static constexpr bool kFp32 = false;
  GbmConst c;
  double s;
  bool neg;
  __device__ __forceinline__ Gbm64(const GbmConst& k, bool negate) : c(k), s(k.s0), neg(negate) {}
  __device__ __forceinline__ void apply(double n) {
    n = neg ? -n : n;
    s *= exp(c.drift + c.diffusion * n);
  }
  __device__ __forceinline__ double step(Pcg& g, const U128 inc, const ZigLds* z) {
    apply(zig_normal64(g, inc, z));
    return s;
  }
  __device__ __forceinline__ double terminal() const { return s; }
};

// max(x, 0) as one v_max instruction: fmax's lowering first quiets x with a
// second v_max.  Same result for every x (v_max returns 0 for a NaN x, as
// fmax does).
__device__ __forceinline__ float max0(float x) {
  float r;
  asm("v_max_f32 %0, 0, %1" : "=v"(r) : "v"(x));
  return r;
}
__device__ __forceinline__ double max0(double x) {
  double r;
  asm("v_max_f64 %0, 0, %1" : "=v"(r) : "v"(x));
  return r;
}

struct Heston64 {
  static constexpr bool kFp32 = false;
  HestonConst c;
  double s, v;
  __device__ __forceinline__ Heston64(const HestonConst& k, bool) : c(k), s(k.s0), v(k.v0) {}
  __device__ __forceinline__ void apply(double dw1, double dz) {
    double dw2 = c.rho * dw1 + c.rho_comp * dz;
    double v_plus = max0(v);
    double sqrt_v = sqrt(v_plus);
    v += c.kappa * (c.theta - v_plus) * c.dt + c.xi * sqrt_v * c.sqrt_dt * dw2;
    s *= exp((c.mu - 0.5 * v_plus) * c.dt + sqrt_v * c.sqrt_dt * dw1);
  }
  __device__ __forceinline__ double step(Pcg& g, const U128 inc, const ZigLds* z) {
    double dw1 = zig_normal64(g, inc, z);
    double dz = zig_normal64(g, inc, z);
    apply(dw1, dz);
    return s;
  }
  __device__ __forceinline__ double terminal() const { return s; }
};

// Heston QE, log-price part (heston.rs:152-157); the variance recursion runs
// in the draw source (HestonQeSrc), which hands over V(t+dt) with the
// step's asset normal.  v tracks the same recursion (v0, then each V(t+dt)).
// kRoundTrip (the path matrices): the state is S and every step takes
// ln S and S = exp(clamp(log_s)) as the reference does, so each stored value
// is the reference's to the device's log/exp.  Without it (terminal prices,
// fused payoffs, greeks_fd) the state is ln S and the pair is dropped, S
// being formed only when read: log(exp(x)) differs from x by at most half an
// ulp of ln S per step, so S_T moves by at most steps x 1e-16 relative.
template <bool kRoundTrip>
struct HestonQeModel {
  static constexpr bool kFp32 = false;
  HestonQeConst c;
  double x, v;  // kRoundTrip: S; otherwise ln S
  __device__ __forceinline__ HestonQeModel(const HestonQeConst& k, bool)
      : c(k), x(kRoundTrip ? k.s0 : log(k.s0)), v(k.v0) {}
  __device__ __forceinline__ void apply_qe(double v_next, double z1) {
    const double v_plus = fmax(v, 0.0);
    double var_term = c.k3s * v_plus + c.k4s * v_next;
    var_term = fmax(var_term, 0.0);
    const double ls = kRoundTrip ? log(x) : x;
    double log_s = ls + c.k0s + c.k1s * v_plus + c.k2s * v_next + sqrt(var_term) * z1;
    // f64::clamp(-500, 500) keeps a NaN
    log_s = log_s < -500.0 ? -500.0 : (log_s > 500.0 ? 500.0 : log_s);
    x = kRoundTrip ? exp(log_s) : log_s;
    v = v_next;
  }
  __device__ __forceinline__ double terminal() const { return kRoundTrip ? x : exp(x); }
};
using HestonQe64 = HestonQeModel<false>;
using HestonQe64Path = HestonQeModel<true>;

// hull_white.rs:47-49: r(t+dt) = r(t) e + mean_reversion + vol z
struct HullWhite64 {
  static constexpr bool kFp32 = false;
  HullWhiteConst c;
  double s;  // the short rate
  __device__ __forceinline__ HullWhite64(const HullWhiteConst& k, bool) : c(k), s(k.s0) {}
  __device__ __forceinline__ void apply(double z) { s = s * c.e + c.mean_reversion + c.vol * z; }
  __device__ __forceinline__ double terminal() const { return s; }
};

struct Merton64 {
  static constexpr bool kFp32 = false;
  MertonConst c;
  double s;
  __device__ __forceinline__ Merton64(const MertonConst& k, bool) : c(k), s(k.s0) {}
  __device__ __forceinline__ void apply(double zz, bool jump, double zj) {
    double diffusion_return = c.drift + c.diffusion * zz;
    double jump_return = 0.0;
    if (jump) jump_return = c.mu_j + c.sigma_j * zj;
    s *= exp(diffusion_return + jump_return);
  }
  __device__ __forceinline__ double step(Pcg& g, const U128 inc, const ZigLds* z) {
    double zz = zig_normal64(g, inc, z);
    bool jump = (g.next(inc) >> 11) < c.jump_t53;  // next_f64() < jump_prob
    double zj = jump ? zig_normal64(g, inc, z) : 0.0;
    apply(zz, jump, zj);
    return s;
  }
  // jump_return precomputed: 0.0, or mu_j + sigma_j * z_j (jump_diffusion.rs:80-85)
  __device__ __forceinline__ double jump_return(double zj) const { return c.mu_j + c.sigma_j * zj; }
  __device__ __forceinline__ void apply_jr(double zz, double jr) {
    double diffusion_return = c.drift + c.diffusion * zz;
    s *= exp(diffusion_return + jr);
  }
  __device__ __forceinline__ double terminal() const { return s; }
};

// ---------------------------------------------------------------- FP32 ----
// step() returns the log-price increment's running sum; s0 * exp(ls) is S.

struct Gbm32 {
  static constexpr bool kFp32 = true;
  GbmConst c;
  float ls;
  bool neg;
  __device__ __forceinline__ Gbm32(const GbmConst& k, bool negate) : c(k), ls(0.f), neg(negate) {}
  __device__ __forceinline__ void apply(float n) {
    n = neg ? -n : n;
    ls += fmaf(c.diffusion_f, n, c.drift_f);
  }
  __device__ __forceinline__ float step(Pcg& g, const U128 inc, const ZigLds* z) {
    apply(zig_normal32(g, inc, z));
    return ls;
  }
  __device__ __forceinline__ double terminal() const { return c.s0 * exp((double)ls); }
};

struct Heston32 {
  static constexpr bool kFp32 = true;
  HestonConst c;
  float ls, v;
  __device__ __forceinline__ Heston32(const HestonConst& k, bool) : c(k), ls(0.f), v(k.v0_f) {}
  __device__ __forceinline__ void apply(float dw1, float dz) {
    float dw2 = fmaf(c.rho_f, dw1, c.rho_comp_f * dz);
    float v_plus = max0(v);
    // v_sqrt_f32 (1 ulp); the FP32 form's error bar is 1e-4 of the FP64 price
    float sqrt_v = __builtin_amdgcn_sqrtf(v_plus);
    // v += kappa*(theta - v+)*dt + xi*sqrt(v+)*sqrt(dt)*dw2
    v = fmaf(c.kappa_dt_f, c.theta_f - v_plus, fmaf(c.xi_sqdt_f * sqrt_v, dw2, v));
    // log S += (mu - v+/2)*dt + sqrt(v+)*sqrt(dt)*dw1
    ls += fmaf(-c.half_dt_f, v_plus, fmaf(sqrt_v * c.sqrt_dt_f, dw1, c.mu_dt_f));
  }
  __device__ __forceinline__ float step(Pcg& g, const U128 inc, const ZigLds* z) {
    float dw1 = zig_normal32(g, inc, z);
    float dz = zig_normal32(g, inc, z);
    apply(dw1, dz);
    return ls;
  }
  __device__ __forceinline__ double terminal() const { return c.s0 * exp((double)ls); }
};

struct Merton32 {
  static constexpr bool kFp32 = true;
  MertonConst c;
  float ls;
  __device__ __forceinline__ Merton32(const MertonConst& k, bool) : c(k), ls(0.f) {}
  __device__ __forceinline__ void apply(float zz, bool jump, float zj) {
    float r = fmaf(c.diffusion_f, zz, c.drift_f);
    if (jump) r += fmaf(c.sigma_j_f, zj, c.mu_j_f);
    ls += r;
  }
  __device__ __forceinline__ float step(Pcg& g, const U128 inc, const ZigLds* z) {
    float zz = zig_normal32(g, inc, z);
    bool jump = (g.next(inc) >> 11) < c.jump_t53;  // same comparison as FP64
    float zj = jump ? zig_normal32(g, inc, z) : 0.f;
    apply(zz, jump, zj);
    return ls;
  }
  __device__ __forceinline__ float jump_return(float zj) const {
    return fmaf(c.sigma_j_f, zj, c.mu_j_f);
  }
  __device__ __forceinline__ void apply_jr(float zz, float jr) {
    ls += fmaf(c.diffusion_f, zz, c.drift_f) + jr;
  }
  __device__ __forceinline__ double terminal() const { return c.s0 * exp((double)ls); }
};

}  // namespace stocha
