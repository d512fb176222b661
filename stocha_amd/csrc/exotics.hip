// exotics.hip -- path-dependent options on GBM: barrier (barrier.rs), Asian
// (asian.rs) and lookback (lookback.rs), Monte Carlo on the device and the
// reference's closed forms on the host.
//
// The Monte Carlo forms all step GBM paths on the block-split streams of
// gbm_paths (block (steps + 1024) * 2, s *= exp(drift + diffusion z)) and
// evaluate a functional of the path: the first step at or beyond a barrier,
// the running maximum / minimum, the arithmetic and log sums.  The device
// runs them with the queue schedule (GbmSrc, FP64 arithmetic of the
// reference) and a per-step observer in the model; payoffs are summed in
// registers and reduced in a fixed order (sums.hpp), or, for the Asian
// control variate, stored per path and reduced twice (means, then the
// centred cross sums the reference's estimator uses).
//
// Paths are always stepped to T: the reference stops a knocked-out path at
// its first hit, but its payoff no longer depends on the draws after it, and
// every path has its own stream, so the values are the same.
#include <math.h>
#include <string.h>

#include "drivers.hpp"
#include "host.hpp"
#include "sums.hpp"

namespace stocha {
namespace {

// ------------------------------------------------------------ functionals --
// barrier.rs:286-372: first monitoring step j >= 1 with S_j at or beyond the
// barrier (0 when the spot already is), then the payoff and the rebate.
struct BarrierFn {
  double h, k, rebate_t, r, dt, rebate, discount;
  int up, out, call, rebate_mode;  // rebate_mode 0 none, 1 at hit (KO), 2 at T
  int already;
  struct State {
    int hit;
    uint32_t j;
  };
  __device__ __forceinline__ State init() const { return State{already ? 0 : -1, 0u}; }
  __device__ __forceinline__ void observe(State& st, double s) const {
    st.j += 1;
    const bool beyond = up ? s >= h : s <= h;  // check_barrier (barrier.rs:402-407)
    st.hit = (st.hit < 0 && beyond) ? (int)st.j : st.hit;
  }
  __device__ __forceinline__ double value(const State& st, double s) const {
    const bool knocked = st.hit >= 0;
    double intrinsic = call ? s - k : k - s;  // barrier.rs:409-414
    intrinsic = intrinsic > 0.0 ? intrinsic : 0.0;
    const double option = out ? (knocked ? 0.0 : intrinsic) : (knocked ? intrinsic : 0.0);
    // rebate_path_pv (barrier.rs:374-400), divided by the global discount
    double rebate_pv = 0.0;
    if (rebate_mode == 1) {
      rebate_pv = knocked ? rebate * exp(-r * ((double)st.hit * dt)) : 0.0;
    } else if (rebate_mode == 2) {
      rebate_pv = (out ? knocked : !knocked) ? rebate_t : 0.0;
    }
    return option + rebate_pv / discount;
  }
};

// lookback.rs:158-210
struct LookbackFn {
  double k, max0, min0;
  int floating, call;
  struct State {
    double mx, mn;
  };
  __device__ __forceinline__ State init() const { return State{max0, min0}; }
  __device__ __forceinline__ void observe(State& st, double s) const {
    st.mx = s > st.mx ? s : st.mx;
    st.mn = s < st.mn ? s : st.mn;
  }
  __device__ __forceinline__ double value(const State& st, double s) const {
    if (floating) return call ? s - st.mn : st.mx - s;
    double v = call ? st.mx - k : k - st.mn;
    return v > 0.0 ? v : 0.0;
  }
};

// asian.rs:208-250 (and :318-345 for a seasoned floating strike): the
// arithmetic and log sums of S_1..S_n; value2 gives the geometric payoff the
// control variate needs.
struct AsianFn {
  double k, n_div, a_spent, w_spent, w_remain, log_a_spent;
  int fixed, call, geometric, seasoned;
  struct State {
    double sum, sum_log;
  };
  __device__ __forceinline__ State init() const { return State{0.0, 0.0}; }
  __device__ __forceinline__ void observe(State& st, double s) const {
    st.sum += s;
    st.sum_log += log(s);
  }
  __device__ __forceinline__ static double pos(double v) { return v > 0.0 ? v : 0.0; }
  __device__ __forceinline__ double pay(double avg, double s) const {
    if (fixed) return call ? pos(avg - k) : pos(k - avg);
    return call ? pos(s - avg) : pos(avg - s);
  }
  // asian.rs:232-233: averages divide by the step count n_div
  __device__ __forceinline__ double value(const State& st, double s) const {
    if (seasoned) {  // asian.rs:329-341
      double avg_t;
      if (geometric) {
        const double avg_remain = exp(st.sum_log / n_div);
        avg_t = exp(log_a_spent * w_spent + log(avg_remain) * w_remain);
      } else {
        avg_t = w_spent * a_spent + w_remain * (st.sum / n_div);
      }
      return call ? pos(s - avg_t) : pos(avg_t - s);
    }
    return pay(st.sum / n_div, s);
  }
  __device__ __forceinline__ double value2(const State& st, double s) const {
    return pay(exp(st.sum_log / n_div), s);
  }
};

// GBM step of the reference (Gbm64) with the functional observing S_j.
template <class F>
struct GbmObserved {
  static constexpr bool kFp32 = false;
  Gbm64 g;
  const F* f;
  typename F::State st;
  __device__ __forceinline__ GbmObserved(const GbmConst& c, const F* fn)
      : g(c, false), f(fn), st(fn->init()) {}
  __device__ __forceinline__ void apply(double n) {
    g.apply(n);
    f->observe(st, g.s);
  }
  __device__ __forceinline__ double terminal() const { return g.s; }
};

template <class Src, class F, uint32_t kSlots>
__device__ __forceinline__ GbmObserved<F> run_path(const GbmConst& c, const F* f,
                                                   const StreamDesc& sd, uint64_t p,
                                                   uint32_t steps, const ZigLds64* zig,
                                                   typename Src::Rec (*q)[kFusedBlock]) {
  Pcg g{stream_at(&sd, p)};
  Src src(c);
  GbmObserved<F> m(c, f);
  queue_run<Src, kSlots, kFusedBlock>(src, m, g, sd.inc, steps, zig, q);
  return m;
}

// Sum over paths of F::value (one accumulator).
template <class F>
__global__ __launch_bounds__(kFusedBlock) __attribute__((amdgpu_waves_per_eu(4))) void
exotic_sum_kernel(const GbmConst c, const F f, const StreamDesc sd, const uint64_t n_paths,
                  const uint32_t steps, double* __restrict__ partials,
                  unsigned* __restrict__ counter, double* __restrict__ out) {
  using Src = GbmSrc<false>;
  __shared__ ZigLds64 zig;
  __shared__ Src::Rec q[8][kFusedBlock];
  __shared__ double red[kFusedBlock / 64][kMaxAcc];
  zig_load(&zig);
  __syncthreads();
  double acc[kMaxAcc];
#pragma unroll
  for (int i = 0; i < kMaxAcc; ++i) acc[i] = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * kFusedBlock;
  for (uint64_t base = (uint64_t)blockIdx.x * kFusedBlock; base < n_paths; base += stride) {
    const uint64_t p = base + threadIdx.x;
    const bool valid = p < n_paths;
    const GbmObserved<F> m = run_path<Src, F, 8>(c, &f, sd, valid ? p : 0, steps, &zig, q);
    if (valid) acc[0] += f.value(m.st, m.g.s);
  }
  finish_sums(acc, 1, red, partials, counter, out);
}

// Per-path (arithmetic payoff, geometric payoff) for the control variate.
__global__ __launch_bounds__(kFusedBlock) __attribute__((amdgpu_waves_per_eu(4))) void
asian_pairs_kernel(const GbmConst c, const AsianFn f, const StreamDesc sd,
                   const uint64_t n_paths, const uint32_t steps, double2* __restrict__ out) {
  using Src = GbmSrc<false>;
  __shared__ ZigLds64 zig;
  __shared__ Src::Rec q[8][kFusedBlock];
  zig_load(&zig);
  __syncthreads();
  const uint64_t stride = (uint64_t)gridDim.x * kFusedBlock;
  for (uint64_t base = (uint64_t)blockIdx.x * kFusedBlock; base < n_paths; base += stride) {
    const uint64_t p = base + threadIdx.x;
    const bool valid = p < n_paths;
    const GbmObserved<AsianFn> m =
        run_path<Src, AsianFn, 8>(c, &f, sd, valid ? p : 0, steps, &zig, q);
    if (valid) out[p] = make_double2(f.value(m.st, m.g.s), f.value2(m.st, m.g.s));
  }
}

// Fixed-order sums over the stored pairs: mode 0 -> (sum a, sum g); mode 1 ->
// (sum (a - ma)(g - mg), sum (g - mg)^2), asian.rs:266-271.
__global__ __launch_bounds__(256) void pair_sums_kernel(const double2* __restrict__ ag,
                                                        const uint64_t n, const double ma,
                                                        const double mg, const int mode,
                                                        double* __restrict__ partials) {
  __shared__ double red[kFusedBlock / 64][kMaxAcc];
  double acc[2] = {0.0, 0.0};
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const double2 v = ag[i];
    if (mode == 0) {
      acc[0] += v.x;
      acc[1] += v.y;
    } else {
      acc[0] += (v.x - ma) * (v.y - mg);
      acc[1] += (v.y - mg) * (v.y - mg);
    }
  }
  const double s = block_total(acc, 2, red);
  if (threadIdx.x < 2) partials[(uint64_t)blockIdx.x * 2 + threadIdx.x] = s;
}

template <class F>
int run_sum(const GbmConst& c, const F& f, uint64_t seed, uint64_t n_paths, uint64_t steps,
            double* total, const Err& err) {
  if (steps > 0xffffffffull) return err.set(STOCHA_EINVAL, "n_steps must be below 2^32");
  hipStream_t st = globals().stream;
  StreamDesc sd;
  make_stream_desc(seed, block_size(kGbm, steps), &sd);
  const unsigned grid = fused_grid(n_paths);
  Workspace* ws = nullptr;
  if (int rc = workspace(0, sizeof(double) * kMaxAcc, &ws, err)) return rc;
  KernelTimer timer(st);
  timer.start();
  exotic_sum_kernel<F><<<grid, kFusedBlock, 0, st>>>(c, f, sd, n_paths, (uint32_t)steps,
                                                     ws->sums_part, ws->counter, ws->host);
  STOCHA_HIP(hipGetLastError());
  timer.stop(1);
  STOCHA_HIP(hipStreamSynchronize(st));
  *total = ws->host[0];
  return STOCHA_OK;
}

// ------------------------------------------------------- host closed forms --
// bs.rs:3-25: Abramowitz & Stegun 7.1.26 erf and the normal cdf on it.
double as_erf(double x) {
  const double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741, a4 = -1.453152027,
               a5 = 1.061405429, p = 0.3275911;
  const double sign = x < 0.0 ? -1.0 : 1.0;
  x = fabs(x);
  const double t = 1.0 / (1.0 + p * x);
  const double y = 1.0 - (((((a5 * t + a4) * t) + a3) * t + a2) * t + a1) * t * exp(-x * x);
  return sign * y;
}
double norm_cdf(double x) { return 0.5 * (1.0 + as_erf(x * M_SQRT1_2)); }
double norm_pdf(double x) { return (1.0 / sqrt(2.0 * M_PI)) * exp(-0.5 * x * x); }
double fmax0(double v) { return v > 0.0 ? v : 0.0; }  // f64::max(0.0): NaN -> 0

// bs.rs:31-41
double bs_price_div(double s, double k, double r, double q, double t, double sigma, bool call) {
  const double d1 = (log(s / k) + (r - q + 0.5 * sigma * sigma) * t) / (sigma * sqrt(t));
  const double d2 = d1 - sigma * sqrt(t);
  const double df = exp(-r * t);
  const double qf = exp(-q * t);
  if (call) return s * qf * norm_cdf(d1) - k * df * norm_cdf(d2);
  return k * df * norm_cdf(-d2) - s * qf * norm_cdf(-d1);
}

struct BarrierSpec {
  double s, k, r, q, sigma, t, barrier, rebate;
  bool up, out, call, rebate_at_hit;
  bool has_mon;
  uint32_t n_mon;
};

// barrier.rs:49-64
double hit_probability(double s, double h, double r, double q, double sigma, double t, bool up) {
  const double nu = r - q - 0.5 * sigma * sigma;
  const double mu = nu / (sigma * sigma);
  const double sqrt_t = sqrt(t);
  double h1, h2;
  if (up) {
    const double a = log(s / h);
    h1 = (a + nu * t) / (sigma * sqrt_t);
    h2 = (a - nu * t) / (sigma * sqrt_t);
  } else {
    const double a = log(h / s);
    h1 = (a - nu * t) / (sigma * sqrt_t);
    h2 = (a + nu * t) / (sigma * sqrt_t);
  }
  const double p = norm_cdf(h1) + pow(h / s, 2.0 * mu) * norm_cdf(h2);
  return p < 0.0 ? 0.0 : (p > 1.0 ? 1.0 : p);  // f64::clamp keeps a NaN
}

// barrier.rs:69-86
double rebate_paid_at_hit(double s, double h, double rebate, double r, double q, double sigma,
                          double t, bool up) {
  if (rebate == 0.0) return 0.0;
  const double nu = r - q - 0.5 * sigma * sigma;
  const double sigma2 = sigma * sigma;
  const double mu = nu / sigma2;
  const double lambda = sqrt(mu * mu + 2.0 * r / sigma2);
  const double sqrt_t = sqrt(t);
  const double log_ratio = up ? log(h / s) : log(s / h);
  const double d1 = (-log_ratio + lambda * sigma2 * t) / (sigma * sqrt_t);
  const double d2 = (-log_ratio - lambda * sigma2 * t) / (sigma * sqrt_t);
  const double ratio = h / s;
  return rebate * (pow(ratio, mu - lambda) * norm_cdf(d1) + pow(ratio, mu + lambda) * norm_cdf(d2));
}

// barrier.rs:90-127
double rebate_value(const BarrierSpec& p) {
  if (p.rebate == 0.0) return 0.0;
  const bool already = p.up ? p.s >= p.barrier : p.s <= p.barrier;
  if (p.out && p.rebate_at_hit) {
    return already ? p.rebate
                   : rebate_paid_at_hit(p.s, p.barrier, p.rebate, p.r, p.q, p.sigma, p.t, p.up);
  }
  if (p.out) {
    const double prob =
        already ? 1.0 : hit_probability(p.s, p.barrier, p.r, p.q, p.sigma, p.t, p.up);
    return p.rebate * exp(-p.r * p.t) * prob;
  }
  if (already) return 0.0;
  const double prob = hit_probability(p.s, p.barrier, p.r, p.q, p.sigma, p.t, p.up);
  return p.rebate * exp(-p.r * p.t) * (1.0 - prob);
}

// barrier.rs:130-143 (Broadie-Glasserman-Kou)
constexpr double kBgkBeta = 0.5825971579390106;
double bgk_adjusted_barrier(double h, double sigma, double t, uint32_t n, bool up) {
  if (n == 0) return h;
  const double dt = t / (double)n;
  const double shift = kBgkBeta * sigma * sqrt(dt);
  return up ? h * exp(shift) : h * exp(-shift);
}

// barrier.rs:270-284
double knocked_value(const BarrierSpec& p) {
  return p.out ? 0.0 : bs_price_div(p.s, p.k, p.r, p.q, p.t, p.sigma, p.call);
}

// barrier.rs:145-268 (Reiner-Rubinstein); false = None
bool barrier_analytical(const BarrierSpec& p, double* out) {
  const double s = p.s, k = p.k;
  const double h = (p.has_mon && p.n_mon > 0)
                       ? bgk_adjusted_barrier(p.barrier, p.sigma, p.t, p.n_mon, p.up)
                       : p.barrier;
  const double r = p.r, q = p.q, sigma = p.sigma, t = p.t;
  if (t <= 0.0 || sigma <= 0.0) return false;
  if (p.up ? s >= h : s <= h) {
    *out = knocked_value(p) + rebate_value(p);
    return true;
  }
  const double sqrt_t = sqrt(t);
  const double mu = (r - q - 0.5 * sigma * sigma) / (sigma * sigma);
  auto d = [&](double x, double y) { return (log(x) + y * sigma * sigma * t) / (sigma * sqrt_t); };
  auto a = [&](double phi) {
    const double d1 = d(s / k, 1.0 + mu);
    const double d2 = d1 - sigma * sqrt_t;
    return phi * (s * (exp(-q * t)) * norm_cdf(phi * d1) - k * (exp(-r * t)) * norm_cdf(phi * d2));
  };
  auto b = [&](double phi) {
    const double d1 = d(s / h, 1.0 + mu);
    const double d2 = d1 - sigma * sqrt_t;
    return phi * (s * (exp(-q * t)) * norm_cdf(phi * d1) - k * (exp(-r * t)) * norm_cdf(phi * d2));
  };
  auto c = [&](double phi, double eta) {
    const double pw = pow(h / s, 2.0 * (1.0 + mu));
    const double d1 = d(h * h / (s * k), 1.0 + mu);
    const double d2 = d1 - sigma * sqrt_t;
    return phi * (s * (exp(-q * t)) * pw * norm_cdf(eta * d1) -
                  k * (exp(-r * t)) * pow(h / s, 2.0 * mu) * norm_cdf(eta * d2));
  };
  auto dd = [&](double phi, double eta) {
    const double pw = pow(h / s, 2.0 * (1.0 + mu));
    const double d1 = d(h / s, 1.0 + mu);
    const double d2 = d1 - sigma * sqrt_t;
    return phi * (s * (exp(-q * t)) * pw * norm_cdf(eta * d1) -
                  k * (exp(-r * t)) * pow(h / s, 2.0 * mu) * norm_cdf(eta * d2));
  };
  double price;
  if (!p.up && p.out && p.call) {
    price = h <= k ? a(1.0) - c(1.0, 1.0) : b(1.0) - dd(1.0, 1.0);
  } else if (!p.up && !p.out && p.call) {
    price = h <= k ? c(1.0, 1.0) : a(1.0) - b(1.0) + dd(1.0, 1.0);
  } else if (p.up && p.out && p.call) {
    price = h > k ? a(1.0) - b(1.0) + c(1.0, -1.0) - dd(1.0, -1.0) : 0.0;
  } else if (p.up && !p.out && p.call) {
    price = h > k ? b(1.0) - c(1.0, -1.0) + dd(1.0, -1.0) : a(1.0);
  } else if (!p.up && p.out) {  // put
    price = h < k ? a(-1.0) - b(-1.0) + c(-1.0, 1.0) - dd(-1.0, 1.0) : 0.0;
  } else if (!p.up) {
    price = h < k ? b(-1.0) - c(-1.0, 1.0) + dd(-1.0, 1.0) : a(-1.0);
  } else if (p.out) {
    price = h >= k ? a(-1.0) - c(-1.0, -1.0) : b(-1.0) - dd(-1.0, -1.0);
  } else {
    price = h >= k ? c(-1.0, -1.0) : a(-1.0) - b(-1.0) + dd(-1.0, -1.0);
  }
  *out = fmax0(price) + rebate_value(p);
  return true;
}

// barrier.rs:286-372 on the device
int barrier_mc(const BarrierSpec& p, uint64_t n_paths, uint64_t n_steps, uint64_t seed,
               double* price, const Err& err) {
  const double dt = p.t / (double)n_steps;
  const double discount = exp(-p.r * p.t);
  BarrierFn f{};
  f.h = p.barrier;
  f.k = p.k;
  f.up = p.up;
  f.out = p.out;
  f.call = p.call;
  f.already = p.up ? p.s >= p.barrier : p.s <= p.barrier;
  f.rebate = p.rebate;
  f.r = p.r;
  f.dt = dt;
  f.discount = discount;
  f.rebate_mode = p.rebate == 0.0 ? 0 : (p.out && p.rebate_at_hit) ? 1 : 2;
  f.rebate_t = p.rebate * exp(-p.r * ((double)n_steps * dt));
  const GbmConst c = gbm_const(p.s, p.r, p.sigma, p.t, n_steps, p.q);
  double total = 0.0;
  if (int rc = run_sum(c, f, seed, n_paths, n_steps, &total, err)) return rc;
  *price = discount * total / (double)n_paths;
  return STOCHA_OK;
}

// ---- Asian (asian.rs) ----
struct AsianSpec {
  double s, k, r, q, sigma, t;
  uint64_t n_steps;
  bool geometric, fixed, call;
  bool has_avg, has_elapsed;
  double running_avg, time_elapsed;
};

// asian.rs:44-50
bool extract_seasoning(const AsianSpec& p, double* a, double* t1) {
  if (p.has_avg && p.has_elapsed && p.time_elapsed > 0.0 && p.running_avg >= 0.0) {
    *a = p.running_avg;
    *t1 = p.time_elapsed;
    return true;
  }
  return false;
}

// asian.rs:53-59
double expected_geometric_avg(double s, double r, double q, double sigma, double t1,
                              double t_total, double geo_avg_spent) {
  const double remaining = t_total - t1;
  const double log_spent_part = (t1 / t_total) * log(geo_avg_spent);
  const double mean_y =
      (remaining * log(s) + (r - q - 0.5 * sigma * sigma) * remaining * remaining / 2.0) / t_total;
  const double var_y =
      sigma * sigma * (remaining * remaining * remaining) / (3.0 * t_total * t_total);
  return exp(log_spent_part + mean_y + 0.5 * var_y);
}

// asian.rs:63-72
double expected_arithmetic_avg(double s, double r, double q, double t1, double t_total,
                               double a_spent) {
  const double remaining = t_total - t1;
  const double mu = r - q;
  const double e_a_remain = fabs(mu) < 1e-12 ? s : s * (exp(mu * remaining) - 1.0) / (mu * remaining);
  return (t1 / t_total) * a_spent + (remaining / t_total) * e_a_remain;
}

// asian.rs:78-137 (Kemna-Vorst); false = None
bool asian_geometric_fixed_analytical(AsianSpec p, double* out) {
  if (p.t <= 0.0 || p.sigma <= 0.0) return false;
  if (!p.geometric || !p.fixed) return false;
  double a_spent, t1;
  if (extract_seasoning(p, &a_spent, &t1)) {
    if (t1 >= p.t) {
      *out = p.call ? fmax0(a_spent - p.k) : fmax0(p.k - a_spent);
      return true;
    }
    const double remaining = p.t - t1;
    const double k_star = (p.t * p.k - t1 * a_spent) / remaining;
    const double scale = remaining / p.t;
    if (k_star > 0.0) {
      AsianSpec pp = p;
      pp.k = k_star;
      pp.t = remaining;
      pp.has_avg = pp.has_elapsed = false;
      double v;
      if (!asian_geometric_fixed_analytical(pp, &v)) return false;
      *out = scale * v;
      return true;
    }
    if (!p.call) {
      *out = 0.0;
      return true;
    }
    const double e_geo = expected_geometric_avg(p.s, p.r, p.q, p.sigma, t1, p.t, a_spent);
    *out = fmax0(exp(-p.r * remaining) * (e_geo - p.k));
    return true;
  }
  const double sigma_g = p.sigma / sqrt(3.0);
  const double mu_g = 0.5 * (p.r - p.q - 0.5 * p.sigma * p.sigma) + 0.5 * 0.5 * sigma_g * sigma_g;
  const double d1 = (log(p.s / p.k) + (mu_g + 0.5 * sigma_g * sigma_g) * p.t) / (sigma_g * sqrt(p.t));
  const double d2 = d1 - sigma_g * sqrt(p.t);
  const double df = exp(-p.r * p.t);
  const double price =
      p.call ? p.s * exp((mu_g - p.r) * p.t) * norm_cdf(d1) - p.k * df * norm_cdf(d2)
             : p.k * df * norm_cdf(-d2) - p.s * exp((mu_g - p.r) * p.t) * norm_cdf(-d1);
  *out = fmax0(price);
  return true;
}

AsianFn asian_fn(const AsianSpec& p, double n_div) {
  AsianFn f{};
  f.k = p.k;
  f.n_div = n_div;
  f.fixed = p.fixed;
  f.call = p.call;
  f.geometric = p.geometric;
  return f;
}

// asian.rs:143-290 on the device (asian_mc_floating_seasoned :293-355)
int asian_mc(AsianSpec p, uint64_t n_paths, uint64_t seed, double* price, const Err& err) {
  double a_spent, t1;
  if (extract_seasoning(p, &a_spent, &t1)) {
    if (t1 >= p.t) {
      *price = p.fixed ? (p.call ? fmax0(a_spent - p.k) : fmax0(p.k - a_spent)) : 0.0;
      return STOCHA_OK;
    }
    if (p.fixed) {
      const double remaining = p.t - t1;
      const double k_star = (p.t * p.k - t1 * a_spent) / remaining;
      const double scale = remaining / p.t;
      if (k_star > 0.0) {
        AsianSpec pp = p;
        pp.k = k_star;
        pp.t = remaining;
        pp.has_avg = pp.has_elapsed = false;
        double v;
        if (int rc = asian_mc(pp, n_paths, seed, &v, err)) return rc;
        *price = scale * v;
        return STOCHA_OK;
      }
      if (!p.call) {
        *price = 0.0;
        return STOCHA_OK;
      }
      const double e_a_t =
          p.geometric ? expected_geometric_avg(p.s, p.r, p.q, p.sigma, t1, p.t, a_spent)
                      : expected_arithmetic_avg(p.s, p.r, p.q, t1, p.t, a_spent);
      *price = fmax0(exp(-p.r * remaining) * (e_a_t - p.k));
      return STOCHA_OK;
    }
    // floating strike, seasoned (asian.rs:293-355)
    const double remaining = p.t - t1;
    double nr = round((double)p.n_steps * remaining / p.t);
    uint64_t n_rem = nr >= 1.0 ? (uint64_t)nr : 1;
    AsianFn f = asian_fn(p, (double)n_rem);
    f.seasoned = 1;
    f.a_spent = a_spent;
    f.w_spent = t1 / p.t;
    f.w_remain = remaining / p.t;
    f.log_a_spent = log(a_spent);
    const GbmConst c = gbm_const(p.s, p.r, p.sigma, remaining, n_rem, p.q);
    double total = 0.0;
    if (int rc = run_sum(c, f, seed, n_paths, n_rem, &total, err)) return rc;
    *price = exp(-p.r * remaining) * total / (double)n_paths;
    return STOCHA_OK;
  }
  const double discount = exp(-p.r * p.t);
  const bool use_cv = !p.geometric && p.fixed;
  const GbmConst c = gbm_const(p.s, p.r, p.sigma, p.t, p.n_steps, p.q);
  const AsianFn f = asian_fn(p, (double)p.n_steps);
  if (!(use_cv && n_paths > 1)) {
    double total = 0.0;
    if (int rc = run_sum(c, f, seed, n_paths, p.n_steps, &total, err)) return rc;
    *price = fmax0(discount * (total / (double)n_paths));
    return STOCHA_OK;
  }
  double geo_analytical = 0.0;
  {
    AsianSpec g = p;
    g.geometric = true;
    if (!asian_geometric_fixed_analytical(g, &geo_analytical)) geo_analytical = 0.0;
  }
  if (p.n_steps > 0xffffffffull) return err.set(STOCHA_EINVAL, "n_steps must be below 2^32");
  hipStream_t st = globals().stream;
  StreamDesc sd;
  make_stream_desc(seed, block_size(kGbm, p.n_steps), &sd);
  const unsigned grid = fused_grid(n_paths);
  const unsigned rgrid = grid < 1024 ? grid : 1024;
  const size_t pair_bytes = sizeof(double2) * n_paths;
  const size_t part_off = (pair_bytes + 255) & ~(size_t)255;
  Workspace* ws = nullptr;
  if (int rc = workspace(part_off + sizeof(double) * 2 * rgrid, sizeof(double) * kMaxAcc, &ws,
                         err))
    return rc;
  double2* pairs = static_cast<double2*>(ws->dev);
  double* part = reinterpret_cast<double*>(static_cast<char*>(ws->dev) + part_off);
  KernelTimer timer(st);
  timer.start();
  asian_pairs_kernel<<<grid, kFusedBlock, 0, st>>>(c, f, sd, n_paths, (uint32_t)p.n_steps, pairs);
  STOCHA_HIP(hipGetLastError());
  timer.stop(1);
  // means (asian.rs:263-264), then the centred sums (:266-271)
  double m[2], cs[2];
  pair_sums_kernel<<<rgrid, 256, 0, st>>>(pairs, n_paths, 0.0, 0.0, 0, part);
  STOCHA_HIP(hipGetLastError());
  if (int rc = launch_reduce_partials(part, (int)rgrid, 2, ws->host, st, err)) return rc;
  STOCHA_HIP(hipStreamSynchronize(st));
  const double n = (double)n_paths;
  m[0] = ws->host[0] / n;
  m[1] = ws->host[1] / n;
  pair_sums_kernel<<<rgrid, 256, 0, st>>>(pairs, n_paths, m[0], m[1], 1, part);
  STOCHA_HIP(hipGetLastError());
  if (int rc = launch_reduce_partials(part, (int)rgrid, 2, ws->host, st, err)) return rc;
  STOCHA_HIP(hipStreamSynchronize(st));
  cs[0] = ws->host[0];
  cs[1] = ws->host[1];
  const double beta = cs[1] > 1e-30 ? cs[0] / cs[1] : 0.0;
  const double geo_mc_price = discount * m[1];
  const double adjusted = discount * m[0] - beta * (geo_mc_price - geo_analytical);
  *price = fmax0(adjusted);
  return STOCHA_OK;
}

// ---- lookback (lookback.rs) ----
struct LookbackSpec {
  double s, k, r, q, sigma, t;
  uint64_t n_steps;
  bool floating, call;
  bool has_max, has_min;
  double running_max, running_min;
};

// lookback.rs:40-42
double d1_fn(double s, double x, double r, double q, double sigma, double t) {
  return (log(s / x) + (r - q + 0.5 * sigma * sigma) * t) / (sigma * sqrt(t));
}

// lookback.rs:50-65
double t3(double s, double x, double r, double q, double sigma, double t) {
  const double sqrt_t = sqrt(t);
  const double d1 = d1_fn(s, x, r, q, sigma, t);
  const double b = r - q;
  const double exp_rt = exp(-r * t);
  if (fabs(b) < 1e-12) return s * exp_rt * sigma * sqrt_t * (norm_pdf(d1) - d1 * norm_cdf(-d1));
  const double y = 2.0 * b / (sigma * sigma);
  const double arg2 = -d1 + (2.0 * b / sigma) * sqrt_t;
  const double bracket = pow(s / x, -y) * norm_cdf(arg2) - exp(b * t) * norm_cdf(-d1);
  return s * exp_rt * (sigma * sigma / (2.0 * b)) * bracket;
}

// lookback.rs:68-100 (Goldman-Sosin-Gatto); false = None
bool lookback_floating_analytical(const LookbackSpec& p, double* out) {
  if (p.t <= 0.0 || p.sigma <= 0.0 || !p.floating) return false;
  const double exp_qt = exp(-p.q * p.t);
  const double exp_rt = exp(-p.r * p.t);
  double price;
  if (p.call) {
    const double m = p.has_min ? p.running_min : p.s;
    if (m > p.s) return false;
    const double d1 = d1_fn(p.s, m, p.r, p.q, p.sigma, p.t);
    const double d2 = d1 - p.sigma * sqrt(p.t);
    price = p.s * exp_qt * norm_cdf(d1) - m * exp_rt * norm_cdf(d2) + t3(p.s, m, p.r, p.q, p.sigma, p.t);
  } else {
    const double big_m = p.has_max ? p.running_max : p.s;
    if (big_m < p.s) return false;
    const double d1 = d1_fn(p.s, big_m, p.r, p.q, p.sigma, p.t);
    const double d2 = d1 - p.sigma * sqrt(p.t);
    price = big_m * exp_rt * norm_cdf(-d2) - p.s * exp_qt * norm_cdf(-d1) +
            t3(p.s, big_m, p.r, p.q, p.sigma, p.t);
  }
  *out = fmax0(price);
  return true;
}

// lookback.rs:104-155 (Conze-Viswanathan); false = None
bool lookback_fixed_analytical(const LookbackSpec& p, double* out) {
  if (p.t <= 0.0 || p.sigma <= 0.0 || p.floating || p.k <= 0.0) return false;
  const double exp_qt = exp(-p.q * p.t);
  const double exp_rt = exp(-p.r * p.t);
  double price;
  if (p.call) {
    const double big_m = p.has_max ? p.running_max : p.s;
    if (big_m < p.s) return false;
    if (p.k > big_m) {
      const double d1 = d1_fn(p.s, p.k, p.r, p.q, p.sigma, p.t);
      const double d2 = d1 - p.sigma * sqrt(p.t);
      price = p.s * exp_qt * norm_cdf(d1) - p.k * exp_rt * norm_cdf(d2) +
              t3(p.s, p.k, p.r, p.q, p.sigma, p.t);
    } else {
      const double d1 = d1_fn(p.s, big_m, p.r, p.q, p.sigma, p.t);
      const double d2 = d1 - p.sigma * sqrt(p.t);
      price = exp_rt * (big_m - p.k) + p.s * exp_qt * norm_cdf(d1) - big_m * exp_rt * norm_cdf(d2) +
              t3(p.s, big_m, p.r, p.q, p.sigma, p.t);
    }
  } else {
    const double m = p.has_min ? p.running_min : p.s;
    if (m > p.s) return false;
    if (p.k < m) {
      const double d1 = d1_fn(p.s, p.k, p.r, p.q, p.sigma, p.t);
      const double d2 = d1 - p.sigma * sqrt(p.t);
      price = p.k * exp_rt * norm_cdf(-d2) - p.s * exp_qt * norm_cdf(-d1) +
              t3(p.s, p.k, p.r, p.q, p.sigma, p.t);
    } else {
      const double d1 = d1_fn(p.s, m, p.r, p.q, p.sigma, p.t);
      const double d2 = d1 - p.sigma * sqrt(p.t);
      price = exp_rt * (p.k - m) + m * exp_rt * norm_cdf(-d2) - p.s * exp_qt * norm_cdf(-d1) +
              t3(p.s, m, p.r, p.q, p.sigma, p.t);
    }
  }
  *out = fmax0(price);
  return true;
}

// lookback.rs:158-210 on the device
int lookback_mc(const LookbackSpec& p, uint64_t n_paths, uint64_t seed, double* price,
                const Err& err) {
  const double discount = exp(-p.r * p.t);
  LookbackFn f{};
  f.k = p.k;
  f.floating = p.floating;
  f.call = p.call;
  const double rmax = p.has_max ? p.running_max : p.s;  // unwrap_or(s).max(s)
  const double rmin = p.has_min ? p.running_min : p.s;  // unwrap_or(s).min(s)
  f.max0 = rmax != rmax ? p.s : (rmax > p.s ? rmax : p.s);
  f.min0 = rmin != rmin ? p.s : (rmin < p.s ? rmin : p.s);
  const GbmConst c = gbm_const(p.s, p.r, p.sigma, p.t, p.n_steps, p.q);
  double total = 0.0;
  if (int rc = run_sum(c, f, seed, n_paths, p.n_steps, &total, err)) return rc;
  *price = discount * total / (double)n_paths;
  return STOCHA_OK;
}

}  // namespace
}  // namespace stocha

using namespace stocha;

namespace {

int parse_barrier(const char* bt, const char* ot, BarrierSpec* p, const Err& err) {
  if (!strcmp(bt, "up-and-out") || !strcmp(bt, "uo")) {
    p->up = true, p->out = true;
  } else if (!strcmp(bt, "up-and-in") || !strcmp(bt, "ui")) {
    p->up = true, p->out = false;
  } else if (!strcmp(bt, "down-and-out") || !strcmp(bt, "do")) {
    p->up = false, p->out = true;
  } else if (!strcmp(bt, "down-and-in") || !strcmp(bt, "di")) {
    p->up = false, p->out = false;
  } else {
    return err.set(STOCHA_EINVAL,
                   "Unknown barrier_type: '%s'. Supported: 'up-and-out', 'up-and-in', "
                   "'down-and-out', 'down-and-in'",
                   bt);
  }
  if (!strcmp(ot, "call")) {
    p->call = true;
  } else if (!strcmp(ot, "put")) {
    p->call = false;
  } else {
    return err.set(STOCHA_EINVAL, "Unknown option_type: '%s'. Supported: 'call', 'put'", ot);
  }
  return STOCHA_OK;
}

}  // namespace

extern "C" {

// src/lib.rs:1551-1638 -> barrier.rs:145-372
int stocha_barrier_price(double s, double k, double r, double sigma, double t, double barrier,
                         const char* barrier_type, const char* option_type, double q,
                         uint64_t n_paths, uint64_t n_steps, uint64_t seed, const char* method,
                         int has_n_monitoring, uint32_t n_monitoring, double rebate,
                         int rebate_at_hit, double* price, char* errbuf, size_t errlen) {
  Err err{errbuf, errlen};
  if (s <= 0.0) return err.set(STOCHA_EINVAL, "s must be positive");
  if (k <= 0.0) return err.set(STOCHA_EINVAL, "k must be positive");
  if (sigma <= 0.0) return err.set(STOCHA_EINVAL, "sigma must be positive");
  if (t <= 0.0) return err.set(STOCHA_EINVAL, "t must be positive");
  if (barrier <= 0.0) return err.set(STOCHA_EINVAL, "barrier must be positive");
  BarrierSpec p{};
  if (int rc = parse_barrier(barrier_type, option_type, &p, err)) return rc;
  p.s = s, p.k = k, p.r = r, p.q = q, p.sigma = sigma, p.t = t, p.barrier = barrier;
  p.rebate = rebate, p.rebate_at_hit = rebate_at_hit != 0;
  p.has_mon = has_n_monitoring != 0, p.n_mon = n_monitoring;
  const bool analytical = !strcmp(method, "analytical");
  if (!strcmp(method, "auto") || analytical) {
    double v;
    const bool some = barrier_analytical(p, &v);
    if (some && (analytical || isfinite(v))) {
      *price = v;
      return STOCHA_OK;
    }
    if (analytical)
      return err.set(STOCHA_EINVAL, "Analytical solution not available for these parameters");
  } else if (strcmp(method, "mc") != 0) {
    return err.set(STOCHA_EINVAL, "Unknown method: '%s'. Supported: 'auto', 'analytical', 'mc'",
                   method);
  }
  std::lock_guard<std::mutex> lk(globals().mu);
  if (int rc = check_device(err)) return rc;
  return barrier_mc(p, n_paths, n_steps, seed, price, err);
}

// src/lib.rs:1647-1745 -> asian.rs:78-355
int stocha_asian_price(double s, double k, double r, double sigma, double t, uint64_t n_steps,
                       const char* average_type, const char* strike_type,
                       const char* option_type, double q, uint64_t n_paths, uint64_t seed,
                       const char* method, int has_running_avg, double running_avg,
                       int has_time_elapsed, double time_elapsed, double* price, char* errbuf,
                       size_t errlen) {
  Err err{errbuf, errlen};
  if (s <= 0.0) return err.set(STOCHA_EINVAL, "s must be positive");
  if (k <= 0.0) return err.set(STOCHA_EINVAL, "k must be positive");
  if (sigma <= 0.0) return err.set(STOCHA_EINVAL, "sigma must be positive");
  if (t <= 0.0) return err.set(STOCHA_EINVAL, "t must be positive");
  AsianSpec p{};
  if (!strcmp(average_type, "arithmetic")) p.geometric = false;
  else if (!strcmp(average_type, "geometric")) p.geometric = true;
  else
    return err.set(STOCHA_EINVAL,
                   "Unknown average_type: '%s'. Supported: 'arithmetic', 'geometric'",
                   average_type);
  if (!strcmp(strike_type, "fixed")) p.fixed = true;
  else if (!strcmp(strike_type, "floating")) p.fixed = false;
  else
    return err.set(STOCHA_EINVAL, "Unknown strike_type: '%s'. Supported: 'fixed', 'floating'",
                   strike_type);
  if (!strcmp(option_type, "call")) p.call = true;
  else if (!strcmp(option_type, "put")) p.call = false;
  else
    return err.set(STOCHA_EINVAL, "Unknown option_type: '%s'. Supported: 'call', 'put'",
                   option_type);
  p.s = s, p.k = k, p.r = r, p.q = q, p.sigma = sigma, p.t = t, p.n_steps = n_steps;
  p.has_avg = has_running_avg != 0, p.running_avg = running_avg;
  p.has_elapsed = has_time_elapsed != 0, p.time_elapsed = time_elapsed;
  const bool has_analytical = p.geometric && p.fixed;
  if (!strcmp(method, "auto")) {
    if (has_analytical && asian_geometric_fixed_analytical(p, price)) return STOCHA_OK;
  } else if (!strcmp(method, "analytical")) {
    if (!has_analytical)
      return err.set(STOCHA_EINVAL,
                     "Analytical solution only available for geometric average with fixed strike");
    if (asian_geometric_fixed_analytical(p, price)) return STOCHA_OK;
    return err.set(STOCHA_EINVAL, "Analytical solution not available for these parameters");
  } else if (strcmp(method, "mc") != 0) {
    return err.set(STOCHA_EINVAL, "Unknown method: '%s'. Supported: 'auto', 'analytical', 'mc'",
                   method);
  }
  std::lock_guard<std::mutex> lk(globals().mu);
  if (int rc = check_device(err)) return rc;
  return asian_mc(p, n_paths, seed, price, err);
}

// src/lib.rs:1761-1848 -> lookback.rs:68-210
int stocha_lookback_price(double s, double r, double sigma, double t, uint64_t n_steps,
                          const char* strike_type, const char* option_type, double k, double q,
                          uint64_t n_paths, uint64_t seed, const char* method, int has_running_max,
                          double running_max, int has_running_min, double running_min,
                          double* price, char* errbuf, size_t errlen) {
  Err err{errbuf, errlen};
  if (s <= 0.0) return err.set(STOCHA_EINVAL, "s must be positive");
  if (sigma <= 0.0) return err.set(STOCHA_EINVAL, "sigma must be positive");
  if (t <= 0.0) return err.set(STOCHA_EINVAL, "t must be positive");
  LookbackSpec p{};
  if (!strcmp(strike_type, "floating")) {
    p.floating = true;
  } else if (!strcmp(strike_type, "fixed")) {
    if (k <= 0.0) return err.set(STOCHA_EINVAL, "k must be positive for fixed strike lookback");
    p.floating = false;
  } else {
    return err.set(STOCHA_EINVAL, "Unknown strike_type: '%s'. Supported: 'floating', 'fixed'",
                   strike_type);
  }
  if (!strcmp(option_type, "call")) p.call = true;
  else if (!strcmp(option_type, "put")) p.call = false;
  else
    return err.set(STOCHA_EINVAL, "Unknown option_type: '%s'. Supported: 'call', 'put'",
                   option_type);
  p.s = s, p.k = k, p.r = r, p.q = q, p.sigma = sigma, p.t = t, p.n_steps = n_steps;
  p.has_max = has_running_max != 0, p.running_max = running_max;
  p.has_min = has_running_min != 0, p.running_min = running_min;
  const bool analytical = !strcmp(method, "analytical");
  if (!strcmp(method, "auto") || analytical) {
    double v;
    const bool some = p.floating ? lookback_floating_analytical(p, &v)
                                 : lookback_fixed_analytical(p, &v);
    if (some && isfinite(v)) {
      *price = v;
      return STOCHA_OK;
    }
    if (analytical)
      return err.set(STOCHA_EINVAL, "Analytical solution not available for these parameters");
  } else if (strcmp(method, "mc") != 0) {
    return err.set(STOCHA_EINVAL, "Unknown method: '%s'. Supported: 'auto', 'analytical', 'mc'",
                   method);
  }
  std::lock_guard<std::mutex> lk(globals().mu);
  if (int rc = check_device(err)) return rc;
  return lookback_mc(p, n_paths, seed, price, err);
}

}  // extern "C"
