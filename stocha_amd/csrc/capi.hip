// capi.hip -- the extern "C" entry points declared in include/stocha_gpu.h.
//
// Each entry point validates its arguments in the order and with the
// messages of the reference's PyO3 function (src/lib.rs, cited per
// function), computes the per-launch constants with the reference's
// expressions (gbm.rs:43-45, heston.rs:45-48, jump_diffusion.rs:49-64,
// multi_gbm.rs:24-40, lsmc.rs:26-27), sets up the block-split stream table
// and launches the kernels on the work stream.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "host.hpp"
#include "multi.hpp"
#include "sums.hpp"

namespace stocha {

Globals& globals() {
  static Globals g;
  return g;
}

int workspace(size_t dev_bytes, size_t host_bytes, Workspace** out, const Err& err) {
  static std::vector<Workspace> per_dev;
  int dev = 0;
  STOCHA_HIP(hipGetDevice(&dev));
  if ((size_t)dev >= per_dev.size()) per_dev.resize(dev + 1);
  Workspace& w = per_dev[dev];
  if (!w.counter) {
    STOCHA_HIP(hipMalloc((void**)&w.counter, sizeof(unsigned) * kCounters));
    STOCHA_HIP(hipMemset(w.counter, 0, sizeof(unsigned) * kCounters));
  }
  if (!w.sums_part) {
    STOCHA_HIP(hipMalloc((void**)&w.sums_part, sizeof(double) * kSumsPartials));
    if (int rc = fill_empty_slots(w.sums_part, kSumsPartials, globals().stream, err)) return rc;
  }
  if (dev_bytes > w.dev_bytes) {
    if (w.dev) STOCHA_HIP(hipFree(w.dev));
    w.dev = nullptr;
    w.dev_bytes = 0;
    // round up to 1 MiB so small growth does not reallocate every call
    const size_t want = (dev_bytes + (1u << 20) - 1) & ~(size_t)((1u << 20) - 1);
    hipError_t e = hipMalloc(&w.dev, want);
    if (e != hipSuccess) {
      w.dev = nullptr;
      return err.set(STOCHA_ENOMEM, "device allocation of %zu bytes failed (%s)", want,
                     hipGetErrorName(e));
    }
    w.dev_bytes = want;
  }
  if (host_bytes > w.host_bytes) {
    if (w.host) STOCHA_HIP(hipHostFree(w.host));
    w.host = nullptr;
    w.host_bytes = 0;
    const size_t want = host_bytes < 4096 ? 4096 : host_bytes;
    STOCHA_HIP(hipHostMalloc((void**)&w.host, want, hipHostMallocDefault));
    w.host_bytes = want;
  }
  *out = &w;
  return STOCHA_OK;
}

GbmConst gbm_const(double s0, double mu, double sigma, double t, uint64_t steps, double q) {
  GbmConst c;
  const double dt = t / (double)steps;
  c.s0 = s0;
  c.drift = (mu - q - 0.5 * sigma * sigma) * dt;  // gbm.rs:44
  c.diffusion = sigma * sqrt(dt);                  // gbm.rs:45
  c.drift_f = (float)c.drift;
  c.diffusion_f = (float)c.diffusion;
  return c;
}

HestonConst heston_const(const double* p, double t, uint64_t steps) {
  HestonConst c;
  c.s0 = p[0];
  c.v0 = p[1];
  c.mu = p[2];
  c.kappa = p[3];
  c.theta = p[4];
  c.xi = p[5];
  c.rho = p[6];
  c.dt = t / (double)steps;              // heston.rs:45
  c.sqrt_dt = sqrt(c.dt);                // :46
  c.rho_comp = sqrt(1.0 - c.rho * c.rho);  // :47
  c.v0_f = (float)c.v0;
  c.rho_f = (float)c.rho;
  c.rho_comp_f = (float)c.rho_comp;
  c.kappa_dt_f = (float)(c.kappa * c.dt);
  c.theta_f = (float)c.theta;
  c.xi_sqdt_f = (float)(c.xi * c.sqrt_dt);
  c.mu_dt_f = (float)(c.mu * c.dt);
  c.half_dt_f = (float)(0.5 * c.dt);
  c.sqrt_dt_f = (float)c.sqrt_dt;
  return c;
}

// heston.rs:91-97 and the loop-invariant parts of :112-150, same operation
// order as the reference's per-step expressions.
HestonQeConst heston_qe_const(const double* p, double t, uint64_t steps) {
  HestonQeConst c;
  const double mu = p[2], kappa = p[3], theta = p[4], xi = p[5], rho = p[6];
  const double dt = t / (double)steps;     // :92
  const double e_kdt = exp(-kappa * dt);   // :93
  const double k1 = kappa * theta;         // :94
  c.s0 = p[0];
  c.v0 = p[1];
  c.kappa = kappa;
  c.xi = xi;
  c.e_kdt = e_kdt;
  c.one_m_e = 1.0 - e_kdt;
  c.m_const = k1 * (1.0 - e_kdt) / kappa;  // :112
  c.s2_const = k1 * xi * xi * ((1.0 - e_kdt) * (1.0 - e_kdt)) / (2.0 * kappa * kappa);  // :115-116
  const double kappa_rho_xi = kappa * rho / xi;  // :143
  const double gamma1 = 0.5, gamma2 = 0.5;
  c.k0s = (mu - kappa_rho_xi * theta) * dt;                  // :146
  c.k1s = gamma1 * dt * (kappa_rho_xi - 0.5) - rho / xi;     // :147
  c.k2s = gamma2 * dt * (kappa_rho_xi - 0.5) + rho / xi;     // :148
  c.k3s = gamma1 * dt * (1.0 - rho * rho);                   // :149
  c.k4s = gamma2 * dt * (1.0 - rho * rho);                   // :150
  return c;
}

MertonConst merton_const(const double* p, double t, uint64_t steps) {
  MertonConst c;
  const double s0 = p[0], mu = p[1], sigma = p[2], lambda = p[3], mu_j = p[4], sigma_j = p[5];
  const double dt = t / (double)steps;                            // jump_diffusion.rs:49
  const double sqrt_dt = sqrt(dt);                                // :50
  const double m_bar = exp(mu_j + 0.5 * sigma_j * sigma_j) - 1.0;  // :53
  c.s0 = s0;
  c.drift = (mu - 0.5 * sigma * sigma - lambda * m_bar) * dt;  // :56
  c.diffusion = sigma * sqrt_dt;                               // :57
  c.jump_prob = lambda * dt;                                   // :60
  const double scaled = ldexp(c.jump_prob, 53);
  c.jump_t53 = !(scaled > 0.0) ? 0 : scaled >= 0x1p53 ? (1ull << 53) : (uint64_t)ceil(scaled);
  c.mu_j = mu_j;
  c.sigma_j = sigma_j;
  c.drift_f = (float)c.drift;
  c.diffusion_f = (float)c.diffusion;
  c.mu_j_f = (float)mu_j;
  c.sigma_j_f = (float)sigma_j;
  return c;
}

// Rust's `{}` Display for f64: shortest round-trip digits, never an
// exponent (used in the Cholesky error text, copula/gaussian.rs:87-90).
std::string rust_f64(double v) {
  if (isnan(v)) return "NaN";
  if (isinf(v)) return v > 0 ? "inf" : "-inf";
  char buf[64];
  int prec = 1;
  for (; prec <= 17; ++prec) {
    snprintf(buf, sizeof buf, "%.*e", prec - 1, v);
    if (strtod(buf, nullptr) == v) break;
  }
  // buf = [-]d.ddddde[+-]XX -> digits and decimal exponent
  std::string s(buf);
  bool neg = s[0] == '-';
  if (neg) s = s.substr(1);
  size_t epos = s.find('e');
  int e10 = atoi(s.c_str() + epos + 1);
  std::string digits;
  for (size_t i = 0; i < epos; ++i)
    if (s[i] != '.') digits += s[i];
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  std::string out;
  const int nd = (int)digits.size();
  if (v == 0.0) {
    out = "0";
  } else if (e10 >= 0) {
    if (nd <= e10 + 1) {
      out = digits + std::string(e10 + 1 - nd, '0');
    } else {
      out = digits.substr(0, e10 + 1) + "." + digits.substr(e10 + 1);
    }
  } else {
    out = "0." + std::string(-e10 - 1, '0') + digits;
  }
  return (neg ? "-" : "") + out;
}

namespace {

// Runs a path-storing launch into host or device memory.
int store_paths(int model, const void* mc, uint64_t seed, uint64_t block, uint64_t n_rows,
                uint64_t base_rows, uint64_t steps, double* out, int where, const Err& err) {
  size_t bytes = 0;
  if (steps == UINT64_MAX) return err.set(STOCHA_EINVAL, "steps is too large");
  if (int rc = check_output(out, where, n_rows, steps + 1, sizeof(double), "out", &bytes, err))
    return rc;
  if (int rc = check_device(err)) return rc;
  hipStream_t st = globals().stream;
  StreamDesc sd;
  make_stream_desc(seed, block, &sd);
  if (where == STOCHA_DEVICE) {
    if (int rc = launch_paths(model, mc, sd, n_rows, base_rows, steps, out, st, err)) return rc;
    STOCHA_HIP(hipStreamSynchronize(st));
    return STOCHA_OK;
  }
  Workspace* ws = nullptr;
  if (int rc = workspace(bytes, 0, &ws, err)) return rc;
  if (int rc = launch_paths(model, mc, sd, n_rows, base_rows, steps,
                            static_cast<double*>(ws->dev), st, err))
    return rc;
  STOCHA_HIP(hipMemcpyAsync(out, ws->dev, bytes, hipMemcpyDeviceToHost, st));
  STOCHA_HIP(hipStreamSynchronize(st));
  return STOCHA_OK;
}

// ---- greeks_fd host logic (greeks.rs:24-358, lib.rs:958-1038) -------------
// A model spec holds the reference's fields in the order of the Python
// params dict: GBM {s0, r, sigma, t}, Heston {s0, v0, r, kappa, theta, xi,
// rho, t}, Merton {s0, r, sigma, lambda_, mu_j, sigma_j, t} (r is the drift,
// ModelSpec's mu).
enum FdParam : int { kBumpS0 = 0, kBumpSigma = 1, kBumpT = 2, kBumpMu = 3 };
enum FdGreek : int { kDelta = 0, kGamma, kVega, kTheta, kRho };

constexpr uint64_t kFdParams[3] = {4, 8, 7};

// Index of a bump parameter in the spec (ModelSpec::param_value /
// with_bumped, greeks.rs:169-224: Heston's "sigma" is v0).
int fd_index(int model, int param) {
  static const int idx[3][4] = {{0, 2, 3, 1}, {0, 1, 7, 2}, {0, 2, 6, 1}};
  return idx[model][param];
}

struct FdScenario {
  int param;  // -1 = Base
  int dir;    // +1 Up, -1 Down
};

// required_scenarios (greeks.rs:231-265): first-seen order, no duplicates.
std::vector<FdScenario> fd_required(const std::vector<int>& greeks) {
  std::vector<FdScenario> set;
  auto add = [&set](int param, int dir) {
    for (const FdScenario& k : set)
      if (k.param == param && (param < 0 || k.dir == dir)) return;
    set.push_back(FdScenario{param, dir});
  };
  for (int g : greeks) {
    switch (g) {
      case kDelta: add(kBumpS0, 1); add(kBumpS0, -1); break;
      case kGamma: add(kBumpS0, 1); add(-1, 0); add(kBumpS0, -1); break;
      case kVega: add(kBumpSigma, 1); add(kBumpSigma, -1); break;
      case kTheta: add(kBumpT, -1); add(-1, 0); break;
      case kRho: add(kBumpMu, 1); add(kBumpMu, -1); break;
    }
  }
  return set;
}

// compute_bump (greeks.rs:227-229): max(|v| * bump_size, 1e-8); f64::max
// returns 1e-8 for a NaN product, as the comparison below does.
double fd_bump(double v, double bump_size) {
  const double b = fabs(v) * bump_size;
  return b > 1e-8 ? b : 1e-8;
}

// The spec's model constants (ModelSpec::terminal_prices, greeks.rs:91-167;
// GBM with q = 0 and no antithetic paths).  kind: kGbm, kHeston (Euler),
// kMerton or kHestonQe.
struct FdConsts {
  GbmConst g;
  HestonConst h;
  MertonConst m;
  HestonQeConst q;
  const void* ptr(int kind) const {
    return kind == kGbm ? (const void*)&g : kind == kHeston ? (const void*)&h
                                          : kind == kMerton ? (const void*)&m : (const void*)&q;
  }
};

FdConsts fd_consts(int kind, const double* v, uint64_t steps) {
  FdConsts c{};
  if (kind == kGbm) c.g = gbm_const(v[0], v[1], v[2], v[3], steps, 0.0);
  else if (kind == kHeston) c.h = heston_const(v, v[7], steps);
  else if (kind == kHestonQe) c.q = heston_qe_const(v, v[7], steps);
  else c.m = merton_const(v, v[6], steps);
  return c;
}

// Whether two scenarios consume the same draws (one launch can serve both):
// GBM and Heston Euler always (normals only); Merton when lambda*dt's jump
// threshold is the same; Heston QE when the variance recursion is the same
// (its kind of first draw per step follows the variance path).
bool fd_same_draws(int kind, const FdConsts& a, const FdConsts& b) {
  if (kind == kMerton) return a.m.jump_t53 == b.m.jump_t53;
  if (kind == kHestonQe)
    return a.q.v0 == b.q.v0 && a.q.kappa == b.q.kappa && a.q.xi == b.q.xi &&
           a.q.e_kdt == b.q.e_kdt && a.q.m_const == b.q.m_const && a.q.s2_const == b.q.s2_const;
  return true;
}

// Heston's scheme (lib.rs:988-1000; only the Heston spec reads it): sets
// *kind to kHeston or kHestonQe.
int fd_kind(int model, const char* scheme, int* kind, const Err& err) {
  *kind = model;
  if (model != kHeston) return STOCHA_OK;
  const char* sch = scheme ? scheme : "euler";
  if (strcmp(sch, "euler") == 0) return STOCHA_OK;
  if (strcmp(sch, "qe") == 0) {
    *kind = kHestonQe;
    return STOCHA_OK;
  }
  return err.set(STOCHA_EINVAL, "scheme must be 'euler' or 'qe'");
}

}  // namespace
}  // namespace stocha

using namespace stocha;

extern "C" {

const char* stocha_version(void) { return "stocha-gpu 0.1.0 (gfx950); reference stocha 1.7.1"; }

int stocha_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int stocha_set_stream(void* s) {
  std::lock_guard<std::mutex> lk(globals().mu);
  globals().stream = (hipStream_t)s;
  return STOCHA_OK;
}

int stocha_timing_enable(int on) {
  std::lock_guard<std::mutex> lk(globals().mu);
  Globals& g = globals();
  if (on && !g.ev0) {
    if (hipEventCreate(&g.ev0) != hipSuccess || hipEventCreate(&g.ev1) != hipSuccess)
      return STOCHA_EDEVICE;
  }
  g.timing = on != 0;
  return STOCHA_OK;
}

int stocha_last_kernel_ms(double* ms, uint64_t* launches) {
  std::lock_guard<std::mutex> lk(globals().mu);
  Globals& g = globals();
  if (g.timer_pending) {
    g.timer_pending = false;
    float t = 0.f;
    if (hipEventSynchronize(g.ev1) == hipSuccess && hipEventElapsedTime(&t, g.ev0, g.ev1) == hipSuccess)
      g.last_ms = t;
  }
  if (ms) *ms = globals().last_ms;
  if (launches) *launches = globals().last_launches;
  return STOCHA_OK;
}

// src/lib.rs:205-252
int stocha_gbm(double s0, double mu, double sigma, double t, uint64_t steps, uint64_t n_paths,
               uint64_t seed, int antithetic, double q, double* out, int where, char* errbuf,
               size_t errlen) {
  Err err{errbuf, errlen};
  if (s0 <= 0.0) return err.set(STOCHA_EINVAL, "s0 must be positive");
  if (sigma <= 0.0) return err.set(STOCHA_EINVAL, "sigma must be positive");
  if (t <= 0.0) return err.set(STOCHA_EINVAL, "t must be positive");
  if (steps == 0 || n_paths == 0)
    return err.set(STOCHA_EINVAL, "steps and n_paths must be positive");
  if (!out) return STOCHA_OK;  // validate only
  std::lock_guard<std::mutex> lk(globals().mu);
  GbmConst c = gbm_const(s0, mu, sigma, t, steps, q);
  const uint64_t base = antithetic ? (n_paths + 1) / 2 : n_paths;  // gbm.rs:48-52
  return store_paths(kGbm, &c, seed, block_size(kGbm, steps), n_paths, base, steps, out, where,
                     err);
}

// src/lib.rs:327-408
int stocha_heston(double s0, double v0, double mu, double kappa, double theta, double xi,
                  double rho, double t, uint64_t steps, uint64_t n_paths, uint64_t seed,
                  const char* scheme, double* out, int where, char* errbuf, size_t errlen) {
  Err err{errbuf, errlen};
  if (s0 <= 0.0) return err.set(STOCHA_EINVAL, "s0 must be positive");
  if (v0 < 0.0) return err.set(STOCHA_EINVAL, "v0 must be non-negative");
  if (kappa <= 0.0) return err.set(STOCHA_EINVAL, "kappa must be positive");
  if (theta <= 0.0) return err.set(STOCHA_EINVAL, "theta must be positive");
  if (xi <= 0.0) return err.set(STOCHA_EINVAL, "xi must be positive");
  if (rho < -1.0 || rho > 1.0) return err.set(STOCHA_EINVAL, "rho must be in [-1, 1]");
  if (t <= 0.0) return err.set(STOCHA_EINVAL, "t must be positive");
  if (steps == 0 || n_paths == 0)
    return err.set(STOCHA_EINVAL, "steps and n_paths must be positive");
  const char* sch = scheme ? scheme : "euler";
  const bool qe = strcmp(sch, "qe") == 0;
  if (!qe && strcmp(sch, "euler") != 0) return err.set(STOCHA_EINVAL, "scheme must be 'euler' or 'qe'");
  if (!out) return STOCHA_OK;  // validate only
  std::lock_guard<std::mutex> lk(globals().mu);
  const double p[7] = {s0, v0, mu, kappa, theta, xi, rho};
  if (qe) {  // heston.rs:89-174
    const HestonQeConst c = heston_qe_const(p, t, steps);
    return store_paths(kHestonQe, &c, seed, block_size(kHestonQe, steps), n_paths, n_paths, steps,
                       out, where, err);
  }
  HestonConst c = heston_const(p, t, steps);
  return store_paths(kHeston, &c, seed, block_size(kHeston, steps), n_paths, n_paths, steps, out,
                     where, err);
}

// src/lib.rs:430-489
int stocha_merton_jump_diffusion(double s0, double mu, double sigma, double lambda_, double mu_j,
                                 double sigma_j, double t, uint64_t steps, uint64_t n_paths,
                                 uint64_t seed, double* out, int where, char* errbuf,
                                 size_t errlen) {
  Err err{errbuf, errlen};
  if (s0 <= 0.0) return err.set(STOCHA_EINVAL, "s0 must be positive");
  if (sigma <= 0.0) return err.set(STOCHA_EINVAL, "sigma must be positive");
  if (lambda_ < 0.0) return err.set(STOCHA_EINVAL, "lambda_ must be non-negative");
  if (sigma_j < 0.0) return err.set(STOCHA_EINVAL, "sigma_j must be non-negative");
  if (t <= 0.0) return err.set(STOCHA_EINVAL, "t must be positive");
  if (steps == 0 || n_paths == 0)
    return err.set(STOCHA_EINVAL, "steps and n_paths must be positive");
  if (!out) return STOCHA_OK;  // validate only
  std::lock_guard<std::mutex> lk(globals().mu);
  const double p[6] = {s0, mu, sigma, lambda_, mu_j, sigma_j};
  MertonConst c = merton_const(p, t, steps);
  return store_paths(kMerton, &c, seed, block_size(kMerton, steps), n_paths, n_paths, steps, out,
                     where, err);
}

// src/lib.rs:759-827
int stocha_multi_gbm(const double* s0, uint64_t n_s0, const double* mu, uint64_t n_mu,
                     const double* sigma, uint64_t n_sigma, const double* corr,
                     uint64_t corr_rows, uint64_t corr_cols, double t, uint64_t steps,
                     uint64_t n_paths, uint64_t seed, int antithetic, double* out, int where,
                     char* errbuf, size_t errlen) {
  Err err{errbuf, errlen};
  const uint64_t n = n_s0;
  if (n == 0) return err.set(STOCHA_EINVAL, "s0 must not be empty");
  if (n_mu != n || n_sigma != n)
    return err.set(STOCHA_EINVAL, "s0, mu, sigma must have the same length");
  for (uint64_t i = 0; i < n; ++i)
    if (s0[i] <= 0.0) return err.set(STOCHA_EINVAL, "all s0 must be positive");
  for (uint64_t i = 0; i < n; ++i)
    if (sigma[i] <= 0.0) return err.set(STOCHA_EINVAL, "all sigma must be positive");
  if (t <= 0.0) return err.set(STOCHA_EINVAL, "t must be positive");
  if (steps == 0 || n_paths == 0)
    return err.set(STOCHA_EINVAL, "steps and n_paths must be positive");
  if (corr_rows != n || corr_cols != n)
    return err.set(STOCHA_EINVAL, "corr must be (%llu, %llu), got (%llu, %llu)",
                   (unsigned long long)n, (unsigned long long)n,
                   (unsigned long long)corr_rows, (unsigned long long)corr_cols);
  if (!out) {  // validate only (including the Cholesky factorisation)
    std::vector<double> l(n * n);
    return cholesky_host(corr, n, l.data(), err);
  }
  std::lock_guard<std::mutex> lk(globals().mu);
  size_t row_vals = 0, bytes = 0;
  if (steps == UINT64_MAX || !byte_count(steps + 1, n, 1, &row_vals))
    return err.set(STOCHA_EINVAL, "out of %llu x %llu x %llu values does not fit in memory",
                   (unsigned long long)n_paths, (unsigned long long)steps,
                   (unsigned long long)n);
  if (int rc = check_output(out, where, n_paths, row_vals, sizeof(double), "out", &bytes, err))
    return rc;
  if (int rc = check_device(err)) return rc;
  return run_multi_gbm_paths(s0, mu, sigma, corr, n, t, steps, n_paths, seed, antithetic, out,
                             where, globals().stream, err);
}

// src/lib.rs:854-918
static int lsmc_validate(double s0, double k, double sigma, double t, uint64_t steps,
                         uint64_t n_paths, uint64_t poly_degree, const Err& err) {
  if (s0 <= 0.0) return err.set(STOCHA_EINVAL, "s0 must be positive");
  if (k <= 0.0) return err.set(STOCHA_EINVAL, "k must be positive");
  if (sigma <= 0.0) return err.set(STOCHA_EINVAL, "sigma must be positive");
  if (t <= 0.0) return err.set(STOCHA_EINVAL, "t must be positive");
  if (steps == 0 || n_paths == 0)
    return err.set(STOCHA_EINVAL, "steps and n_paths must be positive");
  if (n_paths < 2) return err.set(STOCHA_EINVAL, "n_paths must be at least 2");
  if (poly_degree == 0 || poly_degree > 4)
    return err.set(STOCHA_EINVAL, "poly_degree must be in [1, 4]");
  return STOCHA_OK;
}

int stocha_lsmc_ex(double s0, double k, double r, double sigma, double t, uint64_t steps,
                   uint64_t n_paths, int is_put, uint64_t poly_degree, uint64_t seed,
                   int precision, double* price, double* std_err, char* errbuf, size_t errlen) {
  Err err{errbuf, errlen};
  if (int rc = lsmc_validate(s0, k, sigma, t, steps, n_paths, poly_degree, err)) return rc;
  if (precision != STOCHA_FP64 && precision != STOCHA_FP32)
    return err.set(STOCHA_EINVAL, "precision must be 32 or 64");
  std::lock_guard<std::mutex> lk(globals().mu);
  if (int rc = check_device(err)) return rc;
  GbmConst c = gbm_const(s0, r, sigma, t, steps, 0.0);  // lsmc.rs:29-38
  StreamDesc sd;
  make_stream_desc(seed, block_size(kGbm, steps), &sd);
  const double dt = t / (double)steps;
  const double discount = exp(-r * dt);  // lsmc.rs:27
  return run_lsmc(c, sd, n_paths, steps, k, discount, is_put ? 1 : 0, (int)poly_degree + 1,
                  precision, price, std_err, globals().stream, err);
}

int stocha_lsmc_american_option(double s0, double k, double r, double sigma, double t,
                                uint64_t steps, uint64_t n_paths, int is_put,
                                uint64_t poly_degree, uint64_t seed, double* price,
                                double* std_err, char* errbuf, size_t errlen) {
  return stocha_lsmc_ex(s0, k, r, sigma, t, steps, n_paths, is_put, poly_degree, seed,
                        STOCHA_FP64, price, std_err, errbuf, errlen);
}

// src/lib.rs:1177-1242 (+ parse_greeks :920-938)
int stocha_greeks_pathwise(double s0, double r, double sigma, double t, double strike,
                           int is_call, uint64_t n_paths, uint64_t n_steps,
                           const char* const* greeks, uint64_t n_greeks, uint64_t seed,
                           double* values, char* errbuf, size_t errlen) {
  Err err{errbuf, errlen};
  if (s0 <= 0.0) return err.set(STOCHA_EINVAL, "s0 must be positive");
  if (sigma <= 0.0) return err.set(STOCHA_EINVAL, "sigma must be positive");
  if (t <= 0.0) return err.set(STOCHA_EINVAL, "t must be positive");
  if (n_paths == 0 || n_steps == 0)
    return err.set(STOCHA_EINVAL, "n_paths and n_steps must be positive");
  // parse_greeks: every name must be known before the pathwise check runs
  static const char* known[] = {"delta", "gamma", "vega", "theta", "rho"};
  for (uint64_t i = 0; i < n_greeks; ++i) {
    bool ok = false;
    for (const char* kn : known) ok = ok || strcmp(greeks[i], kn) == 0;
    if (!ok)
      return err.set(STOCHA_EINVAL,
                     "Unknown greek: '%s'. Supported: delta, gamma, vega, theta, rho", greeks[i]);
  }
  for (uint64_t i = 0; i < n_greeks; ++i) {
    if (strcmp(greeks[i], "delta") != 0 && strcmp(greeks[i], "vega") != 0)
      return err.set(STOCHA_EINVAL, "Pathwise method only supports 'delta' and 'vega', got '%s'",
                     greeks[i]);
  }
  std::lock_guard<std::mutex> lk(globals().mu);
  if (int rc = check_device(err)) return rc;
  GbmConst c = gbm_const(s0, r, sigma, t, n_steps, 0.0);  // greeks.rs:368-375 (mu = r)
  StreamDesc sd;
  make_stream_desc(seed, block_size(kGbm, n_steps), &sd);
  double dsum = 0.0, vsum = 0.0;
  if (int rc = run_greeks_pathwise(c, sd, n_paths, n_steps, s0, r, sigma, t, strike, is_call,
                                   &dsum, &vsum, globals().stream, err))
    return rc;
  const double discount = exp(-r * t);  // greeks.rs:377
  const double n = (double)n_paths;
  for (uint64_t i = 0; i < n_greeks; ++i)
    values[i] = strcmp(greeks[i], "delta") == 0 ? discount * dsum / n : discount * vsum / n;
  return STOCHA_OK;
}

// src/lib.rs:606-658 -> hull_white.rs:26-52
int stocha_hull_white(double r0, double a, double theta, double sigma, double t, uint64_t steps,
                      uint64_t n_paths, uint64_t seed, double* out, int where, char* errbuf,
                      size_t errlen) {
  Err err{errbuf, errlen};
  if (a <= 0.0) return err.set(STOCHA_EINVAL, "a must be positive");
  if (sigma <= 0.0) return err.set(STOCHA_EINVAL, "sigma must be positive");
  if (t <= 0.0) return err.set(STOCHA_EINVAL, "t must be positive");
  if (steps == 0 || n_paths == 0)
    return err.set(STOCHA_EINVAL, "steps and n_paths must be positive");
  if (!out) return STOCHA_OK;  // validate only
  std::lock_guard<std::mutex> lk(globals().mu);
  HullWhiteConst c;
  const double dt = t / (double)steps;                     // hull_white.rs:27
  const double e = exp(-a * dt);                           // :28
  c.s0 = r0;
  c.e = e;
  c.mean_reversion = (theta / a) * (1.0 - e);              // :29
  c.vol = sigma * sqrt((1.0 - e * e) / (2.0 * a));         // :30
  return store_paths(kHullWhite, &c, seed, block_size(kHullWhite, steps), n_paths, n_paths, steps,
                     out, where, err);
}

// src/lib.rs:1064-1117 -> greeks.rs:227-358 (greeks_fd, built-in payoff)
int stocha_greeks_fd(int model, const double* params, uint64_t n_params, const char* scheme,
                     const char* payoff, double strike, uint64_t n_paths, uint64_t n_steps,
                     const char* const* greeks, uint64_t n_greeks, uint64_t seed,
                     double bump_size, double* values, char* errbuf, size_t errlen) {
  Err err{errbuf, errlen};
  // parse_greeks (lib.rs:920-938) runs first
  static const char* known[] = {"delta", "gamma", "vega", "theta", "rho"};
  std::vector<int> gl;
  for (uint64_t i = 0; i < n_greeks; ++i) {
    int id = -1;
    for (int k = 0; k < 5; ++k)
      if (strcmp(greeks[i], known[k]) == 0) id = k;
    if (id < 0)
      return err.set(STOCHA_EINVAL,
                     "Unknown greek: '%s'. Supported: delta, gamma, vega, theta, rho", greeks[i]);
    gl.push_back(id);
  }
  if (model < 0 || model > 2) return err.set(STOCHA_EINVAL, "unknown model id %d", model);
  if (n_params != kFdParams[model])
    return err.set(STOCHA_EINVAL, "model %d takes %llu parameters, got %llu", model,
                   (unsigned long long)kFdParams[model], (unsigned long long)n_params);
  int kind;
  if (int rc = fd_kind(model, scheme, &kind, err)) return rc;
  double sign;
  if (payoff && strcmp(payoff, "call") == 0) {
    sign = 1.0;
  } else if (payoff && strcmp(payoff, "put") == 0) {
    sign = -1.0;
  } else {
    return err.set(STOCHA_EINVAL, "Unknown payoff: '%s'. Supported: 'call', 'put', or a callable",
                   payoff ? payoff : "");
  }
  if (gl.empty()) return STOCHA_OK;
  std::lock_guard<std::mutex> lk(globals().mu);
  if (int rc = check_device(err)) return rc;

  // scenarios, bumps (greeks_fd_core, greeks.rs:311-340)
  const std::vector<FdScenario> sc = fd_required(gl);
  double bump[4] = {0, 0, 0, 0};
  bool have[4] = {false, false, false, false};
  for (const FdScenario& k : sc) {
    if (k.param >= 0 && !have[k.param]) {
      bump[k.param] = fd_bump(params[fd_index(model, k.param)], bump_size);
      have[k.param] = true;
    }
  }
  const int ns = (int)sc.size();
  std::vector<std::vector<double>> spec(ns, std::vector<double>(params, params + n_params));
  for (int i = 0; i < ns; ++i)
    if (sc[i].param >= 0) spec[i][fd_index(model, sc[i].param)] += sc[i].dir * bump[sc[i].param];
  std::vector<FdConsts> cs(ns);
  for (int i = 0; i < ns; ++i) cs[i] = fd_consts(kind, spec[i].data(), n_steps);

  // one launch per group of scenarios that consume the same draws
  StreamDesc sd;
  make_stream_desc(seed, block_size(kind, n_steps), &sd);
  std::vector<double> sum(ns, 0.0);
  std::vector<bool> done(ns, false);
  for (int i = 0; i < ns; ++i) {
    if (done[i]) continue;
    std::vector<int> grp;
    for (int j = i; j < ns; ++j)
      if (!done[j] && fd_same_draws(kind, cs[i], cs[j])) grp.push_back(j);
    std::vector<GbmConst> g2;
    std::vector<HestonConst> h2;
    std::vector<MertonConst> m2;
    std::vector<HestonQeConst> q2;
    for (int j : grp) {
      g2.push_back(cs[j].g);
      h2.push_back(cs[j].h);
      m2.push_back(cs[j].m);
      q2.push_back(cs[j].q);
    }
    const void* cp = kind == kGbm      ? (const void*)g2.data()
                     : kind == kHeston ? (const void*)h2.data()
                     : kind == kMerton ? (const void*)m2.data()
                                       : (const void*)q2.data();
    double h[8];
    if (int rc = launch_fd(kind, cp, (int)grp.size(), strike, sign, sd, n_paths, n_steps, h,
                           globals().stream, err))
      return rc;
    for (size_t k = 0; k < grp.size(); ++k) {
      sum[grp[k]] = h[k];
      done[grp[k]] = true;
    }
  }
  // mc_price (greeks.rs:304-311): exp(-r t) * sum / n with the scenario's r, t
  const int it = fd_index(model, kBumpT), imu = fd_index(model, kBumpMu);
  std::vector<double> price(ns);
  for (int i = 0; i < ns; ++i) {
    const double discount = exp(-spec[i][imu] * spec[i][it]);
    price[i] = discount * sum[i] / (double)n_paths;
  }
  auto px = [&](int param, int dir) -> double {
    for (int i = 0; i < ns; ++i)
      if (sc[i].param == param && (param < 0 || sc[i].dir == dir)) return price[i];
    return NAN;
  };
  // compute_greeks_from_prices (greeks.rs:267-302)
  for (uint64_t i = 0; i < n_greeks; ++i) {
    double v = NAN;
    switch (gl[i]) {
      case kDelta: v = (px(kBumpS0, 1) - px(kBumpS0, -1)) / (2.0 * bump[kBumpS0]); break;
      case kGamma:
        v = (px(kBumpS0, 1) - 2.0 * px(-1, 0) + px(kBumpS0, -1)) / (bump[kBumpS0] * bump[kBumpS0]);
        break;
      case kVega: v = (px(kBumpSigma, 1) - px(kBumpSigma, -1)) / (2.0 * bump[kBumpSigma]); break;
      case kTheta: v = (px(kBumpT, -1) - px(-1, 0)) / bump[kBumpT]; break;
      case kRho: v = (px(kBumpMu, 1) - px(kBumpMu, -1)) / (2.0 * bump[kBumpMu]); break;
    }
    values[i] = v;
  }
  return STOCHA_OK;
}

// greeks.rs:91-167 (ModelSpec::terminal_prices), for greeks_fd's callable payoff
// (lib.rs:1118-1150).
int stocha_terminal_prices(int model, const double* params, uint64_t n_params, const char* scheme,
                           uint64_t n_steps, uint64_t n_paths, uint64_t seed, double* out,
                           int where, char* errbuf, size_t errlen) {
  Err err{errbuf, errlen};
  if (model < 0 || model > 2) return err.set(STOCHA_EINVAL, "unknown model id %d", model);
  if (n_params != kFdParams[model])
    return err.set(STOCHA_EINVAL, "model %d takes %llu parameters, got %llu", model,
                   (unsigned long long)kFdParams[model], (unsigned long long)n_params);
  int kind;
  if (int rc = fd_kind(model, scheme, &kind, err)) return rc;
  if (!out || n_paths == 0) return STOCHA_OK;
  std::lock_guard<std::mutex> lk(globals().mu);
  size_t bytes = 0;
  if (int rc = check_output(out, where, n_paths, 1, sizeof(double), "out", &bytes, err))
    return rc;
  if (int rc = check_device(err)) return rc;
  const FdConsts c = fd_consts(kind, params, n_steps);
  const void* cp = c.ptr(kind);
  StreamDesc sd;
  make_stream_desc(seed, block_size(kind, n_steps), &sd);
  hipStream_t st = globals().stream;
  if (where == STOCHA_DEVICE) {
    if (int rc = launch_terminals(kind, cp, sd, n_paths, n_steps, out, st, err)) return rc;
    STOCHA_HIP(hipStreamSynchronize(st));
    return STOCHA_OK;
  }
  Workspace* ws = nullptr;
  if (int rc = workspace(sizeof(double) * n_paths, 0, &ws, err)) return rc;
  if (int rc = launch_terminals(kind, cp, sd, n_paths, n_steps, static_cast<double*>(ws->dev),
                                st, err))
    return rc;
  STOCHA_HIP(hipMemcpyAsync(out, ws->dev, sizeof(double) * n_paths, hipMemcpyDeviceToHost, st));
  STOCHA_HIP(hipStreamSynchronize(st));
  return STOCHA_OK;
}

}  // extern "C"

namespace {

// Validation and per-launch constants shared by the two fused entry points.
struct FusedSetup {
  GbmConst gc;
  HestonConst hc;
  MertonConst mc;
  const void* cptr = nullptr;
  StreamDesc sd;
};

int fused_validate(int model, const double* params, uint64_t n_params, double t,
                   uint64_t steps, uint64_t n_paths, int precision, int schedule,
                   uint64_t n_strikes, const Err& err) {
  static const uint64_t np[3] = {3, 7, 6};
  if (model < 0 || model > 2) return err.set(STOCHA_EINVAL, "unknown model id %d", model);
  if (n_params != np[model])
    return err.set(STOCHA_EINVAL, "model %d takes %llu parameters, got %llu", model,
                   (unsigned long long)np[model], (unsigned long long)n_params);
  if (precision != STOCHA_FP64 && precision != STOCHA_FP32)
    return err.set(STOCHA_EINVAL, "precision must be 32 or 64");
  if (schedule != STOCHA_SCHEDULE_FEED && schedule != STOCHA_SCHEDULE_LOOP)
    return err.set(STOCHA_EINVAL, "schedule must be 0 (feed) or 1 (loop)");
  if (n_strikes == 0 || n_strikes > 8)
    return err.set(STOCHA_EINVAL, "n_strikes must be in [1, 8]");
  if (params[0] <= 0.0) return err.set(STOCHA_EINVAL, "s0 must be positive");
  if (t <= 0.0) return err.set(STOCHA_EINVAL, "t must be positive");
  if (steps == 0 || n_paths == 0)
    return err.set(STOCHA_EINVAL, "steps and n_paths must be positive");
  return STOCHA_OK;
}

void fused_setup(int model, const double* params, double t, uint64_t steps, uint64_t seed,
                 FusedSetup* f) {
  if (model == kGbm) {
    f->gc = gbm_const(params[0], params[1], params[2], t, steps, 0.0);
    f->cptr = &f->gc;
  } else if (model == kHeston) {
    f->hc = heston_const(params, t, steps);
    f->cptr = &f->hc;
  } else {
    f->mc = merton_const(params, t, steps);
    f->cptr = &f->mc;
  }
  make_stream_desc(seed, block_size(model, steps), &f->sd);
}

}  // namespace

extern "C" {

int stocha_european_fused(int model, const double* params, uint64_t n_params, double t,
                          uint64_t steps, uint64_t n_paths, uint64_t seed, int precision,
                          int schedule, const double* strikes, uint64_t n_strikes, double r,
                          double* price, double* std_err, double* sums, char* errbuf,
                          size_t errlen) {
  Err err{errbuf, errlen};
  if (int rc = fused_validate(model, params, n_params, t, steps, n_paths, precision, schedule,
                              n_strikes, err))
    return rc;
  std::lock_guard<std::mutex> lk(globals().mu);
  if (int rc = check_device(err)) return rc;
  FusedSetup f;
  fused_setup(model, params, t, steps, seed, &f);
  double h[16];
  if (int rc = launch_fused(model, precision, schedule, f.cptr, f.sd, n_paths, steps, strikes,
                            (int)n_strikes, h, globals().stream, err))
    return rc;
  const double disc = exp(-r * t);
  const double n = (double)n_paths;
  for (uint64_t k = 0; k < n_strikes; ++k) {
    const double s1 = h[2 * k], s2 = h[2 * k + 1];
    price[k] = disc * s1 / n;
    double var = n > 1 ? (s2 - s1 * s1 / n) / (n - 1.0) : 0.0;
    var = var > 0.0 ? var : 0.0;
    std_err[k] = disc * sqrt(var / n);
    if (sums) {
      sums[2 * k] = s1;
      sums[2 * k + 1] = s2;
    }
  }
  return STOCHA_OK;
}

// stocha_european_fused with nothing allocated: device outputs, partials in
// a static device array of fused.hip.  No reference counterpart.
int stocha_european_fused_noalloc(int model, const double* params, uint64_t n_params, double t,
                                  uint64_t steps, uint64_t n_paths, uint64_t seed, int precision,
                                  int schedule, const double* strikes, uint64_t n_strikes,
                                  double r, double* d_price, double* d_std_err, double* d_sums,
                                  char* errbuf, size_t errlen) {
  Err err{errbuf, errlen};
  if (int rc = fused_validate(model, params, n_params, t, steps, n_paths, precision, schedule,
                              n_strikes, err))
    return rc;
  std::lock_guard<std::mutex> lk(globals().mu);
  if (int rc = check_device(err)) return rc;
  const size_t nb = sizeof(double) * n_strikes;
  if (int rc = check_device_span(d_price, nb, "price", err)) return rc;
  if (int rc = check_device_span(d_std_err, nb, "std_err", err)) return rc;
  if (d_sums)
    if (int rc = check_device_span(d_sums, 2 * nb, "sums", err)) return rc;
  FusedSetup f;
  fused_setup(model, params, t, steps, seed, &f);
  return launch_fused_noalloc(model, precision, schedule, f.cptr, f.sd, n_paths, steps, strikes,
                              (int)n_strikes, exp(-r * t), d_price, d_std_err, d_sums,
                              globals().stream, err);
}

// Basket call on correlated GBM (multi_gbm.rs streams) with pathwise adjoint
// Greeks.  No reference counterpart.
int stocha_basket(const double* s0, const double* mu, const double* sigma, const double* corr,
                  const double* weights, uint64_t d, double strike, double r, double t,
                  uint64_t steps, uint64_t n_paths, uint64_t seed, int precision, int greeks,
                  double* price, double* std_err, double* d_s0, double* d_mu, double* d_sigma,
                  char* errbuf, size_t errlen) {
  Err err{errbuf, errlen};
  if (d == 0) return err.set(STOCHA_EINVAL, "s0 must not be empty");
  if (d > 64) return err.set(STOCHA_EINVAL, "the basket form supports up to 64 assets, got %llu",
                             (unsigned long long)d);
  for (uint64_t i = 0; i < d; ++i)
    if (s0[i] <= 0.0) return err.set(STOCHA_EINVAL, "all s0 must be positive");
  for (uint64_t i = 0; i < d; ++i)
    if (sigma[i] <= 0.0) return err.set(STOCHA_EINVAL, "all sigma must be positive");
  if (t <= 0.0) return err.set(STOCHA_EINVAL, "t must be positive");
  if (steps == 0 || n_paths == 0)
    return err.set(STOCHA_EINVAL, "steps and n_paths must be positive");
  if (steps * d > 0xffffffffull) return err.set(STOCHA_EINVAL, "steps * d must be below 2^32");
  if (precision != STOCHA_FP64 && precision != STOCHA_FP32)
    return err.set(STOCHA_EINVAL, "precision must be 32 or 64");
  std::vector<double> L(d * d);
  if (int rc = cholesky_host(corr, d, L.data(), err)) return rc;
  std::lock_guard<std::mutex> lk(globals().mu);
  if (int rc = check_device(err)) return rc;
  std::vector<double> h(2 + 3 * d);
  // FP32 forms L*W on the matrix cores unless STOCHA_BASKET_VALU is set
  const bool mfma = precision == STOCHA_FP32 && !getenv("STOCHA_BASKET_VALU");
  if (int rc = run_basket(s0, mu, sigma, L.data(), weights, d, strike, t, steps, n_paths, seed,
                          precision, greeks != 0, mfma, h.data(), globals().stream, err))
    return rc;
  const double disc = exp(-r * t);
  const double n = (double)n_paths;
  const double s1 = h[0], s2 = h[1];
  *price = disc * s1 / n;
  double var = n > 1 ? (s2 - s1 * s1 / n) / (n - 1.0) : 0.0;
  var = var > 0.0 ? var : 0.0;
  if (std_err) *std_err = disc * sqrt(var / n);
  if (greeks) {
    for (uint64_t a = 0; a < d; ++a) {
      if (d_s0) d_s0[a] = disc * h[2 + 3 * a] / n;
      if (d_mu) d_mu[a] = disc * h[3 + 3 * a] / n;
      if (d_sigma) d_sigma[a] = disc * h[4 + 3 * a] / n;
    }
  }
  return STOCHA_OK;
}

struct stocha_batch_plan {
  stocha::BatchPlan* p;
};

int stocha_batch_plan_create(const stocha_batch_instance* inst, uint64_t n, uint64_t seed,
                             int block, int paths_per_lane, int waves_per_simd,
                             stocha_batch_plan** plan, char* errbuf, size_t errlen) {
  Err err{errbuf, errlen};
  if (n == 0 || n > 0xffffffffull) return err.set(STOCHA_EINVAL, "n must be in [1, 2^32)");
  for (uint64_t i = 0; i < n; ++i) {
    const stocha_batch_instance& b = inst[i];
    if (b.model < 0 || b.model > 2)
      return err.set(STOCHA_EINVAL, "instance %llu: unknown model %d", (unsigned long long)i,
                     b.model);
    if (b.steps == 0 || b.n_paths == 0)
      return err.set(STOCHA_EINVAL, "instance %llu: steps and n_paths must be positive",
                     (unsigned long long)i);
    if (b.s0 <= 0.0 || b.t <= 0.0)
      return err.set(STOCHA_EINVAL, "instance %llu: s0 and t must be positive",
                     (unsigned long long)i);
    if (b.model != STOCHA_MODEL_HESTON && b.sigma <= 0.0)
      return err.set(STOCHA_EINVAL, "instance %llu: sigma must be positive",
                     (unsigned long long)i);
  }
  std::lock_guard<std::mutex> lk(globals().mu);
  if (int rc = check_device(err)) return rc;
  stocha::BatchPlan* p = nullptr;
  if (int rc = batch_plan_create(inst, n, seed, block, paths_per_lane, waves_per_simd, &p, err))
    return rc;
  *plan = new stocha_batch_plan{p};
  return STOCHA_OK;
}

int stocha_batch_plan_run(stocha_batch_plan* plan, double* price, double* std_err, char* errbuf,
                          size_t errlen) {
  Err err{errbuf, errlen};
  if (!plan) return err.set(STOCHA_EINVAL, "null plan");
  std::lock_guard<std::mutex> lk(globals().mu);
  return batch_plan_run(plan->p, price, std_err, err);
}

int stocha_batch_plan_info(const stocha_batch_plan* plan, uint64_t* n_tiles, uint64_t* grid) {
  if (!plan) return STOCHA_EINVAL;
  if (n_tiles) *n_tiles = batch_plan_tiles(plan->p);
  if (grid) *grid = batch_plan_grid(plan->p);
  return STOCHA_OK;
}

void stocha_batch_plan_destroy(stocha_batch_plan* plan) {
  if (!plan) return;
  std::lock_guard<std::mutex> lk(globals().mu);
  batch_plan_destroy(plan->p);
  delete plan;
}

}  // extern "C"
