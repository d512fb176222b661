/*
 * stocha_gpu.h -- C ABI of the MI355X (gfx950) Monte Carlo hot path.
 *
 * The entry points replace, one for one, the PyO3 functions of the
 * reference's `stocha._stocha` module on the Monte Carlo path.  Each keeps
 * the reference's argument meaning, defaults (applied by the binding),
 * validation order and ValueError texts, and its output layout:
 *
 *   stocha_gbm                   <- src/lib.rs:205-252   (gbm)
 *   stocha_heston                <- src/lib.rs:327-408   (heston, both schemes)
 *   stocha_merton_jump_diffusion <- src/lib.rs:430-489   (merton_jump_diffusion)
 *   stocha_multi_gbm             <- src/lib.rs:759-827   (multi_gbm)
 *   stocha_lsmc_american_option  <- src/lib.rs:854-918   (lsmc_american_option)
 *   stocha_greeks_pathwise       <- src/lib.rs:1177-1242 (greeks_pathwise)
 *   stocha_greeks_fd             <- src/lib.rs:1064-1156 (greeks_fd, built-in payoffs)
 *   stocha_terminal_prices       <- src/finance/greeks.rs:91-167 (greeks_fd, callable payoffs)
 *   stocha_barrier_price         <- src/lib.rs:1551-1638 (barrier_price)
 *   stocha_asian_price           <- src/lib.rs:1647-1745 (asian_price)
 *   stocha_lookback_price        <- src/lib.rs:1761-1848 (lookback_price)
 *   stocha_rng_new / _uniform / _normal <- src/lib.rs:75-145 (class RNG)
 *   stocha_hull_white            <- src/lib.rs:606-658   (hull_white)
 *   stocha_gaussian_copula       <- src/lib.rs:538-558   (gaussian_copula)
 *   stocha_student_t_copula      <- src/lib.rs:563-578   (student_t_copula)
 *
 * Conventions (all entry points):
 *   - Plain C types only.  Integer arguments the reference takes as usize/u64
 *     are uint64_t; negative Python ints are rejected by the binding before
 *     the call (OverflowError, as PyO3 does).
 *   - Return value: STOCHA_OK, or an error code with a NUL-terminated message
 *     written to err[0..errlen) (the reference's ValueError text for
 *     STOCHA_EINVAL).  Nothing is thrown across the ABI.
 *   - Outputs are caller-allocated.  `where` says whether `out` is host
 *     memory (STOCHA_HOST; the call returns after the copy-back) or device
 *     memory on the current HIP device (STOCHA_DEVICE).  Any other `where`,
 *     an output whose byte count overflows 64 bits, and a STOCHA_DEVICE
 *     output that is not device memory of the current device with room for
 *     the result are refused with STOCHA_EINVAL before any launch.  For the
 *     path-returning functions out == NULL validates the arguments only
 *     (no device work), so a binding can raise the reference's error before
 *     it allocates the output.
 *   - Work is issued on the stream set with stocha_set_stream (default: the
 *     null stream) and every call synchronises that stream before it returns.
 *   - Calls are serialised by an internal mutex; results do not depend on
 *     launch configuration (block-split streams, fixed-order reductions).
 */
#ifndef STOCHA_GPU_H
#define STOCHA_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STOCHA_OK 0
#define STOCHA_EINVAL 1  /* invalid argument -> ValueError */
#define STOCHA_EDEVICE 2 /* HIP runtime error -> RuntimeError */
#define STOCHA_ENOMEM 3  /* device allocation failed -> MemoryError */
#define STOCHA_ENOTIMPL 4 /* outside this library's hot path -> NotImplementedError */

#define STOCHA_HOST 0
#define STOCHA_DEVICE 1

/* Precision of the stepping arithmetic for the forms that offer a choice. */
#define STOCHA_FP64 64
#define STOCHA_FP32 32

/* Schedules of the in-register forms (same draws, same values):
 *   FEED  every lane draws one u64 per iteration and advances its own
 *         ziggurat/step state machine (no wave-wide waits on slow lanes)
 *   LOOP  the reference's control flow: one ziggurat loop per normal */
#define STOCHA_SCHEDULE_FEED 0
#define STOCHA_SCHEDULE_LOOP 1

/* Model ids for the fused, batched and Greeks entry points. */
#define STOCHA_MODEL_GBM 0
#define STOCHA_MODEL_HESTON 1
#define STOCHA_MODEL_MERTON 2

/* ---- library ----------------------------------------------------------- */

/* Version string of this library ("stocha-gpu <semver>; reference 1.7.1"). */
const char *stocha_version(void);

/* Number of visible HIP devices (0 if none).  No reference counterpart. */
int stocha_device_count(void);

/* Issue all further work on `hip_stream` (a hipStream_t; NULL = null stream).
 * No reference counterpart (the reference has no device). */
int stocha_set_stream(void *hip_stream);

/* Kernel-time probe: when enabled, each entry point records HIP events
 * around its dominant kernel on the work stream; stocha_last_kernel_ms
 * returns that kernel's duration for the last call (ms), and the number of
 * launches it covered.  No reference counterpart (measurement only). */
int stocha_timing_enable(int on);
int stocha_last_kernel_ms(double *ms, uint64_t *launches);

/* ---- drop-in path functions --------------------------------------------- */

/* src/lib.rs:205-252 -> gbm.rs:42-107.  out: (n_paths, steps+1) f64,
 * row-major; row i < base is path i, rows >= base are the antithetic
 * mirrors of paths 0.. (base = (n_paths+1)/2 when antithetic). */
int stocha_gbm(double s0, double mu, double sigma, double t, uint64_t steps, uint64_t n_paths,
               uint64_t seed, int antithetic, double q, double *out, int where, char *err,
               size_t errlen);

/* src/lib.rs:327-408 -> heston.rs:32-174.  scheme: "euler" (full-truncation
 * Euler, heston.rs:44-87) or "qe" (Andersen's quadratic-exponential scheme,
 * heston.rs:89-174); NULL = "euler".  out: (n_paths, steps+1) f64 row-major. */
int stocha_heston(double s0, double v0, double mu, double kappa, double theta, double xi,
                  double rho, double t, uint64_t steps, uint64_t n_paths, uint64_t seed,
                  const char *scheme, double *out, int where, char *err, size_t errlen);

/* src/lib.rs:430-489 -> jump_diffusion.rs:48-105.  out: (n_paths, steps+1). */
int stocha_merton_jump_diffusion(double s0, double mu, double sigma, double lambda_, double mu_j,
                                 double sigma_j, double t, uint64_t steps, uint64_t n_paths,
                                 uint64_t seed, double *out, int where, char *err,
                                 size_t errlen);

/* src/lib.rs:759-827 -> multi_gbm.rs:22-113 (+ copula/gaussian.rs:76-99).
 * s0/mu/sigma have lengths n_s0/n_mu/n_sigma (checked like the reference);
 * corr is (corr_rows, corr_cols) f64 row-major on the host.
 * out: (n_paths, steps+1, d) f64 row-major. */
int stocha_multi_gbm(const double *s0, uint64_t n_s0, const double *mu, uint64_t n_mu,
                     const double *sigma, uint64_t n_sigma, const double *corr,
                     uint64_t corr_rows, uint64_t corr_cols, double t, uint64_t steps,
                     uint64_t n_paths, uint64_t seed, int antithetic, double *out, int where,
                     char *err, size_t errlen);

/* src/lib.rs:854-918 -> lsmc.rs:25-123.  Paths and the per-date solve in
 * FP64 on the device; returns (price, std_err). */
int stocha_lsmc_american_option(double s0, double k, double r, double sigma, double t,
                                uint64_t steps, uint64_t n_paths, int is_put,
                                uint64_t poly_degree, uint64_t seed, double *price,
                                double *std_err, char *err, size_t errlen);

/* src/lib.rs:1177-1242 -> greeks.rs:360-424.  greeks: n_greeks C strings
 * ("delta"/"vega"; other names raise the reference's messages, lib.rs:920-938
 * and :1213-1221).  values[i] receives the i-th requested greek. */
int stocha_greeks_pathwise(double s0, double r, double sigma, double t, double strike,
                           int is_call, uint64_t n_paths, uint64_t n_steps,
                           const char *const *greeks, uint64_t n_greeks, uint64_t seed,
                           double *values, char *err, size_t errlen);

/* src/lib.rs:1064-1117 -> greeks.rs:227-358 (greeks_fd with the built-in
 * payoffs "call" / "put"; other text raises the reference's "Unknown payoff"
 * message).  model: STOCHA_MODEL_*; params, in the order of the reference's
 * params dict (lib.rs:958-1038):
 *   GBM    {s0, r, sigma, t}                          (n_params 4)
 *   Heston {s0, v0, r, kappa, theta, xi, rho, t}      (n_params 8)
 *   Merton {s0, r, sigma, lambda_, mu_j, sigma_j, t}  (n_params 7)
 * scheme (Heston only; NULL = "euler", or "qe").  greeks:
 * n_greeks names among delta/gamma/vega/theta/rho (checked first, lib.rs:920);
 * values[i] receives the i-th.  Bumps max(|p| * bump_size, 1e-8); every
 * scenario on the same seed (common random numbers); scenarios that draw
 * alike share one launch and one stream of draws.  FP64 reference
 * arithmetic throughout. */
int stocha_greeks_fd(int model, const double *params, uint64_t n_params, const char *scheme,
                     const char *payoff, double strike, uint64_t n_paths, uint64_t n_steps,
                     const char *const *greeks, uint64_t n_greeks, uint64_t seed,
                     double bump_size, double *values, char *err, size_t errlen);

/* greeks.rs:91-167 (ModelSpec::terminal_prices): S_T of paths 0..n_paths-1
 * of `model` (params and scheme as stocha_greeks_fd; GBM with q = 0 and no
 * antithetic paths), FP64, into out (n_paths, host or device per `where`;
 * out == NULL validates only).  The callable-payoff form of greeks_fd
 * (lib.rs:1118-1150) evaluates the payoff on these. */
int stocha_terminal_prices(int model, const double *params, uint64_t n_params, const char *scheme,
                           uint64_t n_steps, uint64_t n_paths, uint64_t seed, double *out,
                           int where, char *err, size_t errlen);

/* ---- path-dependent options on GBM (barrier.rs, asian.rs, lookback.rs) ----
 * Same validation, dispatch and message texts as the #[pyfunction]s.  method
 * "auto" / "analytical" evaluate the reference's closed forms on the host
 * (Reiner-Rubinstein with the Broadie-Glasserman-Kou shift and rebates;
 * Kemna-Vorst; Goldman-Sosin-Gatto and Conze-Viswanathan); "mc" -- and
 * "auto" where no finite closed form exists -- simulates on the device on
 * gbm_paths' block-split streams (block (n_steps + 1024) * 2).  Optional
 * arguments of the reference (Option<...>) come as a has_ flag and a value.
 * price receives the result. */
/* src/lib.rs:1551-1638 -> barrier.rs:145-372 */
int stocha_barrier_price(double s, double k, double r, double sigma, double t, double barrier,
                         const char *barrier_type, const char *option_type, double q,
                         uint64_t n_paths, uint64_t n_steps, uint64_t seed, const char *method,
                         int has_n_monitoring, uint32_t n_monitoring, double rebate,
                         int rebate_at_hit, double *price, char *err, size_t errlen);
/* src/lib.rs:1647-1745 -> asian.rs:78-355 (arithmetic fixed strike: geometric
 * control variate, per-path pairs reduced in two fixed-order passes) */
int stocha_asian_price(double s, double k, double r, double sigma, double t, uint64_t n_steps,
                       const char *average_type, const char *strike_type,
                       const char *option_type, double q, uint64_t n_paths, uint64_t seed,
                       const char *method, int has_running_avg, double running_avg,
                       int has_time_elapsed, double time_elapsed, double *price, char *err,
                       size_t errlen);
/* src/lib.rs:1761-1848 -> lookback.rs:68-210 */
int stocha_lookback_price(double s, double r, double sigma, double t, uint64_t n_steps,
                          const char *strike_type, const char *option_type, double k, double q,
                          uint64_t n_paths, uint64_t seed, const char *method,
                          int has_running_max, double running_max, int has_running_min,
                          double running_min, double *price, char *err, size_t errlen);

/* ---- forms beside the drop-in (no reference counterpart; they extend the
 *      same path and use the same streams) ------------------------------- */

/* European calls priced in-register at n_strikes strikes (1..8) on the
 * terminal prices of `model`, without storing paths.  params:
 *   GBM    {s0, mu, sigma}                        (n_params 3)
 *   Heston {s0, v0, mu, kappa, theta, xi, rho}    (n_params 7)
 *   Merton {s0, mu, sigma, lambda, mu_j, sigma_j} (n_params 6)
 * precision: STOCHA_FP64 (reference arithmetic, s *= exp(.) per step) or
 * STOCHA_FP32 (log-space FP32 stepping, FP64 payoff sums).  Draws are the
 * reference's block-split streams in both cases.  schedule:
 * STOCHA_SCHEDULE_FEED or STOCHA_SCHEDULE_LOOP (identical results).  price[k] =
 * exp(-r t) * mean payoff, std_err[k] = exp(-r t) * sqrt(var_{n-1} / n);
 * sums (optional, 2*n_strikes) receives the raw FP64 sum and sum of squares of
 * the undiscounted payoffs per strike. */
int stocha_european_fused(int model, const double *params, uint64_t n_params, double t,
                          uint64_t steps, uint64_t n_paths, uint64_t seed, int precision,
                          int schedule, const double *strikes, uint64_t n_strikes, double r,
                          double *price, double *std_err, double *sums, char *err,
                          size_t errlen);

/* stocha_european_fused with nothing allocated.  No counterpart in the
 * reference: it is for callers that cannot allocate device or pinned host
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Arguments and results are those of
 * stocha_european_fused, except that price, std_err (n_strikes doubles each)
 * and sums (2*n_strikes, may be NULL) are DEVICE pointers of the current
 * device, checked before any launch.  The block partials, totals and the
 * last-block counter live in a static device array of 524,544 bytes per
 * device (room for 4096 blocks x 16 sums: every n_paths and n_strikes the
 * validation admits fits; anything larger is refused with STOCHA_EINVAL).
 * All calls on one device share that array, so calls that overlap on one
 * device race.  Returns after the results are written (stream synchronized).
 */
int stocha_european_fused_noalloc(int model, const double *params, uint64_t n_params, double t,
                                  uint64_t steps, uint64_t n_paths, uint64_t seed, int precision,
                                  int schedule, const double *strikes, uint64_t n_strikes,
                                  double r, double *price, double *std_err, double *sums,
                                  char *err, size_t errlen);

/* LSMC with a choice of path precision (FP32 paths: compensated FP64
 * moment sums, FP64 solve).  Same validation as stocha_lsmc_american_option. */
int stocha_lsmc_ex(double s0, double k, double r, double sigma, double t, uint64_t steps,
                   uint64_t n_paths, int is_put, uint64_t poly_degree, uint64_t seed,
                   int precision, double *price, double *std_err, char *err, size_t errlen);

/* Weighted basket call max(sum_a w_a S_T,a - K, 0) on correlated GBM with the
 * streams and draw order of multi_gbm (multi_gbm.rs:40-64; d <= 64), priced
 * in-register, and -- when greeks != 0 -- the pathwise adjoint of the price
 * with respect to each asset's s0, mu and sigma (d_s0/d_mu/d_sigma, length d,
 * host).  corr is d x d (Cholesky as copula/gaussian.rs:76-99, same error).
 * price = exp(-r t) mean payoff.  precision STOCHA_FP32 sums the normals and
 * forms L*W in FP32, STOCHA_FP64 in FP64.  No reference counterpart. */
int stocha_basket(const double *s0, const double *mu, const double *sigma, const double *corr,
                  const double *weights, uint64_t d, double strike, double r, double t,
                  uint64_t steps, uint64_t n_paths, uint64_t seed, int precision, int greeks,
                  double *price, double *std_err, double *d_s0, double *d_mu, double *d_sigma,
                  char *err, size_t errlen);

/* ---- batched pricing: many uneven instances in one persistent launch ------
 * Instance i prices a European call at `strike` on `n_paths` paths of
 * `steps` steps of its model, on the reference's block-split streams of the
 * plan's seed (the paths the single-instance call with that seed would
 * draw), FP32 log-space stepping, FP64 sums; price = exp(-r t) * mean.
 * Fields used per model: GBM s0, mu, sigma; Heston s0, v0, mu, kappa, theta,
 * xi, rho; Merton s0, mu, sigma, lambda_, mu_j, sigma_j.  No reference
 * counterpart (the reference prices one instance per call). */
typedef struct {
  int model; /* STOCHA_MODEL_* */
  uint32_t steps;
  uint64_t n_paths;
  double s0, strike, r, t;
  double mu, sigma;
  double v0, kappa, theta, xi, rho;
  double lambda_, mu_j, sigma_j;
} stocha_batch_instance;

typedef struct stocha_batch_plan stocha_batch_plan;

/* Build a plan: device tables, cost-ordered path tiles of block *
 * paths_per_lane paths, a persistent grid (waves_per_simd <= 0: as many
 * blocks as are resident), all captured in a hipGraph.  block: 64/128/256. */
int stocha_batch_plan_create(const stocha_batch_instance *inst, uint64_t n, uint64_t seed,
                             int block, int paths_per_lane, int waves_per_simd,
                             stocha_batch_plan **plan, char *err, size_t errlen);
/* Replay the plan's graph; price/std_err have n entries (host). */
int stocha_batch_plan_run(stocha_batch_plan *plan, double *price, double *std_err, char *err,
                          size_t errlen);
/* Tiles and persistent grid size of a plan (for tuning reports). */
int stocha_batch_plan_info(const stocha_batch_plan *plan, uint64_t *n_tiles, uint64_t *grid);
void stocha_batch_plan_destroy(stocha_batch_plan *plan);

/* src/lib.rs:606-658 -> hull_white.rs:26-52: short-rate paths, exact
 * transition, one normal per step on block-split streams (block
 * (steps + 1024) * 4).  out: (n_paths, steps+1) f64 row-major, column 0 = r0. */
int stocha_hull_white(double r0, double a, double theta, double sigma, double t, uint64_t steps,
                      uint64_t n_paths, uint64_t seed, double *out, int where, char *err,
                      size_t errlen);

/* src/lib.rs:538-558 -> copula/gaussian.rs:15-73: n_samples rows of
 * Phi(L z), z the next dim normals of ONE stream Pcg64Dxsm::new(seed) (the
 * RNG class's stream, decoded in parallel), L = cholesky(corr); corr is
 * (rows, cols) row-major on the host, checked like the reference (square,
 * unit diagonal, symmetric, positive definite).  out: (n_samples, dim) f64;
 * out == NULL validates only. */
int stocha_gaussian_copula(const double *corr, uint64_t rows, uint64_t cols, uint64_t n_samples,
                           uint64_t seed, double *out, int where, char *err, size_t errlen);

/* src/lib.rs:563-578 -> copula/student_t.rs:135-169: n_samples rows of
 * t_nu((L z) sqrt(nu / chi2)), per sample a chi-square(nu) (Marsaglia-Tsang
 * gamma(nu/2), doubled) and then dim normals, all from ONE stream
 * Pcg64Dxsm::new(seed) decoded in parallel; t_nu is the reference's
 * incomplete-beta CDF.  Checks: nu > 2 ("nu must be > 2", also for a NaN nu,
 * on which the reference's sampler would never return), square corr,
 * cholesky(corr).  out: (n_samples, dim) f64; out == NULL validates only. */
int stocha_student_t_copula(const double *corr, uint64_t rows, uint64_t cols, double nu,
                            uint64_t n_samples, uint64_t seed, double *out, int where, char *err,
                            size_t errlen);

/* ---- the RNG class (src/lib.rs:75-184): one sequential stream ------------
 * A generator is its 128-bit LCG state and increment, each as two uint64_t
 * (low word first), kept by the caller.  stocha_rng_new: Pcg64Dxsm::new for a
 * 128-bit seed (pcg64.rs:23-30; host only).  stocha_rng_uniform: n values of
 * next_f64 (RNG.uniform, lib.rs:108-112).  stocha_rng_normal: n ziggurat
 * normals in the stream's order (RNG.standard_normal, lib.rs:114-119), or
 * loc + scale z with scaled != 0 (RNG.normal, lib.rs:121-145; "scale must be
 * positive").  Both write new_state, the state after the draws consumed
 * (*consumed, optional, for normals), and fill out (host or device per
 * `where`) from parallel kernels whose values are the sequential stream's. */
int stocha_rng_new(uint64_t seed_lo, uint64_t seed_hi, uint64_t *state, uint64_t *inc);
int stocha_rng_uniform(const uint64_t *state, const uint64_t *inc, uint64_t n, double *out,
                       int where, uint64_t *new_state, char *err, size_t errlen);
int stocha_rng_normal(const uint64_t *state, const uint64_t *inc, uint64_t n, double loc,
                      double scale, int scaled, double *out, int where, uint64_t *new_state,
                      uint64_t *consumed, char *err, size_t errlen);

/* ---- random-number primitives (pcg64.rs:23-43, normal.rs:18-74) ---------
 * Draws of the block-split streams of paths path0 .. path0+n_paths-1, where
 * path i starts at Pcg64Dxsm::new(seed) advanced by i*block (the steppers'
 * layout, e.g. gbm.rs:55-59).  out is (n_paths, n_draws) row-major, host.
 * stocha_stream_normals: precision STOCHA_FP64 (the reference's values) or
 * STOCHA_FP32 (same draws consumed, FP32 values widened to double);
 * consumed (optional, n_paths) receives the u64 draws each path used. */
int stocha_stream_u64(uint64_t seed, uint64_t block, uint64_t path0, uint64_t n_paths,
                      uint64_t n_draws, uint64_t *out, char *err, size_t errlen);
int stocha_stream_normals(uint64_t seed, uint64_t block, uint64_t path0, uint64_t n_paths,
                          uint64_t n_draws, int precision, double *out, uint64_t *consumed,
                          char *err, size_t errlen);
/* The same normals drawn on the steppers' queue schedule (lane-mask ziggurat
 * state machine, per-lane rings, deferred slow block) instead of the
 * per-normal loop; no reference counterpart (test support: the schedule on
 * chosen stream positions).  n_draws < 2^32. */
int stocha_stream_normals_queue(uint64_t seed, uint64_t block, uint64_t path0, uint64_t n_paths,
                                uint64_t n_draws, int precision, double *out, char *err,
                                size_t errlen);

#ifdef __cplusplus
}
#endif
#endif
