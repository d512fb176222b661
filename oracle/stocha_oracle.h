/*
 * stocha_oracle.h -- CPU restatement of the reference's Monte Carlo hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else.  The product (stocha_amd/, libstocha_gpu.so) never links,
 * loads or calls it.
 *
 * It restates, in the reference's evaluation order (compiled with
 * -ffp-contract=off, glibc exp/log, IEEE sqrt):
 *   pcg64.rs:23-43 + rand_pcg 0.10.2 Lcg128CmDxsm64 (new/next_u64/advance),
 *   normal.rs:18-74 (ziggurat, tail), gbm.rs:42-107, heston.rs:44-87,
 *   jump_diffusion.rs:48-105, multi_gbm.rs:22-113, copula/gaussian.rs:76-99,
 *   lsmc.rs:25-123 (least squares by Householder QR in place of faer's QR),
 *   greeks.rs:360-424.
 * Paths run in parallel with OpenMP; every path owns its block-split stream,
 * so results do not depend on the thread count.
 *
 * Return codes: 0 ok, 1 invalid argument (message in err), 2 allocation.
 */
#ifndef STOCHA_ORACLE_H
#define STOCHA_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- streams (pcg64.rs, normal.rs) ---- */
void orc_stream_state(uint64_t seed, uint64_t block, uint64_t path, uint64_t out4[4]);
void orc_raw_u64(uint64_t seed, uint64_t block, uint64_t path, uint64_t n, uint64_t *out);
void orc_normals(uint64_t seed, uint64_t block, uint64_t path, uint64_t n, double *out,
                 uint64_t *consumed);
void orc_advance(uint64_t st[4], uint64_t delta_lo, uint64_t delta_hi);

/* ---- path-storing steppers: out is (n_paths, steps+1) row-major ---- */
int orc_gbm(double s0, double mu, double sigma, double t, uint64_t steps, uint64_t n_paths,
            uint64_t seed, int antithetic, double q, double *out);
int orc_heston_euler(double s0, double v0, double mu, double kappa, double theta, double xi,
                     double rho, double t, uint64_t steps, uint64_t n_paths, uint64_t seed,
                     double *out);
/* heston.rs:89-174 (QE); same layout.  _variance: the V(t_j) paths instead
 * of S (a test hook for the variance scheme's conditional moments). */
int orc_heston_qe_variance(double s0, double v0, double mu, double kappa, double theta,
                           double xi, double rho, double t, uint64_t steps, uint64_t n_paths,
                           uint64_t seed, double *out);
int orc_heston_qe(double s0, double v0, double mu, double kappa, double theta, double xi,
                  double rho, double t, uint64_t steps, uint64_t n_paths, uint64_t seed,
                  double *out);
int orc_merton(double s0, double mu, double sigma, double lambda, double mu_j, double sigma_j,
               double t, uint64_t steps, uint64_t n_paths, uint64_t seed, double *out);
/* out is (n_paths, steps+1, d) row-major */
int orc_multi_gbm(uint64_t d, const double *s0, const double *mu, const double *sigma,
                  const double *corr, double t, uint64_t steps, uint64_t n_paths, uint64_t seed,
                  int antithetic, double *out, char *err, size_t errlen);
int orc_cholesky(uint64_t n, const double *a, double *l, char *err, size_t errlen);
/* Terminal prices only, same arithmetic as orc_multi_gbm: out is (n_paths, d). */
int orc_multi_terminals(uint64_t d, const double *s0, const double *mu, const double *sigma,
                        const double *corr, double t, uint64_t steps, uint64_t n_paths,
                        uint64_t seed, double *out, char *err, size_t errlen);

/* ---- terminal-only steppers (same arithmetic as above, last column only) ---- */
/* model: 0 GBM (p = s0, mu, sigma, q), 1 Heston (s0, v0, mu, kappa, theta, xi, rho),
 *        2 Merton (s0, mu, sigma, lambda, mu_j, sigma_j), 3 Heston QE (as 1) */
int orc_terminals(int model, const double *p, double t, uint64_t steps, uint64_t n_paths,
                  uint64_t seed, double *out);
/* European calls on terminals: sum[k] = sum_i max(S_T,i - K_k, 0), sumsq likewise,
 * accumulated in path order. */
void orc_call_sums(const double *terminals, uint64_t n, const double *strikes, int n_strikes,
                   double *sum, double *sumsq);

/* ---- LSMC (lsmc.rs) ---- */
int orc_lsmc(double s0, double k, double r, double sigma, double t, uint64_t steps,
             uint64_t n_paths, int is_put, uint64_t poly_degree, uint64_t seed, double *price,
             double *std_err);
/* Per-date regression coefficients from the same backward induction, for
 * parity of the solve itself: coeffs is (steps, poly_degree+1); dates that the
 * reference skips are left NaN.  Also returns the ITM count per date. */
int orc_lsmc_trace(double s0, double k, double r, double sigma, double t, uint64_t steps,
                   uint64_t n_paths, int is_put, uint64_t poly_degree, uint64_t seed,
                   double *coeffs, uint64_t *itm_count, double *price, double *std_err);
/* Least squares min ||X b - y|| by Householder QR; X is (m, n) row-major. */
int orc_lstsq(uint64_t m, uint64_t n, const double *X, const double *y, double *b);

/* ---- pathwise Greeks (greeks.rs:360-424) ---- */
int orc_greeks_pathwise(double s0, double r, double sigma, double t, double strike, int is_call,
                        uint64_t n_paths, uint64_t steps, uint64_t seed, double *delta,
                        double *vega);

/* ---- bump-and-revalue Greeks (greeks.rs:227-358) ----
 * spec in the reference's params-dict order per model (GBM 4, Heston 8,
 * Merton 7 values; model 3 = Heston QE with Heston's spec); sign +1 call,
 * -1 put; greeks ids 0..4 = delta, gamma,
 * vega, theta, rho.  Prices with left-to-right payoff sums (mc_price). */
int orc_greeks_fd(int model, const double *spec, double sign, double strike, uint64_t n_paths,
                  uint64_t steps, const int *greeks, int n_greeks, uint64_t seed,
                  double bump_size, double *values);

/* ---- path-dependent options on GBM (Monte Carlo) ----
 * barrier.rs:286-372, lookback.rs:158-210, asian.rs:143-290 (unseasoned) and
 * :293-355 (seasoned floating strike).  Flags are 0/1. */
double orc_barrier_mc(double s0, double k, double r, double q, double sigma, double t,
                      double barrier, int up, int out, int call, double rebate, int rebate_at_hit,
                      uint64_t n_paths, uint64_t n_steps, uint64_t seed);
double orc_lookback_mc(double s0, double k, double r, double q, double sigma, double t,
                       int floating, int call, int has_max, double running_max, int has_min,
                       double running_min, uint64_t n_paths, uint64_t n_steps, uint64_t seed);
double orc_asian_mc(double s0, double k, double r, double q, double sigma, double t,
                    uint64_t n_steps, int geometric, int fixed, int call, uint64_t n_paths,
                    uint64_t seed);
double orc_asian_floating_seasoned(double s0, double r, double q, double sigma, double t,
                                   uint64_t n_steps, int geometric, int call, double a_spent,
                                   double t1, uint64_t n_paths, uint64_t seed);

/* ---- the RNG class (lib.rs:75-184): one sequential stream ----
 * st[4] = state lo, state hi, increment lo, increment hi; uniform / normal
 * advance st in place. */
void orc_rng_new(uint64_t seed_lo, uint64_t seed_hi, uint64_t st[4]);
void orc_rng_uniform(uint64_t st[4], uint64_t n, double *out);
void orc_rng_normal(uint64_t st[4], uint64_t n, double loc, double scale, int scaled,
                    double *out);

/* hull_white.rs:26-52; out (n_paths, steps+1). */
int orc_hull_white(double r0, double a, double theta, double sigma, double t, uint64_t steps,
                   uint64_t n_paths, uint64_t seed, double *out);
/* copula/gaussian.rs:15-73 given the Cholesky factor L (d x d). */
void orc_gaussian_copula(const double *L, uint64_t d, uint64_t n_samples, uint64_t seed,
                         double *out);
/* copula/student_t.rs: log-gamma (Lanczos), t CDF and the copula samples
 * (z_out, s_out, used optional: the normals, scales and draws consumed). */
double orc_lgamma_lanczos(double x);
double orc_student_t_cdf(double x, double nu);
void orc_student_t_copula(const double *L, uint64_t d, double nu, uint64_t n_samples,
                          uint64_t seed, double *out, double *z_out, double *s_out,
                          uint64_t *used);

/* Number of OpenMP threads the oracle will use. */
void orc_set_threads(int n);
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
