/*
 * stocha_oracle.c -- CPU restatement of the reference's Monte Carlo hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; see stocha_oracle.h).  Compile with
 * -O2 -ffp-contract=off so that every a*b+c is two roundings, as in Rust.
 * Each function cites the reference file:line it follows.
 */
#include "stocha_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "zig_tables.h"

typedef unsigned __int128 u128;

static const double ZX[257] = {ZIG_X_VALUES};
static const double ZY[257] = {ZIG_Y_VALUES};
static const double ZR = ZIG_R_VALUE;

/* ------------------------------------------------------------------------ */
/* PCG64DXSM: rand_pcg 0.10.2 Lcg128CmDxsm64 as wrapped by pcg64.rs:23-43.    */
/* ------------------------------------------------------------------------ */
#define PCG_CHEAP_MUL 0xda942042e4dd58b5ull

typedef struct {
    u128 state;
    u128 inc;
} pcg_t;

/* pcg64.rs:23-30: stream = seed + 0xda3e...4f5a; Lcg128CmDxsm64::new(state=seed, stream):
 * inc = (stream << 1) | 1, state = seed + inc, then one LCG step. */
static void pcg_new(pcg_t *g, uint64_t seed) {
    u128 s = (u128)seed;
    u128 stream = s + (((u128)0xda3e39cb94b95bdbull << 64) | 0x7ef55da5c0374f5aull);
    g->inc = (stream << 1) | 1;
    g->state = s + g->inc;
    g->state = g->state * (u128)PCG_CHEAP_MUL + g->inc;
}

/* pcg64.rs:33-35 -> DXSM output of the pre-step state, then the LCG step. */
static inline uint64_t pcg_next_u64(pcg_t *g) {
    uint64_t hi = (uint64_t)(g->state >> 64);
    uint64_t lo = (uint64_t)g->state | 1u;
    hi ^= hi >> 32;
    hi *= PCG_CHEAP_MUL;
    hi ^= hi >> 48;
    hi *= lo;
    g->state = g->state * (u128)PCG_CHEAP_MUL + g->inc;
    return hi;
}

/* pcg64.rs:37-39 */
static inline double pcg_next_f64(pcg_t *g) {
    return (double)(pcg_next_u64(g) >> 11) * (1.0 / 9007199254740992.0);
}

/* pcg64.rs:41-43 -> rand_pcg advance: jump delta steps by square-and-multiply
 * of the affine map s -> a*s + c (Brown, "Random number generation with
 * arbitrary strides", 1994). */
static void pcg_advance(pcg_t *g, u128 delta) {
    u128 acc_mult = 1, acc_plus = 0;
    u128 cur_mult = (u128)PCG_CHEAP_MUL, cur_plus = g->inc;
    while (delta > 0) {
        if (delta & 1) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1) * cur_plus;
        cur_mult *= cur_mult;
        delta >>= 1;
    }
    g->state = acc_mult * g->state + acc_plus;
}

static void pcg_for_path(pcg_t *g, uint64_t seed, uint64_t block, uint64_t path) {
    pcg_new(g, seed);
    pcg_advance(g, (u128)path * (u128)block);
}

/* ------------------------------------------------------------------------ */
/* Ziggurat: normal.rs:18-67                                                 */
/* ------------------------------------------------------------------------ */
static double zig_tail(pcg_t *g) {
    for (;;) {
        double u1 = pcg_next_f64(g);
        double u2 = pcg_next_f64(g);
        if (u1 <= 0.0 || u2 <= 0.0) continue;
        double x = -log(u1) / ZR;
        if (-2.0 * log(u2) >= x * x) return x + ZR;
    }
}

static inline double zig_normal(pcg_t *g) {
    for (;;) {
        uint64_t u = pcg_next_u64(g);
        int layer = (int)(u & 0xFF);
        double sign = (double)((u >> 8) & 1) * -2.0 + 1.0;
        double u_frac = (double)(u >> 11) * (1.0 / 9007199254740992.0);
        double x = u_frac * ZX[layer];
        if (x < ZX[layer + 1]) return sign * x;
        if (layer == 0) return sign * zig_tail(g);
        double y = ZY[layer] + pcg_next_f64(g) * (ZY[layer + 1] - ZY[layer]);
        if (y < exp(-0.5 * x * x)) return sign * x;
    }
}

/* ------------------------------------------------------------------------ */
/* Stream probes                                                             */
/* ------------------------------------------------------------------------ */
void orc_stream_state(uint64_t seed, uint64_t block, uint64_t path, uint64_t out4[4]) {
    pcg_t g;
    pcg_for_path(&g, seed, block, path);
    out4[0] = (uint64_t)g.state;
    out4[1] = (uint64_t)(g.state >> 64);
    out4[2] = (uint64_t)g.inc;
    out4[3] = (uint64_t)(g.inc >> 64);
}

void orc_advance(uint64_t st[4], uint64_t delta_lo, uint64_t delta_hi) {
    pcg_t g;
    g.state = ((u128)st[1] << 64) | st[0];
    g.inc = ((u128)st[3] << 64) | st[2];
    pcg_advance(&g, ((u128)delta_hi << 64) | delta_lo);
    st[0] = (uint64_t)g.state;
    st[1] = (uint64_t)(g.state >> 64);
}

void orc_raw_u64(uint64_t seed, uint64_t block, uint64_t path, uint64_t n, uint64_t *out) {
    pcg_t g;
    pcg_for_path(&g, seed, block, path);
    for (uint64_t i = 0; i < n; i++) out[i] = pcg_next_u64(&g);
}

void orc_normals(uint64_t seed, uint64_t block, uint64_t path, uint64_t n, double *out,
                 uint64_t *consumed) {
    pcg_t g, g0;
    pcg_for_path(&g, seed, block, path);
    g0 = g;
    for (uint64_t i = 0; i < n; i++) out[i] = zig_normal(&g);
    if (consumed) {
        /* count draws by stepping a copy until the states meet (n is small) */
        uint64_t c = 0;
        while (g0.state != g.state) {
            g0.state = g0.state * (u128)PCG_CHEAP_MUL + g0.inc;
            c++;
        }
        *consumed = c;
    }
}

/* OpenMP threads of the following calls (the CPU baseline's single-thread
 * and all-core runs); n <= 0 keeps the current setting. */
void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* GBM: gbm.rs:42-107                                                        */
/* ------------------------------------------------------------------------ */
typedef struct {
    double drift, diffusion, s0;
    uint64_t steps, block;
} gbm_k;

static void gbm_k_init(gbm_k *k, double s0, double mu, double sigma, double t, uint64_t steps,
                       double q) {
    double dt = t / (double)steps;
    k->drift = (mu - q - 0.5 * sigma * sigma) * dt; /* gbm.rs:44 */
    k->diffusion = sigma * sqrt(dt);                 /* gbm.rs:45 */
    k->s0 = s0;
    k->steps = steps;
    k->block = (steps + 1024) * 2; /* gbm.rs:55 */
}

/* gbm.rs:57-75: one path; row (steps+1) or NULL (terminal only). */
static double gbm_one(const gbm_k *k, uint64_t seed, uint64_t path, int negate, double *row) {
    pcg_t g;
    pcg_for_path(&g, seed, k->block, path);
    double s = k->s0;
    if (row) row[0] = s;
    for (uint64_t j = 0; j < k->steps; j++) {
        double z = negate ? -zig_normal(&g) : zig_normal(&g);
        s *= exp(k->drift + k->diffusion * z);
        if (row) row[j + 1] = s;
    }
    return s;
}

int orc_gbm(double s0, double mu, double sigma, double t, uint64_t steps, uint64_t n_paths,
            uint64_t seed, int antithetic, double q, double *out) {
    gbm_k k;
    gbm_k_init(&k, s0, mu, sigma, t, steps, q);
    uint64_t base = antithetic ? (n_paths + 1) / 2 : n_paths;
    uint64_t w = steps + 1;
    /* gbm.rs:83-97: rows [base paths..., antithetic mirrors...] truncated to n. */
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t r = 0; r < (int64_t)n_paths; r++) {
        uint64_t row = (uint64_t)r;
        if (row < base)
            gbm_one(&k, seed, row, 0, out + row * w);
        else
            gbm_one(&k, seed, row - base, 1, out + row * w);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Heston full-truncation Euler: heston.rs:44-87                             */
/* ------------------------------------------------------------------------ */
typedef struct {
    double s0, v0, mu, kappa, theta, xi, rho, dt, sqrt_dt, rho_comp;
    uint64_t steps, block;
} heston_k;

static void heston_k_init(heston_k *k, const double *p, double t, uint64_t steps) {
    k->s0 = p[0];
    k->v0 = p[1];
    k->mu = p[2];
    k->kappa = p[3];
    k->theta = p[4];
    k->xi = p[5];
    k->rho = p[6];
    k->dt = t / (double)steps;
    k->sqrt_dt = sqrt(k->dt);
    k->rho_comp = sqrt(1.0 - k->rho * k->rho);
    k->steps = steps;
    k->block = (steps + 1024) * 6; /* heston.rs:48 */
}

static double heston_one(const heston_k *k, uint64_t seed, uint64_t path, double *row) {
    pcg_t g;
    pcg_for_path(&g, seed, k->block, path);
    double s = k->s0, v = k->v0;
    if (row) row[0] = s;
    for (uint64_t j = 0; j < k->steps; j++) { /* heston.rs:59-71 */
        double dw1 = zig_normal(&g);
        double dz = zig_normal(&g);
        double dw2 = k->rho * dw1 + k->rho_comp * dz;
        double v_plus = v > 0.0 ? v : 0.0;
        double sqrt_v = sqrt(v_plus);
        v += k->kappa * (k->theta - v_plus) * k->dt + k->xi * sqrt_v * k->sqrt_dt * dw2;
        s *= exp((k->mu - 0.5 * v_plus) * k->dt + sqrt_v * k->sqrt_dt * dw1);
        if (row) row[j + 1] = s;
    }
    return s;
}

int orc_heston_euler(double s0, double v0, double mu, double kappa, double theta, double xi,
                     double rho, double t, uint64_t steps, uint64_t n_paths, uint64_t seed,
                     double *out) {
    double p[7] = {s0, v0, mu, kappa, theta, xi, rho};
    heston_k k;
    heston_k_init(&k, p, t, steps);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n_paths; i++)
        heston_one(&k, seed, (uint64_t)i, out + (uint64_t)i * (steps + 1));
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Heston QE (Andersen 2008): heston.rs:89-174, every expression per step as  */
/* the reference writes it (no hoisting).                                     */
/* ------------------------------------------------------------------------ */
static double heston_qe_one(const heston_k *k, double t, uint64_t seed, uint64_t path,
                            double *row, double *vrow) {
    const double dt = t / (double)k->steps;               /* heston.rs:92 */
    const double e_kdt = exp(-k->kappa * dt);              /* :93 */
    const double k1 = k->kappa * k->theta;                 /* :94 */
    pcg_t g;
    pcg_for_path(&g, seed, k->block, path);                /* block :97 */
    double s = k->s0, v = k->v0;
    if (row) row[0] = s;
    if (vrow) vrow[0] = v;
    for (uint64_t j = 0; j < k->steps; j++) {
        double v_plus = v > 0.0 ? v : 0.0;                 /* :109 (max, NaN -> 0) */
        double m = k1 * (1.0 - e_kdt) / k->kappa + v_plus * e_kdt;   /* :112 */
        m = m > 1e-15 ? m : 1e-15;                                    /* :113 */
        double one_m = 1.0 - e_kdt;
        double s2 = v_plus * k->xi * k->xi * e_kdt * (1.0 - e_kdt) / k->kappa
                    + k1 * k->xi * k->xi * (one_m * one_m)
                          / (2.0 * k->kappa * k->kappa);               /* :114-116 */
        double psi = s2 / (m * m);                                    /* :117 */
        double v_next;
        if (psi <= 1.5) {                                             /* :119-126 */
            double b2 = 2.0 / psi - 1.0 + sqrt(2.0 / psi * (2.0 / psi - 1.0));
            b2 = b2 > 0.0 ? b2 : 0.0;
            double a = m / (1.0 + b2);
            double b = sqrt(b2);
            double z = zig_normal(&g);
            v_next = a * (b + z) * (b + z);
        } else {                                                      /* :127-136 */
            double p = (psi - 1.0) / (psi + 1.0);
            double beta = (1.0 - p) / m;
            double u = pcg_next_f64(&g);
            v_next = u <= p ? 0.0 : log((1.0 - p) / (1.0 - u)) / beta;
        }
        v_next = v_next > 0.0 ? v_next : 0.0;                         /* :139 */
        double z1 = zig_normal(&g);                                   /* :142 */
        double kappa_rho_xi = k->kappa * k->rho / k->xi;              /* :143 */
        double gamma1 = 0.5, gamma2 = 0.5;
        double k0_s = (k->mu - kappa_rho_xi * k->theta) * dt;
        double k1_s = gamma1 * dt * (kappa_rho_xi - 0.5) - k->rho / k->xi;
        double k2_s = gamma2 * dt * (kappa_rho_xi - 0.5) + k->rho / k->xi;
        double k3_s = gamma1 * dt * (1.0 - k->rho * k->rho);
        double k4_s = gamma2 * dt * (1.0 - k->rho * k->rho);
        double var_term = k3_s * v_plus + k4_s * v_next;              /* :152 */
        var_term = var_term > 0.0 ? var_term : 0.0;
        double log_s = log(s) + k0_s + k1_s * v_plus + k2_s * v_next + sqrt(var_term) * z1;
        if (log_s < -500.0) log_s = -500.0;                           /* :155 clamp */
        else if (log_s > 500.0) log_s = 500.0;
        s = exp(log_s);
        v = v_next;
        if (row) row[j + 1] = s;
        if (vrow) vrow[j + 1] = v;
    }
    return s;
}

/* The variance paths V(t_j) of the QE scheme (test hook: pins the variance
 * construction by its conditional moments).  out: (n_paths, steps+1). */
int orc_heston_qe_variance(double s0, double v0, double mu, double kappa, double theta,
                           double xi, double rho, double t, uint64_t steps, uint64_t n_paths,
                           uint64_t seed, double *out) {
    double p[7] = {s0, v0, mu, kappa, theta, xi, rho};
    heston_k k;
    heston_k_init(&k, p, t, steps);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n_paths; i++)
        heston_qe_one(&k, t, seed, (uint64_t)i, NULL, out + (uint64_t)i * (steps + 1));
    return 0;
}

int orc_heston_qe(double s0, double v0, double mu, double kappa, double theta, double xi,
                  double rho, double t, uint64_t steps, uint64_t n_paths, uint64_t seed,
                  double *out) {
    double p[7] = {s0, v0, mu, kappa, theta, xi, rho};
    heston_k k;
    heston_k_init(&k, p, t, steps);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n_paths; i++)
        heston_qe_one(&k, t, seed, (uint64_t)i, out + (uint64_t)i * (steps + 1), NULL);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Merton jump-diffusion: jump_diffusion.rs:48-105                           */
/* ------------------------------------------------------------------------ */
typedef struct {
    double s0, drift, diffusion, jump_prob, mu_j, sigma_j;
    uint64_t steps, block;
} merton_k;

static void merton_k_init(merton_k *k, const double *p, double t, uint64_t steps) {
    double s0 = p[0], mu = p[1], sigma = p[2], lambda = p[3], mu_j = p[4], sigma_j = p[5];
    double dt = t / (double)steps;
    double sqrt_dt = sqrt(dt);
    double m_bar = exp(mu_j + 0.5 * sigma_j * sigma_j) - 1.0;    /* :53 */
    k->drift = (mu - 0.5 * sigma * sigma - lambda * m_bar) * dt; /* :56 */
    k->diffusion = sigma * sqrt_dt;
    k->jump_prob = lambda * dt;
    k->mu_j = mu_j;
    k->sigma_j = sigma_j;
    k->s0 = s0;
    k->steps = steps;
    k->block = (steps + 1024) * 8; /* :64 */
}

static double merton_one(const merton_k *k, uint64_t seed, uint64_t path, double *row) {
    pcg_t g;
    pcg_for_path(&g, seed, k->block, path);
    double s = k->s0;
    if (row) row[0] = s;
    for (uint64_t j = 0; j < k->steps; j++) { /* :74-89 */
        double z = zig_normal(&g);
        double diffusion_return = k->drift + k->diffusion * z;
        double u = pcg_next_f64(&g);
        double jump_return = 0.0;
        if (u < k->jump_prob) {
            double z_j = zig_normal(&g);
            jump_return = k->mu_j + k->sigma_j * z_j;
        }
        s *= exp(diffusion_return + jump_return);
        if (row) row[j + 1] = s;
    }
    return s;
}

int orc_merton(double s0, double mu, double sigma, double lambda, double mu_j, double sigma_j,
               double t, uint64_t steps, uint64_t n_paths, uint64_t seed, double *out) {
    double p[6] = {s0, mu, sigma, lambda, mu_j, sigma_j};
    merton_k k;
    merton_k_init(&k, p, t, steps);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n_paths; i++)
        merton_one(&k, seed, (uint64_t)i, out + (uint64_t)i * (steps + 1));
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Terminal-only forms and payoff sums                                       */
/* ------------------------------------------------------------------------ */
int orc_terminals(int model, const double *p, double t, uint64_t steps, uint64_t n_paths,
                  uint64_t seed, double *out) {
    if (model == 0) {
        gbm_k k;
        gbm_k_init(&k, p[0], p[1], p[2], t, steps, p[3]);
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t i = 0; i < (int64_t)n_paths; i++)
            out[i] = gbm_one(&k, seed, (uint64_t)i, 0, NULL);
    } else if (model == 1) {
        heston_k k;
        heston_k_init(&k, p, t, steps);
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t i = 0; i < (int64_t)n_paths; i++)
            out[i] = heston_one(&k, seed, (uint64_t)i, NULL);
    } else if (model == 2) {
        merton_k k;
        merton_k_init(&k, p, t, steps);
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t i = 0; i < (int64_t)n_paths; i++)
            out[i] = merton_one(&k, seed, (uint64_t)i, NULL);
    } else if (model == 3) { /* Heston QE */
        heston_k k;
        heston_k_init(&k, p, t, steps);
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t i = 0; i < (int64_t)n_paths; i++)
            out[i] = heston_qe_one(&k, t, seed, (uint64_t)i, NULL, NULL);
    } else {
        return 1;
    }
    return 0;
}

/* greeks.rs:12-23 (Payoff::eval for a call) summed in path order. */
void orc_call_sums(const double *terminals, uint64_t n, const double *strikes, int n_strikes,
                   double *sum, double *sumsq) {
    for (int k = 0; k < n_strikes; k++) {
        double s1 = 0.0, s2 = 0.0;
        for (uint64_t i = 0; i < n; i++) {
            double pay = terminals[i] - strikes[k];
            pay = pay > 0.0 ? pay : 0.0;
            s1 += pay;
            s2 += pay * pay;
        }
        sum[k] = s1;
        sumsq[k] = s2;
    }
}

/* ------------------------------------------------------------------------ */
/* Cholesky: copula/gaussian.rs:76-99 (row-wise, lower triangle only)        */
/* ------------------------------------------------------------------------ */
int orc_cholesky(uint64_t n, const double *a, double *l, char *err, size_t errlen) {
    memset(l, 0, sizeof(double) * n * n);
    for (uint64_t i = 0; i < n; i++) {
        for (uint64_t j = 0; j <= i; j++) {
            double sum = a[i * n + j];
            for (uint64_t k = 0; k < j; k++) sum -= l[i * n + k] * l[j * n + k];
            if (i == j) {
                if (sum <= 0.0) {
                    if (err)
                        snprintf(err, errlen,
                                 "matrix is not positive-definite (diagonal element %llu = %.17g)",
                                 (unsigned long long)i, sum);
                    return 1;
                }
                l[i * n + j] = sqrt(sum);
            } else {
                l[i * n + j] = sum / l[j * n + j];
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Multi-asset GBM: multi_gbm.rs:22-113                                      */
/* ------------------------------------------------------------------------ */
/* multi_gbm.rs:22-113.  Stores the full path (out != NULL) or only the
 * terminal row (term != NULL). */
static int multi_core(uint64_t d, const double *s0, const double *mu, const double *sigma,
                      const double *corr, double t, uint64_t steps, uint64_t n_paths,
                      uint64_t seed, int antithetic, double *out, double *term, char *err,
                      size_t errlen) {
    double *L = (double *)malloc(sizeof(double) * d * d);
    double *drift = (double *)malloc(sizeof(double) * d);
    double *diff = (double *)malloc(sizeof(double) * d);
    if (!L || !drift || !diff) {
        free(L);
        free(drift);
        free(diff);
        return 2;
    }
    if (orc_cholesky(d, corr, L, err, errlen)) {
        free(L);
        free(drift);
        free(diff);
        return 1;
    }
    double dt = t / (double)steps;
    double sqrt_dt = sqrt(dt);
    for (uint64_t i = 0; i < d; i++) {
        drift[i] = (mu[i] - 0.5 * sigma[i] * sigma[i]) * dt; /* :29-31 */
        diff[i] = sigma[i] * sqrt_dt;                          /* :32 */
    }
    uint64_t base = antithetic ? (n_paths + 1) / 2 : n_paths;
    uint64_t block = (steps * d + 1024) * 2; /* multi_gbm.rs:40 */
    uint64_t row_len = (steps + 1) * d;
    int rc = 0;
#pragma omp parallel
    {
        double *z = (double *)malloc(sizeof(double) * d);
        double *s = (double *)malloc(sizeof(double) * d);
        if (!z || !s) {
#pragma omp atomic write
            rc = 2;
        } else {
#pragma omp for schedule(dynamic, 16)
            for (int64_t r = 0; r < (int64_t)n_paths; r++) {
                uint64_t row = (uint64_t)r;
                int negate = row >= base;
                uint64_t path = negate ? row - base : row;
                pcg_t g;
                pcg_for_path(&g, seed, block, path);
                double *o = out ? out + row * row_len : NULL;
                for (uint64_t i = 0; i < d; i++) {
                    s[i] = s0[i];
                    if (o) o[i] = s0[i];
                }
                for (uint64_t j = 0; j < steps; j++) {
                    for (uint64_t i = 0; i < d; i++) z[i] = zig_normal(&g); /* :59 */
                    if (negate)
                        for (uint64_t i = 0; i < d; i++) z[i] = -z[i]; /* :60-64 */
                    for (uint64_t i = 0; i < d; i++) {                /* :67-74 */
                        double eps = 0.0;
                        for (uint64_t k = 0; k <= i; k++) eps += L[i * d + k] * z[k];
                        s[i] *= exp(drift[i] + diff[i] * eps);
                        if (o) o[(j + 1) * d + i] = s[i];
                    }
                }
                if (term)
                    for (uint64_t i = 0; i < d; i++) term[row * d + i] = s[i];
            }
        }
        free(z);
        free(s);
    }
    free(L);
    free(drift);
    free(diff);
    return rc;
}

int orc_multi_gbm(uint64_t d, const double *s0, const double *mu, const double *sigma,
                  const double *corr, double t, uint64_t steps, uint64_t n_paths, uint64_t seed,
                  int antithetic, double *out, char *err, size_t errlen) {
    return multi_core(d, s0, mu, sigma, corr, t, steps, n_paths, seed, antithetic, out, NULL, err,
                      errlen);
}

int orc_multi_terminals(uint64_t d, const double *s0, const double *mu, const double *sigma,
                        const double *corr, double t, uint64_t steps, uint64_t n_paths,
                        uint64_t seed, double *out, char *err, size_t errlen) {
    return multi_core(d, s0, mu, sigma, corr, t, steps, n_paths, seed, 0, NULL, out, err, errlen);
}

/* ------------------------------------------------------------------------ */
/* Least squares by Householder QR (the role of faer's qr().solve_lstsq at   */
/* lsmc.rs:83-84).  X is m x n row-major, m >= n.                            */
/* ------------------------------------------------------------------------ */
int orc_lstsq(uint64_t m, uint64_t n, const double *X, const double *y, double *b) {
    double *A = (double *)malloc(sizeof(double) * m * n);
    double *r = (double *)malloc(sizeof(double) * m);
    if (!A || !r) {
        free(A);
        free(r);
        return 2;
    }
    memcpy(A, X, sizeof(double) * m * n);
    memcpy(r, y, sizeof(double) * m);
    for (uint64_t k = 0; k < n; k++) {
        double norm = 0.0;
        for (uint64_t i = k; i < m; i++) norm += A[i * n + k] * A[i * n + k];
        norm = sqrt(norm);
        if (norm == 0.0) continue;
        double alpha = A[k * n + k] > 0 ? -norm : norm;
        /* v = x - alpha e1, stored in column k below the diagonal */
        double v0 = A[k * n + k] - alpha;
        A[k * n + k] = v0;
        double vnorm2 = v0 * v0;
        for (uint64_t i = k + 1; i < m; i++) vnorm2 += A[i * n + k] * A[i * n + k];
        if (vnorm2 == 0.0) {
            A[k * n + k] = alpha;
            continue;
        }
        for (uint64_t j = k + 1; j < n; j++) {
            double dot = 0.0;
            for (uint64_t i = k; i < m; i++) dot += A[i * n + k] * A[i * n + j];
            double f = 2.0 * dot / vnorm2;
            for (uint64_t i = k; i < m; i++) A[i * n + j] -= f * A[i * n + k];
        }
        double dot = 0.0;
        for (uint64_t i = k; i < m; i++) dot += A[i * n + k] * r[i];
        double f = 2.0 * dot / vnorm2;
        for (uint64_t i = k; i < m; i++) r[i] -= f * A[i * n + k];
        A[k * n + k] = alpha; /* R diagonal */
    }
    for (int64_t k = (int64_t)n - 1; k >= 0; k--) {
        double acc = r[k];
        for (uint64_t j = (uint64_t)k + 1; j < n; j++) acc -= A[k * n + j] * b[j];
        b[k] = acc / A[k * n + k];
    }
    free(A);
    free(r);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* LSMC: lsmc.rs:25-123                                                      */
/* ------------------------------------------------------------------------ */
static double intrinsic(double s, double k, int is_put) { /* lsmc.rs:117-123 */
    double v = is_put ? k - s : s - k;
    return v > 0.0 ? v : 0.0;
}

/* compiler-rt __powidf2: the lowering of Rust's f64::powi (lsmc.rs:78,94). */
static double powi_rt(double a, int b) {
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return r;
}

int orc_lsmc_trace(double s0, double k, double r, double sigma, double t, uint64_t steps,
                   uint64_t n_paths, int is_put, uint64_t poly_degree, uint64_t seed,
                   double *coeffs_out, uint64_t *itm_out, double *price, double *std_err) {
    uint64_t n = n_paths, w = steps + 1, ncols = poly_degree + 1;
    double dt = t / (double)steps;
    double discount = exp(-r * dt);
    double *paths = (double *)malloc(sizeof(double) * n * w);
    double *cf = (double *)malloc(sizeof(double) * n);
    uint64_t *itm = (uint64_t *)malloc(sizeof(uint64_t) * n);
    double *X = (double *)malloc(sizeof(double) * n * ncols);
    double *y = (double *)malloc(sizeof(double) * n);
    double coeffs[8];
    if (!paths || !cf || !itm || !X || !y) {
        free(paths);
        free(cf);
        free(itm);
        free(X);
        free(y);
        return 2;
    }
    orc_gbm(s0, r, sigma, t, steps, n, seed, 0, 0.0, paths); /* lsmc.rs:29-39 */
    for (uint64_t i = 0; i < n; i++) cf[i] = intrinsic(paths[i * w + steps], k, is_put);
    if (coeffs_out)
        for (uint64_t i = 0; i < steps * ncols; i++) coeffs_out[i] = NAN;
    for (uint64_t step = steps - 1; step >= 1; step--) { /* lsmc.rs:50 */
        for (uint64_t i = 0; i < n; i++) cf[i] *= discount;
        uint64_t m = 0;
        for (uint64_t i = 0; i < n; i++)
            if (intrinsic(paths[i * w + step], k, is_put) > 0.0) itm[m++] = i;
        if (itm_out) itm_out[step] = m;
        if (m >= ncols + 1) {
            double s_mean = 0.0;
            for (uint64_t j = 0; j < m; j++) s_mean += paths[itm[j] * w + step];
            s_mean = s_mean / (double)m;
            double v = 0.0;
            for (uint64_t j = 0; j < m; j++) {
                double dlt = paths[itm[j] * w + step] - s_mean;
                v += powi_rt(dlt, 2);
            }
            v = v / (double)m;
            double s_std = sqrt(v);
            if (s_std < 1e-12) s_std = 1e-12;
            for (uint64_t j = 0; j < m; j++) {
                double x = (paths[itm[j] * w + step] - s_mean) / s_std;
                for (uint64_t c = 0; c < ncols; c++) X[j * ncols + c] = powi_rt(x, (int)c);
                y[j] = cf[itm[j]];
            }
            orc_lstsq(m, ncols, X, y, coeffs);
            if (coeffs_out)
                for (uint64_t c = 0; c < ncols; c++) coeffs_out[step * ncols + c] = coeffs[c];
            for (uint64_t j = 0; j < m; j++) { /* lsmc.rs:91-98 */
                double sp = paths[itm[j] * w + step];
                double x = (sp - s_mean) / s_std;
                double cont = 0.0;
                for (uint64_t c = 0; c < ncols; c++) cont += coeffs[c] * powi_rt(x, (int)c);
                double ex = intrinsic(sp, k, is_put);
                if (ex > (cont > 0.0 ? cont : 0.0)) cf[itm[j]] = ex;
            }
        }
        if (step == 1) break;
    }
    for (uint64_t i = 0; i < n; i++) cf[i] *= discount;
    double mean = 0.0;
    for (uint64_t i = 0; i < n; i++) mean += cf[i];
    mean = mean / (double)n;
    double se = 0.0;
    if (n > 1) {
        double var = 0.0;
        for (uint64_t i = 0; i < n; i++) var += powi_rt(cf[i] - mean, 2);
        var = var / (double)(n - 1);
        se = sqrt(var / (double)n);
    }
    *price = mean;
    *std_err = se;
    free(paths);
    free(cf);
    free(itm);
    free(X);
    free(y);
    return 0;
}

int orc_lsmc(double s0, double k, double r, double sigma, double t, uint64_t steps,
             uint64_t n_paths, int is_put, uint64_t poly_degree, uint64_t seed, double *price,
             double *std_err) {
    return orc_lsmc_trace(s0, k, r, sigma, t, steps, n_paths, is_put, poly_degree, seed, NULL,
                          NULL, price, std_err);
}

/* ------------------------------------------------------------------------ */
/* Pathwise Greeks: greeks.rs:360-424 (GBM with mu = r, terminal prices)     */
/* ------------------------------------------------------------------------ */
int orc_greeks_pathwise(double s0, double r, double sigma, double t, double strike, int is_call,
                        uint64_t n_paths, uint64_t steps, uint64_t seed, double *delta,
                        double *vega) {
    double *term = (double *)malloc(sizeof(double) * n_paths);
    if (!term) return 2;
    double p[4] = {s0, r, sigma, 0.0};
    orc_terminals(0, p, t, steps, n_paths, seed, term);
    double discount = exp(-r * t);
    double n = (double)n_paths;
    double sd = 0.0, sv = 0.0;
    for (uint64_t i = 0; i < n_paths; i++) {
        double s_t = term[i];
        int in_money = is_call ? s_t > strike : s_t < strike;
        double sign = is_call ? 1.0 : -1.0;
        if (in_money) sd += sign * s_t / s0;
    }
    for (uint64_t i = 0; i < n_paths; i++) {
        double s_t = term[i];
        int in_money = is_call ? s_t > strike : s_t < strike;
        double sign = is_call ? 1.0 : -1.0;
        if (in_money) {
            double log_return = log(s_t / s0);
            sv += sign * s_t * (log_return - (r + 0.5 * sigma * sigma) * t) / sigma;
        }
    }
    *delta = discount * sd / n;
    *vega = discount * sv / n;
    free(term);
    return 0;
}

/* ---- bump-and-revalue Greeks (greeks.rs:227-358) ----
 * spec: the ModelSpec fields in the reference's params-dict order (GBM: s0, r,
 * sigma, t; Heston: s0, v0, r, kappa, theta, xi, rho, t; Merton: s0, r,
 * sigma, lambda, mu_j, sigma_j, t).  sign +1 call / -1 put (greeks.rs:12-17).
 * greeks: ids 0 delta, 1 gamma, 2 vega, 3 theta, 4 rho. */
enum { ORC_S0 = 0, ORC_SIGMA = 1, ORC_T = 2, ORC_MU = 3 };

/* ModelSpec::param_value / with_bumped (greeks.rs:169-224): field index of
 * S0, Sigma (Heston: v0), T, Mu (the rate) */
static int orc_fd_index(int model, int param) {
    static const int idx[4][4] = {{0, 2, 3, 1}, {0, 1, 7, 2}, {0, 2, 6, 1}, {0, 1, 7, 2}};
    return idx[model][param];
}

/* mc_price (greeks.rs:304-311): terminal_prices (greeks.rs:91-167, GBM with
 * q = 0), Payoff::eval, exp(-r t) * (left-to-right sum) / n. */
static double orc_mc_price(int model, const double *spec, double sign, double strike,
                           uint64_t steps, uint64_t n_paths, uint64_t seed, double *term) {
    double p[8];
    double t;
    if (model == 0) {
        p[0] = spec[0]; p[1] = spec[1]; p[2] = spec[2]; p[3] = 0.0;
        t = spec[3];
    } else if (model == 1 || model == 3) {
        for (int i = 0; i < 7; i++) p[i] = spec[i];
        t = spec[7];
    } else {
        for (int i = 0; i < 6; i++) p[i] = spec[i];
        t = spec[6];
    }
    orc_terminals(model, p, t, steps, n_paths, seed, term);
    double sum = 0.0;
    for (uint64_t i = 0; i < n_paths; i++) {
        double v = sign > 0 ? term[i] - strike : strike - term[i];
        sum += v > 0.0 ? v : 0.0;   /* f64::max(0.0); a NaN payoff counts 0 */
    }
    double r = spec[orc_fd_index(model, ORC_MU)];
    double discount = exp(-r * t);
    return discount * sum / (double)n_paths;
}

int orc_greeks_fd(int model, const double *spec, double sign, double strike, uint64_t n_paths,
                  uint64_t steps, const int *greeks, int n_greeks, uint64_t seed,
                  double bump_size, double *values) {
    if (model < 0 || model > 3) return 1;
    const int nspec = model == 0 ? 4 : (model == 1 || model == 3) ? 8 : 7;
    /* required_scenarios (greeks.rs:231-265): key = param * 2 + (dir < 0), base = 8 */
    int keys[9], nk = 0;
#define ORC_ADD(k) do { int _k = (k), _f = 0; for (int _i = 0; _i < nk; _i++) _f |= keys[_i] == _k; \
                        if (!_f) keys[nk++] = _k; } while (0)
    for (int g = 0; g < n_greeks; g++) {
        switch (greeks[g]) {
            case 0: ORC_ADD(ORC_S0 * 2); ORC_ADD(ORC_S0 * 2 + 1); break;
            case 1: ORC_ADD(ORC_S0 * 2); ORC_ADD(8); ORC_ADD(ORC_S0 * 2 + 1); break;
            case 2: ORC_ADD(ORC_SIGMA * 2); ORC_ADD(ORC_SIGMA * 2 + 1); break;
            case 3: ORC_ADD(ORC_T * 2 + 1); ORC_ADD(8); break;
            case 4: ORC_ADD(ORC_MU * 2); ORC_ADD(ORC_MU * 2 + 1); break;
            default: return 1;
        }
    }
#undef ORC_ADD
    /* compute_bump (greeks.rs:227-229), first occurrence of each parameter */
    double bump[4] = {0, 0, 0, 0};
    for (int i = 0; i < nk; i++) {
        if (keys[i] == 8) continue;
        int prm = keys[i] / 2;
        double b = fabs(spec[orc_fd_index(model, prm)]) * bump_size;
        bump[prm] = b > 1e-8 ? b : 1e-8;
    }
    double price[9];
    double *term = (double *)malloc(sizeof(double) * (n_paths ? n_paths : 1));
    if (!term) return 1;
    for (int i = 0; i < nk; i++) {
        double v[8];
        for (int j = 0; j < nspec; j++) v[j] = spec[j];
        if (keys[i] != 8) {
            int prm = keys[i] / 2;
            double delta = (keys[i] & 1) ? -bump[prm] : bump[prm];
            v[orc_fd_index(model, prm)] += delta;  /* with_bumped */
        }
        price[keys[i]] = orc_mc_price(model, v, sign, strike, steps, n_paths, seed, term);
    }
    free(term);
    /* compute_greeks_from_prices (greeks.rs:267-302) */
    for (int g = 0; g < n_greeks; g++) {
        double h;
        switch (greeks[g]) {
            case 0: h = bump[ORC_S0];
                values[g] = (price[0] - price[1]) / (2.0 * h); break;
            case 1: h = bump[ORC_S0];
                values[g] = (price[0] - 2.0 * price[8] + price[1]) / (h * h); break;
            case 2: h = bump[ORC_SIGMA];
                values[g] = (price[2] - price[3]) / (2.0 * h); break;
            case 3: h = bump[ORC_T];
                values[g] = (price[5] - price[8]) / h; break;
            default: h = bump[ORC_MU];
                values[g] = (price[6] - price[7]) / (2.0 * h); break;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Path-dependent options on GBM: barrier.rs:286-414, asian.rs:44-355,        */
/* lookback.rs:158-210.  The loops follow the reference: a knocked-out       */
/* barrier path stops drawing at its hit; sums run in path order.            */
/* ------------------------------------------------------------------------ */
/* bs.rs:3-25 (Abramowitz & Stegun 7.1.26) */
static double orc_erf(double x) {
    double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741, a4 = -1.453152027,
           a5 = 1.061405429, p = 0.3275911;
    double sign = x < 0.0 ? -1.0 : 1.0;
    x = fabs(x);
    double t = 1.0 / (1.0 + p * x);
    double y = 1.0 - (((((a5 * t + a4) * t) + a3) * t + a2) * t + a1) * t * exp(-x * x);
    return sign * y;
}
static double orc_norm_cdf(double x) { return 0.5 * (1.0 + orc_erf(x * 0.70710678118654752440)); }
static double orc_max0(double v) { return v > 0.0 ? v : 0.0; }

/* barrier.rs:286-372; up/out/call flags; rebate_at_hit as in BarrierParams. */
double orc_barrier_mc(double s0, double k, double r, double q, double sigma, double t,
                      double barrier, int up, int out, int call, double rebate, int rebate_at_hit,
                      uint64_t n_paths, uint64_t n_steps, uint64_t seed) {
    double dt = t / (double)n_steps;
    double drift = (r - q - 0.5 * sigma * sigma) * dt;
    double diffusion = sigma * sqrt(dt);
    double discount = exp(-r * t);
    uint64_t block = (n_steps + 1024) * 2;
    int already_hit = up ? s0 >= barrier : s0 <= barrier;
    double *vals = (double *)malloc(sizeof(double) * (n_paths ? n_paths : 1));
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n_paths; i++) {
        pcg_t g;
        pcg_for_path(&g, seed, block, (uint64_t)i);
        double s = s0;
        int64_t hit = -1;
        if (already_hit) {
            hit = 0;
        } else {
            for (uint64_t step = 1; step <= n_steps; step++) {
                double z = zig_normal(&g);
                s *= exp(drift + diffusion * z);
                if (up ? s >= barrier : s <= barrier) {
                    hit = (int64_t)step;
                    break;
                }
            }
        }
        int knocked = hit >= 0;
        if (knocked && !out && (uint64_t)hit < n_steps) {
            for (uint64_t j = (uint64_t)hit; j < n_steps; j++) {
                double z = zig_normal(&g);
                s *= exp(drift + diffusion * z);
            }
        }
        double intrinsic = orc_max0(call ? s - k : k - s);
        double option = out ? (knocked ? 0.0 : intrinsic) : (knocked ? intrinsic : 0.0);
        double rebate_pv = 0.0; /* rebate_path_pv, barrier.rs:374-400 */
        if (rebate != 0.0) {
            if (out && rebate_at_hit) {
                rebate_pv = knocked ? rebate * exp(-r * ((double)hit * dt)) : 0.0;
            } else if (out) {
                rebate_pv = knocked ? rebate * exp(-r * ((double)n_steps * dt)) : 0.0;
            } else {
                rebate_pv = knocked ? 0.0 : rebate * exp(-r * ((double)n_steps * dt));
            }
        }
        vals[i] = option + rebate_pv / discount;
    }
    double total = 0.0;
    for (uint64_t i = 0; i < n_paths; i++) total += vals[i];
    free(vals);
    return discount * total / (double)n_paths;
}

/* lookback.rs:158-210; has_max/has_min select the Option values. */
double orc_lookback_mc(double s0, double k, double r, double q, double sigma, double t,
                       int floating, int call, int has_max, double running_max, int has_min,
                       double running_min, uint64_t n_paths, uint64_t n_steps, uint64_t seed) {
    double dt = t / (double)n_steps;
    double drift = (r - q - 0.5 * sigma * sigma) * dt;
    double diffusion = sigma * sqrt(dt);
    double discount = exp(-r * t);
    uint64_t block = (n_steps + 1024) * 2;
    double m0 = has_max ? running_max : s0, n0 = has_min ? running_min : s0;
    m0 = (m0 != m0) ? s0 : (m0 > s0 ? m0 : s0); /* f64::max / min ignore a NaN */
    n0 = (n0 != n0) ? s0 : (n0 < s0 ? n0 : s0);
    double *vals = (double *)malloc(sizeof(double) * (n_paths ? n_paths : 1));
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n_paths; i++) {
        pcg_t g;
        pcg_for_path(&g, seed, block, (uint64_t)i);
        double s = s0, s_max = m0, s_min = n0;
        for (uint64_t j = 0; j < n_steps; j++) {
            double z = zig_normal(&g);
            s *= exp(drift + diffusion * z);
            if (s > s_max) s_max = s;
            if (s < s_min) s_min = s;
        }
        double pay;
        if (floating) pay = call ? s - s_min : s_max - s;
        else pay = call ? orc_max0(s_max - k) : orc_max0(k - s_min);
        vals[i] = pay;
    }
    double total = 0.0;
    for (uint64_t i = 0; i < n_paths; i++) total += vals[i];
    free(vals);
    return discount * total / (double)n_paths;
}

/* asian.rs:78-137 without seasoning (Kemna-Vorst as the reference writes it) */
static double orc_asian_geo_analytical(double s, double k, double r, double q, double sigma,
                                       double t, int call) {
    double sigma_g = sigma / sqrt(3.0);
    double mu_g = 0.5 * (r - q - 0.5 * sigma * sigma) + 0.5 * 0.5 * sigma_g * sigma_g;
    double d1 = (log(s / k) + (mu_g + 0.5 * sigma_g * sigma_g) * t) / (sigma_g * sqrt(t));
    double d2 = d1 - sigma_g * sqrt(t);
    double df = exp(-r * t);
    double price = call ? s * exp((mu_g - r) * t) * orc_norm_cdf(d1) - k * df * orc_norm_cdf(d2)
                        : k * df * orc_norm_cdf(-d2) - s * exp((mu_g - r) * t) * orc_norm_cdf(-d1);
    return orc_max0(price);
}

/* asian.rs:143-290 for an unseasoned option (the seasoned forms reduce to it
 * or to asian_mc_floating_seasoned, below). */
double orc_asian_mc(double s0, double k, double r, double q, double sigma, double t,
                    uint64_t n_steps, int geometric, int fixed, int call, uint64_t n_paths,
                    uint64_t seed) {
    double dt = t / (double)n_steps;
    double drift = (r - q - 0.5 * sigma * sigma) * dt;
    double diffusion = sigma * sqrt(dt);
    double discount = exp(-r * t);
    uint64_t block = (n_steps + 1024) * 2;
    int use_cv = !geometric && fixed;
    double geo_an = use_cv ? orc_asian_geo_analytical(s0, k, r, q, sigma, t, call) : 0.0;
    double *a = (double *)malloc(sizeof(double) * (n_paths ? n_paths : 1));
    double *gv = (double *)malloc(sizeof(double) * (n_paths ? n_paths : 1));
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n_paths; i++) {
        pcg_t g;
        pcg_for_path(&g, seed, block, (uint64_t)i);
        double s = s0, sum_arith = 0.0, sum_log = 0.0;
        for (uint64_t j = 0; j < n_steps; j++) {
            double z = zig_normal(&g);
            s *= exp(drift + diffusion * z);
            sum_arith += s;
            sum_log += log(s);
        }
        double avg_arith = sum_arith / (double)n_steps;
        double avg_geo = exp(sum_log / (double)n_steps);
        double pa, pg;
        if (fixed) {
            pa = call ? orc_max0(avg_arith - k) : orc_max0(k - avg_arith);
            pg = call ? orc_max0(avg_geo - k) : orc_max0(k - avg_geo);
        } else {
            pa = call ? orc_max0(s - avg_arith) : orc_max0(avg_arith - s);
            pg = call ? orc_max0(s - avg_geo) : orc_max0(avg_geo - s);
        }
        a[i] = pa;
        gv[i] = pg;
    }
    double res;
    if (use_cv && n_paths > 1) {
        double n = (double)n_paths, sa = 0.0, sg = 0.0;
        for (uint64_t i = 0; i < n_paths; i++) sa += a[i];
        for (uint64_t i = 0; i < n_paths; i++) sg += gv[i];
        double mean_arith = sa / n, mean_geo = sg / n, cov = 0.0, var_geo = 0.0;
        for (uint64_t i = 0; i < n_paths; i++) {
            cov += (a[i] - mean_arith) * (gv[i] - mean_geo);
            var_geo += (gv[i] - mean_geo) * (gv[i] - mean_geo);
        }
        double beta = var_geo > 1e-30 ? cov / var_geo : 0.0;
        double geo_mc_price = discount * mean_geo;
        res = orc_max0(discount * mean_arith - beta * (geo_mc_price - geo_an));
    } else {
        double sa = 0.0;
        for (uint64_t i = 0; i < n_paths; i++) sa += a[i];
        res = orc_max0(discount * (sa / (double)n_paths));
    }
    free(a);
    free(gv);
    return res;
}

/* asian.rs:293-355 */
double orc_asian_floating_seasoned(double s0, double r, double q, double sigma, double t,
                                   uint64_t n_steps, int geometric, int call, double a_spent,
                                   double t1, uint64_t n_paths, uint64_t seed) {
    double remaining = t - t1;
    double nr = round((double)n_steps * remaining / t);
    uint64_t n_rem = nr >= 1.0 ? (uint64_t)nr : 1;
    double dt = remaining / (double)n_rem;
    double drift = (r - q - 0.5 * sigma * sigma) * dt;
    double diffusion = sigma * sqrt(dt);
    double discount = exp(-r * remaining);
    double w_spent = t1 / t, w_remain = remaining / t;
    uint64_t block = (n_rem + 1024) * 2;
    double *vals = (double *)malloc(sizeof(double) * (n_paths ? n_paths : 1));
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n_paths; i++) {
        pcg_t g;
        pcg_for_path(&g, seed, block, (uint64_t)i);
        double s = s0, sum_arith = 0.0, sum_log = 0.0;
        for (uint64_t j = 0; j < n_rem; j++) {
            double z = zig_normal(&g);
            s *= exp(drift + diffusion * z);
            sum_arith += s;
            sum_log += log(s);
        }
        double avg_remain = geometric ? exp(sum_log / (double)n_rem) : sum_arith / (double)n_rem;
        double avg_t = geometric ? exp(log(a_spent) * w_spent + log(avg_remain) * w_remain)
                                 : w_spent * a_spent + w_remain * avg_remain;
        vals[i] = call ? orc_max0(s - avg_t) : orc_max0(avg_t - s);
    }
    double total = 0.0;
    for (uint64_t i = 0; i < n_paths; i++) total += vals[i];
    free(vals);
    return discount * total / (double)n_paths;
}

/* ------------------------------------------------------------------------ */
/* The RNG class (src/lib.rs:75-184): one sequential stream.                  */
/* st[4] = state lo, state hi, increment lo, increment hi (in/out).           */
/* ------------------------------------------------------------------------ */
void orc_rng_new(uint64_t seed_lo, uint64_t seed_hi, uint64_t st[4]) {
    /* pcg64.rs:23-30 with a 128-bit seed */
    u128 s = ((u128)seed_hi << 64) | seed_lo;
    u128 stream = s + (((u128)0xda3e39cb94b95bdbull << 64) | 0x7ef55da5c0374f5aull);
    u128 inc = (stream << 1) | 1;
    u128 state = (s + inc) * (u128)PCG_CHEAP_MUL + inc;
    st[0] = (uint64_t)state;
    st[1] = (uint64_t)(state >> 64);
    st[2] = (uint64_t)inc;
    st[3] = (uint64_t)(inc >> 64);
}

static void orc_rng_load(pcg_t *g, const uint64_t st[4]) {
    g->state = ((u128)st[1] << 64) | st[0];
    g->inc = ((u128)st[3] << 64) | st[2];
}

static void orc_rng_store(const pcg_t *g, uint64_t st[4]) {
    st[0] = (uint64_t)g->state;
    st[1] = (uint64_t)(g->state >> 64);
}

/* lib.rs:108-112 */
void orc_rng_uniform(uint64_t st[4], uint64_t n, double *out) {
    pcg_t g;
    orc_rng_load(&g, st);
    for (uint64_t i = 0; i < n; i++) out[i] = pcg_next_f64(&g);
    orc_rng_store(&g, st);
}

/* lib.rs:114-145: sample_into, or sample_scaled (normal.rs:70-79) */
void orc_rng_normal(uint64_t st[4], uint64_t n, double loc, double scale, int scaled,
                    double *out) {
    pcg_t g;
    orc_rng_load(&g, st);
    for (uint64_t i = 0; i < n; i++) {
        double z = zig_normal(&g);
        out[i] = scaled ? loc + scale * z : z;
    }
    orc_rng_store(&g, st);
}

/* ------------------------------------------------------------------------ */
/* Hull-White one-factor, exact transition: hull_white.rs:26-52              */
/* ------------------------------------------------------------------------ */
int orc_hull_white(double r0, double a, double theta, double sigma, double t, uint64_t steps,
                   uint64_t n_paths, uint64_t seed, double *out) {
    double dt = t / (double)steps;
    double e = exp(-a * dt);
    double mean_reversion = (theta / a) * (1.0 - e);
    double vol = sigma * sqrt((1.0 - e * e) / (2.0 * a));
    uint64_t block = (steps + 1024) * 4;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n_paths; i++) {
        pcg_t g;
        pcg_for_path(&g, seed, block, (uint64_t)i);
        double *row = out + (uint64_t)i * (steps + 1);
        row[0] = r0;
        for (uint64_t s = 0; s < steps; s++) {
            double z = zig_normal(&g);
            row[s + 1] = row[s] * e + mean_reversion + vol * z;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Gaussian copula: copula/gaussian.rs:15-73 on one stream new(seed);         */
/* L is the lower Cholesky factor (d x d row-major), computed by the caller  */
/* with orc_cholesky.                                                        */
/* ------------------------------------------------------------------------ */
static double orc_norm_cdf_2617(double x) { /* dist/normal_cdf.rs:3-21 */
    const double P = 0.2316419;
    const double B[5] = {0.319381530, -0.356563782, 1.781477937, -1.821255978, 1.330274429};
    double t = 1.0 / (1.0 + P * fabs(x));
    double poly = t * (B[0] + t * (B[1] + t * (B[2] + t * (B[3] + t * B[4]))));
    double pdf = exp(-0.5 * x * x) / sqrt(2.0 * 3.14159265358979323846);
    double cdf_pos = 1.0 - pdf * poly;
    return x >= 0.0 ? cdf_pos : 1.0 - cdf_pos;
}

void orc_gaussian_copula(const double *L, uint64_t d, uint64_t n_samples, uint64_t seed,
                         double *out) {
    pcg_t g;
    pcg_new(&g, seed);
    double *z = (double *)malloc(sizeof(double) * (d ? d : 1));
    for (uint64_t i = 0; i < n_samples; i++) {
        for (uint64_t k = 0; k < d; k++) z[k] = zig_normal(&g);
        for (uint64_t j = 0; j < d; j++) {
            double xj = 0.0;
            for (uint64_t k = 0; k <= j; k++) xj += L[j * d + k] * z[k];
            out[i * d + j] = orc_norm_cdf_2617(xj);
        }
    }
    free(z);
}

/* ------------------------------------------------------------------------ */
/* Student-t copula: copula/student_t.rs on one stream new(seed).            */
/* ------------------------------------------------------------------------ */
/* student_t.rs:67-91: log-gamma, Lanczos g = 7 with 9 coefficients */
double orc_lgamma_lanczos(double x) {
    static const double C[9] = {0.99999999999980993,  676.5203681218851,     -1259.1392167224028,
                                771.32342877765313,   -176.61502916214059,   12.507343278686905,
                                -0.13857109526572012, 9.9843695780195716e-6, 1.5056327351493116e-7};
    const double PI = 3.14159265358979323846;
    if (x < 0.5) return log(PI) - log(sin(PI * x)) - orc_lgamma_lanczos(1.0 - x);
    x -= 1.0;
    double acc = C[0];
    for (int i = 1; i < 9; i++) acc += C[i] / (x + (double)i);
    double t = x + 7.0 + 0.5;
    return 0.5 * log(2.0 * PI) + (x + 0.5) * log(t) - t + log(acc);
}

/* student_t.rs:32-63: continued fraction (modified Lentz), EPS 3e-7 */
static double orc_beta_cf(double x, double a, double b) {
    const double EPS = 3e-7, FPMIN = 1e-30;
    double qab = a + b, qap = a + 1.0, qam = a - 1.0;
    double c = 1.0;
    double d = fmax(1.0 - qab * x / qap, FPMIN);
    d = 1.0 / d;
    double h = d;
    for (int m = 1; m <= 200; m++) {
        double mf = (double)m, m2 = 2.0 * mf;
        double aa = mf * (b - mf) * x / ((qam + m2) * (a + m2));
        d = fmax(fabs(1.0 + aa * d), FPMIN);
        c = fmax(fabs(1.0 + aa / c), FPMIN);
        d = 1.0 / d;
        h *= d * c;
        aa = -(a + mf) * (qab + mf) * x / ((a + m2) * (qap + m2));
        d = fmax(fabs(1.0 + aa * d), FPMIN);
        c = fmax(fabs(1.0 + aa / c), FPMIN);
        d = 1.0 / d;
        double del = d * c;
        h *= del;
        if (fabs(del - 1.0) < EPS) break;
    }
    return h;
}

/* student_t.rs:19-30: regularized incomplete beta I_x(a, b) */
static double orc_ibeta(double x, double a, double b) {
    if (x <= 0.0) return 0.0;
    if (x >= 1.0) return 1.0;
    if (x > (a + 1.0) / (a + b + 2.0)) return 1.0 - orc_ibeta(1.0 - x, b, a);
    double log_beta = orc_lgamma_lanczos(a) + orc_lgamma_lanczos(b) - orc_lgamma_lanczos(a + b);
    double front = exp(a * log(x) + b * log(1.0 - x) - log_beta) / a;
    return front * orc_beta_cf(x, a, b);
}

/* student_t.rs:8-16 */
double orc_student_t_cdf(double x, double nu) {
    double z = nu / (nu + x * x);
    double ib = orc_ibeta(z, nu / 2.0, 0.5);
    return x >= 0.0 ? 1.0 - 0.5 * ib : 0.5 * ib;
}

/* student_t.rs:94-119: Marsaglia-Tsang gamma(shape), shape < 1 by the
 * Ahrens-Dieter boost (a uniform first) */
static double orc_gamma_mt(pcg_t *g, double shape) {
    if (shape < 1.0) {
        double u = pcg_next_f64(g);
        return orc_gamma_mt(g, shape + 1.0) * pow(u, 1.0 / shape);
    }
    double d = shape - 1.0 / 3.0;
    double c = 1.0 / sqrt(9.0 * d);
    for (;;) {
        double x = zig_normal(g);
        double v_inner = 1.0 + c * x;
        if (v_inner <= 0.0) continue;
        double v = v_inner * v_inner * v_inner;
        double u = pcg_next_f64(g);
        if (u < 1.0 - 0.0331 * (x * x) * (x * x)) return d * v;
        if (log(u) < 0.5 * x * x + d * (1.0 - v + log(v))) return d * v;
    }
}

/* student_t.rs:135-169: per sample a chi-square(nu) = 2 gamma(nu/2), then d
 * normals; u_ij = t_cdf((L z_i)_j sqrt(nu / chi2), nu).  Also returns the
 * stream draws consumed (*used) and, when z_out / s_out are given, the
 * normals and the scale of every sample. */
void orc_student_t_copula(const double *L, uint64_t d, double nu, uint64_t n_samples,
                          uint64_t seed, double *out, double *z_out, double *s_out,
                          uint64_t *used) {
    pcg_t g;
    pcg_new(&g, seed);
    double *z = (double *)malloc(sizeof(double) * (d ? d : 1));
    for (uint64_t i = 0; i < n_samples; i++) {
        double chi_sq = 2.0 * orc_gamma_mt(&g, nu / 2.0);
        double scale = sqrt(nu / chi_sq);
        for (uint64_t k = 0; k < d; k++) z[k] = zig_normal(&g);
        if (z_out)
            for (uint64_t k = 0; k < d; k++) z_out[i * d + k] = z[k];
        if (s_out) s_out[i] = scale;
        for (uint64_t j = 0; j < d; j++) {
            double xj = 0.0;
            for (uint64_t k = 0; k <= j; k++) xj += L[j * d + k] * z[k];
            out[i * d + j] = orc_student_t_cdf(xj * scale, nu);
        }
    }
    free(z);
    if (used) {
        /* draws consumed: count the LCG steps from s0 to the final state */
        pcg_t h;
        pcg_new(&h, seed);
        uint64_t k = 0;
        while (h.state != g.state) {
            h.state = h.state * (u128)PCG_CHEAP_MUL + h.inc;
            k++;
        }
        *used = k;
    }
}
