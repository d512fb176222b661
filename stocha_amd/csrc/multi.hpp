// multi.hpp -- correlated multi-asset GBM entry points (multi.hip).
#pragma once
#include <string>

#include "host.hpp"

namespace stocha {

// copula/gaussian.rs:76-99 on the host, FP64; returns STOCHA_EINVAL with the
// reference's message when a pivot is not positive.
int cholesky_host(const double* a, uint64_t n, double* l, const Err& err);

// multi_gbm.rs:22-113, FP64, path-storing: out (n_paths, steps+1, d).
int run_multi_gbm_paths(const double* s0, const double* mu, const double* sigma,
                        const double* corr, uint64_t d, double t, uint64_t steps,
                        uint64_t n_paths, uint64_t seed, int antithetic, double* out, int where,
                        hipStream_t st, const Err& err);

std::string rust_f64(double v);

}  // namespace stocha
