"""ctypes loader for the CPU oracle (libstocha_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  The functions
restate the reference's algorithms; see stocha_oracle.c for the file:line
each follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libstocha_oracle.so")
_lib = None

_d = C.c_double
_u = C.c_uint64
_i = C.c_int
_pd = C.POINTER(C.c_double)
_pu = C.POINTER(C.c_uint64)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_stream_state.argtypes = [_u, _u, _u, _pu]
        L.orc_advance.argtypes = [_pu, _u, _u]
        L.orc_raw_u64.argtypes = [_u, _u, _u, _u, _pu]
        L.orc_normals.argtypes = [_u, _u, _u, _u, _pd, _pu]
        L.orc_gbm.argtypes = [_d, _d, _d, _d, _u, _u, _u, _i, _d, _pd]
        L.orc_heston_euler.argtypes = [_d] * 8 + [_u, _u, _u, _pd]
        L.orc_merton.argtypes = [_d] * 7 + [_u, _u, _u, _pd]
        L.orc_heston_qe.argtypes = [_d] * 8 + [_u, _u, _u, _pd]
        L.orc_heston_qe_variance.argtypes = [_d] * 8 + [_u, _u, _u, _pd]
        L.orc_multi_gbm.argtypes = [_u, _pd, _pd, _pd, _pd, _d, _u, _u, _u, _i, _pd,
                                    C.c_char_p, C.c_size_t]
        L.orc_cholesky.argtypes = [_u, _pd, _pd, C.c_char_p, C.c_size_t]
        L.orc_multi_terminals.argtypes = [_u, _pd, _pd, _pd, _pd, _d, _u, _u, _u, _pd,
                                          C.c_char_p, C.c_size_t]
        L.orc_terminals.argtypes = [_i, _pd, _d, _u, _u, _u, _pd]
        L.orc_call_sums.argtypes = [_pd, _u, _pd, _i, _pd, _pd]
        L.orc_call_sums.restype = None
        L.orc_lsmc.argtypes = [_d, _d, _d, _d, _d, _u, _u, _i, _u, _u, _pd, _pd]
        L.orc_lsmc_trace.argtypes = [_d, _d, _d, _d, _d, _u, _u, _i, _u, _u, _pd, _pu, _pd,
                                     _pd]
        L.orc_lstsq.argtypes = [_u, _u, _pd, _pd, _pd]
        L.orc_greeks_pathwise.argtypes = [_d, _d, _d, _d, _d, _i, _u, _u, _u, _pd, _pd]
        L.orc_greeks_fd.argtypes = [_i, _pd, _d, _d, _u, _u, C.POINTER(C.c_int), _i, _u, _d,
                                    _pd]
        L.orc_barrier_mc.argtypes = [_d] * 7 + [_i, _i, _i, _d, _i, _u, _u, _u]
        L.orc_barrier_mc.restype = _d
        L.orc_lookback_mc.argtypes = [_d] * 6 + [_i, _i, _i, _d, _i, _d, _u, _u, _u]
        L.orc_lookback_mc.restype = _d
        L.orc_asian_mc.argtypes = [_d] * 6 + [_u, _i, _i, _i, _u, _u]
        L.orc_asian_mc.restype = _d
        L.orc_asian_floating_seasoned.argtypes = [_d] * 5 + [_u, _i, _i, _d, _d, _u, _u]
        L.orc_asian_floating_seasoned.restype = _d
        L.orc_rng_new.argtypes = [_u, _u, _pu]
        L.orc_rng_new.restype = None
        L.orc_rng_uniform.argtypes = [_pu, _u, _pd]
        L.orc_rng_uniform.restype = None
        L.orc_rng_normal.argtypes = [_pu, _u, _d, _d, _i, _pd]
        L.orc_rng_normal.restype = None
        L.orc_hull_white.argtypes = [_d] * 5 + [_u, _u, _u, _pd]
        L.orc_gaussian_copula.argtypes = [_pd, _u, _u, _u, _pd]
        L.orc_gaussian_copula.restype = None
        L.orc_lgamma_lanczos.argtypes = [_d]
        L.orc_lgamma_lanczos.restype = _d
        L.orc_student_t_cdf.argtypes = [_d, _d]
        L.orc_student_t_cdf.restype = _d
        L.orc_student_t_copula.argtypes = [_pd, _u, _d, _u, _u, _pd, _pd, _pd,
                                           C.POINTER(C.c_uint64)]
        L.orc_student_t_copula.restype = None
        L.orc_num_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = None
        _lib = L
    return _lib


def _p(a, t=_pd):
    return a.ctypes.data_as(t)


# Block sizes per model (gbm.rs:55, heston.rs:48, jump_diffusion.rs:64, multi_gbm.rs:40)
def block_gbm(steps):
    return (steps + 1024) * 2


def block_heston(steps):
    return (steps + 1024) * 6


def block_merton(steps):
    return (steps + 1024) * 8


def block_multi(steps, d):
    return (steps * d + 1024) * 2


def stream_state(seed, block, path):
    out = np.zeros(4, np.uint64)
    lib().orc_stream_state(seed, block, path, _p(out, _pu))
    return out


def raw_u64(seed, block, path, n):
    out = np.zeros(n, np.uint64)
    lib().orc_raw_u64(seed, block, path, n, _p(out, _pu))
    return out


def normals(seed, block, path, n):
    out = np.zeros(n, np.float64)
    used = np.zeros(1, np.uint64)
    lib().orc_normals(seed, block, path, n, _p(out), _p(used, _pu))
    return out, int(used[0])


def gbm(s0, mu, sigma, t, steps, n_paths, seed=42, antithetic=False, q=0.0):
    out = np.empty((n_paths, steps + 1), np.float64)
    lib().orc_gbm(s0, mu, sigma, t, steps, n_paths, seed, int(antithetic), q, _p(out))
    return out


def heston(s0, v0, mu, kappa, theta, xi, rho, t, steps, n_paths, seed=42):
    out = np.empty((n_paths, steps + 1), np.float64)
    lib().orc_heston_euler(s0, v0, mu, kappa, theta, xi, rho, t, steps, n_paths, seed, _p(out))
    return out


def heston_qe(s0, v0, mu, kappa, theta, xi, rho, t, steps, n_paths, seed=42):
    out = np.empty((n_paths, steps + 1), np.float64)
    rc = lib().orc_heston_qe(s0, v0, mu, kappa, theta, xi, rho, t, steps, n_paths, seed,
                             _p(out))
    assert rc == 0
    return out


def heston_qe_variance(s0, v0, mu, kappa, theta, xi, rho, t, steps, n_paths, seed=42):
    out = np.empty((n_paths, steps + 1), np.float64)
    rc = lib().orc_heston_qe_variance(s0, v0, mu, kappa, theta, xi, rho, t, steps, n_paths,
                                      seed, _p(out))
    assert rc == 0
    return out


def merton(s0, mu, sigma, lambda_, mu_j, sigma_j, t, steps, n_paths, seed=42):
    out = np.empty((n_paths, steps + 1), np.float64)
    lib().orc_merton(s0, mu, sigma, lambda_, mu_j, sigma_j, t, steps, n_paths, seed, _p(out))
    return out


def cholesky(a):
    a = np.ascontiguousarray(a, np.float64)
    n = a.shape[0]
    l = np.zeros_like(a)
    err = C.create_string_buffer(256)
    rc = lib().orc_cholesky(n, _p(a), _p(l), err, 256)
    if rc:
        raise ValueError(err.value.decode())
    return l


def multi_gbm(s0, mu, sigma, corr, t, steps, n_paths, seed=42, antithetic=False):
    d = len(s0)
    s0 = np.ascontiguousarray(s0, np.float64)
    mu = np.ascontiguousarray(mu, np.float64)
    sigma = np.ascontiguousarray(sigma, np.float64)
    corr = np.ascontiguousarray(corr, np.float64)
    out = np.empty((n_paths, steps + 1, d), np.float64)
    err = C.create_string_buffer(256)
    rc = lib().orc_multi_gbm(d, _p(s0), _p(mu), _p(sigma), _p(corr), t, steps, n_paths, seed,
                             int(antithetic), _p(out), err, 256)
    if rc:
        raise ValueError(err.value.decode())
    return out


def multi_terminals(s0, mu, sigma, corr, t, steps, n_paths, seed=42):
    d = len(s0)
    s0 = np.ascontiguousarray(s0, np.float64)
    mu = np.ascontiguousarray(mu, np.float64)
    sigma = np.ascontiguousarray(sigma, np.float64)
    corr = np.ascontiguousarray(corr, np.float64)
    out = np.empty((n_paths, d), np.float64)
    err = C.create_string_buffer(256)
    rc = lib().orc_multi_terminals(d, _p(s0), _p(mu), _p(sigma), _p(corr), t, steps, n_paths,
                                   seed, _p(out), err, 256)
    if rc:
        raise ValueError(err.value.decode())
    return out


MODEL_GBM, MODEL_HESTON, MODEL_MERTON = 0, 1, 2


def terminals(model, params, t, steps, n_paths, seed=42):
    p = np.ascontiguousarray(params, np.float64)
    out = np.empty(n_paths, np.float64)
    rc = lib().orc_terminals(model, _p(p), t, steps, n_paths, seed, _p(out))
    assert rc == 0
    return out


def call_sums(term, strikes):
    term = np.ascontiguousarray(term, np.float64)
    k = np.ascontiguousarray(strikes, np.float64)
    s1 = np.zeros(len(k))
    s2 = np.zeros(len(k))
    lib().orc_call_sums(_p(term), len(term), _p(k), len(k), _p(s1), _p(s2))
    return s1, s2


def lsmc(s0, k, r, sigma, t, steps, n_paths, is_put=True, poly_degree=3, seed=42):
    price = np.zeros(1)
    se = np.zeros(1)
    rc = lib().orc_lsmc(s0, k, r, sigma, t, steps, n_paths, int(is_put), poly_degree, seed,
                        _p(price), _p(se))
    assert rc == 0
    return float(price[0]), float(se[0])


def lsmc_trace(s0, k, r, sigma, t, steps, n_paths, is_put=True, poly_degree=3, seed=42):
    coeffs = np.zeros((steps, poly_degree + 1))
    itm = np.zeros(steps, np.uint64)
    price = np.zeros(1)
    se = np.zeros(1)
    rc = lib().orc_lsmc_trace(s0, k, r, sigma, t, steps, n_paths, int(is_put), poly_degree, seed,
                              _p(coeffs), _p(itm, _pu), _p(price), _p(se))
    assert rc == 0
    return coeffs, itm, float(price[0]), float(se[0])


def lstsq(X, y):
    X = np.ascontiguousarray(X, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    b = np.zeros(X.shape[1])
    lib().orc_lstsq(X.shape[0], X.shape[1], _p(X), _p(y), _p(b))
    return b


def greeks_pathwise(s0, r, sigma, t, strike, is_call, n_paths, n_steps, seed=42):
    d = np.zeros(1)
    v = np.zeros(1)
    lib().orc_greeks_pathwise(s0, r, sigma, t, strike, int(is_call), n_paths, n_steps, seed,
                              _p(d), _p(v))
    return float(d[0]), float(v[0])


_GREEK_IDS = {"delta": 0, "gamma": 1, "vega": 2, "theta": 3, "rho": 4}


def greeks_fd(model, spec, payoff, strike, n_paths, n_steps, greeks, seed=42, bump_size=0.01):
    """greeks_fd_core (greeks.rs:227-358) with a built-in payoff; model id
    0/1/2, spec in the params-dict order.  Returns {name: value}."""
    sp = np.ascontiguousarray(spec, np.float64)
    ids = (C.c_int * max(1, len(greeks)))(*[_GREEK_IDS[g] for g in greeks])
    out = np.zeros(max(1, len(greeks)))
    rc = lib().orc_greeks_fd(model, _p(sp), 1.0 if payoff == "call" else -1.0, strike, n_paths,
                             n_steps, ids, len(greeks), seed, bump_size, _p(out))
    assert rc == 0
    return {g: float(out[i]) for i, g in enumerate(greeks)}


_BARRIER = {"up-and-out": (1, 1), "up-and-in": (1, 0), "down-and-out": (0, 1),
            "down-and-in": (0, 0)}


def barrier_mc(s, k, r, sigma, t, barrier, barrier_type, option_type, q=0.0, n_paths=100_000,
               n_steps=252, seed=42, rebate=0.0, rebate_at_hit=False):
    up, out = _BARRIER[barrier_type]
    return lib().orc_barrier_mc(s, k, r, q, sigma, t, barrier, up, out,
                                int(option_type == "call"), rebate, int(rebate_at_hit), n_paths,
                                n_steps, seed)


def lookback_mc(s, r, sigma, t, n_steps, strike_type, option_type, k=0.0, q=0.0,
                n_paths=100_000, seed=42, running_max=None, running_min=None):
    return lib().orc_lookback_mc(s, k, r, q, sigma, t, int(strike_type == "floating"),
                                 int(option_type == "call"), int(running_max is not None),
                                 running_max or 0.0, int(running_min is not None),
                                 running_min or 0.0, n_paths, n_steps, seed)


def asian_mc(s, k, r, sigma, t, n_steps, average_type, strike_type, option_type, q=0.0,
             n_paths=100_000, seed=42):
    return lib().orc_asian_mc(s, k, r, q, sigma, t, n_steps, int(average_type == "geometric"),
                              int(strike_type == "fixed"), int(option_type == "call"), n_paths,
                              seed)


def asian_floating_seasoned(s, r, sigma, t, n_steps, average_type, option_type, a_spent, t1,
                            q=0.0, n_paths=100_000, seed=42):
    return lib().orc_asian_floating_seasoned(s, r, q, sigma, t, n_steps,
                                             int(average_type == "geometric"),
                                             int(option_type == "call"), a_spent, t1, n_paths,
                                             seed)


class Rng:
    """The reference's RNG class as a sequential CPU stream (lib.rs:75-184)."""

    def __init__(self, seed=0):
        self.st = np.zeros(4, np.uint64)
        lib().orc_rng_new(seed & (2**64 - 1), seed >> 64, _p(self.st, _pu))

    def state(self):
        s = [int(v) for v in self.st]
        return s[0] | (s[1] << 64), s[2] | (s[3] << 64)

    def uniform(self, n):
        out = np.empty(n, np.float64)
        lib().orc_rng_uniform(_p(self.st, _pu), n, _p(out))
        return out

    def normal(self, n, loc=0.0, scale=1.0, scaled=False):
        out = np.empty(n, np.float64)
        lib().orc_rng_normal(_p(self.st, _pu), n, loc, scale, int(scaled), _p(out))
        return out


def hull_white(r0, a, theta, sigma, t, steps, n_paths, seed=42):
    out = np.empty((n_paths, steps + 1), np.float64)
    assert lib().orc_hull_white(r0, a, theta, sigma, t, steps, n_paths, seed, _p(out)) == 0
    return out


def gaussian_copula(corr, n_samples, seed=42):
    corr = np.ascontiguousarray(corr, np.float64)
    d = corr.shape[0]
    L = cholesky(corr) if d else np.zeros((0, 0))
    L = np.ascontiguousarray(L, np.float64)
    out = np.empty((n_samples, d), np.float64)
    lib().orc_gaussian_copula(_p(L), d, n_samples, seed, _p(out))
    return out


def lgamma_lanczos(x):
    """copula/student_t.rs:67-91 (Lanczos g = 7, 9 coefficients)."""
    return lib().orc_lgamma_lanczos(float(x))


def student_t_cdf(x, nu):
    """copula/student_t.rs:8-63 (incomplete beta by modified Lentz); x scalar
    or array."""
    xs = np.asarray(x, np.float64)
    f = lib().orc_student_t_cdf
    out = np.array([f(float(v), float(nu)) for v in xs.ravel()], np.float64)
    return out.reshape(xs.shape) if xs.shape else float(out[0])


def student_t_copula(corr, nu, n_samples, seed=42, details=False):
    """copula/student_t.rs:135-169 on the stream Pcg64Dxsm::new(seed).  With
    details=True also returns the normals (n, d), the scales sqrt(nu / chi2)
    (n,) and the draws consumed."""
    corr = np.ascontiguousarray(corr, np.float64)
    d = corr.shape[0]
    L = cholesky(corr) if d else np.zeros((0, 0))
    L = np.ascontiguousarray(L, np.float64)
    out = np.empty((n_samples, d), np.float64)
    z = np.empty((n_samples, d), np.float64)
    s = np.empty(n_samples, np.float64)
    used = C.c_uint64(0)
    lib().orc_student_t_copula(_p(L), d, float(nu), n_samples, seed, _p(out), _p(z), _p(s),
                               C.byref(used))
    if details:
        return out, z, s, used.value
    return out


def num_threads():
    return lib().orc_num_threads()


def set_threads(n):
    lib().orc_set_threads(int(n))
