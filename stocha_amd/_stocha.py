"""Binding layer: the hot-path functions of the reference's `stocha._stocha`.

Mirrors the PyO3 functions of src/lib.rs (signatures, keyword names,
defaults, argument extraction and error types) and calls the C ABI of
libstocha_gpu.so.  Argument extraction follows PyO3's rules: floats accept
anything with __float__ (PyFloat_AsDouble), usize/u64 arguments take Python
ints only (TypeError otherwise) and raise OverflowError when negative or
too large, bool arguments take bool only, `scheme` takes str, `greeks` takes a
list.  The range checks and their ValueError texts live in the C library and
run before any device work.
"""
from __future__ import annotations

import ctypes as C
import json as _json
import operator

import numpy as np

from . import _lib as L

__version__ = "1.7.1"  # the reference version whose API this module mirrors


# ------------------------------------------------------------ extraction --
def _f64(x, name):
    if isinstance(x, (bool, np.bool_)):
        return float(x)
    try:
        return float(x) if isinstance(x, (int, float, np.floating, np.integer)) else float(
            x.__float__())
    except AttributeError:
        raise TypeError(f"argument '{name}': must be real number, not {type(x).__name__}")


def _uint(x, name, bits=64):
    if isinstance(x, (float, np.floating)) or not hasattr(x, "__index__"):
        raise TypeError(f"argument '{name}': '{type(x).__name__}' object cannot be interpreted "
                        "as an integer")
    v = operator.index(x)
    if v < 0:
        raise OverflowError(f"argument '{name}': can't convert negative int to unsigned")
    if v >= 1 << bits:
        raise OverflowError(f"argument '{name}': int too big to convert")
    return v


def _bool(x, name):
    if isinstance(x, (bool, np.bool_)):
        return bool(x)
    raise TypeError(f"argument '{name}': '{type(x).__name__}' object is not an instance of "
                    "'bool'")


def _vec_f64(x, name):
    if isinstance(x, (str, bytes)):
        raise TypeError(f"argument '{name}': Can't extract `str` to `Vec`")
    return np.ascontiguousarray([_f64(v, name) for v in x], dtype=np.float64)


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ----------------------------------------------------------- entry points --
def gbm(s0, mu, sigma, t, steps, n_paths, seed=42, antithetic=False, q=0.0):
    """src/lib.rs:205-252 (signature :206).  Returns (n_paths, steps+1) float64."""
    s0, mu, sigma, t = (_f64(s0, "s0"), _f64(mu, "mu"), _f64(sigma, "sigma"), _f64(t, "t"))
    steps, n_paths = _uint(steps, "steps"), _uint(n_paths, "n_paths")
    seed, antithetic, q = _uint(seed, "seed"), _bool(antithetic, "antithetic"), _f64(q, "q")
    lib, eb = L.load(), L.ErrBuf()
    # out=NULL validates only, so a bad argument never allocates the output
    L.check(lib.stocha_gbm(s0, mu, sigma, t, steps, n_paths, seed, int(antithetic), q, None,
                           L.STOCHA_HOST, eb.buf, eb.n), eb)
    out = L.host_array((n_paths, steps + 1))
    rc = lib.stocha_gbm(s0, mu, sigma, t, steps, n_paths, seed, int(antithetic), q,
                        out.ctypes.data, L.STOCHA_HOST, eb.buf, eb.n)
    L.check(rc, eb)
    return out


def heston(s0, v0, mu, kappa, theta, xi, rho, t, steps, n_paths, seed=42, scheme="euler"):
    """src/lib.rs:327-408 (signature :328).  Returns (n_paths, steps+1) float64."""
    vals = [_f64(v, n) for v, n in ((s0, "s0"), (v0, "v0"), (mu, "mu"), (kappa, "kappa"),
                                     (theta, "theta"), (xi, "xi"), (rho, "rho"), (t, "t"))]
    steps, n_paths, seed = _uint(steps, "steps"), _uint(n_paths, "n_paths"), _uint(seed, "seed")
    if not isinstance(scheme, str):
        raise TypeError(f"argument 'scheme': '{type(scheme).__name__}' object cannot be "
                        "converted to 'PyString'")
    lib, eb = L.load(), L.ErrBuf()
    sch = scheme.encode()
    L.check(lib.stocha_heston(*vals, steps, n_paths, seed, sch, None, L.STOCHA_HOST, eb.buf,
                              eb.n), eb)
    out = L.host_array((n_paths, steps + 1))
    rc = lib.stocha_heston(*vals, steps, n_paths, seed, sch, out.ctypes.data, L.STOCHA_HOST,
                           eb.buf, eb.n)
    L.check(rc, eb)
    return out


def merton_jump_diffusion(s0, mu, sigma, lambda_, mu_j, sigma_j, t, steps, n_paths, seed=42):
    """src/lib.rs:430-489 (signature :431).  Returns (n_paths, steps+1) float64."""
    vals = [_f64(v, n) for v, n in ((s0, "s0"), (mu, "mu"), (sigma, "sigma"),
                                     (lambda_, "lambda_"), (mu_j, "mu_j"),
                                     (sigma_j, "sigma_j"), (t, "t"))]
    steps, n_paths, seed = _uint(steps, "steps"), _uint(n_paths, "n_paths"), _uint(seed, "seed")
    lib, eb = L.load(), L.ErrBuf()
    L.check(lib.stocha_merton_jump_diffusion(*vals, steps, n_paths, seed, None, L.STOCHA_HOST,
                                             eb.buf, eb.n), eb)
    out = L.host_array((n_paths, steps + 1))
    rc = lib.stocha_merton_jump_diffusion(*vals, steps, n_paths, seed, out.ctypes.data,
                                          L.STOCHA_HOST, eb.buf, eb.n)
    L.check(rc, eb)
    return out


def multi_gbm(s0, mu, sigma, corr, t, steps, n_paths, seed=42, antithetic=False):
    """src/lib.rs:759-827 (signature :760).  Returns (n_paths, steps+1, d) float64."""
    s0, mu, sigma = _vec_f64(s0, "s0"), _vec_f64(mu, "mu"), _vec_f64(sigma, "sigma")
    if not isinstance(corr, np.ndarray) or corr.dtype != np.float64:
        raise TypeError("argument 'corr': expected a numpy.ndarray of float64")
    if corr.ndim != 2:
        raise TypeError(f"argument 'corr': expected a 2-dimensional array, got {corr.ndim}")
    corr_c = np.ascontiguousarray(corr)
    t = _f64(t, "t")
    steps, n_paths = _uint(steps, "steps"), _uint(n_paths, "n_paths")
    seed, antithetic = _uint(seed, "seed"), _bool(antithetic, "antithetic")
    lib, eb = L.load(), L.ErrBuf()
    d = len(s0)

    def call(ptr):
        return lib.stocha_multi_gbm(_pd(s0), len(s0), _pd(mu), len(mu), _pd(sigma), len(sigma),
                                    _pd(corr_c), corr_c.shape[0], corr_c.shape[1], t, steps,
                                    n_paths, seed, int(antithetic), ptr, L.STOCHA_HOST, eb.buf,
                                    eb.n)

    L.check(call(None), eb)
    out = L.host_array((n_paths, steps + 1, d))
    L.check(call(out.ctypes.data), eb)
    return out


def lsmc_american_option(s0, k, r, sigma, t, steps, n_paths, is_put=True, poly_degree=3,
                         seed=42):
    """src/lib.rs:854-918 (signature :855).  Returns (price, std_err)."""
    return _lsmc(s0, k, r, sigma, t, steps, n_paths, is_put, poly_degree, seed, L.STOCHA_FP64)


def _lsmc(s0, k, r, sigma, t, steps, n_paths, is_put, poly_degree, seed, precision):
    vals = [_f64(v, n) for v, n in ((s0, "s0"), (k, "k"), (r, "r"), (sigma, "sigma"),
                                     (t, "t"))]
    steps, n_paths = _uint(steps, "steps"), _uint(n_paths, "n_paths")
    is_put = _bool(is_put, "is_put")
    poly_degree, seed = _uint(poly_degree, "poly_degree"), _uint(seed, "seed")
    lib, eb = L.load(), L.errbuf()
    price, se = C.c_double(), C.c_double()
    rc = lib.stocha_lsmc_ex(*vals, steps, n_paths, int(is_put), poly_degree, seed, precision,
                            C.byref(price), C.byref(se), eb.buf, eb.n)
    L.check(rc, eb)
    return float(price.value), float(se.value)


def greeks_pathwise(s0, r, sigma, t, strike, is_call, n_paths, n_steps, greeks, seed=42):
    """src/lib.rs:1177-1242 (signature :1178).  Returns {name: value} in request order."""
    vals = [_f64(v, n) for v, n in ((s0, "s0"), (r, "r"), (sigma, "sigma"), (t, "t"),
                                     (strike, "strike"))]
    is_call = _bool(is_call, "is_call")
    n_paths, n_steps = _uint(n_paths, "n_paths"), _uint(n_steps, "n_steps")
    if not isinstance(greeks, list):
        raise TypeError(f"argument 'greeks': '{type(greeks).__name__}' object cannot be "
                        "converted to 'PyList'")
    names = []
    for g in greeks:
        if not isinstance(g, str):
            raise TypeError(f"'{type(g).__name__}' object cannot be converted to 'PyString'")
        names.append(g)
    seed = _uint(seed, "seed")
    lib, eb = L.load(), L.ErrBuf()
    arr = (C.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
    out = np.zeros(max(1, len(names)), dtype=np.float64)
    rc = lib.stocha_greeks_pathwise(*vals, int(is_call), n_paths, n_steps, arr, len(names), seed,
                                    _pd(out), eb.buf, eb.n)
    L.check(rc, eb)
    res = {}
    for i, n in enumerate(names):
        res[n] = float(out[i])
    return res


# ------------------------------------------------------------- greeks_fd --
# lib.rs:958-1038 (build_model_spec): the params dict keys per model, in the
# order the C ABI takes them.
_FD_MODELS = {"gbm": (L.MODEL_GBM, ("s0", "r", "sigma", "t")),
              "heston": (L.MODEL_HESTON, ("s0", "v0", "r", "kappa", "theta", "xi", "rho", "t")),
              "merton": (L.MODEL_MERTON, ("s0", "r", "sigma", "lambda_", "mu_j", "sigma_j",
                                          "t"))}
_GREEKS = ("delta", "gamma", "vega", "theta", "rho")
# ModelSpec::param_value / with_bumped (greeks.rs:169-224): index of the
# bumped field in the spec (S0, Sigma -- Heston's v0 --, T, Mu -- the rate r)
_FD_INDEX = {"gbm": (0, 2, 3, 1), "heston": (0, 1, 7, 2), "merton": (0, 2, 6, 1)}
_S0, _SIGMA, _T, _MU = 0, 1, 2, 3


def _parse_greeks(greeks):
    """lib.rs:920-938."""
    if not isinstance(greeks, list):
        raise TypeError(f"argument 'greeks': '{type(greeks).__name__}' object cannot be "
                        "converted to 'PyList'")
    out = []
    for g in greeks:
        if not isinstance(g, str):
            raise TypeError(f"'{type(g).__name__}' object cannot be converted to 'PyString'")
        if g not in _GREEKS:
            raise ValueError(f"Unknown greek: '{g}'. Supported: delta, gamma, vega, theta, rho")
        out.append(g)
    return out


def _model_spec(model, params):
    """lib.rs:958-1038: (model id, field values, scheme)."""
    if model not in _FD_MODELS:
        raise ValueError(f"Unknown model: '{model}'. Supported: gbm, heston, merton")
    mid, keys = _FD_MODELS[model]
    vals = []
    for k in keys:
        if k not in params:
            raise ValueError(f"Missing required param: '{k}'")
        vals.append(_f64(params[k], k))
    scheme = None
    if model == "heston":
        scheme = params.get("scheme", "euler")
        if not isinstance(scheme, str):
            raise TypeError(f"'{type(scheme).__name__}' object cannot be converted to 'PyString'")
        if scheme not in ("euler", "qe"):
            raise ValueError("scheme must be 'euler' or 'qe'")
    return mid, np.array(vals, dtype=np.float64), scheme


def _required_scenarios(greeks):
    """greeks.rs:231-265: (param, dir) keys, None = Base, first-seen order."""
    keys = []

    def add(k):
        if k not in keys:
            keys.append(k)
    for g in greeks:
        if g == "delta":
            add((_S0, 1)); add((_S0, -1))
        elif g == "gamma":
            add((_S0, 1)); add(None); add((_S0, -1))
        elif g == "vega":
            add((_SIGMA, 1)); add((_SIGMA, -1))
        elif g == "theta":
            add((_T, -1)); add(None)
        else:
            add((_MU, 1)); add((_MU, -1))
    return keys


def _greeks_from_prices(greeks, prices, bumps):
    """greeks.rs:267-302."""
    res = {}
    for g in greeks:
        if g == "delta":
            h = bumps[_S0]
            v = (prices[(_S0, 1)] - prices[(_S0, -1)]) / (2.0 * h)
        elif g == "gamma":
            h = bumps[_S0]
            v = (prices[(_S0, 1)] - 2.0 * prices[None] + prices[(_S0, -1)]) / (h * h)
        elif g == "vega":
            h = bumps[_SIGMA]
            v = (prices[(_SIGMA, 1)] - prices[(_SIGMA, -1)]) / (2.0 * h)
        elif g == "theta":
            h = bumps[_T]
            v = (prices[(_T, -1)] - prices[None]) / h
        else:
            h = bumps[_MU]
            v = (prices[(_MU, 1)] - prices[(_MU, -1)]) / (2.0 * h)
        res[g] = float(v)
    return res


def _terminal_prices(mid, vals, scheme, n_steps, n_paths, seed):
    out = L.host_array((n_paths,))
    lib, eb = L.load(), L.ErrBuf()
    rc = lib.stocha_terminal_prices(mid, _pd(vals), len(vals),
                                    None if scheme is None else scheme.encode(), n_steps,
                                    n_paths, seed, out.ctypes.data, L.STOCHA_HOST, eb.buf, eb.n)
    L.check(rc, eb)
    return out


def greeks_fd(model, params, payoff, strike, n_paths, n_steps, greeks, seed=42, bump_size=0.01):
    """src/lib.rs:1064-1156 (signature :1063).  Returns {name: value}."""
    if not isinstance(model, str):
        raise TypeError(f"argument 'model': '{type(model).__name__}' object cannot be converted "
                        "to 'PyString'")
    if not isinstance(params, dict):
        raise TypeError(f"argument 'params': '{type(params).__name__}' object cannot be "
                        "converted to 'PyDict'")
    strike = _f64(strike, "strike")
    n_paths, n_steps = _uint(n_paths, "n_paths"), _uint(n_steps, "n_steps")
    if not isinstance(greeks, list):
        raise TypeError(f"argument 'greeks': '{type(greeks).__name__}' object cannot be "
                        "converted to 'PyList'")
    seed, bump_size = _uint(seed, "seed"), _f64(bump_size, "bump_size")
    names = _parse_greeks(greeks)
    mid, vals, scheme = _model_spec(model, params)
    sch = None if scheme is None else scheme.encode()
    lib, eb = L.load(), L.ErrBuf()
    if isinstance(payoff, str):
        arr = (C.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
        out = np.zeros(max(1, len(names)), dtype=np.float64)
        rc = lib.stocha_greeks_fd(mid, _pd(vals), len(vals), sch, payoff.encode(), strike,
                                  n_paths, n_steps, arr, len(names), seed, bump_size, _pd(out),
                                  eb.buf, eb.n)
        L.check(rc, eb)
        return {n: float(out[i]) for i, n in enumerate(names)}
    if not callable(payoff):
        raise ValueError("payoff must be 'call', 'put', or a callable")
    # lib.rs:1118-1150: terminals per scenario, the callable's payoffs, the
    # price exp(-r t) * sum / len with the reference's left-to-right sum
    rc = lib.stocha_terminal_prices(mid, _pd(vals), len(vals), sch, n_steps, 0, seed, None,
                                    L.STOCHA_HOST, eb.buf, eb.n)
    L.check(rc, eb)
    keys = _required_scenarios(names)
    idx = _FD_INDEX[model]
    bumps = {}
    for k in keys:
        if k is not None and k[0] not in bumps:
            b = abs(float(vals[idx[k[0]]])) * bump_size
            bumps[k[0]] = b if b > 1e-8 else 1e-8  # compute_bump, greeks.rs:227-229
    prices = {}
    for k in keys:
        v = vals.copy()
        if k is not None:
            v[idx[k[0]]] += k[1] * bumps[k[0]]
        r_rate, mat = float(v[idx[_MU]]), float(v[idx[_T]])
        terminals = _terminal_prices(mid, v, scheme, n_steps, n_paths, seed)
        pv = np.asarray(payoff(terminals), dtype=np.float64).reshape(-1)
        total = float(np.cumsum(pv)[-1]) if pv.size else 0.0
        discount = np.exp(np.float64(-r_rate * mat))
        with np.errstate(divide="ignore", invalid="ignore"):
            prices[k] = float(np.float64(discount * total) / np.float64(pv.size))
    return _greeks_from_prices(names, prices, bumps)


# ---------------------------------------------- path-dependent options --
def _str(x, name):
    if not isinstance(x, str):
        raise TypeError(f"argument '{name}': '{type(x).__name__}' object cannot be converted to "
                        "'PyString'")
    return x.encode()


def _opt_f64(x, name):
    return (0, 0.0) if x is None else (1, _f64(x, name))


def _call_price(fn, *args):
    lib, eb = L.load(), L.ErrBuf()
    out = np.zeros(1, dtype=np.float64)
    L.check(getattr(lib, fn)(*args, _pd(out), eb.buf, eb.n), eb)
    return float(out[0])


def barrier_price(s, k, r, sigma, t, barrier, barrier_type="up-and-out", option_type="call",
                  q=0.0, n_paths=100_000, n_steps=252, seed=42, method="auto",
                  n_monitoring=None, rebate=0.0, rebate_at_hit=False):
    """src/lib.rs:1551-1638 (signature :1549)."""
    vals = [_f64(v, n) for v, n in ((s, "s"), (k, "k"), (r, "r"), (sigma, "sigma"), (t, "t"),
                                     (barrier, "barrier"))]
    bt, ot = _str(barrier_type, "barrier_type"), _str(option_type, "option_type")
    q = _f64(q, "q")
    n_paths, n_steps, seed = _uint(n_paths, "n_paths"), _uint(n_steps, "n_steps"), \
        _uint(seed, "seed")
    m = _str(method, "method")
    has_mon, mon = (0, 0) if n_monitoring is None else (1, _uint(n_monitoring, "n_monitoring",
                                                                 bits=32))
    rebate, rebate_at_hit = _f64(rebate, "rebate"), _bool(rebate_at_hit, "rebate_at_hit")
    return _call_price("stocha_barrier_price", *vals, bt, ot, q, n_paths, n_steps, seed, m,
                       has_mon, mon, rebate, int(rebate_at_hit))


def asian_price(s, k, r, sigma, t, n_steps=252, average_type="arithmetic", strike_type="fixed",
                option_type="call", q=0.0, n_paths=100_000, seed=42, method="auto",
                running_avg=None, time_elapsed=None):
    """src/lib.rs:1647-1745 (signature :1645)."""
    vals = [_f64(v, n) for v, n in ((s, "s"), (k, "k"), (r, "r"), (sigma, "sigma"), (t, "t"))]
    n_steps = _uint(n_steps, "n_steps")
    at, st = _str(average_type, "average_type"), _str(strike_type, "strike_type")
    ot, q = _str(option_type, "option_type"), _f64(q, "q")
    n_paths, seed, m = _uint(n_paths, "n_paths"), _uint(seed, "seed"), _str(method, "method")
    ha, a = _opt_f64(running_avg, "running_avg")
    he, e = _opt_f64(time_elapsed, "time_elapsed")
    return _call_price("stocha_asian_price", *vals, n_steps, at, st, ot, q, n_paths, seed, m,
                       ha, a, he, e)


def lookback_price(s, r, sigma, t, n_steps=252, strike_type="floating", option_type="call",
                   k=0.0, q=0.0, n_paths=100_000, seed=42, method="auto", running_max=None,
                   running_min=None):
    """src/lib.rs:1761-1848 (signature :1759)."""
    vals = [_f64(v, n) for v, n in ((s, "s"), (r, "r"), (sigma, "sigma"), (t, "t"))]
    n_steps = _uint(n_steps, "n_steps")
    st, ot = _str(strike_type, "strike_type"), _str(option_type, "option_type")
    k, q = _f64(k, "k"), _f64(q, "q")
    n_paths, seed, m = _uint(n_paths, "n_paths"), _uint(seed, "seed"), _str(method, "method")
    hx, mx = _opt_f64(running_max, "running_max")
    hn, mn = _opt_f64(running_min, "running_min")
    return _call_price("stocha_lookback_price", *vals, n_steps, st, ot, k, q, n_paths, seed, m,
                       hx, mx, hn, mn)


# -------------------------------------------------------------- RNG class --
_M64 = (1 << 64) - 1


def _u128_field(obj, key):
    v = obj[key]
    if isinstance(v, bool) or not isinstance(v, int) or v < 0 or v >= 1 << 128:
        raise ValueError(f"invalid value for '{key}': expected u128")
    return v


class RNG:
    """src/lib.rs:75-184: a PCG64DXSM generator (one sequential stream) whose
    bulk draws are made on the GPU.  The stream position is kept on the host
    as the 128-bit LCG state; save_state / from_state use the reference's
    JSON (pcg64.rs:57-77: {"version":2,"rng":{"state":..,"increment":..},
    "seed":..}, or the legacy {"seed":..})."""

    def __init__(self, seed=0, algorithm="pcg64dxsm"):
        seed = _uint(seed, "seed", bits=128)
        if not isinstance(algorithm, str):
            raise TypeError(f"argument 'algorithm': '{type(algorithm).__name__}' object cannot "
                            "be converted to 'PyString'")
        if algorithm not in ("pcg64dxsm", "default"):
            raise ValueError(f"Unknown algorithm: '{algorithm}'. Supported: 'pcg64dxsm'")
        st, inc = (C.c_uint64 * 2)(), (C.c_uint64 * 2)()
        L.load().stocha_rng_new(seed & _M64, seed >> 64, st, inc)
        self._seed = seed
        self._state = (st[0] | (st[1] << 64))
        self._inc = (inc[0] | (inc[1] << 64))

    @property
    def seed(self):
        return self._seed

    def _args(self):
        st = (C.c_uint64 * 2)(self._state & _M64, self._state >> 64)
        inc = (C.c_uint64 * 2)(self._inc & _M64, self._inc >> 64)
        return st, inc, (C.c_uint64 * 2)()

    def uniform(self, size):
        """lib.rs:108-112: size draws of next_f64."""
        size = _uint(size, "size")
        out = L.host_array((size,))
        st, inc, new = self._args()
        lib, eb = L.load(), L.ErrBuf()
        L.check(lib.stocha_rng_uniform(st, inc, size, out.ctypes.data, L.STOCHA_HOST, new,
                                       eb.buf, eb.n), eb)
        self._state = new[0] | (new[1] << 64)
        return out

    def _normals(self, size, loc, scale, scaled):
        out = L.host_array((size,))
        st, inc, new = self._args()
        lib, eb = L.load(), L.ErrBuf()
        L.check(lib.stocha_rng_normal(st, inc, size, loc, scale, int(scaled), out.ctypes.data,
                                      L.STOCHA_HOST, new, None, eb.buf, eb.n), eb)
        self._state = new[0] | (new[1] << 64)
        return out

    def standard_normal(self, size):
        """lib.rs:114-119: size ziggurat normals (NormalSampler::sample_into)."""
        return self._normals(_uint(size, "size"), 0.0, 1.0, False)

    def normal(self, size, loc=0.0, scale=1.0):
        """lib.rs:121-145: loc + scale * z per value (sample_scaled)."""
        size, loc, scale = _uint(size, "size"), _f64(loc, "loc"), _f64(scale, "scale")
        if scale <= 0.0:
            raise ValueError("scale must be positive")
        return self._normals(size, loc, scale, True)

    def save_state(self):
        """lib.rs:147-157 -> pcg64.rs:57-65."""
        return _json.dumps({"version": 2, "rng": {"state": self._state, "increment": self._inc},
                            "seed": self._seed}, separators=(",", ":"))

    @staticmethod
    def from_state(json):
        """lib.rs:159-178 -> pcg64.rs:67-77: the full state, else the legacy
        seed-only form; anything else raises ValueError."""
        if not isinstance(json, str):
            raise TypeError(f"argument 'json': '{type(json).__name__}' object cannot be "
                            "converted to 'PyString'")
        try:
            obj = _json.loads(json)
        except ValueError as e:
            raise ValueError(str(e)) from None
        if not isinstance(obj, dict):
            raise ValueError("invalid type: expected a JSON object")
        try:
            v = obj["version"]
            if isinstance(v, bool) or not isinstance(v, int) or not 0 <= v < 256:
                raise ValueError("invalid value for 'version': expected u8")
            rng = obj["rng"]
            if not isinstance(rng, dict):
                raise ValueError("invalid type for 'rng'")
            state, inc, seed = _u128_field(rng, "state"), _u128_field(rng, "increment"), \
                _u128_field(obj, "seed")
            r = RNG.__new__(RNG)
            r._seed, r._state, r._inc = seed, state, inc
            return r
        except (KeyError, ValueError):
            pass
        try:
            seed = _u128_field(obj, "seed")
        except KeyError:
            raise ValueError("missing field `seed`") from None
        return RNG(seed)

    def __repr__(self):
        return f"RNG(seed={self._seed}, algorithm='pcg64dxsm')"


# ----------------------------------------------- Hull-White, Gaussian copula --
def hull_white(r0, a, theta, sigma, t, steps, n_paths, seed=42):
    """src/lib.rs:606-658 (signature :605).  Returns (n_paths, steps+1) float64."""
    vals = [_f64(v, n) for v, n in ((r0, "r0"), (a, "a"), (theta, "theta"), (sigma, "sigma"),
                                     (t, "t"))]
    steps, n_paths, seed = _uint(steps, "steps"), _uint(n_paths, "n_paths"), _uint(seed, "seed")
    lib, eb = L.load(), L.ErrBuf()
    L.check(lib.stocha_hull_white(*vals, steps, n_paths, seed, None, L.STOCHA_HOST, eb.buf, eb.n),
            eb)
    out = L.host_array((n_paths, steps + 1))
    L.check(lib.stocha_hull_white(*vals, steps, n_paths, seed, out.ctypes.data, L.STOCHA_HOST,
                                  eb.buf, eb.n), eb)
    return out


def student_t_copula(corr, nu, n_samples, seed=42):
    """src/lib.rs:563-578 (signature :562).  Returns (n_samples, dim) float64."""
    if not isinstance(corr, np.ndarray) or corr.dtype != np.float64:
        raise TypeError("argument 'corr': expected a numpy.ndarray of float64")
    if corr.ndim != 2:
        raise TypeError(f"argument 'corr': expected a 2-dimensional array, got {corr.ndim}")
    nu = _f64(nu, "nu")
    n_samples, seed = _uint(n_samples, "n_samples"), _uint(seed, "seed")
    c = np.ascontiguousarray(corr)
    lib, eb = L.load(), L.ErrBuf()
    L.check(lib.stocha_student_t_copula(_pd(c), c.shape[0], c.shape[1], nu, n_samples, seed,
                                        None, L.STOCHA_HOST, eb.buf, eb.n), eb)
    out = L.host_array((n_samples, c.shape[0]))
    L.check(lib.stocha_student_t_copula(_pd(c), c.shape[0], c.shape[1], nu, n_samples, seed,
                                        out.ctypes.data, L.STOCHA_HOST, eb.buf, eb.n), eb)
    return out


def gaussian_copula(corr, n_samples, seed=42):
    """src/lib.rs:538-558 (signature :537).  Returns (n_samples, dim) float64."""
    if not isinstance(corr, np.ndarray) or corr.dtype != np.float64:
        raise TypeError("argument 'corr': expected a numpy.ndarray of float64")
    if corr.ndim != 2:
        raise TypeError(f"argument 'corr': expected a 2-dimensional array, got {corr.ndim}")
    n_samples, seed = _uint(n_samples, "n_samples"), _uint(seed, "seed")
    c = np.ascontiguousarray(corr)
    lib, eb = L.load(), L.ErrBuf()
    L.check(lib.stocha_gaussian_copula(_pd(c), c.shape[0], c.shape[1], n_samples, seed, None,
                                       L.STOCHA_HOST, eb.buf, eb.n), eb)
    out = L.host_array((n_samples, c.shape[0]))
    L.check(lib.stocha_gaussian_copula(_pd(c), c.shape[0], c.shape[1], n_samples, seed,
                                       out.ctypes.data, L.STOCHA_HOST, eb.buf, eb.n), eb)
    return out
