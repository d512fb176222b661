"""stocha_amd -- the Monte Carlo hot path of stocha, on MI355X (gfx950).

A drop-in for these functions of the reference's Python API
(python/stocha/__init__.py), same signatures, defaults, return types and
errors:

    gbm                    (__init__.py:180-231)
    heston                 (__init__.py:276-324; schemes "euler" and "qe")
    merton_jump_diffusion  (__init__.py:327-373)
    multi_gbm              (__init__.py:574-627)
    lsmc_american_option   (__init__.py:630-671)
    greeks_pathwise        (__init__.py:725-769)
    greeks_fd              (__init__.py:674-722)
    barrier_price          (__init__.py:983-1060)
    asian_price            (__init__.py:1063-1127)
    lookback_price         (__init__.py:1130-1190)
    RNG                    (the class: uniform / standard_normal / normal on one
                            sequential stream, drawn on the GPU; JSON state)
    hull_white             (__init__.py:451-492)
    gaussian_copula        (__init__.py:400-423)
    student_t_copula       (__init__.py:426-448)

They run on the GPU through the C ABI of libstocha_gpu.so (hand-written HIP
kernels); there is no CPU fallback.  Forms beside the drop-in API:

    european_price         in-register European calls, FP64 or FP32 stepping
    lsmc_american_option_ex  LSMC with FP32 paths
    basket_price           multi-asset basket call with pathwise adjoint Greeks
                           (stocha_amd.autograd wraps it as a torch.autograd.Function)

The reference's other functions (sobol, halton, SABR, Heston COS /
calibration, SSVI, var_cvar) are not part of this package: calling them
raises NotImplementedError.
"""
from __future__ import annotations

from typing import Literal, Sequence

import ctypes as _C

import numpy as np

from . import _lib
from . import _stocha as _S
from ._stocha import RNG, __version__

__all__ = [
    "gbm",
    "heston",
    "merton_jump_diffusion",
    "multi_gbm",
    "lsmc_american_option",
    "greeks_pathwise",
    "greeks_fd",
    "RNG",
    "hull_white",
    "gaussian_copula",
    "student_t_copula",
    "barrier_price",
    "asian_price",
    "lookback_price",
    "european_price",
    "lsmc_american_option_ex",
    "basket_price",
    "__version__",
]


def gbm(s0: float, mu: float, sigma: float, t: float, steps: int, n_paths: int,
        seed: int = 42, antithetic: bool = False, q: float = 0.0) -> np.ndarray:
    """Simulate GBM paths (reference: python/stocha/__init__.py:180-231).

    Returns a (n_paths, steps + 1) float64 array; column 0 is s0.  Path i
    uses the reference's block-split PCG64DXSM stream i, so the values equal
    the reference's for the same seed.
    """
    return _S.gbm(s0=s0, mu=mu, sigma=sigma, t=t, steps=steps, n_paths=n_paths, seed=seed,
                  antithetic=antithetic, q=q)


def heston(s0: float, v0: float, mu: float, kappa: float, theta: float, xi: float, rho: float,
           t: float, steps: int, n_paths: int, seed: int = 42,
           scheme: Literal["euler", "qe"] = "euler") -> np.ndarray:
    """Simulate Heston paths, full-truncation Euler or Andersen's QE scheme
    (reference: python/stocha/__init__.py:276-324)."""
    return _S.heston(s0=s0, v0=v0, mu=mu, kappa=kappa, theta=theta, xi=xi, rho=rho, t=t,
                     steps=steps, n_paths=n_paths, seed=seed, scheme=scheme)


def merton_jump_diffusion(s0: float, mu: float, sigma: float, lambda_: float, mu_j: float,
                          sigma_j: float, t: float, steps: int, n_paths: int,
                          seed: int = 42) -> np.ndarray:
    """Simulate Merton jump-diffusion paths (reference: python/stocha/__init__.py:327-373)."""
    return _S.merton_jump_diffusion(s0=s0, mu=mu, sigma=sigma, lambda_=lambda_, mu_j=mu_j,
                                    sigma_j=sigma_j, t=t, steps=steps, n_paths=n_paths,
                                    seed=seed)


def multi_gbm(s0: list[float], mu: list[float], sigma: list[float], corr: np.ndarray, t: float,
              steps: int, n_paths: int, seed: int = 42, antithetic: bool = False) -> np.ndarray:
    """Simulate correlated multi-asset GBM (reference: python/stocha/__init__.py:574-627).

    Returns (n_paths, steps + 1, n_assets) float64.
    """
    return _S.multi_gbm(s0=list(s0), mu=list(mu), sigma=list(sigma),
                        corr=np.asarray(corr, dtype=np.float64), t=t, steps=steps,
                        n_paths=n_paths, seed=seed, antithetic=antithetic)


def lsmc_american_option(s0: float, k: float, r: float, sigma: float, t: float, steps: int,
                         n_paths: int, is_put: bool = True, poly_degree: int = 3,
                         seed: int = 42) -> tuple[float, float]:
    """Longstaff-Schwartz American option, FP64 paths and solve (reference:
    python/stocha/__init__.py:630-671).  Returns (price, std_err)."""
    return _S.lsmc_american_option(s0=s0, k=k, r=r, sigma=sigma, t=t, steps=steps,
                                   n_paths=n_paths, is_put=is_put, poly_degree=poly_degree,
                                   seed=seed)


def greeks_pathwise(s0: float, r: float, sigma: float, t: float, strike: float, is_call: bool,
                    n_paths: int, n_steps: int, greeks: list[str],
                    seed: int = 42) -> dict[str, float]:
    """Pathwise (IPA) delta and vega for GBM (reference: python/stocha/__init__.py:725-769)."""
    return _S.greeks_pathwise(s0=s0, r=r, sigma=sigma, t=t, strike=strike, is_call=is_call,
                              n_paths=n_paths, n_steps=n_steps, greeks=greeks, seed=seed)


def greeks_fd(model: str, params: dict, payoff, strike: float, n_paths: int, n_steps: int,
              greeks: list[str], seed: int = 42, bump_size: float = 0.01) -> dict[str, float]:
    """Bump-and-revalue Greeks with common random numbers (reference:
    python/stocha/__init__.py:674-722).

    model "gbm" / "heston" / "merton"; params as the reference's (GBM: s0, r,
    sigma, t; Heston: s0, v0, r, kappa, theta, xi, rho, t [, scheme]; Merton:
    s0, r, sigma, lambda_, mu_j, sigma_j, t); payoff "call", "put" or a callable
    f(terminals) -> values.  Built-in payoffs price every bump scenario in
    one launch on shared draws; a callable gets each scenario's FP64 terminal
    prices from the GPU.
    """
    return _S.greeks_fd(model=model, params=params, payoff=payoff, strike=strike,
                        n_paths=n_paths, n_steps=n_steps, greeks=greeks, seed=seed,
                        bump_size=bump_size)


def barrier_price(s: float, k: float, r: float, sigma: float, t: float, barrier: float,
                  barrier_type: str = "up-and-out", option_type: str = "call", q: float = 0.0,
                  n_paths: int = 100_000, n_steps: int = 252, seed: int = 42,
                  method: str = "auto", n_monitoring: int | None = None, rebate: float = 0.0,
                  rebate_at_hit: bool = False) -> float:
    """Barrier option, analytical (Reiner-Rubinstein, BGK discrete-monitoring
    shift, rebates) or Monte Carlo on the GPU (reference:
    python/stocha/__init__.py:983-1060)."""
    return _S.barrier_price(s=s, k=k, r=r, sigma=sigma, t=t, barrier=barrier,
                            barrier_type=barrier_type, option_type=option_type, q=q,
                            n_paths=n_paths, n_steps=n_steps, seed=seed, method=method,
                            n_monitoring=n_monitoring, rebate=rebate, rebate_at_hit=rebate_at_hit)


def asian_price(s: float, k: float, r: float, sigma: float, t: float, n_steps: int = 252,
                average_type: str = "arithmetic", strike_type: str = "fixed",
                option_type: str = "call", q: float = 0.0, n_paths: int = 100_000,
                seed: int = 42, method: str = "auto", running_avg: float | None = None,
                time_elapsed: float | None = None) -> float:
    """Asian option: Kemna-Vorst for the geometric fixed strike, Monte Carlo on
    the GPU otherwise (geometric control variate for the arithmetic fixed
    strike) (reference: python/stocha/__init__.py:1063-1127)."""
    return _S.asian_price(s=s, k=k, r=r, sigma=sigma, t=t, n_steps=n_steps,
                          average_type=average_type, strike_type=strike_type,
                          option_type=option_type, q=q, n_paths=n_paths, seed=seed, method=method,
                          running_avg=running_avg, time_elapsed=time_elapsed)


def lookback_price(s: float, r: float, sigma: float, t: float, n_steps: int = 252,
                   strike_type: str = "floating", option_type: str = "call", k: float = 0.0,
                   q: float = 0.0, n_paths: int = 100_000, seed: int = 42, method: str = "auto",
                   running_max: float | None = None, running_min: float | None = None) -> float:
    """Lookback option: Goldman-Sosin-Gatto / Conze-Viswanathan closed forms or
    Monte Carlo on the GPU (reference: python/stocha/__init__.py:1130-1190)."""
    return _S.lookback_price(s=s, r=r, sigma=sigma, t=t, n_steps=n_steps,
                             strike_type=strike_type, option_type=option_type, k=k, q=q,
                             n_paths=n_paths, seed=seed, method=method, running_max=running_max,
                             running_min=running_min)


def hull_white(r0: float, a: float, theta: float, sigma: float, t: float, steps: int,
               n_paths: int, seed: int = 42) -> np.ndarray:
    """Hull-White one-factor short-rate paths, exact transition; (n_paths,
    steps + 1), column 0 = r0 (reference: python/stocha/__init__.py:451-492)."""
    return _S.hull_white(r0=r0, a=a, theta=theta, sigma=sigma, t=t, steps=steps,
                         n_paths=n_paths, seed=seed)


def gaussian_copula(corr: np.ndarray, n_samples: int, seed: int = 42) -> np.ndarray:
    """Gaussian copula samples Phi(L z) from one PCG64DXSM stream; (n_samples,
    dim) in (0, 1) (reference: python/stocha/__init__.py:400-423)."""
    return _S.gaussian_copula(corr=np.asarray(corr, dtype=np.float64), n_samples=n_samples,
                              seed=seed)


def student_t_copula(corr: np.ndarray, nu: float, n_samples: int, seed: int = 42) -> np.ndarray:
    """Student-t copula samples t_nu((L z) sqrt(nu / chi2)) from one PCG64DXSM
    stream; (n_samples, dim) in (0, 1) (reference: python/stocha/__init__.py:426-448)."""
    return _S.student_t_copula(corr=np.asarray(corr, dtype=np.float64), nu=nu,
                               n_samples=n_samples, seed=seed)


# ------------------------------------------------------------------ extras --
_MODELS = {"gbm": (_lib.MODEL_GBM, ("s0", "mu", "sigma")),
           "heston": (_lib.MODEL_HESTON, ("s0", "v0", "mu", "kappa", "theta", "xi", "rho")),
           "merton": (_lib.MODEL_MERTON, ("s0", "mu", "sigma", "lambda_", "mu_j", "sigma_j"))}


def european_price(model: str, params: dict, t: float, steps: int, n_paths: int,
                   strikes: Sequence[float], r: float, seed: int = 42,
                   precision: int = 64, schedule: str = "feed") -> dict:
    """European calls at up to 8 strikes, priced in-register on the GPU.

    model: "gbm" | "heston" | "merton"; params holds the model's parameters
    under the names of the corresponding path function.  precision 64 steps
    with the reference's FP64 arithmetic; 32 steps the log-price in FP32 with
    FP64 payoff sums.  Both use the reference's block-split streams, so they
    price the same paths.  schedule "feed" (per-lane state machine, default)
    and "loop" (the reference's per-normal loop) give identical results.
    Returns {"price": array, "std_err": array, "sum": array, "sumsq": array}
    (one entry per strike).
    """
    if schedule not in ("feed", "loop"):
        raise ValueError("schedule must be 'feed' or 'loop'")
    sched = _lib.STOCHA_SCHEDULE_FEED if schedule == "feed" else _lib.STOCHA_SCHEDULE_LOOP
    if model not in _MODELS:
        raise ValueError(f"Unknown model: '{model}'. Supported: gbm, heston, merton")
    mid, names = _MODELS[model]
    missing = [n for n in names if n not in params]
    if missing:
        raise ValueError(f"missing parameter(s) for {model}: {', '.join(missing)}")
    p = np.array([_S._f64(params[n], n) for n in names], dtype=np.float64)
    k = np.ascontiguousarray([_S._f64(v, "strikes") for v in strikes], dtype=np.float64)
    steps, n_paths, seed = _S._uint(steps, "steps"), _S._uint(n_paths, "n_paths"), \
        _S._uint(seed, "seed")
    lib, eb = _lib.load(), _lib.errbuf()
    ns = len(k)
    m = max(ns, 1)
    out = np.zeros(4 * m)  # price | std_err | (sum, sumsq) pairs
    a = out.ctypes.data  # raw addresses: cheaper per call than ctypes pointer objects
    rc = lib.stocha_european_fused(mid, p.ctypes.data, len(p), _S._f64(t, "t"), steps, n_paths,
                                   seed, int(precision), sched, k.ctypes.data, ns,
                                   _S._f64(r, "r"), a, a + 8 * m, a + 16 * m, eb.buf, eb.n)
    _lib.check(rc, eb)
    return {"price": out[:ns], "std_err": out[m:m + ns], "sum": out[2 * m:2 * m + 2 * ns:2],
            "sumsq": out[2 * m + 1:2 * m + 2 * ns:2]}


def lsmc_american_option_ex(s0: float, k: float, r: float, sigma: float, t: float, steps: int,
                            n_paths: int, is_put: bool = True, poly_degree: int = 3,
                            seed: int = 42, precision: int = 32) -> tuple[float, float]:
    """lsmc_american_option with a choice of path precision (32: FP32 paths,
    FP64 moment sums and solve)."""
    return _S._lsmc(s0, k, r, sigma, t, steps, n_paths, is_put, poly_degree, seed,
                    int(precision))


def last_kernel_ms() -> tuple[float, int]:
    """Duration of the dominant kernel(s) of the last call, when timing is on."""
    lib = _lib.load()
    ms, n = _C.c_double(), _C.c_uint64()
    lib.stocha_last_kernel_ms(_C.byref(ms), _C.byref(n))
    return ms.value, n.value


def enable_kernel_timing(on: bool = True) -> None:
    if _lib.load().stocha_timing_enable(int(on)) != 0:
        raise RuntimeError("could not create HIP events")


def device_count() -> int:
    return int(_lib.load().stocha_device_count())


def _not_here(name):
    def f(*_a, **_k):
        raise NotImplementedError(
            f"stocha.{name} is not part of the GPU Monte Carlo hot path implemented by "
            "stocha_amd")
    f.__name__ = name
    return f


for _name in ("sobol", "halton", "var_cvar", "sabr_implied_vol", "sabr_calibrate", "heston_price",
              "heston_calibrate", "ssvi_calibrate", "ssvi_implied_vol", "ssvi_local_vol"):
    globals()[_name] = _not_here(_name)
del _name


def stream_u64(seed: int, block: int, path0: int, n_paths: int, n_draws: int) -> np.ndarray:
    """Raw PCG64DXSM draws of block-split streams path0.. (device-generated)."""
    out = np.zeros((n_paths, n_draws), dtype=np.uint64)
    lib, eb = _lib.load(), _lib.ErrBuf()
    rc = lib.stocha_stream_u64(seed, block, path0, n_paths, n_draws,
                               out.ctypes.data_as(_C.POINTER(_C.c_uint64)), eb.buf, eb.n)
    _lib.check(rc, eb)
    return out


def stream_normals(seed: int, block: int, path0: int, n_paths: int, n_draws: int,
                   precision: int = 64) -> tuple[np.ndarray, np.ndarray]:
    """Ziggurat normals of block-split streams (device-generated) and the
    number of u64 draws each path consumed."""
    out = np.zeros((n_paths, n_draws), dtype=np.float64)
    used = np.zeros(n_paths, dtype=np.uint64)
    lib, eb = _lib.load(), _lib.ErrBuf()
    rc = lib.stocha_stream_normals(seed, block, path0, n_paths, n_draws, int(precision),
                                   out.ctypes.data_as(_C.POINTER(_C.c_double)),
                                   used.ctypes.data_as(_C.POINTER(_C.c_uint64)), eb.buf, eb.n)
    _lib.check(rc, eb)
    return out, used


def stream_normals_queue(seed: int, block: int, path0: int, n_paths: int, n_draws: int,
                         precision: int = 64) -> np.ndarray:
    """The normals of stream_normals drawn on the steppers' queue schedule
    (same values, another schedule)."""
    out = np.zeros((n_paths, n_draws), dtype=np.float64)
    lib, eb = _lib.load(), _lib.errbuf()
    rc = lib.stocha_stream_normals_queue(seed, block, path0, n_paths, n_draws, int(precision),
                                         out.ctypes.data_as(_C.POINTER(_C.c_double)), eb.buf,
                                         eb.n)
    _lib.check(rc, eb)
    return out


def basket_price(s0: Sequence[float], mu: Sequence[float], sigma: Sequence[float],
                 corr: np.ndarray, weights: Sequence[float], strike: float, r: float, t: float,
                 steps: int, n_paths: int, seed: int = 42, precision: int = 32,
                 greeks: bool = False) -> dict:
    """Basket call on correlated GBM, priced in-register on the GPU.

    Paths use multi_gbm's streams and draw order (so they are multi_gbm's
    paths); the payoff is max(sum_a w_a S_T,a - strike, 0), discounted at r.
    greeks=True also returns the pathwise adjoint of the price with respect
    to every asset's s0 ("delta"), mu ("dmu") and sigma ("vega").
    """
    s0 = _S._vec_f64(s0, "s0")
    mu = _S._vec_f64(mu, "mu")
    sigma = _S._vec_f64(sigma, "sigma")
    w = _S._vec_f64(weights, "weights")
    d = len(s0)
    if len(mu) != d or len(sigma) != d or len(w) != d:
        raise ValueError("s0, mu, sigma, weights must have the same length")
    corr = np.ascontiguousarray(corr, dtype=np.float64)
    if corr.shape != (d, d):
        raise ValueError(f"corr must be ({d}, {d}), got {corr.shape}")
    steps, n_paths, seed = _S._uint(steps, "steps"), _S._uint(n_paths, "n_paths"), \
        _S._uint(seed, "seed")
    lib, eb = _lib.load(), _lib.ErrBuf()
    price, se = _C.c_double(), _C.c_double()
    gd, gm, gv = np.zeros(max(d, 1)), np.zeros(max(d, 1)), np.zeros(max(d, 1))
    rc = lib.stocha_basket(_S._pd(s0), _S._pd(mu), _S._pd(sigma), _S._pd(corr), _S._pd(w), d,
                           _S._f64(strike, "strike"), _S._f64(r, "r"), _S._f64(t, "t"), steps,
                           n_paths, seed, int(precision), int(bool(greeks)), _C.byref(price),
                           _C.byref(se), _S._pd(gd), _S._pd(gm), _S._pd(gv), eb.buf, eb.n)
    _lib.check(rc, eb)
    out = {"price": price.value, "std_err": se.value}
    if greeks:
        out.update(delta=gd[:d], dmu=gm[:d], vega=gv[:d])
    return out
