"""GPU parity of greeks_fd (greeks.rs:227-358) against the CPU oracle.

The device prices every bump scenario in FP64 with the reference's
arithmetic; only the payoff sum's order differs (a fixed tree against the
reference's left-to-right sum), so prices agree to ~1e-15 relative and a
Greek, a difference quotient with bump h, to ~1e-15 * price / h: the bar is
1e-9 relative + 1e-9 absolute.  Terminal prices (the callable-payoff form)
agree to 1e-12 relative (the device exp and glibc's differ in the last bit
now and then).  Also the reference's own checks (tests/test_greeks.py):
Black-Scholes at 200k x 252, custom callable == built-in payoff, common
random numbers reproduce, Heston / Merton deltas in (0, 1).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ALL = ["delta", "gamma", "vega", "theta", "rho"]
GBM_P = {"s0": 100.0, "r": 0.05, "sigma": 0.2, "t": 1.0}
HESTON_P = {"s0": 100.0, "v0": 0.04, "r": 0.05, "kappa": 2.0, "theta": 0.04, "xi": 0.3,
            "rho": -0.7, "t": 1.0}
MERTON_P = {"s0": 100.0, "r": 0.05, "sigma": 0.2, "lambda_": 1.0, "mu_j": -0.05,
            "sigma_j": 0.1, "t": 1.0}
SPEC = {"gbm": (0, ["s0", "r", "sigma", "t"]),
        "heston": (1, ["s0", "v0", "r", "kappa", "theta", "xi", "rho", "t"]),
        "merton": (2, ["s0", "r", "sigma", "lambda_", "mu_j", "sigma_j", "t"])}
PARAMS = {"gbm": GBM_P, "heston": HESTON_P, "merton": MERTON_P}


def _oracle(orc, model, params, payoff, strike, n, steps, greeks, seed, bump=0.01):
    mid, keys = SPEC[model]
    return orc.greeks_fd(mid, [params[k] for k in keys], payoff, strike, n, steps, greeks,
                         seed=seed, bump_size=bump)


@pytest.mark.parametrize("model", ["gbm", "heston", "merton"])
@pytest.mark.parametrize("payoff,strike", [("call", 100.0), ("put", 95.0)])
def test_fd_matches_oracle(gpu, orc, model, payoff, strike):
    n, steps, seed = 6000, 40, 7
    print(f"greeks_fd {model} {payoff} n={n} steps={steps}", flush=True)
    dev = gpu.greeks_fd(model=model, params=PARAMS[model], payoff=payoff, strike=strike,
                        n_paths=n, n_steps=steps, greeks=ALL, seed=seed)
    ref = _oracle(orc, model, PARAMS[model], payoff, strike, n, steps, ALL, seed)
    for g in ALL:
        assert dev[g] == pytest.approx(ref[g], rel=1e-9, abs=1e-9), g


@pytest.mark.parametrize("model", ["gbm", "heston", "merton"])
def test_fd_subsets_bumps_and_duplicates(gpu, orc, model):
    # single greeks (one or two scenarios per launch), a repeated name, a
    # larger bump, the theta-only Merton launch of its own
    n, steps = 3000, 25
    for greeks, bump in ((["theta"], 0.01), (["rho", "rho"], 0.05), (["gamma"], 0.002),
                         (["vega", "delta"], 0.03)):
        print(f"greeks_fd {model} {greeks} bump={bump}", flush=True)
        dev = gpu.greeks_fd(model=model, params=PARAMS[model], payoff="call", strike=100.0,
                            n_paths=n, n_steps=steps, greeks=greeks, seed=3, bump_size=bump)
        ref = _oracle(orc, model, PARAMS[model], "call", 100.0, n, steps, greeks, 3, bump)
        for g in greeks:
            assert dev[g] == pytest.approx(ref[g], rel=1e-9, abs=1e-9), (g, bump)


@pytest.mark.parametrize("model", ["gbm", "heston", "merton"])
def test_terminal_prices_match_oracle(gpu, orc, model):
    from stocha_amd import _stocha as S
    mid, keys = SPEC[model]
    vals = np.array([PARAMS[model][k] for k in keys])
    n, steps = 700, 33
    dev = S._terminal_prices(mid, vals, "euler" if model == "heston" else None, steps, n, 42)
    if model == "gbm":
        p, t = [vals[0], vals[1], vals[2], 0.0], vals[3]
    elif model == "heston":
        p, t = list(vals[:7]), vals[7]
    else:
        p, t = list(vals[:6]), vals[6]
    ref = orc.terminals(mid, p, t, steps, n, seed=42)
    np.testing.assert_allclose(dev, ref, rtol=1e-12, atol=0)


def test_fd_vs_black_scholes_reference_sizes(gpu):
    # tests/test_greeks.py:60-128 (200k x 252, seed 42, same tolerances)
    s, k, r, sig, t = 100.0, 100.0, 0.05, 0.2, 1.0
    g = gpu.greeks_fd(model="gbm", params=GBM_P, payoff="call", strike=k, n_paths=200_000,
                      n_steps=252, greeks=ALL, seed=42)
    d1 = (math.log(s / k) + (r + 0.5 * sig * sig) * t) / (sig * math.sqrt(t))
    d2 = d1 - sig * math.sqrt(t)
    cdf = lambda x: 0.5 * (1 + math.erf(x / math.sqrt(2)))
    pdf = lambda x: math.exp(-0.5 * x * x) / math.sqrt(2 * math.pi)
    assert abs(g["delta"] - cdf(d1)) < 0.02
    assert abs(g["gamma"] - pdf(d1) / (s * sig * math.sqrt(t))) < 0.005
    assert abs(g["vega"] - s * pdf(d1) * math.sqrt(t)) < 2.0
    theta = -s * pdf(d1) * sig / (2 * math.sqrt(t)) - r * k * math.exp(-r * t) * cdf(d2)
    assert abs(g["theta"] - theta) < 1.0
    assert abs(g["rho"] - k * t * math.exp(-r * t) * cdf(d2)) < 2.0
    p = gpu.greeks_fd(model="gbm", params=GBM_P, payoff="put", strike=k, n_paths=200_000,
                      n_steps=252, greeks=["delta"], seed=42)
    assert abs((g["delta"] - p["delta"]) - 1.0) < 0.03
    pw = gpu.greeks_pathwise(s0=s, r=r, sigma=sig, t=t, strike=k, is_call=True,
                             n_paths=200_000, n_steps=252, greeks=["delta"], seed=42)
    assert abs(g["delta"] - pw["delta"]) < 0.03


def test_fd_custom_payoff_matches_builtin(gpu):
    # test_greeks.py:199-211
    builtin = gpu.greeks_fd(model="gbm", params=GBM_P, payoff="call", strike=100.0,
                            n_paths=50_000, n_steps=50, greeks=["delta", "gamma"], seed=42)
    custom = gpu.greeks_fd(model="gbm", params=GBM_P,
                           payoff=lambda x: np.maximum(x - 100.0, 0.0), strike=100.0,
                           n_paths=50_000, n_steps=50, greeks=["delta", "gamma"], seed=42)
    assert abs(builtin["delta"] - custom["delta"]) < 1e-10
    assert abs(builtin["gamma"] - custom["gamma"]) < 1e-9


def test_fd_crn_reproducible_and_model_smoke(gpu):
    a = gpu.greeks_fd(model="gbm", params=GBM_P, payoff="call", strike=100.0, n_paths=10_000,
                      n_steps=50, greeks=["delta", "vega"], seed=123)
    b = gpu.greeks_fd(model="gbm", params=GBM_P, payoff="call", strike=100.0, n_paths=10_000,
                      n_steps=50, greeks=["delta", "vega"], seed=123)
    assert a == b
    for model in ("heston", "merton"):
        r = gpu.greeks_fd(model=model, params=PARAMS[model], payoff="call", strike=100.0,
                          n_paths=50_000, n_steps=100, greeks=["delta"], seed=42)
        assert 0 < r["delta"] < 1


def test_fd_empty_inputs(gpu):
    # no paths: every price is 0/0 (NaN), as in the reference; no greeks: {}
    r = gpu.greeks_fd(model="gbm", params=GBM_P, payoff="call", strike=100.0, n_paths=0,
                      n_steps=10, greeks=["delta"])
    assert math.isnan(r["delta"])
    assert gpu.greeks_fd(model="gbm", params=GBM_P, payoff="call", strike=100.0, n_paths=100,
                         n_steps=10, greeks=[]) == {}
