"""Pins the oracle's bump-and-revalue Greeks (greeks.rs:227-358) on the CPU,
and the binding's argument checks (lib.rs:920-1156), which run before any
device work.

The oracle is checked against what the mathematics fixes, at the reference's
own sizes and tolerances (tests/test_greeks.py:60-128): GBM call / put delta,
gamma, vega, theta and rho against Black-Scholes; put-call delta parity; and
two structural identities that a slip in the scenario bookkeeping breaks --
the FD delta of a forward-like payoff (the call at strike 0, S_T itself) is
exactly e^{-rt} E[S_T]/s0 scaled path by path (S_T is linear in s0 under GBM,
so the central difference has no truncation error), and gamma of that payoff
vanishes; a Merton T bump must price with the T-bumped lambda*dt (checked
against a direct oracle price of the bumped spec).
"""
import math

import numpy as np
import pytest

S0, K, R, SIGMA, T = 100.0, 100.0, 0.05, 0.2, 1.0
GBM_SPEC = [S0, R, SIGMA, T]


def norm_cdf(x):
    return 0.5 * (1.0 + math.erf(x / math.sqrt(2.0)))


def norm_pdf(x):
    return math.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def _d1(s, k, r, sigma, t):
    return (math.log(s / k) + (r + 0.5 * sigma ** 2) * t) / (sigma * math.sqrt(t))


def bs_greeks(s, k, r, sigma, t):
    d1 = _d1(s, k, r, sigma, t)
    d2 = d1 - sigma * math.sqrt(t)
    return {
        "delta": norm_cdf(d1),
        "put_delta": norm_cdf(d1) - 1.0,
        "gamma": norm_pdf(d1) / (s * sigma * math.sqrt(t)),
        "vega": s * norm_pdf(d1) * math.sqrt(t),
        "theta": -s * norm_pdf(d1) * sigma / (2.0 * math.sqrt(t))
                 - r * k * math.exp(-r * t) * norm_cdf(d2),
        "rho": k * t * math.exp(-r * t) * norm_cdf(d2),
    }


ALL = ["delta", "gamma", "vega", "theta", "rho"]


def test_oracle_fd_gbm_vs_black_scholes(orc):
    # test_greeks.py:60-113: 200k x 252, seed 42, the reference's tolerances
    g = orc.greeks_fd(0, GBM_SPEC, "call", K, 200_000, 252, ALL, seed=42)
    bs = bs_greeks(S0, K, R, SIGMA, T)
    assert abs(g["delta"] - bs["delta"]) < 0.02
    assert abs(g["gamma"] - bs["gamma"]) < 0.005
    assert abs(g["vega"] - bs["vega"]) < 2.0
    assert abs(g["theta"] - bs["theta"]) < 1.0
    assert abs(g["rho"] - bs["rho"]) < 2.0
    p = orc.greeks_fd(0, GBM_SPEC, "put", K, 200_000, 252, ["delta"], seed=42)
    assert abs(p["delta"] - bs["put_delta"]) < 0.02
    # test_greeks.py:171-181: put-call delta parity
    assert abs((g["delta"] - p["delta"]) - 1.0) < 0.03


def test_oracle_fd_linear_payoff_is_exact(orc):
    # call at strike 0 pays S_T > 0; S_T is s0 times a factor that does not
    # depend on s0, so the central difference equals sum(S_T)/s0 exactly up
    # to rounding, and the second difference vanishes
    n, steps = 4000, 16
    g = orc.greeks_fd(0, GBM_SPEC, "call", 0.0, n, steps, ["delta", "gamma"], seed=5)
    term = orc.terminals(0, [S0, R, SIGMA, 0.0], T, steps, n, seed=5)
    want = math.exp(-R * T) * float(np.sum(term)) / n / S0
    assert abs(g["delta"] - want) < 1e-10 * want
    assert abs(g["gamma"]) < 1e-8


def test_oracle_fd_scenarios_price_bumped_specs(orc):
    # theta of the Merton spec = (price(t - h) - price(t)) / h with h = t/100,
    # each price from the oracle's terminals of that spec (lambda*dt moves)
    spec = [100.0, 0.05, 0.2, 3.0, -0.05, 0.1, 1.0]
    n, steps = 3000, 24
    g = orc.greeks_fd(2, spec, "call", 100.0, n, steps, ["theta"], seed=11)

    def price(t):
        term = orc.terminals(2, spec[:6], t, steps, n, seed=11)
        pay = np.maximum(term - 100.0, 0.0)
        s = 0.0
        for v in pay:
            s += v
        return math.exp(-0.05 * t) * s / n
    h = max(abs(1.0) * 0.01, 1e-8)
    want = (price(1.0 - h) - price(1.0)) / h
    assert g["theta"] == pytest.approx(want, rel=1e-12, abs=1e-12)


# ---------------------------------------------- binding checks (no device) --
def test_fd_validation_messages():
    import stocha_amd as S
    P = {"s0": S0, "r": R, "sigma": SIGMA, "t": T}
    with pytest.raises(ValueError, match="Unknown model"):
        S.greeks_fd(model="invalid", params=P, payoff="call", strike=K, n_paths=1000,
                    n_steps=10, greeks=["delta"])
    with pytest.raises(ValueError, match="Unknown greek"):
        S.greeks_fd(model="gbm", params=P, payoff="call", strike=K, n_paths=1000, n_steps=10,
                    greeks=["invalid"])
    with pytest.raises(ValueError, match="Unknown payoff"):
        S.greeks_fd(model="gbm", params=P, payoff="binary", strike=K, n_paths=1000,
                    n_steps=10, greeks=["delta"])
    with pytest.raises(ValueError, match="Missing required param: 'sigma'"):
        S.greeks_fd(model="gbm", params={"s0": S0, "r": R, "t": T}, payoff="call", strike=K,
                    n_paths=1000, n_steps=10, greeks=["delta"])
    with pytest.raises(ValueError, match="payoff must be 'call', 'put', or a callable"):
        S.greeks_fd(model="gbm", params=P, payoff=3.0, strike=K, n_paths=1000, n_steps=10,
                    greeks=["delta"])
    H = {"s0": S0, "v0": 0.04, "r": R, "kappa": 2.0, "theta": 0.04, "xi": 0.3, "rho": -0.7,
         "t": T, "scheme": "milstein"}
    with pytest.raises(ValueError, match="scheme must be 'euler' or 'qe'"):
        S.greeks_fd(model="heston", params=H, payoff="call", strike=K, n_paths=1000,
                    n_steps=10, greeks=["delta"])
    # the greek names are checked before the model (lib.rs:1076-1077)
    with pytest.raises(ValueError, match="Unknown greek"):
        S.greeks_fd(model="invalid", params=P, payoff="call", strike=K, n_paths=1000,
                    n_steps=10, greeks=["nope"])
    with pytest.raises(OverflowError):
        S.greeks_fd(model="gbm", params=P, payoff="call", strike=K, n_paths=-1, n_steps=10,
                    greeks=["delta"])
