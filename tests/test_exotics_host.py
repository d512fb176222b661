"""Path-dependent options on the host (no device): the closed forms that
barrier_price / asian_price / lookback_price evaluate for method "auto" /
"analytical" (barrier.rs:49-268, asian.rs:44-137, lookback.rs:35-155), their
dispatch and their argument checks (src/lib.rs:1551-1848).

Pins: Haug's published barrier table for the knock-in types (The Complete
Guide to Option Pricing Formulas, 2nd ed., Table 4-13: S = 100, T = 0.5, r =
0.08, b = 0.04, sigma = 0.25, cash rebate 3 paid at expiry when the barrier
is not hit) -- the knock-out rows of that table pay the rebate at the hit
with Haug's F term, which barrier.rs:69-86 evaluates with the two powers of
H/S paired the other way round, so those rows are not this library's values;
the knock-outs are pinned through in + out = vanilla instead -- and the
reference's own hand-computed rebate values
(tests/test_exotics.py:146-157); in + out = vanilla for all 8 types (an
identity of the formulas, checked against an independent Black-Scholes with
scipy's normal cdf at the A&S erf's 1.5e-7 accuracy); a barrier far away
giving the vanilla; the Asian and lookback identities the reference tests
(deep in-the-money seasoned calls, realized payoffs at expiry, running
extrema equal to the spot changing nothing).
"""
import math

import pytest
from scipy.stats import norm

import stocha_amd as S


# the reference's normal cdf is Abramowitz & Stegun 7.1.26 (|error| <= 1.5e-7
# in erf), so a price built on it is within ~(S + K) * 1.5e-7 of the exact one
AS_TOL = 3e-5


def bs(s, k, r, q, sigma, t, call):
    d1 = (math.log(s / k) + (r - q + 0.5 * sigma * sigma) * t) / (sigma * math.sqrt(t))
    d2 = d1 - sigma * math.sqrt(t)
    if call:
        return s * math.exp(-q * t) * norm.cdf(d1) - k * math.exp(-r * t) * norm.cdf(d2)
    return k * math.exp(-r * t) * norm.cdf(-d2) - s * math.exp(-q * t) * norm.cdf(-d1)


# Haug Table 4-13 (sigma = 0.25), knock-in rows: {(type, H): prices at K =
# 90, 100, 110}, rebate 3 (H = 100 is already breached: the vanilla)
HAUG = {
    ("down-and-in", "call", 95): (7.7627, 4.0109, 2.0576),
    ("down-and-in", "call", 100): (13.8333, 7.8494, 3.9795),
    ("up-and-in", "call", 105): (14.1112, 8.4482, 4.5910),
    ("down-and-in", "put", 95): (2.9586, 6.5677, 11.9752),
    ("down-and-in", "put", 100): (2.2845, 5.9085, 11.6465),
    ("up-and-in", "put", 105): (1.4653, 3.3721, 7.0846),
}


@pytest.mark.parametrize("key", list(HAUG))
def test_barrier_haug_table(key):
    bt, ot, h = key
    for k, want in zip((90.0, 100.0, 110.0), HAUG[key]):
        got = S.barrier_price(s=100.0, k=k, r=0.08, sigma=0.25, t=0.5, barrier=float(h),
                              barrier_type=bt, option_type=ot, q=0.04, rebate=3.0,
                              method="analytical")
        assert got == pytest.approx(want, abs=6e-4), (bt, ot, h, k)


def test_barrier_rebates_match_reference_values():
    base = dict(s=100, k=100, r=0.05, sigma=0.2, t=1.0, barrier=120, barrier_type="up-and-out")
    p0 = S.barrier_price(**base)
    assert abs(S.barrier_price(**base, rebate=5, rebate_at_hit=True) - p0 - 2.0133) < 1e-3
    assert abs(S.barrier_price(**base, rebate=5, rebate_at_hit=False) - p0 - 1.9629) < 1e-3
    hi = dict(s=125, k=100, r=0.05, sigma=0.2, t=1.0, barrier=120, barrier_type="up-and-out")
    assert abs(S.barrier_price(**hi, rebate=5, rebate_at_hit=True) - 5.0) < 1e-10
    assert abs(S.barrier_price(**hi, rebate=5) - 5.0 * math.exp(-0.05)) < 1e-10
    ki = dict(hi, barrier_type="up-and-in")
    assert S.barrier_price(**ki, rebate=5) == S.barrier_price(**ki)
    assert S.barrier_price(**ki) == pytest.approx(bs(125, 100, 0.05, 0.0, 0.2, 1.0, True),
                                                  abs=AS_TOL)


@pytest.mark.parametrize("direction,h", [("up", 115.0), ("down", 85.0)])
@pytest.mark.parametrize("ot", ["call", "put"])
@pytest.mark.parametrize("k", [80.0, 100.0, 125.0])
@pytest.mark.parametrize("q", [0.0, 0.03])
def test_barrier_in_out_parity(direction, h, ot, k, q):
    kw = dict(s=100.0, k=k, r=0.05, sigma=0.3, t=0.75, barrier=h, option_type=ot, q=q,
              method="analytical")
    out = S.barrier_price(**kw, barrier_type=f"{direction}-and-out")
    inn = S.barrier_price(**kw, barrier_type=f"{direction}-and-in")
    assert out + inn == pytest.approx(bs(100.0, k, 0.05, q, 0.3, 0.75, ot == "call"),
                                      abs=AS_TOL)


def test_barrier_far_away_is_vanilla_and_bgk():
    v = S.barrier_price(s=100, k=100, r=0.05, sigma=0.2, t=1.0, barrier=1e6,
                        barrier_type="up-and-out")
    assert v == pytest.approx(bs(100, 100, 0.05, 0.0, 0.2, 1.0, True), abs=AS_TOL)
    base = dict(s=100, k=100, r=0.05, sigma=0.2, t=1.0, barrier=120, barrier_type="up-and-out")
    cont = S.barrier_price(**base)
    assert S.barrier_price(**base, n_monitoring=None) == cont
    assert S.barrier_price(**base, n_monitoring=12) > cont
    assert abs(S.barrier_price(**base, n_monitoring=100_000) - cont) < 1e-2
    assert S.barrier_price(**{**base, "barrier_type": "uo"}) == cont


def test_asian_closed_forms():
    g = S.asian_price(s=100, k=100, r=0.05, sigma=0.2, t=1.0, average_type="geometric")
    assert 0 < g < bs(100, 100, 0.05, 0.0, 0.2, 1.0, True)
    assert S.asian_price(s=100, k=100, r=0.05, sigma=0.2, t=1.0, average_type="geometric",
                         method="analytical") == g
    # asian.rs:97-103: at expiry the realized payoff, no discount
    assert S.asian_price(s=100, k=90, r=0.05, sigma=0.2, t=1.0, average_type="geometric",
                         running_avg=104.0, time_elapsed=1.0) == pytest.approx(14.0)
    # deep in the money (k* <= 0): exp(-r tau)(E[G_T] - K), independent of the vol
    a = S.asian_price(s=100, k=10, r=0.05, sigma=0.2, t=1.0, average_type="geometric",
                      running_avg=100.0, time_elapsed=0.95)
    b = S.asian_price(s=100, k=10, r=0.05, sigma=0.4, t=1.0, average_type="geometric",
                      running_avg=100.0, time_elapsed=0.95)
    assert a > 80 and abs(a - b) < 0.05
    assert S.asian_price(s=100, k=10, r=0.05, sigma=0.2, t=1.0, average_type="geometric",
                         option_type="put", running_avg=100.0, time_elapsed=0.95) == 0.0
    with pytest.raises(ValueError, match="only available for geometric average with fixed"):
        S.asian_price(s=100, k=100, r=0.05, sigma=0.2, t=1.0, method="analytical")


def test_lookback_closed_forms():
    fc = S.lookback_price(s=100, r=0.05, sigma=0.2, t=1.0)
    assert fc > bs(100, 100, 0.05, 0.0, 0.2, 1.0, True)
    assert S.lookback_price(s=100, r=0.05, sigma=0.2, t=1.0, running_min=100.0) == fc
    fp = S.lookback_price(s=100, r=0.05, sigma=0.2, t=1.0, option_type="put")
    assert S.lookback_price(s=100, r=0.05, sigma=0.2, t=1.0, option_type="put",
                            running_max=100.0) == fp
    assert S.lookback_price(s=100, r=0.05, sigma=0.2, t=1.0, option_type="put",
                            running_max=120.0) > fp
    # fixed call with K <= M: e^{-rT}(M - K) + the lookback at strike M
    fx = S.lookback_price(s=100, r=0.05, sigma=0.2, t=1.0, strike_type="fixed", k=100.0,
                          running_max=110.0)
    fm = S.lookback_price(s=100, r=0.05, sigma=0.2, t=1.0, strike_type="fixed", k=110.0,
                          running_max=110.0)
    assert fx == pytest.approx(fm + math.exp(-0.05) * 10.0, rel=1e-12)
    with pytest.raises(ValueError, match="Analytical solution not available"):
        S.lookback_price(s=100, r=0.05, sigma=0.2, t=1.0, option_type="put", running_max=90.0,
                         method="analytical")


@pytest.mark.parametrize("fn,kw,msg", [
    (S.barrier_price, dict(s=-1, k=100, r=0.05, sigma=0.2, t=1, barrier=120),
     "s must be positive"),
    (S.barrier_price, dict(s=100, k=100, r=0.05, sigma=0.2, t=1, barrier=0),
     "barrier must be positive"),
    (S.barrier_price, dict(s=100, k=100, r=0.05, sigma=0.2, t=1, barrier=120,
                           barrier_type="invalid"), "Unknown barrier_type: 'invalid'"),
    (S.barrier_price, dict(s=100, k=100, r=0.05, sigma=0.2, t=1, barrier=120,
                           option_type="digital"), "Unknown option_type: 'digital'"),
    (S.barrier_price, dict(s=100, k=100, r=0.05, sigma=0.2, t=1, barrier=120, method="x"),
     "Unknown method: 'x'. Supported: 'auto', 'analytical', 'mc'"),
    (S.asian_price, dict(s=100, k=0, r=0.05, sigma=0.2, t=1), "k must be positive"),
    (S.asian_price, dict(s=100, k=100, r=0.05, sigma=0.2, t=1, average_type="harmonic"),
     "Unknown average_type: 'harmonic'"),
    (S.asian_price, dict(s=100, k=100, r=0.05, sigma=0.2, t=1, strike_type="x"),
     "Unknown strike_type: 'x'. Supported: 'fixed', 'floating'"),
    (S.lookback_price, dict(s=100, r=0.05, sigma=0.2, t=1, strike_type="fixed"),
     "k must be positive for fixed strike lookback"),
    (S.lookback_price, dict(s=100, r=0.05, sigma=0.0, t=1), "sigma must be positive"),
])
def test_exotic_validation(fn, kw, msg):
    with pytest.raises(ValueError, match=msg):
        fn(**kw)


def test_exotic_extraction_like_pyo3():
    with pytest.raises(OverflowError):
        S.barrier_price(s=100, k=100, r=0.05, sigma=0.2, t=1, barrier=120, n_monitoring=-1)
    with pytest.raises(TypeError):
        S.barrier_price(s=100, k=100, r=0.05, sigma=0.2, t=1, barrier=120, barrier_type=3)
    with pytest.raises(TypeError):
        S.asian_price(s=100, k=100, r=0.05, sigma=0.2, t=1, running_avg="x")
