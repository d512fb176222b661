"""Pins the oracle's path-dependent Monte Carlo (barrier.rs:286-400,
lookback.rs:158-210, asian.rs:143-355) on the CPU.

Each functional is recomputed by a different route: from the oracle's GBM
path matrix (gbm.rs, same streams and block size) with NumPy -- the first
column at or beyond the barrier, the running extrema including s0, the
arithmetic and geometric means of columns 1..n, the control-variate
estimator of asian.rs:260-280 in two passes (and, as in the reference, the
arithmetic payoff where no control variate is used, whatever the average
type) -- so a slip in the hit index,
an extremum that skips s0, an average over the wrong columns or a sign in
the estimator fails.  Identities: knock-in + knock-out = the vanilla on the
same paths, and the two barrier rebates at expiry add up to R e^{-rT}.
"""
import math

import numpy as np
import pytest

R, SIG, T = 0.05, 0.25, 0.75


def _paths(orc, s, steps, n, q=0.0, t=T, seed=7):
    return orc.gbm(s, R, SIG, t, steps, n, seed=seed, q=q)


@pytest.mark.parametrize("bt", ["up-and-out", "up-and-in", "down-and-out", "down-and-in"])
@pytest.mark.parametrize("ot", ["call", "put"])
def test_barrier_mc_from_path_matrix(orc, bt, ot):
    s, k, n, steps, q = 100.0, 100.0, 3000, 40, 0.02
    h = 115.0 if bt.startswith("up") else 88.0
    p = _paths(orc, s, steps, n, q=q)
    up, out = bt.startswith("up"), bt.endswith("out")
    beyond = p[:, 1:] >= h if up else p[:, 1:] <= h
    hit = np.where(beyond.any(axis=1), beyond.argmax(axis=1) + 1, -1)
    st = p[:, -1]
    intr = np.maximum(st - k, 0.0) if ot == "call" else np.maximum(k - st, 0.0)
    knocked = hit >= 0
    pay = np.where(knocked, 0.0, intr) if out else np.where(knocked, intr, 0.0)
    disc = math.exp(-R * T)
    want = disc * pay.sum() / n
    got = orc.barrier_mc(s, k, R, SIG, T, h, bt, ot, q=q, n_paths=n, n_steps=steps, seed=7)
    assert got == pytest.approx(want, rel=1e-12, abs=1e-12)
    if out:  # rebate paid at the hit: R e^{-r tau}, tau = hit * dt, through the discount
        reb = np.where(knocked, 4.0 * np.exp(-R * (hit * (T / steps))), 0.0)
        want_h = disc * (pay + reb / disc).sum() / n
        got_h = orc.barrier_mc(s, k, R, SIG, T, h, bt, ot, q=q, n_paths=n, n_steps=steps,
                               seed=7, rebate=4.0, rebate_at_hit=True)
        assert got_h == pytest.approx(want_h, rel=1e-12)


@pytest.mark.parametrize("ot", ["call", "put"])
def test_barrier_in_plus_out_is_vanilla(orc, ot):
    s, k, h, n, steps = 100.0, 95.0, 120.0, 4000, 50
    kw = dict(q=0.0, n_paths=n, n_steps=steps, seed=3)
    ko = orc.barrier_mc(s, k, R, SIG, T, h, "up-and-out", ot, **kw)
    ki = orc.barrier_mc(s, k, R, SIG, T, h, "up-and-in", ot, **kw)
    st = _paths(orc, s, steps, n, seed=3)[:, -1]
    intr = np.maximum(st - k, 0.0) if ot == "call" else np.maximum(k - st, 0.0)
    assert ko + ki == pytest.approx(math.exp(-R * T) * intr.sum() / n, rel=1e-12)
    rko = orc.barrier_mc(s, k, R, SIG, T, h, "up-and-out", ot, rebate=2.0, **kw) - ko
    rki = orc.barrier_mc(s, k, R, SIG, T, h, "up-and-in", ot, rebate=2.0, **kw) - ki
    assert rko + rki == pytest.approx(2.0 * math.exp(-R * T), rel=1e-10)


@pytest.mark.parametrize("st_,ot", [("floating", "call"), ("floating", "put"),
                                    ("fixed", "call"), ("fixed", "put")])
@pytest.mark.parametrize("season", [None, (112.0, 91.0)])
def test_lookback_mc_from_path_matrix(orc, st_, ot, season):
    s, k, n, steps = 100.0, 104.0, 2500, 30
    p = _paths(orc, s, steps, n, seed=11)
    mx, mn = p.max(axis=1), p.min(axis=1)
    rmax = rmin = None
    if season:
        rmax, rmin = season
        mx, mn = np.maximum(mx, rmax), np.minimum(mn, rmin)
    st = p[:, -1]
    if st_ == "floating":
        pay = st - mn if ot == "call" else mx - st
    else:
        pay = np.maximum(mx - k, 0.0) if ot == "call" else np.maximum(k - mn, 0.0)
    want = math.exp(-R * T) * pay.sum() / n
    got = orc.lookback_mc(s, R, SIG, T, steps, st_, ot, k=k, n_paths=n, seed=11,
                          running_max=rmax, running_min=rmin)
    assert got == pytest.approx(want, rel=1e-12)


@pytest.mark.parametrize("avg", ["arithmetic", "geometric"])
@pytest.mark.parametrize("strike", ["fixed", "floating"])
@pytest.mark.parametrize("ot", ["call", "put"])
def test_asian_mc_from_path_matrix(orc, avg, strike, ot):
    s, k, n, steps = 100.0, 98.0, 3000, 24
    p = _paths(orc, s, steps, n, seed=5)[:, 1:]
    st = p[:, -1]
    a_avg = p.mean(axis=1)
    g_avg = np.exp(np.log(p).mean(axis=1))

    def pay(x):
        if strike == "fixed":
            return np.maximum(x - k, 0.0) if ot == "call" else np.maximum(k - x, 0.0)
        return np.maximum(st - x, 0.0) if ot == "call" else np.maximum(x - st, 0.0)
    disc = math.exp(-R * T)
    # asian.rs:274-282: without the control variate the price is the mean of
    # the first element of each path's pair -- the arithmetic-average payoff --
    # for either average type (the reference's behaviour, kept)
    pa = pay(a_avg)
    if avg == "arithmetic" and strike == "fixed":  # control variate, asian.rs:260-280
        pg = pay(g_avg)
        ma, mg = pa.mean(), pg.mean()
        beta = ((pa - ma) * (pg - mg)).sum() / ((pg - mg) ** 2).sum()
        geo_an = _kemna_vorst(s, k, R, 0.0, SIG, T, ot == "call")
        want = max(disc * ma - beta * (disc * mg - geo_an), 0.0)
    else:
        want = max(disc * pa.mean(), 0.0)
    got = orc.asian_mc(s, k, R, SIG, T, steps, avg, strike, ot, n_paths=n, seed=5)
    assert got == pytest.approx(want, rel=1e-10)


def _kemna_vorst(s, k, r, q, sigma, t, call):
    # asian.rs:124-136 as the reference writes it (A&S normal cdf), for the CV mean
    def erf(x):
        a1, a2, a3, a4, a5, pp = (0.254829592, -0.284496736, 1.421413741, -1.453152027,
                                  1.061405429, 0.3275911)
        sg = -1.0 if x < 0 else 1.0
        x = abs(x)
        tt = 1.0 / (1.0 + pp * x)
        return sg * (1.0 - (((((a5 * tt + a4) * tt) + a3) * tt + a2) * tt + a1) * tt
                     * math.exp(-x * x))
    cdf = lambda x: 0.5 * (1.0 + erf(x * math.sqrt(0.5)))
    sg = sigma / math.sqrt(3.0)
    mug = 0.5 * (r - q - 0.5 * sigma * sigma) + 0.25 * sg * sg
    d1 = (math.log(s / k) + (mug + 0.5 * sg * sg) * t) / (sg * math.sqrt(t))
    d2 = d1 - sg * math.sqrt(t)
    if call:
        return max(s * math.exp((mug - r) * t) * cdf(d1) - k * math.exp(-r * t) * cdf(d2), 0.0)
    return max(k * math.exp(-r * t) * cdf(-d2) - s * math.exp((mug - r) * t) * cdf(-d1), 0.0)


@pytest.mark.parametrize("avg", ["arithmetic", "geometric"])
def test_asian_floating_seasoned_from_path_matrix(orc, avg):
    s, n, steps, t1, a_spent = 100.0, 2000, 52, 0.3, 97.0
    rem = T - t1
    n_rem = max(round(steps * rem / T), 1)
    p = _paths(orc, s, n_rem, n, t=rem, seed=2)[:, 1:]
    st = p[:, -1]
    ws, wr = t1 / T, rem / T
    if avg == "arithmetic":
        avg_t = ws * a_spent + wr * p.mean(axis=1)
    else:
        avg_t = np.exp(math.log(a_spent) * ws + np.log(p).mean(axis=1) * wr)
    want = math.exp(-R * rem) * np.maximum(st - avg_t, 0.0).sum() / n
    got = orc.asian_floating_seasoned(s, R, SIG, T, steps, avg, "call", a_spent, t1,
                                      n_paths=n, seed=2)
    assert got == pytest.approx(want, rel=1e-11)
