"""GPU parity of the path-dependent Monte Carlo (barrier_price /
asian_price / lookback_price with method="mc") against the CPU oracle.

FP64 with the reference's arithmetic on gbm_paths' streams; the payoff sums
run in a different order (fixed tree against path order) and the device exp
/ log may differ from glibc's in the last bit, so prices agree to 1e-10
relative.  Covers the 8 barrier types with both rebate timings and a spot
already beyond the barrier, the 4 lookback types with and without running
extrema, the Asian types (the arithmetic fixed strike through the control
variate, whose per-path pairs are reduced on the device in two passes),
seasoned fixed strikes (k* > 0 recursion and the deep in-the-money branch)
and a seasoned floating strike; and the reference's own Monte Carlo checks
(tests/test_exotics.py).
"""
import math

import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-10
R, SIG, T = 0.05, 0.25, 0.75


@pytest.mark.parametrize("bt", ["up-and-out", "up-and-in", "down-and-out", "down-and-in"])
@pytest.mark.parametrize("ot", ["call", "put"])
@pytest.mark.parametrize("rebate,at_hit", [(0.0, False), (3.0, True), (3.0, False)])
def test_barrier_mc_matches_oracle(gpu, orc, bt, ot, rebate, at_hit):
    h = 118.0 if bt.startswith("up") else 85.0
    kw = dict(s=100.0, k=100.0, r=R, sigma=SIG, t=T, barrier=h, barrier_type=bt,
              option_type=ot, q=0.01, n_paths=20_000, n_steps=60, seed=9, rebate=rebate,
              rebate_at_hit=at_hit)
    print(f"barrier mc {bt} {ot} rebate={rebate} at_hit={at_hit}", flush=True)
    dev = gpu.barrier_price(**kw, method="mc")
    ref = orc.barrier_mc(kw["s"], kw["k"], R, SIG, T, h, bt, ot, q=0.01, n_paths=20_000,
                         n_steps=60, seed=9, rebate=rebate, rebate_at_hit=at_hit)
    assert dev == pytest.approx(ref, rel=RTOL, abs=1e-12)


@pytest.mark.parametrize("bt,s", [("up-and-out", 125.0), ("up-and-in", 125.0),
                                  ("down-and-in", 70.0)])
def test_barrier_mc_already_breached(gpu, orc, bt, s):
    h = 120.0 if bt.startswith("up") else 80.0
    dev = gpu.barrier_price(s=s, k=100.0, r=R, sigma=SIG, t=T, barrier=h, barrier_type=bt,
                            n_paths=5000, n_steps=20, seed=4, method="mc", rebate=2.0,
                            rebate_at_hit=True)
    ref = orc.barrier_mc(s, 100.0, R, SIG, T, h, bt, "call", n_paths=5000, n_steps=20, seed=4,
                         rebate=2.0, rebate_at_hit=True)
    assert dev == pytest.approx(ref, rel=RTOL, abs=1e-12)


@pytest.mark.parametrize("st_,ot", [("floating", "call"), ("floating", "put"),
                                    ("fixed", "call"), ("fixed", "put")])
@pytest.mark.parametrize("season", [None, (115.0, 93.0)])
def test_lookback_mc_matches_oracle(gpu, orc, st_, ot, season):
    rmax, rmin = season if season else (None, None)
    print(f"lookback mc {st_} {ot} season={season}", flush=True)
    dev = gpu.lookback_price(s=100.0, r=R, sigma=SIG, t=T, n_steps=50, strike_type=st_,
                             option_type=ot, k=103.0, q=0.02, n_paths=15_000, seed=8,
                             method="mc", running_max=rmax, running_min=rmin)
    ref = orc.lookback_mc(100.0, R, SIG, T, 50, st_, ot, k=103.0, q=0.02, n_paths=15_000,
                          seed=8, running_max=rmax, running_min=rmin)
    assert dev == pytest.approx(ref, rel=RTOL)


@pytest.mark.parametrize("avg", ["arithmetic", "geometric"])
@pytest.mark.parametrize("strike", ["fixed", "floating"])
@pytest.mark.parametrize("ot", ["call", "put"])
def test_asian_mc_matches_oracle(gpu, orc, avg, strike, ot):
    print(f"asian mc {avg} {strike} {ot}", flush=True)
    dev = gpu.asian_price(s=100.0, k=97.0, r=R, sigma=SIG, t=T, n_steps=40, average_type=avg,
                          strike_type=strike, option_type=ot, q=0.015, n_paths=12_000, seed=6,
                          method="mc")
    ref = orc.asian_mc(100.0, 97.0, R, SIG, T, 40, avg, strike, ot, q=0.015, n_paths=12_000,
                       seed=6)
    assert dev == pytest.approx(ref, rel=RTOL, abs=1e-12)


def test_asian_mc_seasoned(gpu, orc):
    # fixed strike, k* = (T K - t1 A) / (T - t1) > 0: the unseasoned price on
    # [t1, T] at k*, scaled by (T - t1) / T (asian.rs:159-171)
    t1, a = 0.25, 101.0
    k_star = (T * 100.0 - t1 * a) / (T - t1)
    dev = gpu.asian_price(s=100.0, k=100.0, r=R, sigma=SIG, t=T, n_steps=40, n_paths=8000,
                          seed=3, method="mc", running_avg=a, time_elapsed=t1)
    ref = (T - t1) / T * orc.asian_mc(100.0, k_star, R, SIG, T - t1, 40, "arithmetic", "fixed",
                                      "call", n_paths=8000, seed=3)
    assert dev == pytest.approx(ref, rel=RTOL)
    # floating strike, seasoned (asian.rs:293-355)
    for avg in ("arithmetic", "geometric"):
        dev = gpu.asian_price(s=100.0, k=100.0, r=R, sigma=SIG, t=T, n_steps=52,
                              average_type=avg, strike_type="floating", n_paths=6000, seed=2,
                              method="mc", running_avg=97.0, time_elapsed=0.3)
        ref = orc.asian_floating_seasoned(100.0, R, SIG, T, 52, avg, "call", 97.0, 0.3,
                                          n_paths=6000, seed=2)
        assert dev == pytest.approx(ref, rel=RTOL)
    # deep in the money: no simulation
    v = gpu.asian_price(s=100.0, k=10.0, r=R, sigma=SIG, t=T, n_paths=1000, method="mc",
                        running_avg=100.0, time_elapsed=0.7)
    assert v > 80


def test_exotics_reference_mc_checks(gpu):
    # tests/test_exotics.py:51-62, 123-133, 199-206, 510-515, 619-625, 637-643
    ana = gpu.barrier_price(s=100, k=100, r=0.05, sigma=0.2, t=1.0, barrier=120,
                            barrier_type="up-and-out", method="analytical")
    mc = gpu.barrier_price(s=100, k=100, r=0.05, sigma=0.2, t=1.0, barrier=120,
                           barrier_type="up-and-out", method="mc", n_paths=200_000,
                           n_steps=1000)
    assert abs(mc - ana) / ana < 0.15
    m1 = gpu.barrier_price(s=100, k=100, r=0.05, sigma=0.2, t=1.0, barrier=120,
                           barrier_type="up-and-out", method="mc", n_paths=20_000, seed=1)
    m2 = gpu.barrier_price(s=100, k=100, r=0.05, sigma=0.2, t=1.0, barrier=120,
                           barrier_type="up-and-out", method="mc", n_paths=20_000, seed=1,
                           n_monitoring=12)
    assert m1 == m2
    base = dict(s=100, k=100, r=0.05, sigma=0.2, t=1.0, barrier=120, barrier_type="up-and-out",
                rebate=5, rebate_at_hit=True)
    ana_r = gpu.barrier_price(**base)
    mc_r = gpu.barrier_price(**base, method="mc", n_paths=100_000, n_steps=2000, seed=1)
    assert abs(mc_r - ana_r) / ana_r < 0.05
    a1 = gpu.asian_price(s=100, k=100, r=0.05, sigma=0.2, t=1.0, n_paths=10_000, seed=42)
    assert a1 == gpu.asian_price(s=100, k=100, r=0.05, sigma=0.2, t=1.0, n_paths=10_000,
                                 seed=42)
    geo = gpu.asian_price(s=100, k=100, r=0.05, sigma=0.2, t=1.0, average_type="geometric")
    assert a1 >= geo - 0.1
    lb_ana = gpu.lookback_price(s=100, r=0.05, sigma=0.2, t=1.0)
    lb_mc = gpu.lookback_price(s=100, r=0.05, sigma=0.2, t=1.0, method="mc", n_paths=50_000,
                               n_steps=252)
    assert lb_mc < lb_ana
    # no paths: the reference's 0/0
    assert math.isnan(gpu.lookback_price(s=100, r=0.05, sigma=0.2, t=1.0, method="mc",
                                         n_paths=0))
