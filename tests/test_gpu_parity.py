"""GPU parity: the HIP kernels against the CPU oracle on the same seeds.

Bars (fixed by the reference's arithmetic):
  * raw u64 streams: bit-exact;
  * FP64 normals: bit-exact, except tail values (|x| > R), which go through
    log() and may differ by <= 2 ulp between the device libm and glibc;
  * FP64 paths, prices, Greeks: relative error <= 1e-12;
  * FP32 forms: same u64 consumption as FP64; prices within 1e-4 relative.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL64 = 1e-12
R_ZIG = 3.654152885361008796


def _ulps(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.spacing(np.maximum(np.abs(a), np.abs(b)))


# ------------------------------------------------------------------ streams --
@pytest.mark.parametrize("seed", [0, 42, 2**64 - 1])
@pytest.mark.parametrize("block", [2552, 7656, 10208, 34304])
def test_raw_u64_bit_exact(gpu, orc, seed, block):
    paths = [0, 1, 2, 255, 256, 4097, 99_999, 1_048_575]
    for p in paths:
        dev = gpu.stream_u64(seed, block, p, 1, 64)[0]
        ref = orc.raw_u64(seed, block, p, 64)
        np.testing.assert_array_equal(dev, ref, err_msg=f"path {p}")


def test_raw_u64_many_paths(gpu, orc):
    dev = gpu.stream_u64(42, 7656, 1000, 300, 8)
    for i in [0, 1, 63, 64, 150, 299]:
        np.testing.assert_array_equal(dev[i], orc.raw_u64(42, 7656, 1000 + i, 8))


@pytest.mark.parametrize("prec", [64, 32])
def test_normals_match_oracle(gpu, orc, prec):
    n_paths, n = 512, 400
    dev, used = gpu.stream_normals(42, 2552, 0, n_paths, n, precision=prec)
    tails = 0
    for i in range(0, n_paths, 7):
        ref, ref_used = orc.normals(42, 2552, i, n)
        assert int(used[i]) == ref_used, f"path {i}: consumption differs"
        if prec == 64:
            tail = np.abs(ref) > R_ZIG
            tails += int(tail.sum())
            np.testing.assert_array_equal(dev[i][~tail], ref[~tail])
            assert np.all(_ulps(dev[i][tail], ref[tail]) <= 2)
        else:
            # fast path: top 32 of the 53 mantissa bits -> absolute error
            # <= 2^-32 * X[l] (< 1e-9), plus FP32 rounding (rel 2^-23)
            np.testing.assert_allclose(dev[i], ref, rtol=2.4e-7, atol=1e-9)


# -------------------------------------------------------------------- paths --
@pytest.mark.parametrize("n_paths,steps,anti", [(1000, 37, False), (2048, 252, False),
                                                (999, 50, True), (1, 1, False),
                                                (64, 9, False), (65, 8, True),
                                                (193, 16, False), (257, 7, True)])
def test_gbm_paths(gpu, orc, n_paths, steps, anti):
    # each wavefront writes its own rows: 64 / 65 / 193 / 257 rows leave whole
    # wavefronts without rows and one ragged wavefront; 8 / 16 steps end on a
    # full 8-column tile, 7 / 9 on a partial one
    kw = dict(s0=100.0, mu=0.05, sigma=0.2, t=1.0, steps=steps, n_paths=n_paths, seed=42,
              antithetic=anti, q=0.01)
    dev = gpu.gbm(**kw)
    ref = orc.gbm(**kw)
    assert dev.shape == (n_paths, steps + 1) and dev.dtype == np.float64
    np.testing.assert_array_equal(dev[:, 0], 100.0)
    np.testing.assert_allclose(dev, ref, rtol=RTOL64, atol=0)


@pytest.mark.parametrize("n_paths,steps", [(2000, 252), (300, 17), (64, 9), (130, 16)])
def test_heston_paths(gpu, orc, n_paths, steps):
    kw = dict(s0=100.0, v0=0.04, mu=0.05, kappa=2.0, theta=0.04, xi=0.3, rho=-0.7, t=1.0,
              steps=steps, n_paths=n_paths, seed=42)
    dev = gpu.heston(**kw)
    ref = orc.heston(**kw)
    np.testing.assert_allclose(dev, ref, rtol=RTOL64, atol=0)


def test_heston_feller_violated_paths(gpu, orc):
    # xi large: v hits zero often, exercising the full-truncation branch
    kw = dict(s0=100.0, v0=0.01, mu=0.0, kappa=0.5, theta=0.02, xi=1.5, rho=-0.9, t=2.0,
              steps=100, n_paths=1500, seed=7)
    np.testing.assert_allclose(gpu.heston(**kw), orc.heston(**kw), rtol=RTOL64, atol=0)


@pytest.mark.parametrize("lam", [1.0, 5.0, 0.0])
def test_merton_paths(gpu, orc, lam):
    kw = dict(s0=100.0, mu=0.05, sigma=0.2, lambda_=lam, mu_j=-0.05, sigma_j=0.1, t=1.0,
              steps=252, n_paths=1500, seed=42)
    np.testing.assert_allclose(gpu.merton_jump_diffusion(**kw), orc.merton(**kw), rtol=RTOL64,
                               atol=0)


def test_multi_gbm_paths(gpu, orc):
    corr = np.array([[1.0, 0.6, 0.3], [0.6, 1.0, 0.5], [0.3, 0.5, 1.0]])
    kw = dict(s0=[100.0, 50.0, 200.0], mu=[0.05, 0.08, 0.03], sigma=[0.2, 0.3, 0.15], corr=corr,
              t=1.0, steps=50, n_paths=700, seed=3)
    for anti in (False, True):
        dev = gpu.multi_gbm(**kw, antithetic=anti)
        ref = orc.multi_gbm(**kw, antithetic=anti)
        assert dev.shape == (700, 51, 3)
        np.testing.assert_allclose(dev, ref, rtol=RTOL64, atol=0)


def _corr(d, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((d, d))
    c = a @ a.T / d + 0.1 * np.eye(d)
    dg = np.sqrt(np.diag(c))
    return c / np.outer(dg, dg), rng


@pytest.mark.parametrize("d,n,steps,anti", [(2, 190, 41, True), (4, 65, 33, False),
                                            (5, 301, 37, False), (6, 777, 17, True),
                                            (7, 64, 23, False), (8, 1000, 20, True),
                                            (12, 257, 9, False), (16, 513, 11, True),
                                            (17, 300, 6, False), (24, 129, 5, True)])
def test_multi_gbm_tile_shapes(gpu, orc, d, n, steps, anti):
    # d <= 8: one register kernel per d (each instantiation here or in the
    # tests above); d <= 16 stages whole steps in the LDS tile (d = 12, 16:
    # one step per flush; 2 to 8: several, 7: a 14-value tile); d > 16 writes
    # rows directly; ragged last wavefronts and blocks, and n = 64: one
    # wavefront with rows, three without
    corr, rng = _corr(d, d)
    kw = dict(s0=list(rng.uniform(50, 150, d)), mu=list(rng.uniform(0.0, 0.08, d)),
              sigma=list(rng.uniform(0.1, 0.4, d)), corr=corr, t=1.0, steps=steps, n_paths=n,
              seed=d + 1, antithetic=anti)
    dev = gpu.multi_gbm(**kw)
    ref = orc.multi_gbm(**kw)
    assert dev.shape == (n, steps + 1, d)
    np.testing.assert_allclose(dev, ref, rtol=RTOL64, atol=0)


def test_multi_gbm_d8_100k(gpu, orc):
    corr, rng = _corr(8, 42)
    kw = dict(s0=list(rng.uniform(50, 150, 8)), mu=list(rng.uniform(0.0, 0.08, 8)),
              sigma=list(rng.uniform(0.1, 0.4, 8)), corr=corr, t=1.0, steps=252,
              n_paths=100_000, seed=42)
    np.testing.assert_allclose(gpu.multi_gbm(**kw), orc.multi_gbm(**kw), rtol=RTOL64, atol=0)


def test_multi_gbm_d1_equals_gbm(gpu):
    kw = dict(t=1.0, steps=60, n_paths=500, seed=11)
    m = gpu.multi_gbm(s0=[100.0], mu=[0.05], sigma=[0.2], corr=np.eye(1), **kw)
    g = gpu.gbm(s0=100.0, mu=0.05, sigma=0.2, **kw)
    np.testing.assert_array_equal(m[:, :, 0], g)


# ---------------------------------------------------------- fused payoffs --
HESTON = dict(s0=100.0, v0=0.04, mu=0.05, kappa=2.0, theta=0.04, xi=0.3, rho=-0.7)
MERTON = dict(s0=100.0, mu=0.05, sigma=0.2, lambda_=1.0, mu_j=-0.05, sigma_j=0.1)
GBM = dict(s0=100.0, mu=0.05, sigma=0.2)
STRIKES = [90.0, 100.0, 110.0]


def _oracle_prices(orc, model, params, steps, n, seed, r=0.05, t=1.0):
    mid = {"gbm": 0, "heston": 1, "merton": 2}[model]
    if model == "gbm":
        p = [params["s0"], params["mu"], params["sigma"], 0.0]
    else:
        p = list(params.values())
    term = orc.terminals(mid, p, t, steps, n, seed)
    s1, s2 = orc.call_sums(term, STRIKES)
    return math.exp(-r * t) * s1 / n, s1, s2


@pytest.mark.parametrize("model,params", [("gbm", GBM), ("heston", HESTON),
                                          ("merton", MERTON)])
def test_fused_fp64_matches_oracle(gpu, orc, model, params):
    n, steps = 20_000, 252
    res = gpu.european_price(model, params, 1.0, steps, n, STRIKES, r=0.05, seed=42,
                             precision=64)
    ref, s1, s2 = _oracle_prices(orc, model, params, steps, n, 42)
    np.testing.assert_allclose(res["price"], ref, rtol=RTOL64)
    np.testing.assert_allclose(res["sum"], s1, rtol=RTOL64)
    np.testing.assert_allclose(res["sumsq"], s2, rtol=RTOL64)


@pytest.mark.parametrize("model,params", [("gbm", GBM), ("heston", HESTON)])
def test_fused_fp64_paths_past_a_million_steps(gpu, orc, model, params):
    """One wavefront of paths 2^20 + 3 steps long through the queue schedule
    (its ring head wraps ~2^17 times): same draws, same sums as the oracle."""
    n, steps = 64, (1 << 20) + 3
    res = gpu.european_price(model, params, 1.0, steps, n, STRIKES, r=0.05, seed=7,
                             precision=64)
    ref, s1, s2 = _oracle_prices(orc, model, params, steps, n, 7)
    np.testing.assert_allclose(res["sum"], s1, rtol=RTOL64)
    np.testing.assert_allclose(res["sumsq"], s2, rtol=RTOL64)


@pytest.mark.parametrize("model,params", [("gbm", GBM), ("heston", HESTON),
                                          ("merton", MERTON)])
def test_fused_fp32_within_1e4(gpu, orc, model, params):
    n, steps = 20_000, 252
    res = gpu.european_price(model, params, 1.0, steps, n, STRIKES, r=0.05, seed=42,
                             precision=32)
    ref, _, _ = _oracle_prices(orc, model, params, steps, n, 42)
    rel = np.abs(res["price"] - ref) / ref
    assert np.all(rel <= 1e-4), rel


# --------------------------------------------------------------------- LSMC --
@pytest.mark.parametrize("is_put,deg", [(True, 3), (False, 2), (True, 4), (True, 1)])
def test_lsmc_fp64_matches_oracle(gpu, orc, is_put, deg):
    kw = dict(s0=100.0, k=100.0, r=0.05, sigma=0.2, t=1.0, steps=50, n_paths=20_000,
              is_put=is_put, poly_degree=deg, seed=42)
    p, se = gpu.lsmc_american_option(**kw)
    rp, rse = orc.lsmc(**kw)
    assert p == pytest.approx(rp, rel=1e-9)
    assert se == pytest.approx(rse, rel=1e-9)


@pytest.mark.parametrize("multikernel", [False, True])
def test_lsmc_kernels_match_oracle(gpu, orc, monkeypatch, multikernel):
    # the one-launch cooperative induction (default) and the per-date kernels
    if multikernel:
        monkeypatch.setenv("STOCHA_LSMC_MULTIKERNEL", "1")
    kw = dict(s0=100.0, k=100.0, r=0.05, sigma=0.2, t=1.0, steps=30, n_paths=30_000,
              is_put=True, poly_degree=3, seed=7)
    p, se = gpu.lsmc_american_option(**kw)
    rp, rse = orc.lsmc(**kw)
    assert p == pytest.approx(rp, rel=1e-9)
    assert se == pytest.approx(rse, rel=1e-9)


@pytest.mark.parametrize("n_paths", [300_000, 600_000, 1_200_000, 2_200_000])
def test_lsmc_paths_per_thread_match_oracle(gpu, orc, n_paths):
    # 256 CUs x 1024 threads: 2, 4, 8 and 16 paths per thread in the
    # cooperative kernel (16 re-reads the previous date's spots)
    kw = dict(s0=100.0, k=100.0, r=0.05, sigma=0.2, t=1.0, steps=6, n_paths=n_paths,
              is_put=True, poly_degree=3, seed=11)
    p, se = gpu.lsmc_american_option(**kw)
    rp, rse = orc.lsmc(**kw)
    assert p == pytest.approx(rp, rel=1e-9)
    assert se == pytest.approx(rse, rel=1e-9)


@pytest.mark.parametrize("n_paths", [262_144, 262_145, 524_289, 1_048_577, 4_194_304])
def test_lsmc_paths_per_thread_boundaries(gpu, orc, n_paths):
    # on 256 CUs: 1 path per thread exactly; one path over (2 per thread, the
    # second half partly empty); just over 2 and 4 per thread; 16 per thread
    # with every slot used -- the per-path checks of the second half of a
    # thread's paths and the exact fit
    kw = dict(s0=100.0, k=100.0, r=0.05, sigma=0.2, t=1.0, steps=4, n_paths=n_paths,
              is_put=True, poly_degree=3, seed=13)
    p, se = gpu.lsmc_american_option(**kw)
    rp, rse = orc.lsmc(**kw)
    assert p == pytest.approx(rp, rel=1e-9)
    assert se == pytest.approx(rse, rel=1e-9)


@pytest.mark.parametrize("steps", [1, 2, 3])
def test_lsmc_fewest_dates(gpu, orc, steps):
    # one step: no regression date (the per-date kernels); two: one date in
    # the cooperative induction
    kw = dict(s0=100.0, k=100.0, r=0.05, sigma=0.2, t=1.0, steps=steps, n_paths=70_001,
              is_put=False, poly_degree=2, seed=5)
    p, se = gpu.lsmc_american_option(**kw)
    rp, rse = orc.lsmc(**kw)
    assert p == pytest.approx(rp, rel=1e-9)
    assert se == pytest.approx(rse, rel=1e-9)


@pytest.mark.parametrize("multikernel", [False, True])
@pytest.mark.parametrize("seed", [0, 1, 4])
def test_lsmc_dates_at_the_regression_threshold(gpu, orc, monkeypatch, multikernel, seed):
    # 12-path put books whose dates hold exactly ncols (skipped, lsmc.rs:62)
    # and ncols + 1 (solved) paths in the money
    if multikernel:
        monkeypatch.setenv("STOCHA_LSMC_MULTIKERNEL", "1")
    kw = dict(s0=100.0, k=92.0, r=0.05, sigma=0.3, t=1.0, steps=20, n_paths=12,
              is_put=True, poly_degree=3, seed=seed)
    _, itm, _, _ = orc.lsmc_trace(**kw)
    assert 4 in [int(m) for m in itm[1:]]
    p, se = gpu.lsmc_american_option(**kw)
    rp, rse = orc.lsmc(**kw)
    assert p == pytest.approx(rp, rel=1e-9)
    assert se == pytest.approx(rse, rel=1e-9)


@pytest.mark.parametrize("n_paths", [262_145, 4_194_304])
def test_lsmc_fp32_boundaries_agree_with_fp64(gpu, n_paths):
    kw = dict(s0=100.0, k=100.0, r=0.05, sigma=0.2, t=1.0, steps=4, n_paths=n_paths,
              is_put=True, poly_degree=3, seed=13)
    p32, se32 = gpu.lsmc_american_option_ex(**kw, precision=32)
    p64, se64 = gpu.lsmc_american_option(**kw)
    assert abs(p32 - p64) < 0.05 * se64, (p32, p64, se64)


def test_lsmc_fp32_close_to_oracle(gpu, orc):
    kw = dict(s0=100.0, k=100.0, r=0.05, sigma=0.2, t=1.0, steps=50, n_paths=50_000,
              is_put=True, poly_degree=3, seed=42)
    p, se = gpu.lsmc_american_option_ex(**kw, precision=32)
    rp, rse = orc.lsmc(**kw)
    assert abs(p - rp) < 0.1 * rse, (p, rp, rse)


# ------------------------------------------------------------------- Greeks --
@pytest.mark.parametrize("is_call", [True, False])
def test_greeks_pathwise_matches_oracle(gpu, orc, is_call):
    kw = dict(s0=100.0, r=0.05, sigma=0.2, t=1.0, strike=100.0, is_call=is_call,
              n_paths=20_000, n_steps=252, seed=42)
    g = gpu.greeks_pathwise(**kw, greeks=["delta", "vega"])
    d, v = orc.greeks_pathwise(**kw)
    assert list(g) == ["delta", "vega"]
    assert g["delta"] == pytest.approx(d, rel=RTOL64)
    assert g["vega"] == pytest.approx(v, rel=RTOL64)


@pytest.mark.parametrize("model,params", [("gbm", GBM), ("heston", HESTON),
                                          ("merton", MERTON)])
@pytest.mark.parametrize("prec", [64, 32])
def test_feed_schedule_equals_loop(gpu, model, params, prec):
    # the per-lane state machine consumes each path's draws in the same order
    # as the per-normal loop, so both schedules give the same bits
    kw = dict(t=1.0, steps=100, n_paths=3000, strikes=STRIKES, r=0.05, seed=7, precision=prec)
    a = gpu.european_price(model, params, schedule="feed", **kw)
    b = gpu.european_price(model, params, schedule="loop", **kw)
    np.testing.assert_array_equal(a["sum"], b["sum"])
    np.testing.assert_array_equal(a["sumsq"], b["sumsq"])


def test_feed_multiple_paths_per_thread(gpu, orc):
    # more paths than one grid holds: lanes start new paths inside the loop
    n, steps = 1_200_000, 4
    res = gpu.european_price("heston", HESTON, 1.0, steps, n, STRIKES, r=0.05, seed=3,
                             precision=64)
    term = orc.terminals(1, list(HESTON.values()), 1.0, steps, n, 3)
    s1, _ = orc.call_sums(term, STRIKES)
    np.testing.assert_allclose(res["sum"], s1, rtol=RTOL64)
