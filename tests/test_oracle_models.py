"""Pins the oracle's steppers, LSMC and pathwise Greeks (CPU only).

Each check is chosen so a plausible slip (a dropped term, wrong sign, wrong
stream index or block size, swapped normals) fails it:
  * identities against the NumPy-pinned normals: log S_T of a GBM path is
    log s0 + steps*drift + diffusion*sum(z) for that path's stream; the same
    for Heston with vanishing vol-of-vol (first normal of each pair drives S);
  * d=1 multi-asset GBM is bit-identical to GBM (same block size, L=1);
  * closed forms at the reference's own tolerances (tests/test_models.py,
    tests/test_greeks.py): E[S_T], the Merton Poisson-weighted BS sum, the
    LSMC benchmark put, BS delta/vega;
  * LSMC backward induction against an independent NumPy implementation
    using numpy.linalg.lstsq on the oracle's paths;
  * pathwise Greeks against central finite differences with common random
    numbers (the streams do not depend on s0 or sigma).
"""
import math

import numpy as np
import pytest


def norm_cdf(x):
    return 0.5 * (1.0 + math.erf(x / math.sqrt(2.0)))


def norm_pdf(x):
    return math.exp(-0.5 * x * x) / math.sqrt(2 * math.pi)


def bs_call(s, k, r, sigma, t):
    d1 = (math.log(s / k) + (r + 0.5 * sigma * sigma) * t) / (sigma * math.sqrt(t))
    return s * norm_cdf(d1) - k * math.exp(-r * t) * norm_cdf(d1 - sigma * math.sqrt(t))


def bs_put(s, k, r, sigma, t):
    return bs_call(s, k, r, sigma, t) - s + k * math.exp(-r * t)


# ---------------------------------------------------------------- GBM ----
@pytest.mark.parametrize("anti", [False, True])
def test_gbm_log_sum_identity(orc, anti):
    s0, mu, sigma, t, steps, n, q = 100.0, 0.05, 0.2, 1.0, 40, 9, 0.01
    paths = orc.gbm(s0, mu, sigma, t, steps, n, seed=42, antithetic=anti, q=q)
    dt = t / steps
    drift = (mu - q - 0.5 * sigma * sigma) * dt
    diff = sigma * math.sqrt(dt)
    base = (n + 1) // 2 if anti else n
    for row in range(n):
        path, sgn = (row, 1.0) if row < base else (row - base, -1.0)
        z, _ = orc.normals(42, orc.block_gbm(steps), path, steps)
        ls = np.log(s0) + np.cumsum(drift + diff * sgn * z)
        np.testing.assert_allclose(np.log(paths[row, 1:]), ls, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(paths[:, 0], s0)


def test_gbm_expected_terminal(orc):
    p = orc.gbm(100.0, 0.05, 0.2, 1.0, 252, 100_000, seed=0)
    assert p[:, -1].mean() == pytest.approx(100.0 * math.exp(0.05), rel=5e-3)  # :48-54


def test_gbm_antithetic_reduces_variance(orc):
    kw = dict(s0=100.0, mu=0.05, sigma=0.2, t=1.0, steps=252, n_paths=50_000, seed=42)
    assert orc.gbm(**kw, antithetic=True)[:, -1].std() < orc.gbm(**kw)[:, -1].std()


# -------------------------------------------------------------- Heston ----
def test_heston_constant_variance_identity(orc):
    # v0 = theta and xi -> 0 keeps v at theta: S follows GBM driven by the first
    # normal of each (dw1, dz) pair of the Heston block stream.
    s0, v0, mu, kappa, theta, xi, rho, t, steps = 100.0, 0.04, 0.05, 2.0, 0.04, 1e-13, -0.7, 1.0, 30
    paths = orc.heston(s0, v0, mu, kappa, theta, xi, rho, t, steps, 5, seed=9)
    dt = t / steps
    for i in range(5):
        z, _ = orc.normals(9, orc.block_heston(steps), i, 2 * steps)
        dw1 = z[0::2]
        ls = math.log(s0) + np.cumsum((mu - 0.5 * theta) * dt + math.sqrt(theta * dt) * dw1)
        np.testing.assert_allclose(np.log(paths[i, 1:]), ls, rtol=0, atol=1e-9)


def test_heston_expected_terminal(orc):
    p = orc.heston(100.0, 0.04, 0.05, 2.0, 0.04, 0.3, -0.7, 1.0, 252, 100_000, seed=42)
    assert p[:, -1].mean() == pytest.approx(100.0 * math.exp(0.05), rel=2e-2)  # :112-117
    assert (p > 0).all()


def test_heston_rho_sign_changes_skew(orc):
    # negative rho gives a left-skewed terminal log-price, positive rho right-skewed
    from scipy import stats
    kw = dict(s0=100.0, v0=0.04, mu=0.05, kappa=2.0, theta=0.04, xi=0.6, t=1.0, steps=100,
              n_paths=40_000, seed=1)
    neg = stats.skew(np.log(orc.heston(rho=-0.8, **kw)[:, -1]))
    pos = stats.skew(np.log(orc.heston(rho=0.8, **kw)[:, -1]))
    assert neg < -0.1 and pos > 0.1


# -------------------------------------------------------------- Merton ----
def _merton_call(S, K, r, sigma, T, lam, mu_j, sigma_j):
    if lam <= 0:
        return bs_call(S, K, r, sigma, T)
    m_bar = math.exp(mu_j + 0.5 * sigma_j * sigma_j) - 1.0
    lam_t = lam * (1.0 + m_bar) * T
    price, cum = 0.0, 0.0
    for n in range(200):
        w = math.exp(-lam_t + n * math.log(lam_t) - math.lgamma(n + 1))
        var_n = sigma * sigma + n * sigma_j * sigma_j / T
        r_n = r - lam * m_bar + n * (mu_j + 0.5 * sigma_j * sigma_j) / T
        price += w * bs_call(S, K, r_n, math.sqrt(var_n), T)
        cum += w
        if n >= 20 and cum > 1 - 1e-12:
            break
    return price


@pytest.mark.parametrize("S,K,r,sigma,T,lam,mu_j,sigma_j", [
    (100, 100, 0.05, 0.20, 1.0, 1.0, -0.05, 0.10),
    (100, 100, 0.05, 0.20, 1.0, 5.0, -0.05, 0.10),
    (100, 140, 0.05, 0.20, 1.0, 1.0, -0.05, 0.10),
    (100, 100, 0.05, 0.15, 1.0, 2.0, -0.30, 0.20),
])
def test_merton_call_vs_poisson_bs(orc, S, K, r, sigma, T, lam, mu_j, sigma_j):
    # tests/test_models.py:290-333 (atol 0.25, rtol 0.04 at 200k paths)
    steps = max(252, int(math.ceil(lam * T / 0.005)))
    p = orc.merton(S, r, sigma, lam, mu_j, sigma_j, T, steps, 200_000, seed=42)
    mc = math.exp(-r * T) * np.maximum(p[:, -1] - K, 0).mean()
    assert mc == pytest.approx(_merton_call(S, K, r, sigma, T, lam, mu_j, sigma_j), abs=0.25,
                               rel=0.04)


def test_merton_lambda_zero_has_no_jumps(orc):
    # lambda = 0: every step draws a normal and an unused uniform; the log
    # increments are then Gaussian with the GBM drift (no jump outliers).
    p = orc.merton(100.0, 0.05, 0.2, 0.0, -0.5, 0.5, 1.0, 50, 20_000, seed=3)
    inc = np.diff(np.log(p), axis=1).ravel()
    dt = 1.0 / 50
    assert np.abs(inc - (0.05 - 0.02) * dt).max() < 0.2 * math.sqrt(dt) * 6


# --------------------------------------------------------- multi-asset ----
def test_multi_d1_is_gbm_bitwise(orc):
    m = orc.multi_gbm([100.0], [0.05], [0.2], np.eye(1), 1.0, 60, 300, seed=5, antithetic=True)
    g = orc.gbm(100.0, 0.05, 0.2, 1.0, 60, 300, seed=5, antithetic=True)
    np.testing.assert_array_equal(m[:, :, 0], g)


def test_cholesky_matches_numpy(orc):
    rng = np.random.default_rng(0)
    a = rng.standard_normal((12, 12))
    c = a @ a.T / 12 + 0.1 * np.eye(12)
    np.testing.assert_allclose(orc.cholesky(c), np.linalg.cholesky(c), rtol=1e-12, atol=1e-14)


def test_cholesky_not_pd_message(orc):
    with pytest.raises(ValueError, match=r"not positive-definite \(diagonal element 1 = -1.25"):
        orc.cholesky(np.array([[1.0, 1.5], [1.5, 1.0]]))


def test_multi_correlation_structure(orc):
    corr = np.array([[1.0, 0.6], [0.6, 1.0]])
    p = orc.multi_gbm([100.0, 50.0], [0.05, 0.08], [0.2, 0.3], corr, 1.0, 252, 20_000, seed=0)
    lr = np.log(p[:, 1:, :] / p[:, :-1, :]).reshape(-1, 2)
    assert abs(np.corrcoef(lr.T)[0, 1] - 0.6) < 0.05  # tests/test_models.py:678-688


# ---------------------------------------------------------------- LSMC ----
def _lsmc_numpy(paths, k, r, dt, is_put, deg):
    """Longstaff-Schwartz backward induction (lsmc.rs:41-112) in NumPy, with
    numpy.linalg.lstsq as the solver."""
    n, w = paths.shape
    steps = w - 1
    disc = math.exp(-r * dt)
    intr = (lambda s: np.maximum(k - s, 0.0)) if is_put else (lambda s: np.maximum(s - k, 0.0))
    cf = intr(paths[:, steps])
    for step in range(steps - 1, 0, -1):
        cf = cf * disc
        spot = paths[:, step]
        itm = np.nonzero(intr(spot) > 0)[0]
        if len(itm) < deg + 2:
            continue
        s = spot[itm]
        mean = s.mean()
        std = max(math.sqrt(((s - mean) ** 2).mean()), 1e-12)
        x = (s - mean) / std
        X = np.vander(x, deg + 1, increasing=True)
        coef = np.linalg.lstsq(X, cf[itm], rcond=None)[0]
        cont = X @ coef
        ex = intr(s)
        hit = ex > np.maximum(cont, 0.0)
        cf[itm[hit]] = ex[hit]
    cf = cf * disc
    return cf.mean(), math.sqrt(cf.var(ddof=1) / n)


@pytest.mark.parametrize("is_put,deg", [(True, 3), (False, 2), (True, 1), (True, 4)])
def test_lsmc_matches_numpy_induction(orc, is_put, deg):
    kw = dict(s0=100.0, k=105.0 if is_put else 95.0, r=0.05, sigma=0.25, t=1.0, steps=20,
              n_paths=3000, seed=42)
    price, se = orc.lsmc(**kw, is_put=is_put, poly_degree=deg)
    paths = orc.gbm(kw["s0"], kw["r"], kw["sigma"], kw["t"], kw["steps"], kw["n_paths"], seed=42)
    p2, se2 = _lsmc_numpy(paths, kw["k"], kw["r"], kw["t"] / kw["steps"], is_put, deg)
    assert price == pytest.approx(p2, rel=1e-10)
    assert se == pytest.approx(se2, rel=1e-8)


def test_lstsq_matches_numpy(orc):
    rng = np.random.default_rng(1)
    for m, n in [(50, 4), (1000, 5), (7, 2)]:
        X = rng.standard_normal((m, n))
        y = rng.standard_normal(m)
        np.testing.assert_allclose(orc.lstsq(X, y), np.linalg.lstsq(X, y, rcond=None)[0],
                                   rtol=1e-10, atol=1e-12)


def test_lsmc_reference_benchmarks(orc):
    p, se = orc.lsmc(100.0, 100.0, 0.05, 0.2, 1.0, 50, 50_000, True, 3, 42)
    assert abs(p - 6.07) < 0.30  # tests/test_models.py:766-772
    p, se = orc.lsmc(100.0, 100.0, 0.05, 0.2, 1.0, 50, 30_000, True, 3, 0)
    assert p >= bs_put(100.0, 100.0, 0.05, 0.2, 1.0) - 3 * se  # :756-764
    assert orc.lsmc(80.0, 100.0, 0.01, 0.1, 1.0, 50, 10_000, True, 3, 42)[0] > 15.0
    assert orc.lsmc(150.0, 100.0, 0.05, 0.2, 1.0, 50, 10_000, True, 3, 42)[0] < 2.0


# --------------------------------------------------------------- Greeks ----
def test_pathwise_vs_black_scholes(orc):
    # tests/test_greeks.py:131-156 (200k x 252, seed 42)
    d, v = orc.greeks_pathwise(100.0, 0.05, 0.2, 1.0, 100.0, True, 200_000, 252, seed=42)
    d1 = (math.log(1.0) + (0.05 + 0.02)) / 0.2
    assert abs(d - norm_cdf(d1)) < 0.015
    assert abs(v - 100.0 * norm_pdf(d1)) < 1.5
    dp, _ = orc.greeks_pathwise(100.0, 0.05, 0.2, 1.0, 100.0, False, 200_000, 252, seed=42)
    assert abs(dp - (norm_cdf(d1) - 1.0)) < 0.015


@pytest.mark.parametrize("is_call", [True, False])
def test_pathwise_vs_crn_finite_difference(orc, is_call):
    s0, r, sigma, t, k, n, steps = 100.0, 0.05, 0.2, 1.0, 100.0, 50_000, 50
    d, v = orc.greeks_pathwise(s0, r, sigma, t, k, is_call, n, steps, seed=7)

    def price(s, sg):
        term = orc.terminals(0, [s, r, sg, 0.0], t, steps, n, 7)
        pay = np.maximum(term - k, 0) if is_call else np.maximum(k - term, 0)
        return math.exp(-r * t) * pay.mean()

    h = 1e-4 * s0
    fd_d = (price(s0 + h, sigma) - price(s0 - h, sigma)) / (2 * h)
    hs = 1e-4 * sigma
    fd_v = (price(s0, sigma + hs) - price(s0, sigma - hs)) / (2 * hs)
    assert d == pytest.approx(fd_d, rel=2e-3, abs=1e-4)
    assert v == pytest.approx(fd_v, rel=2e-3, abs=1e-3)
