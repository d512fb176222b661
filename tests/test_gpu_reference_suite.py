"""The reference's own property checks of the hot-path functions, run on the
GPU through the reference's import path (`import stocha`), so that a user of
the reference sees its test suite hold on this build:
tests/test_models.py:31-86 (GBM), :88-135 (Heston), :186-232 (Merton),
:621-743 (multi_gbm), :745-800 (LSMC); tests/test_greeks.py:131-178, 275-280
(pathwise Greeks).  Sizes, seeds and tolerances are the
reference's; the cases are grouped here by property instead of by class.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GBM = dict(s0=100.0, mu=0.05, sigma=0.2, t=1.0, steps=252)
HESTON = dict(s0=100.0, v0=0.04, mu=0.05, kappa=2.0, theta=0.04, xi=0.3, rho=-0.7, t=1.0,
              steps=252, seed=42)
MERTON = dict(s0=100.0, mu=0.05, sigma=0.2, lambda_=1.0, mu_j=-0.05, sigma_j=0.1, t=1.0,
              steps=252, seed=42)
C2 = np.array([[1.0, 0.6], [0.6, 1.0]])
C3 = np.array([[1.0, 0.6, 0.3], [0.6, 1.0, 0.5], [0.3, 0.5, 1.0]])
M2 = dict(s0=[100.0, 50.0], mu=[0.05, 0.08], sigma=[0.2, 0.3], corr=C2, t=1.0, steps=252)


@pytest.fixture(scope="module")
def stocha(gpu):
    import stocha as s
    return s


def _bs_put(s, k, r, sigma, t):
    cdf = lambda x: 0.5 * (1.0 + math.erf(x / math.sqrt(2.0)))  # noqa: E731
    d1 = (math.log(s / k) + (r + 0.5 * sigma * sigma) * t) / (sigma * math.sqrt(t))
    return k * math.exp(-r * t) * cdf(-(d1 - sigma * math.sqrt(t))) - s * cdf(-d1)


# ---------------------------------------------------- shape, s0, positivity --
def test_single_asset_shapes_start_and_positivity(stocha):
    cases = [  # (paths, s0 column) -- test_models.py:32-59, 95-106, 194-205
        (stocha.gbm(**GBM, n_paths=1_000, seed=0), 100.0),
        (stocha.gbm(s0=50.0, mu=0.0, sigma=0.1, t=1.0, steps=10, n_paths=100, seed=0), 50.0),
        (stocha.heston(**HESTON, n_paths=1000), 100.0),
        (stocha.merton_jump_diffusion(**MERTON, n_paths=1000), 100.0),
    ]
    for paths, s0 in cases:
        assert paths.dtype == np.float64
        assert paths.shape[1] in (253, 11)
        np.testing.assert_array_equal(paths[:, 0], s0)
        assert (paths > 0).all()
    assert cases[0][0].shape == (1_000, 253)
    assert (stocha.gbm(s0=100.0, mu=0.05, sigma=0.3, t=1.0, steps=252, n_paths=10_000,
                       seed=0) > 0).all()


def test_multi_gbm_shapes_start_and_positivity(stocha):
    # test_models.py:626-656, 700-707
    p = stocha.multi_gbm(**M2, n_paths=1000, seed=42)
    assert p.shape == (1000, 253, 2) and p.dtype == np.float64
    s0 = [100.0, 50.0, 200.0]
    p3 = stocha.multi_gbm(s0=s0, mu=[0.05, 0.08, 0.03], sigma=[0.2, 0.3, 0.15], corr=C3,
                          t=1.0, steps=252, n_paths=500, seed=0)
    assert p3.shape == (500, 253, 3)
    p3s = stocha.multi_gbm(s0=s0, mu=[0.05, 0.08, 0.03], sigma=[0.2, 0.3, 0.15], corr=C3,
                           t=1.0, steps=10, n_paths=100, seed=0)
    for i, s in enumerate(s0):
        np.testing.assert_array_equal(p3s[:, 0, i], s)
    assert (stocha.multi_gbm(**M2, n_paths=10_000, seed=0) > 0).all()
    assert stocha.multi_gbm(**M2, n_paths=1000, seed=42, antithetic=True).shape == (1000, 253, 2)


# ------------------------------------------------------------ reproducible --
def test_reproducibility(stocha):
    # test_models.py:43-46, 108-110, 207-212, 658-661
    np.testing.assert_array_equal(stocha.gbm(**GBM, n_paths=1_000, seed=42),
                                  stocha.gbm(**GBM, n_paths=1_000, seed=42))
    np.testing.assert_array_equal(stocha.heston(**HESTON, n_paths=1000),
                                  stocha.heston(**HESTON, n_paths=1000))
    np.testing.assert_array_equal(stocha.merton_jump_diffusion(**MERTON, n_paths=1000),
                                  stocha.merton_jump_diffusion(**MERTON, n_paths=1000))
    np.testing.assert_array_equal(stocha.multi_gbm(**M2, n_paths=1000, seed=42),
                                  stocha.multi_gbm(**M2, n_paths=1000, seed=42))


# -------------------------------------------------- martingale drift checks --
def test_expected_terminal_prices(stocha):
    e = 100.0 * math.exp(0.05)
    # test_models.py:48-54 (5e-3), :112-117 (2e-2), :214-219 (2e-2)
    assert stocha.gbm(**GBM, n_paths=100_000, seed=0)[:, -1].mean() == pytest.approx(e, rel=5e-3)
    assert stocha.heston(**HESTON, n_paths=100_000)[:, -1].mean() == pytest.approx(e, rel=2e-2)
    assert stocha.merton_jump_diffusion(**MERTON, n_paths=200_000)[:, -1].mean() == \
        pytest.approx(e, rel=2e-2)
    # test_models.py:663-676
    p = stocha.multi_gbm(**M2, n_paths=100_000, seed=0)
    for i, (s0, mu) in enumerate(zip(M2["s0"], M2["mu"])):
        assert abs(p[:, -1, i].mean() - s0 * math.exp(mu)) / (s0 * math.exp(mu)) < 0.02


def test_antithetic_reduces_variance(stocha):
    # test_models.py:61-66, 687-698
    plain = stocha.gbm(**GBM, n_paths=50_000, seed=42, antithetic=False)
    anti = stocha.gbm(**GBM, n_paths=50_000, seed=42, antithetic=True)
    assert anti[:, -1].std() < plain[:, -1].std()
    mp = stocha.multi_gbm(**M2, n_paths=50_000, seed=42, antithetic=False)
    ma = stocha.multi_gbm(**M2, n_paths=50_000, seed=42, antithetic=True)
    assert ma[:, -1, :].sum(axis=1).std() < mp[:, -1, :].sum(axis=1).std()


@pytest.mark.parametrize("corr,want,tol", [(C2, 0.6, 0.05), (np.eye(2), 0.0, 0.02)])
def test_multi_gbm_correlation(stocha, corr, want, tol):
    # test_models.py:678-685, 709-717
    if want == 0.0:
        p = stocha.multi_gbm(s0=[100.0, 100.0], mu=[0.05, 0.05], sigma=[0.2, 0.2], corr=corr,
                             t=1.0, steps=252, n_paths=50_000, seed=0)
    else:
        p = stocha.multi_gbm(**M2, n_paths=50_000, seed=0)
    lr = np.log(p[:, 1:, :] / p[:, :-1, :]).reshape(-1, 2)
    assert abs(np.corrcoef(lr.T)[0, 1] - want) < tol


# ------------------------------------------------------------- validation --
@pytest.mark.parametrize("fn,kw", [
    ("gbm", dict(GBM, s0=-1.0, steps=10, n_paths=10, seed=0)),        # :68-71
    ("gbm", dict(GBM, sigma=0.0, steps=10, n_paths=5, seed=0)),       # :73-76
    ("gbm", dict(GBM, t=-1.0, steps=10, n_paths=10, seed=0)),         # :78-81
    ("heston", dict(HESTON, s0=-1.0, n_paths=1000)),                  # :119-123
    ("heston", dict(HESTON, rho=2.0, n_paths=1000)),                  # :125-129
    ("heston", dict(HESTON, n_paths=1000, scheme="bad")),             # :131-134
    ("merton_jump_diffusion", dict(MERTON, sigma=0.0, n_paths=1000)),  # :221-225
    ("merton_jump_diffusion", dict(MERTON, s0=-1.0, n_paths=1000)),   # :227-231
    ("multi_gbm", dict(s0=[100.0], mu=[0.05, 0.08], sigma=[0.2], corr=np.eye(1), t=1.0,
                       steps=10, n_paths=10)),                          # :719-722
    ("multi_gbm", dict(M2, s0=[-1.0, 50.0], steps=10, n_paths=10)),   # :724-727
    ("multi_gbm", dict(M2, corr=np.array([[1.0, 1.5], [1.5, 1.0]]), steps=10,
                       n_paths=10)),                                    # :729-733
    ("multi_gbm", dict(M2, corr=np.eye(3), steps=10, n_paths=10)),    # :735-739
])
def test_invalid_arguments_raise(stocha, fn, kw):
    with pytest.raises(ValueError):
        getattr(stocha, fn)(**kw)


# ------------------------------------------------------------------- LSMC --
def test_lsmc_reference_cases(stocha):
    # test_models.py:746-755: types and signs
    p, se = stocha.lsmc_american_option(s0=100.0, k=100.0, r=0.05, sigma=0.20, t=1.0,
                                        steps=10, n_paths=1000)
    assert isinstance(p, float) and isinstance(se, float) and p > 0 and se > 0
    # :757-764: American put >= European put
    p, se = stocha.lsmc_american_option(s0=100.0, k=100.0, r=0.05, sigma=0.20, t=1.0,
                                        steps=50, n_paths=30_000, seed=0)
    assert p >= _bs_put(100.0, 100.0, 0.05, 0.20, 1.0) - 3 * se
    # :766-772: the ATM benchmark 6.07
    p, _ = stocha.lsmc_american_option(s0=100.0, k=100.0, r=0.05, sigma=0.20, t=1.0,
                                       steps=50, n_paths=50_000, seed=42)
    assert abs(p - 6.07) < 0.30
    # :774-786: deep in and out of the money
    assert stocha.lsmc_american_option(s0=80.0, k=100.0, r=0.01, sigma=0.10, t=1.0, steps=50,
                                       n_paths=10_000)[0] > 15.0
    assert stocha.lsmc_american_option(s0=150.0, k=100.0, r=0.05, sigma=0.20, t=1.0, steps=50,
                                       n_paths=10_000)[0] < 2.0
    # :788-793: the call form
    assert stocha.lsmc_american_option(s0=100.0, k=100.0, r=0.05, sigma=0.20, t=1.0, steps=50,
                                       n_paths=20_000, is_put=False)[0] > 0


# ------------------------------------------------------- pathwise Greeks --
def _bs(s, k, r, sigma, t):
    d1 = (math.log(s / k) + (r + 0.5 * sigma * sigma) * t) / (sigma * math.sqrt(t))
    cdf = 0.5 * (1.0 + math.erf(d1 / math.sqrt(2.0)))
    pdf = math.exp(-0.5 * d1 * d1) / math.sqrt(2.0 * math.pi)
    return cdf, s * pdf * math.sqrt(t)  # call delta, vega


def test_greeks_pathwise_reference_cases(stocha):
    # tests/test_greeks.py:131-156 (200,000 x 252, seed 42, same tolerances)
    kw = dict(s0=100.0, r=0.05, sigma=0.2, t=1.0, strike=100.0, n_paths=200_000, n_steps=252,
              seed=42)
    delta, vega = _bs(100.0, 100.0, 0.05, 0.2, 1.0)
    assert abs(stocha.greeks_pathwise(**kw, is_call=True, greeks=["delta"])["delta"]
               - delta) < 0.015
    assert abs(stocha.greeks_pathwise(**kw, is_call=False, greeks=["delta"])["delta"]
               - (delta - 1.0)) < 0.015
    pw = stocha.greeks_pathwise(**kw, is_call=True, greeks=["vega"])
    assert abs(pw["vega"] - vega) < 1.5
    # :169-178: pathwise and bump-and-revalue deltas agree
    fd = stocha.greeks_fd(model="gbm", params={"s0": 100.0, "r": 0.05, "sigma": 0.2, "t": 1.0},
                          payoff="call", strike=100.0, n_paths=200_000, n_steps=252,
                          greeks=["delta"], seed=42)
    pd = stocha.greeks_pathwise(**kw, is_call=True, greeks=["delta"])
    assert abs(fd["delta"] - pd["delta"]) < 0.03


def test_greeks_pathwise_reference_errors(stocha):
    base = dict(r=0.05, sigma=0.2, t=1.0, strike=100.0, is_call=True, n_paths=1000, n_steps=10)
    with pytest.raises(ValueError, match="only supports"):  # tests/test_greeks.py:158-164
        stocha.greeks_pathwise(s0=100.0, **base, greeks=["gamma"])
    with pytest.raises(ValueError, match="s0 must be positive"):  # :275-280
        stocha.greeks_pathwise(s0=-1.0, **base, greeks=["delta"])
