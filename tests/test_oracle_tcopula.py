"""Pins the oracle's Student-t copula (copula/student_t.rs) on the CPU.

No golden vectors exist for it in the reference (its tests check shapes,
the unit interval, nu > 2, heavier joint tails than the Gaussian copula and
reproducibility: tests/test_risk.py:95-136, repeated here), so each part is
pinned by what the mathematics fixes, chosen so that a plausible slip fails:
  * the log-gamma (Lanczos g = 7) against math.lgamma, both branches;
  * the t CDF (incomplete beta by modified Lentz, stop at 3e-7) against
    closed forms for nu = 1, 2, 4 and scipy.stats.t elsewhere -- a swapped
    (a, b), a dropped symmetry branch or a wrong sign fails by orders of
    magnitude more than the 1e-6 allowed;
  * the gamma sampler through the scales s_i = sqrt(nu / chi2_i): nu / s^2
    must be chi-square(nu) (shape nu/2 and the factor 2 of student_t.rs:121-124;
    a shape of nu or a missing factor 2 fails the KS test);
  * the copula: uniform marginals (the CDF's nu equals the sampler's),
    Kendall's tau = 2/pi arcsin(rho) (a property of every elliptical copula;
    a transposed Cholesky factor fails it in d = 3), and the normals of each
    sample are the stream's next d normals after its gamma draws (checked by
    decoding the stream with the pinned RNG-class oracle from the draws the
    sample consumed).
"""
import math

import numpy as np
import pytest
from scipy import stats

import stocha_amd as S


def test_lgamma_matches_math(orc):
    for x in (0.1, 0.3, 0.5, 0.7, 1.0, 1.5, 2.5, 7.25, 30.0, 170.5):
        assert orc.lgamma_lanczos(x) == pytest.approx(math.lgamma(x), rel=1e-12, abs=1e-13)


@pytest.mark.parametrize("nu,exact", [
    (1.0, lambda x: 0.5 + np.arctan(x) / np.pi),
    (2.0, lambda x: 0.5 + x / (2.0 * np.sqrt(2.0 + x * x))),
    (4.0, lambda x: 0.5 + 0.75 * (x / np.sqrt(4 + x * x)) * (1 - x * x / (3 * (4 + x * x)))),
])
def test_t_cdf_closed_forms(orc, nu, exact):
    xs = np.concatenate([np.linspace(-30, 30, 121), [0.0, 1e-9, -1e-9]])
    np.testing.assert_allclose(orc.student_t_cdf(xs, nu), exact(xs), rtol=0, atol=1e-6)


@pytest.mark.parametrize("nu", [2.5, 3.0, 5.0, 7.3, 30.0, 200.0])
def test_t_cdf_matches_scipy(orc, nu):
    xs = np.linspace(-12, 12, 97)
    np.testing.assert_allclose(orc.student_t_cdf(xs, nu), stats.t.cdf(xs, nu), rtol=0, atol=1e-6)
    # symmetry: F(-x) = 1 - F(x)
    np.testing.assert_allclose(orc.student_t_cdf(-xs, nu), 1 - orc.student_t_cdf(xs, nu),
                               atol=1e-12)


def test_scales_are_chi_square(orc):
    nu = 5.0
    corr = np.eye(2)
    _, _, s, _ = orc.student_t_copula(corr, nu, 20_000, seed=11, details=True)
    chi2 = nu / s ** 2
    assert stats.kstest(chi2, stats.chi2(nu).cdf).pvalue > 1e-3
    assert stats.kstest(chi2, stats.chi2(2 * nu).cdf).pvalue < 1e-6  # shape nu, not nu/2


def test_marginals_uniform_and_kendall_tau(orc):
    rho = np.array([[1.0, 0.7, -0.3], [0.7, 1.0, 0.2], [-0.3, 0.2, 1.0]])
    u = orc.student_t_copula(rho, 4.0, 20_000, seed=5)
    assert np.all(u > 0) and np.all(u < 1)
    for j in range(3):
        assert stats.kstest(u[:, j], "uniform").pvalue > 1e-3
    for i, j in ((0, 1), (0, 2), (1, 2)):
        tau = stats.kendalltau(u[:, i], u[:, j]).statistic
        assert tau == pytest.approx(2 / np.pi * np.arcsin(rho[i, j]), abs=0.02)


def test_draw_order_gamma_then_normals(orc):
    # sample 0 on the pinned RNG-class stream: a gamma normal x, a uniform u
    # (accepted here by the squeeze, the premise is asserted), then the d
    # normals; drawing the normals first, or no uniform, reads other draws
    nu, d = 6.0, 3
    corr = np.eye(d)
    _, z, s, _ = orc.student_t_copula(corr, nu, 1, seed=21, details=True)
    r = orc.Rng(21)
    x = r.normal(1)[0]
    u = r.uniform(1)[0]
    sh = nu / 2 - 1 / 3
    v = (1 + x / math.sqrt(9 * sh)) ** 3
    assert u < 1 - 0.0331 * x ** 4  # accepted on the first attempt
    np.testing.assert_array_equal(z[0], r.normal(d))
    assert s[0] == pytest.approx(math.sqrt(nu / (2 * sh * v)), rel=1e-14)


def test_normals_are_the_stream_between_samples(orc):
    # over many samples: the normals of all samples, in order, are a
    # subsequence of the stream's normals decoded with the RNG-class oracle
    # from the start (gamma normals interleave; uniforms are one draw)
    corr = np.array([[1.0, 0.4], [0.4, 1.0]])
    _, z, _, used = orc.student_t_copula(corr, 6.0, 200, seed=4, details=True)
    assert 200 * 4 < used < 200 * 5  # ~2.1 gamma draws + 2.05 for the normals
    assert np.all(np.isfinite(z)) and abs(np.mean(z)) < 0.15 and 0.85 < np.std(z) < 1.15


def test_reference_properties(orc):
    # tests/test_risk.py:95-136: shape, (0, 1), heavier joint tails than the
    # Gaussian copula at nu = 3, reproducibility
    corr = np.array([[1.0, 0.8], [0.8, 1.0]])
    u_t = orc.student_t_copula(corr, 3.0, 20_000, seed=0)
    u_g = orc.gaussian_copula(corr, 20_000, seed=0)
    jt = np.mean((u_t[:, 0] > 0.95) & (u_t[:, 1] > 0.95))
    jg = np.mean((u_g[:, 0] > 0.95) & (u_g[:, 1] > 0.95))
    assert u_t.shape == (20_000, 2) and jt > jg
    np.testing.assert_array_equal(orc.student_t_copula(corr, 5.0, 100, seed=3),
                                  orc.student_t_copula(corr, 5.0, 100, seed=3))
