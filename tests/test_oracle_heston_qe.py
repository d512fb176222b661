"""Pins the oracle's Heston QE scheme (heston.rs:89-174) on the CPU.

QE draws V(t+dt) so that its conditional mean and variance are exactly
m = theta (1 - e) + V+ e and s2 = V+ xi^2 e (1 - e) / kappa +
theta xi^2 (1 - e)^2 / (2 kappa), e = exp(-kappa dt) (Andersen 2008, eqs.
17-18); the quadratic branch (psi <= 1.5) and the exponential branch
(psi > 1.5, a point mass p = (psi - 1)/(psi + 1) at zero) are checked
separately on one step from a fixed V(0), so a slip in b^2, a, p or beta
moves a moment.  With rho = 0 the log-price update is an exact martingale
step (k1 = k2 = -dt/4, k3 = k4 = dt/2: the drift cancels half the variance
of the normal term), so E[S_T] = s0 e^{mu T} up to Monte Carlo error -- a
dropped 0.5 or a wrong sign in the k's fails it.  Also the reference's own
checks (tests/test_models.py:137-179).
"""
import math

import numpy as np
import pytest


def _moments(v0, theta, kappa, xi, dt):
    e = math.exp(-kappa * dt)
    m = theta * (1 - e) + v0 * e
    s2 = v0 * xi * xi * e * (1 - e) / kappa + theta * xi * xi * (1 - e) ** 2 / (2 * kappa)
    return m, s2


@pytest.fixture(scope="module")
def z_moments(orc):
    # E[Z^2], E[Z^4] of the reference's ziggurat (its normals have standard
    # deviation 0.99278, not 1: a property of its tables, kept on purpose)
    z, _ = orc.normals(42, 7656, 0, 2_000_000)
    return float(np.mean(z * z)), float(np.mean(z ** 4))


@pytest.mark.parametrize("v0,theta,kappa,xi,dt", [
    (0.04, 0.04, 2.0, 0.3, 1 / 252),
    (0.09, 0.04, 1.5, 0.8, 1 / 12),
])
def test_qe_quadratic_step_matches_conditional_moments(orc, z_moments, v0, theta, kappa, xi,
                                                       dt):
    m, s2 = _moments(v0, theta, kappa, xi, dt)
    psi = s2 / (m * m)
    assert psi <= 1.5
    # Andersen's quadratic branch V = a (b + Z)^2 with b^2 = 2/psi - 1 +
    # sqrt(2/psi (2/psi - 1)), a = m / (1 + b^2) matches m and s2 for a unit
    # normal: E V = a (b^2 + 1), Var V = a^2 (4 b^2 + 2)
    b2 = 2 / psi - 1 + math.sqrt(2 / psi * (2 / psi - 1))
    a = m / (1 + b2)
    assert a * (b2 + 1) == pytest.approx(m, rel=1e-12)
    assert a * a * (4 * b2 + 2) == pytest.approx(s2, rel=1e-9)
    # sampled with the reference's normals (E Z^2 = mu2, E Z^4 = mu4, odd
    # moments 0): E V = a (b^2 + mu2), Var V = a^2 (4 b^2 mu2 + mu4 - mu2^2)
    mu2, mu4 = z_moments
    n = 1_000_000
    v = orc.heston_qe_variance(100.0, v0, 0.05, kappa, theta, xi, -0.7, dt, 1, n, seed=17)
    assert np.all(v[:, 0] == v0)
    v1 = v[:, 1]
    assert np.all(v1 > 0.0)
    mean_want = a * (b2 + mu2)
    var_want = a * a * (4 * b2 * mu2 + mu4 - mu2 * mu2)
    assert abs(v1.mean() - mean_want) < 5 * math.sqrt(var_want / n)
    dev2 = (v1 - v1.mean()) ** 2
    assert abs(dev2.mean() - var_want) < 5 * dev2.std() / math.sqrt(n)


def test_qe_exponential_step_matches_conditional_moments(orc):
    # psi > 1.5: V = 0 with probability p = (psi-1)/(psi+1), else an
    # exponential of rate beta = (1-p)/m; the draw is a uniform, so the
    # moments are m and s2 exactly
    v0, theta, kappa, xi, dt = 1e-4, 1e-4, 1.0, 0.5, 1 / 52
    m, s2 = _moments(v0, theta, kappa, xi, dt)
    psi = s2 / (m * m)
    assert psi > 1.5
    n = 1_000_000
    v1 = orc.heston_qe_variance(100.0, v0, 0.05, kappa, theta, xi, -0.7, dt, 1, n, seed=17)[:, 1]
    assert abs(v1.mean() - m) < 5 * math.sqrt(s2 / n)
    dev2 = (v1 - m) ** 2
    assert abs(dev2.mean() - s2) < 5 * dev2.std() / math.sqrt(n)
    p = (psi - 1) / (psi + 1)
    assert abs(np.mean(v1 == 0.0) - p) < 5 * math.sqrt(p * (1 - p) / n)


def test_qe_rho_zero_is_a_martingale(orc):
    n, steps, mu, t = 200_000, 50, 0.05, 1.0
    paths = orc.heston_qe(100.0, 0.04, mu, 2.0, 0.04, 0.5, 0.0, t, steps, n, seed=5)
    st = paths[:, -1]
    se = st.std() / math.sqrt(n)
    assert abs(st.mean() - 100.0 * math.exp(mu * t)) < 4 * se


def test_qe_reference_checks(orc):
    # tests/test_models.py:137-179
    kw = dict(s0=100.0, v0=0.04, mu=0.05, kappa=2.0, theta=0.04, xi=0.3, rho=-0.7, t=1.0,
              steps=252)
    a = orc.heston_qe(**kw, n_paths=1000, seed=42)
    assert a.shape == (1000, 253)
    assert (a > 0).all()
    np.testing.assert_array_equal(a, orc.heston_qe(**kw, n_paths=1000, seed=42))
    big = orc.heston_qe(**kw, n_paths=100_000, seed=42)
    assert big[:, -1].mean() == pytest.approx(100.0 * math.exp(0.05), rel=2e-2)
    feller = orc.heston_qe(**{**kw, "kappa": 1.0}, n_paths=5000, seed=77)
    assert (feller > 0).all()
    # the QE stream uses the same block-split layout as Euler: path 0's first
    # step takes one normal for V (quadratic branch here) then z1
    assert np.all(a[:, 0] == 100.0)
