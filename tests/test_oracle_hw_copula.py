"""Pins the oracle's Hull-White paths (hull_white.rs:26-52) and Gaussian
copula (copula/gaussian.rs:15-73) on the CPU, and the bindings' argument
checks (src/lib.rs:538-658), which run before any device work.

Hull-White: each step is r' = r e + m + v z with the path's own normal z, so
(r' - r e - m) / v recovers the normals of the path's stream -- equal to the
oracle's ziggurat normals for block (steps + 1024) * 4 to rounding: this
pins the recursion, its constants and the block size at once; the terminal
mean is the exact transition's mean.  Copula: the uniforms are Phi(L z) with
z the stream's normals in order; Phi (Abramowitz & Stegun 26.2.17) is within
its published 7.5e-8 of scipy's normal cdf, and the rows recover L z.
"""
import math

import numpy as np
import pytest
from scipy.stats import norm

import stocha_amd as S


def test_hull_white_recovers_the_stream_normals(orc):
    r0, a, theta, sigma, t, steps, n = 0.03, 0.4, 0.02, 0.015, 2.0, 24, 40
    p = orc.hull_white(r0, a, theta, sigma, t, steps, n, seed=9)
    dt = t / steps
    e = math.exp(-a * dt)
    m = theta / a * (1 - e)
    v = sigma * math.sqrt((1 - e * e) / (2 * a))
    assert np.all(p[:, 0] == r0)
    z = (p[:, 1:] - p[:, :-1] * e - m) / v
    for i in (0, 1, n - 1):
        want, _ = orc.normals(9, (steps + 1024) * 4, i, steps)
        np.testing.assert_allclose(z[i], want, rtol=0, atol=1e-9)


def test_hull_white_terminal_mean(orc):
    # hull_white.rs:82-98 and the exact transition's mean
    r0, a, theta, sigma, t = 0.10, 1.0, 0.05, 0.001, 10.0
    p = orc.hull_white(r0, a, theta, sigma, t, 100, 1000, seed=99)
    want = r0 * math.exp(-a * t) + theta / a * (1 - math.exp(-a * t))
    assert abs(p[:, -1].mean() - want) < 1e-4


def test_gaussian_copula_rows_are_phi_of_l_z(orc):
    corr = np.array([[1.0, 0.6, -0.2], [0.6, 1.0, 0.3], [-0.2, 0.3, 1.0]])
    n = 2000
    u = orc.gaussian_copula(corr, n, seed=5)
    rng = orc.Rng(5)
    z = rng.normal(n * 3).reshape(n, 3)
    L = orc.cholesky(corr)
    x = np.einsum("jk,ik->ij", L, z)
    assert np.max(np.abs(u - norm.cdf(x))) < 7.5e-8
    assert np.all((u > 0) & (u < 1))
    emp = np.corrcoef(norm.ppf(u).T)
    assert np.max(np.abs(emp - corr)) < 0.06


def test_hw_and_copula_validation():
    with pytest.raises(ValueError, match="a must be positive"):
        S.hull_white(r0=0.05, a=0.0, theta=0.05, sigma=0.01, t=1.0, steps=12, n_paths=10)
    with pytest.raises(ValueError, match="sigma must be positive"):
        S.hull_white(r0=0.05, a=0.1, theta=0.05, sigma=-0.01, t=1.0, steps=12, n_paths=10)
    with pytest.raises(ValueError, match="steps and n_paths must be positive"):
        S.hull_white(r0=0.05, a=0.1, theta=0.05, sigma=0.01, t=1.0, steps=0, n_paths=10)
    with pytest.raises(ValueError, match="correlation matrix must be square"):
        S.gaussian_copula(np.eye(2)[:, :1].copy(), 10)
    with pytest.raises(ValueError, match=r"diagonal element \[1\]\[1\] must be 1.0, got 2"):
        S.gaussian_copula(np.array([[1.0, 0.0], [0.0, 2.0]]), 10)
    with pytest.raises(ValueError, match=r"not symmetric: \[0,1\]=0.5 vs \[1,0\]=0.4"):
        S.gaussian_copula(np.array([[1.0, 0.5], [0.4, 1.0]]), 10)
    with pytest.raises(ValueError, match="not positive-definite"):
        S.gaussian_copula(np.array([[1.0, 1.5], [1.5, 1.0]]), 10)
    # the extension function takes float64 arrays only (PyReadonlyArray2<f64>,
    # lib.rs:540); the package wrapper converts first (np.asarray, as
    # python/stocha/__init__.py:421-423 does), so a float32 matrix gets as far
    # as the checks
    from stocha_amd import _stocha
    with pytest.raises(TypeError):
        _stocha.gaussian_copula(np.eye(2, dtype=np.float32), 10)
    with pytest.raises(ValueError, match="not positive-definite"):
        S.gaussian_copula(np.array([[1.0, 1.5], [1.5, 1.0]], dtype=np.float32), 10)
