"""GPU parity of hull_white (hull_white.rs:26-52) and gaussian_copula
(copula/gaussian.rs:15-73) against the oracle.

Hull-White steps are additions and products of the path's normals, so the
paths are the oracle's to 1e-15 (tail normals go through log: <= 2 ulp).
The copula's normals come from ONE stream decoded in parallel (the RNG
class's machinery), then L z and the A&S normal cdf (one exp): agreement to
1e-15 absolute.  Also the reference's own Hull-White checks
(hull_white.rs:58-120) and copula shapes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,steps", [(1000, 252), (257, 12), (3, 1)])
def test_hull_white_matches_oracle(gpu, orc, n, steps):
    kw = dict(r0=-0.01, a=0.5, theta=0.02, sigma=0.005, t=5.0, steps=steps, n_paths=n, seed=7)
    dev = gpu.hull_white(**kw)
    ref = orc.hull_white(kw["r0"], kw["a"], kw["theta"], kw["sigma"], kw["t"], steps, n, seed=7)
    np.testing.assert_allclose(dev, ref, rtol=1e-12, atol=1e-15)
    assert dev.shape == (n, steps + 1) and np.all(dev[:, 0] == -0.01)


@pytest.mark.parametrize("d,n", [(1, 10), (3, 5000), (8, 40_000), (5, 0)])
def test_gaussian_copula_matches_oracle(gpu, orc, d, n):
    rng = np.random.default_rng(d)
    a = rng.standard_normal((d, d))
    c = a @ a.T + d * np.eye(d)
    s = np.sqrt(np.diag(c))
    corr = c / np.outer(s, s)
    np.fill_diagonal(corr, 1.0)
    corr = 0.5 * (corr + corr.T)
    dev = gpu.gaussian_copula(corr, n, seed=11)
    ref = orc.gaussian_copula(corr, n, seed=11)
    assert dev.shape == (n, d)
    np.testing.assert_allclose(dev, ref, rtol=0, atol=1e-15)


def test_reference_checks(gpu):
    p = gpu.hull_white(r0=0.05, a=0.1, theta=0.05, sigma=0.01, t=1.0, steps=12, n_paths=100)
    assert p.shape == (100, 13) and abs(p[0, 0] - 0.05) < 1e-12
    q = gpu.hull_white(r0=0.10, a=1.0, theta=0.05, sigma=0.001, t=10.0, steps=100,
                       n_paths=1000, seed=99)
    assert abs(q[:, 100].mean() - 0.05) < 0.01
    u = gpu.gaussian_copula(np.array([[1.0, 0.8], [0.8, 1.0]]), n_samples=1000)
    assert u.shape == (1000, 2) and np.all((u > 0) & (u < 1))
