"""GPU parity of student_t_copula (copula/student_t.rs:135-169) against the
oracle's sequential restatement.

The device decodes ONE stream in parallel: every position's sample length,
chunk maps composed by a scan, then the samples.  A wrong start anywhere
shifts every later sample, so agreement on all rows at once checks the
decoding end to end; the values then differ from the oracle only by the
device's log/exp in the t CDF (1e-15 relative; the continued fraction stops
at 3e-7, so a stop one step apart would move a value by ~1e-7: allowed for
no element here).  Also the pointer-doubling path for samples longer than a
chunk's 64 draws (d >= 60), empty requests, and the reference's own properties
(tests/test_risk.py:95-136) at a size the oracle does not run.
"""
import numpy as np
import pytest
from scipy import stats

pytestmark = pytest.mark.gpu


def _corr(d, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((d, d))
    c = a @ a.T + d * np.eye(d)
    s = np.sqrt(np.diag(c))
    c = c / s[:, None] / s[None, :]
    np.fill_diagonal(c, 1.0)
    return (c + c.T) / 2


@pytest.mark.parametrize("d,n,nu,seed", [(1, 1, 5.0, 42), (2, 500, 5.0, 3), (3, 7, 2.5, 0),
                                         (4, 20_000, 3.0, 11), (8, 3_000, 30.0, 7),
                                         (2, 40_000, 2.000001, 5)])
def test_matches_oracle(gpu, orc, d, n, nu, seed):
    corr = _corr(d, seed) if d > 1 else np.eye(1)
    dev = gpu.student_t_copula(corr, nu=nu, n_samples=n, seed=seed)
    ref = orc.student_t_copula(corr, nu, n, seed=seed)
    assert dev.shape == (n, d)
    np.testing.assert_allclose(dev, ref, rtol=0, atol=1e-12)


# d >= 60: every sample takes > 64 draws, beyond the chunk maps, so the
# starts come from pointer doubling (jumps of 64 and 4096 samples): 300
# samples use one top-level jump chain entry, 9000 three, and 100,000 at
# d = 64 is the size where a serial walk of the starts took ~0.1 s
@pytest.mark.parametrize("d,n,nu,seed", [(60, 300, 6.0, 9), (61, 9_000, 4.0, 2),
                                         (200, 700, 7.5, 4)])
def test_long_samples_take_pointer_doubling(gpu, orc, d, n, nu, seed):
    corr = _corr(d, seed)
    dev = gpu.student_t_copula(corr, nu=nu, n_samples=n, seed=seed)
    ref = orc.student_t_copula(corr, nu, n, seed=seed)
    np.testing.assert_allclose(dev, ref, rtol=0, atol=1e-12)


def test_long_samples_d64_100k_timed(gpu, orc):
    import time
    corr = _corr(64, 12)
    gpu.student_t_copula(corr, nu=5.0, n_samples=1000, seed=1)  # workspace warm-up
    t0 = time.perf_counter()
    dev = gpu.student_t_copula(corr, nu=5.0, n_samples=100_000, seed=1)
    dt = time.perf_counter() - t0
    ref = orc.student_t_copula(corr, 5.0, 100_000, seed=1)
    np.testing.assert_allclose(dev, ref, rtol=0, atol=1e-12)
    print(f"student_t_copula d=64 100k: {dt * 1e3:.2f} ms")
    assert dt < 0.05, dt


def test_empty(gpu):
    corr = np.array([[1.0, 0.5], [0.5, 1.0]])
    assert gpu.student_t_copula(corr, nu=5.0, n_samples=0).shape == (0, 2)
    assert gpu.student_t_copula(np.zeros((0, 0)), nu=5.0, n_samples=4).shape == (4, 0)


def test_reference_properties_large(gpu):
    # tests/test_risk.py:95-136 at 1M samples: (0, 1), uniform marginals,
    # heavier joint tails than the Gaussian copula, reproducible
    corr = np.array([[1.0, 0.8], [0.8, 1.0]])
    u = gpu.student_t_copula(corr, nu=3.0, n_samples=1_000_000, seed=0)
    g = gpu.gaussian_copula(corr, n_samples=1_000_000, seed=0)
    assert np.all(u > 0) and np.all(u < 1)
    for j in range(2):
        assert stats.kstest(u[::10, j], "uniform").pvalue > 1e-4
    jt = np.mean((u[:, 0] > 0.95) & (u[:, 1] > 0.95))
    jg = np.mean((g[:, 0] > 0.95) & (g[:, 1] > 0.95))
    assert jt > jg
    tau = stats.kendalltau(u[:20_000, 0], u[:20_000, 1]).statistic
    assert tau == pytest.approx(2 / np.pi * np.arcsin(0.8), abs=0.02)
    np.testing.assert_array_equal(u[:1000], gpu.student_t_copula(corr, nu=3.0, n_samples=1000,
                                                                  seed=0))
