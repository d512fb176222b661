"""GPU parity of the RNG class's bulk draws against the oracle's sequential
stream (src/lib.rs:108-145).

Uniforms are bit-exact (integer work).  Normals go through the parallel
chunk decoding (speculative chains resolved to the true one); they must be
the reference's values in the reference's order -- bit-exact, except tail
values, which pass through log (device libm vs glibc, <= 2 ulp) -- and the
generator must end where the sequential one ends, which checks the number of
draws consumed exactly.  Sizes cover empty, one value, chunk edges (64
positions) and millions; the draws interleave across calls as the
reference's do.  Also the reference's own checks (tests/test_rng.py:19-112).
"""
import numpy as np
import pytest
from scipy import stats

pytestmark = pytest.mark.gpu

R_ZIG = 3.654152885361008796


def _check_normals(dev, ref):
    assert dev.shape == ref.shape
    tail = np.abs(ref) > R_ZIG
    np.testing.assert_array_equal(dev[~tail], ref[~tail])
    if tail.any():
        ulps = np.abs(dev[tail] - ref[tail]) / np.spacing(np.abs(ref[tail]))
        assert ulps.max() <= 2


@pytest.mark.parametrize("seed", [0, 42, 2**128 - 1])
@pytest.mark.parametrize("n", [1, 2, 15, 16, 17, 4095, 4096, 4097, 100_003])
def test_uniform_bit_exact(gpu, orc, seed, n):
    r, o = gpu.RNG(seed), orc.Rng(seed)
    np.testing.assert_array_equal(r.uniform(n), o.uniform(n))
    np.testing.assert_array_equal(r.uniform(7), o.uniform(7))  # the position moved by n
    assert (r._state, r._inc) == o.state()


@pytest.mark.parametrize("seed", [0, 7, 2**100 + 3])
@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 127, 1000, 100_000, 2_000_000])
def test_standard_normal_matches_sequential_stream(gpu, orc, seed, n):
    print(f"rng normals seed={seed} n={n}", flush=True)
    r, o = gpu.RNG(seed), orc.Rng(seed)
    _check_normals(r.standard_normal(n), o.normal(n))
    assert r._state == o.state()[0]  # consumed exactly the same draws
    _check_normals(r.standard_normal(33), o.normal(33))
    assert r._state == o.state()[0]


def test_interleaved_calls_and_scaling(gpu, orc):
    r, o = gpu.RNG(2024), orc.Rng(2024)
    np.testing.assert_array_equal(r.uniform(10), o.uniform(10))
    _check_normals(r.normal(5000, loc=1.5, scale=0.25),
                   o.normal(5000, loc=1.5, scale=0.25, scaled=True))
    np.testing.assert_array_equal(r.uniform(3), o.uniform(3))
    _check_normals(r.standard_normal(77), o.normal(77))
    _check_normals(r.normal(1), o.normal(1, scaled=True))
    assert r._state == o.state()[0]


def test_reference_rng_checks(gpu):
    u = gpu.RNG(seed=0).uniform(size=10_000)
    assert u.dtype == np.float64 and u.min() >= 0.0 and u.max() < 1.0
    x = gpu.RNG(seed=42).normal(size=1_000_000)
    assert abs(x.mean()) < 0.01 and abs(x.std() - 1.0) < 0.01
    y = gpu.RNG(seed=1).normal(size=100_000, loc=5.0, scale=2.0)
    assert y.mean() == pytest.approx(5.0, rel=1e-2) and y.std() == pytest.approx(2.0, rel=1e-2)
    a = gpu.RNG(seed=99).normal(size=1_000)
    np.testing.assert_array_equal(a, gpu.RNG(seed=99).normal(size=1_000))
    assert not np.array_equal(gpu.RNG(seed=1).normal(size=100), gpu.RNG(seed=2).normal(size=100))
    rng = gpu.RNG(seed=42)
    _ = rng.normal(size=500)
    state = rng.save_state()
    expected = rng.normal(size=100)
    np.testing.assert_array_equal(gpu.RNG.from_state(state).normal(size=100), expected)
    np.testing.assert_array_equal(gpu.RNG.from_state('{"seed":42}').normal(size=50),
                                  gpu.RNG(seed=42).normal(size=50))
    z = gpu.RNG(seed=0).standard_normal(size=500_000)
    assert abs(z.mean()) < 0.005 and abs(z.std() - 1.0) < 0.01
    _, p = stats.kstest(gpu.RNG(seed=123).standard_normal(size=100_000), "norm")
    assert p > 0.01


def test_standard_normal_many_scan_tiles(gpu, orc):
    # ~1.1M decoding chunks: the device prefix sum of the chunk counts has
    # more than one tile (1024) of tile sums, so its carry across tiles is used
    n = 70_000_000
    print(f"rng normals n={n}", flush=True)
    r, o = gpu.RNG(31), orc.Rng(31)
    _check_normals(r.standard_normal(n), o.normal(n))
    assert r._state == o.state()[0]
