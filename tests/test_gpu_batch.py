"""Batched persistent launch (GPU) against the oracle, per instance."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small_table(n=45, seed=1):
    from stocha_amd import batch
    inst = batch.config4_instances(n, seed=seed)
    for i in inst:  # keep the oracle quick
        i.n_paths = int(300 + (i.n_paths % 4000))
        i.steps = min(i.steps, 52)
    return inst


def _oracle_price(orc, i, seed):
    if i.model == 0:
        p = [i.s0, i.mu, i.sigma, 0.0]
    elif i.model == 1:
        p = [i.s0, i.v0, i.mu, i.kappa, i.theta, i.xi, i.rho]
    else:
        p = [i.s0, i.mu, i.sigma, i.lambda_, i.mu_j, i.sigma_j]
    term = orc.terminals(i.model, p, i.t, i.steps, i.n_paths, seed)
    s1, _ = orc.call_sums(term, [i.strike])
    return math.exp(-i.r * i.t) * s1[0] / i.n_paths


def test_batch_matches_oracle_per_instance(gpu, orc):
    from stocha_amd import batch
    inst = _small_table()
    plan = batch.Plan(inst, seed=42, shape=(128, 1, 0))
    price, se = plan.run()
    for k, i in enumerate(inst):
        ref = _oracle_price(orc, i, 42)
        if ref > 1e-2:
            assert abs(price[k] - ref) / ref < 1e-4, (k, i.model, price[k], ref)
        else:
            assert abs(price[k] - ref) < 1e-6, (k, price[k], ref)
        assert se[k] >= 0


def test_batch_equals_single_instance_fp32(gpu):
    # one instance of the batch sees the same paths as european_price(FP32)
    from stocha_amd import batch
    inst = _small_table(9, seed=3)
    price, _ = batch.Plan(inst, seed=42, shape=(64, 2, 0)).run()
    names = {0: ("gbm", ("s0", "mu", "sigma")),
             1: ("heston", ("s0", "v0", "mu", "kappa", "theta", "xi", "rho")),
             2: ("merton", ("s0", "mu", "sigma", "lambda_", "mu_j", "sigma_j"))}
    for k, i in enumerate(inst):
        m, keys = names[i.model]
        r = gpu.european_price(m, {n: getattr(i, n) for n in keys}, i.t, i.steps, i.n_paths,
                               [i.strike], r=i.r, seed=42, precision=32)
        assert price[k] == pytest.approx(r["price"][0], rel=1e-12, abs=1e-14)


def test_batch_graph_replays_are_identical(gpu):
    from stocha_amd import batch
    inst = _small_table(30, seed=5)
    plan = batch.Plan(inst, seed=42, shape=(256, 1, 2))
    a, _ = plan.run()
    # eager work in between replays
    gpu.european_price("gbm", dict(s0=100.0, mu=0.05, sigma=0.2), 1.0, 10, 5000, [100.0], r=0.05)
    b, _ = plan.run()
    gpu.gbm(s0=100.0, mu=0.05, sigma=0.2, t=1.0, steps=5, n_paths=100)
    c, _ = plan.run()
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(a, c)


@pytest.mark.parametrize("shape", [(64, 1, 0), (128, 2, 2), (256, 1, 0)])
def test_batch_shape_does_not_change_results(gpu, shape):
    from stocha_amd import batch
    inst = _small_table(24, seed=7)
    ref, _ = batch.Plan(inst, seed=42, shape=(128, 1, 0)).run()
    got, _ = batch.Plan(inst, seed=42, shape=shape).run()
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-14)


def test_autotuner_caches_choice(gpu, tmp_path, monkeypatch):
    from stocha_amd import batch
    monkeypatch.setenv("STOCHA_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    inst = _small_table(40, seed=9)
    shape = batch.tune(inst, candidates=[(64, 1, 0), (256, 1, 0)], reps=1)
    assert shape in [(64, 1, 0), (256, 1, 0)]
    assert (tmp_path / "tune.json").exists()
    assert batch.tune(inst, candidates=[(128, 2, 2)]) == shape  # served from the cache
