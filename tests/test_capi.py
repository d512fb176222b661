"""The C ABI without a GPU: the library loads, exports every entry point that
include/stocha_gpu.h declares, and rejects bad arguments with the reference's
exception types and messages (src/lib.rs) before any device work."""
import ctypes
import os
import re

import numpy as np
import pytest

import stocha_amd as S
from stocha_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"\b(stocha_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _declared()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(_lib.exported_names())


def test_version_string():
    assert b"stocha" in _lib.load().stocha_version()
    assert S.__version__ == "1.7.1"


GBM = dict(s0=100.0, mu=0.05, sigma=0.2, t=1.0, steps=10, n_paths=10)


@pytest.mark.parametrize("bad,msg", [
    (dict(s0=-1.0), "s0 must be positive"),
    (dict(s0=0.0), "s0 must be positive"),
    (dict(sigma=0.0), "sigma must be positive"),
    (dict(t=-1.0), "t must be positive"),
    (dict(steps=0), "steps and n_paths must be positive"),
    (dict(n_paths=0), "steps and n_paths must be positive"),
])
def test_gbm_validation(bad, msg):
    with pytest.raises(ValueError, match=re.escape(msg)):
        S.gbm(**{**GBM, **bad})


def test_gbm_validation_order():
    # s0 is checked before sigma before t (src/lib.rs:219-238)
    with pytest.raises(ValueError, match="s0 must be positive"):
        S.gbm(s0=-1.0, mu=0.05, sigma=-1.0, t=-1.0, steps=0, n_paths=0)
    with pytest.raises(ValueError, match="sigma must be positive"):
        S.gbm(s0=1.0, mu=0.05, sigma=-1.0, t=-1.0, steps=0, n_paths=0)


def test_integer_extraction_like_pyo3():
    with pytest.raises(OverflowError):
        S.gbm(**{**GBM, "steps": -1})
    with pytest.raises(OverflowError):
        S.gbm(**{**GBM, "seed": -5})
    with pytest.raises(OverflowError):
        S.gbm(**{**GBM, "seed": 2**64})
    with pytest.raises(TypeError):
        S.gbm(**{**GBM, "steps": 10.0})
    with pytest.raises(TypeError):
        S.gbm(**{**GBM, "antithetic": 1})
    with pytest.raises(TypeError):
        S.gbm(**{**GBM, "s0": "100"})


HESTON = dict(s0=100.0, v0=0.04, mu=0.05, kappa=2.0, theta=0.04, xi=0.3, rho=-0.7, t=1.0,
              steps=10, n_paths=10)


@pytest.mark.parametrize("bad,msg", [
    (dict(s0=-1.0), "s0 must be positive"),
    (dict(v0=-0.01), "v0 must be non-negative"),
    (dict(kappa=0.0), "kappa must be positive"),
    (dict(theta=0.0), "theta must be positive"),
    (dict(xi=0.0), "xi must be positive"),
    (dict(rho=2.0), "rho must be in [-1, 1]"),
    (dict(rho=-1.5), "rho must be in [-1, 1]"),
    (dict(t=0.0), "t must be positive"),
    (dict(steps=0), "steps and n_paths must be positive"),
    (dict(scheme="bad"), "scheme must be 'euler' or 'qe'"),
])
def test_heston_validation(bad, msg):
    with pytest.raises(ValueError, match=re.escape(msg)):
        S.heston(**{**HESTON, **bad})


def test_heston_qe_validates_like_euler():
    # scheme "qe" takes the same checks, in the same order (lib.rs:344-392)
    with pytest.raises(ValueError, match="rho must be in"):
        S.heston(**{**HESTON, "rho": 2.0}, scheme="qe")
    with pytest.raises(ValueError, match="scheme must be 'euler' or 'qe'"):
        S.heston(**HESTON, scheme="QE")


MERTON = dict(s0=100.0, mu=0.05, sigma=0.2, lambda_=1.0, mu_j=-0.05, sigma_j=0.1, t=1.0,
              steps=10, n_paths=10)


@pytest.mark.parametrize("bad,msg", [
    (dict(s0=-1.0), "s0 must be positive"),
    (dict(sigma=0.0), "sigma must be positive"),
    (dict(lambda_=-1.0), "lambda_ must be non-negative"),
    (dict(sigma_j=-0.1), "sigma_j must be non-negative"),
    (dict(t=0.0), "t must be positive"),
    (dict(n_paths=0), "steps and n_paths must be positive"),
])
def test_merton_validation(bad, msg):
    with pytest.raises(ValueError, match=re.escape(msg)):
        S.merton_jump_diffusion(**{**MERTON, **bad})


MULTI = dict(s0=[100.0, 50.0], mu=[0.05, 0.08], sigma=[0.2, 0.3],
             corr=np.array([[1.0, 0.6], [0.6, 1.0]]), t=1.0, steps=10, n_paths=10)


@pytest.mark.parametrize("bad,msg", [
    (dict(s0=[], mu=[], sigma=[], corr=np.eye(0)), "s0 must not be empty"),
    (dict(mu=[0.05]), "s0, mu, sigma must have the same length"),
    (dict(s0=[-1.0, 50.0]), "all s0 must be positive"),
    (dict(sigma=[0.2, 0.0]), "all sigma must be positive"),
    (dict(t=0.0), "t must be positive"),
    (dict(steps=0), "steps and n_paths must be positive"),
    (dict(corr=np.eye(3)), "corr must be (2, 2), got (3, 3)"),
    (dict(corr=np.array([[1.0, 1.5], [1.5, 1.0]])),
     "matrix is not positive-definite (diagonal element 1 = -1.25)"),
])
def test_multi_gbm_validation(bad, msg):
    with pytest.raises(ValueError, match=re.escape(msg)):
        S.multi_gbm(**{**MULTI, **bad})


def test_cholesky_message_formats_like_rust():
    # Rust's Display for f64: shortest round-trip digits, never an exponent
    pivot = 1.0 - 1.00000005 * 1.00000005  # copula/gaussian.rs:82-84
    txt = np.format_float_positional(pivot)
    assert txt == "-0.00000010000000227883277"
    with pytest.raises(ValueError, match=re.escape(f"(diagonal element 1 = {txt})")):
        S.multi_gbm(**{**MULTI, "corr": np.array([[1.0, 1.00000005], [1.00000005, 1.0]])})


LSMC = dict(s0=100.0, k=100.0, r=0.05, sigma=0.2, t=1.0, steps=10, n_paths=100)


@pytest.mark.parametrize("bad,msg", [
    (dict(s0=-1.0), "s0 must be positive"),
    (dict(k=0.0), "k must be positive"),
    (dict(sigma=0.0), "sigma must be positive"),
    (dict(t=0.0), "t must be positive"),
    (dict(steps=0), "steps and n_paths must be positive"),
    (dict(n_paths=1), "n_paths must be at least 2"),
    (dict(poly_degree=0), "poly_degree must be in [1, 4]"),
    (dict(poly_degree=5), "poly_degree must be in [1, 4]"),
])
def test_lsmc_validation(bad, msg):
    with pytest.raises(ValueError, match=re.escape(msg)):
        S.lsmc_american_option(**{**LSMC, **bad})


GP = dict(s0=100.0, r=0.05, sigma=0.2, t=1.0, strike=100.0, is_call=True, n_paths=1000,
          n_steps=10)


@pytest.mark.parametrize("bad,greeks,msg", [
    (dict(s0=-1.0), ["delta"], "s0 must be positive"),
    (dict(sigma=0.0), ["delta"], "sigma must be positive"),
    (dict(t=0.0), ["delta"], "t must be positive"),
    (dict(n_paths=0), ["delta"], "n_paths and n_steps must be positive"),
    (dict(), ["gamma"], "Pathwise method only supports 'delta' and 'vega', got 'gamma'"),
    (dict(), ["delta", "rho"], "Pathwise method only supports 'delta' and 'vega', got 'rho'"),
    (dict(), ["bogus"], "Unknown greek: 'bogus'. Supported: delta, gamma, vega, theta, rho"),
    (dict(), ["theta", "bogus"], "Unknown greek: 'bogus'"),
])
def test_greeks_pathwise_validation(bad, greeks, msg):
    with pytest.raises(ValueError, match=re.escape(msg)):
        S.greeks_pathwise(**{**GP, **bad}, greeks=greeks)


def test_out_of_scope_functions_say_so():
    for name in ("sobol", "halton", "sabr_implied_vol"):
        with pytest.raises(NotImplementedError, match="hot path"):
            getattr(S, name)()


TC = dict(corr=np.array([[1.0, 0.5], [0.5, 1.0]]), nu=5.0, n_samples=10)


@pytest.mark.parametrize("bad,msg", [
    (dict(nu=2.0), "nu must be > 2"),
    (dict(nu=1.0), "nu must be > 2"),
    (dict(nu=float("nan")), "nu must be > 2"),
    (dict(corr=np.ones((2, 3))), "correlation matrix must be square"),
    (dict(corr=np.array([[1.0, 1.5], [1.5, 1.0]])),
     "matrix is not positive-definite (diagonal element 1 = -1.25)"),
])
def test_student_t_copula_validation(bad, msg):
    # lib.rs:570-572 checks nu first, then student_t.rs:146-153 and the
    # Cholesky of gaussian.rs:75-95; no diagonal or symmetry check here
    with pytest.raises(ValueError, match=re.escape(msg)):
        S.student_t_copula(**{**TC, **bad})


def test_student_t_copula_takes_lists_like_the_wrapper():
    # python/stocha/__init__.py:446-448 converts corr with np.asarray; an
    # invalid nu is still reported before any device work
    with pytest.raises(ValueError, match="nu must be > 2"):
        S.student_t_copula([[1.0, 0.5], [0.5, 1.0]], nu=2.0, n_samples=3)


def test_no_device_is_a_loud_error():
    # On a machine without a GPU the product raises; it never falls back.
    if S.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(RuntimeError, match="no HIP device"):
        S.gbm(**GBM)


def test_european_price_argument_checks():
    p = dict(s0=100.0, mu=0.05, sigma=0.2)
    with pytest.raises(ValueError, match="Unknown model"):
        S.european_price("bad", p, 1.0, 10, 10, [100.0], r=0.05)
    with pytest.raises(ValueError, match="n_strikes"):
        S.european_price("gbm", p, 1.0, 10, 10, [100.0] * 9, r=0.05)
    with pytest.raises(ValueError, match="precision"):
        S.european_price("gbm", p, 1.0, 10, 10, [100.0], r=0.05, precision=16)


def test_stream_probe_signature():
    # declared argtypes match the header's parameter counts
    lib = _lib.load()
    assert len(lib.stocha_stream_normals.argtypes) == 10
    assert lib.stocha_stream_u64.restype is ctypes.c_int
