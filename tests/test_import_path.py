"""The reference's import path (python/stocha/__init__.py, module-name
stocha._stocha in pyproject.toml:43-47) resolves to this build (CPU only)."""
import pytest

REFERENCE_ALL = [
    "RNG", "gbm", "sobol", "halton", "heston", "merton_jump_diffusion", "var_cvar",
    "gaussian_copula", "student_t_copula", "hull_white", "sabr_implied_vol", "sabr_calibrate",
    "multi_gbm", "lsmc_american_option", "greeks_fd", "greeks_pathwise", "heston_price",
    "heston_calibrate", "ssvi_calibrate", "ssvi_implied_vol", "ssvi_local_vol",
    "barrier_price", "asian_price", "lookback_price", "__version__",
]


def test_all_names_resolve():
    import stocha
    import stocha_amd
    assert stocha.__all__ == REFERENCE_ALL  # python/stocha/__init__.py:46-72
    for name in REFERENCE_ALL:
        assert getattr(stocha, name) is getattr(stocha_amd, name), name
    assert stocha.__version__ == "1.7.1"


def test_native_module_name():
    import stocha._stocha as native
    import stocha_amd._stocha as binding
    for name in ("gbm", "heston", "merton_jump_diffusion", "multi_gbm", "lsmc_american_option",
                 "greeks_pathwise", "greeks_fd", "RNG"):
        assert getattr(native, name) is getattr(binding, name), name


def test_errors_through_the_reference_path():
    import stocha
    with pytest.raises(ValueError, match="sigma must be positive"):
        stocha.gbm(s0=100.0, mu=0.05, sigma=0.0, t=1.0, steps=10, n_paths=10)
    with pytest.raises(NotImplementedError):
        stocha.sobol(2, 8)
