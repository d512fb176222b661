"""The reference's import path: ``import stocha`` (python/stocha/__init__.py)
with ``stocha._stocha`` as the native module (pyproject.toml:43-47), served by
this build's GPU path (stocha_amd): the same public names (``__all__`` below
is the reference's), signatures, defaults and errors.  The functions outside
the Monte Carlo path (sobol, halton, var_cvar, sabr_*, heston_price,
heston_calibrate, ssvi_*) raise NotImplementedError.  The GPU-only extras
(european_price, basket_price, lsmc_american_option_ex, batch, autograd) stay
under stocha_amd.
"""
from stocha_amd import (  # noqa: F401
    RNG,
    __version__,
    asian_price,
    barrier_price,
    gaussian_copula,
    gbm,
    greeks_fd,
    greeks_pathwise,
    halton,
    heston,
    heston_calibrate,
    heston_price,
    hull_white,
    lookback_price,
    lsmc_american_option,
    merton_jump_diffusion,
    multi_gbm,
    sabr_calibrate,
    sabr_implied_vol,
    sobol,
    ssvi_calibrate,
    ssvi_implied_vol,
    ssvi_local_vol,
    student_t_copula,
    var_cvar,
)

from . import _stocha  # noqa: F401,E402

__all__ = [
    "RNG", "gbm", "sobol", "halton", "heston", "merton_jump_diffusion", "var_cvar",
    "gaussian_copula", "student_t_copula", "hull_white", "sabr_implied_vol", "sabr_calibrate",
    "multi_gbm", "lsmc_american_option", "greeks_fd", "greeks_pathwise", "heston_price",
    "heston_calibrate", "ssvi_calibrate", "ssvi_implied_vol", "ssvi_local_vol",
    "barrier_price", "asian_price", "lookback_price", "__version__",
]
