"""``stocha._stocha``: the name of the reference's native module
(pyproject.toml:43-47; the PyO3 functions of src/lib.rs), here the binding of
libstocha_gpu.so (stocha_amd._stocha)."""
from stocha_amd._stocha import *  # noqa: F401,F403
from stocha_amd._stocha import (  # noqa: F401
    RNG,
    __version__,
    asian_price,
    barrier_price,
    gaussian_copula,
    gbm,
    greeks_fd,
    greeks_pathwise,
    heston,
    hull_white,
    lookback_price,
    lsmc_american_option,
    merton_jump_diffusion,
    multi_gbm,
    student_t_copula,
)
