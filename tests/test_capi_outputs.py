"""Output-placement checks of the C ABI, without a GPU.

Entry points that fill a caller array take `where` (STOCHA_HOST or
STOCHA_DEVICE).  Any other value, and an output whose byte count does not fit
in 64 bits, is refused with STOCHA_EINVAL before any device work, so these
run on a machine without a GPU.  (Device pointers themselves are checked on
the GPU: tests/test_gpu_device_out.py.)
"""
import ctypes as C

import numpy as np
import pytest

from stocha_amd import _lib as L


def _call(fn, *args):
    eb = L.ErrBuf()
    rc = fn(*args, eb.buf, eb.n)
    return rc, eb.msg


@pytest.fixture(scope="module")
def lib():
    return L.load()


BUF = np.zeros(16)


@pytest.mark.parametrize("where", [2, -1, 7])
def test_bad_placement_is_refused(lib, where):
    p = BUF.ctypes.data
    st = np.zeros(2, np.uint64)
    inc = np.ones(2, np.uint64)
    new = np.zeros(2, np.uint64)
    pu = C.POINTER(C.c_uint64)
    corr = np.eye(2)
    calls = [
        (lib.stocha_gbm, (100.0, 0.05, 0.2, 1.0, 3, 2, 42, 0, 0.0, p, where)),
        (lib.stocha_heston, (100.0, 0.04, 0.05, 2.0, 0.04, 0.3, -0.7, 1.0, 3, 2, 42, b"euler", p,
                             where)),
        (lib.stocha_merton_jump_diffusion, (100.0, 0.05, 0.2, 1.0, -0.05, 0.1, 1.0, 3, 2, 42, p,
                                            where)),
        (lib.stocha_hull_white, (0.03, 0.1, 0.04, 0.01, 2.0, 3, 2, 42, p, where)),
        (lib.stocha_rng_uniform, (st.ctypes.data_as(pu), inc.ctypes.data_as(pu), 4, p, where,
                                  new.ctypes.data_as(pu))),
        (lib.stocha_rng_normal, (st.ctypes.data_as(pu), inc.ctypes.data_as(pu), 4, 0.0, 1.0, 0, p,
                                 where, new.ctypes.data_as(pu), None)),
        (lib.stocha_gaussian_copula, (corr.ctypes.data_as(C.POINTER(C.c_double)), 2, 2, 3, 42, p,
                                      where)),
        (lib.stocha_student_t_copula, (corr.ctypes.data_as(C.POINTER(C.c_double)), 2, 2, 5.0, 3,
                                       42, p, where)),
        (lib.stocha_terminal_prices, (0, np.array([100.0, 0.05, 0.2, 1.0]).ctypes.data_as(
            C.POINTER(C.c_double)), 4, None, 3, 2, 42, p, where)),
    ]
    for fn, args in calls:
        rc, msg = _call(fn, *args)
        assert rc == L.STOCHA_EINVAL, (fn.__name__, rc, msg)
        assert msg == f"where must be STOCHA_HOST (0) or STOCHA_DEVICE (1), got {where}", msg


def test_multi_gbm_bad_placement(lib):
    pd = C.POINTER(C.c_double)
    v = np.array([100.0, 50.0])
    m = np.array([0.05, 0.03])
    s = np.array([0.2, 0.3])
    corr = np.eye(2)
    rc, msg = _call(lib.stocha_multi_gbm, v.ctypes.data_as(pd), 2, m.ctypes.data_as(pd), 2,
                    s.ctypes.data_as(pd), 2, corr.ctypes.data_as(pd), 2, 2, 1.0, 3, 2, 42, 0,
                    BUF.ctypes.data, 3)
    assert rc == L.STOCHA_EINVAL and msg.startswith("where must be"), msg


@pytest.mark.parametrize("where", [L.STOCHA_HOST, L.STOCHA_DEVICE])
def test_output_size_overflow_is_refused(lib, where):
    p = BUF.ctypes.data
    # 2^61 paths x 11 columns x 8 bytes wraps a 64-bit byte count
    rc, msg = _call(lib.stocha_gbm, 100.0, 0.05, 0.2, 1.0, 10, 1 << 61, 42, 0, 0.0, p, where)
    assert rc == L.STOCHA_EINVAL and "does not fit in memory" in msg, msg
    rc, msg = _call(lib.stocha_heston, 100.0, 0.04, 0.05, 2.0, 0.04, 0.3, -0.7, 1.0,
                    (1 << 64) - 1, 2, 42, b"euler", p, where)
    assert rc == L.STOCHA_EINVAL and "too large" in msg, msg
    pd = C.POINTER(C.c_double)
    v, m, s = np.array([100.0]), np.array([0.05]), np.array([0.2])
    one = np.eye(1)
    rc, msg = _call(lib.stocha_multi_gbm, v.ctypes.data_as(pd), 1, m.ctypes.data_as(pd), 1,
                    s.ctypes.data_as(pd), 1, one.ctypes.data_as(pd), 1, 1, 1.0, 1 << 40, 1 << 30,
                    42, 0, p, where)
    assert rc == L.STOCHA_EINVAL and "does not fit in memory" in msg, msg
    st = np.zeros(2, np.uint64)
    pu = C.POINTER(C.c_uint64)
    rc, msg = _call(lib.stocha_rng_uniform, st.ctypes.data_as(pu), st.ctypes.data_as(pu), 1 << 62,
                    p, where, st.copy().ctypes.data_as(pu))
    assert rc == L.STOCHA_EINVAL and "does not fit in memory" in msg, msg
