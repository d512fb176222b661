"""stocha_european_fused_noalloc (device outputs, partials in a static device
array) against stocha_european_fused, through the C ABI.

Both run the same kernels with the same grid and the same fixed-order
reduction, so the sums must be the same bits and the prices and standard
errors the same values.  Two calls in a row, the second with a smaller
input, check that the shared scratch carries nothing over; guard words
around the device outputs check that nothing is written outside them.
"""
import ctypes as C

import numpy as np
import pytest

from stocha_amd import _lib as L

pytestmark = pytest.mark.gpu

_pd = C.POINTER(C.c_double)
HESTON = [100.0, 0.04, 0.05, 2.0, 0.04, 0.3, -0.7]
GBM = [100.0, 0.05, 0.2]
MERTON = [100.0, 0.05, 0.2, 1.0, -0.05, 0.1]
GUARD = 32


def _reference(model, params, t, steps, n, strikes, r, seed, prec):
    lib, eb = L.load(), L.ErrBuf()
    p = np.array(params, dtype=np.float64)
    k = np.array(strikes, dtype=np.float64)
    ns = len(k)
    price, se, sums = np.zeros(ns), np.zeros(ns), np.zeros(2 * ns)
    rc = lib.stocha_european_fused(model, p.ctypes.data_as(_pd), len(p), t, steps, n, seed, prec,
                                   L.STOCHA_SCHEDULE_FEED, k.ctypes.data_as(_pd), ns, r,
                                   price.ctypes.data_as(_pd), se.ctypes.data_as(_pd),
                                   sums.ctypes.data_as(_pd), eb.buf, eb.n)
    L.check(rc, eb)
    return price, se, sums


def _noalloc(model, params, t, steps, n, strikes, r, seed, prec, with_sums=True):
    import torch
    lib, eb = L.load(), L.ErrBuf()
    p = np.array(params, dtype=np.float64)
    k = np.array(strikes, dtype=np.float64)
    ns = len(k)
    # one device buffer: guard | price | guard | std_err | guard | sums | guard
    sizes = [GUARD, ns, GUARD, ns, GUARD, 2 * ns, GUARD]
    offs = np.cumsum([0] + sizes)
    buf = torch.full((int(offs[-1]),), -12345.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    base = buf.data_ptr()
    ptr = [base + 8 * int(o) for o in offs]
    rc = lib.stocha_european_fused_noalloc(model, p.ctypes.data_as(_pd), len(p), t, steps, n,
                                           seed, prec, L.STOCHA_SCHEDULE_FEED,
                                           k.ctypes.data_as(_pd), ns, r, ptr[1], ptr[3],
                                           ptr[5] if with_sums else None, eb.buf, eb.n)
    L.check(rc, eb)
    h = buf.cpu().numpy()
    for g in (0, 2, 4, 6):
        assert (h[offs[g]:offs[g + 1]] == -12345.0).all(), "write outside the outputs"
    sums = h[offs[5]:offs[6]]
    if not with_sums:
        assert (sums == -12345.0).all()
    return h[offs[1]:offs[2]], h[offs[3]:offs[4]], sums


CASES = [
    # the headline workload, then a smaller input on the same scratch
    (L.MODEL_HESTON, HESTON, 1.0, 252, 1_048_576, [90.0, 100.0, 110.0], 0.05, 42, 32, True),
    (L.MODEL_GBM, GBM, 1.0, 50, 10_000, [100.0], 0.05, 7, 64, True),
    (L.MODEL_MERTON, MERTON, 0.5, 30, 3_001, [80.0, 85.0, 90.0, 95.0, 100.0, 105.0, 110.0, 120.0],
     0.03, 9, 64, False),
]


def test_noalloc_equals_fused_two_calls_in_a_row(gpu):
    for model, params, t, steps, n, strikes, r, seed, prec, ws in CASES:
        ref_p, ref_se, ref_s = _reference(model, params, t, steps, n, strikes, r, seed, prec)
        p, se, s = _noalloc(model, params, t, steps, n, strikes, r, seed, prec, with_sums=ws)
        if ws:
            np.testing.assert_array_equal(s.view(np.uint64), ref_s.view(np.uint64))
        np.testing.assert_allclose(p, ref_p, rtol=1e-15, atol=0)
        np.testing.assert_allclose(se, ref_se, rtol=1e-15, atol=0)


def test_noalloc_repeated_call_is_identical(gpu):
    a = _noalloc(*CASES[1][:-1])
    b = _noalloc(*CASES[1][:-1])
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.view(np.uint64), y.view(np.uint64))


def test_noalloc_refuses_host_outputs(gpu):
    lib, eb = L.load(), L.ErrBuf()
    p = np.array(GBM)
    k = np.array([100.0])
    host = np.zeros(4)
    rc = lib.stocha_european_fused_noalloc(L.MODEL_GBM, p.ctypes.data_as(_pd), 3, 1.0, 10, 100, 1,
                                           64, 0, k.ctypes.data_as(_pd), 1, 0.05,
                                           host.ctypes.data, host.ctypes.data + 8, None, eb.buf,
                                           eb.n)
    assert rc == L.STOCHA_EINVAL and "price is not device memory" in eb.msg, eb.msg
    assert (host == 0.0).all()
