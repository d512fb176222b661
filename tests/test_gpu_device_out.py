"""Device-resident outputs through the C ABI (where = STOCHA_DEVICE).

Every entry point that fills a caller array can write straight into device
memory on the current HIP device instead of copying back to the host.  Each
case below calls the entry point twice with the same arguments, once into a
numpy array (STOCHA_HOST) and once into a torch device tensor, and requires
the two results to be the same bits, with guard words on both sides of the
device region left untouched (no write outside [out, out + n)).  The host
results themselves are checked against the oracle by the parity tests.
"""
import ctypes as C

import numpy as np
import pytest

from stocha_amd import _lib as L

pytestmark = pytest.mark.gpu

_pd = C.POINTER(C.c_double)
GUARD = 64
SENTINEL = np.uint64(0x7FF4DEADBEEF0001)  # a signalling-NaN bit pattern


def _host_and_device(call, n):
    """call(ptr, where, errbuf, errlen) -> rc; returns (host, device) arrays of n f64."""
    import torch
    eb = L.ErrBuf()
    host = np.empty(n, dtype=np.float64)
    L.check(call(host.ctypes.data if n else None, L.STOCHA_HOST, eb.buf, eb.n), eb)
    words = torch.full((n + 2 * GUARD,), int(SENTINEL), dtype=torch.int64,
                       device="cuda")
    torch.cuda.synchronize()
    base = words.data_ptr() + 8 * GUARD
    L.check(call(base, L.STOCHA_DEVICE, eb.buf, eb.n), eb)
    torch.cuda.synchronize()
    got = words.cpu().numpy().view(np.uint64)
    assert (got[:GUARD] == SENTINEL).all(), "write below the output region"
    assert (got[GUARD + n:] == SENTINEL).all(), "write past the output region"
    return host, got[GUARD:GUARD + n].view(np.float64)


def _same_bits(a, b):
    assert a.shape == b.shape
    np.testing.assert_array_equal(a.view(np.uint64), b.view(np.uint64))


CORR3 = np.array([[1.0, 0.5, 0.2], [0.5, 1.0, 0.3], [0.2, 0.3, 1.0]])


def _cases():
    lib = L.load()
    c3 = np.ascontiguousarray(CORR3)
    s0, mu, sg = (np.array(v, dtype=np.float64) for v in
                  ([100.0, 50.0, 75.0], [0.05, 0.03, 0.04], [0.2, 0.3, 0.25]))
    hw = (0.03, 0.1, 0.04, 0.01, 2.0)
    return {
        "gbm": (lambda p, w, e, n: lib.stocha_gbm(100.0, 0.05, 0.2, 1.0, 50, 1001, 7, 0, 0.0,
                                                  p, w, e, n), 1001 * 51),
        "gbm_antithetic": (lambda p, w, e, n: lib.stocha_gbm(100.0, 0.05, 0.2, 1.0, 37, 999, 8,
                                                             1, 0.01, p, w, e, n), 999 * 38),
        "heston_euler": (lambda p, w, e, n: lib.stocha_heston(
            100.0, 0.04, 0.05, 2.0, 0.04, 0.3, -0.7, 1.0, 40, 777, 9, b"euler", p, w, e, n),
            777 * 41),
        "heston_qe": (lambda p, w, e, n: lib.stocha_heston(
            100.0, 0.04, 0.05, 2.0, 0.04, 0.3, -0.7, 1.0, 40, 777, 9, b"qe", p, w, e, n),
            777 * 41),
        "merton": (lambda p, w, e, n: lib.stocha_merton_jump_diffusion(
            100.0, 0.05, 0.2, 1.0, -0.05, 0.1, 1.0, 30, 1500, 10, p, w, e, n), 1500 * 31),
        "multi_gbm": (lambda p, w, e, n: lib.stocha_multi_gbm(
            s0.ctypes.data_as(_pd), 3, mu.ctypes.data_as(_pd), 3, sg.ctypes.data_as(_pd), 3,
            c3.ctypes.data_as(_pd), 3, 3, 1.0, 20, 513, 11, 1, p, w, e, n), 513 * 21 * 3),
        "terminal_prices_gbm": (lambda p, w, e, n: lib.stocha_terminal_prices(
            L.MODEL_GBM, np.array([100.0, 0.05, 0.2, 1.0]).ctypes.data_as(_pd), 4, None, 25,
            3001, 12, p, w, e, n), 3001),
        "hull_white": (lambda p, w, e, n: lib.stocha_hull_white(*hw, 60, 640, 13, p, w, e, n),
                       640 * 61),
        "gaussian_copula": (lambda p, w, e, n: lib.stocha_gaussian_copula(
            c3.ctypes.data_as(_pd), 3, 3, 4097, 14, p, w, e, n), 4097 * 3),
        "student_t_copula": (lambda p, w, e, n: lib.stocha_student_t_copula(
            c3.ctypes.data_as(_pd), 3, 3, 5.0, 4097, 15, p, w, e, n), 4097 * 3),
    }


@pytest.mark.parametrize("name", ["gbm", "gbm_antithetic", "heston_euler", "heston_qe",
                                  "merton", "multi_gbm", "terminal_prices_gbm", "hull_white",
                                  "gaussian_copula", "student_t_copula"])
def test_device_output_equals_host_output(gpu, name):
    call, n = _cases()[name]
    host, dev = _host_and_device(call, n)
    assert np.isfinite(host).all()
    _same_bits(host, dev)


def _rng(seed):
    st, inc = (C.c_uint64 * 2)(), (C.c_uint64 * 2)()
    L.load().stocha_rng_new(seed, 0, st, inc)
    return st, inc


@pytest.mark.parametrize("n", [1, 1000, 65537])
def test_rng_device_output_and_state(gpu, n):
    # both the values and the advanced state agree between placements
    lib = L.load()
    st, inc = _rng(2024)
    states = {}

    def uni(p, w, e, m):
        new = (C.c_uint64 * 2)()
        rc = lib.stocha_rng_uniform(st, inc, n, p, w, new, e, m)
        states.setdefault(("u", w), tuple(new))
        return rc

    def nor(p, w, e, m):
        new, used = (C.c_uint64 * 2)(), (C.c_uint64 * 2)()
        rc = lib.stocha_rng_normal(st, inc, n, 1.5, 0.25, 1, p, w, new, used, e, m)
        states.setdefault(("n", w), tuple(new))
        return rc

    for f, tag in ((uni, "u"), (nor, "n")):
        host, dev = _host_and_device(f, n)
        _same_bits(host, dev)
        assert states[(tag, L.STOCHA_HOST)] == states[(tag, L.STOCHA_DEVICE)]


@pytest.mark.parametrize("name", ["gbm", "heston_euler", "multi_gbm", "terminal_prices_gbm",
                                  "gaussian_copula"])
def test_device_output_must_be_device_memory(gpu, name):
    # a host array, or pinned host memory, given as a STOCHA_DEVICE output is
    # refused before any kernel runs (a kernel writing it would fault)
    import torch
    call, n = _cases()[name]
    eb = L.ErrBuf()
    host = np.full(n, 7.0)
    rc = call(host.ctypes.data, L.STOCHA_DEVICE, eb.buf, eb.n)
    assert rc == L.STOCHA_EINVAL and "not device memory" in eb.msg, eb.msg
    assert (host == 7.0).all()
    pinned = torch.full((n,), 7.0, dtype=torch.float64).pin_memory()
    rc = call(pinned.data_ptr(), L.STOCHA_DEVICE, eb.buf, eb.n)
    assert rc == L.STOCHA_EINVAL and "not device memory" in eb.msg, eb.msg
    assert bool((pinned == 7.0).all())
    # the refusal leaves no error behind: the next call runs normally
    h2, d2 = _host_and_device(call, n)
    _same_bits(h2, d2)


def _segment_of(ptr):
    """(address, size) of the device allocation torch's caching allocator
    carved `ptr` from (requests between 1 and 10 MiB share 20 MiB
    segments, so a tensor's own size says nothing about it)."""
    import torch
    for seg in torch.cuda.memory_snapshot():
        if seg["address"] <= ptr < seg["address"] + seg["total_size"]:
            return seg["address"], seg["total_size"]
    return None


def test_device_output_too_small_is_refused(gpu):
    import torch
    call, n = _cases()["gbm"]
    buf = torch.full((16 << 17,), 7.0, dtype=torch.float64, device="cuda")  # 16 MiB
    torch.cuda.synchronize()
    seg = _segment_of(buf.data_ptr())
    if seg is None:
        pytest.skip("the caching allocator's segment of the buffer is not listed")
    addr, size = seg
    end = addr + size
    assert buf.data_ptr() + buf.numel() * 8 <= end
    eb = L.ErrBuf()
    # 80 bytes before the end of the allocation: far less than n doubles
    rc = call(end - 80, L.STOCHA_DEVICE, eb.buf, eb.n)
    assert rc == L.STOCHA_EINVAL and "the result needs" in eb.msg, eb.msg
    torch.cuda.synchronize()
    assert bool((buf == 7.0).all())
    # at the start of the buffer it fits and matches the host result
    assert 8 * n <= buf.numel() * 8
    eb2 = L.ErrBuf()
    L.check(call(buf.data_ptr(), L.STOCHA_DEVICE, eb2.buf, eb2.n), eb2)
    host = np.empty(n)
    L.check(call(host.ctypes.data, L.STOCHA_HOST, eb2.buf, eb2.n), eb2)
    _same_bits(host, buf[:n].cpu().numpy())
