"""ctypes loader for libstocha_gpu.so (the C ABI in include/stocha_gpu.h).

The library is loaded after torch when torch is installed: torch ships its
own HIP runtime, and the soname both share (libamdhip64.so.7) then resolves
to that one copy, so device pointers from torch tensors and from this library
live in the same HIP context.  Without torch the system ROCm runtime is used.

STOCHA_AMD_TORCH=0 skips torch: the library then binds the ROCm runtime it
was linked against (its RUNPATH), the only HIP runtime in the process --
what a profiler of that ROCm version expects.

There is no fallback: if the library is missing the import fails with an
error that says how to build it.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

torch = None
if os.environ.get("STOCHA_AMD_TORCH", "1") != "0":
    try:  # share torch's HIP runtime when it is present (see module docstring)
        import torch  # noqa: F401
    except Exception:  # pragma: no cover - torch is optional for the C ABI
        torch = None

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("STOCHA_AMD_LIB") or os.path.join(_HERE, "libstocha_gpu.so")

STOCHA_OK, STOCHA_EINVAL, STOCHA_EDEVICE, STOCHA_ENOMEM, STOCHA_ENOTIMPL = 0, 1, 2, 3, 4
STOCHA_HOST, STOCHA_DEVICE = 0, 1
STOCHA_FP64, STOCHA_FP32 = 64, 32
STOCHA_SCHEDULE_FEED, STOCHA_SCHEDULE_LOOP = 0, 1
MODEL_GBM, MODEL_HESTON, MODEL_MERTON = 0, 1, 2

_d = C.c_double
_u = C.c_uint64
_i = C.c_int
_pd = C.POINTER(C.c_double)
_pu = C.POINTER(C.c_uint64)
_ps = C.c_char_p
_sz = C.c_size_t
_vp = C.c_void_p

# name -> argtypes (restype int unless noted)
_SIGS = {
    "stocha_version": ([], C.c_char_p),
    "stocha_device_count": ([], C.c_int),
    "stocha_set_stream": ([_vp], C.c_int),
    "stocha_timing_enable": ([_i], C.c_int),
    "stocha_last_kernel_ms": ([_pd, _pu], C.c_int),
    "stocha_gbm": ([_d, _d, _d, _d, _u, _u, _u, _i, _d, _vp, _i, _ps, _sz], C.c_int),
    "stocha_heston": ([_d] * 8 + [_u, _u, _u, _ps, _vp, _i, _ps, _sz], C.c_int),
    "stocha_merton_jump_diffusion": ([_d] * 7 + [_u, _u, _u, _vp, _i, _ps, _sz], C.c_int),
    "stocha_multi_gbm": ([_pd, _u, _pd, _u, _pd, _u, _pd, _u, _u, _d, _u, _u, _u, _i, _vp, _i,
                          _ps, _sz], C.c_int),
    "stocha_lsmc_american_option": ([_d] * 5 + [_u, _u, _i, _u, _u, _pd, _pd, _ps, _sz],
                                    C.c_int),
    "stocha_greeks_pathwise": ([_d] * 5 + [_i, _u, _u, C.POINTER(C.c_char_p), _u, _u, _pd, _ps,
                                           _sz], C.c_int),
    "stocha_greeks_fd": ([_i, _pd, _u, _ps, _ps, _d, _u, _u, C.POINTER(C.c_char_p), _u, _u, _d,
                          _pd, _ps, _sz], C.c_int),
    "stocha_terminal_prices": ([_i, _pd, _u, _ps, _u, _u, _u, _vp, _i, _ps, _sz], C.c_int),
    "stocha_barrier_price": ([_d] * 6 + [_ps, _ps, _d, _u, _u, _u, _ps, _i, C.c_uint32, _d, _i,
                                          _pd, _ps, _sz], C.c_int),
    "stocha_asian_price": ([_d] * 5 + [_u, _ps, _ps, _ps, _d, _u, _u, _ps, _i, _d, _i, _d, _pd,
                                        _ps, _sz], C.c_int),
    "stocha_lookback_price": ([_d] * 4 + [_u, _ps, _ps, _d, _d, _u, _u, _ps, _i, _d, _i, _d, _pd,
                                           _ps, _sz], C.c_int),
    "stocha_rng_new": ([_u, _u, _pu, _pu], C.c_int),
    "stocha_rng_uniform": ([_pu, _pu, _u, _vp, _i, _pu, _ps, _sz], C.c_int),
    "stocha_rng_normal": ([_pu, _pu, _u, _d, _d, _i, _vp, _i, _pu, _pu, _ps, _sz], C.c_int),
    "stocha_hull_white": ([_d] * 5 + [_u, _u, _u, _vp, _i, _ps, _sz], C.c_int),
    "stocha_gaussian_copula": ([_pd, _u, _u, _u, _u, _vp, _i, _ps, _sz], C.c_int),
    "stocha_student_t_copula": ([_pd, _u, _u, _d, _u, _u, _vp, _i, _ps, _sz], C.c_int),
    "stocha_european_fused": ([_i, _vp, _u, _d, _u, _u, _u, _i, _i, _vp, _u, _d, _vp, _vp, _vp,
                               _ps, _sz], C.c_int),
    "stocha_european_fused_noalloc": ([_i, _pd, _u, _d, _u, _u, _u, _i, _i, _pd, _u, _d, _vp,
                                       _vp, _vp, _ps, _sz], C.c_int),
    "stocha_lsmc_ex": ([_d] * 5 + [_u, _u, _i, _u, _u, _i, _pd, _pd, _ps, _sz], C.c_int),
    "stocha_basket": ([_pd, _pd, _pd, _pd, _pd, _u, _d, _d, _d, _u, _u, _u, _i, _i, _pd, _pd,
                       _pd, _pd, _pd, _ps, _sz], C.c_int),
    "stocha_batch_plan_create": ([_vp, _u, _u, _i, _i, _i, C.POINTER(_vp), _ps, _sz], C.c_int),
    "stocha_batch_plan_run": ([_vp, _pd, _pd, _ps, _sz], C.c_int),
    "stocha_batch_plan_info": ([_vp, _pu, _pu], C.c_int),
    "stocha_batch_plan_destroy": ([_vp], None),
    "stocha_stream_u64": ([_u, _u, _u, _u, _u, _pu, _ps, _sz], C.c_int),
    "stocha_stream_normals": ([_u, _u, _u, _u, _u, _i, _pd, _pu, _ps, _sz], C.c_int),
    "stocha_stream_normals_queue": ([_u, _u, _u, _u, _u, _i, _pd, _ps, _sz], C.c_int),
}

_lib = None


def load():
    """Load the library and declare every exported signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m stocha_amd.build` (hipcc, gfx950)")
    lib = C.CDLL(LIB_PATH)
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


# Results below this size are copied through a pageable buffer: cheaper than
# the pinned allocator's bookkeeping.
_PINNED_MIN = 1 << 20
_pinned = None


def _pinned_available() -> bool:
    global _pinned
    if _pinned is None:
        try:
            _pinned = torch is not None and bool(torch.cuda.is_available())
        except Exception:  # pragma: no cover
            _pinned = False
    return _pinned


def host_array(shape, dtype=np.float64):
    """A C-contiguous float64 host array to receive a result from the device.

    Large results go to page-locked memory from torch's caching host allocator
    when torch's HIP runtime is the one in use: the copy back then runs at
    PCIe DMA speed instead of through a pageable staging buffer, and the
    allocator reuses freed blocks instead of pinning pages again.  The array
    keeps its block alive.  Plain NumPy memory otherwise."""
    shape = tuple(int(v) for v in shape)
    n = 1
    for v in shape:
        n *= v
    if (np.dtype(dtype) == np.float64 and n * 8 >= _PINNED_MIN and _pinned_available()):
        return torch.empty(shape, dtype=torch.float64, pin_memory=True).numpy()
    return np.empty(shape, dtype=dtype)


def exported_names():
    return list(_SIGS)


_tls = threading.local()


def errbuf() -> "ErrBuf":
    """This thread's reusable error buffer (entry points write it only when
    they return an error)."""
    eb = getattr(_tls, "eb", None)
    if eb is None:
        eb = _tls.eb = ErrBuf()
    return eb


class ErrBuf:
    """Caller-side error buffer passed to every entry point."""

    def __init__(self, n: int = 512):
        self.buf = C.create_string_buffer(n)
        self.n = n

    @property
    def msg(self) -> str:
        return self.buf.value.decode("utf-8", "replace")


def check(rc: int, eb: ErrBuf):
    """Map a status code to the Python exception the reference raises."""
    if rc == STOCHA_OK:
        return
    msg = eb.msg
    if rc == STOCHA_EINVAL:
        raise ValueError(msg)
    if rc == STOCHA_ENOMEM:
        raise MemoryError(msg)
    if rc == STOCHA_ENOTIMPL:
        raise NotImplementedError(msg)
    raise RuntimeError(msg or f"stocha_gpu error {rc}")
