"""Batched pricing: thousands of uneven European-call instances in one
persistent, hipGraph-captured launch, with an autotuned launch shape.

    from stocha_amd import batch
    inst = batch.config4_instances(4096, seed=42)      # or build your own
    plan = batch.Plan(inst, seed=42)                    # autotuned shape
    price, std_err = plan.run()

An instance prices one European call on GBM, Heston (full-truncation Euler)
or Merton paths on the reference's block-split streams of the plan's seed,
with FP32 log-space stepping and FP64 sums.

The autotuner times candidate shapes (block size, paths per lane, resident
waves per SIMD) on a cost-preserving subsample of the instance table and
caches the winner on disk, keyed by device and workload signature
(``$STOCHA_AMD_TUNE_CACHE`` or ``~/.cache/stocha_amd/tune.json``).
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import time

import numpy as np

from . import _lib

MODEL_GBM, MODEL_HESTON, MODEL_MERTON = 0, 1, 2


class Instance(C.Structure):
    """Mirror of stocha_batch_instance (include/stocha_gpu.h)."""
    _fields_ = [("model", C.c_int), ("steps", C.c_uint32), ("n_paths", C.c_uint64),
                ("s0", C.c_double), ("strike", C.c_double), ("r", C.c_double),
                ("t", C.c_double), ("mu", C.c_double), ("sigma", C.c_double),
                ("v0", C.c_double), ("kappa", C.c_double), ("theta", C.c_double),
                ("xi", C.c_double), ("rho", C.c_double), ("lambda_", C.c_double),
                ("mu_j", C.c_double), ("sigma_j", C.c_double)]


def _declare(lib):
    if getattr(lib, "_batch_declared", False):
        return
    PI = C.POINTER(Instance)
    lib.stocha_batch_plan_create.argtypes = [PI, C.c_uint64, C.c_uint64, C.c_int, C.c_int,
                                             C.c_int, C.POINTER(C.c_void_p), C.c_char_p,
                                             C.c_size_t]
    lib.stocha_batch_plan_create.restype = C.c_int
    lib.stocha_batch_plan_run.argtypes = [C.c_void_p, C.POINTER(C.c_double),
                                          C.POINTER(C.c_double), C.c_char_p, C.c_size_t]
    lib.stocha_batch_plan_run.restype = C.c_int
    lib.stocha_batch_plan_info.argtypes = [C.c_void_p, C.POINTER(C.c_uint64),
                                           C.POINTER(C.c_uint64)]
    lib.stocha_batch_plan_info.restype = C.c_int
    lib.stocha_batch_plan_destroy.argtypes = [C.c_void_p]
    lib.stocha_batch_plan_destroy.restype = None
    lib._batch_declared = True


def config4_instances(n: int = 4096, seed: int = 42) -> list[Instance]:
    """The uneven workload of the batched benchmark, drawn from NumPy's
    PCG64DXSM generator with `seed`, per instance in this order:
    model ~ U{GBM, Heston, Merton}; n_paths = round(10**U(3, 6));
    steps ~ U{12, 52, 252}; strike ~ U(80, 120); sigma ~ U(0.10, 0.50);
    t ~ U(0.25, 2.0); kappa ~ U(0.5, 4); theta ~ U(0.01, 0.09);
    xi ~ U(0.1, 0.8); rho ~ U(-0.9, 0).
    Fixed: s0 = 100, r = mu = 0.05, Heston v0 = theta, Merton lambda = 1,
    mu_j = -0.05, sigma_j = 0.10 (the single-model benchmark's jump law)."""
    g = np.random.Generator(np.random.PCG64DXSM(seed))
    out = []
    for _ in range(n):
        model = int(g.integers(0, 3))
        n_paths = int(round(10 ** g.uniform(3.0, 6.0)))
        steps = int((12, 52, 252)[g.integers(0, 3)])
        strike = g.uniform(80, 120)
        sigma = g.uniform(0.10, 0.50)
        t = g.uniform(0.25, 2.0)
        kappa, theta = g.uniform(0.5, 4.0), g.uniform(0.01, 0.09)
        xi, rho = g.uniform(0.1, 0.8), g.uniform(-0.9, 0.0)
        out.append(Instance(model=model, steps=steps, n_paths=n_paths, s0=100.0, strike=strike,
                            r=0.05, t=t, mu=0.05, sigma=sigma, v0=theta, kappa=kappa,
                            theta=theta, xi=xi, rho=rho, lambda_=1.0, mu_j=-0.05, sigma_j=0.10))
    return out


def path_steps(instances) -> int:
    return int(sum(i.n_paths * i.steps for i in instances))


class Plan:
    """A captured batch launch.  shape = (block, paths_per_lane, waves_per_simd);
    None asks the autotuner."""

    def __init__(self, instances, seed: int = 42, shape=None):
        self.lib = _lib.load()
        _declare(self.lib)
        self.n = len(instances)
        self._arr = (Instance * self.n)(*instances)
        self.shape = tuple(shape) if shape is not None else tune(instances, seed)
        eb = _lib.ErrBuf()
        h = C.c_void_p()
        rc = self.lib.stocha_batch_plan_create(self._arr, self.n, seed, *self.shape, C.byref(h),
                                               eb.buf, eb.n)
        _lib.check(rc, eb)
        self._h = h

    def run(self):
        price = np.zeros(self.n)
        se = np.zeros(self.n)
        eb = _lib.ErrBuf()
        rc = self.lib.stocha_batch_plan_run(self._h, price.ctypes.data_as(C.POINTER(C.c_double)),
                                            se.ctypes.data_as(C.POINTER(C.c_double)), eb.buf,
                                            eb.n)
        _lib.check(rc, eb)
        return price, se

    def info(self):
        t, g = C.c_uint64(), C.c_uint64()
        self.lib.stocha_batch_plan_info(self._h, C.byref(t), C.byref(g))
        return {"tiles": t.value, "grid": g.value, "shape": self.shape}

    def close(self):
        if getattr(self, "_h", None):
            self.lib.stocha_batch_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------- autotuner --
CANDIDATES = [(b, p, w) for b in (64, 128, 256) for p in (1, 2) for w in (0, 2)]


def _cache_path():
    return os.environ.get("STOCHA_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "stocha_amd",
                                       "tune.json"))


def _signature(instances):
    h = hashlib.sha1()
    ms = np.array([(i.model, i.steps, i.n_paths) for i in instances], dtype=np.int64)
    # shape of the workload, not its exact parameters
    for m in range(3):
        sel = ms[ms[:, 0] == m]
        h.update(np.array([len(sel), sel[:, 1].sum() if len(sel) else 0,
                           int(np.log2(max(1, sel[:, 2].sum()))) if len(sel) else 0]).tobytes())
    return h.hexdigest()[:16]


def _device_key():
    if os.environ.get("STOCHA_AMD_TORCH", "1") == "0":
        return "hip-device"  # keep torch's HIP runtime out of the process
    try:
        import torch
        return torch.cuda.get_device_name(0)
    except Exception:
        return "hip-device"


def subsample(instances, k: int = 384, seed: int = 0):
    """Instances drawn with probability proportional to cost, n_paths scaled
    down so the sample runs in well under a second."""
    rng = np.random.default_rng(seed)
    cost = np.array([i.n_paths * i.steps * (2.0 if i.model == 1 else 1.0) for i in instances])
    idx = rng.choice(len(instances), size=min(k, len(instances)), replace=False,
                     p=cost / cost.sum())
    out = []
    for j in idx:
        src = instances[int(j)]
        c = Instance()
        C.memmove(C.byref(c), C.byref(src), C.sizeof(Instance))
        c.n_paths = max(256, src.n_paths // 16)
        out.append(c)
    return out


def tune(instances, seed: int = 42, candidates=None, reps: int = 2, force: bool = False):
    """Pick (block, paths_per_lane, waves_per_simd) by timing candidates on a
    subsample; cached on disk."""
    key = f"{_device_key()}|{_signature(instances)}"
    path = _cache_path()
    cache = {}
    if os.path.exists(path):
        try:
            cache = json.load(open(path))
        except Exception:
            cache = {}
    if not force and key in cache:
        return tuple(cache[key]["shape"])
    sample = subsample(instances)
    best, best_t = None, float("inf")
    timings = {}
    for shape in candidates or CANDIDATES:
        p = Plan(sample, seed=seed, shape=shape)
        p.run()  # warm
        t0 = time.perf_counter()
        for _ in range(reps):
            p.run()
        dt = (time.perf_counter() - t0) / reps
        p.close()
        timings[str(shape)] = dt
        if dt < best_t:
            best, best_t = shape, dt
    cache[key] = {"shape": list(best), "seconds": best_t, "timings": timings}
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as fh:
            json.dump(cache, fh, indent=1)
    except OSError:
        pass
    return tuple(best)
