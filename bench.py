"""Benchmark: Heston path-steps/s (1,048,576 paths x 252 steps), FP32 fused
form vs the plain FP64 form, error vs the FP64 oracle; LSMC ms.

A "step" is one pricing pass of the hot path: simulate 1,048,576 Heston
full-truncation Euler paths of 252 steps on their block-split PCG64DXSM
streams and price European calls at K = 90, 100, 110 in-register (no path
storage).  `value` is the FP32 form's path-steps/s over the timed steps.

Printed as one JSON line (rank 0, one GPU).  Extra keys carry the FP64 form,
the per-strike error against the FP64 CPU oracle, Merton, LSMC timings, the
roofline of the dominant kernel (its HIP-event duration averaged over the
timed calls, on the stream it runs on) and the CPU baselines: the C oracle
with OpenMP on the host cores, timed like examples/benchmark.py:24-31 (a
warm-up, then min and median of 5), all cores on the full Heston workload
and one core on a 1/16 sample, and LSMC on a 262,144-path sample plus one
1,048,576-path run for the oracle's price.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extras]

Under rocprofv3, run with STOCHA_AMD_TORCH=0 so that the process maps one
HIP runtime (the library's, the profiler's version) instead of torch's too.
The default --warmup 30 covers the GPU clock's ramp (the headline kernel
runs 1.13 ms per call at first, 0.99 ms from about the 25th call on).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("Heston path-steps/s (1M×252) FP32 vs FP64 GPU at ≤1e-4 rel price err; "
          "LSMC ms")
N_PATHS = 1_048_576
STEPS = 252
STRIKES = [90.0, 100.0, 110.0]
HESTON = dict(s0=100.0, v0=0.04, mu=0.05, kappa=2.0, theta=0.04, xi=0.3, rho=-0.7)
MERTON = dict(s0=100.0, mu=0.05, sigma=0.20, lambda_=1.0, mu_j=-0.05, sigma_j=0.10)
R = 0.05
T = 1.0
SEED = 42

# MI355X peaks: HBM 8 TB/s; VALU issue 256 CU x 4 SIMD x 16 lanes x 2.4 GHz
# = 39.3 T lane-instructions/s (one wave64 VALU instruction per 4 clocks per
# SIMD).  Measured on this kernel: kernel time = SQ_INSTS_VALU x 4 clocks per
# SIMD, i.e. the 32-bit integer multiplies of PCG issue at the full rate too.
PEAK_VALU_TOPS = 256 * 64 * 2.4e9 / 1e12
PEAK_HBM_GBS = 8000.0

# Algorithmic VALU work per Heston FP32 path-step, in lane-instructions of the
# minimal gfx950 code for each primitive:
#   per u64 draw  PCG64DXSM DXSM output 11 (6 multiplies) + 128x64-bit LCG
#                 step 14 (7 multiply-adds, 7 add-with-carry)              = 25
#                 ziggurat fast path: index, 32-bit test, u32->f32, scale    =  5
#   draws/step    2 normals x 1.0296 draws per normal (the ziggurat's rate)
#   per step      full-truncation Euler in log space (6 FMA, max, sqrt, ...) = 14
DRAWS_PER_HESTON_STEP = 2 * 1.0296
# HBM bytes per launch of that kernel: rocprofv3 FETCH_SIZE and WRITE_SIZE
# (KiB per dispatch) scaled by the same counters on a 256 MiB device copy
# (FETCH_SIZE reads 131092.625 KiB for 262144 KiB moved -> x1.99970;
# WRITE_SIZE reads 262144 -> x1.0000): 251.19 KiB x 1.99970 + 4736.53 KiB =
# 5238.8 KiB.  No path data: the block partials (and the last block's read of
# them), the tables, the arguments, and the scratch stores of the two dwords
# per thread the 7-wave build keeps for its epilogue (~4.4 MB).
TRAFFIC_BYTES = int((251.1875 * 262144.0 / 131092.625 + 4736.53125 * 1.0) * 1024)
OPS_PER_STEP_HESTON32 = DRAWS_PER_HESTON_STEP * (25 + 5) + 14
def time_calls(fn, steps, warmup, S=None):
    """Wall time per call over `steps` calls after `warmup` untimed ones.

    Every entry point synchronises its stream before it returns, so the wall
    clock brackets finished device work without a separate sync.  With S
    (the package) the HIP-event probe is on for the timed calls and the
    second value is the dominant kernel's average duration over exactly
    those calls (ms), measured on the stream the kernel runs on."""
    for _ in range(warmup):
        fn()
    if S is not None:
        S.enable_kernel_timing(True)
    kern = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
        if S is not None:
            kern += S.last_kernel_ms()[0]
    dt = (time.perf_counter() - t0) / steps
    if S is not None:
        S.enable_kernel_timing(False)
    return dt, (kern / steps if S is not None else None)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def min_median(fn, n_repeat=5):
    """examples/benchmark.py:24-31: one warm-up, then n_repeat timed runs;
    min and median wall time (s)."""
    fn()
    ts = []
    for _ in range(n_repeat):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return ts[0], ts[len(ts) // 2]


def log(msg):
    """Progress on stderr (the JSON line alone goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baselines(S, out, r32, r64):
    """The C oracle (-O2, OpenMP over paths) on the host cores, timed like
    examples/benchmark.py:24-31 (one warm-up, min and median of 5)."""
    from oracle import oracle as O
    O.build()
    nproc = os.cpu_count() or 1
    threads = min(nproc, 16)
    p = list(HESTON.values())
    ps = N_PATHS * STEPS
    res = {}

    def heston_cpu(n):
        term = O.terminals(1, p, T, STEPS, n, SEED)
        res["s1"], _ = O.call_sums(term, STRIKES)

    O.set_threads(threads)
    t_min, t_med = min_median(lambda: heston_cpu(N_PATHS))
    ref = math.exp(-R * T) * res["s1"] / N_PATHS
    # single thread on a 1/16 sample of the paths (same per-path work)
    n1 = N_PATHS // 16
    O.set_threads(1)
    t1_min, t1_med = min_median(lambda: heston_cpu(n1))
    O.set_threads(threads)
    out["cpu_baseline"] = {
        "value": ps / t_min, "unit": "path-steps/s", "cores": threads, "kind": "port",
        "sample": f"full workload: {N_PATHS} Heston paths x {STEPS} steps, FP64 oracle, calls at "
                  f"90/100/110; min of 5 after a warm-up",
        "median_path_steps_per_s": ps / t_med, "seconds_min": t_min, "seconds_median": t_med,
        "single_thread": {"path_steps_per_s_min": n1 * STEPS / t1_min,
                          "path_steps_per_s_median": n1 * STEPS / t1_med,
                          "sample": f"{n1} paths x {STEPS} steps (1/16 of the workload)"},
        "nproc": nproc, "cpu_model": cpu_model()}
    out["rel_err_vs_fp64_oracle"] = {
        "fp32": (np.abs(r32["price"] - ref) / ref).tolist(),
        "fp64": (np.abs(r64["price"] - ref) / ref).tolist(),
        "oracle_price": ref.tolist()}
    log(f"cpu heston: {ps / t_min:.3e} path-steps/s on {threads} threads")

    # LSMC put (50 steps, degree 3): min/median of 5 on a 262,144-path
    # sample, and one run at 1,048,576 paths for the oracle's price / SE
    kw = dict(s0=100.0, k=100.0, r=0.05, sigma=0.20, t=1.0, steps=50, is_put=True,
              poly_degree=3, seed=SEED)
    ns = 262_144
    l_min, l_med = min_median(lambda: O.lsmc(**kw, n_paths=ns))
    t0 = time.perf_counter()
    p1, se1 = O.lsmc(**kw, n_paths=N_PATHS)
    l1 = time.perf_counter() - t0
    out["cpu_baseline_lsmc"] = {
        "ms_min": l_min * 1e3, "ms_median": l_med * 1e3, "cores": threads, "kind": "port",
        "sample": f"{ns} paths x 50 steps (min/median of 5 after a warm-up)",
        "ms_1048576_one_run": l1 * 1e3, "price_1048576": p1, "std_err_1048576": se1}
    log(f"cpu lsmc: {l_min * 1e3:.1f} ms at {ns} paths, {l1 * 1e3:.0f} ms at {N_PATHS}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # The GPU clock ramps up over the first ~25 calls of the headline kernel
    # (1.13 -> 0.99 ms per call); 30 warm-up calls time the steady state.
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-extras", action="store_true", help="only the headline workload")
    ap.add_argument("--calib-copy", action="store_true",
                    help="end with one 256 MiB device-to-device copy (torch), the reference "
                         "kernel for scaling rocprof FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this build "
              "has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import stocha_amd as S

    def heston(prec):
        return S.european_price("heston", HESTON, T, STEPS, N_PATHS, STRIKES, r=R, seed=SEED,
                                precision=prec)

    # ---- timed: FP32 fused (the headline); kernel time over the same calls ----
    t32, k32 = time_calls(lambda: heston(32), args.steps, args.warmup, S)
    r32 = heston(32)
    log(f"heston fp32: {t32 * 1e3:.4f} ms/call, kernel {k32:.4f} ms")
    # ---- FP64 plain form (reference arithmetic), same paths ----
    t64, k64 = time_calls(lambda: heston(64), max(2, args.steps // 2), 1, S)
    r64 = heston(64)
    log(f"heston fp64: {t64 * 1e3:.4f} ms/call, kernel {k64:.4f} ms")

    ps = N_PATHS * STEPS
    value = ps / t32
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "path-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded PCG64DXSM block-split streams, seed 42)",
        "config": {"workload": "heston_full_truncation_euler_fused_european_call",
                   "n_paths": N_PATHS, "path_steps": STEPS, "strikes": STRIKES,
                   "params": HESTON, "r": R, "t": T, "seed": SEED,
                   "payoff_sums": "fp64"},
        "fp64_plain": {"value": ps / t64, "ms_per_step": t64 * 1e3, "kernel_ms": k64,
                       "price": r64["price"].tolist()},
        "fp32_fused": {"kernel_ms": k32, "price": r32["price"].tolist(),
                       "std_err": r32["std_err"].tolist()},
        "speedup_fp32_vs_fp64": t64 / t32,
    }
    achieved = OPS_PER_STEP_HESTON32 * ps / (k32 * 1e-3) / 1e12
    out["roofline"] = {"bound": "valu", "achieved": achieved, "peak": PEAK_VALU_TOPS,
                       "unit": "T lane-instr/s", "frac": achieved / PEAK_VALU_TOPS,
                       "traffic": TRAFFIC_BYTES,
                       "kernel": "terminal_feed_kernel<HestonSrc<fp32>, Heston32, CallPayoff>",
                       "ops_per_path_step": OPS_PER_STEP_HESTON32,
                       "kernel_ms": k32, "kernel_ms_over": f"the {args.steps} timed calls"}

    if not args.no_cpu:
        cpu_baselines(S, out, r32, r64)

    if not args.no_extras:
        extras(S, out)

    if args.calib_copy:
        import torch
        x = torch.empty(1 << 26, dtype=torch.float32, device="cuda")  # 256 MiB
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        out["calib_copy_bytes"] = 1 << 28

    print(json.dumps(out))


def extras(S, out):
    """Beside the headline: Merton, LSMC, the GBM path matrix, greeks_fd,
    Heston QE, exotics, the RNG class, copulas, the batch and the basket."""
    ps = N_PATHS * STEPS

    def merton(prec):
        return S.european_price("merton", MERTON, T, STEPS, N_PATHS, STRIKES, r=R,
                                seed=SEED, precision=prec)
    m32, _ = time_calls(lambda: merton(32), 8, 2)
    m64, _ = time_calls(lambda: merton(64), 6, 2)
    out["merton"] = {"fp32_path_steps_per_s": ps / m32, "fp64_path_steps_per_s": ps / m64,
                     "speedup": m64 / m32}
    log("merton done")

    # LSMC put, 50 steps, degree 3: wall time per call and the two kernels'
    # HIP-event time (path generation + cooperative induction)
    lsmc = {}
    for n in (1_048_576, 4_194_304):
        kw = dict(s0=100.0, k=100.0, r=0.05, sigma=0.20, t=1.0, steps=50, n_paths=n,
                  is_put=True, poly_degree=3, seed=SEED)
        tl, kl = time_calls(lambda: S.lsmc_american_option_ex(**kw, precision=32), 5, 1, S)
        price, se = S.lsmc_american_option_ex(**kw, precision=32)
        tl64, kl64 = time_calls(lambda: S.lsmc_american_option(**kw), 5, 1, S)
        p64, se64 = S.lsmc_american_option(**kw)
        # algorithmic HBM bytes: the FP32 path store written once (n x 50 x 4 B)
        # and read once by the induction (n x 49 x 4 B)
        hbm = n * (50 + 49) * 4
        lsmc[str(n)] = {"fp32_ms": tl * 1e3, "fp32_kernels_ms": kl, "fp32_price": price,
                        "fp32_std_err": se, "fp64_ms": tl64 * 1e3, "fp64_kernels_ms": kl64,
                        "fp64_price": p64, "fp64_std_err": se64,
                        "roofline_fp32": {"bound": "hbm", "achieved": hbm / (kl * 1e-3) / 1e9,
                                          "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                          "frac": hbm / (kl * 1e-3) / 1e9 / PEAK_HBM_GBS,
                                          "bytes": hbm}}
    out["lsmc_ms"] = lsmc
    log("lsmc done")

    # GBM path matrix, 100,000 x 253 FP64 (gbm.rs): kernel time from the
    # HIP-event probe, wall time including the copy to host memory
    gkw = dict(s0=100.0, mu=0.05, sigma=0.2, t=1.0, steps=252, n_paths=100_000, seed=SEED)
    tg, kg = time_calls(lambda: S.gbm(**gkw), 5, 1, S)
    gsteps = gkw["n_paths"] * gkw["steps"]
    gbytes = gkw["n_paths"] * 253 * 8
    out["gbm_paths_100k"] = {"kernel_ms": kg, "wall_ms_incl_d2h": tg * 1e3,
                             "path_steps_per_s_kernel": gsteps / (kg * 1e-3),
                             "path_steps_per_s_wall": gsteps / tg,
                             "roofline": {"bound": "hbm",
                                          "achieved": gbytes / (kg * 1e-3) / 1e9,
                                          "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                          "frac": gbytes / (kg * 1e-3) / 1e9 / PEAK_HBM_GBS,
                                          "bytes": gbytes}}
    log("gbm paths done")

    # multi_gbm path tensor, d = 8 correlated assets, 100,000 x 252, FP64
    # (multi_gbm.rs): 100,000 x 253 x 8 values written
    mrng = np.random.default_rng(SEED)
    ma = mrng.standard_normal((8, 8))
    mcv = ma @ ma.T / 8 + 0.1 * np.eye(8)
    mdg = np.sqrt(np.diag(mcv))
    mkw = dict(s0=list(mrng.uniform(50, 150, 8)), mu=list(mrng.uniform(0.0, 0.08, 8)),
               sigma=list(mrng.uniform(0.1, 0.4, 8)), corr=mcv / np.outer(mdg, mdg), t=1.0,
               steps=252, n_paths=100_000, seed=SEED)
    tm, km = time_calls(lambda: S.multi_gbm(**mkw), 3, 1, S)
    mbytes = 100_000 * 253 * 8 * 8
    out["multi_gbm_d8_100k"] = {"kernel_ms": km, "wall_ms_incl_d2h": tm * 1e3,
                                "asset_steps_per_s_kernel": 100_000 * 252 * 8 / (km * 1e-3),
                                "roofline": {"bound": "hbm",
                                             "achieved": mbytes / (km * 1e-3) / 1e9,
                                             "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                             "frac": mbytes / (km * 1e-3) / 1e9 / PEAK_HBM_GBS,
                                             "bytes": mbytes}}
    log("multi_gbm done")

    # bump-and-revalue Greeks, GBM call 1,048,576 x 252, all five Greeks:
    # 8 bump scenarios in FP64 on shared draws (one launch), against the
    # same 8 scenarios priced one launch each (FP64 fused European kernel)
    fkw = dict(model="gbm", params={"s0": 100.0, "r": R, "sigma": 0.2, "t": T},
               payoff="call", strike=100.0, n_paths=N_PATHS, n_steps=STEPS,
               greeks=["delta", "gamma", "vega", "theta", "rho"], seed=SEED)
    tf, kf = time_calls(lambda: S.greeks_fd(**fkw), 3, 1, S)

    def one_per_launch():
        for _ in range(8):
            S.european_price("gbm", {"s0": 100.0, "mu": R, "sigma": 0.2}, T, STEPS, N_PATHS,
                             [100.0], r=R, seed=SEED, precision=64)
    t1, _ = time_calls(one_per_launch, 2, 1)
    out["greeks_fd_gbm_1Mx252"] = {
        "ms": tf * 1e3, "kernel_ms": kf, "scenarios": 8,
        "scenario_path_steps_per_s": 8 * N_PATHS * STEPS / tf,
        "ms_one_launch_per_scenario": t1 * 1e3, "speedup": t1 / tf,
        "values": S.greeks_fd(**fkw)}

    # Heston QE (Andersen), 1,048,576 x 252 terminal prices, FP64
    from stocha_amd import _stocha as SS
    qv = np.array([100.0, 0.04, 0.05, 2.0, 0.04, 0.3, -0.7, T])
    tq, kq = time_calls(lambda: SS._terminal_prices(1, qv, "qe", STEPS, N_PATHS, SEED), 3, 1, S)
    te, ke = time_calls(lambda: SS._terminal_prices(1, qv, "euler", STEPS, N_PATHS, SEED), 3, 1,
                        S)
    out["heston_qe_terminals_1Mx252"] = {
        "ms": tq * 1e3, "kernel_ms": kq, "path_steps_per_s_kernel": ps / (kq * 1e-3),
        "euler_same_api_ms": te * 1e3, "euler_kernel_ms": ke}
    log("greeks_fd / qe done")

    # path-dependent Monte Carlo on GBM, 1,048,576 x 252, FP64 (method="mc")
    ex = {}
    ekw = dict(s=100.0, r=R, sigma=0.2, t=T, n_paths=N_PATHS, seed=SEED, method="mc")
    cases = {
        "barrier_uo_call_rebate": lambda: S.barrier_price(k=100.0, barrier=120.0,
                                                          n_steps=STEPS, rebate=3.0,
                                                          rebate_at_hit=True, **ekw),
        "lookback_floating_call": lambda: S.lookback_price(n_steps=STEPS, **ekw),
        "asian_arithmetic_call_cv": lambda: S.asian_price(k=100.0, n_steps=STEPS, **ekw),
    }
    for name, fn in cases.items():
        te, ke = time_calls(fn, 3, 1, S)
        ex[name] = {"ms": te * 1e3, "kernel_ms": ke, "price": fn(),
                    "path_steps_per_s": ps / te}
    out["exotics_mc_1Mx252"] = ex

    # RNG class bulk draws of one sequential stream, 100M values each
    rng_ex = {}
    nbulk = 100_000_000
    for name, fn in (("uniform", lambda: S.RNG(SEED).uniform(nbulk)),
                     ("standard_normal", lambda: S.RNG(SEED).standard_normal(nbulk))):
        tw, kb = time_calls(fn, 2, 1, S)
        rng_ex[name] = {"values": nbulk, "wall_ms": tw * 1e3, "kernel_ms": kb,
                        "values_per_s_kernel": nbulk / (kb * 1e-3),
                        "hbm_gbs_written": nbulk * 8 / (kb * 1e-3) / 1e9}
    out["rng_class_100M"] = rng_ex

    # single-stream copulas, 1M samples of dimension 4
    cop = {}
    ccorr = np.array([[1.0, 0.5, 0.2, 0.1], [0.5, 1.0, 0.3, 0.2],
                      [0.2, 0.3, 1.0, 0.4], [0.1, 0.2, 0.4, 1.0]])
    ncop = 1_000_000
    for name, fn in (("gaussian", lambda: S.gaussian_copula(ccorr, n_samples=ncop, seed=SEED)),
                     ("student_t_nu5", lambda: S.student_t_copula(ccorr, nu=5.0,
                                                                  n_samples=ncop,
                                                                  seed=SEED))):
        tw, kb = time_calls(fn, 2, 1, S)
        cop[name] = {"samples": ncop, "dim": 4, "wall_ms": tw * 1e3, "kernel_ms": kb,
                     "samples_per_s_kernel": ncop / (kb * 1e-3)}
    out["copula_1Mx4"] = cop
    log("exotics / rng / copulas done")

    # batched launch: 4096 uneven GBM / Heston / Merton instances (FP32)
    from stocha_amd import batch as B
    inst = B.config4_instances(4096, seed=SEED)
    plan = B.Plan(inst, seed=SEED)
    tb, kb = time_calls(plan.run, 3, 1, S)
    out["batch_4096"] = {"ms": tb * 1e3, "kernel_ms": kb, "path_steps": B.path_steps(inst),
                         "path_steps_per_s": B.path_steps(inst) / tb, **plan.info()}
    plan.close()

    # basket call, d = 64 correlated GBM, 262,144 paths x 252 steps, FP32
    # (L * sum z on MFMA) with the pathwise adjoint for all 64 assets
    rng = np.random.default_rng(SEED)
    a = rng.standard_normal((64, 64))
    c = a @ a.T / 64 + 0.1 * np.eye(64)
    dg = np.sqrt(np.diag(c))
    corr = c / np.outer(dg, dg)
    bs0 = rng.uniform(50, 200, 64)
    bmu = rng.uniform(0.01, 0.08, 64)
    bsig = rng.uniform(0.1, 0.4, 64)
    w = np.full(64, 1.0 / 64)
    bkw = dict(s0=bs0, mu=bmu, sigma=bsig, corr=corr, weights=w, strike=float(w @ bs0),
               r=R, t=T, steps=252, n_paths=262_144, seed=SEED, precision=32, greeks=True)
    tk, kk = time_calls(lambda: S.basket_price(**bkw), 3, 1, S)
    out["basket_d64"] = {"ms": tk * 1e3, "kernels_ms": kk,
                         "normals_per_s": 262_144 * 252 * 64 / tk,
                         "price": S.basket_price(**bkw)["price"]}
    log("batch / basket done")


if __name__ == "__main__":
    main()
