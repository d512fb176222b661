"""Benchmark: Heston path-steps/s (1,048,576 paths x 252 steps), FP32 fused
form vs the plain FP64 form, error vs the FP64 oracle; LSMC ms.

A "step" is one pricing pass of the hot path: simulate 1,048,576 Heston
full-truncation Euler paths of 252 steps on their block-split PCG64DXSM
streams and price European calls at K = 90, 100, 110 in-register (no path
storage).  `value` is the FP32 form's path-steps/s over the timed steps.

Printed as one JSON line (rank 0, one GPU).  Extra keys carry the FP64 form,
the per-strike error against the FP64 CPU oracle, Merton, LSMC timings, the
roofline of the dominant kernel (timed with HIP events on the stream it runs
on) and the CPU baseline (the C oracle with OpenMP on the host cores, on the
same workload).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extras]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

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
# (FETCH_SIZE reads 131092.75 KiB for 262144 KiB moved -> x1.9997; WRITE_SIZE
# reads 262144 -> x1.0000): 219.89 KiB x 1.9997 + 384.09 KiB = 823.8 KiB.
# The kernel's only traffic is its block partials (and the last block's read
# of them), the tables and the arguments.
TRAFFIC_BYTES = int((219.8875 * 262144.0 / 131092.75 + 384.09375 * 1.0) * 1024)
OPS_PER_STEP_HESTON32 = DRAWS_PER_HESTON_STEP * (25 + 5) + 14
def _sync():
    import torch
    torch.cuda.synchronize()


def time_calls(fn, steps, warmup):
    for _ in range(warmup):
        fn()
    _sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    _sync()
    return (time.perf_counter() - t0) / steps


def kernel_ms(S, fn, reps=3):
    """Average duration of the dominant kernel (HIP events on its stream)."""
    S.enable_kernel_timing(True)
    tot = 0.0
    for _ in range(reps):
        fn()
        ms, _ = S.last_kernel_ms()
        tot += ms
    S.enable_kernel_timing(False)
    return tot / reps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="skip Merton and LSMC")
    ap.add_argument("--calib-copy", action="store_true",
                    help="end with one 256 MiB device-to-device copy (torch), the reference "
                         "kernel for scaling rocprof FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this build "
              "has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import stocha_amd as S

    def heston(prec):
        return S.european_price("heston", HESTON, T, STEPS, N_PATHS, STRIKES, r=R, seed=SEED,
                                precision=prec)

    # ---- timed: FP32 fused (the headline) ----
    t32 = time_calls(lambda: heston(32), args.steps, args.warmup)
    k32 = kernel_ms(S, lambda: heston(32))
    r32 = heston(32)
    # ---- FP64 plain form (reference arithmetic), same paths ----
    t64 = time_calls(lambda: heston(64), max(2, args.steps // 2), 1)
    k64 = kernel_ms(S, lambda: heston(64))
    r64 = heston(64)

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
                       "kernel_ms": k32}

    # ---- CPU baseline: the C oracle (OpenMP) on the same workload ----
    if not args.no_cpu:
        from oracle import oracle as O
        O.build()
        p = list(HESTON.values())
        t0 = time.perf_counter()
        term = O.terminals(1, p, T, STEPS, N_PATHS, SEED)
        s1, _ = O.call_sums(term, STRIKES)
        tc = time.perf_counter() - t0
        ref = math.exp(-R * T) * s1 / N_PATHS
        out["cpu_baseline"] = {"value": ps / tc, "unit": "path-steps/s",
                               "cores": O.num_threads(), "kind": "port",
                               "sample": f"full workload: {N_PATHS} Heston paths x {STEPS} "
                                         "steps, FP64, calls at 90/100/110 (one run)",
                               "seconds": tc}
        out["rel_err_vs_fp64_oracle"] = {
            "fp32": (np.abs(r32["price"] - ref) / ref).tolist(),
            "fp64": (np.abs(r64["price"] - ref) / ref).tolist(),
            "oracle_price": ref.tolist()}

    if not args.no_extras:
        def merton(prec):
            return S.european_price("merton", MERTON, T, STEPS, N_PATHS, STRIKES, r=R,
                                    seed=SEED, precision=prec)
        m32 = time_calls(lambda: merton(32), 8, 2)
        m64 = time_calls(lambda: merton(64), 6, 2)
        out["merton"] = {"fp32_path_steps_per_s": ps / m32, "fp64_path_steps_per_s": ps / m64,
                         "speedup": m64 / m32}
        lsmc = {}
        for n in (1_048_576, 4_194_304):
            kw = dict(s0=100.0, k=100.0, r=0.05, sigma=0.20, t=1.0, steps=50, n_paths=n,
                      is_put=True, poly_degree=3, seed=SEED)
            tl = time_calls(lambda: S.lsmc_american_option_ex(**kw, precision=32), 2, 1)
            price, se = S.lsmc_american_option_ex(**kw, precision=32)
            tl64 = time_calls(lambda: S.lsmc_american_option(**kw), 2, 1)
            p64, se64 = S.lsmc_american_option(**kw)
            lsmc[str(n)] = {"fp32_ms": tl * 1e3, "fp32_price": price, "fp32_std_err": se,
                            "fp64_ms": tl64 * 1e3, "fp64_price": p64, "fp64_std_err": se64}
        out["lsmc_ms"] = lsmc

        # GBM path matrix, 100,000 x 253 FP64 (gbm.rs); kernel time from the
        # HIP-event probe, wall time including the copy to host memory
        gkw = dict(s0=100.0, mu=0.05, sigma=0.2, t=1.0, steps=252, n_paths=100_000, seed=SEED)
        tg = time_calls(lambda: S.gbm(**gkw), 2, 1)
        kg = kernel_ms(S, lambda: S.gbm(**gkw))
        gsteps = gkw["n_paths"] * gkw["steps"]
        out["gbm_paths_100k"] = {"kernel_ms": kg, "wall_ms_incl_d2h": tg * 1e3,
                                 "path_steps_per_s_kernel": gsteps / (kg * 1e-3),
                                 "hbm_gbs_kernel": gkw["n_paths"] * 253 * 8 / (kg * 1e-3) / 1e9}

        # bump-and-revalue Greeks, GBM call 1,048,576 x 252, all five Greeks:
        # 8 bump scenarios in FP64 on shared draws (one launch), against the
        # same 8 scenarios priced one launch each (FP64 fused European kernel)
        fkw = dict(model="gbm", params={"s0": 100.0, "r": R, "sigma": 0.2, "t": T},
                   payoff="call", strike=100.0, n_paths=N_PATHS, n_steps=STEPS,
                   greeks=["delta", "gamma", "vega", "theta", "rho"], seed=SEED)
        tf = time_calls(lambda: S.greeks_fd(**fkw), 3, 1)
        kf = kernel_ms(S, lambda: S.greeks_fd(**fkw))

        def one_per_launch():
            for _ in range(8):
                S.european_price("gbm", {"s0": 100.0, "mu": R, "sigma": 0.2}, T, STEPS, N_PATHS,
                                 [100.0], r=R, seed=SEED, precision=64)
        t1 = time_calls(one_per_launch, 2, 1)
        out["greeks_fd_gbm_1Mx252"] = {
            "ms": tf * 1e3, "kernel_ms": kf, "scenarios": 8,
            "scenario_path_steps_per_s": 8 * N_PATHS * STEPS / tf,
            "ms_one_launch_per_scenario": t1 * 1e3, "speedup": t1 / tf,
            "values": S.greeks_fd(**fkw)}

        # Heston QE (Andersen), 1,048,576 x 252 terminal prices, FP64: the
        # variance recursion in the draw source, the log-price step per lane
        from stocha_amd import _stocha as SS
        qv = np.array([100.0, 0.04, 0.05, 2.0, 0.04, 0.3, -0.7, T])
        tq = time_calls(lambda: SS._terminal_prices(1, qv, "qe", STEPS, N_PATHS, SEED), 3, 1)
        kq = kernel_ms(S, lambda: SS._terminal_prices(1, qv, "qe", STEPS, N_PATHS, SEED))
        te = time_calls(lambda: SS._terminal_prices(1, qv, "euler", STEPS, N_PATHS, SEED), 3, 1)
        ke = kernel_ms(S, lambda: SS._terminal_prices(1, qv, "euler", STEPS, N_PATHS, SEED))
        out["heston_qe_terminals_1Mx252"] = {
            "ms": tq * 1e3, "kernel_ms": kq, "path_steps_per_s_kernel": ps / (kq * 1e-3),
            "euler_same_api_ms": te * 1e3, "euler_kernel_ms": ke}

        # path-dependent Monte Carlo on GBM, 1,048,576 x 252, FP64 (method="mc"):
        # up-and-out call with a rebate at the hit, floating lookback call,
        # arithmetic Asian call with the geometric control variate
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
            te = time_calls(fn, 3, 1)
            ke = kernel_ms(S, fn)
            ex[name] = {"ms": te * 1e3, "kernel_ms": ke, "price": fn(),
                        "path_steps_per_s": ps / te}
        out["exotics_mc_1Mx252"] = ex

        # RNG class bulk draws of one sequential stream, 100M values each
        # (kernel time: the GPU work of the call; wall time adds the copy of
        # 800 MB into pageable host memory)
        rng_ex = {}
        nbulk = 100_000_000
        for name, fn in (("uniform", lambda: S.RNG(SEED).uniform(nbulk)),
                         ("standard_normal", lambda: S.RNG(SEED).standard_normal(nbulk))):
            tw = time_calls(fn, 2, 1)
            kb = kernel_ms(S, fn, reps=2)
            rng_ex[name] = {"values": nbulk, "wall_ms": tw * 1e3, "kernel_ms": kb,
                            "values_per_s_kernel": nbulk / (kb * 1e-3),
                            "hbm_gbs_written": nbulk * 8 / (kb * 1e-3) / 1e9}
        out["rng_class_100M"] = rng_ex

        # single-stream copulas, 1M samples of dimension 4 (kernel time: the
        # GPU work; wall time adds the 32 MB copy to pageable host memory)
        cop = {}
        ccorr = np.array([[1.0, 0.5, 0.2, 0.1], [0.5, 1.0, 0.3, 0.2],
                          [0.2, 0.3, 1.0, 0.4], [0.1, 0.2, 0.4, 1.0]])
        ncop = 1_000_000
        for name, fn in (("gaussian", lambda: S.gaussian_copula(ccorr, n_samples=ncop, seed=SEED)),
                         ("student_t_nu5", lambda: S.student_t_copula(ccorr, nu=5.0,
                                                                      n_samples=ncop,
                                                                      seed=SEED))):
            tw = time_calls(fn, 2, 1)
            kb = kernel_ms(S, fn, reps=2)
            cop[name] = {"samples": ncop, "dim": 4, "wall_ms": tw * 1e3, "kernel_ms": kb,
                         "samples_per_s_kernel": ncop / (kb * 1e-3)}
        out["copula_1Mx4"] = cop

        # batched launch: 4096 uneven GBM / Heston / Merton instances (FP32)
        from stocha_amd import batch as B
        inst = B.config4_instances(4096, seed=SEED)
        plan = B.Plan(inst, seed=SEED)
        tb = time_calls(plan.run, 2, 1)
        out["batch_4096"] = {"ms": tb * 1e3, "path_steps": B.path_steps(inst),
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
        tk = time_calls(lambda: S.basket_price(**bkw), 2, 1)
        out["basket_d64"] = {"ms": tk * 1e3, "normals_per_s": 262_144 * 252 * 64 / tk,
                             "price": S.basket_price(**bkw)["price"]}

    if args.calib_copy:
        import torch
        x = torch.empty(1 << 26, dtype=torch.float32, device="cuda")  # 256 MiB
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        out["calib_copy_bytes"] = 1 << 28

    print(json.dumps(out))


if __name__ == "__main__":
    main()
