"""Benchmark: VB steps/s + IS draws/s of the SGDLM VB step at m=400,
N=5e4 importance draws (k ~ U{0..20} simultaneous parents per series),
FP32-MFMA mixed precision vs plain FP64, on one MI355X.

A "step" is one computeVBPosterior on the GPU with the model state resident
in HBM: Philox lambda/theta draws, the batched LU log-determinants of
I - Gamma_i, log-sum-exp importance weights, weighted moments and the
digamma/trigamma moment-match.  Results stay on the device (no PCIe copy in
the timed region); the PCIe-inclusive rate through the full API is reported
separately as `api_steps_per_s`.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "VB steps/s + IS draws/s at m=400,N=5e4: FP32-MFMA mixed vs plain FP64"
PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def lu_flops(m: int) -> float:
    """Algorithmic flops of one dense LU of I - Gamma (m x m)."""
    return 2.0 / 3.0 * m ** 3


def run_gpu(mod, prec, N, steps, warmup, T_filter):
    from gpusgdlm_amd import SGDLM
    from gpusgdlm_amd.workloads import configure, filter_steps
    g = SGDLM(strict=True)
    g.set_rng_state(1234, 0)
    g.set_precision(prec)
    configure(g, mod)
    filter_steps(g, mod, 1, T_filter)
    g.computePosterior(mod.y[T_filter], mod.F(T_filter))
    for _ in range(warmup):
        g.vb_device(N, N)
    g.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        g.vb_device(N, N)
    dt = time.perf_counter() - t0
    stage = {s: g.get_timing(s) for s in ("draw", "lu", "weights", "moments")}
    g.set_timing(False)
    res = g.get_vb_result()
    fallback = g.lu_fallback_count()
    # PCIe-inclusive figure through the full API (outputs copied to host)
    t1 = time.perf_counter()
    out = g.computeVBPosterior(N, N)
    api = time.perf_counter() - t1
    cfg = g.lu_config()
    g.close()
    return {"ms_per_step": 1e3 * dt / steps, "stage_ms": {k: v[0] / max(v[1], 1) for k, v in stage.items()},
            "lu_launches": stage["lu"][1], "ess": res["ess"], "lu_fallback_draws": fallback, "api_ms": 1e3 * api, "lu_config": cfg,
            "vb": out}


def cpu_baseline(mod, T_filter, budget_s=12.0):
    """FP64 oracle (OpenMP over draws) on a bounded sample of the same VB step:
    the LU log-determinants dominate (>99% of flops), so the sample is the
    draws + LU of n_cpu importance draws, scaled to one VB step of N draws."""
    from oracle import oracle as orc
    from gpusgdlm_amd.workloads import configure, filter_steps
    o = orc.OracleSGDLM(seed=1234)
    configure(o, mod)
    filter_steps(o, mod, 1, T_filter)
    o.computePosterior(mod.y[T_filter], mod.F(T_filter))
    prm = o.getParameters()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    n = max(threads, 8)
    while True:
        t0 = time.perf_counter()
        lam, th = orc.sample(mod.m, mod.max_p, mod.p, prm["m"], prm["C"], prm["n"], prm["s"], n, 1234, 1)
        orc.logdet(mod.m, mod.max_p, mod.p, mod.sp_u32, th, nthreads=threads)
        dt = time.perf_counter() - t0
        if dt > budget_s / 4 or n >= 4096:
            break
        n = int(n * min(8.0, max(2.0, budget_s / 4 / max(dt, 1e-3))))
    return n, dt, threads


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nsim", type=int, default=50000)
    ap.add_argument("--m", type=int, default=400)
    ap.add_argument("--kmax", type=int, default=20)
    ap.add_argument("--filter-steps", type=int, default=250)
    ap.add_argument("--precisions", default="fp32,fp64,bf16x3")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy kernel of MIB MiB (counter calibration)")
    a = ap.parse_args()
    if a.copy_calibration:
        import torch
        n = a.copy_calibration * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_mib": a.copy_calibration, "kernels": 3}))
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this engine has no multi-GPU path "
              "(draw sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    from gpusgdlm_amd.workloads import random_model
    log(f"generating workload m={a.m} kmax={a.kmax} T={a.filter_steps + 1}")
    mod = random_model(a.m, a.kmax, T=a.filter_steps + 2, seed=20250101)
    N = a.nsim
    results = {}
    for prec in a.precisions.split(","):
        log(f"[{prec}] {a.warmup} warm-up + {a.steps} timed VB steps, N={N}")
        results[prec] = run_gpu(mod, prec, N, a.steps, a.warmup, a.filter_steps)
        log(f"[{prec}] {results[prec]['ms_per_step']:.2f} ms/step stages={results[prec]['stage_ms']} "
            f"fallback={results[prec]['lu_fallback_draws']}")

    head = "fp32" if "fp32" in results else next(iter(results))
    r = results[head]
    steps_s = 1e3 / r["ms_per_step"]
    lu_ms = r["stage_ms"]["lu"]
    achieved = lu_flops(a.m) * N / (lu_ms * 1e-3) / 1e12
    line = {
        "metric": METRIC,
        "value": steps_s,
        "unit": "VB steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (LU on MFMA) + f64 (draws, log-det sums, weights, moments)",
        "data": "synthetic: random SGDLM (seed 20250101), 250 naive filter steps, then VB",
        "config": {"workload": f"VB step, m={a.m} series, k~U{{0..{a.kmax}}} simultaneous parents, "
                               f"max_p={mod.max_p}, N={N} IS draws",
                   "m": a.m, "max_p": mod.max_p, "nsim": N, "lu": r["lu_config"]},
        "is_draws_per_s": steps_s * N,
        "variants": {p: {"vb_steps_per_s": 1e3 / v["ms_per_step"], "is_draws_per_s": 1e3 / v["ms_per_step"] * N,
                         "ms_per_step": v["ms_per_step"], "stage_ms": v["stage_ms"], "ess": v["ess"],
                         "lu_fallback_draws": v["lu_fallback_draws"],
                         "api_steps_per_s": 1e3 / v["api_ms"]}
                     for p, v in results.items()},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK[head], "unit": "TFLOP/s",
                     "frac": achieved / PEAK[head], "traffic": None,
                     "kernel": f"k_lu ({head}), {lu_flops(a.m) * N / 1e12:.3f} TFLOP algorithmic per launch"},
    }
    if "fp64" in results and head != "fp64":
        line["speedup_vs_fp64"] = results["fp64"]["ms_per_step"] / r["ms_per_step"]
    if not a.no_cpu:
        n, dt, thr = cpu_baseline(mod, a.filter_steps)
        cpu_steps = 1.0 / (dt * N / n)
        line["cpu_baseline"] = {"value": cpu_steps, "unit": "VB steps/s", "cores": thr, "kind": "port",
                                "sample": f"FP64 oracle draws + LU log-dets of {n} of the {N} draws "
                                          f"({dt:.2f} s), scaled to one VB step",
                                "is_draws_per_s": n / dt}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
