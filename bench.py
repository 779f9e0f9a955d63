"""Benchmark: VB steps/s + IS draws/s of the SGDLM VB step at m=400,
N=5e4 importance draws (k ~ U{0..20} simultaneous parents per series),
FP32-MFMA mixed precision vs plain FP64, on one MI355X.

A "step" is one computeVBPosterior on the GPU with the model state resident
in HBM: Philox lambda/theta draws, the batched LU log-determinants of
I - Gamma_i, log-sum-exp importance weights, weighted moments and the
digamma/trigamma moment-match.  Results stay on the device (no PCIe copy in
the timed region); the PCIe-inclusive rate through the full API is reported
separately as `api_steps_per_s`.  Also reported: 1- and 10-step forecasts
with N=1e5 draws (device-resident), and the FP64 oracle on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "VB steps/s + IS draws/s at m=400,N=5e4: FP32-MFMA mixed vs plain FP64"
PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0}
COUNTERS_FILE = os.path.join(ROOT, "profiles", "vb_m400_n50k_lu_counters.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def dense_lu_flops(n: int) -> float:
    """Flops of one dense LU of an n x n matrix."""
    return 2.0 / 3.0 * n ** 3


def run_gpu(mod, prec, N, steps, warmup, T_filter, fc_N, fc_steps, fc_check=0):
    from gpusgdlm_amd import SGDLM
    from gpusgdlm_amd.workloads import configure, filter_steps
    g = SGDLM(strict=True)
    g.set_rng_state(1234, 0)
    g.set_precision(prec)
    configure(g, mod)
    filter_steps(g, mod, 1, T_filter)
    g.computePosterior(mod.y[T_filter], mod.F(T_filter))
    t0 = time.perf_counter()
    for _ in range(max(warmup, 1)):
        g.vb_device(N, N)          # the first call tunes the LU configuration
    warm_s = time.perf_counter() - t0
    tune = g.lu_autotune()
    g.set_timing(True)
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        g.vb_device(N, N)          # synchronous: returns when the step is done
        times.append(time.perf_counter() - t0)
    stage = {s: g.get_timing(s) for s in ("draw", "lu", "weights", "moments")}
    g.set_timing(False)
    res = g.get_vb_result()
    lu_fb = g.lu_fallback_count()
    # one more VB step at epoch 1 (the oracle's first VB call) for the error
    # table, outside the timed region
    g.set_rng_state(1234, 0)
    g.vb_device(N, N)
    chk = g.get_vb_result()
    out = {"ms_per_step": 1e3 * float(np.mean(times)), "median_ms": 1e3 * float(np.median(times)),
           "p99_ms": 1e3 * float(np.percentile(times, 99)),
           "stage_ms": {k: v[0] / max(v[1], 1) for k, v in stage.items()},
           "ess": res["ess"], "lu_fallback_draws": lu_fb,
           "lu_config": g.lu_config(), "elimination": g.lu_elimination(), "_check": chk,
           "autotune": tune, "warmup_s": warm_s}
    # PCIe-inclusive figure through the full API (outputs copied to host)
    t1 = time.perf_counter()
    g.computeVBPosterior(N, N)
    out["api_ms"] = 1e3 * (time.perf_counter() - t1)
    if fc_N:
        g.computePrior()
        Ff = mod.F_forecast(T_filter + 1)
        g.forecast_device(fc_N, Ff, 1)          # warm-up (allocation, plan)
        g.set_timing(True)
        t0 = time.perf_counter()
        g.forecast_device(fc_N, Ff, 1)
        f1 = time.perf_counter() - t0
        fl = g.get_timing("forecast_lu")
        t0 = time.perf_counter()
        g.forecast_device(fc_N, Ff, fc_steps)
        fk = time.perf_counter() - t0
        g.set_timing(False)
        out["forecast"] = {"nsim": fc_N, "1step_ms": 1e3 * f1, f"{fc_steps}step_ms": 1e3 * fk,
                           "lu_solve_ms_per_step": fl[0] / max(fl[1], 1)}
        if fc_check:
            # the first fc_check draws of a 1-step forecast at fixed epochs
            # (outside the timed region), for the error table
            g.set_rng_state(1234, 100)
            out["_fc_check"] = g.computeForecast(fc_N, fc_N, Ff)[:, :fc_check].copy()
    g.close()
    return out


def run_instances(K, N, steps, T_filter=250, seed=20250102):
    """Many independent models (K instances, m log-uniform on [20, 400],
    k_j ~ U{0..min(20, m-1)}, T_filter naive filter steps each, N draws,
    FP32 LU): VB steps of every instance per second through one batched
    call per step (one launch per stage over all models, replayed from a
    captured hipGraph; also eagerly with stage times, and with each model
    on its own stream), vs one call per instance; batch step times and
    per-instance step times (one call each, synchronous) as p50 / p99."""
    from gpusgdlm_amd import SGDLM, batch_vb
    from gpusgdlm_amd.workloads import configure, filter_steps, instance_models
    mods = instance_models(K, T=T_filter + 2, seed=seed)
    gs = []
    for mod in mods:
        g = SGDLM(strict=True)
        g.set_precision("fp32")
        configure(g, mod)
        filter_steps(g, mod, 1, T_filter)
        g.computePosterior(mod.y[T_filter], mod.F(T_filter))
        gs.append(g)
    t0 = time.perf_counter()
    batch_vb(gs, N)                    # first call: each structure's LU tuning
    first_s = time.perf_counter() - t0
    tunes = [g.lu_autotune() for g in gs]
    kinds = {}
    for t in tunes:
        kinds[t["kernel"]] = kinds.get(t["kernel"], 0) + 1
    def timed_batches(n):
        out = []
        for _ in range(n):
            t1 = time.perf_counter()
            batch_vb(gs, N)
            out.append(time.perf_counter() - t1)
        return out
    # one launch per stage over all models (the default batch mode)
    batch_vb(gs, N)
    tb = timed_batches(steps)
    lead = gs[0]
    replays = lead.graph_stats()["replays"]
    # the same stages run eagerly (stage timing on), with HIP-event stage times
    lead.set_timing(True)
    t1 = time.perf_counter()
    batch_vb(gs, N)
    eager_ms = 1e3 * (time.perf_counter() - t1)
    stages = {k: round(lead.get_timing(k)[0], 3) for k in
              ("batch_draw", "batch_lu", "batch_weights", "batch_moments")}
    lead.set_timing(False)
    # each model's stages on its own stream, enqueued before any wait
    lead.set_batch_mode("streams")
    batch_vb(gs, N)
    tsm = timed_batches(steps)
    lead.set_batch_mode("group")
    bstats = lead.batch_stats()
    for g in gs:
        g.vb_device(N, N)
    per = []
    t0 = time.perf_counter()
    for _ in range(steps):
        for g in gs:
            t1 = time.perf_counter()
            g.vb_device(N, N)
            per.append(time.perf_counter() - t1)
    ts = time.perf_counter() - t0
    for g in gs:
        g.close()
    ms = [md.m for md in mods]
    return {"instances": K, "nsim": N, "m_distribution": "log-uniform on [20, 400]",
            "m_quantiles": [int(np.percentile(ms, q)) for q in (0, 50, 90, 100)], "steps": steps,
            "batched_instance_vb_steps_per_s": K * steps / sum(tb),
            "batched_ms_per_step": 1e3 * sum(tb) / steps,
            "batched_step_ms_p50": 1e3 * float(np.percentile(tb, 50)),
            "batched_step_ms_p99": 1e3 * float(np.percentile(tb, 99)),
            "batched_graph_replays": replays,
            "batched_eager_ms_per_step": eager_ms,
            "batched_eager_stage_ms": stages,
            "streams_instance_vb_steps_per_s": K * steps / sum(tsm),
            "sequential_instance_vb_steps_per_s": K * steps / ts,
            "batched_vs_sequential": ts / sum(tb),
            "batch_calls": bstats,
            "instance_step_ms_p50": 1e3 * float(np.percentile(per, 50)),
            "instance_step_ms_p99": 1e3 * float(np.percentile(per, 99)),
            "first_batched_call_s": first_s,
            "lu_tuning_s_total": 1e-3 * sum(t["tune_ms"] for t in tunes),
            "lu_tuned_kernels": kinds}


def host_cpu():
    """CPU model and the cores this process may use."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"model": name, "logical_cpus": os.cpu_count(), "available_to_process": avail}


def lu_ms(mod, prec, N, T_filter, steps=5, setup=None):
    """Mean LU stage time (HIP events) of `steps` VB steps after one
    warm-up, with the context configured by setup(g) first."""
    from gpusgdlm_amd import SGDLM
    from gpusgdlm_amd.workloads import configure, filter_steps
    g = SGDLM(strict=True)
    g.set_rng_state(1234, 0)
    g.set_precision(prec)
    configure(g, mod)
    if setup:
        setup(g)
    filter_steps(g, mod, 1, T_filter)
    g.computePosterior(mod.y[T_filter], mod.F(T_filter))
    g.vb_device(N, N)
    g.set_timing(True)
    for _ in range(steps):
        g.vb_device(N, N)
    t = g.get_timing("lu")
    out = {"lu_ms": t[0] / max(t[1], 1), "lu_fallback_draws": g.lu_fallback_count(),
           "lu": g.lu_config()}
    g.close()
    return out


def vb_graph_ms(mod, prec, N, T_filter, steps=10):
    """Wall time per device-resident VB step in graph mode (the step
    captured once, then replayed), after two warm-up calls."""
    from gpusgdlm_amd import SGDLM
    from gpusgdlm_amd.workloads import configure, filter_steps
    g = SGDLM(strict=True)
    g.set_rng_state(1234, 0)
    g.set_precision(prec)
    g.set_graph_mode(True)
    configure(g, mod)
    filter_steps(g, mod, 1, T_filter)
    g.computePosterior(mod.y[T_filter], mod.F(T_filter))
    for _ in range(2):
        g.vb_device(N, N)
    t = []
    for _ in range(steps):
        t0 = time.perf_counter()
        g.vb_device(N, N)
        t.append(time.perf_counter() - t0)
    st = g.graph_stats()
    g.close()
    return {"ms_per_step": 1e3 * float(np.mean(t)), "median_ms": 1e3 * float(np.median(t)), **st}


def run_early(m, kmax, N, T_early, T_late, precs):
    """Early-series LU cost: after T_early naive filter steps the posterior
    of Gamma is wide and most draws need row interchanges (pivoting kernel
    on the Schur complement); LU time against the same model after T_late
    steps."""
    from gpusgdlm_amd.workloads import random_model
    mod = random_model(m, kmax, T=T_late + 2, seed=20250101)
    out = {"m": m, "nsim": N, "T_early": T_early, "T_late": T_late}
    for prec in precs:
        e = lu_ms(mod, prec, N, T_early)
        l = lu_ms(mod, prec, N, T_late)
        out[prec] = {"early_lu_ms": e["lu_ms"], "late_lu_ms": l["lu_ms"],
                     "ratio": e["lu_ms"] / l["lu_ms"], "early_fallback_draws": e["lu_fallback_draws"],
                     "late_fallback_draws": l["lu_fallback_draws"], "pivot_dim": e["lu"]["pivot_dim"]}
        log(f"[early m={m} {prec}] {out[prec]}")
    return out


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def vb_errors(res, vo):
    """Max errors of one GPU VB result against the oracle's (same draws)."""
    with np.errstate(divide="ignore"):
        lwo = np.log(vo["IS_weights"])
    lw = res["log_weights"] - res["log_sum_det"]
    fin = np.isfinite(lwo)
    e = {k: relerr(res[k], vo[k]) for k in ("m", "C", "n", "s")}
    e["ess_rel"] = abs(res["ess"] / vo["ess"] - 1.0)
    e["log_weight_abs"] = float(np.max(np.abs(lw[fin] - lwo[fin])))
    return e


def oracle_vb(mod, T_filter, N):
    from oracle import oracle as orc
    from gpusgdlm_amd.workloads import configure, filter_steps
    o = orc.OracleSGDLM(seed=1234, epoch=0)
    configure(o, mod)
    filter_steps(o, mod, 1, T_filter)
    o.computePosterior(mod.y[T_filter], mod.F(T_filter))
    t0 = time.perf_counter()
    vo = o.computeVBPosterior(N, N, diag=True)
    return vo, time.perf_counter() - t0


def oracle_forecast(mod, T_filter, n, Ff):
    """The oracle's 1-step forecast of n draws at the epochs the GPU check
    uses (posterior -> prior, draws at epoch 101, noise at 102)."""
    from oracle import oracle as orc
    from gpusgdlm_amd.workloads import configure, filter_steps
    o = orc.OracleSGDLM(seed=1234, epoch=100)
    configure(o, mod)
    filter_steps(o, mod, 1, T_filter)
    o.computePosterior(mod.y[T_filter], mod.F(T_filter))
    o.computePrior()
    return o.computeForecast(n, n, Ff)


def forecast_errors(yg, yo):
    q = [0.05, 0.5, 0.95]
    return {"y_rel": relerr(yg, yo),
            "quantiles_5_50_95_rel": relerr(np.quantile(yg, q, axis=1), np.quantile(yo, q, axis=1))}


def run_stretch(m, kmax, N, T_filter, precs, fc_check):
    """Precision stretch: m series, k ~ U{0..kmax}, N draws; per LU variant
    the VB step time and the max errors of log-weights, ESS, VB m/C/n/s and
    1-step forecast 5/50/95% quantiles (first fc_check draws) against the
    FP64 oracle on the same draws."""
    from gpusgdlm_amd.workloads import random_model
    mod = random_model(m, kmax, T=T_filter + 2, seed=20250103)
    out = {"m": m, "max_p": mod.max_p, "nsim": N, "variants": {}}
    res = {}
    for prec in precs:
        log(f"[m={m} {prec}] VB steps, N={N}")
        r = run_gpu(mod, prec, N, 3, 1, T_filter, N if prec != "bf16x3" else 0, 1,
                    fc_check=fc_check if prec != "bf16x3" else 0)
        res[prec] = r
    vo, dt = oracle_vb(mod, T_filter, N)
    yo = oracle_forecast(mod, T_filter, fc_check, mod.F_forecast(T_filter + 1)) if fc_check else None
    for prec, r in res.items():
        e = vb_errors(r["_check"], vo)
        if yo is not None and "_fc_check" in r:
            e.update({"forecast_" + k: v for k, v in forecast_errors(r["_fc_check"], yo).items()})
        out["variants"][prec] = {"ms_per_step": r["ms_per_step"], "vb_steps_per_s": 1e3 / r["ms_per_step"],
                                 "stage_ms": r["stage_ms"], "lu": r["lu_config"],
                                 "lu_fallback_draws": r["lu_fallback_draws"],
                                 "errors_vs_oracle": e}
        log(f"[m={m} {prec}] {r['ms_per_step']:.2f} ms/step, errors {e}")
    out["oracle_vb_s"] = dt
    out["forecast_check_draws"] = fc_check
    return out


def cpu_baseline(mod, T_filter, N_cap, budget_s=12.0):
    """FP64 oracle (OpenMP over draws) on a bounded sample of the same VB step:
    the LU log-determinants dominate (>99% of flops), so the sample is the
    draws + LU of n_cpu importance draws, scaled to one VB step of N draws.
    The sample grows until it takes a quarter of the budget or covers the
    whole step.  Also a 1-thread figure on 256 draws."""
    from oracle import oracle as orc
    from gpusgdlm_amd.workloads import configure, filter_steps
    o = orc.OracleSGDLM(seed=1234)
    configure(o, mod)
    filter_steps(o, mod, 1, T_filter)
    o.computePosterior(mod.y[T_filter], mod.F(T_filter))
    prm = o.getParameters()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()

    def timed(n, nthreads):
        t0 = time.perf_counter()
        lam, th = orc.sample(mod.m, mod.max_p, mod.p, prm["m"], prm["C"], prm["n"], prm["s"], n, 1234, 1)
        orc.logdet(mod.m, mod.max_p, mod.p, mod.sp_u32, th, nthreads=nthreads)
        return time.perf_counter() - t0

    n = max(threads, 8)
    while True:
        dt = timed(n, threads)
        if dt > budget_s / 4 or n >= N_cap:
            break
        n = min(N_cap, int(n * min(8.0, max(2.0, budget_s / 4 / max(dt, 1e-3)))))
    n1 = 256
    dt1 = timed(n1, 1)
    return {"n": n, "dt": dt, "threads": threads, "n1": n1, "dt1": dt1}


def lu_counters(prec, m, N, cfg):
    """Counter figures of the LU kernel of this workload (L2<->fabric bytes,
    MFMA busy cycles and executed MFMA flops per launch) from the committed
    rocprofv3 profile of the same workload and kernel (None if it does not
    match)."""
    try:
        with open(COUNTERS_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    e = t.get("precisions", {}).get(prec)
    if not e or t.get("m") != m or t.get("nsim") != N:
        return None
    if e.get("kind") != cfg.get("kernel") or e.get("dense_dim") != cfg.get("dense_dim"):
        return None
    return e


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nsim", type=int, default=50000)
    ap.add_argument("--m", type=int, default=400)
    ap.add_argument("--kmax", type=int, default=20)
    ap.add_argument("--filter-steps", type=int, default=250)
    ap.add_argument("--precisions", default="fp32,fp64,bf16x3")
    ap.add_argument("--forecast-nsim", type=int, default=100000)
    ap.add_argument("--forecast-steps", type=int, default=10)
    ap.add_argument("--instances", type=int, default=256,
                    help="also time K independent models (m ~ U{20..400}, N=1e4) batched vs one call each "
                         "(0 = skip)")
    ap.add_argument("--instances-nsim", type=int, default=10000)
    ap.add_argument("--no-stretch", action="store_true",
                    help="skip the m=1000, k~U{0..30}, N=1e4 precision-stretch leg")
    ap.add_argument("--no-errors", action="store_true",
                    help="skip the FP64-oracle error table of the headline workload")
    ap.add_argument("--no-early", action="store_true",
                    help="skip the early-series leg (m=400 after 30 vs 250 filter steps, N=1e4)")
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
        fcN = a.forecast_nsim if prec in ("fp32", "fp64") else 0
        results[prec] = run_gpu(mod, prec, N, a.steps, a.warmup, a.filter_steps, fcN, a.forecast_steps,
                                fc_check=0 if a.no_errors else 2000)
        r = results[prec]
        log(f"[{prec}] {r['ms_per_step']:.2f} ms/step (median {r['median_ms']:.2f}, p99 {r['p99_ms']:.2f}) "
            f"stages={ {k: round(v, 2) for k, v in r['stage_ms'].items()} } fallback={r['lu_fallback_draws']} "
            f"forecast={r.get('forecast')}")

    head = "fp32" if "fp32" in results else next(iter(results))
    r = results[head]
    steps_s = 1e3 / r["ms_per_step"]
    lu_s = r["stage_ms"]["lu"] * 1e-3
    el = r["elimination"]
    dense_eq = dense_lu_flops(a.m) * N                      # per launch, as the reference computes it
    M2 = (el["n_dense"] + 15) // 16 * 16
    executed = (dense_lu_flops(M2) + 2.0 * el["prefix_fma"]) * N
    achieved = dense_eq / lu_s / 1e12
    cnt = lu_counters(head, a.m, N, r["lu_config"])
    line = {
        "metric": METRIC,
        "value": steps_s,
        "unit": "VB steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"],
        "median_ms": r["median_ms"],
        "p99_ms": r["p99_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic: random SGDLM (seed 20250101), 250 naive filter steps, then VB",
        "config": {"workload": f"VB step, m={a.m} series, k~U{{0..{a.kmax}}} simultaneous parents, "
                               f"max_p={mod.max_p}, N={N} IS draws, LU in FP32 on MFMA",
                   "m": a.m, "max_p": mod.max_p, "nsim": N, "lu": r["lu_config"], "elimination": el},
        "is_draws_per_s": steps_s * N,
        "variants": {p: {"vb_steps_per_s": 1e3 / v["ms_per_step"], "is_draws_per_s": 1e3 / v["ms_per_step"] * N,
                         "ms_per_step": v["ms_per_step"], "median_ms": v["median_ms"], "p99_ms": v["p99_ms"],
                         "stage_ms": v["stage_ms"], "ess": v["ess"],
                         "lu_fallback_draws": v["lu_fallback_draws"], "lu": v["lu_config"],
                         "lu_autotune": v["autotune"],
                         "api_steps_per_s": 1e3 / v["api_ms"], "forecast": v.get("forecast")}
                     for p, v in results.items()},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK[head], "unit": "TFLOP/s",
                     "frac": achieved / PEAK[head],
                     "traffic": cnt["fetch_bytes"] + cnt["write_bytes"] if cnt else None,
                     "traffic_kind": "L2<->fabric bytes per LU launch (FETCH_SIZE x2 + WRITE_SIZE, scaled "
                                     "by a 1 GiB copy; Infinity-Cache hits included, not HBM alone)",
                     "mfma_util": cnt["mfma_util"] if cnt else None,
                     "mfma_util_def": "SQ_VALU_MFMA_BUSY_CYCLES / (4 SIMD x 256 CU x GRBM_GUI_ACTIVE / 8 XCD), "
                                      "profiles/vb_m400_n50k_lu_counters.md",
                     "mfma_executed_tflops": cnt["mfma_flop"] / lu_s / 1e12 if cnt else None,
                     "kernel": f"{cnt['kernel'] if cnt else r['lu_config']['kernel']} ({head}); achieved = "
                               f"dense-equivalent (2/3) m^3 x N = {dense_eq / 1e12:.3f} TFLOP per launch / "
                               f"LU time (HIP events, bench stream)",
                     "executed_tflops": executed / lu_s / 1e12,
                     "executed_tflop_per_launch": executed / 1e12},
    }
    if "fp64" in results and head != "fp64":
        line["speedup_vs_fp64"] = results["fp64"]["ms_per_step"] / r["ms_per_step"]
    if not a.no_errors:
        log(f"[errors] FP64 oracle VB at m={a.m}, N={N}")
        vo, dt = oracle_vb(mod, a.filter_steps, N)
        yo = oracle_forecast(mod, a.filter_steps, 2000, mod.F_forecast(a.filter_steps + 1)) \
            if a.forecast_nsim else None
        for p, v in results.items():
            e = vb_errors(v["_check"], vo)
            if yo is not None and "_fc_check" in v:
                e.update({"forecast_" + k: x for k, x in forecast_errors(v["_fc_check"], yo).items()})
            line["variants"][p]["errors_vs_oracle"] = e
            log(f"[errors] {p}: {e}")
        line["errors_note"] = (f"max error vs the FP64 oracle on the same draws: VB m/C/n/s relative to "
                               f"the array's largest entry, ESS relative, normalised log-weights "
                               f"absolute (oracle VB {dt:.1f} s); forecasts: first 2000 of "
                               f"{a.forecast_nsim} 1-step draws")
    # the autotuner's pick against the untuned default plan on the headline
    # workload (LU stage only)
    if not a.no_errors:
        d = lu_ms(mod, head, N, a.filter_steps, setup=lambda g: g.set_lu_autotune(False))
        line["autotune"] = {"tuned": r["autotune"], "tuned_lu_ms": r["stage_ms"]["lu"],
                            "default_plan_lu_ms": d["lu_ms"], "default_plan": d["lu"],
                            "tuned_vs_default": r["stage_ms"]["lu"] / d["lu_ms"]}
        log(f"[autotune] {line['autotune']}")
    line["graph_mode"] = {p: vb_graph_ms(mod, p, N, a.filter_steps) for p in results}
    log(f"[graph] {line['graph_mode']}")
    if not a.no_early:
        line["early_series"] = run_early(a.m, a.kmax, 10000, 30, a.filter_steps,
                                         [p for p in ("fp32", "fp64") if p in results])
    if not a.no_stretch:
        line["precision_stretch"] = run_stretch(1000, 30, 10000, a.filter_steps, a.precisions.split(","),
                                                0 if a.no_errors else 1000)
    if a.instances:
        log(f"[instances] {a.instances} models, N={a.instances_nsim}")
        line["many_instances"] = run_instances(a.instances, a.instances_nsim, 5)
        log(f"[instances] {line['many_instances']}")
    if not a.no_cpu:
        cb = cpu_baseline(mod, a.filter_steps, N)
        cpu_steps = 1.0 / (cb["dt"] * N / cb["n"])
        line["cpu_baseline"] = {"value": cpu_steps, "unit": "VB steps/s", "cores": cb["threads"], "kind": "port",
                                "cpu": host_cpu(),
                                "sample": f"FP64 oracle (OpenMP) draws + LU log-dets of {cb['n']} of the {N} draws "
                                          f"({cb['dt']:.2f} s), scaled to one VB step",
                                "is_draws_per_s": cb["n"] / cb["dt"],
                                "one_core_vb_steps_per_s": 1.0 / (cb["dt1"] * N / cb["n1"]),
                                "one_core_sample": f"{cb['n1']} draws in {cb['dt1']:.2f} s on 1 thread"}
    print(json.dumps(line, default=float), flush=True)


if __name__ == "__main__":
    main()
