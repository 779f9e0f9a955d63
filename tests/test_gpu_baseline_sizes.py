"""GPU parity against the FP64 oracle at the benchmark sizes themselves.

Same Philox seed and epochs on both sides, so both consume identical
lambda / theta draws; only the LU arithmetic differs.  Tolerances (max
error over the array divided by its largest magnitude, "rel" below):

  workload                                   FP64      FP32-MFMA   BF16x3
  m=400, k~U{0..20}, N=5e4, VB m/C/n/s       1e-10     1e-4        1e-3
    ESS (relative)                           1e-10     1%          1%
    log-weights (absolute, on |l| ~ 1e2)     1e-9      5e-3        5e-3
  m=400, 10-step forecast, first 2000 of     1e-9      1e-8 (FP32 + one FP64-residual refinement)
    N=1e5 draws; 5/50/95% quantiles
  m=1000, k~U{0..30}, N=1e4: as m=400        1e-10     1e-4        1e-3
  8 models, m log-uniform on [20, 400],      -         1e-4, ESS 1%
    N=1e4, one batched call
  README 6-series loop, T=500, VB + accept   1e-8 (after 500 coupled steps);
    if ESS > 5000 + forecast every step      forecast draws condition-aware

The quantile error is bounded by the per-draw error (a quantile is
1-Lipschitz in the sup norm), so the forecast quantiles over the 2000-draw
subset carry the per-draw bar.  Draws are keyed by draw index, so the first
2000 forecast draws of an N = 1e5 run are the oracle's N = 2000 draws.
Reference stage order: SGDLM.cu:181-503 (VB), SGDLM.cu:505-621 and
RcppWrapper.cpp:457-472 (forecast loop), README.md:103-154 (the loop).
"""
import numpy as np
import pytest

from gpusgdlm_amd import SGDLM, batch_vb
from gpusgdlm_amd.workloads import (configure, filter_steps, instance_models, random_model,
                                    readme_model)
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SEED = 1234
VB_TOL = {"fp64": 1e-10, "fp32": 1e-4, "bf16x3": 1e-3}
ESS_TOL = {"fp64": 1e-10, "fp32": 0.01, "bf16x3": 0.01}
LW_TOL = {"fp64": 1e-9, "fp32": 5e-3, "bf16x3": 5e-3}


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    den = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / den)


def _gpu_posterior(mod, prec, T):
    g = SGDLM(strict=True)
    g.set_rng_state(SEED, 0)
    g.set_precision(prec)
    configure(g, mod)
    filter_steps(g, mod, 1, T)
    g.computePosterior(mod.y[T], mod.F(T))
    return g


def _oracle_posterior(mod, T):
    o = orc.OracleSGDLM(seed=SEED, epoch=0)
    configure(o, mod)
    filter_steps(o, mod, 1, T)
    o.computePosterior(mod.y[T], mod.F(T))
    return o


def _vb_errors(g, vo, N):
    g.vb_device(N, N)
    r = g.get_vb_result()
    lw = r["log_weights"] - r["log_sum_det"]          # log of the normalised weight
    with np.errstate(divide="ignore"):
        lwo = np.log(vo["IS_weights"])
    fin = np.isfinite(lwo)
    err = {k: relerr(r[k], vo[k]) for k in ("m", "C", "n", "s")}
    err["ess"] = abs(r["ess"] / vo["ess"] - 1.0)
    err["log_w"] = float(np.max(np.abs(lw[fin] - lwo[fin])))
    err["w"] = relerr(r["IS_weights"], vo["IS_weights"])
    assert np.all(np.isfinite(lw[fin]))
    return err, r


def _check(err, prec, what):
    print(f"{what} {prec}: " + " ".join(f"{k}={v:.2e}" for k, v in err.items()), flush=True)
    for k in ("m", "C", "n", "s"):
        assert err[k] <= VB_TOL[prec], f"{what} {prec} {k}: {err[k]:.3e} > {VB_TOL[prec]}"
    assert err["ess"] <= ESS_TOL[prec], f"{what} {prec} ESS: {err['ess']:.3e}"
    assert err["log_w"] <= LW_TOL[prec], f"{what} {prec} log-weights: {err['log_w']:.3e}"


# ---------------------------------------------------------------- m = 400
@pytest.fixture(scope="module")
def c400():
    mod = random_model(400, 20, T=262, seed=20250101)
    T = 250
    print(f"m=400 oracle: {T} filter steps + VB N=50000", flush=True)
    o = _oracle_posterior(mod, T)
    vo = o.computeVBPosterior(50000, 50000, diag=True)
    return mod, T, vo


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_vb_m400_n50k_vs_oracle(hiplib, c400, prec):
    mod, T, vo = c400
    print(f"m=400 N=50000 VB {prec}", flush=True)
    g = _gpu_posterior(mod, prec, T)
    err, _ = _vb_errors(g, vo, 50000)
    assert g.lu_fallback_count() <= 50000 // 8
    g.close()
    _check(err, prec, "m=400 N=5e4")


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", 1e-8)])
def test_forecast_10step_m400_n100k_vs_oracle(hiplib, c400, prec, tol):
    mod, T, _ = c400
    K, N, n_or = 10, 100000, 2000
    Fk = np.stack([mod.F_forecast(T + 1 + k) for k in range(K)], axis=-1)
    print(f"m=400 10-step forecast N={N} {prec}", flush=True)
    g = _gpu_posterior(mod, prec, T)
    g.computePrior()
    yg = g.computeForecast(N, N, Fk)[:, :n_or, :]
    g.close()
    o = _oracle_posterior(mod, T)
    o.computePrior()
    yo = o.computeForecast(n_or, n_or, Fk)
    assert yg.shape == yo.shape == (mod.m, n_or, K)
    for k in range(K):
        e = relerr(yg[:, :, k], yo[:, :, k])
        qg = np.quantile(yg[:, :, k], [0.05, 0.5, 0.95], axis=1)
        qo = np.quantile(yo[:, :, k], [0.05, 0.5, 0.95], axis=1)
        eq = relerr(qg, qo)
        print(f"  step {k + 1}: y rel {e:.2e}, quantiles rel {eq:.2e}", flush=True)
        assert e <= tol and eq <= tol


# --------------------------------------------------------------- m = 1000
@pytest.fixture(scope="module")
def c1000():
    mod = random_model(1000, 30, T=252, seed=20250103)
    T = 250
    print(f"m=1000 oracle: {T} filter steps + VB N=10000", flush=True)
    o = _oracle_posterior(mod, T)
    vo = o.computeVBPosterior(10000, 10000, diag=True)
    return mod, T, vo


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_vb_m1000_n10k_vs_oracle(hiplib, c1000, prec):
    mod, T, vo = c1000
    print(f"m=1000 N=10000 VB {prec}", flush=True)
    g = _gpu_posterior(mod, prec, T)
    err, _ = _vb_errors(g, vo, 10000)
    g.close()
    _check(err, prec, "m=1000 N=1e4")


def test_forecast_m1000_quantiles_vs_oracle(hiplib, c1000):
    """1-step forecast at m = 1000 in FP32 with refinement: the first 1000 of
    N = 1e4 draws and their 5/50/95% quantiles against the oracle."""
    mod, T, _ = c1000
    N, n_or = 10000, 1000
    g = _gpu_posterior(mod, "fp32", T)
    g.computePrior()
    yg = g.computeForecast(N, N, mod.F_forecast(T + 1))[:, :n_or]
    g.close()
    o = _oracle_posterior(mod, T)
    o.computePrior()
    yo = o.computeForecast(n_or, n_or, mod.F_forecast(T + 1))
    e = relerr(yg, yo)
    eq = relerr(np.quantile(yg, [0.05, 0.5, 0.95], axis=1), np.quantile(yo, [0.05, 0.5, 0.95], axis=1))
    print(f"m=1000 forecast: y rel {e:.2e}, quantiles rel {eq:.2e}", flush=True)
    assert e <= 1e-8 and eq <= 1e-8


# ------------------------------------------------ many independent models
def test_batched_instances_vs_oracle(hiplib):
    """8 models with m log-uniform on [20, 400] through
    sgdlm_batch_compute_vb_posterior (FP32 LU, one launch per stage over all
    models), each against the oracle's VB of the same model."""
    T, N = 100, 10000
    mods = instance_models(8, T=T + 2, seed=77)
    print("instances m =", [md.m for md in mods], flush=True)
    gs = []
    for i, mod in enumerate(mods):
        g = SGDLM(strict=True)
        g.set_rng_state(SEED + i, 0)
        g.set_precision("fp32")
        configure(g, mod)
        filter_steps(g, mod, 1, T)
        g.computePosterior(mod.y[T], mod.F(T))
        gs.append(g)
    # the first batched call tunes each structure's LU (one stream per
    # model); the step compared runs as one launch per stage over all models
    assert batch_vb(gs, N) == [0] * len(gs)
    for i, g in enumerate(gs):
        g.set_rng_state(SEED + i, 0)
    assert batch_vb(gs, N) == [0] * len(gs)
    assert gs[0].batch_stats()["group"] == 1
    for i, (g, mod) in enumerate(zip(gs, mods)):
        o = orc.OracleSGDLM(seed=SEED + i, epoch=0)
        configure(o, mod)
        filter_steps(o, mod, 1, T)
        o.computePosterior(mod.y[T], mod.F(T))
        vo = o.computeVBPosterior(N, N, diag=True)
        r = g.get_vb_result()
        err = {k: relerr(r[k], vo[k]) for k in ("m", "C", "n", "s")}
        err["ess"] = abs(r["ess"] / vo["ess"] - 1)
        print(f"  model {i} m={mod.m}: " + " ".join(f"{k}={v:.1e}" for k, v in err.items()), flush=True)
        for k in ("m", "C", "n", "s"):
            assert err[k] <= 1e-4, (i, k)
        assert err["ess"] <= 0.01, i
        g.close()


# ------------------------------------------------ the README loop, T = 500
def _kappas(mod, th):
    """2-norm condition numbers of I - Gamma for draws th (max_p, m, n)."""
    n = th.shape[-1]
    A = np.broadcast_to(np.eye(mod.m), (n, mod.m, mod.m)).copy()
    spf = mod.sp_u32.ravel(order="F")
    for q, v in enumerate(spf):
        if v < mod.m * mod.m:
            k, j = q % mod.max_p, q // mod.max_p
            A[:, v % mod.m, v // mod.m] = -th[k, j, :]
    return np.linalg.cond(A)


def test_readme_loop_T500(hiplib):
    """README.md:103-154 end to end, FP64: naive posterior, VB with 1e4
    draws, accept the VB posterior when ESS > 5000, step-ahead prior and a
    1e4-draw forecast with its 5/95% quantiles, at every step, GPU and oracle
    each carrying their own state.  Bars: VB m/C/n/s 1e-8 relative after up
    to 499 coupled steps; the accept decisions identical; forecast quantiles
    1e-8; per draw, relative error <= 1e-10 max(1, kappa): the two states
    drift apart by up to ~1e-10 in single entries over the coupled steps
    (reported as the parameter drift) and the solve amplifies a perturbation
    of the draw by the condition number kappa of its I - Gamma; the README
    model has 2-cycles, so ill-conditioned draws occur."""
    mod = readme_model(T=500)
    T, N = 500, 10000
    g = SGDLM(strict=True)
    g.set_rng_state(SEED, 0)
    o = orc.OracleSGDLM(seed=SEED, epoch=0)
    for api in (g, o):
        configure(api, mod)
    accepted = 0
    worst = {"vb": 0.0, "q": 0.0, "y_easy": 0.0, "kappa_max": 0.0, "drift": 0.0}
    n_hard = 0
    for t in range(1, T):
        Ft = mod.F(t)
        for api in (g, o):
            api.computePosterior(mod.y[t], Ft)
        vg = g.computeVBPosterior(N, N)
        vo = o.computeVBPosterior(N, N, diag=True)
        eg = 1.0 / np.sum(vg["IS_weights"] ** 2)
        eo = vo["ess"]
        for k in ("m", "C", "n", "s"):
            worst["vb"] = max(worst["vb"], relerr(vg[k], vo[k]))
        assert (eg > 5000) == (eo > 5000), t
        if eg > 5000:
            accepted += 1
            g.setPosteriorParameters(vg["m"], vg["C"], vg["n"], vg["s"])
            o.setPosteriorParameters(vo["m"], vo["C"], vo["n"], vo["s"])
        for api in (g, o):
            api.computePrior()
        pg, po = g.getParameters(), o.getParameters()
        with np.errstate(divide="ignore", invalid="ignore"):
            d = max(float(np.nanmax(np.abs(pg[k] - po[k]) / np.abs(po[k]))) for k in ("m", "C", "n", "s"))
        worst["drift"] = max(worst["drift"], d)
        if t < T - 1:
            Fn = mod.F_forecast(t + 1)
            rg = g.computeForecastDebug(N, N, Fn)
            yg = rg["y"]
            yo = o.computeForecast(N, N, Fn)
            err = np.linalg.norm(yg - yo, axis=0) / np.linalg.norm(yo, axis=0)
            easy = err <= 1e-10
            worst["y_easy"] = max(worst["y_easy"], float(err[easy].max(initial=0.0)))
            if not easy.all():
                hard = np.nonzero(~easy)[0]
                n_hard += len(hard)
                kap = _kappas(mod, rg["thetas"][:, :, hard])
                worst["kappa_max"] = max(worst["kappa_max"], float(kap.max()))
                assert np.all(err[hard] <= 1e-10 * kap), (t, err[hard].max(), kap.min())
            qg = np.quantile(yg, [0.05, 0.95], axis=1)
            qo = np.quantile(yo, [0.05, 0.95], axis=1)
            worst["q"] = max(worst["q"], relerr(qg, qo))
        if t % 100 == 0:
            print(f"  t={t}: accepted {accepted}, worst {worst}, draws above 1e-10: {n_hard}", flush=True)
    g.close()
    print(f"README loop: {accepted} of {T - 1} VB posteriors accepted; worst {worst}; "
          f"{n_hard} draws above 1e-10", flush=True)
    assert accepted > 0
    assert worst["vb"] <= 1e-8 and worst["q"] <= 1e-8
    assert n_hard <= (T - 2) * N // 1000
