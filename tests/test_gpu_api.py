"""API behaviour on the GPU: the call-order rules of the reference
(checkInitialized / checkSimInitialized / checkPrior /
checkUseStateEvolutionMatrix, rSGDLM/src/SGDLM/HostWrapperImpl.cu:1239-1286,
and the is_prior transitions of HostWrapperImpl.cu:803-806,954-956,
1027-1028,1065-1067), argument validation, and the evolution-matrix path
(computeEvoPrior / computeEvoForecast[Debug], RcppWrapper.cpp:377-409,
576-825) against the oracle."""
import numpy as np
import pytest

from gpusgdlm_amd import SGDLM
from gpusgdlm_amd import _native as nat
from gpusgdlm_amd.sgdlm import SGDLMError
from gpusgdlm_amd.workloads import configure, filter_steps, readme_model
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _code(fn, *a):
    with pytest.raises(SGDLMError) as e:
        fn(*a)
    return e.value.code


def test_call_order_rules(hiplib):
    mod = readme_model(T=20)
    g = SGDLM(strict=True)
    # nothing initialised yet
    assert _code(g.computePosterior, mod.y[1], mod.F(1)) in (nat.SGDLM_ERR_STATE, nat.SGDLM_ERR_ARG)
    configure(g, mod)                               # prior set
    assert _code(g.computeVBPosterior, 100, 100) == nat.SGDLM_ERR_STATE   # needs a posterior
    assert _code(g.computePrior) == nat.SGDLM_ERR_STATE                   # needs a posterior
    g.computePosterior(mod.y[1], mod.F(1))          # now a posterior
    assert _code(g.computePosterior, mod.y[2], mod.F(2)) == nat.SGDLM_ERR_STATE  # needs a prior
    assert _code(g.computeForecast, 100, 100, mod.F_forecast(2)) == nat.SGDLM_ERR_STATE
    g.computeVBPosterior(200, 200)                  # allowed on a posterior
    g.computePrior()
    g.computeForecast(100, 100, mod.F_forecast(2))  # allowed on a prior
    prm = g.getParameters()
    g.setPosteriorParameters(prm["m"], prm["C"], prm["n"], prm["s"])
    assert _code(g.computeForecast, 100, 100, mod.F_forecast(2)) == nat.SGDLM_ERR_STATE
    g.clearGPUMemory()
    assert _code(g.computeVBPosterior, 100, 100) in (nat.SGDLM_ERR_STATE, nat.SGDLM_ERR_ARG)
    g.close()


def test_parental_set_validation(hiplib):
    mod = readme_model(T=10)
    m = mod.m
    bad = []
    sp = mod.sp.copy()
    sp[2, 0] = 0 + 0 * m                            # diagonal position (self-parent)
    bad.append(sp)
    sp = mod.sp.copy()
    sp[4, 3] = sp[3, 3]                             # the same position twice
    bad.append(sp)
    for s in bad:
        g = SGDLM(strict=True)
        assert _code(g.setSimultaneousParents, mod.p, s) == nat.SGDLM_ERR_ARG
        g.close()
    # a simultaneous position beyond p_j
    g = SGDLM(strict=True)
    p = mod.p.copy()
    p[3] = 2                                        # series 3 has parents at k = 3, 4
    assert _code(g.setSimultaneousParents, p, mod.sp) == nat.SGDLM_ERR_ARG
    g.close()
    g = SGDLM(strict=True)
    assert _code(g.set_lu_tuning, 128, 32) == nat.SGDLM_ERR_ARG
    assert _code(g.set_lu_max_front, -1) == nat.SGDLM_ERR_ARG
    g.close()


def _evo_G(mod, seed, steps):
    rng = np.random.default_rng(seed)
    G = np.zeros((mod.max_p, mod.max_p, mod.m, steps), order="F")
    for k in range(steps):
        for j in range(mod.m):
            P = int(mod.p[j])
            G[:P, :P, j, k] = np.eye(P) + rng.uniform(-0.05, 0.05, size=(P, P))
    return G


def test_evolution_matrix_prior_and_forecast(hiplib):
    mod = readme_model(T=60)
    G = _evo_G(mod, 3, 3)
    g = SGDLM(strict=True)
    g.set_rng_state(1234, 0)
    o = orc.OracleSGDLM(seed=1234, epoch=0)
    configure(g, mod)
    configure(o, mod)
    for t in range(1, 25):
        for api in (g, o):
            api.computePosterior(mod.y[t], mod.F(t))
            api.computeEvoPrior(G[..., t % 3])
        pg, po = g.getParameters(), o.getParameters()
        for k in ("m", "C", "n", "s"):
            den = max(np.max(np.abs(po[k])), 1e-300)
            assert np.max(np.abs(pg[k] - po[k])) / den <= 1e-12, (t, k)
    Fk = np.stack([mod.F_forecast(25 + k) for k in range(3)], axis=-1)
    rg = g.computeEvoForecastDebug(1500, 1500, Fk, G)
    ro = o.computeEvoForecastDebug(1500, 1500, Fk, G)
    assert np.array_equal(rg["lambdas"], ro["lambdas"])
    assert np.array_equal(rg["thetas"], ro["thetas"])
    den = np.max(np.abs(ro["y"]))
    assert np.max(np.abs(rg["y"] - ro["y"])) / den <= 1e-9
    # a single G (3-d) is reused at every step
    rg1 = g.computeEvoForecast(500, 500, Fk, G[..., 0])
    ro1 = o.computeEvoForecast(500, 500, Fk, G[..., 0])
    assert np.max(np.abs(rg1 - ro1)) / np.max(np.abs(ro1)) <= 1e-9
    g.close()


def test_batched_vb_equals_sequential(hiplib):
    """sgdlm_batch_compute_vb_posterior over independent models gives, bit
    for bit, what one computeVBPosterior per model gives (same seeds and
    epochs; one stream per context)."""
    from gpusgdlm_amd import batch_vb
    from gpusgdlm_amd.workloads import random_model
    mods = [random_model(m, min(k, m - 1), T=30, seed=100 + i)
            for i, (m, k) in enumerate([(20, 4), (60, 8), (150, 12), (33, 5), (90, 20), (7, 3)])]
    N = 2000

    def build():
        gs = []
        for mod in mods:
            g = SGDLM(strict=True)
            g.set_rng_state(77, 0)
            g.set_precision("fp32")
            configure(g, mod)
            filter_steps(g, mod, 1, 20)
            g.computePosterior(mod.y[20], mod.F(20))
            gs.append(g)
        return gs

    ga, gb = build(), build()
    print(f"batched VB over {len(mods)} models, N={N}", flush=True)
    codes = batch_vb(ga, N)
    assert codes == [0] * len(mods)
    codes = batch_vb(ga, N)            # a second batched step (epochs advance alike)
    assert codes == [0] * len(mods)
    for g in gb:
        g.vb_device(N, N)
        g.vb_device(N, N)
    for x, y in zip(ga, gb):
        rx, ry = x.get_vb_result(), y.get_vb_result()
        for k in ("m", "C", "n", "s", "IS_weights"):
            assert np.array_equal(rx[k], ry[k]), k
        assert x.get_rng_state() == y.get_rng_state()
    for g in ga + gb:
        g.close()
