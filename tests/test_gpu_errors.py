"""Failure handling on the GPU, forced from the host with ordinary inputs
(every kernel still runs to completion):

  - a singular I - Gamma in every draw (2-cycle with gamma_12 gamma_21 = 1
    and zero posterior variance): all weights zero -> SGDLM_ERR_NUMERIC, on
    the host-copy path and the device-resident path alike, as the oracle's
    sum of dets is 0 (the reference divides by it, SGDLM.cu:327-335);
  - NaN in y_t of a simultaneous parent: NaN posteriors where the oracle has
    them, then every det NaN -> SGDLM_ERR_NUMERIC;
  - NaN in an external regressor of a series outside Gamma: NaN VB moments
    of that series only, the other series equal to the oracle's;
  - n_t < 2 (gamma shape below 1, the reference's make_gamma2 skips such
    series, curand_manager.cuh:111-141): draws through the shape-boost
    sampler; the truncated psi-series Newton then fails to converge in 50
    iterations and n falls back to n_t (kernel_functions.cuh:1024-1068),
    as in the oracle;
  - a device allocation that cannot succeed (nsim = 2e9 at m = 400):
    SGDLM_ERR_ALLOC, and the context still works afterwards;
  - the Philox epoch counter at its 2^24 limit: SGDLM_ERR_STATE.
"""
import numpy as np
import pytest

from gpusgdlm_amd import SGDLM
from gpusgdlm_amd import _native as nat
from gpusgdlm_amd.sgdlm import SGDLMError
from gpusgdlm_amd.workloads import configure, filter_steps, random_model, readme_model
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _code(fn, *a):
    with pytest.raises(SGDLMError) as e:
        fn(*a)
    return e.value.code


def _two_cycle():
    """m = 2, each series the other's simultaneous parent, p = (2, 2)."""
    p = np.array([2, 2], dtype=np.uint32)
    sp = np.full((2, 2), None, dtype=object)
    sp[1, 0] = 0 + 1 * 2            # Gamma[0, 1]: series 0 regresses on series 1
    sp[1, 1] = 1 + 0 * 2            # Gamma[1, 0]
    m_t = np.array([[0.1, 1.0], [-0.2, 1.0]]).T.copy(order="F")   # (max_p, m)
    C_t = np.zeros((2, 2, 2), order="F")
    n_t, s_t = np.array([10.0, 12.0]), np.array([0.5, 0.7])
    return p, sp, m_t, C_t, n_t, s_t


def test_singular_two_cycle_all_weights_zero(hiplib):
    p, sp, m_t, C_t, n_t, s_t = _two_cycle()
    g = SGDLM(strict=True)
    g.setSimultaneousParents(p, sp)
    g.setDiscountFactors(np.ones(2), np.ones((2, 2, 2), order="F"))
    g.setPosteriorParameters(m_t, C_t, n_t, s_t)
    assert _code(g.computeVBPosterior, 500, 500) == nat.SGDLM_ERR_NUMERIC
    assert _code(g.vb_device, 500) == nat.SGDLM_ERR_NUMERIC
    lw = g.get_vb_result()["log_weights"]
    assert np.all(np.isneginf(lw))
    o = orc.OracleSGDLM(seed=1234)
    o.setSimultaneousParents(p, sp)
    o.setDiscountFactors(np.ones(2), np.ones((2, 2, 2), order="F"))
    o.setPosteriorParameters(m_t, C_t, n_t, s_t)
    vo = o.computeVBPosterior(500, 500, diag=True)
    assert vo["sum_det_weights"] == 0.0
    g.close()


def test_nan_observation_of_a_parent(hiplib):
    mod = readme_model(T=30)
    g, o = SGDLM(strict=True), orc.OracleSGDLM(seed=1234)
    g.set_rng_state(1234, 0)
    for api in (g, o):
        configure(api, mod)
        filter_steps(api, mod, 1, 10)
    y = mod.y[10].copy()
    y[5] = np.nan                   # series 5 is a simultaneous parent of series 0 and 1
    F = mod.F(10)
    F[2, 0] = np.nan                # the parent slot that holds y_t[5]
    for k in range(5):
        if mod.sp_u32[k, 1] == 1 + 5 * mod.m:
            F[k, 1] = np.nan
    for api in (g, o):
        api.computePosterior(y, F)
    pg, po = g.getParameters(), o.getParameters()
    for k in ("m", "C", "n", "s"):
        assert np.array_equal(np.isnan(pg[k]), np.isnan(po[k])), k
        fin = ~np.isnan(po[k])
        assert np.allclose(pg[k][fin], po[k][fin], rtol=1e-12, atol=0), k
    assert np.isnan(pg["m"][:, 0]).any() and np.isnan(pg["m"][:, 5]).any()
    assert _code(g.vb_device, 1000) == nat.SGDLM_ERR_NUMERIC
    assert _code(g.computeVBPosterior, 1000, 1000) == nat.SGDLM_ERR_NUMERIC
    vo = o.computeVBPosterior(1000, 1000, diag=True)
    assert not (vo["sum_det_weights"] > 0)
    g.close()


def test_nan_regressor_outside_gamma(hiplib):
    """Series 2 of the README model has no simultaneous parents and is no
    series' parent: a NaN intercept regressor makes its posterior and VB
    moments NaN, weights stay finite and every other series matches."""
    mod = readme_model(T=30)
    g, o = SGDLM(strict=True), orc.OracleSGDLM(seed=1234)
    g.set_rng_state(1234, 0)
    for api in (g, o):
        configure(api, mod)
        filter_steps(api, mod, 1, 10)
    F = mod.F(10)
    F[0, 2] = np.nan
    for api in (g, o):
        api.computePosterior(mod.y[10], F)
    vg = g.computeVBPosterior(2000, 2000)
    vo = o.computeVBPosterior(2000, 2000, diag=True)
    assert np.all(np.isfinite(vg["IS_weights"]))
    assert np.allclose(vg["IS_weights"], vo["IS_weights"], rtol=1e-10, atol=0)
    for k in ("m", "C", "n", "s"):
        a, b = vg[k], vo[k]
        assert np.array_equal(np.isnan(a), np.isnan(b)), k
        fin = ~np.isnan(b)
        assert np.allclose(a[fin], b[fin], rtol=1e-10, atol=1e-300), k
    assert np.isnan(vg["m"][0, 2]) and np.isfinite(vg["m"][:, [0, 1, 3, 4, 5]]).all()
    g.close()


def test_small_n_gamma_boost_and_newton_fallback(hiplib):
    """n_t = 1.5 and 0.5 for series 2 and 4 (gamma shapes 0.75 and 0.25;
    neither is in Gamma) and 1.9 for series 0 (in Gamma): the VB matches the
    oracle at 1e-10; for series 2 and 4 the Newton moment-match does not
    converge within 50 iterations and returns n_t."""
    mod = readme_model(T=30)
    g, o = SGDLM(strict=True), orc.OracleSGDLM(seed=1234)
    g.set_rng_state(1234, 0)
    for api in (g, o):
        configure(api, mod)
        filter_steps(api, mod, 1, 12)
        api.computePosterior(mod.y[12], mod.F(12))
    prm = o.getParameters()
    n = prm["n"].copy()
    n[0], n[2], n[4] = 1.9, 1.5, 0.5
    for api in (g, o):
        api.setPosteriorParameters(prm["m"], prm["C"], n, prm["s"])
    vg = g.computeVBPosterior(4000, 4000)
    vo = o.computeVBPosterior(4000, 4000, diag=True)
    print("newton iterations (oracle):", vo["newton_iter"].tolist(), "n:", vg["n"].tolist(), flush=True)
    assert vo["newton_iter"][2] >= 50 and vo["newton_iter"][4] >= 50
    assert vg["n"][2] == 1.5 and vg["n"][4] == 0.5
    for k in ("m", "C", "n", "s"):
        den = max(np.max(np.abs(vo[k])), 1e-300)
        assert np.max(np.abs(vg[k] - vo[k])) / den <= 1e-10, k
    g.close()


def test_impossible_allocation_reports_alloc(hiplib):
    mod = random_model(400, 20, T=3, seed=7)
    g = SGDLM(strict=True)
    configure(g, mod)
    g.computePosterior(mod.y[1], mod.F(1))
    assert _code(g.computeVBPosterior, 2_000_000_000, 2_000_000_000) == nat.SGDLM_ERR_ALLOC
    v = g.computeVBPosterior(200, 200)          # the context still works
    assert v is not None and abs(np.sum(v["IS_weights"]) - 1) < 1e-12
    g.close()


def test_epoch_counter_limit(hiplib):
    mod = readme_model(T=10)
    g = SGDLM(strict=True)
    configure(g, mod)
    g.computePosterior(mod.y[1], mod.F(1))
    assert _code(g.set_rng_state, 5, 1 << 24) == nat.SGDLM_ERR_ARG
    g.set_rng_state(5, (1 << 24) - 1)
    assert _code(g.vb_device, 100) == nat.SGDLM_ERR_STATE
    assert g.get_rng_state() == (5, (1 << 24) - 1)
    g.set_rng_state(5, (1 << 24) - 2)
    assert g.vb_device(100)
    assert g.get_rng_state() == (5, (1 << 24) - 1)
    g.close()
