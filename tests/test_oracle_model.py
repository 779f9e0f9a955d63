"""Pins of the oracle's model recurrences against closed forms, invariants
and library routines (numpy / LAPACK), chosen so that a dropped term, a
wrong sign, index or transposition fails."""
import ctypes as C
import math

import numpy as np
import pytest

from gpusgdlm_amd.workloads import configure, filter_steps, random_model, readme_model
from oracle import oracle as orc

NA = None


def one_series(a, R, r, c, p=None, max_p=None):
    a = np.atleast_1d(np.asarray(a, float))
    R = np.atleast_2d(np.asarray(R, float))
    max_p = max_p or a.shape[0]
    p = p or a.shape[0]
    mt = np.zeros((max_p, 1), order="F")
    mt[: a.shape[0], 0] = a
    Ct = np.zeros((max_p, max_p, 1), order="F")
    Ct[: R.shape[0], : R.shape[1], 0] = R
    return mt, Ct, np.array([r], float), np.array([c], float), np.array([p], np.uint32)


def posterior(mt, Ct, nt, st, p, y, F):
    max_p = mt.shape[0]
    Fa = np.zeros((max_p, 1), order="F")
    Fa[: len(F), 0] = F
    orc.lib().sgo_compute_posterior(1, max_p, orc.u(p), orc.d(mt), orc.d(Ct), orc.d(nt), orc.d(st),
                                    orc.d(np.array([y], float)), orc.d(Fa))


def test_posterior_closed_form_scalar(oracle_lib):
    a, R, r, c, y, F = 0.3, 0.5, 3.0, 0.2, 1.1, 2.0
    mt, Ct, nt, st, p = one_series(a, R, r, c)
    posterior(mt, Ct, nt, st, p, y, [F])
    Q = c + F * R * F
    A = R * F / Q
    e = y - F * a
    s = c * (r + e * e / Q) / (r + 1)
    assert nt[0] == r + 1
    assert math.isclose(st[0], s, rel_tol=1e-14)
    assert math.isclose(mt[0, 0], a + A * e, rel_tol=1e-14)
    assert math.isclose(Ct[0, 0, 0], (R - A * A * Q) * s / c, rel_tol=1e-13)


def test_posterior_closed_form_matrix(oracle_lib):
    rng = np.random.default_rng(3)
    p = 4
    a = rng.normal(size=p)
    X = rng.normal(size=(p, p))
    R = X @ X.T + np.eye(p)
    r, c, y = 5.0, 0.7, 0.4
    F = rng.normal(size=p)
    mt, Ct, nt, st, pp = one_series(a, R, r, c, max_p=6)
    posterior(mt, Ct, nt, st, pp, y, F)
    Q = c + F @ R @ F
    A = R @ F / Q
    e = y - F @ a
    s = c * (r + e * e / Q) / (r + 1)
    assert np.allclose(mt[:p, 0], a + A * e, rtol=1e-13)
    Cp = (R - np.outer(A, A) * Q) * s / c
    assert np.allclose(Ct[:p, :p, 0], Cp, rtol=1e-12)
    assert np.allclose(Ct[:p, :p, 0], Ct[:p, :p, 0].T, rtol=0, atol=1e-14)
    assert np.all(np.linalg.eigvalsh(Ct[:p, :p, 0]) > 0)
    assert np.all(mt[p:, 0] == 0) and np.all(Ct[p:, p:, 0] == 0)


def test_posterior_zero_regressors(oracle_lib):
    a, R, r, c, y = [0.5, -1.0], np.diag([0.3, 0.4]), 4.0, 0.5, 0.9
    mt, Ct, nt, st, p = one_series(a, R, r, c)
    posterior(mt, Ct, nt, st, p, y, [0.0, 0.0])
    assert np.array_equal(mt[:, 0], np.array(a))
    assert np.allclose(Ct[:, :, 0], R * (r + y * y / c) / (r + 1), rtol=1e-14)


def test_prior_discount_and_G(oracle_lib):
    rng = np.random.default_rng(4)
    m, max_p = 2, 4
    p = np.array([3, 4], np.uint32)
    X = rng.normal(size=(max_p, max_p))
    C0 = X @ X.T
    Ct = np.asfortranarray(np.stack([C0, C0 + np.eye(max_p)], axis=2))
    Ct[3, 3, 0] = 0  # beyond p_0 (row>=p and col>=p) is zero after filtering
    mt = np.asfortranarray(rng.normal(size=(max_p, m)))
    mt[3, 0] = 0
    nt = np.array([10.0, 20.0])
    W = np.zeros((max_p, max_p, m), order="F")
    beta = np.array([0.9, 0.8])
    delta = np.asfortranarray(rng.uniform(0.7, 0.99, size=(max_p, max_p, m)))
    C_before, m_before = Ct.copy(), mt.copy()
    orc.lib().sgo_compute_prior(m, max_p, orc.u(p), orc.d(mt), orc.d(Ct), orc.d(nt), orc.d(W),
                                orc.d(beta), orc.d(delta), None)
    assert np.allclose(nt, [9.0, 16.0])
    for j in range(m):
        P = p[j]
        R = C_before[:P, :P, j] / delta[:P, :P, j]
        assert np.allclose(Ct[:P, :P, j], R, rtol=1e-15)
        assert np.allclose(W[:P, :P, j], R - C_before[:P, :P, j], rtol=1e-12, atol=1e-14)
    assert np.array_equal(mt, m_before)
    # with G
    G = np.asfortranarray(rng.normal(size=(max_p, max_p, m)))
    Ct, mt = C_before.copy(order="F"), m_before.copy(order="F")
    orc.lib().sgo_compute_prior(m, max_p, orc.u(p), orc.d(mt), orc.d(Ct), orc.d(nt), orc.d(W),
                                orc.d(beta), orc.d(delta), orc.d(G))
    for j in range(m):
        P = p[j]
        R = C_before[:, :, j].copy()
        R[:P, :P] = C_before[:P, :P, j] / delta[:P, :P, j]
        Rg = G[:, :, j] @ R @ G[:, :, j].T
        Rg[P:, P:] = 0
        assert np.allclose(Ct[:, :, j], Rg, rtol=1e-12, atol=1e-12)
        a = G[:, :, j] @ m_before[:, j]
        a[P:] = 0
        assert np.allclose(mt[:, j], a, rtol=1e-13, atol=1e-15)


def test_lu_logdet_against_lapack(oracle_lib):
    rng = np.random.default_rng(6)
    for n in (2, 7, 33, 64, 130):
        A = np.eye(n) - rng.uniform(-0.3, 0.3, size=(n, n)) * (rng.uniform(size=(n, n)) < 0.1)
        A[0, 0] = 0.0  # forces a row interchange at the first column
        A[1, 0] = 0.5
        A[0, 1] = 0.3
        ref = np.linalg.slogdet(A)[1]
        assert abs(orc.lu_logabsdet(A) - ref) <= 1e-12 * max(1, abs(ref))


def test_logdet_closed_forms(oracle_lib):
    assert orc.lu_logabsdet(np.array([[1.0, -0.7], [-0.4, 1.0]])) == pytest.approx(math.log(1 - 0.28), abs=1e-15)
    assert orc.lu_logabsdet(np.array([[1.0, -1.0], [-1.0, 1.0]])) == -math.inf
    rng = np.random.default_rng(1)
    L = np.tril(rng.normal(size=(20, 20)), -1) + np.eye(20)   # DAG: det(I - Gamma) = 1
    P = np.eye(20)[rng.permutation(20)]
    assert orc.lu_logabsdet(P @ L @ P.T) == pytest.approx(0.0, abs=1e-12)


def _posterior_state(mod, steps):
    o = orc.OracleSGDLM(seed=1234)
    configure(o, mod)
    filter_steps(o, mod, 1, steps)
    o.computePosterior(mod.y[steps], mod.F(steps))
    return o


def _dense_I_minus_Gamma(mod, th_i):
    A = np.eye(mod.m)
    spf = mod.sp_u32.ravel(order="F")
    thf = th_i.ravel(order="F")
    for q, v in enumerate(spf):
        if v < mod.m * mod.m:
            A[v % mod.m, v // mod.m] = -thf[q]
    return A


def test_vb_against_bruteforce_numpy(oracle_lib):
    """sgo_vb == an independent numpy evaluation of the VB formulas
    (SGDLM.cu:325-503) on the same draws."""
    mod = readme_model(T=40)
    o = _posterior_state(mod, 25)
    prm = o.getParameters()
    N = 3000
    v = o.computeVBPosterior(N, diag=True)
    lam, th = orc.sample(mod.m, mod.max_p, mod.p, prm["m"], prm["C"], prm["n"], prm["s"], N, 1234, 1)
    ld = np.array([np.linalg.slogdet(_dense_I_minus_Gamma(mod, th[:, :, i]))[1] for i in range(N)])
    w = np.exp(ld - ld.max())
    w /= w.sum()
    assert np.allclose(v["IS_weights"], w, rtol=1e-11, atol=0)
    assert v["ess"] == pytest.approx(1 / np.sum(w ** 2), rel=1e-11)
    for j in range(mod.m):
        P = int(mod.p[j])
        L = lam[j]
        T = th[:P, j, :]
        lb = np.sum(w * L)
        llb = np.sum(w * np.log(L))
        mb = (T * (w * L)).sum(axis=1) / lb
        D = T - mb[:, None]
        V = (D * (w * L)) @ D.T
        assert v["mean_lambda"][j] == pytest.approx(lb, rel=1e-12)
        assert v["mean_log_lambda"][j] == pytest.approx(llb, rel=1e-12)
        assert np.allclose(v["m"][:P, j], mb, rtol=1e-11, atol=1e-14)
        assert np.all(v["m"][P:, j] == 0)
        Q = np.sum(w * L * np.einsum("ki,kl,li->i", D, np.linalg.inv(V), D))
        assert v["mean_Q"][j] == pytest.approx(Q, rel=1e-10)
        assert v["mean_Q"][j] == pytest.approx(P, rel=1e-10)   # trace identity
        n = v["n"][j]
        # Newton root: |f(n)| <= 1e-6 with the restated series (or fallback n_t)
        if v["newton_iter"][j] < 50:
            Lb = orc.lib()
            f = (math.log(n + P - Q) - Lb.sgo_psi(n / 2) - (P - Q) / n - math.log(2 * lb) + llb)
            assert abs(f) <= 1e-6
        s = (n + P - v["mean_Q"][j]) / (n * lb)
        assert v["s"][j] == pytest.approx(s if s > 0 else prm["s"][j], rel=1e-14)
        assert np.allclose(v["C"][:P, :P, j], v["s"][j] * V, rtol=1e-10, atol=1e-16)


def test_equal_logdets_give_full_ess(oracle_lib):
    """No simultaneous parents: every det(I - Gamma) = 1, ESS = N."""
    mod = readme_model(T=30)
    mod.sp = np.full(mod.sp.shape, None, dtype=object)
    mod.sp_u32[:] = 2 ** 31
    o = _posterior_state(mod, 12)
    v = o.computeVBPosterior(500, diag=True)
    assert np.allclose(v["IS_weights"], 1 / 500, rtol=1e-14)
    assert v["ess"] == pytest.approx(500, rel=1e-12)
    assert v["sum_det_weights"] == pytest.approx(500, rel=1e-14)


def test_forecast_solves_the_system(oracle_lib):
    mod = random_model(20, 4, T=30, seed=5)
    o = orc.OracleSGDLM(seed=99)
    configure(o, mod)
    filter_steps(o, mod, 1, 20)
    x = mod.F_forecast(20)
    r = o.computeForecastDebug(200, 200, x)
    L = orc.lib()
    e_nu = o.epoch
    for i in range(0, 200, 37):
        th = r["thetas"][:, :, i]
        A = _dense_I_minus_Gamma(mod, th)
        rhs = np.zeros(mod.m)
        for j in range(mod.m):
            nu = L.sgo_normal_k(99, e_nu, 3, i, j, 0) / math.sqrt(r["lambdas"][j, i])
            ext = mod.sp_u32[:, j] >= mod.m * mod.m
            rhs[j] = nu + np.sum(x[ext, j] * th[ext, j])
        assert np.allclose(A @ r["y"][:, i], rhs, rtol=1e-12, atol=1e-12)
        assert np.allclose(r["y"][:, i], np.linalg.solve(A, rhs), rtol=1e-10, atol=1e-12)


def test_multistep_evolution_identities(oracle_lib):
    """beta = 1 keeps lambda (eta = 1); W = 0 (Delta = 1) keeps theta (the
    semi-definite Cholesky gives a zero factor instead of the reference's
    0 * (1/0) = NaN)."""
    mod = readme_model(T=40)
    mod.beta[:] = 1.0
    mod.delta[:] = 1.0
    o = orc.OracleSGDLM(seed=3)
    configure(o, mod)
    filter_steps(o, mod, 1, 20)
    Fk = np.stack([mod.F_forecast(20), mod.F_forecast(21)], axis=-1)
    r = o.computeForecastDebug(300, 300, Fk)
    assert np.array_equal(r["lambdas"][:, :, 0], r["lambdas"][:, :, 1])
    assert np.allclose(r["thetas"][:, :, :, 0], r["thetas"][:, :, :, 1], rtol=0, atol=0)


def test_theta_transform_moments(oracle_lib):
    """sgo_sample: lambda ~ Gamma(n/2, rate n s/2) gives E[lambda] = 1/s,
    and sqrt(s lambda) (theta - m) ~ N(0, C) -- a transposed Cholesky factor
    (U U' instead of U'U), a missing sqrt or a dropped s fails this."""
    rng = np.random.default_rng(4)
    P, mp = 3, 4
    X = rng.normal(size=(P, P + 2))
    Cb = X @ X.T / 4 + 0.2 * np.eye(P)
    Cb[0, 2] = Cb[2, 0] = 0.6 * math.sqrt(Cb[0, 0] * Cb[2, 2])
    mt, Ct, nt, st, p = one_series([0.3, -1.2, 2.0], Cb, 12.0, 0.7, p=P, max_p=mp)
    N = 60000
    lam, th = orc.sample(1, mp, p, mt, Ct, nt, st, N, 77, 5)
    assert np.mean(lam) == pytest.approx(1 / 0.7, rel=0.01)
    z = (th[:P, 0, :] - mt[:P, :1]) * np.sqrt(0.7 * lam[0])
    assert np.allclose(np.cov(z), Cb, rtol=0.03, atol=0.03 * np.max(np.abs(Cb)))
    assert np.all(th[P:, 0, :] == mt[P:, :1])


def test_evolution_moments(oracle_lib):
    """sgo_evolve: lambda' = lambda * eta with E[eta] = 1 (Beta(beta r/2,
    (1-beta) r/2) / beta), and theta' - theta = chol(W)' omega / sqrt(c
    lambda') ~ N(0, W / (c lambda')) on the p_j block, zero beyond it."""
    rng = np.random.default_rng(6)
    P, mp, N = 3, 4, 60000
    X = rng.normal(size=(P, P + 2))
    Wb = X @ X.T / 5 + 0.1 * np.eye(P)
    W = np.zeros((mp, mp, 1), order="F")
    W[:P, :P, 0] = Wb
    lam = np.full((1, N), 2.0, order="F")
    th = np.zeros((mp, 1, N), order="F")
    r_t, c_t, beta = np.array([20.0]), np.array([0.5]), np.array([0.9])
    orc.lib().sgo_evolve(1, mp, orc.u(np.array([P], np.uint32)), orc.d(W), orc.d(r_t), orc.d(c_t),
                         orc.d(beta), None, N, 11, 3, orc.d(lam), orc.d(th))
    eta = lam[0] / 2.0
    assert np.mean(eta) == pytest.approx(1.0, rel=0.005)
    a, b = 0.9 * 20 / 2, 0.1 * 20 / 2            # Beta(a, b) / beta
    var_eta = a * b / ((a + b) ** 2 * (a + b + 1)) / 0.81
    assert np.var(eta) == pytest.approx(var_eta, rel=0.05)
    z = th[:P, 0, :] * np.sqrt(0.5 * lam[0])
    assert np.allclose(np.cov(z), Wb, rtol=0.03, atol=0.03 * np.max(np.abs(Wb)))
    assert np.all(th[P:, 0, :] == 0)


def test_evolution_with_G_is_G_theta(oracle_lib):
    """sgo_evolve with G (SGDLM.cu:729-740): beta = 1 (eta = 1) and W = 0
    (no noise) leave exactly theta' = G_j theta on every series, G_j read
    column-major ([j*max_p^2 + r + c*max_p]) -- a transposed G, a skipped
    column or a dropped padding row fails against numpy's product."""
    rng = np.random.default_rng(8)
    m, mp, N = 3, 4, 50
    p = np.array([4, 2, 3], dtype=np.uint32)
    G = np.zeros((mp, mp, m), order="F")
    for j in range(m):
        G[:, :, j] = rng.normal(size=(mp, mp))
    W = np.zeros((mp, mp, m), order="F")
    lam0 = rng.uniform(0.5, 2.0, size=(m, N))
    th0 = rng.normal(size=(mp, m, N))
    lam, th = np.asfortranarray(lam0.copy()), np.asfortranarray(th0.copy())
    r_t, c_t, beta = np.full(m, 10.0), np.full(m, 0.5), np.ones(m)
    orc.lib().sgo_evolve(m, mp, orc.u(p), orc.d(W), orc.d(r_t), orc.d(c_t), orc.d(beta),
                         orc.d(G), N, 11, 3, orc.d(lam), orc.d(th))
    assert np.array_equal(lam, lam0)
    for j in range(m):
        want = G[:, :, j] @ th0[:, j, :]
        assert np.allclose(th[:, j, :], want, rtol=1e-14, atol=1e-14), j
    # the transpose would differ (G_j is not symmetric)
    assert not np.allclose(G[:, :, 0].T @ th0[:, 0, :], th[:, 0, :])
