"""Synthetic SGDLM workloads (seeded), in the reference's R layouts.

readme_model()     -- the 6-series example of the reference README
                      (README.md:37-154): parents, discounts, prior, data and
                      the per-step regressor matrices F_t / F_tp1.
random_model(...)  -- m series with 3 external regressors (intercept, AR1,
                      one N(0,1) covariate) and k_j ~ U{0..kmax} simultaneous
                      parents drawn without replacement from the other series;
                      data from a true SGDLM with Gamma entries U(-0.3, 0.3)
                      rescaled to spectral radius < 0.8, phi ~ N(0, 0.1^2),
                      lambda ~ Gamma(shape 20, rate 20*400) (mean precision
                      1/400, the shape/rate convention of
                      kernel_functions.cuh:579-581).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

NA = None


@dataclass
class Model:
    m: int
    max_p: int
    p: np.ndarray                 # (m,) uint32
    sp: np.ndarray                # (max_p, m) object/int with NA = None
    sp_u32: np.ndarray            # (max_p, m) uint32 Fortran, NA -> 2^31
    beta: np.ndarray              # (m,)
    delta: np.ndarray             # (max_p, max_p, m)
    a0: np.ndarray                # (max_p, m)
    R0: np.ndarray                # (max_p, max_p, m)
    r0: np.ndarray                # (m,)
    c0: np.ndarray                # (m,)
    y: np.ndarray                 # (T, m)
    parents: list = field(default_factory=list)   # simultaneous parents (0-based) per series
    n_ext: np.ndarray = None      # external regressor count per series
    X: np.ndarray = None          # (T, m, 3) external regressors (random models)
    gamma_true: np.ndarray = None

    def F(self, t: int) -> np.ndarray:
        """Regressors of time t (simultaneous parent slots hold y_t)."""
        return self._F(t, self.y[t])

    def F_forecast(self, t: int) -> np.ndarray:
        """Regressors for forecasting time t from t-1 (parent slots 0)."""
        return self._F(t, None)

    def _F(self, t, yt):
        raise NotImplementedError


class ReadmeModel(Model):
    def _F(self, t, yt):
        y = self.y
        Ft = np.zeros((self.max_p, self.m), order="F")
        Ft[0:2, 0] = (1.0, y[t - 1, 0])
        Ft[0, 2] = 1.0
        Ft[0:3, 3] = (1.0, y[t - 1, 3], self.true_means[t])
        Ft[0, 4] = 1.0
        Ft[0:2, 5] = (1.0, y[t - 1, 0])
        if yt is not None:
            for j, par in enumerate(self.parents):
                for i, q in enumerate(par):
                    Ft[self.n_ext[j] + i, j] = yt[q]
        return Ft


class RandomModel(Model):
    def _F(self, t, yt):
        Ft = np.zeros((self.max_p, self.m), order="F")
        Ft[0:3, :] = self.X[t].T
        Ft[1, :] = self.y[t - 1]
        if yt is not None:
            for j, par in enumerate(self.parents):
                for i, q in enumerate(par):
                    Ft[3 + i, j] = yt[q]
        return Ft


def _encode(sp):
    from .sgdlm import encode_sp
    return encode_sp(sp)


def readme_model(T: int = 500, seed: int = 20250101) -> ReadmeModel:
    m = 6
    count = np.array([2, 0, 1, 3, 1, 2])
    sim = [[5], [0, 5, 3], [], [0, 1], [], [1, 3]]   # README lists, 0-based
    p = (count + np.array([len(s) for s in sim])).astype(np.uint32)
    max_p = int(p.max())
    sp = np.full((max_p, m), None, dtype=object)
    for j in range(m):
        for i, q in enumerate(sim[j]):
            sp[count[j] + i, j] = j + q * m          # helper_pos_Gamma[j, q]
    beta = np.full(m, 0.95)
    delta = np.full((max_p, max_p, m), 0.95, order="F")
    delta[:, :, 2] = 0.7
    delta[:, :, 4] = 0.9999
    a0 = np.zeros((max_p, m), order="F")
    R0 = np.asfortranarray(np.repeat((0.05 * np.eye(max_p))[:, :, None], m, axis=2))
    rng = np.random.default_rng(seed)
    tm = np.concatenate([np.full(T // 5, v) for v in (0, 0.1, 0.2, 0.15, 0.05)])
    y = rng.normal(tm[:, None], 0.025, size=(T, m))
    mod = ReadmeModel(m=m, max_p=max_p, p=p, sp=sp, sp_u32=_encode(sp), beta=beta, delta=delta,
                      a0=a0, R0=R0, r0=np.ones(m), c0=np.full(m, 0.01), y=y, parents=sim,
                      n_ext=count)
    mod.true_means = tm
    return mod


def random_model(m: int, kmax: int, T: int, seed: int = 20250101, beta: float = 0.97,
                 delta: float = 0.98) -> RandomModel:
    rng = np.random.default_rng(seed)
    parents = []
    for j in range(m):
        k = int(rng.integers(0, min(kmax, m - 1) + 1))
        others = np.delete(np.arange(m), j)
        parents.append(sorted(rng.choice(others, size=k, replace=False).tolist()))
    p = np.array([3 + len(q) for q in parents], dtype=np.uint32)
    max_p = int(p.max())
    sp = np.full((max_p, m), None, dtype=object)
    G = np.zeros((m, m))
    for j, par in enumerate(parents):
        for i, q in enumerate(par):
            sp[3 + i, j] = j + q * m
            G[j, q] = rng.uniform(-0.3, 0.3)
    rho = max(abs(np.linalg.eigvals(G))) if m > 1 else 0.0
    if rho >= 0.8:
        G *= 0.79 / rho
    phi = rng.normal(0.0, 0.1, size=(m, 3))
    lam = rng.gamma(20.0, 1.0 / (20.0 * 400.0), size=m)
    X = np.empty((T, m, 3))
    X[:, :, 0] = 1.0
    X[:, :, 2] = rng.normal(size=(T, m))
    y = np.zeros((T, m))
    IG = np.eye(m) - G
    for t in range(T):
        X[t, :, 1] = y[t - 1] if t > 0 else 0.0
        mu = np.einsum("jk,jk->j", X[t], phi)
        y[t] = np.linalg.solve(IG, mu + rng.normal(size=m) / np.sqrt(lam))
    a0 = np.zeros((max_p, m), order="F")
    R0 = np.asfortranarray(np.repeat((0.05 * np.eye(max_p))[:, :, None], m, axis=2))
    mod = RandomModel(m=m, max_p=max_p, p=p, sp=sp, sp_u32=_encode(sp), beta=np.full(m, beta),
                      delta=np.full((max_p, max_p, m), delta, order="F"), a0=a0, R0=R0,
                      r0=np.ones(m), c0=np.full(m, 0.01), y=y, parents=parents,
                      n_ext=np.full(m, 3), X=X, gamma_true=G)
    return mod


def configure(model_api, mod: Model):
    """setSimultaneousParents / setDiscountFactors / setPriorParameters."""
    model_api.setSimultaneousParents(mod.p, mod.sp)
    model_api.setDiscountFactors(mod.beta, mod.delta)
    model_api.setPriorParameters(mod.a0, mod.R0, mod.r0, mod.c0)


def filter_steps(model_api, mod: Model, t0: int, t1: int):
    """Naive filtering (computePosterior + computePrior) over t0..t1-1."""
    for t in range(t0, t1):
        model_api.computePosterior(mod.y[t], mod.F(t))
        model_api.computePrior()


def instance_sizes(K: int, seed: int = 20250102, mmin: int = 20, mmax: int = 400) -> list:
    """Series counts of K independent models, log-uniform on [mmin, mmax]
    (m = round(exp(U(log mmin, log mmax))))."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(np.log(mmin), np.log(mmax), size=K)
    return [int(min(mmax, max(mmin, round(float(np.exp(x)))))) for x in u]


def instance_models(K: int, T: int, seed: int = 20250102, mmin: int = 20, mmax: int = 400,
                    kmax: int = 20) -> list:
    """K independent random models (each with its own data seed), m
    log-uniform on [mmin, mmax], k_j ~ U{0..min(kmax, m - 1)} parents."""
    return [random_model(m, min(kmax, m - 1), T=T, seed=seed + 1 + i)
            for i, m in enumerate(instance_sizes(K, seed, mmin, mmax))]
