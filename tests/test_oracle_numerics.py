"""Pins of the shared deterministic numerics (Philox, elementary functions,
samplers, Cholesky) used by both the GPU kernels and the oracle.

Independent references: the Random123 known-answer vectors for
Philox4x32-10, rocRAND's host engine (tests/golden/philox_rocrand_kat.json),
libm / numpy for log, exp, sin, cos, scipy.stats distributions for the
samplers, numpy.linalg.cholesky for the factor."""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest
from scipy import special, stats

from oracle import oracle as orc

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_philox_random123_kat(oracle_lib):
    # kat_vectors of Random123 (philox4x32, 10 rounds)
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, out in kat:
        assert orc.philox(ctr, key) == list(out)


def test_philox_rocrand_fixture(oracle_lib):
    with open(os.path.join(GOLDEN, "philox_rocrand_kat.json")) as f:
        cases = json.load(f)
    assert len(cases) >= 5
    for c in cases:
        assert orc.philox(c["ctr"], c["key"]) == c["out"], c


def test_log_exp_sincospi_accuracy(oracle_lib):
    L = orc.lib()
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.uniform(1e-16, 1, 3000), rng.uniform(1, 1e6, 1000),
                         10.0 ** rng.uniform(-300, 300, 1000), [2.0 ** -53, 1.0, 2.0, 4e-320]])
    for x in xs:
        ref = math.log(x)
        assert abs(L.sgo_log(x) - ref) <= 2.5 * np.spacing(abs(ref)) + 1e-300, x
    assert L.sgo_log(0.0) == -math.inf and math.isnan(L.sgo_log(-1.0))
    for x in np.concatenate([rng.uniform(-700, 700, 3000), [0.0, -745.0, 709.0, 1e-10]]):
        ref = math.exp(x)
        assert abs(L.sgo_exp(x) - ref) <= 2.0 * np.spacing(ref) + 5e-324, x
    s, c = C.c_double(), C.c_double()
    for x in np.concatenate([rng.uniform(0, 2, 3000), [0.0, 0.25, 0.5, 0.75, 1.0, 1.5, 2.0]]):
        L.sgo_sincospi(x, C.byref(s), C.byref(c))
        assert abs(s.value - math.sin(math.pi * x)) <= 1e-15
        assert abs(c.value - math.cos(math.pi * x)) <= 1e-15
    L.sgo_sincospi(0.5, C.byref(s), C.byref(c))
    assert s.value == 1.0 and c.value == 0.0


def _normals(n, purpose=2, epoch=3):
    L = orc.lib()
    return np.array([L.sgo_normal_k(1234, epoch, purpose, i // 7, 3, i % 7) for i in range(n)])


def test_normal_distribution(oracle_lib):
    z = _normals(40000)
    assert stats.kstest(z, "norm").pvalue > 1e-3
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1) < 0.02
    # the two Box-Muller branches are uncorrelated
    L = orc.lib()
    a = np.array([L.sgo_normal_k(9, 1, 2, i, 0, 0) for i in range(20000)])
    b = np.array([L.sgo_normal_k(9, 1, 2, i, 0, 1) for i in range(20000)])
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.03


@pytest.mark.parametrize("shape", [0.3, 0.8, 1.0, 2.5, 8.0, 60.0])
def test_gamma_sampler_ks(oracle_lib, shape):
    L = orc.lib()
    rate = 2.0
    g = np.array([L.sgo_gamma(shape, rate, 77, 5, 1, i, 4) for i in range(20000)])
    assert np.all(g > 0)
    assert stats.kstest(g, stats.gamma(shape, scale=1 / rate).cdf).pvalue > 1e-3


def test_gamma_invalid_parameters_nan(oracle_lib):
    L = orc.lib()
    assert math.isnan(L.sgo_gamma(-1.0, 1.0, 1, 1, 1, 0, 0))
    assert math.isnan(L.sgo_gamma(2.0, 0.0, 1, 1, 1, 0, 0))


def test_beta_eta_distribution(oracle_lib):
    L = orc.lib()
    r, beta = 30.0, 0.9
    eta = np.array([L.sgo_beta_eta(r, beta, 11, 2, i, 1) for i in range(20000)])
    # eta * beta ~ Beta(beta r / 2, (1 - beta) r / 2)  (kernel_functions.cuh:655,737,754-787)
    assert stats.kstest(eta * beta, stats.beta(beta * r / 2, (1 - beta) * r / 2).cdf).pvalue > 1e-3
    assert L.sgo_beta_eta(r, 1.0, 11, 2, 0, 1) == 1.0


def test_cholesky_matches_numpy(oracle_lib):
    rng = np.random.default_rng(2)
    for p, max_p in [(1, 1), (3, 5), (7, 7), (12, 20)]:
        X = rng.normal(size=(p, p + 3))
        S = X @ X.T / 10
        C_ = np.zeros((max_p, max_p))
        C_[:p, :p] = S
        C_[p:, :] = 7.0  # junk outside the block must be ignored
        U = np.zeros((max_p, max_p))
        orc.lib().sgo_cholesky_upper(p, max_p, orc.d(np.asfortranarray(C_)), orc.d(U))
        U = U.reshape((max_p, max_p), order="C").T  # buffer is column-major
        assert np.allclose(U[:p, :p], np.linalg.cholesky(S).T, rtol=1e-13, atol=1e-14)
        assert np.all(U[p:, :] == 0) and np.all(U[:, p:] == 0)


def test_psi_series_error_reported(oracle_lib):
    """The truncated asymptotic psi / psi1 series of the reference
    (kernel_functions.cuh:1006-1022) are replicated, not corrected: they are
    accurate only for moderate arguments.  scipy quantifies the error; the
    moment-match uses n/2 >= ~1 in practice."""
    L = orc.lib()
    for x, tol in [(5.0, 1e-9), (10.0, 1e-12), (50.0, 1e-14)]:
        assert abs(L.sgo_psi(x) - special.digamma(x)) < tol
        assert abs(L.sgo_psi1(x) - special.polygamma(1, x)) < tol * 10
    assert abs(L.sgo_psi(0.5) - special.digamma(0.5)) > 1e-3  # documented inaccuracy
