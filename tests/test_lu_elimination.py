"""Symbolic elimination plan of the log-determinant LU (lu_symbolic.cpp).

The plan (minimum-degree prefix, LDS slot tables, Schur-complement gather)
is executed on the CPU in FP64 by tests/cpp/elim_check.cpp in the kernel's
order, and its log|det(I - Gamma)| is compared with LAPACK's slogdet on
dense matrices.  A wrong slot, a missing fill entry, a dropped contribution
or a mis-ordered wave changes the determinant by far more than rounding.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

from gpusgdlm_amd import build as B
from gpusgdlm_amd.workloads import random_model, readme_model

LIB = os.path.join(B.LIB, "libelim_check.so")


@pytest.fixture(scope="module")
def chk():
    if not os.path.exists(LIB):
        pytest.fail("libelim_check.so not built: run python -m gpusgdlm_amd.build")
    lib = C.CDLL(LIB)
    ip = C.POINTER(C.c_int)
    dp = C.POINTER(C.c_double)
    lib.elim_check_logdet.restype = C.c_int
    lib.elim_check_logdet.argtypes = [C.c_int, C.c_int, ip, ip, ip, dp, C.c_int, C.c_int, dp, dp, ip]
    return lib


def entries_from_parents(m, parents, n_ext):
    """(row, col, idx) of -theta in I - Gamma, packed theta index
    off[j] + n_ext[j] + i for the i-th simultaneous parent of series j."""
    rows, cols, idx = [], [], []
    off = 0
    for j, par in enumerate(parents):
        for i, q in enumerate(par):
            rows.append(j)
            cols.append(q)
            idx.append(off + n_ext[j] + i)
        off += n_ext[j] + len(par)
    return np.array(rows, np.int32), np.array(cols, np.int32), np.array(idx, np.int32), off


def run(chk, m, parents, n_ext, theta, max_front=64, max_slots=65536):
    r, c, ix, _ = entries_from_parents(m, parents, n_ext)
    ld = C.c_double()
    amax = C.c_double()
    info = (C.c_int * 8)()
    ip = C.POINTER(C.c_int)
    bad = chk.elim_check_logdet(m, len(r), r.ctypes.data_as(ip), c.ctypes.data_as(ip), ix.ctypes.data_as(ip),
                                theta.ctypes.data_as(C.POINTER(C.c_double)), max_front, max_slots,
                                C.byref(ld), C.byref(amax), info)
    A = np.eye(m)
    A[r, c] = -theta[ix]
    sign, ref = np.linalg.slogdet(A)
    return bad, ld.value, ref, amax.value, list(info)[:7]


def _theta(P, scale, seed):
    return np.random.default_rng(seed).uniform(-scale, scale, size=max(P, 1))


@pytest.mark.parametrize("m,kmax,max_front", [(400, 20, 64), (400, 20, 32), (400, 20, 0), (100, 5, 64),
                                              (1000, 30, 64), (60, 59, 64), (60, 59, 8)])
def test_random_structures(chk, m, kmax, max_front):
    mod = random_model(m, kmax, T=3, seed=7)
    _, _, _, P = entries_from_parents(m, mod.parents, mod.n_ext)
    th = _theta(P, 0.3, 1)
    bad, ld, ref, amax, info = run(chk, m, mod.parents, mod.n_ext, th, max_front)
    n1, n2, M2, ns, nw, nst, nct = info
    assert n1 + n2 == m and M2 % 16 == 0 and M2 >= n2
    if max_front == 0:
        assert n1 == 0 and nw == 0
    assert bad == 0
    assert abs(ld - ref) <= 1e-10 * max(1.0, abs(ref)), (ld, ref, info)


def test_prefix_shrinks_c3_dense_block(chk):
    """m = 400, k ~ U{0..20}: the prefix removes about a third of the series
    from the dense LU (n2 ~ 264), i.e. ~3x fewer dense flops."""
    mod = random_model(400, 20, T=3, seed=20250101)
    _, _, _, P = entries_from_parents(400, mod.parents, mod.n_ext)
    bad, ld, ref, amax, info = run(chk, 400, mod.parents, mod.n_ext, _theta(P, 0.3, 2))
    assert info[1] <= 300 and info[4] <= 16
    assert abs(ld - ref) <= 1e-10 * abs(ref) + 1e-12


def test_readme_fully_sparse(chk):
    mod = readme_model(T=10)
    _, _, _, P = entries_from_parents(mod.m, mod.parents, mod.n_ext)
    th = _theta(P, 0.9, 3)
    bad, ld, ref, amax, info = run(chk, mod.m, mod.parents, mod.n_ext, th)
    assert info[0] == 6 and info[1] == 0 and info[2] == 0
    assert abs(ld - ref) <= 1e-13


@pytest.mark.parametrize("kind", ["none", "chain", "cycle", "star_in", "star_out", "single"])
def test_special_graphs(chk, kind):
    m = {"single": 1}.get(kind, 50)
    par = [[] for _ in range(m)]
    if kind == "chain":
        for j in range(1, m):
            par[j] = [j - 1]
    elif kind == "cycle":
        for j in range(m):
            par[j] = [(j - 1) % m]
    elif kind == "star_in":          # series 0 has every other series as parent
        par[0] = list(range(1, m))
    elif kind == "star_out":         # every series has series 0 as parent
        for j in range(1, m):
            par[j] = [0]
    n_ext = [1] * m
    P = sum(len(p) + 1 for p in par)
    th = _theta(P, 0.9, 4)
    for mf in (64, 4, 1, 0):
        bad, ld, ref, amax, info = run(chk, m, par, n_ext, th, mf)
        assert bad == 0
        assert abs(ld - ref) <= 1e-12 * max(1.0, abs(ref)), (kind, mf, ld, ref, info)
    if kind == "cycle":
        # det(I - Gamma) of a directed m-cycle = 1 - prod(gamma)
        g = th[[sum(len(p) + 1 for p in par[:j]) + 1 for j in range(m)]]
        assert abs(ld - np.log(abs(1 - np.prod(g)))) < 1e-12


def test_slot_limit_truncates_prefix(chk):
    mod = random_model(400, 20, T=3, seed=11)
    _, _, _, P = entries_from_parents(400, mod.parents, mod.n_ext)
    th = _theta(P, 0.3, 5)
    full = run(chk, 400, mod.parents, mod.n_ext, th, 64, 65536)
    small = run(chk, 400, mod.parents, mod.n_ext, th, 64, 1500)
    assert small[4][3] <= 1500 and small[4][0] < full[4][0]
    for res in (full, small):
        assert abs(res[1] - res[2]) <= 1e-10 * abs(res[2]) + 1e-12


def test_zero_pivot_is_flagged(chk):
    """(I - Gamma) with gamma_01 = gamma_10 = 1 has a zero pivot after the
    first elimination: the plan reports it (the kernel hands such a draw to
    the partial-pivoting path)."""
    par = [[1], [0], []]
    th = np.array([0.0, 1.0, 0.0, 1.0, 0.0])
    bad, ld, ref, amax, info = run(chk, 3, par, [1, 1, 1], th)
    assert bad == 1


@pytest.mark.parametrize("m,kmax,max_front", [(400, 20, 32), (400, 20, 0), (100, 5, 64), (60, 59, 8),
                                              (1000, 30, 32)])
def test_solve_through_plan(chk, m, kmax, max_front):
    """Forecast solves through the plan's prefix tables (forward over the
    waves, dense LU of S, backward over the waves in reverse) give
    numpy.linalg.solve's y."""
    ip = C.POINTER(C.c_int)
    dp = C.POINTER(C.c_double)
    chk.elim_check_solve.restype = C.c_int
    chk.elim_check_solve.argtypes = [C.c_int, C.c_int, ip, ip, ip, dp, C.c_int, dp, dp]
    mod = random_model(m, kmax, T=3, seed=12)
    r, c, ix, P = entries_from_parents(m, mod.parents, mod.n_ext)
    th = _theta(P, 0.3, 7)
    b = np.random.default_rng(8).normal(size=m)
    y = np.zeros(m)
    rc = chk.elim_check_solve(m, len(r), r.ctypes.data_as(ip), c.ctypes.data_as(ip), ix.ctypes.data_as(ip),
                              th.ctypes.data_as(dp), max_front, b.ctypes.data_as(dp), y.ctypes.data_as(dp))
    assert rc == 0
    A = np.eye(m)
    A[r, c] = -th[ix]
    ref = np.linalg.solve(A, b)
    assert np.max(np.abs(y - ref)) <= 1e-11 * max(1.0, np.max(np.abs(ref)))
