"""CPU-side checks of the product: the C-ABI library loads and exports every
function include/sgdlm.h declares, the Python mirror's host helpers encode
the reference layouts, and the workload generators build the reference's
README structure.  No compute call is made (no GPU here)."""
import ctypes as C
import os
import re

import numpy as np

from gpusgdlm_amd import NA_INDEX, encode_sp
from gpusgdlm_amd import _native as nat
from gpusgdlm_amd.workloads import random_model, readme_model

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "sgdlm.h")).read()
    return sorted(set(re.findall(r"\b(sgdlm_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(hiplib):
    names = header_functions()
    assert len(names) >= 35
    for n in names:
        assert hasattr(hiplib, n), n
    assert set(names) == set(nat.SIGNATURES), set(names) ^ set(nat.SIGNATURES)


def test_encode_sp_reference_layout():
    sp = [[None, 3], [7, float("nan")], [-1, 12]]
    a = encode_sp(sp)
    assert a.shape == (3, 2) and a.flags.f_contiguous
    assert a.ravel(order="F").tolist() == [NA_INDEX, 7, NA_INDEX, 3, NA_INDEX, 12]


def test_readme_structure_positions():
    mod = readme_model(T=10)
    assert mod.p.tolist() == [3, 3, 1, 5, 1, 4] and mod.max_p == 5
    # README.md:66-77: sp[count_j + i, j] = helper_pos_Gamma[j, parent] = j + parent*m
    assert mod.sp[2, 0] == 0 + 5 * 6
    assert [mod.sp[k, 3] for k in (3, 4)] == [3 + 0 * 6, 3 + 1 * 6]
    flat = mod.sp_u32.ravel(order="F")
    sim = flat[flat < 36]
    assert sorted(sim.tolist()) == sorted([30, 1 + 0, 1 + 30, 1 + 18, 3, 9, 5 + 6, 5 + 18])
    F = mod.F(5)
    assert F[2, 0] == mod.y[5, 5] and F[0, 0] == 1.0 and F[1, 0] == mod.y[4, 0]


def test_random_model_shapes_and_stability():
    mod = random_model(40, 6, T=20, seed=2)
    assert mod.max_p == int(mod.p.max()) and np.all(mod.p >= 3)
    assert max(abs(np.linalg.eigvals(mod.gamma_true))) < 0.8
    flat = mod.sp_u32.ravel(order="F")
    pos = flat[flat < 40 * 40]
    assert len(pos) == len(set(pos.tolist()))
    assert np.all(pos % 40 != pos // 40)
