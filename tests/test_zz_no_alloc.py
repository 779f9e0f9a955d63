"""The allocation-free VB entry point (sgdlm_compute_vb_posterior_noalloc,
no counterpart in the reference) against sgdlm_compute_vb_posterior, called
through the C-ABI on the same model and RNG state: two calls in a row, the
second with fewer draws, each bit-identical to the allocating entry point;
a size beyond the static scratch array fails with SGDLM_ERR_ARG."""
import ctypes as C

import numpy as np
import pytest

from gpusgdlm_amd import SGDLM
from gpusgdlm_amd import _native as nat
from gpusgdlm_amd.workloads import configure, filter_steps, random_model, readme_model

pytestmark = pytest.mark.gpu


def _outputs(m, mp, N):
    return (np.zeros((mp, m), order="F"), np.zeros((mp, mp, m), order="F"), np.zeros(m),
            np.zeros(m), np.zeros(N))


def _call(lib, g, name, N, epoch):
    m, mp = g._dims()
    mt, Ct, nt, st, w = _outputs(m, mp, N)
    sd = C.c_double()
    g.set_rng_state(77, epoch)
    if name == "noalloc":
        need = C.c_size_t()
        rc = lib.sgdlm_compute_vb_posterior_noalloc(g._h, C.c_size_t(N), nat.dptr(mt), nat.dptr(Ct),
                                                    nat.dptr(nt), nat.dptr(st), nat.dptr(w),
                                                    C.byref(sd), C.byref(need))
    else:
        assert g._ensure_sim(N, N)
        rc = lib.sgdlm_compute_vb_posterior(g._h, nat.dptr(mt), nat.dptr(Ct), nat.dptr(nt),
                                            nat.dptr(st), nat.dptr(w), C.byref(sd))
    assert rc == nat.SGDLM_OK, lib.sgdlm_last_error(g._h).decode()
    return {"m": mt, "C": Ct, "n": nt, "s": st, "w": w, "sd": sd.value, "state": g.get_rng_state()}


@pytest.mark.parametrize("kind,prec", [("readme", "fp64"), ("m100", "fp32")])
def test_noalloc_matches_allocating_entry_point(hiplib, kind, prec):
    mod = readme_model(T=30) if kind == "readme" else random_model(100, 5, T=30, seed=11)
    g = SGDLM(strict=True)
    g.set_precision(prec)
    # the first allocating call would otherwise tune the LU configuration
    # (the allocation-free one never does): both sides on the same plan
    g.set_lu_autotune(False)
    configure(g, mod)
    filter_steps(g, mod, 1, 20)
    g.computePosterior(mod.y[20], mod.F(20))
    sizes = (6000, 2500) if kind == "readme" else (4000, 1500)
    for i, N in enumerate(sizes):
        print(f"noalloc {kind} {prec} call {i + 1}: N={N}", flush=True)
        a = _call(hiplib, g, "noalloc", N, 5 + i)
        b = _call(hiplib, g, "alloc", N, 5 + i)
        for k in ("m", "C", "n", "s", "w"):
            assert np.array_equal(a[k], b[k]), (N, k)
        assert a["sd"] == b["sd"] and a["state"] == b["state"]
        assert abs(np.sum(a["w"]) - 1.0) < 1e-12
    g.close()


def test_noalloc_rejects_oversized_request(hiplib):
    mod = random_model(100, 5, T=10, seed=11)
    g = SGDLM(strict=True)
    configure(g, mod)
    g.computePosterior(mod.y[1], mod.F(1))
    m, mp = g._dims()
    N = 400000
    mt, Ct, nt, st, w = _outputs(m, mp, N)
    need = C.c_size_t()
    rc = hiplib.sgdlm_compute_vb_posterior_noalloc(g._h, C.c_size_t(N), nat.dptr(mt), nat.dptr(Ct),
                                                   nat.dptr(nt), nat.dptr(st), nat.dptr(w), None,
                                                   C.byref(need))
    assert rc == nat.SGDLM_ERR_ARG
    assert need.value > 64 << 20
    assert "scratch" in hiplib.sgdlm_last_error(g._h).decode()
    g.close()
