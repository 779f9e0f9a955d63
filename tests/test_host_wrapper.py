"""The C++ plugin interface SGDLM::HostWrapper<double> (libsgdlm_host.so).

The reference's R package reaches the engine only through
HostWrapperFactory<double>::create + the virtual methods of HostWrapper
(rSGDLM/src/SGDLM/HostWrapper.hpp:13-99, RcppWrapper.cpp).  The driver
program tests/cpp/host_wrapper_driver.cpp makes the Rcpp layer's call
sequence through that interface; its outputs must equal, bit for bit, those
of the Python binding (which calls the C-ABI directly) on the same inputs.
"""
from __future__ import annotations

import os
import struct
import subprocess

import numpy as np
import pytest

from gpusgdlm_amd import build as B
from gpusgdlm_amd.workloads import readme_model

DRIVER = os.path.join(B.LIB, "host_wrapper_driver")
HOSTLIB = os.path.join(B.LIB, "libsgdlm_host.so")


def _require_built():
    if not (os.path.exists(DRIVER) and os.path.exists(HOSTLIB)):
        pytest.fail("libsgdlm_host.so / host_wrapper_driver not built: run python -m gpusgdlm_amd.build")


def write_input(path, mod, T, nsim):
    m, mp = mod.m, mod.max_p
    with open(path, "wb") as f:
        f.write(struct.pack("<4Q", m, mp, T, nsim))
        f.write(np.asarray(mod.p, dtype="<u4").tobytes())
        f.write(np.asarray(mod.sp_u32, dtype="<u4").tobytes(order="F"))
        for a in (mod.beta, mod.delta, mod.a0, mod.R0, mod.r0, mod.c0):
            f.write(np.asarray(a, dtype="<f8").tobytes(order="F"))
        for t in range(1, T + 1):
            f.write(np.asarray(mod.y[t], dtype="<f8").tobytes())
            f.write(np.asarray(mod.F(t), dtype="<f8").tobytes(order="F"))
        f.write(np.asarray(mod.F_forecast(T + 1), dtype="<f8").tobytes(order="F"))


def read_output(path, m, mp, nsim):
    raw = open(path, "rb").read()
    sizes = [("pm", mp * m), ("pC", mp * mp * m), ("pn", m), ("ps", m),
             ("vm", mp * m), ("vC", mp * mp * m), ("vn", m), ("vs", m), ("w", nsim), ("sd", 1),
             ("qm", mp * m), ("qC", mp * mp * m), ("qn", m), ("qs", m),
             ("fy", nsim * m), ("fl", nsim * m), ("ft", nsim * m * mp)]
    out, off = {}, 0
    for k, n in sizes:
        out[k] = np.frombuffer(raw, dtype="<f8", count=n, offset=off).copy()
        off += 8 * n
    out["meta"] = np.frombuffer(raw, dtype="<u8", count=5, offset=off).copy()
    assert off + 40 == len(raw)
    return out


def test_host_library_exports_factory():
    """libsgdlm_host.so defines the three HostWrapperFactory<double>::create
    overloads of HostWrapperFactory.hpp:15-24 (Itanium-mangled)."""
    _require_built()
    r = subprocess.run(["nm", "-D", "--defined-only", HOSTLIB], capture_output=True, text=True, check=True)
    syms = set(l.split()[-1] for l in r.stdout.splitlines() if l.strip())
    for s in ("_ZN5SGDLM18HostWrapperFactoryIdE6createEm",
              "_ZN5SGDLM18HostWrapperFactoryIdE6createEmmm",
              "_ZN5SGDLM18HostWrapperFactoryIdE6createEmmmb"):
        assert s in syms, s


def test_driver_without_device_logs_and_returns(tmp_path):
    """With no usable GPU every method logs an error and returns (the
    reference's myAssert behaviour, logger.hpp:160-165); nothing crashes."""
    _require_built()
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    mod = readme_model(T=20)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(inp, mod, 3, 64)
    r = subprocess.run([DRIVER, str(inp), str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "[SGDLM] ERROR" in r.stderr
    res = read_output(out, mod.m, mod.max_p, 64)
    assert not res["fy"].any() and res["meta"][0] == 0


@pytest.mark.gpu
def test_host_wrapper_matches_python_binding(tmp_path):
    _require_built()
    from gpusgdlm_amd import SGDLM
    from gpusgdlm_amd.workloads import configure
    mod = readme_model(T=60)
    T, nsim = 40, 3000
    m, mp = mod.m, mod.max_p
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(inp, mod, T, nsim)
    print(f"host_wrapper_driver readme T={T} nsim={nsim}", flush=True)
    # bit-for-bit comparison across two processes: the LU configuration must
    # be the same on both sides, so neither tunes it by timing
    env = dict(os.environ, SGDLM_LU_AUTOTUNE="0")
    r = subprocess.run([DRIVER, str(inp), str(out)], capture_output=True, text=True, timeout=120,
                       env=env)
    assert r.returncode == 0 and "[SGDLM] ERROR" not in r.stderr, r.stderr
    res = read_output(out, m, mp, nsim)
    assert list(res["meta"]) == [m, mp, nsim, nsim, 1]

    g = SGDLM(strict=True)
    g.set_lu_autotune(False)
    configure(g, mod)
    for t in range(1, T + 1):
        g.computePosterior(mod.y[t], mod.F(t))
        if t < T:
            g.computePrior()
    post = g.getParameters()
    vb = g.computeVBPosterior(nsim, nsim)
    g.computePrior()
    prior = g.getParameters()
    fc = g.computeForecastDebug(nsim, nsim, mod.F_forecast(T + 1))
    g.close()

    f = lambda a: np.asarray(a).reshape(-1, order="F")
    for k, ref in (("pm", post["m"]), ("pC", post["C"]), ("pn", post["n"]), ("ps", post["s"]),
                   ("vm", vb["m"]), ("vC", vb["C"]), ("vn", vb["n"]), ("vs", vb["s"]),
                   ("w", vb["IS_weights"]), ("sd", [vb["sum_det_weights"]]),
                   ("qm", prior["m"]), ("qC", prior["C"]), ("qn", prior["n"]), ("qs", prior["s"]),
                   ("fy", fc["y"]), ("fl", fc["lambdas"]), ("ft", fc["thetas"])):
        np.testing.assert_array_equal(res[k], f(ref), err_msg=k)
    assert np.isfinite(res["fy"]).all() and res["w"].sum() == pytest.approx(1.0, rel=1e-12)
