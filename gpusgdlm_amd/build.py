"""Builds the native libraries of gpusgdlm_amd in-tree.

  gpusgdlm_amd/lib/libsgdlm_hip.so   -- HIP kernels + C-ABI (include/sgdlm.h)
  gpusgdlm_amd/lib/libsgdlm_host.so  -- C++ SGDLM::HostWrapper<double> over it
  gpusgdlm_amd/lib/host_wrapper_driver -- test program of the C++ interface
                                          (tests/cpp/host_wrapper_driver.cpp)
  gpusgdlm_amd/lib/libelim_check.so    -- CPU execution of the LU elimination
                                          plan (tests/cpp/elim_check.cpp)

Usage: python -m gpusgdlm_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(LIB, "obj")

ARCH = "gfx950"
HIP_SOURCES = ["k_filter.hip", "k_draws.hip", "k_lu.hip", "k_lu_rs.hip", "k_lu_wave.hip", "k_vb.hip",
               "sgdlm_capi.hip", "lu_symbolic.cpp"]
HOST_SOURCES = ["host_wrapper.cpp"]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # bit-reproducible draws: no implicit FMA contraction anywhere
    "-ffp-contract=off",
    # no packed FP32 instructions (v_pk_*_f32) in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", os.path.join(ROOT, "include"),
]


def _hipcc() -> str:
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found")
    return p


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        sys.stderr.write(out + "\n")


def _stale(src: str, obj: str) -> bool:
    """obj older than src or than any header it may include."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    heads = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    heads.append(os.path.join(ROOT, "include", "sgdlm.h"))
    return any(os.path.getmtime(f) > t for f in [src] + heads)


def build(jobs: int = 4, verbose: bool = True, incremental: bool = False) -> str:
    """Compiles every source (incremental: only those older than their
    object file or a header) and links the libraries."""
    os.makedirs(OBJ, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        o = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        objs.append(o)
        if incremental and not _stale(os.path.join(CSRC, src), o):
            continue
        cmds.append([hipcc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, cmds))
    so = os.path.join(LIB, "libsgdlm_hip.so")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", so + ".tmp", *objs])
    os.replace(so + ".tmp", so)
    host = os.path.join(LIB, "libsgdlm_host.so")
    srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if srcs:
        _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-I", os.path.join(ROOT, "include"),
              *srcs, "-o", host + ".tmp", "-L", LIB, "-lsgdlm_hip", "-Wl,-rpath,$ORIGIN"])
        os.replace(host + ".tmp", host)
        drv_src = os.path.join(ROOT, "tests", "cpp", "host_wrapper_driver.cpp")
        if os.path.exists(drv_src):
            drv = os.path.join(LIB, "host_wrapper_driver")
            _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), drv_src,
                  "-o", drv + ".tmp", "-L", LIB, "-lsgdlm_host", "-Wl,-rpath,$ORIGIN"])
            os.replace(drv + ".tmp", drv)
    chk_src = os.path.join(ROOT, "tests", "cpp", "elim_check.cpp")
    if os.path.exists(chk_src):
        chk = os.path.join(LIB, "libelim_check.so")
        _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", chk_src,
              os.path.join(CSRC, "lu_symbolic.cpp"), "-o", chk + ".tmp"])
        os.replace(chk + ".tmp", chk)
    if verbose:
        print(f"built {so}")
    return so


def build_oracle() -> str:
    d = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", d])
    return os.path.join(d, "_build", "libsgdlm_oracle.so")


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--incremental", action="store_true",
                    help="compile only sources older than their object or a header")
    a = ap.parse_args()
    build(a.jobs, incremental=a.incremental)
