"""VB steps of many independent models as one launch per stage
(sgdlm_batch_compute_vb_posterior, group path) against one
computeVBPosterior per model, on the GPU.

The group path runs every model's stages from device tables (one launch per
stage; one per kernel variant for the LU, workgroups split by work), so its
results must equal the one-call-per-model results bit for bit: the same
Philox seeds and epochs, the same LU plan per model, the same kernels'
arithmetic.  Covered: models of uneven m (20..400) and uneven draw counts,
every LU kernel kind (one-wavefront, register, blocked + fallback,
pivoting for every draw), early-series draws that mostly take the fallback,
all three precisions, the one-stream-per-model mode and graph replays.
"""
import ctypes as C

import numpy as np
import pytest

from gpusgdlm_amd import SGDLM, batch_vb
from gpusgdlm_amd import _native as nat
from gpusgdlm_amd.workloads import configure, filter_steps, instance_models, random_model

pytestmark = pytest.mark.gpu

KEYS = ("m", "C", "n", "s", "IS_weights", "log_weights")


def _ctx(mod, prec, T, seed, kind="auto"):
    g = SGDLM(strict=True)
    g.set_rng_state(seed, 0)
    g.set_precision(prec)
    configure(g, mod)
    if kind != "auto":
        g.set_lu_kernel(kind)
    filter_steps(g, mod, 1, T)
    g.computePosterior(mod.y[T], mod.F(T))
    return g


def _batch(models, nsims):
    """The batched entry point with a draw count per model."""
    for g, n in zip(models, nsims):
        assert g._ensure_sim(int(n), int(n))
    lib = nat.load()
    arr = (C.c_void_p * len(models))(*[g._h.value for g in models])
    st = (C.c_int * len(models))()
    rc = lib.sgdlm_batch_compute_vb_posterior(arr, C.c_size_t(len(models)), st)
    return rc, list(st)


def _same(a, b, what):
    """Bit for bit, NaN in the same places included (a degenerate early
    model can give NaN moments in both)."""
    for k in KEYS:
        assert np.array_equal(a[k], b[k], equal_nan=True), (what, k)
    assert a["ess"] == b["ess"] and a["log_sum_det"] == b["log_sum_det"], what


def _check_steps(batched, single, nsims, steps, label):
    for step in range(steps):
        print(f"{label}: batched step {step}", flush=True)
        rc, codes = _batch(batched, nsims)
        assert rc == 0 and codes == [0] * len(batched), codes
        for g, n in zip(single, nsims):
            g.vb_device(n, n)
        for i, (a, b) in enumerate(zip(batched, single)):
            _same(a.get_vb_result(), b.get_vb_result(), f"{label} model {i} step {step}")
            assert a.get_rng_state() == b.get_rng_state()
            assert a.lu_fallback_count() == b.lu_fallback_count(), (label, i, step)


@pytest.mark.parametrize("prec", ["fp32", "fp64", "bf16x3"])
def test_group_step_equals_one_call_per_model(hiplib, prec):
    """Eight models with m log-uniform on [20, 400] plus two small ones (the
    one-wavefront kernel's sizes), uneven draw counts, three batched steps
    after the one that tunes the LU plans."""
    T = 40
    mods = instance_models(8, T=T + 2, seed=41) + [random_model(24, 4, T=T + 2, seed=43),
                                                   random_model(60, 8, T=T + 2, seed=44)]
    print(f"{prec}: m = {[md.m for md in mods]}", flush=True)
    nsims = [2000 + 250 * (i % 4) for i in range(len(mods))]
    ga = [_ctx(md, prec, T, 100 + i) for i, md in enumerate(mods)]
    gb = [_ctx(md, prec, T, 100 + i) for i, md in enumerate(mods)]
    # first call: the LU autotuner times each structure (one stream per model)
    rc, codes = _batch(ga, nsims)
    assert rc == 0, codes
    for g, n in zip(gb, nsims):
        g.vb_device(n, n)
    _check_steps(ga, gb, nsims, 3, prec)
    st = ga[0].batch_stats()
    print(f"  batch stats {st}, kernels {[g.lu_kernel()['kernel'] for g in ga]}", flush=True)
    assert st["group"] == 3 and st["streams"] == 1
    for g in ga + gb:
        g.close()


def test_group_step_every_lu_kind(hiplib):
    """Each model pins an LU kernel kind (one-wavefront on the whole matrix
    or after the prefix, register with and without interchanges, blocked,
    pivoting for every draw); some
    models are early in their series, so many of their draws are refactored
    by the fallback kernel."""
    T = 40
    spec = [(40, 6, "wave", T), (120, 12, "blocked", T), (200, 16, "register", T),
            (70, 10, "pivoting", T), (160, 12, "blocked", 3), (300, 20, "register", 3),
            (150, 14, "wave_prefix", T), (120, 12, "wave_prefix", 3),
            (200, 16, "register_pivoting", 3)]
    mods = [random_model(m, k, T=t + 2, seed=60 + i) for i, (m, k, _, t) in enumerate(spec)]
    nsims = [1500, 2500, 2000, 1000, 2000, 1500, 2000, 1500, 1500]
    ga = [_ctx(md, "fp32", t, 200 + i, kind) for i, (md, (_, _, kind, t)) in enumerate(zip(mods, spec))]
    gb = [_ctx(md, "fp32", t, 200 + i, kind) for i, (md, (_, _, kind, t)) in enumerate(zip(mods, spec))]
    print("kinds", [g.lu_kernel() for g in ga], flush=True)
    _check_steps(ga, gb, nsims, 2, "pinned kinds")
    fb = [g.lu_fallback_count() for g in ga]
    print(f"  fallback draws {fb}", flush=True)
    assert fb[3] == nsims[3]            # pivoting: every draw counts
    assert fb[0] == 0                    # the one-wavefront kernel pivots itself
    assert fb[4] > nsims[4] // 4 and fb[5] > nsims[5] // 4   # early in the series
    assert ga[0].batch_stats()["group"] == 2
    for g in ga + gb:
        g.close()


def test_stream_mode_equals_group(hiplib):
    """sgdlm_set_batch_mode(SGDLM_BATCH_STREAMS) runs each model on its own
    stream; the results are the group path's, bit for bit."""
    T = 30
    mods = instance_models(5, T=T + 2, seed=81)
    nsims = [2000] * len(mods)
    ga = [_ctx(md, "fp32", T, 300 + i) for i, md in enumerate(mods)]
    gb = [_ctx(md, "fp32", T, 300 + i) for i, md in enumerate(mods)]
    gb[0].set_batch_mode("streams")
    for step in range(3):
        print(f"streams vs group step {step}", flush=True)
        assert _batch(ga, nsims)[0] == 0
        assert _batch(gb, nsims)[0] == 0
        for i, (a, b) in enumerate(zip(ga, gb)):
            _same(a.get_vb_result(), b.get_vb_result(), f"model {i} step {step}")
    assert ga[0].batch_stats()["group"] == 2
    assert gb[0].batch_stats()["group"] == 0 and gb[0].batch_stats()["streams"] == 3
    for g in ga + gb:
        g.close()


def test_group_tables_rebuilt_on_change(hiplib):
    """The cached device tables follow the contexts: another draw count, a
    new seed and a different list each give a fresh build whose results
    equal single calls."""
    T = 30
    mods = instance_models(4, T=T + 2, seed=91)
    ga = [_ctx(md, "fp32", T, 400 + i) for i, md in enumerate(mods)]
    gb = [_ctx(md, "fp32", T, 400 + i) for i, md in enumerate(mods)]
    _check_steps(ga, gb, [1500] * 4, 2, "first list")
    _check_steps(ga, gb, [2500, 1500, 1000, 3000], 1, "new draw counts")
    for i, (a, b) in enumerate(zip(ga, gb)):
        a.set_rng_state(999 + i, 5)
        b.set_rng_state(999 + i, 5)
    _check_steps(ga, gb, [2500, 1500, 1000, 3000], 1, "new seeds")
    _check_steps(ga[1:], gb[1:], [1500, 1000, 3000], 1, "sub-list")
    assert ga[0].batch_stats()["group"] >= 3
    for g in ga + gb:
        g.close()


def test_tuned_front_is_the_built_one(hiplib):
    """A structure whose LU tables do not fit LDS at the autotuner's front 48
    (the builder halves it to 24): the tuned configuration records the front
    the plan was built with, so the structure counts as tuned and its batches
    run as one launch per stage."""
    mods = instance_models(5, T=32, seed=20250102)[3:]   # m = 48 and 66
    print("m =", [md.m for md in mods], flush=True)
    gs = [_ctx(md, "fp32", 30, 500 + i) for i, md in enumerate(mods)]
    for _ in range(3):
        rc, codes = _batch(gs, [2000, 2000])
        assert rc == 0, codes
    for g in gs:
        t, e = g.lu_autotune(), g.lu_elimination()
        print(f"  tuned {t} built front {e['max_front']}", flush=True)
        assert t["kernel"] is not None and t["max_front"] == e["max_front"]
    assert gs[0].batch_stats()["group"] == 2
    for g in gs:
        g.close()


def test_eager_group_step_equals_graph(hiplib):
    """With stage timing on in the leading context the group step runs
    eagerly (no graph) and records its stage times; results equal the
    replayed graph's and single calls', bit for bit."""
    T = 30
    mods = instance_models(6, T=T + 2, seed=97)
    nsims = [2000] * len(mods)
    ga = [_ctx(md, "fp32", T, 600 + i) for i, md in enumerate(mods)]
    gb = [_ctx(md, "fp32", T, 600 + i) for i, md in enumerate(mods)]
    _check_steps(ga, gb, nsims, 2, "graph")
    ga[0].set_timing(True)
    _check_steps(ga, gb, nsims, 2, "eager")
    t = {k: ga[0].get_timing(k) for k in ("batch_draw", "batch_lu", "batch_weights",
                                            "batch_moments")}
    print(f"  eager stage times {t}", flush=True)
    assert all(v[1] == 2 and v[0] > 0.0 for v in t.values())
    assert ga[0].batch_stats()["group"] == 3
    for g in ga + gb:
        g.close()
