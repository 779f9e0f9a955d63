"""Graph mode (GPU): the device part of a VB call captured once as a
hipGraph and replayed gives bit-identical results to the direct path.

Each replay is checked against a direct call of a second context on the
same model, RNG seed and epoch; the direct context's work and the host
synchronisations of both calls run between the replays.  Epochs repeat
and go backwards (the epoch reaches the replayed draw kernels through a
device word), the draw count changes (re-capture), and a batched call over
graph-mode contexts matches direct ones."""
import numpy as np
import pytest

from gpusgdlm_amd import SGDLM, batch_vb
from gpusgdlm_amd.workloads import configure, filter_steps, instance_models, random_model, readme_model

pytestmark = pytest.mark.gpu

KEYS = ("m", "C", "n", "s", "IS_weights", "log_weights")


def _ctx(mod, prec, T, graph):
    g = SGDLM(strict=True)
    g.set_precision(prec)
    g.set_graph_mode(graph)
    configure(g, mod)
    filter_steps(g, mod, 1, T)
    g.computePosterior(mod.y[T], mod.F(T))
    return g


def _same(a, b, what):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), (what, k)
    assert a["ess"] == b["ess"] and a["log_sum_det"] == b["log_sum_det"], what


@pytest.mark.parametrize("kind,prec,N", [("readme", "fp64", 3000), ("m100", "fp32", 4000),
                                         ("m400", "fp32", 6000), ("m400", "bf16x3", 6000),
                                         ("m400", "fp64", 3000)])
def test_graph_replays_match_direct_calls(hiplib, kind, prec, N):
    if kind == "readme":
        mod, T = readme_model(T=40), 30
    elif kind == "m100":
        mod, T = random_model(100, 8, T=60, seed=21), 50
    else:
        mod, T = random_model(400, 20, T=60, seed=22), 50
    direct, graph = _ctx(mod, prec, T, False), _ctx(mod, prec, T, True)
    epochs = [3, 4, 9, 4, 2, 15]
    for i, ep in enumerate(epochs):
        print(f"graph {kind} {prec} N={N} call {i} epoch {ep}", flush=True)
        direct.set_rng_state(77, ep)
        graph.set_rng_state(77, ep)
        direct.vb_device(N, N)
        a = direct.get_vb_result()
        graph.vb_device(N, N)
        b = graph.get_vb_result()
        _same(a, b, f"call {i} epoch {ep}")
        assert graph.get_rng_state() == direct.get_rng_state()
    st = graph.graph_stats()
    print(f"  {st}", flush=True)
    assert st["captures"] == 1 and st["replays"] >= len(epochs) - 1
    # another draw count: re-captured, still identical
    N2 = N // 2
    direct.set_rng_state(77, 5)
    graph.set_rng_state(77, 5)
    direct.vb_device(N2, N2)
    graph.vb_device(N2, N2)
    _same(direct.get_vb_result(), graph.get_vb_result(), "N/2")
    assert graph.graph_stats()["captures"] == 2
    # the full API (host copies of the outputs after the replay)
    direct.set_rng_state(77, 6)
    graph.set_rng_state(77, 6)
    va = direct.computeVBPosterior(N2, N2)
    vb = graph.computeVBPosterior(N2, N2)
    for k in ("m", "C", "n", "s", "IS_weights"):
        assert np.array_equal(va[k], vb[k]), k
    assert va["sum_det_weights"] == vb["sum_det_weights"]
    direct.close()
    graph.close()


def test_graph_batched_instances_match_direct(hiplib):
    mods = instance_models(6, T=62, seed=31)
    N = 3000
    direct = [_ctx(m, "fp32", 60, False) for m in mods]
    graph = [_ctx(m, "fp32", 60, True) for m in mods]
    for step, ep in enumerate([7, 8, 8, 3]):
        print(f"graph batched 6 instances step {step} epoch {ep}", flush=True)
        for g in direct + graph:
            g.set_rng_state(11, ep)
        assert batch_vb(direct, N) == [0] * len(mods)
        assert batch_vb(graph, N) == [0] * len(mods)
        for i, (a, b) in enumerate(zip(direct, graph)):
            _same(a.get_vb_result(), b.get_vb_result(), f"instance {i} step {step}")
    # the first call tunes each structure's LU (one stream per model); the
    # others run as one launch per stage, replayed from one captured graph
    # (batches always run as a graph)
    st = graph[0].batch_stats()
    print(f"  leader batch stats {st}, graph {graph[0].graph_stats()}", flush=True)
    assert st["group"] == 3 and direct[0].batch_stats()["group"] == 3
    assert graph[0].graph_stats()["replays"] >= 3 and direct[0].graph_stats()["replays"] >= 3
    for g in direct + graph:
        g.close()
