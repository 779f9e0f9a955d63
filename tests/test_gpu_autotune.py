"""VB LU kernel kinds and the autotuner (GPU): every kernel kind a caller
can pin gives the oracle's log-determinants on the same draws, and the
autotuner's choice (timed on the first VB call of a structure, then taken
from the process-wide table by any context with that structure) gives the
same results as the configuration it names.

Bars: FP64 1e-10 absolute on the log-determinants, FP32 / BF16x3 5e-3
absolute (the LU bars of test_gpu_parity.py)."""
import numpy as np
import pytest

from gpusgdlm_amd import SGDLM
from gpusgdlm_amd.workloads import configure, filter_steps, random_model
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SEED = 1234
TOL = {"fp64": 1e-10, "fp32": 5e-3, "bf16x3": 5e-3}


def _model():
    return random_model(400, 20, T=45, seed=9)


def _oracle_logdets(mod, N, t_post=40):
    o = orc.OracleSGDLM(seed=SEED, epoch=0)
    configure(o, mod)
    filter_steps(o, mod, 1, t_post)
    o.computePosterior(mod.y[t_post], mod.F(t_post))
    prm = o.getParameters()
    lam, th = orc.sample(mod.m, mod.max_p, mod.p, prm["m"], prm["C"], prm["n"], prm["s"], N, SEED, 1)
    return orc.logdet(mod.m, mod.max_p, mod.p, mod.sp_u32, th)


def _gpu(mod, prec, N, setup=None, t_post=40):
    g = SGDLM(strict=True)
    g.set_rng_state(SEED, 0)
    g.set_precision(prec)
    configure(g, mod)
    if setup:
        setup(g)
    filter_steps(g, mod, 1, t_post)
    g.computePosterior(mod.y[t_post], mod.F(t_post))
    g.computeVBPosterior(N, N)
    out = {"lw": g.get_vb_result()["log_weights"], "cfg": g.lu_config(), "tune": g.lu_autotune(),
           "el": g.lu_elimination()}
    g.close()
    return out


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "fp64"])
def test_pinned_kernel_kinds_vs_oracle(hiplib, prec):
    mod = _model()
    N = 192
    ld = _oracle_logdets(mod, N)
    kinds = ["blocked", "pivoting"] + ([] if prec == "fp64" else ["register"])
    # (one wavefront per draw: m = 400 is beyond it; small models below)
    for kind in kinds:
        print(f"pinned LU kernel {kind} {prec} m=400 N={N}", flush=True)
        r = _gpu(mod, prec, N, lambda g: g.set_lu_kernel(kind))
        assert r["cfg"]["kernel"] == kind
        assert r["tune"]["kernel"] is None  # pinned: no tuning
        err = np.max(np.abs(r["lw"] - ld))
        print(f"  {r['cfg']}: max |err| {err:.2e}", flush=True)
        assert err <= TOL[prec]


@pytest.mark.parametrize("prec,kind", [("fp64", "register"), ("fp32", "wave")])
def test_infeasible_kernel_kind_refused(hiplib, prec, kind):
    """FP64 has no register-resident kernel; m = 400 is beyond one
    wavefront per draw."""
    mod = _model()
    g = SGDLM(strict=True)
    g.set_precision(prec)
    configure(g, mod)
    g.set_lu_kernel(kind)
    with pytest.raises(Exception):
        g.lu_config()
    g.close()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_autotune_then_table(hiplib, prec):
    """First context: tuned on its first VB call (timed candidates); second
    context, same structure: the table's entry, bit-identical results; both
    match the oracle, as does the configuration pinned by hand."""
    mod = _model()
    N = 4096
    ld = _oracle_logdets(mod, N)
    print(f"autotune {prec} m=400 N={N}", flush=True)
    a = _gpu(mod, prec, N)
    t = a["tune"]
    print(f"  tuned {t}, plan {a['cfg']}, elimination {a['el']}", flush=True)
    assert t["kernel"] in ("register", "blocked") and t["candidates"] >= 3
    assert a["cfg"]["kernel"] == t["kernel"] and a["el"]["max_front"] == t["max_front"]
    b = _gpu(mod, prec, N)
    assert b["tune"]["from_cache"] and b["tune"]["tune_ms"] == 0.0
    assert {k: b["tune"][k] for k in ("kernel", "threads", "nb", "max_front")} == \
        {k: t[k] for k in ("kernel", "threads", "nb", "max_front")}
    assert np.array_equal(a["lw"], b["lw"])

    def pin(g):
        g.set_lu_max_front(t["max_front"])
        if t["kernel"] == "register":
            g.set_lu_kernel("register")
        else:
            g.set_lu_tuning(t["threads"], t["nb"])
    c = _gpu(mod, prec, N, pin)
    assert np.array_equal(a["lw"], c["lw"])
    assert np.max(np.abs(a["lw"] - ld)) <= TOL[prec]


def _dense(mod, th):
    A = np.eye(mod.m)
    spf = mod.sp_u32.ravel(order="F")
    thf = th.ravel(order="F")
    for q, v in enumerate(spf):
        if v < mod.m * mod.m:
            A[v % mod.m, v // mod.m] = -thf[q]
    return A


@pytest.mark.parametrize("m,prec,T", [(20, "fp64", 40), (20, "fp32", 3), (48, "fp64", 3),
                                      (48, "fp32", 40), (100, "fp32", 40), (100, "bf16x3", 3)])
def test_wave_kernel_small_models_vs_oracle(hiplib, m, prec, T):
    """One wavefront per draw, the whole matrix in registers, partial
    pivoting (the oracle's algorithm): log-dets within 1e-10 (FP64) or
    5e-3 (FP32) of the oracle, or within 1e-15 kappa / 6e-8 kappa for an
    ill-conditioned draw; early in a series (T = 3) as well, where the
    no-interchange kernels send most draws to the fallback."""
    mod = random_model(m, min(20, m - 1), T=T + 2, seed=40 + m)
    N = 512
    print(f"wave kernel m={m} {prec} T={T} N={N}", flush=True)
    r = _gpu(mod, prec, N, lambda g: g.set_lu_kernel("wave"), t_post=T)
    assert r["cfg"]["kernel"] == "wave" and r["cfg"]["dense_dim"] in (32, 64, 128)
    ld = _oracle_logdets(mod, N, t_post=T)
    o = orc.OracleSGDLM(seed=SEED, epoch=0)
    configure(o, mod)
    filter_steps(o, mod, 1, T)
    o.computePosterior(mod.y[T], mod.F(T))
    prm = o.getParameters()
    lam, th = orc.sample(mod.m, mod.max_p, mod.p, prm["m"], prm["C"], prm["n"], prm["s"], N, SEED, 1)
    err = np.abs(r["lw"] - ld)
    easy = 1e-10 if prec == "fp64" else 5e-3
    hard = np.nonzero(~(err <= easy))[0]
    print(f"  |err| median {np.median(err):.2e} max {err.max():.2e}; {len(hard)} above {easy:g}",
          flush=True)
    assert len(hard) <= N // 20
    for i in hard:
        kappa = np.linalg.cond(_dense(mod, th[:, :, i]))
        print(f"  draw {i}: kappa {kappa:.2e} err {err[i]:.2e}", flush=True)
        assert err[i] <= (1e-15 if prec == "fp64" else 6e-8) * kappa


@pytest.mark.parametrize("m,k,prec,T", [(100, 8, "fp32", 40), (120, 12, "fp32", 3),
                                        (160, 16, "fp32", 40), (140, 14, "bf16x3", 40),
                                        (90, 10, "fp64", 40), (100, 14, "fp64", 3)])
def test_wave_prefix_kernel_vs_oracle(hiplib, m, k, prec, T):
    """One wavefront per draw after the sparse prefix: the prefix on the
    wavefront's LDS slots (multipliers up to 64), the Schur complement of
    48..112 rows gathered into registers, partial pivoting on it.  Bars as
    for the whole-matrix wavefront kernel; draws whose prefix fails go to
    the partial-pivoting kernel (counted as fallback draws)."""
    mod = random_model(m, k, T=T + 2, seed=70 + m + k)
    N = 512
    print(f"wave-prefix kernel m={m} k={k} {prec} T={T} N={N}", flush=True)
    r = _gpu(mod, prec, N, lambda g: g.set_lu_kernel("wave_prefix"), t_post=T)
    print(f"  {r['cfg']} elimination {r['el']}", flush=True)
    assert r["cfg"]["kernel"] == "wave_prefix" and r["cfg"]["dense_dim"] in (32, 64, 128)
    assert 0 < r["el"]["n_dense"] <= r["cfg"]["dense_dim"]
    ld = _oracle_logdets(mod, N, t_post=T)
    o = orc.OracleSGDLM(seed=SEED, epoch=0)
    configure(o, mod)
    filter_steps(o, mod, 1, T)
    o.computePosterior(mod.y[T], mod.F(T))
    prm = o.getParameters()
    lam, th = orc.sample(mod.m, mod.max_p, mod.p, prm["m"], prm["C"], prm["n"], prm["s"], N, SEED, 1)
    err = np.abs(r["lw"] - ld)
    easy = 1e-10 if prec == "fp64" else 5e-3
    hard = np.nonzero(~(err <= easy))[0]
    print(f"  |err| median {np.median(err):.2e} max {err.max():.2e}; {len(hard)} above {easy:g}",
          flush=True)
    assert len(hard) <= N // 20
    for i in hard:
        kappa = np.linalg.cond(_dense(mod, th[:, :, i]))
        print(f"  draw {i}: kappa {kappa:.2e} err {err[i]:.2e}", flush=True)
        assert err[i] <= (1e-15 if prec == "fp64" else 6e-8) * kappa


@pytest.mark.parametrize("m,k,T", [(400, 20, 30), (400, 20, 40), (200, 16, 3), (150, 14, 40)])
def test_register_pivoting_kernel_vs_oracle(hiplib, m, k, T):
    """The register-resident LU with partial pivoting (each 16-wide column
    panel factored by one wavefront, its row exchanges applied to the
    register tiles): FP32 log-dets within 5e-3 of the oracle (or 6e-8 kappa
    for an ill-conditioned draw), early in a series too, where the
    no-interchange kernels hand almost every draw to the fallback; no draw
    is flagged (no multiplier bound, the prefix within 64)."""
    mod = random_model(m, k, T=T + 2, seed=20250101 if m == 400 else 77 + m)
    N = 256
    print(f"register_pivoting m={m} k={k} T={T} N={N}", flush=True)

    def pin(g):
        g.set_lu_autotune(False)
        g.set_lu_kernel("register_pivoting")
    r = _gpu(mod, "fp32", N, pin, t_post=T)
    print(f"  {r['cfg']}", flush=True)
    assert r["cfg"]["kernel"] == "register_pivoting"
    ld = _oracle_logdets(mod, N, t_post=T)
    o = orc.OracleSGDLM(seed=SEED, epoch=0)
    configure(o, mod)
    filter_steps(o, mod, 1, T)
    o.computePosterior(mod.y[T], mod.F(T))
    prm = o.getParameters()
    lam, th = orc.sample(mod.m, mod.max_p, mod.p, prm["m"], prm["C"], prm["n"], prm["s"], N, SEED, 1)
    err = np.abs(r["lw"] - ld)
    hard = np.nonzero(~(err <= 5e-3))[0]
    print(f"  |err| median {np.median(err):.2e} max {err.max():.2e}; {len(hard)} above 5e-3",
          flush=True)
    assert np.all(np.isfinite(r["lw"]))
    assert len(hard) <= N // 20
    for i in hard:
        kappa = np.linalg.cond(_dense(mod, th[:, :, i]))
        print(f"  draw {i}: kappa {kappa:.2e} err {err[i]:.2e}", flush=True)
        assert err[i] <= 6e-8 * kappa


def test_adaptive_lu_early_then_late(hiplib):
    """Kind not pinned, FP32, m = 400 early in a series (every draw needs
    row interchanges): after the call that tunes the structure, calls pivot
    every draw (the fallback count is the draw count) except every 8th,
    which tries the no-interchange kernel again; log-dets stay within the
    oracle's bars.  The tuned kind is not one that pivots every draw, so
    late in the series, once a retry finds few flagged draws (within 8
    calls), the no-interchange kernel runs again."""
    mod = random_model(400, 20, T=252, seed=20250101)
    N = 2048
    g = SGDLM(strict=True)
    g.set_rng_state(SEED, 0)
    g.set_precision("fp32")
    configure(g, mod)
    filter_steps(g, mod, 1, 30)
    g.computePosterior(mod.y[30], mod.F(30))
    counts = []
    for call in range(9):
        g.set_rng_state(SEED, 0)
        g.computeVBPosterior(N, N)
        counts.append(g.lu_fallback_count())
        if call == 1:
            lw_early = g.get_vb_result()["log_weights"].copy()
    t = g.lu_autotune()
    print(f"  early: tuned {t}, fallback per call {counts}", flush=True)
    assert t["kernel"] not in ("pivoting", "register_pivoting")
    assert counts[0] > 0.9 * N
    assert all(c == N for c in counts[1:8])
    assert counts[8] > 0.9 * N and counts[8] < N or counts[8] == N
    ld = _oracle_logdets(mod, N, t_post=30)
    err = np.abs(lw_early - ld)
    print(f"  early |err| median {np.median(err):.2e} max {err.max():.2e}", flush=True)
    assert np.median(err) <= 5e-5 and np.sum(err > 5e-3) <= N // 20
    g.computePrior()
    filter_steps(g, mod, 31, 250)
    g.computePosterior(mod.y[250], mod.F(250))
    late = []
    for call in range(9):
        g.computeVBPosterior(N, N)
        late.append(g.lu_fallback_count())
    print(f"  late: fallback per call {late}", flush=True)
    # within 8 calls the no-interchange kernel is tried again, finds few
    # flagged draws, and the calls after it keep it
    assert min(late) < N // 10 and late[-1] < N // 10
    g.close()


@pytest.mark.parametrize("m,k,prec,T", [(150, 14, "fp32", 40), (250, 20, "fp32", 40),
                                        (200, 16, "bf16x3", 40), (250, 20, "fp32", 5)])
def test_register_four_wave_shape_vs_oracle(hiplib, m, k, prec, T):
    """Dense blocks of up to 12 tiles run the register kernel with four
    wavefronts and two workgroups per CU (prefix slots, tables and staging
    band aliasing the panel region): log-dets within the LU bars of the
    oracle; early in a series (T = 5) the flagged draws come back from the
    partial-pivoting kernel."""
    mod = random_model(m, k, T=T + 2, seed=5 if m == 250 else 77 + m)
    N = 512

    def pin(g):
        g.set_lu_autotune(False)
        g.set_lu_kernel("register")
    r = _gpu(mod, prec, N, pin, t_post=T)
    print(f"register m={m} {prec} T={T}: {r['cfg']}", flush=True)
    assert r["cfg"]["kernel"] == "register" and r["cfg"]["threads"] == 256
    assert r["cfg"]["dense_dim"] <= 192
    ld = _oracle_logdets(mod, N, t_post=T)
    err = np.abs(r["lw"] - ld)
    print(f"  |err| median {np.median(err):.2e} max {err.max():.2e}", flush=True)
    assert np.median(err) <= 5e-5 and np.sum(~(err <= TOL[prec])) <= N // 50
