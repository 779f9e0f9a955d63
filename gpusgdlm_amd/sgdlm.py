"""Python mirror of the Rcpp module ``gpuSGDLM``, class ``SGDLM``.

Same method names, argument meaning, array layouts and error behaviour as
rSGDLM/src/SGDLM/RcppWrapper.{hpp,cpp} and rSGDLM/src/R_main.cpp:16-39:
arrays use R's dimensions in column-major (Fortran) order, dimension
mismatches are logged and the method returns early (None / empty), and
``sp`` entries that are NA (None / NaN / negative) mean "external
regressor".  ``strict=True`` turns every logged error into an exception.

The compute happens in libsgdlm_hip.so on the GPU (see include/sgdlm.h).
"""
from __future__ import annotations

import ctypes as C
import logging

import numpy as np

from . import _native as nat

log = logging.getLogger("gpusgdlm_amd")

NA_INDEX = 2147483648  # R's NA_integer_ (INT_MIN) read as unsigned int


class SGDLMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def encode_sp(sp) -> np.ndarray:
    """R integer matrix (max_p x m) with NA -> uint32 Fortran-ordered array in
    the reference layout sp[j*max_p + k]; NA (None / NaN / negative) -> 2^31."""
    a = np.array(sp, dtype=object)
    if a.ndim != 2:
        raise ValueError("sp must be a 2-d (max_p x m) array")
    out = np.full(a.shape, NA_INDEX, dtype=np.uint32, order="F")
    for idx, v in np.ndenumerate(a):
        if v is None:
            continue
        fv = float(v)
        if np.isnan(fv) or fv < 0:
            continue
        out[idx] = int(fv)
    return out


def _f(a, shape=None):
    a = np.asarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape, order="F")
    return np.asfortranarray(a)


class SGDLM:
    """new(M$SGDLM[, no_gpus]) -- rSGDLM/src/SGDLM/RcppWrapper.cpp:9-19."""

    _loglevel = 20  # LOGRESULT, rSGDLM/src/logger.cpp:11

    def __init__(self, no_gpus: int = 1, strict: bool = False):
        self._lib = nat.load()
        self.strict = strict
        h = C.c_void_p()
        rc = self._lib.sgdlm_create(C.byref(h), C.c_size_t(no_gpus))
        if rc != nat.SGDLM_OK:
            raise SGDLMError(rc, "could not create a context on a GPU (no HIP device?)")
        self._h = h

    # ------------------------------------------------------------------ utils
    def close(self):
        if getattr(self, "_h", None):
            self._lib.sgdlm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self, msg: str, code: int = nat.SGDLM_ERR_ARG):
        if self.strict:
            raise SGDLMError(code, msg)
        log.error(msg)

    def _call(self, name, *args) -> bool:
        rc = getattr(self._lib, name)(self._h, *args)
        if rc != nat.SGDLM_OK:
            msg = self._lib.sgdlm_last_error(self._h).decode()
            self._err(f"{name}: {msg}", rc)
            return False
        return True

    def _dims(self):
        m, mp = C.c_size_t(), C.c_size_t()
        self._lib.sgdlm_get_no_series(self._h, C.byref(m))
        self._lib.sgdlm_get_max_p(self._h, C.byref(mp))
        return m.value, mp.value

    def _nsim(self):
        n, nb = C.c_size_t(), C.c_size_t()
        self._lib.sgdlm_get_nsim(self._h, C.byref(n))
        self._lib.sgdlm_get_nsim_batch(self._h, C.byref(nb))
        return n.value, nb.value

    # ----------------------------------------------------- memory management
    def clearGPUMemory(self):  # RcppWrapper.cpp:29-34
        self._call("sgdlm_clear_memory")

    def clearSimulationMemory(self):  # RcppWrapper.cpp:36-41
        self._call("sgdlm_clear_sim_memory")

    # --------------------------------------------------------------- parents
    def getSimultaneousParents(self):  # RcppWrapper.cpp:43-71
        m, mp = self._dims()
        p = np.zeros(m, dtype=np.uint32)
        sp = np.zeros((mp, m), dtype=np.uint32, order="F")
        if not self._call("sgdlm_get_parental_sets", nat.uptr(p), nat.uptr(sp)):
            return None
        return {"p": p, "sp": sp}

    def setSimultaneousParents(self, p, sp):  # RcppWrapper.cpp:73-121
        spa = encode_sp(sp)
        p = np.ascontiguousarray(np.asarray(p, dtype=np.uint32))
        if spa.shape[1] != p.shape[0]:
            return self._err("setSimultaneousParents: length(p) != ncol(sp)")
        m, mp = self._dims()
        if m != p.shape[0] or mp != spa.shape[0]:
            if not self._call("sgdlm_init_memory", C.c_size_t(p.shape[0]), C.c_size_t(spa.shape[0])):
                return None
        self._call("sgdlm_set_parental_sets", nat.uptr(p), nat.uptr(spa))

    # ------------------------------------------------------------- discounts
    def getDiscountFactors(self):  # RcppWrapper.cpp:123-147
        m, mp = self._dims()
        beta = np.zeros(m)
        delta = np.zeros((mp, mp, m), order="F")
        if not self._call("sgdlm_get_discount_factors", nat.dptr(beta), nat.dptr(delta)):
            return None
        return {"beta": beta, "delta": delta}

    def setDiscountFactors(self, beta, delta):  # RcppWrapper.cpp:150-183
        beta = np.ascontiguousarray(np.asarray(beta, dtype=np.float64))
        delta = _f(delta)
        m, mp = self._dims()
        if delta.ndim != 3:
            return self._err("setDiscountFactors: Delta must be a 3-d array")
        if beta.shape[0] != m or delta.shape != (mp, mp, m):
            return self._err("setDiscountFactors: dimension mismatch")
        self._call("sgdlm_set_discount_factors", nat.dptr(beta), nat.dptr(delta))

    # ------------------------------------------------------------ parameters
    def getParameters(self):  # RcppWrapper.cpp:185-219
        m, mp = self._dims()
        mt = np.zeros((mp, m), order="F")
        Ct = np.zeros((mp, mp, m), order="F")
        nt = np.zeros(m)
        st = np.zeros(m)
        if not self._call("sgdlm_get_parameters", nat.dptr(mt), nat.dptr(Ct), nat.dptr(nt), nat.dptr(st)):
            return None
        return {"m": mt, "C": Ct, "n": nt, "s": st}

    def _set_parameters(self, m_t, C_t, n_t, s_t, is_prior):  # RcppWrapper.cpp:235-283
        m_t, C_t = _f(m_t), _f(C_t)
        n_t = np.ascontiguousarray(np.asarray(n_t, dtype=np.float64))
        s_t = np.ascontiguousarray(np.asarray(s_t, dtype=np.float64))
        m, mp = self._dims()
        if m_t.ndim != 2 or C_t.ndim != 3:
            return self._err("setParameters: m must be 2-d and C 3-d")
        if m_t.shape != (mp, m) or C_t.shape != (mp, mp, m) or n_t.shape != (m,) or s_t.shape != (m,):
            return self._err("setParameters: dimension mismatch")
        if self._call("sgdlm_set_parameters", nat.dptr(m_t), nat.dptr(C_t), nat.dptr(n_t), nat.dptr(s_t)):
            self._call("sgdlm_set_is_prior", C.c_int(1 if is_prior else 0))

    def setPriorParameters(self, m_t, C_t, n_t, s_t):
        self._set_parameters(m_t, C_t, n_t, s_t, True)

    def setPosteriorParameters(self, m_t, C_t, n_t, s_t):
        self._set_parameters(m_t, C_t, n_t, s_t, False)

    # ------------------------------------------------------------- filtering
    def computePosterior(self, y_t, F_t):  # RcppWrapper.cpp:285-314
        y_t = np.ascontiguousarray(np.asarray(y_t, dtype=np.float64))
        F_t = _f(F_t)
        m, mp = self._dims()
        if F_t.ndim != 2 or y_t.shape != (m,) or F_t.shape != (mp, m):
            return self._err("computePosterior: dimension mismatch")
        self._call("sgdlm_compute_posterior", nat.dptr(y_t), nat.dptr(F_t))

    def _ensure_sim(self, nsim, nsim_batch):
        n, nb = self._nsim()
        if nsim != n or nsim_batch != nb:  # RcppWrapper.cpp:322-325
            return self._call("sgdlm_init_sim_memory", C.c_size_t(nsim), C.c_size_t(nsim_batch))
        return True

    def computeVBPosterior(self, nsim, nsim_batch):  # RcppWrapper.cpp:316-369
        if not self._ensure_sim(int(nsim), int(nsim_batch)):
            return None
        m, mp = self._dims()
        n, _ = self._nsim()
        mt = np.zeros((mp, m), order="F")
        Ct = np.zeros((mp, mp, m), order="F")
        nt = np.zeros(m)
        st = np.zeros(m)
        w = np.zeros(n)
        sd = C.c_double(0.0)
        if not self._call("sgdlm_compute_vb_posterior", nat.dptr(mt), nat.dptr(Ct), nat.dptr(nt),
                          nat.dptr(st), nat.dptr(w), C.byref(sd)):
            return None
        return {"m": mt, "C": Ct, "n": nt, "s": st, "IS_weights": w, "sum_det_weights": sd.value}

    def computePrior(self):  # RcppWrapper.cpp:371-375
        self._call("sgdlm_compute_prior", None, C.c_int(0))

    def computeEvoPrior(self, G_tp1):  # RcppWrapper.cpp:377-409
        G = _f(G_tp1)
        m, mp = self._dims()
        if G.ndim != 3 or G.shape != (mp, mp, m):
            return self._err("computeEvoPrior: dimension mismatch")
        if self._call("sgdlm_manage_evo_memory", C.c_int(1)):
            self._call("sgdlm_compute_prior", nat.dptr(G), C.c_int(0))

    # ------------------------------------------------------------ forecasting
    def _forecast(self, nsim, nsim_batch, F_tp1, G_tp1=None, debug=False):
        F = _f(F_tp1)
        if F.ndim not in (2, 3):
            return self._err("computeForecast: F_tp1 must be 2-d or 3-d")
        m, mp = self._dims()
        nsteps = F.shape[2] if F.ndim == 3 else 1
        if F.shape[0] != mp or F.shape[1] != m:
            return self._err("computeForecast: dimension mismatch")
        if not self._ensure_sim(int(nsim), int(nsim_batch)):
            return None
        G = None
        G_steps = 1
        if G_tp1 is not None:
            G = _f(G_tp1)
            if G.ndim not in (3, 4):
                return self._err("computeEvoForecast: G_tp1 must be 3-d or 4-d")
            if G.ndim == 4:
                G_steps = G.shape[3]
                if G_steps > 1 and G_steps != nsteps:
                    return self._err("computeEvoForecast: G steps != F steps")
            if G.shape[0] != mp or G.shape[1] != mp or G.shape[2] != m:
                return self._err("computeEvoForecast: dimension mismatch")
            if not self._call("sgdlm_manage_evo_memory", C.c_int(1)):
                return None
        n, _ = self._nsim()
        y = np.zeros((m, n, nsteps), order="F")
        lam = np.zeros((m, n, nsteps), order="F") if debug else None
        th = np.zeros((mp, m, n, nsteps), order="F") if debug else None
        if not self._call("sgdlm_init_forecast_priors"):
            return None
        Gflat = G.reshape(-1, order="F") if G is not None else None
        for k in range(nsteps):  # RcppWrapper.cpp:457-472
            if k > 0:
                if G is not None:
                    off = (k if G_steps > 1 else 0) * m * mp * mp
                    if not self._call("sgdlm_evolve_forecast_samples", nat.dptr(Gflat[off:])):
                        return None
                elif not self._call("sgdlm_evolve_forecast_samples", None):
                    return None
            Fk = np.asfortranarray(F[:, :, k] if F.ndim == 3 else F)
            yk = np.zeros((m, n), order="F")
            lk = np.zeros((m, n), order="F") if debug else None
            tk = np.zeros((mp, m, n), order="F") if debug else None
            if not self._call("sgdlm_compute_forecast", nat.dptr(yk), nat.dptr(Fk), nat.dptr(lk),
                              nat.dptr(tk), C.c_int(1 if k > 0 else 0)):
                return None
            y[:, :, k] = yk
            if debug:
                lam[:, :, k] = lk
                th[:, :, :, k] = tk
        if F.ndim == 2:
            y = y[:, :, 0]
            if debug:
                lam = lam[:, :, 0]
                th = th[:, :, :, 0]
        if debug:
            return {"y": y, "lambdas": lam, "thetas": th}
        return y

    def computeForecast(self, nsim, nsim_batch, F_tp1):  # RcppWrapper.cpp:411-479
        return self._forecast(nsim, nsim_batch, F_tp1)

    def computeForecastDebug(self, nsim, nsim_batch, F_tp1):  # RcppWrapper.cpp:481-574
        return self._forecast(nsim, nsim_batch, F_tp1, debug=True)

    def computeEvoForecast(self, nsim, nsim_batch, F_tp1, G_tp1):  # RcppWrapper.cpp:576-688
        return self._forecast(nsim, nsim_batch, F_tp1, G_tp1)

    def computeEvoForecastDebug(self, nsim, nsim_batch, F_tp1, G_tp1):  # RcppWrapper.cpp:690-825
        return self._forecast(nsim, nsim_batch, F_tp1, G_tp1, debug=True)

    # ---------------------------------------------------------------- logging
    def getLogLevel(self):  # RcppWrapper.cpp:827-832
        return SGDLM._loglevel

    def setLogLevel(self, loglevel):  # RcppWrapper.cpp:834-839
        SGDLM._loglevel = int(loglevel)
        log.setLevel(logging.DEBUG if loglevel >= 50 else logging.INFO if loglevel >= 20 else logging.ERROR)

    # ------------------------------------------------ extensions (no R twin)
    def set_precision(self, prec):
        code = nat.PRECISIONS[prec] if isinstance(prec, str) else int(prec)
        self._call("sgdlm_set_precision", C.c_int(code))

    def set_rng_state(self, seed: int, epoch: int = 0):
        self._call("sgdlm_set_rng_state", C.c_uint64(seed), C.c_uint32(epoch))

    def get_rng_state(self):
        s, e = C.c_uint64(), C.c_uint32()
        self._call("sgdlm_get_rng_state", C.byref(s), C.byref(e))
        return s.value, e.value

    def vb_device(self, nsim, nsim_batch=None):
        """VB step whose results stay on the device (no host copies)."""
        if not self._ensure_sim(int(nsim), int(nsim_batch or nsim)):
            return False
        return self._call("sgdlm_compute_vb_posterior", None, None, None, None, None, None)

    def forecast_device(self, nsim, F_tp1, nsteps=1):
        """computeForecast's device work (draws, evolution for k > 0, LU
        solves) for nsteps steps ahead with the same F_tp1 at every step;
        the forecasts stay on the device (no PCIe copy)."""
        F = _f(F_tp1)
        if not self._ensure_sim(int(nsim), int(nsim)):
            return False
        if not self._call("sgdlm_init_forecast_priors"):
            return False
        for k in range(int(nsteps)):
            if k > 0 and not self._call("sgdlm_evolve_forecast_samples", None):
                return False
            if not self._call("sgdlm_compute_forecast", None, nat.dptr(F), None, None,
                              C.c_int(1 if k > 0 else 0)):
                return False
        return True

    def get_vb_result(self):
        m, mp = self._dims()
        n, _ = self._nsim()
        mt = np.zeros((mp, m), order="F")
        Ct = np.zeros((mp, mp, m), order="F")
        nt, st = np.zeros(m), np.zeros(m)
        w, lw = np.zeros(n), np.zeros(n)
        ess, lsd = C.c_double(), C.c_double()
        if not self._call("sgdlm_get_vb_result", nat.dptr(mt), nat.dptr(Ct), nat.dptr(nt), nat.dptr(st),
                          nat.dptr(w), nat.dptr(lw), C.byref(ess), C.byref(lsd)):
            return None
        return {"m": mt, "C": Ct, "n": nt, "s": st, "IS_weights": w, "log_weights": lw,
                "ess": ess.value, "log_sum_det": lsd.value}

    def set_timing(self, on: bool):
        self._call("sgdlm_set_timing", C.c_int(1 if on else 0))

    def get_timing(self, stage: str):
        ms, n = C.c_double(), C.c_long()
        self._call("sgdlm_get_timing", stage.encode(), C.byref(ms), C.byref(n))
        return ms.value, n.value

    def lu_fallback_count(self):
        n = C.c_long()
        self._call("sgdlm_get_lu_fallback_count", C.byref(n))
        return n.value

    LU_KINDS = {0: "auto", 1: "blocked", 2: "register", 3: "pivoting", 4: "wave", 5: "wave_prefix",
                6: "register_pivoting"}

    def lu_config(self):
        """Launch shape of the VB LU kernel for the current structure and
        precision (with the tuned or pinned choices), and which kernel it is
        ("register": the dense block resident in registers, nb = 16-wide
        tiles; "blocked": LDS panels of nb columns; "pivoting": row
        interchanges for every draw)."""
        g, l, nb, th = C.c_int(), C.c_size_t(), C.c_int(), C.c_int()
        self._call("sgdlm_get_lu_config", C.byref(g), C.byref(l), C.byref(nb), C.byref(th))
        out = {"grid": g.value, "lds": l.value, "nb": nb.value, "threads": th.value}
        out.update(self.lu_kernel())
        return out

    def lu_kernel(self):
        """Kernel kind of the VB LU, the dimension of the dense block it
        factors and the dimension the partial-pivoting kernel factors for a
        draw that needs row interchanges."""
        k, dd, pd = C.c_int(), C.c_int(), C.c_int()
        self._call("sgdlm_get_lu_kernel", C.byref(k), C.byref(dd), C.byref(pd))
        return {"kernel": self.LU_KINDS[k.value], "dense_dim": dd.value, "pivot_dim": pd.value}

    def set_lu_kernel(self, kind: str = "auto"):
        """Pin the VB LU kernel kind ("auto", "blocked", "register", "pivoting",
        "wave", "wave_prefix", "register_pivoting")."""
        inv = {v: k for k, v in self.LU_KINDS.items()}
        self._call("sgdlm_set_lu_kernel", C.c_int(inv[kind]))

    def lu_solve_kernel(self):
        """Kernel kind of the forecast solves ("register", "blocked" or
        "pivoting")."""
        k = C.c_int()
        self._call("sgdlm_get_lu_solve_kernel", C.byref(k))
        return self.LU_KINDS[k.value]

    def set_lu_solve_kernel(self, kind: str = "auto"):
        """Pin the forecast-solve kernel kind ("auto", "blocked", "register",
        "pivoting")."""
        inv = {v: k for k, v in self.LU_KINDS.items()}
        self._call("sgdlm_set_lu_solve_kernel", C.c_int(inv[kind]))

    def set_graph_mode(self, on: bool = True):
        """Capture the device part of a VB call as a hipGraph once and replay
        it on later calls (bit-identical results)."""
        self._call("sgdlm_set_graph_mode", C.c_int(1 if on else 0))

    def graph_stats(self):
        """VB graphs captured and VB calls replayed from one."""
        cap, rep = C.c_long(), C.c_long()
        self._call("sgdlm_get_graph_stats", C.byref(cap), C.byref(rep))
        return {"captures": cap.value, "replays": rep.value}

    def set_batch_mode(self, mode: str = "group"):
        """How a batch_vb led by this model (first in the list) runs: "group"
        (one launch per stage over all models) or "streams" (each model's
        stages on its own stream)."""
        modes = {"group": nat.BATCH_GROUP, "streams": nat.BATCH_STREAMS}
        self._call("sgdlm_set_batch_mode", C.c_int(modes[mode]))

    def batch_stats(self):
        """Batches this model led: as one launch per stage, as one stream per
        model."""
        g, s = C.c_long(), C.c_long()
        self._call("sgdlm_get_batch_stats", C.byref(g), C.byref(s))
        return {"group": g.value, "streams": s.value}

    def set_lu_autotune(self, on: bool = True):
        """Time the feasible VB LU configurations on the first VB call of a
        structure and keep the fastest (default on)."""
        self._call("sgdlm_set_lu_autotune", C.c_int(1 if on else 0))

    def lu_autotune(self):
        """The tuned configuration of the current structure (kernel None when
        none is), the tuning time in ms (0 when taken from the process-wide
        table), the configurations timed and whether it came from the table."""
        k, th, nb, mf, n, fc = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        ms = C.c_double()
        self._call("sgdlm_get_lu_autotune", C.byref(k), C.byref(th), C.byref(nb), C.byref(mf),
                   C.byref(ms), C.byref(n), C.byref(fc))
        kind = self.LU_KINDS[k.value]
        return {"kernel": None if kind == "auto" else kind, "threads": th.value, "nb": nb.value,
                "max_front": mf.value, "tune_ms": ms.value, "candidates": n.value,
                "from_cache": bool(fc.value)}

    def lu_solve_config(self):
        """Configuration of the forecast-solve LU kernels at the current
        precision (nb = 0: partial pivoting for every draw)."""
        g, l, nb, th, sw = C.c_int(), C.c_size_t(), C.c_int(), C.c_int(), C.c_int()
        self._call("sgdlm_get_lu_solve_config", C.byref(g), C.byref(l), C.byref(nb), C.byref(th),
                   C.byref(sw))
        return {"grid": g.value, "lds": l.value, "nb": nb.value, "threads": th.value,
                "slots_in_workspace": bool(sw.value)}

    def set_lu_tuning(self, threads: int = 0, nb: int = 0):
        """Pin the fast LU kernel's workgroup size (256 / 512) and panel width
        (16 / 32 / 64); 0 = automatic."""
        self._call("sgdlm_set_lu_tuning", C.c_int(int(threads)), C.c_int(int(nb)))

    def set_lu_max_front(self, max_front: int):
        """Largest pivot degree of the sparse LU prefix (0 = dense LU only)."""
        self._call("sgdlm_set_lu_max_front", C.c_int(int(max_front)))

    def lu_elimination(self):
        mf, n1, n2, nw = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        fma = C.c_long()
        self._call("sgdlm_get_lu_elimination", C.byref(mf), C.byref(n1), C.byref(n2), C.byref(nw),
                   C.byref(fma))
        return {"max_front": mf.value, "n_prefix": n1.value, "n_dense": n2.value, "waves": nw.value,
                "prefix_fma": fma.value}


def batch_vb(models, nsim, nsim_batch=None):
    """computeVBPosterior (results left on the device) for a list of SGDLM
    objects in one batched call: every stage is one launch over all models
    (or, with set_batch_mode("streams") on the first model, each model's
    stages on its own stream).  Returns the list of per-model status codes;
    results via get_vb_result()."""
    lib = nat.load()
    nb = int(nsim_batch or nsim)
    for g in models:
        if not g._ensure_sim(int(nsim), nb):
            raise SGDLMError(nat.SGDLM_ERR_STATE, "could not initialise simulation memory")
    arr = (C.c_void_p * len(models))(*[g._h.value for g in models])
    st = (C.c_int * len(models))()
    rc = lib.sgdlm_batch_compute_vb_posterior(arr, C.c_size_t(len(models)), st)
    codes = list(st)
    if rc != nat.SGDLM_OK and any(g.strict for g in models):
        bad = [i for i, c in enumerate(codes) if c]
        raise SGDLMError(rc, f"batched VB failed for models {bad[:8]}")
    return codes

