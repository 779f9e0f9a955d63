"""Python binding of the FP64 CPU oracle (oracle/_build/libsgdlm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline of bench.py, as the checker.  ``OracleSGDLM`` replays the same
public API as gpusgdlm_amd.SGDLM (state machine of
rSGDLM/src/SGDLM/HostWrapperImpl.cu and the multi-step loop of
RcppWrapper.cpp:457-472) and consumes Philox epochs in the same order as the
GPU context: +1 per VB call, +1 for the draws of a non-reusing forecast,
+1 for the forecast noise, +1 per evolution step.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libsgdlm_oracle.so")

_dp = C.POINTER(C.c_double)
_up = C.POINTER(C.c_uint)
_ip = C.POINTER(C.c_int)


class VbDiag(C.Structure):
    _fields_ = [("sum_det_weights", C.c_double), ("ess", C.c_double), ("max_logdet", C.c_double),
                ("mean_lambda", _dp), ("mean_log_lambda", _dp), ("mean_Q", _dp),
                ("newton_iter", _ip)]


_SIG = {
    "sgo_compute_posterior": (None, [C.c_int, C.c_int, _up, _dp, _dp, _dp, _dp, _dp, _dp]),
    "sgo_compute_prior": (None, [C.c_int, C.c_int, _up, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "sgo_sample": (None, [C.c_int, C.c_int, _up, _dp, _dp, _dp, _dp, C.c_long, C.c_uint64,
                          C.c_uint32, _dp, _dp]),
    "sgo_logdet": (None, [C.c_int, C.c_int, _up, _up, C.c_long, _dp, _dp, C.c_int]),
    "sgo_lu_logabsdet": (C.c_double, [C.c_int, _dp]),
    "sgo_vb": (None, [C.c_int, C.c_int, _up, _up, _dp, _dp, _dp, _dp, C.c_long, C.c_uint64,
                      C.c_uint32, _dp, _dp, _dp, _dp, _dp, C.POINTER(VbDiag), C.c_int]),
    "sgo_forecast": (None, [C.c_int, C.c_int, _up, _up, _dp, _dp, _dp, _dp, C.c_long, C.c_uint64,
                            C.c_uint32, C.c_uint32, _dp, _dp, _dp, C.c_int, _dp, C.c_int]),
    "sgo_evolve": (None, [C.c_int, C.c_int, _up, _dp, _dp, _dp, _dp, _dp, C.c_long, C.c_uint64,
                          C.c_uint32, _dp, _dp]),
    "sgo_mean_n": (C.c_double, [C.c_uint, C.c_double, C.c_double, C.c_double, C.c_double, _ip]),
    "sgo_psi": (C.c_double, [C.c_double]),
    "sgo_psi1": (C.c_double, [C.c_double]),
    "sgo_philox": (None, [C.c_uint32] * 6 + [C.POINTER(C.c_uint32)]),
    "sgo_log": (C.c_double, [C.c_double]),
    "sgo_exp": (C.c_double, [C.c_double]),
    "sgo_sincospi": (None, [C.c_double, _dp, _dp]),
    "sgo_gamma": (C.c_double, [C.c_double, C.c_double, C.c_uint64, C.c_uint32, C.c_uint32,
                               C.c_uint32, C.c_uint32]),
    "sgo_normal_k": (C.c_double, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.c_uint32]),
    "sgo_beta_eta": (C.c_double, [C.c_double, C.c_double, C.c_uint64, C.c_uint32, C.c_uint32,
                                  C.c_uint32]),
    "sgo_cholesky_upper": (None, [C.c_int, C.c_int, _dp, _dp]),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = C.CDLL(LIB)
        for n, (r, a) in _SIG.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def u(a):
    return None if a is None else a.ctypes.data_as(_up)


def F(a, dtype=np.float64):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


# ---------------------------------------------------------------- functions
def philox(ctr, key):
    out = (C.c_uint32 * 4)()
    lib().sgo_philox(*[C.c_uint32(v) for v in list(ctr) + list(key)], out)
    return list(out)


def sample(m, max_p, p, m_t, C_t, n_t, s_t, N, seed, epoch, nthreads=0):
    lam = np.zeros((m, N), order="F")
    th = np.zeros((max_p, m, N), order="F")
    lib().sgo_sample(m, max_p, u(p), d(m_t), d(C_t), d(n_t), d(s_t), N, seed, epoch, d(lam), d(th))
    return lam, th


def logdet(m, max_p, p, sp, thetas, nthreads=0):
    N = thetas.shape[-1]
    out = np.zeros(N)
    lib().sgo_logdet(m, max_p, u(p), u(sp), N, d(F(thetas)), d(out), nthreads)
    return out


def lu_logabsdet(A):
    A = np.ascontiguousarray(np.asarray(A, dtype=np.float64))
    return lib().sgo_lu_logabsdet(A.shape[0], d(A))


# ------------------------------------------------------------- state machine
class OracleSGDLM:
    """FP64 host replay of the SGDLM API (same method names as
    gpusgdlm_amd.SGDLM)."""

    def __init__(self, seed: int = 1234, epoch: int = 0, nthreads: int = 0):
        self.seed, self.epoch, self.nthreads = seed, epoch, nthreads
        self.m = self.max_p = 0
        self.is_prior = False
        self.G = None

    def setSimultaneousParents(self, p, sp):
        from gpusgdlm_amd.sgdlm import encode_sp  # pure host helper (no GPU)
        spa = encode_sp(sp)
        p = np.ascontiguousarray(np.asarray(p, dtype=np.uint32))
        if (self.m, self.max_p) != (p.shape[0], spa.shape[0]):
            self.m, self.max_p = p.shape[0], spa.shape[0]
            m, mp = self.m, self.max_p
            self.mt = np.zeros((mp, m), order="F")
            self.Ct = np.zeros((mp, mp, m), order="F")
            self.nt = np.zeros(m)
            self.st = np.zeros(m)
            self.Wt = np.zeros((mp, mp, m), order="F")
            self.beta = np.zeros(m)
            self.delta = np.zeros((mp, mp, m), order="F")
        self.p = p
        self.sp = spa

    def setDiscountFactors(self, beta, delta):
        self.beta = np.ascontiguousarray(np.asarray(beta, dtype=np.float64))
        self.delta = F(delta)

    def _set(self, m_t, C_t, n_t, s_t, prior):
        self.mt, self.Ct = F(m_t).copy(order="F"), F(C_t).copy(order="F")
        self.nt = np.array(n_t, dtype=np.float64)
        self.st = np.array(s_t, dtype=np.float64)
        self.is_prior = prior

    def setPriorParameters(self, *a):
        self._set(*a, True)

    def setPosteriorParameters(self, *a):
        self._set(*a, False)

    def getParameters(self):
        return {"m": self.mt.copy(order="F"), "C": self.Ct.copy(order="F"),
                "n": self.nt.copy(), "s": self.st.copy()}

    def computePosterior(self, y_t, F_t):
        assert self.is_prior
        y = np.ascontiguousarray(np.asarray(y_t, dtype=np.float64))
        lib().sgo_compute_posterior(self.m, self.max_p, u(self.p), d(self.mt), d(self.Ct), d(self.nt),
                                    d(self.st), d(y), d(F(F_t)))
        self.is_prior = False

    def computePrior(self, G=None):
        assert not self.is_prior
        lib().sgo_compute_prior(self.m, self.max_p, u(self.p), d(self.mt), d(self.Ct), d(self.nt),
                                d(self.Wt), d(self.beta), d(self.delta), d(None if G is None else F(G)))
        self.is_prior = True

    def computeEvoPrior(self, G):
        self.computePrior(G)

    def computeVBPosterior(self, nsim, nsim_batch=None, diag=False):
        assert not self.is_prior
        m, mp, N = self.m, self.max_p, int(nsim)
        self.epoch += 1
        mm = np.zeros((mp, m), order="F")
        CC = np.zeros((mp, mp, m), order="F")
        nn, ss, w = np.zeros(m), np.zeros(m), np.zeros(N)
        lb, llb, Q = np.zeros(m), np.zeros(m), np.zeros(m)
        it = np.zeros(m, dtype=np.int32)
        dg = VbDiag(0, 0, 0, d(lb), d(llb), d(Q), it.ctypes.data_as(_ip))
        lib().sgo_vb(m, mp, u(self.p), u(self.sp), d(self.mt), d(self.Ct), d(self.nt), d(self.st), N,
                     self.seed, self.epoch, d(mm), d(CC), d(nn), d(ss), d(w), C.byref(dg), self.nthreads)
        out = {"m": mm, "C": CC, "n": nn, "s": ss, "IS_weights": w, "sum_det_weights": dg.sum_det_weights}
        if diag:
            out.update(ess=dg.ess, max_logdet=dg.max_logdet, mean_lambda=lb, mean_log_lambda=llb,
                       mean_Q=Q, newton_iter=it)
        return out

    def _forecast(self, nsim, F_tp1, G_tp1=None, debug=False):
        assert self.is_prior
        m, mp, N = self.m, self.max_p, int(nsim)
        Fa = F(F_tp1)
        nsteps = Fa.shape[2] if Fa.ndim == 3 else 1
        fm, fC, fn, fs, fW = (self.mt.copy(order="F"), self.Ct.copy(order="F"), self.nt.copy(),
                              self.st.copy(), self.Wt.copy(order="F"))
        lam = np.zeros((m, N), order="F")
        th = np.zeros((mp, m, N), order="F")
        ys, ls, ts = [], [], []
        Ga = None if G_tp1 is None else F(G_tp1)
        for k in range(nsteps):
            if k > 0:
                self.epoch += 1
                Gk = None
                if Ga is not None:
                    Gk = np.asfortranarray(Ga[..., k] if (Ga.ndim == 4 and Ga.shape[3] > 1) else
                                           (Ga[..., 0] if Ga.ndim == 4 else Ga))
                lib().sgo_evolve(m, mp, u(self.p), d(fW), d(fn), d(fs), d(self.beta), d(Gk), N,
                                 self.seed, self.epoch, d(lam), d(th))
            e_draw = 0
            if k == 0:
                self.epoch += 1
                e_draw = self.epoch
            self.epoch += 1
            e_nu = self.epoch
            x = np.asfortranarray(Fa[:, :, k] if Fa.ndim == 3 else Fa)
            y = np.zeros((m, N), order="F")
            lib().sgo_forecast(m, mp, u(self.p), u(self.sp), d(fm), d(fC), d(fn), d(fs), N, self.seed,
                               e_draw, e_nu, d(x), d(lam), d(th), 1 if k > 0 else 0, d(y), self.nthreads)
            ys.append(y)
            if debug:
                ls.append(lam.copy(order="F"))
                ts.append(th.copy(order="F"))
        stack = (lambda v: v[0]) if Fa.ndim == 2 else (lambda v: np.stack(v, axis=-1))
        if debug:
            return {"y": stack(ys), "lambdas": stack(ls), "thetas": stack(ts)}
        return stack(ys)

    def computeForecast(self, nsim, nsim_batch, F_tp1):
        return self._forecast(nsim, F_tp1)

    def computeForecastDebug(self, nsim, nsim_batch, F_tp1):
        return self._forecast(nsim, F_tp1, debug=True)

    def computeEvoForecast(self, nsim, nsim_batch, F_tp1, G_tp1):
        return self._forecast(nsim, F_tp1, G_tp1)

    def computeEvoForecastDebug(self, nsim, nsim_batch, F_tp1, G_tp1):
        return self._forecast(nsim, F_tp1, G_tp1, debug=True)
