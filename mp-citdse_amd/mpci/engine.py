"""Python mirror of the tdse propagation interface over the C ABI.

`System` holds the device-resident operator (what tdrd::readEnergies /
readDipoles load, src/tdse/td_read.cpp:50-120); `Propagator` runs P pulses
on it.  `Propagator.prop_fixed` is the reference mode of tdse::propV /
propL (src/tdse/tdse.cpp:146-310): fixed steps, field frozen at t + dt,
renormalisation after every step.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import MpciPulse, MpciStats, check, dptr


def to_au(Io_W: float, w_eV: float):
    a, b = C.c_double(), C.c_double()
    _lib.tdse_lib().mpci_to_au(Io_W, w_eV, C.byref(a), C.byref(b))
    return a.value, b.value


def period(w: float, cycles: int) -> float:
    return _lib.tdse_lib().mpci_period(w, cycles)


class Pulse:
    """pulse::params (src/tdse/include/pulse.hpp:20-26) + PWL extension."""

    def __init__(self, c: MpciPulse, env=None):
        self.c = c
        self._env = env

    @staticmethod
    def sine(Io_au, w_au, cycles, gauge="v", cepd=0.0):
        c = MpciPulse()
        check(_lib.tdse_lib().mpci_pulse_sine(Io_au, w_au, cepd, cycles,
                                              gauge.encode(), C.byref(c)))
        return Pulse(c)

    @staticmethod
    def gauss(Io_au, w_au, tau, cycles, gauge="v", cepd=0.0):
        c = MpciPulse()
        check(_lib.tdse_lib().mpci_pulse_gauss(Io_au, w_au, tau, cycles, cepd,
                                               gauge.encode(), C.byref(c)))
        return Pulse(c)

    @staticmethod
    def pwl(env, w_au, t_end, gauge="v", cepd=0.0):
        env = np.ascontiguousarray(env, dtype=np.float64)
        c = MpciPulse()
        c.kind = _lib.PULSE_PWL
        c.w = w_au
        c.cepd = cepd
        c.t_end = t_end
        c.scale = 1.0 / w_au if gauge == "v" else 1.0
        c.n_env = len(env)
        c.env = dptr(env)
        return Pulse(c, env)

    def __call__(self, t: float) -> float:
        return _lib.tdse_lib().mpci_pulse_eval(C.byref(self.c), t)


class System:
    """Device-resident block-tridiagonal operator."""

    def __init__(self, E, D, gauge="v"):
        lib = _lib.tdse_lib()
        self.E = [np.ascontiguousarray(e, dtype=np.float64) for e in E]
        D = [np.ascontiguousarray(d, dtype=np.float64) for d in D]
        self.s = np.array([len(e) for e in self.E], dtype=np.int32)
        nb = len(self.E)
        for L, d in enumerate(D):
            if d.shape != (self.s[L], self.s[L + 1]):
                raise ValueError(f"dipole block {L} has shape {d.shape}")
        self.N = int(self.s.sum())
        self.offs = np.concatenate([[0], np.cumsum(self.s)]).astype(np.int64)
        self.gauge = gauge
        Ep = (_lib._dp * nb)(*[dptr(e) for e in self.E])
        Dp = (_lib._dp * max(nb - 1, 1))(*[dptr(d) for d in D])
        h = C.c_void_p()
        check(lib.mpci_system_create(nb, self.s.ctypes.data_as(_lib._ip),
                                     gauge.encode(), Ep, Dp, C.byref(h)))
        self.h = h

    @property
    def dipole_elements(self) -> int:
        return int(sum(int(self.s[L]) * int(self.s[L + 1])
                       for L in range(len(self.s) - 1)))

    def set_gemv_grid(self, workgroups: int):
        """Workgroups of the persistent single-pulse coupling kernel (0:
        the default, 2 per compute unit)."""
        check(_lib.tdse_lib().mpci_system_set_gemv_grid(self.h, workgroups))

    def close(self):
        if self.h:
            _lib.tdse_lib().mpci_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Propagator:
    def __init__(self, system: System, P: int = 1):
        self.sys = system
        self.P = P
        h = C.c_void_p()
        check(_lib.tdse_lib().mpci_prop_create(system.h, P, C.byref(h)))
        self.h = h
        self._pulses = [None] * P

    def close(self):
        if self.h:
            _lib.tdse_lib().mpci_prop_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- state ------------------------------------------------------------
    def set_state(self, ct, p: int = 0):
        ct = np.ascontiguousarray(ct, dtype=np.float64)
        if ct.shape != (2 * self.sys.N,):
            raise ValueError("state must hold 2N doubles")
        check(_lib.tdse_lib().mpci_prop_set_state(self.h, p, dptr(ct)))

    def get_state(self, p: int = 0) -> np.ndarray:
        out = np.empty(2 * self.sys.N)
        check(_lib.tdse_lib().mpci_prop_get_state(self.h, p, dptr(out)))
        return out

    def set_pulse(self, pulse: Pulse, p: int = 0):
        self._pulses[p] = pulse
        check(_lib.tdse_lib().mpci_prop_set_pulse(self.h, p, C.byref(pulse.c)))

    # --- compute ----------------------------------------------------------
    def rhs(self, F, x) -> np.ndarray:
        F = np.ascontiguousarray(np.broadcast_to(F, (self.P,)), dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        out = np.empty_like(x)
        check(_lib.tdse_lib().mpci_prop_rhs(self.h, dptr(F), dptr(x),
                                            dptr(out)))
        return out

    def prop_fixed(self, t0, dt, F, track=0):
        """F: (P, steps) frozen per-step fields. Returns pop (P, steps)."""
        F = np.ascontiguousarray(np.atleast_2d(F), dtype=np.float64)
        steps = F.shape[1]
        pop = np.empty((self.P, steps))
        check(_lib.tdse_lib().mpci_prop_fixed(self.h, t0, dt, steps, dptr(F),
                                              track, dptr(pop)))
        return pop

    def prop_adaptive(self, t0, t1, h0, atol, rtol, max_attempts=10_000_000,
                      track=0, log_cap=0, allow_limit=False):
        """Adaptive propagation; with allow_limit a run that stops at
        max_attempts returns normally (stats[p].status reports it)."""
        P = self.P
        t0 = np.ascontiguousarray(np.broadcast_to(t0, (P,)), dtype=np.float64)
        t1 = np.ascontiguousarray(np.broadcast_to(t1, (P,)), dtype=np.float64)
        h0 = np.ascontiguousarray(np.broadcast_to(h0, (P,)), dtype=np.float64)
        st = (MpciStats * P)()
        log = np.zeros((P, max(log_cap, 1), 3))
        rc = _lib.tdse_lib().mpci_prop_adaptive(
            self.h, dptr(t0), dptr(t1), dptr(h0), atol, rtol, max_attempts,
            track, st, dptr(log) if log_cap > 0 else None, log_cap)
        if not (allow_limit and rc == _lib.MPCI_ERR_LIMIT):
            check(rc)
        return list(st), log[:, :log_cap]

    def prop_schedule(self, t0, hs):
        hs = np.ascontiguousarray(np.atleast_2d(hs), dtype=np.float64)
        t0 = np.ascontiguousarray(np.broadcast_to(t0, (self.P,)),
                                  dtype=np.float64)
        check(_lib.tdse_lib().mpci_prop_schedule(self.h, dptr(t0), hs.shape[1],
                                                 dptr(hs)))

    def adjoint(self, t0, t1, h0, W, atol=1e-10, rtol=1e-10, max_steps=100000,
                ckpt_every=32, hlog_cap=0):
        """Discrete-adjoint gradient of J_p = sum_k W_k |c_{k,p}(t1)|^2 with
        respect to each pulse's PWL envelope samples.  Returns (J, grad,
        nsteps, hlog, stats)."""
        P = self.P
        t0 = np.ascontiguousarray(np.broadcast_to(t0, (P,)), dtype=np.float64)
        t1 = np.ascontiguousarray(np.broadcast_to(t1, (P,)), dtype=np.float64)
        h0 = np.ascontiguousarray(np.broadcast_to(h0, (P,)), dtype=np.float64)
        W = np.ascontiguousarray(W, dtype=np.float64)
        if W.shape != (self.sys.N,):
            raise ValueError("W must hold N weights")
        nenv = max(int(p.c.n_env) for p in self._pulses)
        J = np.empty(P)
        grad = np.zeros((P, nenv))
        nst = np.zeros(P, dtype=np.int32)
        cap = max(hlog_cap, 1)
        hlog = np.zeros((P, cap))
        st = _lib.MpciAdjointStats()
        check(_lib.tdse_lib().mpci_prop_adjoint(
            self.h, dptr(t0), dptr(t1), dptr(h0), atol, rtol, max_steps,
            dptr(W), ckpt_every, dptr(J), dptr(grad), nenv,
            nst.ctypes.data_as(_lib._ip), dptr(hlog) if hlog_cap else None,
            hlog_cap, C.byref(st)))
        return J, grad, nst, hlog[:, :hlog_cap], st

    def populations(self, W) -> np.ndarray:
        W = np.ascontiguousarray(np.atleast_2d(W), dtype=np.float64)
        out = np.empty((W.shape[0], self.P))
        check(_lib.tdse_lib().mpci_prop_populations(self.h, W.shape[0],
                                                    dptr(W), dptr(out)))
        return out

    def observables(self, electrons: int, threshold: float = 0.0):
        """pes's scalars for every pulse's current state, on the device:
        an (OBS_ROWS, P) array with rows ground population, yield, excited
        population and norm (src/pes/pes.cpp:92-110, 138-146, 199-213)."""
        out = np.empty((_lib.OBS_ROWS, self.P))
        check(_lib.tdse_lib().mpci_prop_observables(self.h, int(electrons),
                                                    float(threshold),
                                                    dptr(out)))
        return out

    def set_graph(self, on: bool):
        check(_lib.tdse_lib().mpci_prop_set_graph(self.h, int(on)))

    def set_instrument(self, on: bool):
        check(_lib.tdse_lib().mpci_prop_set_instrument(self.h, int(on)))

    def kernel_time(self):
        """(total ms, launches, bytes per launch, flops per launch) of the
        coupling kernel over the instrumented attempts."""
        ms = C.c_double()
        n = C.c_longlong()
        b = C.c_longlong()
        f = C.c_longlong()
        check(_lib.tdse_lib().mpci_prop_kernel_time(self.h, C.byref(ms),
                                                    C.byref(n), C.byref(b),
                                                    C.byref(f)))
        return ms.value, n.value, b.value, f.value

    def gemm_config(self):
        wm, wn, ks, kind, tk, fm, fn = (C.c_int() for _ in range(7))
        ms = C.c_double()
        check(_lib.tdse_lib().mpci_prop_gemm_config(
            self.h, C.byref(wm), C.byref(wn), C.byref(ks), C.byref(ms),
            C.byref(kind), C.byref(tk), C.byref(fm), C.byref(fn)))
        return {"kernel": ("k_gemm_coupling", "k_gemm_db", "k_gemm_lds",
                           "k_gemm_pair")[kind.value],
                "wm": wm.value, "wn": wn.value, "fm": fm.value,
                "fn": fn.value, "tk": tk.value, "ksplit": ks.value,
                "ms": ms.value}

    @staticmethod
    def gemm_candidates():
        """Number of candidate tiles the batched-form autotuner searches."""
        n = C.c_int()
        check(_lib.tdse_lib().mpci_prop_gemm_candidates(C.byref(n)))
        return n.value

    def set_small_path(self, mode=-1):
        """One pulse on a small system: fused coupling + stage launches.
        mode -1 automatic, 0 off, 1 on where it applies.  Returns whether
        the path is active."""
        on = C.c_int()
        check(_lib.tdse_lib().mpci_prop_set_small_path(self.h, int(mode),
                                                       C.byref(on)))
        return bool(on.value)

    def set_gemm_candidate(self, index):
        """Pin candidate tile `index` in place of the autotuned one."""
        check(_lib.tdse_lib().mpci_prop_set_gemm_candidate(self.h, int(index)))
        return self.gemm_config()


def multi_adaptive(props, t0, t1, h0, atol, rtol, max_attempts=10_000_000,
                   allow_limit=False, grouped=False):
    """Adaptive propagation of several propagators (different systems and
    batch sizes) concurrently on one device; equal to separate
    Propagator.prop_adaptive calls.  t0, t1, h0: one scalar or P-vector per
    member.  grouped: every member's attempt in the same launches
    (batched members only); otherwise each member on its own stream.
    Returns one list of per-pulse stats per member."""
    n = len(props)
    keep, st = [], []

    def vec(v, P):
        a = np.ascontiguousarray(np.broadcast_to(v, (P,)), dtype=np.float64)
        keep.append(a)
        return dptr(a)

    hs = (_lib._vp * n)(*[p.h.value for p in props])
    a0 = (_lib._dp * n)(*[vec(t0[i], props[i].P) for i in range(n)])
    a1 = (_lib._dp * n)(*[vec(t1[i], props[i].P) for i in range(n)])
    ah = (_lib._dp * n)(*[vec(h0[i], props[i].P) for i in range(n)])
    for p in props:
        st.append((MpciStats * p.P)())
    sp = (C.POINTER(MpciStats) * n)(
        *[C.cast(s, C.POINTER(MpciStats)) for s in st])
    fn = (_lib.tdse_lib().mpci_multi_adaptive_grouped if grouped
          else _lib.tdse_lib().mpci_multi_adaptive)
    rc = fn(n, hs, a0, a1, ah, atol, rtol, max_attempts, sp)
    if not (allow_limit and rc == _lib.MPCI_ERR_LIMIT):
        check(rc)
    return [list(s) for s in st]


def obs_weights(E, electrons: int, threshold: float = 0.0) -> np.ndarray:
    """Per-state weights of pes's scalars (mpci_obs_weights): an
    (OBS_ROWS, N) array; observable o of a state c is W[o] @ |c|^2.  E: the
    energies of each block as tdse reads them.  Host code, no GPU needed."""
    E = [np.ascontiguousarray(e, dtype=np.float64) for e in E]
    s = np.array([len(e) for e in E], dtype=np.int32)
    Ep = (_lib._dp * len(E))(*[dptr(e) for e in E])
    W = np.empty((_lib.OBS_ROWS, int(s.sum())))
    check(_lib.tdse_lib().mpci_obs_weights(len(E), s.ctypes.data_as(_lib._ip),
                                           Ep, int(electrons),
                                           float(threshold), dptr(W)))
    return W


def device_count() -> int:
    n = C.c_int()
    check(_lib.tdse_lib().mpci_device_count(C.byref(n)))
    return n.value


def probe_hbm_read(nbytes: int = 384 << 20, reps: int = 20) -> float:
    """Best HBM read rate in GB/s of a non-temporal streaming kernel over
    nbytes (mpci_probe_hbm_read): the ceiling reached in practice."""
    v = C.c_double()
    check(_lib.tdse_lib().mpci_probe_hbm_read(int(nbytes), int(reps),
                                              C.byref(v)))
    return v.value


def probe_mfma_f64(iters: int = 4096) -> float:
    """Best FP64 MFMA rate in TFLOP/s of back-to-back v_mfma_f64_16x16x4f64
    chains with no memory traffic (mpci_probe_mfma_f64)."""
    v = C.c_double()
    check(_lib.tdse_lib().mpci_probe_mfma_f64(int(iters), C.byref(v)))
    return v.value
